"""The outside of [0,1]^3 for 3D tessellations with zero_boundary=False.

The reference gives a point outside the domain the field of the tetrahedron
its lookup returns after pushing the point onto the boundary
(libcpab/numpy/findcellidx.py:84-109): v(x) = A_c [x; 1] with c =
lookup(push(x)).  For the closed-form integrator that assignment is turned
into convex regions with facets, like the cells inside:

* face prisms: beyond one domain face, over one boundary triangle (the face
  of a boundary tetrahedron), bounded by the face and the three planes
  through the triangle's edges perpendicular to the face;
* half-wedges: beyond two faces, over one cell-long segment of the domain
  edge they meet at, split by the plane where both excesses are equal;
* cones: beyond three faces at a domain corner, one per ordering of the
  three excesses (the push rule depends on that ordering).

Each region is a "virtual cell" numbered after the nC tetrahedra, with up to
4 facets, the cells across them, and the tetrahedron whose affine map it
uses (amap) -- the lookup of the pushed position of a point inside it.  The
boundary faces of the tetrahedra lead to the face prisms instead of being
walls.  The excess of axis a beyond side s is e_a = x_a - 1 (s = 1) or -x_a
(s = 0).
"""
import itertools

import numpy as np


# ---------------------------------------------------------------------------
# the numpy backend's 3D lookup (libcpab/numpy/findcellidx.py:84-149), FP64
# ---------------------------------------------------------------------------
def push_onto_boundary(p, nc):
    """Points [3, n] outside [0,1]^3 moved onto the boundary as the reference
    does before its lookup: beyond a face -> onto the face, except that the
    axis whose distance from the centre is the smallest of the three (by the
    reference's comparisons, which for y and z compare against x) stops half
    a cell inside the face."""
    p = np.array(p, dtype=np.float64, copy=True)
    out = np.any((p < 0.0) | (p > 1.0), axis=0)
    if not out.any():
        return p
    inc = 1.0 / np.asarray(nc, dtype=np.float64)
    d = p[:, out] - 0.5
    ad = np.abs(d)
    push = np.zeros_like(d)
    push[0] = np.where((ad[0] < ad[1]) & (ad[0] < ad[2]), 0.5 * inc[0], 0.0)
    push[1] = np.where((ad[1] < ad[0]) & (ad[0] < ad[2]), 0.5 * inc[1], 0.0)
    push[2] = np.where((ad[2] < ad[0]) & (ad[0] < ad[1]), 0.5 * inc[2], 0.0)
    d = np.where(ad > 0.5, np.copysign(0.5 - push, d), d)
    p[:, out] = d + 0.5
    return p


def lookup3(p, nc):
    """Tetrahedron of each point [3, n] (numpy-backend rules, FP64)."""
    nx, ny, nz = nc
    p = push_onto_boundary(p, nc)
    inc = 1.0 / np.array([nx, ny, nz], dtype=np.float64)
    top = np.array([nx, ny, nz]) * inc - 1e-8
    q = np.minimum(top[:, None], np.maximum(0.0, p))
    m = np.mod(q, inc[:, None])
    idx3 = []
    for a, n in enumerate((nx, ny, nz)):
        b = (q[a] - m[a]) / inc[a]
        idx3.append(np.where(n - 1 < b, n - 1, np.round(b)))
    i, j, k = idx3
    u, v, w = m / inc[:, None]
    odd = ((i + j + k) % 2) == 1
    u, v = np.where(odd, v, u), np.where(odd, 1 - u, v)
    sub = np.select([-u - v + w >= 0, u + v + w - 2 >= 0, -u + v - w >= 0, u - v - w >= 0], [1, 2, 3, 4], 0)
    return (5 * (i + j * nx + k * nx * ny) + sub).astype(np.int64)


# ---------------------------------------------------------------------------
# regions
# ---------------------------------------------------------------------------
def _excess_plane(a, s, sign):
    """Half-space sign * e_a <= 0 as (n, e) with n . x + e <= 0."""
    n = np.zeros(3)
    if s == 1:  # e = x_a - 1
        n[a] = sign
        return n, -sign
    n[a] = -sign  # e = -x_a
    return n, 0.0


def _diff_plane(a, sa, b, sb):
    """Half-space e_a - e_b <= 0."""
    na, ea = _excess_plane(a, sa, 1.0)
    nb, eb = _excess_plane(b, sb, 1.0)
    return na - nb, ea - eb


def build_regions(tess):
    """Virtual cells of a 3D tessellation without zero boundary.

    Returns (planes [nV, 4, 4] in x-space (n0, n1, n2, e; unused slots zero),
    nbr [nV, 4] (cell ids, -2 unused), amap [nV], face_nbr {(tet, facet
    slot): virtual id} for the tetrahedra's boundary facets, keys [nV])."""
    nc = list(tess.nc)
    V = tess.verts[:, :, :3]
    nC = tess.nC
    regions = []  # dicts: key, planes [(n, e, nbr_key)], probe
    index = {}

    def add(key, planes, probe):
        index[key] = len(regions)
        regions.append({"key": key, "planes": planes, "probe": np.asarray(probe, dtype=np.float64)})

    face_tris = {}  # (a, s) -> list of (tet, verts [3, 3])
    for c in range(nC):
        for a in range(3):
            for s in (0, 1):
                on = np.isclose(V[c, :, a], float(s))
                if on.sum() == 3:
                    face_tris.setdefault((a, s), []).append((c, V[c][on]))

    def seg_of(b, lo):
        return int(round(lo * nc[b]))

    # face prisms
    for (a, s), tris in face_tris.items():
        out = np.zeros(3)
        out[a] = 1.0 if s == 1 else -1.0
        for c, T in tris:
            planes = [(*_excess_plane(a, s, -1.0), ("tet", c))]
            cen = T.mean(axis=0)
            for m in range(3):
                u, w = np.delete(T, m, axis=0)
                n = np.cross(w - u, out)
                n /= np.linalg.norm(n)
                e = -float(n @ u)
                if n @ cen + e > 0:
                    n, e = -n, -e
                # the other coordinate both edge ends share at a domain face, if any
                edge_key = None
                for a2 in range(3):
                    if a2 == a:
                        continue
                    for s2 in (0, 1):
                        if np.isclose(u[a2], s2) and np.isclose(w[a2], s2):
                            b = 3 - a - a2
                            lo = min(u[b], w[b])
                            edge_key = ("wedge", frozenset([(a, s), (a2, s2)]), b, seg_of(b, lo), a2)
                if edge_key is None:
                    edge_key = ("prism-edge", (a, s), frozenset([tuple(np.round(u, 9)), tuple(np.round(w, 9))]), c)
                planes.append((n, e, edge_key))
            add(("prism", (a, s), c), planes, cen + 0.5 * out)

    # half-wedges: beyond (a1, s1) and (a2, s2), segment k along b; the half
    # in which e_small <= e_other (crossing e_small = 0 leads to the prism
    # beyond the other face)
    for a1, a2 in ((0, 1), (0, 2), (1, 2)):
        b = 3 - a1 - a2
        for s1 in (0, 1):
            for s2 in (0, 1):
                fs = frozenset([(a1, s1), (a2, s2)])
                side = {a1: s1, a2: s2}
                for k in range(nc[b]):
                    for small in (a1, a2):
                        other = a2 if small == a1 else a1
                        lo, hi = k / nc[b], (k + 1) / nc[b]
                        nlo = np.zeros(3)
                        nlo[b] = -1.0
                        nhi = np.zeros(3)
                        nhi[b] = 1.0
                        planes = [
                            (*_excess_plane(small, side[small], -1.0), ("prism-at-edge", (other, side[other]), fs, b, k)),
                            (*_diff_plane(small, side[small], other, side[other]), ("wedge", fs, b, k, other)),
                            (nlo, lo, ("wedge", fs, b, k - 1, small) if k > 0 else
                             ("cone", frozenset([(a1, s1), (a2, s2), (b, 0)]), (b, small, other))),
                            (nhi, -hi, ("wedge", fs, b, k + 1, small) if k < nc[b] - 1 else
                             ("cone", frozenset([(a1, s1), (a2, s2), (b, 1)]), (b, small, other))),
                        ]
                        probe = np.zeros(3)
                        probe[b] = 0.5 * (lo + hi)
                        for ax, ex in ((small, 0.3), (other, 0.6)):
                            probe[ax] = 1.0 + ex if side[ax] == 1 else -ex
                        add(("wedge", fs, b, k, small), planes, probe)

    # cones: beyond all three faces at a corner, ordering e_p <= e_q <= e_r
    for sides in itertools.product((0, 1), repeat=3):
        fs = frozenset((a, sides[a]) for a in range(3))
        for p, q, r in itertools.permutations(range(3)):
            seg_p = 0 if sides[p] == 0 else nc[p] - 1
            planes = [
                (*_excess_plane(p, sides[p], -1.0), ("wedge", frozenset([(q, sides[q]), (r, sides[r])]), p, seg_p, q)),
                (*_diff_plane(p, sides[p], q, sides[q]), ("cone", fs, (q, p, r))),
                (*_diff_plane(q, sides[q], r, sides[r]), ("cone", fs, (p, r, q))),
            ]
            probe = np.zeros(3)
            for ax, ex in ((p, 0.2), (q, 0.5), (r, 0.8)):
                probe[ax] = 1.0 + ex if sides[ax] == 1 else -ex
            add(("cone", fs, (p, q, r)), planes, probe)

    # prisms adjacent to a domain edge segment: the prism beyond face f whose
    # triangle has an edge on that segment
    edge_prism = {}
    for reg in regions:
        if reg["key"][0] != "prism":
            continue
        for n, e, nk in reg["planes"][1:]:
            if nk[0] == "wedge":
                _, fs, b, k, a2 = nk
                face = reg["key"][1]
                edge_prism[(face, fs, b, k)] = index[reg["key"]]

    # facets shared by two prisms of one face: key by the edge's end points
    edge_owner = {}
    for reg in regions:
        if reg["key"][0] != "prism":
            continue
        for n, e, nk in reg["planes"][1:]:
            if nk[0] == "prism-edge":
                edge_owner.setdefault((nk[1], nk[2]), []).append(index[reg["key"]])

    def resolve(nk, me):
        kind = nk[0]
        if kind == "tet":
            return nk[1]
        if kind == "prism-edge":
            owners = edge_owner[(nk[1], nk[2])]
            other = [o for o in owners if o != me]
            assert len(other) == 1, (nk, owners)
            return nC + other[0]
        if kind == "prism-at-edge":
            _, face, fs, b, k = nk
            return nC + edge_prism[(face, fs, b, k)]
        return nC + index[nk]

    nV = len(regions)
    planes = np.zeros((nV, 4, 4))
    nbr = np.full((nV, 4), -2, dtype=np.int64)
    for v, reg in enumerate(regions):
        for f, (n, e, nk) in enumerate(reg["planes"]):
            planes[v, f, :3] = n
            planes[v, f, 3] = e
            nbr[v, f] = resolve(nk, v)
    probes = np.stack([r["probe"] for r in regions], axis=1)
    amap = lookup3(probes, nc)
    face_nbr = {}
    for (a, s), tris in face_tris.items():
        for c, T in tris:
            face_nbr[(c, a, s)] = nC + index[("prism", (a, s), c)]
    return planes, nbr, amap, face_nbr, [r["key"] for r in regions]


def region_of(points, nc, planes, nC):
    """Virtual cell of each point [3, n] outside [0,1]^3 (-1 inside): the
    region whose largest facet violation (x-space planes, n . x + e) is the
    smallest, i.e. the region containing the point."""
    p = np.asarray(points, dtype=np.float64)
    inside = np.all((p >= 0.0) & (p <= 1.0), axis=0)
    G = np.einsum("vfa,an->vfn", planes[:, :, :3], p) + planes[:, :, 3][:, :, None]
    used = np.any(planes[:, :, :3] != 0, axis=2)
    G = np.where(used[:, :, None], G, -np.inf)
    worst = G.max(axis=1)  # [nV, n]
    v = np.argmin(worst, axis=0)
    return np.where(inside, -1, nC + v)
