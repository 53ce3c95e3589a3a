"""Tessellations of [0,1]^ndim into simplices, the CPA constraint matrix L and
the basis B = null(L).

Restates libcpab/core/tesselation.py:13-442 and libcpab/core/utility.py:27-42.
The rows of L are produced in the same order as the reference (continuity rows
for every pair of cells that share ndim vertices, then the auxiliary
"outside" rows, then zero-boundary rows, then zero-trace rows) and the basis is
the same SVD-based null space with the same threshold, so that a theta vector
means the same velocity field as it does in the reference on the same LAPACK.

Besides L and B this module derives the geometry the integration kernels use:
for every cell a list of facets (outward plane in scaled coordinates
q = x * nc and the id of the cell across the facet).  That table is this
build's own; the reference has no facet structure because it never computes
boundary-hit times.
"""
import itertools
import numpy as np
import scipy.linalg as la

# Largest number of facets a cell can have in the facet table (2D boundary
# triangles grow to 4 facets when their region is extended outside [0,1]^2).
NF_MAX = 4


def null_space(L, eps=1e-6):
    """Null space of L from the SVD (libcpab/core/utility.py:27-42).

    Right-singular vectors whose singular value is <= eps, plus the rows of
    vh beyond len(s) when L has fewer rows than columns.  Returns [D, d].
    """
    _, s, vh = la.svd(L)
    pad = max(0, L.shape[1] - s.shape[0])
    mask = np.concatenate([s <= eps, np.ones(pad, dtype=bool)])
    return vh[mask].T.copy()


def _hashable(arr):
    # tuple-of-tuples form of a vertex array, so that sets of vertices can be
    # intersected exactly as libcpab/core/utility.py:45-49 does
    return tuple(tuple(r.tolist()) for r in arr)


class _TessBase:
    """Shared constraint construction (libcpab/core/tesselation.py:13-162)."""

    ndim = None
    n_params = None

    def __init__(self, nc, domain_min, domain_max, zero_boundary=True,
                 volume_perservation=False):
        self.nc = list(nc)
        self.domain_min = list(domain_min)
        self.domain_max = list(domain_max)
        self.zero_boundary = bool(zero_boundary)
        self.volume_perservation = bool(volume_perservation)
        self.find_verts()
        self.find_shared_verts()
        if not self.zero_boundary:
            self.find_verts_outside()
        L = self.continuity_rows()
        if self.zero_boundary:
            L = np.concatenate([L, self.zero_boundary_rows()], axis=0)
        if self.volume_perservation:
            L = np.concatenate([L, self.zero_trace_rows()], axis=0)
        self.L = L
        self.B = null_space(L)

    # -- shared pieces -----------------------------------------------------
    def get_cell_centers(self):
        return np.mean(self.verts[:, :, :self.ndim], axis=1)

    def find_shared_verts(self):
        # libcpab/core/tesselation.py:109-126: ordered pairs (i, j), i != j,
        # that share exactly ndim vertices, each unordered pair once.
        hv = [_hashable(v) for v in self.verts]
        pairs, shared = [], []
        seen = set()
        for i in range(self.nC):
            si = set(hv[i])
            for j in range(self.nC):
                if i == j:
                    continue
                common = si.intersection(hv[j])
                if len(common) == self.ndim and (j, i) not in seen:
                    shared.append(list(common)[:self.ndim])
                    pairs.append((i, j))
                    seen.add((i, j))
        self.shared_v = np.asarray(shared)
        self.shared_v_idx = pairs

    def continuity_rows(self):
        # libcpab/core/tesselation.py:134-148: for every shared vertex and
        # every output component k, A_i[k] . v~ - A_j[k] . v~ = 0.
        n, p = self.ndim, self.n_params
        rows = []
        for idx, (i, j) in enumerate(self.shared_v_idx):
            for vidx in range(n):
                v = np.asarray(self.shared_v[idx][vidx], dtype=np.float64)
                for k in range(n):
                    r = np.zeros(p * self.nC)
                    a = p * int(i) + k * (n + 1)
                    b = p * int(j) + k * (n + 1)
                    r[a:a + n + 1] = v
                    r[b:b + n + 1] = -v
                    rows.append(r)
        if not rows:
            return np.zeros((0, p * self.nC))
        return np.stack(rows)

    def zero_trace_rows(self):
        # libcpab/core/tesselation.py:155-162
        n, p = self.ndim, self.n_params
        row = np.concatenate([np.eye(n), np.zeros((n, 1))], axis=1).ravel()
        L = np.zeros((self.nC, p * self.nC))
        for c in range(self.nC):
            L[c, p * c:p * (c + 1)] = row
        return L

    # -- geometry for the integrators ---------------------------------------
    def q_verts(self):
        """Cell vertices in scaled coordinates q = (x - min) / (max - min) * nc."""
        lo = np.asarray(self.domain_min, dtype=np.float64)
        hi = np.asarray(self.domain_max, dtype=np.float64)
        nc = np.asarray(self.nc, dtype=np.float64)
        return (self.verts[:, :, :self.ndim] - lo) / (hi - lo) * nc

    def facet_table(self):
        """Per-cell facets for boundary-hit integration.

        Returns (planes [nC + nV, NF_MAX, 4] float64, nbr [nC + nV, NF_MAX]
        int32).  plane = (n0, n1, n2, e) with G(q) = n . q + e, G <= 0 inside
        the cell region and |n|_2 = 1 in q-space.  nbr >= 0 is the cell across
        the facet, -1 a wall (no crossing: zero-velocity domain boundary), -2
        an unused slot.  nV virtual cells (3D without zero boundary: the
        regions outside the domain) follow the nC simplices; self.amap [nV]
        holds the simplex whose affine map each of them uses.
        """
        self.amap = np.zeros(0, dtype=np.int32)
        return _simplex_facets(self)


def _plane_through(points, inside):
    """Unit outward plane through `points` (ndim points in R^ndim)."""
    points = np.asarray(points, dtype=np.float64)
    ndim = points.shape[1]
    if ndim == 1:
        n = np.array([1.0])
    else:
        d = points[1:] - points[0]
        if ndim == 2:
            n = np.array([-d[0, 1], d[0, 0]])
        else:
            n = np.cross(d[0], d[1])
    n = n / np.linalg.norm(n)
    e = -float(n @ points[0])
    if n @ inside + e > 0:
        n, e = -n, -e
    return n, e


def _simplex_facets(tess):
    ndim = tess.ndim
    nC = tess.nC
    qv = tess.q_verts()
    scale = 2.0  # 2*q is integral for every vertex (cell centres sit at +0.5)
    key_of = lambda pts: tuple(sorted(tuple(int(round(c * scale)) for c in p) for p in pts))
    facets = {}
    per_cell = []
    for c in range(nC):
        V = qv[c]
        centroid = V.mean(axis=0)
        fl = []
        for m in range(ndim + 1):
            pts = np.delete(V, m, axis=0)
            n, e = _plane_through(pts, centroid)
            k = key_of(pts)
            facets.setdefault(k, []).append((c, len(fl)))
            fl.append([n, e, -1, k])
        per_cell.append(fl)
    for k, owners in facets.items():
        if len(owners) == 2:
            (c0, f0), (c1, f1) = owners
            per_cell[c0][f0][2] = c1
            per_cell[c1][f1][2] = c0
    if ndim in (1, 2):
        per_cell = _extend_outside(tess, per_cell)
    vplanes = vnbr = None
    if ndim == 3 and not tess.zero_boundary:
        vplanes, vnbr, tess.amap, per_cell = _outside_regions_3d(tess, per_cell, qv)
    nV = 0 if vplanes is None else vplanes.shape[0]
    planes = np.zeros((nC + nV, NF_MAX, 4))
    nbr = np.full((nC + nV, NF_MAX), -2, dtype=np.int32)
    for c, fl in enumerate(per_cell):
        assert len(fl) <= NF_MAX
        for f, (n, e, nb, _) in enumerate(fl):
            planes[c, f, :ndim] = n
            planes[c, f, 3] = e
            nbr[c, f] = nb
    if nV:
        planes[nC:] = vplanes
        nbr[nC:] = vnbr
    return planes, nbr


def _outside_regions_3d(tess, per_cell, qv):
    """Virtual cells outside [0,1]^3 (libcpab_amd.core.exterior3d), planes
    moved to q-space; the simplices' domain faces lead to them instead of
    being walls."""
    from .exterior3d import build_regions
    xplanes, vnbr, amap, face_nbr, _ = build_regions(tess)
    nc = np.asarray(tess.nc, dtype=np.float64)
    vplanes = np.zeros_like(xplanes)
    used = vnbr != -2
    # n . x + e with x = q / nc: (n / nc) . q + e, then unit normal in q
    nq = xplanes[:, :, :3] / nc
    norm = np.linalg.norm(nq, axis=2)
    norm[~used] = 1.0
    vplanes[:, :, :3] = np.where(used[:, :, None], nq / norm[:, :, None], 0.0)
    vplanes[:, :, 3] = np.where(used, xplanes[:, :, 3] / norm, 0.0)
    out = []
    for c, fl in enumerate(per_cell):
        fl2 = []
        for n, e, nb, k in fl:
            if nb < 0:
                # a domain face: all three vertices share a coordinate 0 or nc
                V = np.array(k, dtype=np.float64) / 2.0
                a = [ax for ax in range(3) if np.allclose(V[:, ax], V[0, ax])][0]
                s = 0 if np.isclose(V[0, a], 0.0) else 1
                nb = int(face_nbr[(c, a, s)])
            fl2.append([n, e, nb, k])
        out.append(fl2)
    return vplanes, vnbr.astype(np.int32), np.asarray(amap, dtype=np.int32), out


def _extend_outside(tess, per_cell):
    """Extend boundary cells so that the exterior is covered the way the
    reference's cell lookup assigns it (libcpab/numpy/findcellidx.py:22-27
    clamps in 1D; libcpab/numpy/findcellidx.py:30-80 in 2D: strips beyond
    an edge go to the edge's boundary triangle of that column/row, and the
    corner quadrants are split by the diagonals through the corners).

    The facet on the domain boundary is dropped and replaced by rays along
    the grid lines that separate neighbouring boundary cells.
    """
    ndim = tess.ndim
    out = []
    if ndim == 1:
        nx = tess.nc[0]
        for c, fl in enumerate(per_cell):
            out.append([f for f in fl if f[2] >= 0])
        return out
    nx, ny = tess.nc
    for c, fl in enumerate(per_cell):
        i, j, t = (c // 4) % nx, (c // 4) // nx, c % 4
        keep = [f for f in fl if f[2] >= 0]
        boundary = len(keep) < len(fl)
        if boundary:
            if t in (0, 2):  # top / bottom edge: rays x = i and x = i+1
                if i > 0:
                    keep.append([np.array([-1.0, 0.0]), float(i), 4 * ((i - 1) + j * nx) + t, None])
                if i < nx - 1:
                    keep.append([np.array([1.0, 0.0]), -float(i + 1), 4 * ((i + 1) + j * nx) + t, None])
            else:  # left / right edge: rays y = j and y = j+1
                if j > 0:
                    keep.append([np.array([0.0, -1.0]), float(j), 4 * (i + (j - 1) * nx) + t, None])
                if j < ny - 1:
                    keep.append([np.array([0.0, 1.0]), -float(j + 1), 4 * (i + (j + 1) * nx) + t, None])
        out.append(keep)
    return out


class Tessellation1D(_TessBase):
    ndim = 1
    n_params = 2

    def __init__(self, nc, domain_min, domain_max, zero_boundary=True,
                 volume_perservation=False):
        self.nC = int(np.prod(nc))
        super().__init__(nc, domain_min, domain_max, zero_boundary, volume_perservation)

    def find_verts(self):
        # libcpab/core/tesselation.py:177-191
        Vx = np.linspace(self.domain_min[0], self.domain_max[0], self.nc[0] + 1)
        self.verts = np.asarray([((Vx[i], 1), (Vx[i + 1], 1)) for i in range(self.nc[0])])
        self.cells = list(range(self.nc[0]))

    def find_verts_outside(self):
        pass  # 1D needs no auxiliary points (libcpab/core/tesselation.py:193-194)

    def zero_boundary_rows(self):
        # libcpab/core/tesselation.py:196-200
        L = np.zeros((2, 2 * self.nC))
        L[0, :2] = [self.domain_min[0], 1]
        L[1, -2:] = [self.domain_max[0], 1]
        return L


class Tessellation2D(_TessBase):
    ndim = 2
    n_params = 6

    def __init__(self, nc, domain_min, domain_max, zero_boundary=True,
                 volume_perservation=False):
        self.nC = 4 * int(np.prod(nc))
        super().__init__(nc, domain_min, domain_max, zero_boundary, volume_perservation)

    def find_verts(self):
        # libcpab/core/tesselation.py:215-245: four triangles per square,
        # ordered top, right, bottom, left, each (centre, v_a, v_b).
        Vx = np.linspace(self.domain_min[0], self.domain_max[0], self.nc[0] + 1)
        Vy = np.linspace(self.domain_min[1], self.domain_max[1], self.nc[1] + 1)
        verts, cells = [], []
        for i in range(self.nc[1]):
            for j in range(self.nc[0]):
                ul = (Vx[j], Vy[i], 1)
                ur = (Vx[j + 1], Vy[i], 1)
                ll = (Vx[j], Vy[i + 1], 1)
                lr = (Vx[j + 1], Vy[i + 1], 1)
                ce = ((Vx[j] + Vx[j + 1]) / 2, (Vy[i] + Vy[i + 1]) / 2, 1)
                verts += [(ce, ul, ur), (ce, ur, lr), (ce, lr, ll), (ce, ll, ul)]
                cells += [(j, i, t) for t in range(4)]
        self.verts = np.asarray(verts)
        self.cells = cells

    def find_verts_outside(self):
        # libcpab/core/tesselation.py:246-311: neighbouring boundary triangles
        # of the outer row/column that share one vertex get a second,
        # auxiliary shared point 10 units away along the grid line, which
        # keeps the extended fields continuous outside the domain.
        nx, ny = self.nc
        hv = [_hashable(v) for v in self.verts]
        extra_v, extra_idx = [], []
        seen = set()
        for i in range(self.nC):
            si = set(hv[i])
            mi = self.cells[i]
            for j in range(self.nC):
                mj = self.cells[j]
                common = si.intersection(hv[j])
                lr_side = ((mi[0] == mj[0] == 0 and mi[2] == mj[2] == 3) or
                           (mi[0] == mj[0] == nx - 1 and mi[2] == mj[2] == 1)) and abs(mi[1] - mj[1]) == 1
                tb_side = ((mi[1] == mj[1] == 0 and mi[2] == mj[2] == 0) or
                           (mi[1] == mj[1] == ny - 1 and mi[2] == mj[2] == 2)) and abs(mi[0] - mj[0]) == 1
                if len(common) == 1 and (lr_side or tb_side) and (j, i) not in seen:
                    v = list(common)[0]
                    aux = list(v)
                    if lr_side:
                        aux[0] -= 10
                    else:
                        aux[1] -= 10
                    extra_v.append([tuple(common)[0], tuple(aux)])
                    extra_idx.append((i, j))
                    seen.add((i, j))
        if extra_v:
            self.shared_v = np.concatenate([self.shared_v, np.asarray(extra_v)])
            self.shared_v_idx = list(self.shared_v_idx) + extra_idx

    def zero_boundary_rows(self):
        # libcpab/core/tesselation.py:313-327: the normal velocity component
        # vanishes at every boundary vertex.
        xmin, ymin = self.domain_min
        xmax, ymax = self.domain_max
        rows = []
        for c in range(self.nC):
            for v in self.verts[c]:
                if v[0] == xmin or v[0] == xmax:
                    r = np.zeros(6 * self.nC)
                    r[6 * c:6 * c + 3] = v
                    rows.append(r)
                if v[1] == ymin or v[1] == ymax:
                    r = np.zeros(6 * self.nC)
                    r[6 * c + 3:6 * c + 6] = v
                    rows.append(r)
        return np.stack(rows) if rows else np.zeros((0, 6 * self.nC))


class Tessellation3D(_TessBase):
    ndim = 3
    n_params = 12

    def __init__(self, nc, domain_min, domain_max, zero_boundary=True,
                 volume_perservation=False):
        self.nC = 5 * int(np.prod(nc))
        super().__init__(nc, domain_min, domain_max, zero_boundary, volume_perservation)

    def find_verts(self):
        # libcpab/core/tesselation.py:342-383: five tetrahedra per cube; cubes
        # with odd x+y+z have their corner labels rotated a quarter turn.
        Vx = np.linspace(self.domain_min[0], self.domain_max[0], self.nc[0] + 1)
        Vy = np.linspace(self.domain_min[1], self.domain_max[1], self.nc[1] + 1)
        Vz = np.linspace(self.domain_min[2], self.domain_max[2], self.nc[2] + 1)
        verts, cells = [], []
        for i in range(self.nc[2]):
            for j in range(self.nc[1]):
                for k in range(self.nc[0]):
                    ul0 = (Vx[k], Vy[j], Vz[i], 1)
                    ur0 = (Vx[k + 1], Vy[j], Vz[i], 1)
                    ll0 = (Vx[k], Vy[j + 1], Vz[i], 1)
                    lr0 = (Vx[k + 1], Vy[j + 1], Vz[i], 1)
                    ul1 = (Vx[k], Vy[j], Vz[i + 1], 1)
                    ur1 = (Vx[k + 1], Vy[j], Vz[i + 1], 1)
                    ll1 = (Vx[k], Vy[j + 1], Vz[i + 1], 1)
                    lr1 = (Vx[k + 1], Vy[j + 1], Vz[i + 1], 1)
                    if (i + j + k) % 2 == 1:
                        ul0, ur0, lr0, ll0 = ur0, lr0, ll0, ul0
                        ul1, ur1, lr1, ll1 = ur1, lr1, ll1, ul1
                    verts += [(ll1, ur1, ul0, lr0), (ul1, ur1, ll1, ul0),
                              (lr1, ur1, ll1, lr0), (ll0, ul0, lr0, ll1),
                              (ur0, ul0, lr0, ur1)]
                    cells += [(k, j, i, l) for l in range(5)]
        self.verts = np.asarray(verts)
        self.cells = cells

    def find_verts_outside(self):
        # libcpab/core/tesselation.py:385-422: pairs of tetrahedra of one cube
        # that both have a face on the same domain face get an auxiliary
        # shared point one unit outside, making the extension continuous.
        extra_v, extra_idx = [], []
        seen = set()
        for i in range(self.nC):
            for j in range(self.nC):
                if i == j:
                    continue
                for d in range(self.ndim):
                    vi, vj = self.verts[i], self.verts[j]
                    ci, cj = self.cells[i], self.cells[j]
                    on_min = (np.sum(vi[:, d] == self.domain_min[d]) == 3 and
                              np.sum(vj[:, d] == self.domain_min[d]) == 3)
                    on_max = (np.sum(vi[:, d] == self.domain_max[d]) == 3 and
                              np.sum(vj[:, d] == self.domain_max[d]) == 3)
                    same_cube = tuple(ci[:3]) == tuple(cj[:3])
                    if (on_min or on_max) and same_cube and (j, i) not in seen:
                        hi_, hj_ = _hashable(vi), _hashable(vj)
                        sv = set(hi_).intersection(hj_)
                        centre = [(a + b) / 2.0 for a, b in zip(hi_[0], hj_[0])]
                        centre[d] += -1 if on_min else 1
                        extra_v.append(list(sv.union([tuple(centre)])))
                        extra_idx.append((i, j))
                        seen.add((i, j))
        if extra_v:
            self.shared_v = np.concatenate([self.shared_v, np.asarray(extra_v)])
            self.shared_v_idx = list(self.shared_v_idx) + extra_idx

    def zero_boundary_rows(self):
        # libcpab/core/tesselation.py:424-442
        lo, hi = self.domain_min, self.domain_max
        rows = []
        for c in range(self.nC):
            for v in self.verts[c]:
                for k in range(3):
                    if v[k] == lo[k] or v[k] == hi[k]:
                        r = np.zeros(12 * self.nC)
                        r[12 * c + 4 * k:12 * c + 4 * k + 4] = v
                        rows.append(r)
        return np.stack(rows) if rows else np.zeros((0, 12 * self.nC))


TESSELLATIONS = {1: Tessellation1D, 2: Tessellation2D, 3: Tessellation3D}


def make_tessellation(nc, zero_boundary=True, volume_perservation=False,
                      domain_min=None, domain_max=None):
    ndim = len(nc)
    domain_min = [0.0] * ndim if domain_min is None else list(domain_min)
    domain_max = [1.0] * ndim if domain_max is None else list(domain_max)
    return TESSELLATIONS[ndim](list(nc), domain_min, domain_max, zero_boundary,
                               volume_perservation)
