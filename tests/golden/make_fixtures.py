"""Writes the committed golden fixtures under tests/golden/.

Run in the build container (needs /root/reference for the reference's own
CPU op, compiled by oracle/Makefile into oracle/_ref/):

    make -C oracle && python tests/golden/make_fixtures.py [--force]

Existing files are kept unless --force is given (a rewrite changes the
archives' timestamps even when the numbers are the same).

Fixtures (all .npz, no pickles):
  basis_<name>.npz   L-derived basis B [D, d] (float64), facet planes and
                     neighbours of the tessellations the tests use (and
                     amap, the simplex of each region outside the domain,
                     for 3D without zero boundary).  theta
                     only has a meaning together with B, so every test that
                     compares against stored numbers uses these B.
  ref_step_<name>.npz  inputs and outputs of the reference's compiled
                     cpab_forward_op / cpab_backward_op
                     (libcpab/core/cpab_ops.cpp:224-372) at nstepsolver=50:
                     theta, points, Trels (expm(A~/50), float32), As, Bs,
                     forward output [n_theta, ndim, nP] and the backward
                     Jacobian [d, n_theta, ndim, nP].
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from libcpab_amd.core.tessellation import make_tessellation  # noqa: E402

TESS = {
    "1d_10": ([10], True),
    "1d_50": ([50], True),
    "2d_3x3": ([3, 3], True),
    "2d_4x4": ([4, 4], True),
    "2d_2x2_nzb": ([2, 2], False),
    "3d_2x2x2": ([2, 2, 2], True),
    "3d_3x3x3": ([3, 3, 3], True),
    "3d_2x2x2_nzb": ([2, 2, 2], False),
}

STEP_CASES = {  # name: (tess, n_theta, nP, theta scale)
    "1d_10": ("1d_10", 2, 64, 0.8),
    "2d_3x3": ("2d_3x3", 2, 48, 0.5),
    "3d_2x2x2": ("3d_2x2x2", 2, 32, 0.3),
}


def main():
    if O.ref_lib() is None:
        raise SystemExit("oracle/_ref/libcpab_ref.so missing: run `make -C oracle` with /root/reference present")
    force = "--force" in sys.argv[1:]
    for name, (nc, zb) in TESS.items():
        if os.path.exists(os.path.join(HERE, "basis_%s.npz" % name)) and not force:
            continue
        t = make_tessellation(nc, zero_boundary=zb)
        planes, nbr = t.facet_table()
        np.savez_compressed(os.path.join(HERE, "basis_%s.npz" % name), nc=np.array(nc), zero_boundary=zb,
                            B=t.B, planes=planes, nbr=nbr, amap=t.amap)
    rng = np.random.default_rng(20240601)
    for name, (tname, n_theta, nP, scale) in STEP_CASES.items():
        if os.path.exists(os.path.join(HERE, "ref_step_%s.npz" % name)) and not force:
            continue
        nc, zb = TESS[tname]
        ndim = len(nc)
        f = np.load(os.path.join(HERE, "basis_%s.npz" % tname))
        B = f["B"]
        d = B.shape[1]
        nC = B.shape[0] // (ndim * (ndim + 1))
        theta = (rng.standard_normal((n_theta, d)) * scale).astype(np.float32)
        points = rng.uniform(0, 1, size=(ndim, nP)).astype(np.float32)
        A32 = (B.astype(np.float32) @ theta.T).T.reshape(n_theta, nC, ndim, ndim + 1).astype(np.float32)
        trels = O.trels(A32.astype(np.float64), 50).astype(np.float32)
        Bs = O.bases_Bs(B.astype(np.float32), nC, ndim).astype(np.float32)
        fwd = O.ref_forward(ndim, nc, 50, points, trels)
        bwd = O.ref_backward(ndim, nc, 50, points, A32, Bs)
        np.savez_compressed(os.path.join(HERE, "ref_step_%s.npz" % name), nc=np.array(nc), theta=theta,
                            points=points, trels=trels, As=A32, Bs=Bs, nstep=50, forward=fwd, backward=bwd)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
