"""Writes the LAPACK golden vectors that pin the PLASMA-convention oracle.

PLASMA 2.x (the library behind GEQRT/TSQRT/LARFB/SSRFB at
Hybrid_tile_QR.c:985,1135,1276,1356) is not vendored, and the reference holds
no fixtures.  Its kernels are the ib-blocked LAPACK ones, so the LAPACK
routines dgeqrt / dtpqrt(l=0) / dgemqrt('L','T') / dtpmqrt('L','T', l=0), as
shipped with scipy.linalg.lapack, produce the same reflectors, T blocks and
sign convention to rounding.  This script records their inputs and outputs
as small .npz files (plain arrays, no pickle).

Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy
import scipy.linalg.lapack as L

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20151130)
    out = {}
    # dgeqrt: square, tall, wide, ib not dividing n
    for tag, (m, n, ib) in {"sq": (32, 32, 8), "tall": (40, 24, 5), "wide": (16, 24, 4),
                            "ib1": (12, 12, 1), "ibn": (16, 16, 16)}.items():
        A = rng.uniform(-1, 1, (m, n))
        a, t, info = L.dgeqrt(ib, A)
        assert info == 0
        out[f"geqrt_{tag}_A"] = A
        out[f"geqrt_{tag}_out"] = a
        out[f"geqrt_{tag}_T"] = t
        out[f"geqrt_{tag}_ib"] = np.array([ib])
    # dtpqrt (l=0): triangle on square / on rectangle
    for tag, (m, n, ib) in {"sq": (16, 16, 4), "rect": (20, 12, 5), "ib1": (8, 8, 1)}.items():
        A1 = np.triu(rng.uniform(-1, 1, (n, n)))
        B = rng.uniform(-1, 1, (m, n))
        a, b, t, info = L.dtpqrt(0, ib, A1, B)
        assert info == 0
        out[f"tpqrt_{tag}_A1"] = A1
        out[f"tpqrt_{tag}_B"] = B
        out[f"tpqrt_{tag}_A1out"] = a
        out[f"tpqrt_{tag}_Bout"] = b
        out[f"tpqrt_{tag}_T"] = t
        out[f"tpqrt_{tag}_ib"] = np.array([ib])
        # dtpmqrt('L','T') applying that block to a fresh pair
        C1 = rng.uniform(-1, 1, (n, 10))
        C2 = rng.uniform(-1, 1, (m, 10))
        c1, c2, info = L.dtpmqrt(0, b, t, C1, C2, side="L", trans="T")
        assert info == 0
        out[f"tpmqrt_{tag}_C1"] = C1
        out[f"tpmqrt_{tag}_C2"] = C2
        out[f"tpmqrt_{tag}_C1out"] = c1
        out[f"tpmqrt_{tag}_C2out"] = c2
    # dgemqrt('L','T')
    for tag, (m, ib) in {"sq": (32, 8), "ib3": (24, 3)}.items():
        A = rng.uniform(-1, 1, (m, m))
        v, t, info = L.dgeqrt(ib, A)
        Cm = rng.uniform(-1, 1, (m, 20))
        c, info2 = L.dgemqrt(v, t, Cm, side="L", trans="T")
        assert info == 0 and info2 == 0
        out[f"gemqrt_{tag}_V"] = v
        out[f"gemqrt_{tag}_T"] = t
        out[f"gemqrt_{tag}_C"] = Cm
        out[f"gemqrt_{tag}_Cout"] = c
        out[f"gemqrt_{tag}_ib"] = np.array([ib])
    np.savez_compressed(os.path.join(HERE, "lapack_tile_kernels.npz"), **out)
    with open(os.path.join(HERE, "lapack_tile_kernels.txt"), "w") as f:
        f.write(f"generated by tests/golden/make_golden.py with scipy {scipy.__version__}, "
                f"numpy {np.__version__}, seed 20151130\n")
        for k in sorted(out):
            f.write(f"{k} {out[k].shape}\n")


if __name__ == "__main__":
    main()
