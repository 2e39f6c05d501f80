"""Write golden RU/RD GEMM vectors from the reference's own vblas rmatmul /
udmatmul, built from the reference sources into oracle/_ref by
`make -C oracle ref`.  Inputs are seeded; the fixture holds inputs and
outputs only (data, no reference code).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle_ffi as O  # noqa: E402


def main():
    if O.ref_lib() is None:
        raise SystemExit("oracle/_ref not built (needs the reference tree): make -C oracle ref")
    rng = np.random.default_rng(0x5EED0100)
    out = {}
    cases = [(1, 1, 1), (3, 5, 7), (16, 16, 16), (33, 17, 65), (64, 48, 96)]
    for idx, (m, n, k) in enumerate(cases):
        A = np.asfortranarray(rng.uniform(-1, 1, (m, k)) * 2.0 ** rng.integers(-12, 12, (m, k)))
        B = np.asfortranarray(rng.uniform(-1, 1, (k, n)) * 2.0 ** rng.integers(-12, 12, (k, n)))
        out["A%d" % idx] = A
        out["B%d" % idx] = B
        out["CU%d" % idx] = O.ref_rmatmul(A, B, 1)
        out["CD%d" % idx] = O.ref_rmatmul(A, B, -1)
        out["CN%d" % idx] = O.ref_rmatmul(A, B, 0)
        C0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        CU = C0.copy(order="F")
        CD = C0.copy(order="F")
        if O.ref_udmatmul(A, B, CU, CD):
            out["C0%d" % idx] = C0
            out["UU%d" % idx] = CU
            out["UD%d" % idx] = CD
    # the reference's rounding-check patterns (vblas/Check_rmatmul_rounding.cpp)
    n = 24
    A = np.zeros((n, n), order="F")
    B = np.zeros((n, n), order="F")
    A[:, 1] = 1.0
    B[1, :] = 1.0
    A[:, 7] = 2.0 ** -27
    B[7, :] = 2.0 ** -27
    out["KAT_A"] = A
    out["KAT_B"] = B
    out["KAT_CU"] = O.ref_rmatmul(A, B, 1)
    out["KAT_CD"] = O.ref_rmatmul(A, B, -1)
    out["ncases"] = np.array(len(cases))
    path = os.path.join(HERE, "rmatmul_ref.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "(%d arrays)" % len(out))


if __name__ == "__main__":
    main()
