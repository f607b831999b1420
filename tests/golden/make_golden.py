"""Writes the golden fixtures of the Cholesky path (run in a container with
scipy; the fixtures are committed, this script is kept beside them).

Inputs come from the Chameleon generators (dplgsy/dplrnt, seeds as in
example/lapack_to_chameleon/step1.c:101,104 and the sizes of
testing/input/{potrf,potrs,posv}.in); expected outputs come from LAPACK
(scipy.linalg.lapack dpotrf/dpotrs, OpenBLAS 0.3.29) and BLAS
(scipy.linalg.blas dtrsm/dsyrk/dgemm) -- the libraries Chameleon's
coreblas calls (coreblas/compute/core_z{potrf,trsm,syrk,gemm}.c).

The LCG vectors are computed here by a pure-Python restatement of
coreblas/include/coreblas/random.h (sequential stepping, no jump-ahead), so
they pin the C oracle's jump-ahead independently.
"""
import os
import sys

import numpy as np
import scipy.linalg.blas as sb
import scipy.linalg.lapack as sl

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

MASK = (1 << 64) - 1
A_LCG, C_LCG = 6364136223846793005, 1


def lcg_sequence(seed, count):
    x, out = seed, []
    for _ in range(count):
        out.append(x)
        x = (A_LCG * x + C_LCG) & MASK
    return out


def laran(x):
    return 0.5 - float(x) * 5.4210108624275222e-20


def python_plgsy(bump, n, seed):
    """Full symmetric matrix straight from the element rule: (r,c), r>=c,
    uses stream position r + c*n; diagonal + bump."""
    states = lcg_sequence(seed, n * n)
    A = np.empty((n, n))
    for c in range(n):
        for r in range(c, n):
            v = laran(states[r + c * n])
            A[r, c] = v
            A[c, r] = v
        A[c, c] += bump
    return A


def python_plrnt(m, n, seed):
    states = lcg_sequence(seed, m * n)
    return np.array([[laran(states[r + c * m]) for c in range(n)] for r in range(m)])


def main():
    out = {}
    # --- RNG: states and values at a few positions (pure-Python stepping)
    seeds = [0, 1, 51, 5673, 2**63 + 12345]
    pos = [0, 1, 2, 7, 100, 1000, 4095]
    for s in seeds:
        seq = lcg_sequence(s, max(pos) + 1)
        out[f"lcg_seed{s}_pos"] = np.array(pos, dtype=np.uint64)
        out[f"lcg_seed{s}_state"] = np.array([seq[p] for p in pos], dtype=np.uint64)
    # --- generators from the element rule
    for n, bump, seed in [(15, 15.0, 51), (37, 37.0, 51), (21, 0.0, 3)]:
        out[f"plgsy_n{n}_b{int(bump)}_s{seed}"] = python_plgsy(bump, n, seed)
    for m, n, seed in [(37, 13, 5673), (16, 22, 7)]:
        out[f"plrnt_m{m}_n{n}_s{seed}"] = python_plrnt(m, n, seed)
    # --- factor / solve cases (testing/input/posv.in sizes + a few larger)
    cases = [(15, 1), (19, 13), (21, 22), (35, 33), (37, 13), (64, 16), (100, 7)]
    for n, nrhs in cases:
        A = python_plgsy(float(n), n, 51)
        B = python_plrnt(n, nrhs, 5673)
        for uplo, ch in ((oracle.LOWER, "L"), (oracle.UPPER, "U")):
            F, info = sl.dpotrf(A, lower=(ch == "L"), clean=True)
            assert info == 0
            X, info = sl.dpotrs(F, B, lower=(ch == "L"))
            assert info == 0
            out[f"chol_n{n}_{ch}"] = F
            out[f"sol_n{n}_r{nrhs}_{ch}"] = X
    # --- a non-SPD input: dplgsy with bump 0 (LAPACK reports the first
    #     leading minor that is not positive definite)
    A = python_plgsy(0.0, 21, 3)
    _, info = sl.dpotrf(A, lower=True)
    out["nonspd_n21_info"] = np.array([info])
    A = python_plgsy(2.0, 40, 9)
    _, info = sl.dpotrf(A, lower=True)
    out["nonspd_n40_b2_info"] = np.array([info])
    # --- tile kernels against BLAS
    rng = np.random.default_rng(20240218)
    T = np.tril(rng.uniform(-0.5, 0.5, (17, 17))) + 4 * np.eye(17)
    Bt = rng.uniform(-0.5, 0.5, (13, 17))
    out["trsm_T"] = T
    out["trsm_B"] = Bt
    out["trsm_RLTN"] = sb.dtrsm(1.0, T, Bt, side=1, lower=1, trans_a=1, diag=0)
    Bl = rng.uniform(-0.5, 0.5, (17, 5))
    out["trsm_Bl"] = Bl
    out["trsm_LLNN"] = sb.dtrsm(1.0, T, Bl, side=0, lower=1, trans_a=0, diag=0)
    out["trsm_LLTN"] = sb.dtrsm(1.0, T, Bl, side=0, lower=1, trans_a=1, diag=0)
    Ak = rng.uniform(-0.5, 0.5, (17, 11))
    C = rng.uniform(-0.5, 0.5, (17, 17))
    out["syrk_A"], out["syrk_C"] = Ak, C
    out["syrk_LN"] = sb.dsyrk(-1.0, Ak, beta=1.0, c=C, trans=0, lower=1)
    Ag = rng.uniform(-0.5, 0.5, (13, 11))
    Bg = rng.uniform(-0.5, 0.5, (9, 11))
    Cg = rng.uniform(-0.5, 0.5, (13, 9))
    out["gemm_A"], out["gemm_B"], out["gemm_C"] = Ag, Bg, Cg
    out["gemm_NT"] = sb.dgemm(-1.0, Ag, Bg, beta=1.0, c=Cg, trans_a=0, trans_b=1)
    np.savez_compressed(os.path.join(HERE, "cholesky_golden.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
