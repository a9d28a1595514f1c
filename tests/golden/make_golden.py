"""Generate exact-rational golden vectors for the triangular solve.

The reference ships no golden vectors or known-answer files for trsm (its tests
are random and tolerance-based: clients/include/blas3/testing_trsm.hpp:297-315,
433-448), so these fixtures are computed here from the definition of the
operation (library/include/internal/rocblas-functions.h:19302-19386):

    side L:  op(A) X = alpha B        side R:  X op(A) = alpha B

with op(A) in {A, A^T, A^H=A^T for reals}, only the ``uplo`` triangle of A
referenced, a unit diagonal assumed (and never read) for diag=U, and
alpha == 0 giving X = 0.

Every input is a dyadic rational (exact in binary64), the solution is computed
in exact rational arithmetic with ``fractions.Fraction`` by brute-force
substitution, and each entry of X is stored correctly rounded to binary64
(Python's int/int true division rounds correctly).  Entries of A outside the
referenced triangle, and the diagonal when diag=U, hold NaN so that a solver
reading them fails.  Values are stored as ``float.hex`` strings.

Run:  python tests/golden/make_golden.py   (writes trsm_rational.json beside it)
"""
import itertools
import json
import os
import random
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))


def dyadic(rng, lo, hi, bits=3):
    """random dyadic rational in [lo, hi] with ``bits`` fractional bits"""
    q = 1 << bits
    return Fraction(rng.randint(int(lo * q), int(hi * q)), q)


def op_entry(A, uplo, trans, diag, i, j):
    """op(A)[i][j] as the solve sees it (A is a dict (r,c) -> Fraction)."""
    r, c = (j, i) if trans != "N" else (i, j)
    if r == c:
        return Fraction(1) if diag == "U" else A[(r, c)]
    if (uplo == "U" and r < c) or (uplo == "L" and r > c):
        return A[(r, c)]
    return Fraction(0)


def solve_exact(side, uplo, trans, diag, m, n, alpha, A, B):
    """Exact X with op(A) X = alpha B (L) or X op(A) = alpha B (R)."""
    k = m if side == "L" else n
    M = [[op_entry(A, uplo, trans, diag, i, j) for j in range(k)] for i in range(k)]
    if side == "R":
        # X op(A) = alpha B  <=>  op(A)^T X^T = alpha B^T
        M = [[M[j][i] for j in range(k)] for i in range(k)]
        rhs = [[alpha * B[(i, j)] for i in range(m)] for j in range(n)]  # rows of B^T
        cols = n
        vec_len = m
    else:
        rhs = [[alpha * B[(i, j)] for j in range(n)] for i in range(m)]
        cols = m
        vec_len = n
    lower = all(M[i][j] == 0 for i in range(k) for j in range(i + 1, k))
    order = range(k) if lower else range(k - 1, -1, -1)
    Y = [[Fraction(0)] * vec_len for _ in range(cols)]
    for i in order:
        for v in range(vec_len):
            s = rhs[i][v]
            for p in range(k):
                if p != i and M[i][p] != 0:
                    s -= M[i][p] * Y[p][v]
            Y[i][v] = s / M[i][i]
    if side == "R":
        return {(i, j): Y[j][i] for i in range(m) for j in range(n)}
    return {(i, j): Y[i][j] for i in range(m) for j in range(n)}


def f2hex(x):
    return float(x).hex()


def make_case(rng, side, uplo, trans, diag, m, n, alpha, lda_pad, ldb_pad):
    k = m if side == "L" else n
    lda, ldb = k + lda_pad, m + ldb_pad
    A = {}
    for r in range(k):
        for c in range(k):
            if r == c:
                # well away from zero so the exact solution stays modest
                A[(r, c)] = dyadic(rng, 1, 2) * rng.choice([-1, 1])
            elif (uplo == "U" and r < c) or (uplo == "L" and r > c):
                A[(r, c)] = dyadic(rng, -1, 1)
    B = {(i, j): dyadic(rng, -1, 1) for i in range(m) for j in range(n)}
    if alpha == 0:
        X = {(i, j): Fraction(0) for i in range(m) for j in range(n)}
    else:
        X = solve_exact(side, uplo, trans, diag, m, n, Fraction(alpha), A, B)
    nan = float("nan").hex()
    a_flat = []
    for c in range(k):
        for r in range(lda):
            ref = r < k and (r != c or diag == "N") and (
                r == c or (uplo == "U" and r < c) or (uplo == "L" and r > c))
            a_flat.append(f2hex(A[(r, c)]) if ref else nan)
    b_flat, x_flat = [], []
    for c in range(n):
        for r in range(ldb):
            b_flat.append(f2hex(B[(r, c)]) if r < m else nan)
            x_flat.append(f2hex(X[(r, c)]) if r < m else nan)
    return {
        "side": side, "uplo": uplo, "trans": trans, "diag": diag,
        "m": m, "n": n, "lda": lda, "ldb": ldb, "alpha": float(alpha).hex(),
        "A": a_flat, "B": b_flat, "X": x_flat,
    }


def main():
    rng = random.Random(20260101)
    cases = []
    shapes = [(1, 1), (1, 3), (3, 1), (2, 5), (5, 2), (4, 4), (6, 3), (3, 6)]
    for side, uplo, trans, diag in itertools.product("LR", "UL", "NTC", "NU"):
        for (m, n) in shapes:
            alpha = rng.choice([1.0, -0.5, 3.0, 0.75])
            cases.append(make_case(rng, side, uplo, trans, diag, m, n, alpha,
                                   rng.choice([0, 1, 3]), rng.choice([0, 2])))
        # alpha == 0: B := 0, A unreferenced
        cases.append(make_case(rng, side, uplo, trans, diag, 3, 2, 0.0, 1, 1))
    with open(os.path.join(HERE, "trsm_rational.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "cases": cases}, f,
                  separators=(",", ":"))
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
