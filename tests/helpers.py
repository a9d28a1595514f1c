"""Shared test helpers: golden-fixture loading and problem construction."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

ALL_CASES = [(s, u, t) for s in "LR" for u in "UL" for t in "NT"]


def load_rational_cases():
    with open(os.path.join(GOLDEN, "trsm_rational.json")) as f:
        data = json.load(f)
    out = []
    for c in data["cases"]:
        k = c["m"] if c["side"] == "L" else c["n"]
        A = np.array([float.fromhex(v) for v in c["A"]]).reshape(k, c["lda"]).T
        B = np.array([float.fromhex(v) for v in c["B"]]).reshape(c["n"], c["ldb"]).T
        X = np.array([float.fromhex(v) for v in c["X"]]).reshape(c["n"], c["ldb"]).T
        out.append(dict(c, A=np.asfortranarray(A), B=np.asfortranarray(B),
                        X=np.asfortranarray(X), alpha=float.fromhex(c["alpha"])))
    return out


def op_dense(A, uplo, trans, diag, k):
    """op(A) as a dense k x k float64 array (triangle, unit diagonal honoured)."""
    T = np.array(A[:k, :k], dtype=np.float64)
    T = np.where(np.isnan(T), 0.0, T)
    T = np.triu(T) if uplo == "U" else np.tril(T)
    if diag == "U":
        np.fill_diagonal(T, 1.0)
    return T.T.copy() if trans != "N" else T


def integer_kat(side, uplo, trans, diag, m, n, rng, lda=None, ldb=None, alpha=1.0,
                poison=np.nan):
    """Exactly-solvable problem: integer X and integer off-diagonal A with a
    power-of-two diagonal, B = op(A) X / alpha (side L) or X op(A) / alpha
    (side R) formed in exact integer arithmetic.  Forward or backward
    substitution reproduces X bit-for-bit (every intermediate is a small
    integer or an exact dyadic).  Returns (A, B, X) Fortran arrays."""
    k = m if side == "L" else n
    lda = k if lda is None else lda
    ldb = m if ldb is None else ldb
    Ai = rng.integers(-2, 3, size=(k, k))
    d = 2 ** rng.integers(0, 3, size=k) * rng.choice([-1, 1], size=k)
    Xi = rng.integers(-4, 5, size=(m, n)) * (2 ** rng.integers(0, 3, size=(m, n)))
    T = np.triu(Ai, 1) if uplo == "U" else np.tril(Ai, -1)
    Td = T.astype(np.float64) + (np.diag(np.ones(k)) if diag == "U" else np.diag(d.astype(float)))
    # op(A) X is formed in float64: entries are small integers times dyadics -> exact
    opA = Td.T if trans != "N" else Td
    Xf = Xi.astype(np.float64) * np.float64(alpha)  # X is the solution of op(A) X = alpha B
    Bf = (opA @ Xi.astype(np.float64)) if side == "L" else (Xi.astype(np.float64) @ opA)
    A = np.full((lda, k), poison, order="F")
    mask = np.triu(np.ones((k, k), bool), 1) if uplo == "U" else np.tril(np.ones((k, k), bool), -1)
    A[:k, :k] = np.where(mask, T, poison)
    if diag == "N":
        A[np.arange(k), np.arange(k)] = d
    B = np.zeros((ldb, n), order="F")
    B[:m, :] = Bf
    X = np.zeros((ldb, n), order="F")
    X[:m, :] = Xi.astype(np.float64)
    if alpha != 1.0:
        # op(A) X = alpha B  <=> B = op(A) X / alpha ; keep alpha a power of two
        B[:m, :] = Bf / alpha
    del Xf
    return A, B, X
