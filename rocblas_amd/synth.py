"""Seeded synthetic inputs for the triangular-solve workloads.

Counter-based SplitMix64 (Steele, Lea & Flood, OOPSLA 2014): the i-th value of
stream ``key`` is ``mix(key + (i+1) * 0x9E3779B97F4A7C15)``; a uniform double in
[0, 1) takes its top 53 bits.  Because the stream is counter based, any element
can be generated independently, in any order, in vectorised numpy.

Storage is column-major (rocBLAS convention): a k x k matrix with leading
dimension ``ld`` is a Fortran-ordered numpy array of shape ``(ld, k)``; element
(i, j) draws counter ``i + j*k`` (independent of ``ld``), and instance ``b`` of a
batch draws counters offset by ``b*k*k``.

Distributions (one per workload):

* ``tri_uniform``: triangle with diagonal U[1, 2] and off-diagonal
  U[-1, 1] / k (well conditioned), optionally scaled as a whole.
* ``tri_normal``: diagonal U[1, 2] and off-diagonal N(0, 1) / sqrt(k)
  (Box-Muller on two uniforms), condition number ~1e3 at k = 16384.
* ``rhs_uniform``: U[-1, 1].
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
TAG_MULT = 0xD1B54A32D192ED03
MASK64 = (1 << 64) - 1

TAG_A = 1
TAG_B = 2
TAG_X = 3
TAG_SIZES = 4

_CHUNK = 1 << 22


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * M1
    z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def _mix_int(z):
    z &= MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def stream_key(seed, tag):
    """64-bit key of stream ``tag`` under ``seed``."""
    return _mix_int((int(seed) + int(tag) * TAG_MULT) & MASK64)


def uniform(key, offset, count, lo=0.0, hi=1.0):
    """float64 array: lo + (hi - lo) * u_{offset + i}, i < count."""
    out = np.empty(count, dtype=np.float64)
    k = np.uint64(key)
    scale = 1.0 / 9007199254740992.0
    with np.errstate(over="ignore"):
        for s in range(0, count, _CHUNK):
            e = min(count, s + _CHUNK)
            ctr = np.arange(offset + s + 1, offset + e + 1, dtype=np.uint64)
            z = _mix(k + ctr * GAMMA)
            u = (z >> np.uint64(11)).astype(np.float64) * scale
            out[s:e] = lo + (hi - lo) * u
    return out


def normal(key, offset, count):
    """Standard normals by Box-Muller: value i uses counters 2(offset+i), 2(offset+i)+1."""
    out = np.empty(count, dtype=np.float64)
    for s in range(0, count, _CHUNK):
        e = min(count, s + _CHUNK)
        u = uniform(key, 2 * (offset + s), 2 * (e - s))
        u1, u2 = u[0::2], u[1::2]
        out[s:e] = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    return out


def _tri_mask(k, uplo):
    i = np.arange(k)[:, None]
    j = np.arange(k)[None, :]
    return (i > j) if uplo == "L" else (i < j)


def tri_uniform(k, uplo, seed, ld=None, scale=1.0, poison=0.0, batch=None, dtype=np.float64):
    """Triangular k x k matrix (or a batch of them): diagonal U[1,2], off-diagonal
    U[-1,1]/k, everything times ``scale``; the unreferenced triangle holds
    ``poison``.  Returns (ld, k) Fortran array, or (batch, ld, k) with each
    instance Fortran-ordered (instance stride ld*k)."""
    ld = k if ld is None else ld
    nb = 1 if batch is None else batch
    key = stream_key(seed, TAG_A)
    u = uniform(key, 0, nb * k * k).reshape(nb, k, k)  # [b, j, i]
    u = np.swapaxes(u, 1, 2)  # [b, i, j]
    diag = 1.0 + u
    off = (2.0 * u - 1.0) / k
    mask = _tri_mask(k, uplo)[None]
    eye = np.eye(k, dtype=bool)[None]
    vals = np.where(eye, diag, np.where(mask, off, np.nan))
    vals *= scale
    vals = np.where(eye | mask, vals, poison)
    out = np.full((nb, k, ld), poison, dtype=dtype)  # [b, j, i] -> column-major per instance
    out[:, :, :k] = np.swapaxes(vals, 1, 2).astype(dtype)
    out = np.swapaxes(out, 1, 2)  # [b, i(ld), j]
    if batch is None:
        return np.asfortranarray(out[0])
    return out


def tri_normal(k, uplo, seed, ld=None, poison=0.0, dtype=np.float64):
    """Triangular k x k: diagonal U[1,2] (counter i + i*k of stream A), off-diagonal
    N(0,1)/sqrt(k) (Box-Muller pairs from stream A counters offset by k*k).  Built
    column block by column block to bound host memory."""
    ld = k if ld is None else ld
    key = stream_key(seed, TAG_A)
    A = np.full((ld, k), poison, dtype=dtype, order="F")
    inv = 1.0 / np.sqrt(k)
    cb = max(1, _CHUNK // max(k, 1))
    for j0 in range(0, k, cb):
        j1 = min(k, j0 + cb)
        z = normal(key, k * k + j0 * k, (j1 - j0) * k).reshape(j1 - j0, k).T  # [i, j]
        for jj in range(j0, j1):
            col = z[:, jj - j0] * inv
            if uplo == "U":
                A[:jj, jj] = col[:jj]
            else:
                A[jj + 1:k, jj] = col[jj + 1:k]
    d = uniform(key, 0, k * k)[:: k + 1] if k * k <= (1 << 26) else np.array(
        [uniform(key, i + i * k, 1)[0] for i in range(k)])
    A[np.arange(k), np.arange(k)] = 1.0 + d
    return A


def rhs_uniform(m, n, seed, ld=None, batch=None, dtype=np.float64, tag=TAG_B):
    """m x n right-hand side(s) with entries U[-1,1]; (ld, n) Fortran array, or
    (batch, ld, n) with instance stride ld*n."""
    ld = m if ld is None else ld
    nb = 1 if batch is None else batch
    key = stream_key(seed, tag)
    u = uniform(key, 0, nb * m * n, -1.0, 1.0).reshape(nb, n, m)
    out = np.zeros((nb, n, ld), dtype=dtype)
    out[:, :, :m] = u.astype(dtype)
    out = np.swapaxes(out, 1, 2)
    if batch is None:
        return np.asfortranarray(out[0])
    return out
