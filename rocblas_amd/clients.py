"""Client side of the trsm path: what the reference's rocblas-bench and
rocblas-test (gtest) do for trsm, rebuilt over this library.

Reference:
  clients/benchmarks/client.cpp:1430-1704      rocblas-bench flags
  clients/include/blas3/testing_trsm.hpp       (and _batched, _strided_batched;
  clients/include/blas_ex/testing_trsm_ex.hpp)  the per-case procedure
  clients/include/argument_model.hpp:56-247    the printed result line
  clients/include/rocblas_common.yaml           Defaults, list-expansion rules
  clients/common/rocblas_gentest.py            YAML case expansion
  clients/gtest/trsm_gtest.yaml                the trsm case matrix

One case follows the reference's known-solution procedure
(testing_trsm.hpp:283-315, 433-448):
  A  a diagonally dominant triangle (rocblas_init.hpp:263-327); for a unit
     diagonal each row is divided by its diagonal entry, which is then
     overwritten with a random value the solve must not read
     (client_utility.hpp:419-444);
  X  a general matrix with alternating signs (rocblas_init.hpp:63-84);
  B  = op(A) X / alpha (or X op(A) / alpha), host BLAS trmm;
the library solves for X' in place of B, and the case passes when the forward
error |X - X'|_1 / |X|_1 and the residual |op(A) X' / alpha - B|_1 / |...|_1
are both at most 40 eps M (testing_trsm.hpp:30-31; unit.hpp:578-583).  Those
sums are the reference's matrix_norm_1 (norm.hpp:449-492), whose running
column maxima of cumulative sums reduce to whole-matrix sums.  With alpha = 0
the result must be exactly zero.  The host BLAS is scipy's; the reference
links a host CBLAS for the same role.

Device memory comes from torch (plumbing); every solve is a call into
librocblas_amd.so through its C ABI, or through the C++ templates of
include/rocblas_amd_internal.hpp for api INTERNAL.
"""
import ctypes
import os
import re
import subprocess
import time
from dataclasses import dataclass, fields

import numpy as np

from . import LIB_PATH, lib, rocblas_status

S = rocblas_status
SIDE = {"L": 141, "R": 142}
FILL = {"U": 121, "L": 122}
OP = {"N": 111, "T": 112, "C": 113}
DIAG = {"N": 131, "U": 132}
DTYPE = {"f32_r": 151, "f64_r": 152}  # rocblas_datatype (rocblas-types.h)
BOTH = 143  # rocblas_side_both, used as an invalid value of every enum

# functions this client runs; the reference's others are not on the path
SOLVE_FUNCS = ("trsm", "trsm_batched", "trsm_strided_batched", "trsm_ex", "trsm_batched_ex",
               "trsm_strided_batched_ex")
BAD_ARG_FUNCS = tuple(f + "_bad_arg" for f in SOLVE_FUNCS)
INTERNAL_FUNCS = ("trsm_internal_interfaces", "trsm_batched_internal_interfaces")

PRECISIONS = {"s": "f32_r", "d": "f64_r", "f32_r": "f32_r", "f64_r": "f64_r",
              "single": "f32_r", "double": "f64_r"}
# api values (clients/include/rocblas_arguments.hpp:39-51).  This library has
# no Fortran module, so the FORTRAN forms call the same C entry points.
API = {"C": 0, "C_64": 1, "FORTRAN": 2, "FORTRAN_64": 3, "INTERNAL": 4, "INTERNAL_64": 5}


@dataclass
class Arguments:
    """The fields of the reference's Arguments a trsm case uses
    (clients/include/rocblas_arguments.hpp), with rocblas_common.yaml's
    defaults."""
    function: str = "trsm"
    precision: str = "f64_r"
    side: str = "L"
    uplo: str = "L"
    transA: str = "N"
    diag: str = "N"
    M: int = 128
    N: int = 128
    lda: int = 128
    ldb: int = 128
    ldc: int = 0  # INTERNAL api: element offset of A
    ldd: int = 0  # INTERNAL api: element offset of B
    stride_a: int = 0
    stride_b: int = 0
    stride_scale: float = 1.0
    batch_count: int = 1
    alpha: float = 1.0
    api: int = 0
    pointer_mode_host: bool = True
    pointer_mode_device: bool = True
    user_allocated_workspace: int = 0
    initialization: str = "rand_int"
    unit_check: int = 1
    norm_check: int = 0
    timing: int = 0
    iters: int = 10
    cold_iters: int = 2
    flush_batch_count: int = 1
    flush_memory_size: int = 0
    graph_test: bool = False
    repeatability_check: bool = False
    name: str = "rocblas-test"
    category: str = "nightly"
    seed: int = 0

    @property
    def K(self):
        return self.M if self.side.upper() == "L" else self.N

    @property
    def dtype(self):
        return np.float32 if self.precision == "f32_r" else np.float64

    @property
    def eps(self):
        return float(np.finfo(self.dtype).eps)

    def label(self):
        parts = [self.name, self.function, self.precision, self.side, self.uplo, self.transA,
                 self.diag, self.M, self.N, self.lda, self.ldb, f"{self.alpha:g}",
                 self.batch_count, self.stride_a, self.stride_b,
                 next(k for k, v in API.items() if v == self.api), self.user_allocated_workspace]
        return "_".join(str(p) for p in parts)

    @classmethod
    def from_case(cls, case):
        """An Arguments from one expanded YAML case (a dict of the reference's keys)."""
        a = cls()
        names = {f.name for f in fields(cls)}
        for key, val in case.items():
            if key in names and key not in ("precision", "api"):
                setattr(a, key, type(getattr(a, key))(val) if val is not None else val)
        prec = case.get("compute_type") or case.get("a_type") or case.get("precision")
        a.precision = PRECISIONS.get(prec, prec) if isinstance(prec, str) else "f64_r"
        api = case.get("api", "C")
        a.api = API[api] if isinstance(api, str) else int(api)
        for key in ("side", "uplo", "transA", "diag"):
            v = str(case.get(key, getattr(a, key)))
            setattr(a, key, v)
        return a


# --------------------------------------------------------------------------
# YAML case matrix (clients/common/rocblas_gentest.py's rules, restated)
# --------------------------------------------------------------------------
_INCLUDE = re.compile(r"include\s*:\s*([-.\w/]+)")
_RANGE = re.compile(r"^\s*(-?\d+)\s*\.\.\s*(-?\d+)\s*(?:\.\.\s*(-?\d+))?\s*$")


def read_yaml_source(path, include_dirs=()):
    """The text of a gtest YAML file with its ``include:`` lines replaced by
    the named files (searched next to the file, then in include_dirs), so
    anchors defined in an included file resolve."""
    out = []
    with open(path) as f:
        for line in f:
            m = line.startswith("include") and _INCLUDE.match(line)
            if not m:
                out.append(line)
                continue
            for d in [os.path.dirname(path) or "."] + list(include_dirs):
                p = os.path.join(d, m.group(1))
                if os.path.exists(p):
                    out.append(read_yaml_source(p, include_dirs))
                    break
            else:
                raise FileNotFoundError(f"{path}: cannot find included {m.group(1)}")
    return "".join(out)


def load_yaml_doc(path, include_dirs=()):
    import yaml
    for doc in yaml.safe_load_all(read_yaml_source(path, include_dirs)):
        if doc and doc.get("Tests"):
            return doc
    raise ValueError(f"{path}: no Tests")


def _generate(case, dict_lists, no_expand, out):
    """Expand one test entry into cases: the named dictionary lists first
    (each item merged into the case), then integer ranges and lists of the
    remaining keys in sorted order, as the reference's generator does."""
    for name in dict_lists:
        if isinstance(name, dict):
            ((key, target),) = name.items()
            if isinstance(case.get(key), dict):
                for k, v in sorted(case[key].items(), key=lambda kv: kv[0]):
                    c = dict(case)
                    c[key], c[target] = k, v
                    _generate(c, dict_lists, no_expand, out)
                return
        elif name in case and isinstance(case[name], (list, tuple, dict)):
            items = case[name]
            rest = {k: v for k, v in case.items() if k != name}
            for item in [items] if isinstance(items, dict) else items:
                c = dict(rest)
                c.update(item)
                _generate(c, dict_lists, no_expand, out)
            return
    for key in sorted(case):
        val = case[key]
        if isinstance(val, str):
            m = _RANGE.match(val)
            if m:
                for v in range(int(m.group(1)), int(m.group(2)) + 1, int(m.group(3) or 1)):
                    c = dict(case)
                    c[key] = v
                    _generate(c, dict_lists, no_expand, out)
                return
        elif isinstance(val, (list, tuple)) and key not in no_expand:
            for v in val:
                c = dict(case)
                c[key] = v
                _generate(c, dict_lists, no_expand, out)
            return
    out.append(case)


def _set_strides(case):
    """Strides from stride_scale for the strided-batched trsm forms
    (rocblas_gentest.py setdefaults)."""
    if case.get("function") in ("trsm_strided_batched", "trsm_strided_batched_ex") \
            and "stride_scale" in case:
        sc = case["stride_scale"]
        case["stride_b"] = int(abs(case["N"] * case["ldb"] * sc))
        first = case["M"] if str(case["side"]).upper() == "L" else case["N"]
        case["stride_a"] = int(abs(first * case["lda"] * sc))
    return case


def expand_doc(doc):
    """Every case of a parsed gtest YAML document, with its Defaults."""
    defaults = doc.get("Defaults") or {}
    dict_lists = doc.get("Dictionary lists to expand") or ()
    no_expand = doc.get("Lists to not expand") or ()
    out = []
    for test in doc["Tests"]:
        case = dict(defaults)
        case.update(test)
        cases = []
        _generate(case, dict_lists, no_expand, cases)
        out += [_set_strides(c) for c in cases]
    return out


def case_kind(case):
    """'run', or the reason this library does not run the case."""
    fn = case.get("function")
    if fn not in SOLVE_FUNCS + BAD_ARG_FUNCS + INTERNAL_FUNCS:
        return f"not on the trsm path: {fn}"
    prec = case.get("compute_type") or case.get("a_type")
    if isinstance(prec, str) and prec not in ("f32_r", "f64_r"):
        return f"precision {prec} not built (real single and double only)"
    if case.get("HMM"):
        return "HipManagedMemory cases not run"
    if max(int(case.get("lda", 0)), int(case.get("ldb", 0))) > 0x7FFFFFFF:
        # the _64 entry points solve these (tests/test_ilp64_huge_ld_gpu.py runs
        # the same cases); this client's dense host copies of a 2^31-row
        # leading dimension would take several times 26 GB of host memory
        return "leading dimension beyond INT32_MAX: run by the ILP64 GPU tests, not this client"
    if case.get("threads", 0) or case.get("streams", 0) or case.get("devices", 0):
        return "multi-thread / multi-stream / multi-device cases not run"
    return "run"


# --------------------------------------------------------------------------
# data (the reference's generators, restated)
# --------------------------------------------------------------------------
def _rand(rng, init, shape, dtype, rows=None, cols=None, ld=0):
    """Values of the reference's initialisations: rand_int 1..10, hpl
    U(-0.5, 0.5), trig_float sin of the element's linear index (rows / cols
    give the block's absolute indices for that)."""
    if init == "hpl":
        return rng.uniform(-0.5, 0.5, shape).astype(dtype)
    if init == "trig_float":
        r = np.arange(shape[0]) if rows is None else rows
        c = np.arange(shape[1]) if cols is None else cols
        return np.sin(r[:, None] + c[None, :] * max(ld, shape[0])).astype(dtype)
    return rng.integers(1, 11, shape, dtype=np.int8).astype(dtype)  # rand_int: 1..10


def init_triangle(arg, rng):
    """lda x K column-major diagonally dominant triangle (rocblas_init.hpp:263-327),
    generated in blocks of columns so large K needs little more than A itself."""
    K, lda, dt = arg.K, arg.lda, arg.dtype
    upper = arg.uplo.upper() == "U"
    A = np.zeros((lda, K), dtype=dt, order="F")
    rowsum = np.zeros(K)
    colsum = np.zeros(K)
    step = max(1, (1 << 23) // max(K, 1))
    rows = np.arange(K)
    for j0 in range(0, K, step):
        j1 = min(K, j0 + step)
        cols = np.arange(j0, j1)
        keep = (cols[None, :] >= rows[:, None]) if upper else (cols[None, :] <= rows[:, None])
        v = np.where(keep, _rand(rng, arg.initialization, (K, j1 - j0), dt, rows, cols, lda), 0)
        A[:K, j0:j1] = v
        off = np.abs(np.where(cols[None, :] == rows[:, None], 0, v), dtype=np.float64)
        colsum[j0:j1] += off.sum(axis=0)
        rowsum += off.sum(axis=1)
    # row sum + column sum of the off-diagonal entries, times 1.01
    s = rowsum + colsum
    d = np.where(s == 0, 1.0, s * 1.01).astype(dt)
    A[rows, rows] = d
    if arg.diag.upper() == "U":
        # client_utility.hpp:419-444: scale rows (lower) or columns (upper)
        # by the diagonal, then overwrite the diagonal with random values
        for j0 in range(0, K, step):
            j1 = min(K, j0 + step)
            if upper:
                A[:K, j0:j1] /= d[None, j0:j1]
            else:
                A[:K, j0:j1] /= d[:, None]
        A[rows, rows] = _rand(rng, arg.initialization, (1, K), dt)[0]
    return A


def init_general(arg, rng, m, n, ld):
    """ld x n column-major matrix, alternating signs (rocblas_init.hpp:63-84)."""
    X = np.zeros((ld, n), dtype=arg.dtype, order="F")
    v = _rand(rng, arg.initialization, (m, n), arg.dtype, ld=ld)
    i, j = np.indices((m, n))
    X[:m, :n] = np.where((i ^ j) & 1, v, -v)
    return X


def _blas(arg, name):
    from scipy.linalg import blas
    return getattr(blas, ("s" if arg.precision == "f32_r" else "d") + name)


def host_trmm(arg, A, X, alpha):
    """alpha op(A) X or alpha X op(A) on the host (the reference's ref_trmm)."""
    K, m, n = arg.K, arg.M, arg.N
    fn = _blas(arg, "trmm")
    return fn(arg.dtype(alpha), np.asfortranarray(A[:K, :K]), np.asfortranarray(X[:m, :n]),
              side=int(arg.side.upper() == "R"), lower=int(arg.uplo.upper() == "L"),
              trans_a=int(arg.transA.upper() != "N"), diag=int(arg.diag.upper() == "U"))


def host_trsm(arg, A, B, alpha):
    K, m, n = arg.K, arg.M, arg.N
    fn = _blas(arg, "trsm")
    return fn(arg.dtype(alpha), np.asfortranarray(A[:K, :K]), np.asfortranarray(B[:m, :n]),
              side=int(arg.side.upper() == "R"), lower=int(arg.uplo.upper() == "L"),
              trans_a=int(arg.transA.upper() != "N"), diag=int(arg.diag.upper() == "U"))


def norm1_rel(gold, got):
    """sum |gold - got| / sum |gold| in double (norm.hpp:449-492)."""
    g = np.asarray(gold, dtype=np.float64)
    d = np.abs(g - np.asarray(got, dtype=np.float64)).sum()
    s = np.abs(g).sum()
    return float(d / s) if s else float(d)


# --------------------------------------------------------------------------
# calls into the library
# --------------------------------------------------------------------------
_vp, _i32, _i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64


def _scalar(arg, value):
    return ctypes.c_float(value) if arg.precision == "f32_r" else ctypes.c_double(value)


def _internal_symbol(prefix):
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    dem = subprocess.run(["nm", "-DC", "--defined-only", LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for raw, d in zip(out.splitlines(), dem.splitlines()):
        name = d.split(" ", 2)[2]
        if name.startswith(prefix) or name.startswith("rocblas_status_ " + prefix):
            return getattr(lib(), raw.split()[-1])
    raise KeyError(prefix)


_INTERNAL = {}


def internal_fn(prefix, argtypes, restype=ctypes.c_int32):
    if prefix not in _INTERNAL:
        f = _internal_symbol(prefix)
        f.restype, f.argtypes = restype, argtypes
        _INTERNAL[prefix] = f
    return _INTERNAL[prefix]


_TRSM_INTERNAL_ARGS = [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i64, _i32, _i64, _vp,
                       _i64, _i32, _i64, _i32, ctypes.c_bool, _vp, _vp, _vp, _vp, _vp, _i32, _i64,
                       _i64]


class Call:
    """One invocation of the case's entry point on given device operands.
    ``A`` and ``B`` are device addresses: the matrices (plain), the first
    instance (strided) or device arrays of instance addresses (batched)."""

    def __init__(self, arg):
        self.arg = arg
        p = "s" if arg.precision == "f32_r" else "d"
        fn = arg.function
        self.form = ("batched" if "batched" in fn and "strided" not in fn
                     else "strided" if "strided" in fn else "plain")
        self.ex = fn.endswith("_ex")
        self.api64 = arg.api in (1, 3)
        self.internal = arg.api in (4, 5)
        L = lib()
        if self.internal:
            T = "float" if p == "s" else "double"
            name = ("rocblas_internal_trsm_batched_template" if self.form == "batched"
                    else "rocblas_internal_trsm_template")
            self.fn = internal_fn(f"{name}<{T}>", _TRSM_INTERNAL_ARGS)
        elif self.ex:
            self.fn = getattr(L, {"plain": "rocblas_trsm_ex", "batched": "rocblas_trsm_batched_ex",
                                  "strided": "rocblas_trsm_strided_batched_ex"}[self.form])
        else:
            suffix = {"plain": "", "batched": "_batched", "strided": "_strided_batched"}[self.form]
            self.fn = getattr(L, f"rocblas_{p}trsm{suffix}" + ("_64" if self.api64 else ""))

    def __call__(self, handle, alpha_ptr, A, lda, B, ldb, M=None, N=None, batch=None,
                 stride_a=None, stride_b=None, side=None, uplo=None, trans=None, diag=None,
                 invA=None, invA_size=0):
        a = self.arg
        M = a.M if M is None else M
        N = a.N if N is None else N
        batch = a.batch_count if batch is None else batch
        sa = a.stride_a if stride_a is None else stride_a
        sb = a.stride_b if stride_b is None else stride_b
        e = (SIDE.get(a.side.upper(), BOTH) if side is None else side,
             FILL.get(a.uplo.upper(), BOTH) if uplo is None else uplo,
             OP.get(a.transA.upper(), BOTH) if trans is None else trans,
             DIAG.get(a.diag.upper(), BOTH) if diag is None else diag)
        if self.internal:
            oa, ob = a.ldc, a.ldd
            return self.fn(handle, *e, M, N, alpha_ptr, A, -oa, lda,
                           sa if self.form == "strided" else 0, B, -ob, ldb,
                           sb if self.form == "strided" else 0,
                           1 if self.form == "plain" else batch, True, None, None, None, None,
                           None, 0, 0, 0)
        if self.ex:
            dt = DTYPE[a.precision]
            if self.form == "plain":
                return self.fn(handle, *e, M, N, alpha_ptr, A, lda, B, ldb, invA, invA_size, dt)
            if self.form == "batched":
                return self.fn(handle, *e, M, N, alpha_ptr, A, lda, B, ldb, batch, invA,
                               invA_size, dt)
            return self.fn(handle, *e, M, N, alpha_ptr, A, lda, sa, B, ldb, sb, batch, invA,
                           invA_size, 0, dt)
        if self.form == "plain":
            return self.fn(handle, *e, M, N, alpha_ptr, A, lda, B, ldb)
        if self.form == "batched":
            return self.fn(handle, *e, M, N, alpha_ptr, A, lda, B, ldb, batch)
        return self.fn(handle, *e, M, N, alpha_ptr, A, lda, sa, B, ldb, sb, batch)


class Operands:
    """Host data of a case and its device copies (instance b of A at
    b * stride_a, of B at b * stride_b; pointer arrays for the batched forms;
    INTERNAL api: addresses shifted forward by the element offsets ldc / ldd,
    which the call then passes negated, as testing_trsm.hpp:381-410 does)."""

    def __init__(self, arg, rng):
        import torch
        self.arg = arg
        K, M, N, lda, ldb = arg.K, arg.M, arg.N, arg.lda, arg.ldb
        batch = max(arg.batch_count, 1) if arg.function not in ("trsm", "trsm_ex") else 1
        self.batch = batch
        self.sa = max(arg.stride_a, lda * K) if "strided" in arg.function else lda * K
        self.sb = max(arg.stride_b, ldb * N) if "strided" in arg.function else ldb * N
        self.A = [init_triangle(arg, rng) for _ in range(batch)]
        self.X = [init_general(arg, rng, M, N, ldb) for _ in range(batch)]
        alpha = arg.alpha
        self.B = []
        for b in range(batch):
            Bb = np.zeros((ldb, N), dtype=arg.dtype, order="F")
            if alpha != 0:
                Bb[:M, :N] = host_trmm(arg, self.A[b], self.X[b], 1.0 / alpha)
            else:
                Bb[:M, :N] = self.X[b][:M, :N]
            self.B.append(Bb)
        tdt = torch.float32 if arg.precision == "f32_r" else torch.float64
        ha = np.zeros(self.sa * batch, dtype=arg.dtype)
        hb = np.zeros(self.sb * batch, dtype=arg.dtype)
        for b in range(batch):
            ha[b * self.sa:b * self.sa + lda * K] = self.A[b].ravel(order="F")
            hb[b * self.sb:b * self.sb + ldb * N] = self.B[b].ravel(order="F")
        self.hb = hb
        self.dA = torch.from_numpy(ha).to("cuda")
        self.dB = torch.from_numpy(hb.copy()).to("cuda")
        self.elt = self.dA.element_size()
        self.pa = self.pb = None
        if "batched" in arg.function and "strided" not in arg.function:
            shift_a = arg.ldc if arg.api in (4, 5) else 0
            shift_b = arg.ldd if arg.api in (4, 5) else 0
            self.pa = torch.tensor([self.dA.data_ptr() + (b * self.sa + shift_a) * self.elt
                                    for b in range(batch)], dtype=torch.int64, device="cuda")
            self.pb = torch.tensor([self.dB.data_ptr() + (b * self.sb + shift_b) * self.elt
                                    for b in range(batch)], dtype=torch.int64, device="cuda")

    def reset(self):
        import torch
        self.dB.copy_(torch.from_numpy(self.hb))

    def pointers(self):
        a = self.arg
        if self.pa is not None:
            return self.pa.data_ptr(), self.pb.data_ptr()
        pA, pB = self.dA.data_ptr(), self.dB.data_ptr()
        if a.api in (4, 5):
            pA += a.ldc * self.elt
            pB += a.ldd * self.elt
        return pA, pB

    def result(self):
        """X' per instance, host, ldb x N."""
        h = self.dB.cpu().numpy()
        return [h[b * self.sb:b * self.sb + self.arg.ldb * self.arg.N]
                .reshape(self.arg.N, self.arg.ldb).T for b in range(self.batch)]


# --------------------------------------------------------------------------
# testers
# --------------------------------------------------------------------------
class CaseFailure(AssertionError):
    pass


def _expect(got, want, what):
    if int(got) != int(want):
        raise CaseFailure(f"{what}: got {S(int(got)).name}, want {S(int(want)).name}")


def _invalid_size(arg):
    K = arg.K
    bad = arg.M < 0 or arg.N < 0 or arg.lda < K or arg.ldb < arg.M
    if arg.function not in ("trsm", "trsm_ex"):
        bad = bad or arg.batch_count < 0
    return bad


def _set_workspace(handle, arg):
    """user_allocated_workspace: hand the library a caller-owned buffer of
    that many bytes (the reference's rocblas_local_handle does the same)."""
    import torch
    if not arg.user_allocated_workspace:
        return None
    buf = torch.empty(arg.user_allocated_workspace, dtype=torch.uint8, device="cuda")
    _expect(lib().rocblas_set_workspace(handle, buf.data_ptr(), arg.user_allocated_workspace),
            S.success, "rocblas_set_workspace")
    return buf


def testing_trsm(handle, arg, rng=None):
    """Run one solve case (any of the six functions, any api); raises
    CaseFailure on a wrong status or a result out of tolerance.  Returns the
    host and device-pointer-mode errors (and timings when arg.timing)."""
    import torch
    rng = rng or np.random.default_rng(arg.seed)
    call = Call(arg)
    lib().rocblas_set_pointer_mode(handle, 0)
    if _invalid_size(arg) or (arg.function not in ("trsm", "trsm_ex") and arg.batch_count == 0):
        # checked through the public entry point whatever the api, as the
        # reference does (testing_trsm.hpp:268-279)
        want = S.invalid_size if _invalid_size(arg) else S.success
        public = Call(Arguments(**{**arg.__dict__, "api": 1 if arg.api in (1, 3, 5) else 0}))
        _expect(public(handle, None, None, arg.lda, None, arg.ldb), want, "invalid size / quick")
        return {}
    ws = _set_workspace(handle, arg)
    try:
        return _solve_and_check(handle, arg, call, rng)
    finally:
        if ws is not None:
            lib().rocblas_set_workspace(handle, None, 0)
        del ws


def _solve_and_check(handle, arg, call, rng):
    import torch
    ops = Operands(arg, rng)
    pA, pB = ops.pointers()
    strides = dict(stride_a=ops.sa, stride_b=ops.sb)
    if arg.M == 0 or arg.N == 0:
        alpha_h = _scalar(arg, arg.alpha)
        _expect(call(handle, ctypes.addressof(alpha_h), pA, arg.lda, pB, arg.ldb, **strides),
                S.success, "empty solve")
        return {}
    out = {"err_host": 0.0, "err_device": 0.0}
    tol = 40 * arg.eps * max(arg.M, 1)
    inv = _trsm_ex_inverse_blocks(handle, arg, ops)
    if inv is not None:
        strides = dict(strides, invA=inv[0].data_ptr(), invA_size=inv[1])
    modes = [m for m, on in ((0, arg.pointer_mode_host), (1, arg.pointer_mode_device)) if on]
    if arg.api in (4, 5):
        modes = [m for m in modes if m == 0] or [0]
    for mode in modes if (arg.unit_check or arg.norm_check) else []:
        ops.reset()
        lib().rocblas_set_pointer_mode(handle, mode)
        alpha_h = _scalar(arg, arg.alpha)
        dalpha = torch.tensor([arg.alpha], dtype=ops.dA.dtype, device="cuda")
        ap = ctypes.addressof(alpha_h) if mode == 0 else dalpha.data_ptr()
        if arg.graph_test:
            st = _captured(handle, lambda: call(handle, ap, pA, arg.lda, pB, arg.ldb, **strides))
        else:
            st = call(handle, ap, pA, arg.lda, pB, arg.ldb, **strides)
        lib().rocblas_set_pointer_mode(handle, 0)
        _expect(st, S.success, f"solve (pointer mode {'host' if mode == 0 else 'device'})")
        torch.cuda.synchronize()
        if arg.repeatability_check:
            _repeat(handle, arg, call, ops, ap, mode, strides)
        err = 0.0
        for b, Xg in enumerate(ops.result()):
            Xg = Xg[:arg.M, :arg.N]
            if arg.alpha == 0:
                if arg.unit_check and np.any(Xg != 0):
                    raise CaseFailure("alpha == 0: result is not zero")
                err = max(err, float(np.abs(Xg).sum()))
                continue
            fwd = norm1_rel(ops.X[b][:arg.M, :arg.N], Xg)
            res = norm1_rel(host_trmm(arg, ops.A[b], Xg, 1.0 / arg.alpha), ops.B[b][:arg.M, :arg.N])
            if arg.unit_check and not (fwd <= tol and res <= tol):
                raise CaseFailure(f"instance {b}: forward error {fwd:.3e}, residual {res:.3e} "
                                  f"> {tol:.3e}")
            err = max(err, fwd, res)
        out["err_host" if mode == 0 else "err_device"] = err
    if arg.timing:
        out.update(time_trsm(handle, arg, call, ops))
    return out


TRSM_BLOCK = 128  # the reference clients' diagonal block of invA (testing_trsm_ex.hpp:29)


def _trsm_ex_inverse_blocks(handle, arg, ops):
    """trsm_ex: invA built as the reference's test builds it
    (clients/include/blas_ex/testing_trsm_ex.hpp:398-440): the inverses of A's
    TRSM_BLOCK x TRSM_BLOCK diagonal blocks (and of the remainder block) by
    rocblas_?trtri_strided_batched, passed with invA_size = TRSM_BLOCK * K.
    The library accepts invA and solves without it; the result must meet the
    same tolerance.  Returns (device tensor, invA_size) or None."""
    import torch
    if arg.function != "trsm_ex" or arg.api != 0:
        return None
    K = arg.M if arg.side.upper() == "L" else arg.N
    if K <= 0:
        return None
    p = "s" if arg.precision == "f32_r" else "d"
    fn = getattr(lib(), f"rocblas_{p}trtri_strided_batched")
    blocks = K // TRSM_BLOCK
    dinv = torch.zeros(TRSM_BLOCK * TRSM_BLOCK * (blocks + 1), dtype=ops.dA.dtype, device="cuda")
    es = ops.dA.element_size()
    a0 = ops.dA.data_ptr()
    stride_a = TRSM_BLOCK * arg.lda + TRSM_BLOCK
    stride_inv = TRSM_BLOCK * TRSM_BLOCK
    uplo, diag = FILL[arg.uplo.upper()], DIAG[arg.diag.upper()]
    if blocks > 0:
        _expect(fn(handle, uplo, diag, TRSM_BLOCK, a0, arg.lda, stride_a, dinv.data_ptr(),
                   TRSM_BLOCK, stride_inv, blocks), S.success, "trtri of invA blocks")
    rem = K - TRSM_BLOCK * blocks
    if rem:
        _expect(fn(handle, uplo, diag, rem, a0 + stride_a * blocks * es, arg.lda, stride_a,
                   dinv.data_ptr() + stride_inv * blocks * es, TRSM_BLOCK, stride_inv, 1),
                S.success, "trtri of the invA remainder")
    return dinv, TRSM_BLOCK * K


def _captured(handle, fn):
    """Issue fn's library calls into a HIP graph on a fresh stream and replay
    it once (the reference's graph_test: client_utility.hpp:138-150)."""
    import torch
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    lib().rocblas_set_stream(handle, s.cuda_stream)
    try:
        with torch.cuda.graph(g, stream=s):
            st = fn()
    finally:
        lib().rocblas_set_stream(handle, None)
    if st == S.success:
        g.replay()
        torch.cuda.synchronize()
    return st


def _repeat(handle, arg, call, ops, alpha_ptr, mode, strides):
    """The reference's repeatability_check (testing_trsm.hpp:468-520): the
    same solve run iters more times from the same input gives the same bits."""
    import torch
    first = ops.dB.clone()
    pA, pB = ops.pointers()
    for _ in range(max(arg.iters, 1)):
        ops.reset()
        lib().rocblas_set_pointer_mode(handle, mode)
        st = call(handle, alpha_ptr, pA, arg.lda, pB, arg.ldb, **strides)
        lib().rocblas_set_pointer_mode(handle, 0)
        _expect(st, S.success, "repeated solve")
        torch.cuda.synchronize()
        if not torch.equal(ops.dB, first):
            raise CaseFailure("repeated solve differs from the first")


def time_trsm(handle, arg, call, ops):
    """cold_iters untimed calls, then iters timed ones on the handle's stream
    (testing_trsm.hpp:547-570), rotating through copies of the operands when
    flush_batch_count / flush_memory_size ask for it (client.cpp:1674-1697);
    the host BLAS solve timed once."""
    import torch
    lib().rocblas_set_pointer_mode(handle, 0)
    alpha_h = _scalar(arg, arg.alpha)
    foot = (arg.K * arg.K // 2 + 2 * arg.M * arg.N) * ops.elt * ops.batch
    copies = max(1, arg.flush_batch_count)
    if arg.flush_memory_size:
        copies = max(1, -(-arg.flush_memory_size // max(foot, 1)))
    if arg.api in (4, 5):
        copies = 1
    sets = [(ops.dA, ops.dB)] + [(ops.dA.clone(), ops.dB.clone()) for _ in range(copies - 1)]
    ptrs = [ops.pointers() + (None, None)]
    for dA, dB in sets[1:]:
        if ops.pa is not None:
            pa = torch.tensor([dA.data_ptr() + b * ops.sa * ops.elt for b in range(ops.batch)],
                              dtype=torch.int64, device="cuda")
            pb = torch.tensor([dB.data_ptr() + b * ops.sb * ops.elt for b in range(ops.batch)],
                              dtype=torch.int64, device="cuda")
            ptrs.append((pa.data_ptr(), pb.data_ptr(), pa, pb))
        else:
            ptrs.append((dA.data_ptr(), dB.data_ptr(), None, None))
    ops.reset()
    t0 = 0.0
    for i in range(arg.cold_iters + arg.iters):
        if i == arg.cold_iters:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
        pA, pB = ptrs[i % len(ptrs)][:2]
        _expect(call(handle, ctypes.addressof(alpha_h), pA, arg.lda, pB, arg.ldb,
                     stride_a=ops.sa, stride_b=ops.sb), S.success, "timed solve")
    torch.cuda.synchronize()
    gpu_us = (time.perf_counter() - t0) * 1e6
    t1 = time.perf_counter()
    for b in range(ops.batch):
        host_trsm(arg, ops.A[b], ops.B[b], arg.alpha)
    cpu_us = (time.perf_counter() - t1) * 1e6
    return {"gpu_us": gpu_us, "cpu_us": cpu_us}


def testing_trsm_bad_arg(handle, arg):
    """The reference's bad-argument sequence (testing_trsm.hpp:33-186 and the
    batched / strided-batched / _ex versions) in both pointer modes."""
    import torch
    a = Arguments(**{**arg.__dict__, "function": arg.function[:-len("_bad_arg")]})
    call = Call(a)
    batched = a.function not in ("trsm", "trsm_ex")
    M = N = lda = ldb = 100
    batch = 2 if batched else 1
    for mode in (0, 1):
        lib().rocblas_set_pointer_mode(handle, mode)
        try:
            dt = torch.float32 if a.precision == "f32_r" else torch.float64
            dA = torch.zeros(lda * M * batch, dtype=dt, device="cuda")
            dB = torch.zeros(ldb * N * batch, dtype=dt, device="cuda")
            if call.form == "batched":
                pa_t = torch.tensor([dA.data_ptr() + b * lda * M * dA.element_size()
                                     for b in range(batch)], dtype=torch.int64, device="cuda")
                pb_t = torch.tensor([dB.data_ptr() + b * ldb * N * dB.element_size()
                                     for b in range(batch)], dtype=torch.int64, device="cuda")
                pA, pB = pa_t.data_ptr(), pb_t.data_ptr()
            else:
                pA, pB = dA.data_ptr(), dB.data_ptr()
            one_h, zero_h = _scalar(a, 1.0), _scalar(a, 0.0)
            one_d = torch.ones(1, dtype=dt, device="cuda")
            zero_d = torch.zeros(1, dtype=dt, device="cuda")
            one = ctypes.addressof(one_h) if mode == 0 else one_d.data_ptr()
            zero = ctypes.addressof(zero_h) if mode == 0 else zero_d.data_ptr()
            kw = dict(stride_a=lda * M, stride_b=ldb * N, batch=batch)
            E, V, P, H, OK = S.invalid_value, S.invalid_size, S.invalid_pointer, \
                S.invalid_handle, S.success
            _expect(call(handle, one, pA, lda, pB, ldb, M, N, side=BOTH, **kw), E, "side")
            _expect(call(handle, one, pA, lda, pB, ldb, M, N, uplo=BOTH, **kw), E, "uplo")
            _expect(call(handle, one, pA, lda, pB, ldb, M, N, trans=BOTH, **kw), E, "transA")
            _expect(call(handle, one, pA, lda, pB, ldb, M, N, diag=BOTH, **kw), E, "diag")
            _expect(call(handle, one, pA, lda, pB, ldb, -1, N, **kw), V, "M < 0")
            _expect(call(handle, one, pA, lda, pB, ldb, M, -1, **kw), V, "N < 0")
            _expect(call(handle, one, pA, lda, pB, M - 1, M, N, **kw), V, "ldb < M")
            _expect(call(handle, one, pA, M - 1, pB, ldb, M, N, side=141, **kw), V, "lda < M")
            _expect(call(handle, one, pA, N - 1, pB, ldb, M, N, side=142, **kw), V, "lda < N")
            if batched:
                _expect(call(handle, one, pA, lda, pB, ldb, M, N, **{**kw, "batch": -1}), V,
                        "batch_count < 0")
            _expect(call(None, one, pA, lda, pB, ldb, M, N, **kw), H, "null handle")
            _expect(call(handle, None, pA, lda, pB, ldb, M, N, **kw), P, "null alpha")
            if mode == 0:
                _expect(call(handle, one, None, lda, pB, ldb, M, N, **kw), P, "null A")
            _expect(call(handle, one, pA, lda, None, ldb, M, N, **kw), P, "null B")
            _expect(call(handle, zero, None, lda, pB, ldb, M, N, **kw), OK, "alpha 0, null A")
            _expect(call(handle, None, None, lda, None, ldb, 0, N, **kw), OK, "M == 0")
            _expect(call(handle, None, None, lda, None, ldb, M, 0, **kw), OK, "N == 0")
            if batched:
                _expect(call(handle, None, None, lda, None, ldb, M, N, **{**kw, "batch": 0}), OK,
                        "batch_count == 0")
            torch.cuda.synchronize()
        finally:
            lib().rocblas_set_pointer_mode(handle, 0)
    return {}


_WS_MAX = [ctypes.c_int32, _i64, _i64, _i64] + [ctypes.POINTER(ctypes.c_size_t)] * 5
_WS = [ctypes.c_int32] * 6 + [ctypes.POINTER(ctypes.c_size_t)] * 5


def testing_trsm_internal_interfaces(arg):
    """The reference's check that the max-size workspace query covers every
    smaller size (testing_trsm.hpp:188-240, testing_trsm_batched.hpp:290-350).
    Sizes are walked in strides so the largest cases stay quick; the queries
    are host-only."""
    T = "float" if arg.precision == "f32_r" else "double"
    batched = arg.function.startswith("trsm_batched")
    pre = "rocblas_internal_trsm_batched" if batched else "rocblas_internal_trsm"
    mx = internal_fn(f"{pre}_workspace_max_size_64<{T}>",
                     _WS_MAX if batched else _WS_MAX[:4] + _WS_MAX[4:7])
    ws = internal_fn(f"{pre}_workspace_size<{T}>", _WS)
    side = SIDE[arg.side.upper()]
    top = [ctypes.c_size_t(0) for _ in range(5)]
    refs = [ctypes.byref(x) for x in top]
    st = mx(side, arg.M, arg.N, arg.batch_count, *(refs if batched else [refs[0], refs[2], refs[4]]))
    _expect(st, S.success, "workspace_max_size_64")
    step_m = max(1, arg.M // 64)
    step_n = max(1, arg.N // 64)
    for m in range(arg.M, 0, -step_m):
        for n in range(arg.N, 0, -step_n):
            got = [ctypes.c_size_t(0) for _ in range(5)]
            st = ws(side, 112, m, n, arg.batch_count, 0, *[ctypes.byref(x) for x in got])
            if st not in (S.success, S.continue_):
                raise CaseFailure(f"workspace_size: {S(st).name}")
            if got[0].value > top[0].value or got[2].value > top[2].value \
                    or got[4].value > top[4].value:
                raise CaseFailure(f"m={m} n={n}: workspace above the reported maximum")
    return {}


def run_case(handle, arg):
    """Dispatch one case by function name."""
    if arg.function in BAD_ARG_FUNCS:
        return testing_trsm_bad_arg(handle, arg)
    if arg.function in INTERNAL_FUNCS:
        return testing_trsm_internal_interfaces(arg)
    return testing_trsm(handle, arg)


# --------------------------------------------------------------------------
# the result line (argument_model.hpp:56-247)
# --------------------------------------------------------------------------
def log_args(arg, res, log_function_name=False, log_datatype=False):
    """Header and value lines as rocblas-bench prints them for trsm."""
    names, vals = [], []
    if log_function_name:
        names.append("function")
        vals.append(arg.function)
    if log_datatype:
        for key in ("a_type", "b_type", "c_type", "d_type", "compute_type"):
            names.append(key)
            vals.append(arg.precision)
    fields_ = [("side", arg.side), ("uplo", arg.uplo), ("transA", arg.transA),
               ("diag", arg.diag), ("M", arg.M), ("N", arg.N), ("alpha", f"{arg.alpha:g}"),
               ("lda", arg.lda)]
    strided = "strided" in arg.function
    if strided:
        fields_.append(("stride_a", arg.stride_a))
    fields_.append(("ldb", arg.ldb))
    if strided:
        fields_.append(("stride_b", arg.stride_b))
    if "batched" in arg.function:
        fields_.append(("batch_count", arg.batch_count))
    for k, v in fields_:
        names.append(k)
        vals.append(str(v))
    line_n = ",".join(names)
    line_v = ",".join(vals)
    hot = max(arg.iters, 1)
    gpu_us = res["gpu_us"] / hot
    batch = arg.batch_count if "batched" in arg.function else 1
    gflops = arg.M * arg.N * arg.K / 1e9
    line_n += ",cold_iters,hot_iters,rocblas-Gflops,us"
    line_v += f", {arg.cold_iters}, {arg.iters}, {gflops * batch / gpu_us * 1e6:g}, {gpu_us:g}"
    if arg.unit_check or arg.norm_check:
        cpu = res["cpu_us"]
        line_n += ",CPU-Gflops,CPU-us"
        line_v += f",{gflops * batch / cpu * 1e6:g},{cpu:g}"
        if arg.norm_check:
            line_n += ",norm_error_1,norm_error_2"
            line_v += f",{res.get('err_host', 0.0):g},{res.get('err_device', 0.0):g}"
    return line_n + "\n" + line_v


def cases_from(path=None, include_dirs=(), doc=None):
    """Expanded cases of a gtest YAML file or of a JSON document holding the
    same structure (Defaults, Dictionary lists to expand, Tests)."""
    if doc is None:
        if path.endswith(".json"):
            import json
            with open(path) as f:
                doc = json.load(f)
        else:
            doc = load_yaml_doc(path, include_dirs)
    return expand_doc(doc)


def dedupe(cases):
    """Drop cases that make the same library calls as an earlier one: the
    FORTRAN apis call the C entry points here."""
    seen, out = set(), []
    for c in cases:
        c2 = dict(c)
        api = c2.get("api", "C")
        c2["api"] = {"FORTRAN": "C", "FORTRAN_64": "C_64"}.get(api, api)
        key = repr(sorted((k, repr(v)) for k, v in c2.items() if k != "name"))
        if key in seen:
            continue
        seen.add(key)
        out.append(c2)
    return out


def case_bytes(arg):
    """Device bytes a case allocates (A and B copies), for size limits."""
    elt = 4 if arg.precision == "f32_r" else 8
    batch = max(arg.batch_count, 1)
    sa = max(arg.stride_a, arg.lda * arg.K) if "strided" in arg.function else arg.lda * arg.K
    sb = max(arg.stride_b, arg.ldb * arg.N) if "strided" in arg.function else arg.ldb * arg.N
    return elt * batch * (max(sa, 0) + max(sb, 0))


__all__ = ["Arguments", "cases_from", "expand_doc", "dedupe", "case_kind", "run_case",
           "testing_trsm", "testing_trsm_bad_arg", "testing_trsm_internal_interfaces",
           "log_args", "CaseFailure", "case_bytes"]
