"""ctypes binding of the C ABI in include/hpcgx.h (hpcg_amd/libhpcgx.so).

The library is built in-tree by `python -m hpcg_amd.build` (or
__graft_entry__.build()).  There is no fallback: if the library is missing or
fails to load, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhpcgx.so")

HPCGX_PREC_FP64, HPCGX_PREC_FP32, HPCGX_PREC_FP16 = 0, 1, 2
HPCGX_CG_FP64, HPCGX_CG_MIXED_FP32, HPCGX_CG_MIXED_FP16 = 0, 1, 2
HPCGX_COLS_EXPLICIT, HPCGX_COLS_COMPACT, HPCGX_COLS_COMPACT_ALWAYS = 0, 1, 2

ERRORS = {0: "ok", -1: "EINVAL", -2: "ENOTSTENCIL", -3: "EHIP", -4: "ENOMEM", -5: "ESTATE"}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

# name -> (restype, argtypes); ctx is c_void_p, device pointers are c_void_p
_vp = C.c_void_p
SIGNATURES = {
    "hpcgx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "hpcgx_destroy": (C.c_int, [_vp]),
    "hpcgx_error_string": (C.c_char_p, [_vp]),
    "hpcgx_set_stream": (C.c_int, [_vp, _vp]),
    "hpcgx_sync": (C.c_int, [_vp]),
    "hpcgx_set_colour_order": (C.c_int, [_vp, _ip]),
    "hpcgx_level_upload": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "hpcgx_level_upload_packed": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _u8p, _ip, _dp, _vp]),
    "hpcgx_level_upload_rows": (C.c_int, [_vp, C.c_int, C.c_int64, _u8p, _ip, _dp, _vp, C.c_int64]),
    "hpcgx_level_colouring": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _vp]),
    "hpcgx_colour_rows": (C.c_int, [C.c_int64, _u8p, _ip, _vp, C.c_int64, _ip, _ip, C.POINTER(C.c_int)]),
    "hpcgx_problem_generate": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp]),
    "hpcgx_level_info": (C.c_int, [_vp, C.c_int, _i64p]),
    "hpcgx_level_set_diagonal": (C.c_int, [_vp, C.c_int, _dp]),
    "hpcgx_level_get_diagonal": (C.c_int, [_vp, C.c_int, _dp]),
    "hpcgx_level_export_packed": (C.c_int, [_vp, C.c_int, _u8p, _ip, _dp]),
    "hpcgx_prepare_precision": (C.c_int, [_vp, C.c_int]),
    "hpcgx_set_column_format": (C.c_int, [_vp, C.c_int]),
    "hpcgx_level_column_forms": (C.c_int, [_vp, C.c_int, _i64p]),
    "hpcgx_memory_bytes": (C.c_double, [_vp]),
    "hpcgx_vec_alloc": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_void_p)]),
    "hpcgx_vec_free": (C.c_int, [_vp, _vp]),
    "hpcgx_vec_upload": (C.c_int, [_vp, C.c_int, _vp, _dp, C.c_int64]),
    "hpcgx_vec_download": (C.c_int, [_vp, C.c_int, _vp, _dp, C.c_int64]),
    "hpcgx_vec_fill": (C.c_int, [_vp, C.c_int, _vp, C.c_double]),
    "hpcgx_vec_from_natural": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "hpcgx_vec_to_natural": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "hpcgx_spmv": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "hpcgx_symgs": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "hpcgx_restrict": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    "hpcgx_prolong": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "hpcgx_mg": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "hpcgx_mg_mixed": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "hpcgx_dot": (C.c_int, [_vp, C.c_int64, _vp, _vp, C.POINTER(C.c_double)]),
    "hpcgx_waxpby": (C.c_int, [_vp, C.c_int64, C.c_double, _vp, C.c_double, _vp, _vp]),
    "hpcgx_waxpby_dot": (C.c_int, [_vp, C.c_int64, C.c_double, _vp, C.c_double, _vp, _vp,
                                   C.POINTER(C.c_double)]),
    "hpcgx_dot_noalloc": (C.c_int, [_vp, C.c_int64, _vp, _vp, C.POINTER(C.c_double)]),
    "hpcgx_cg": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, C.c_double, C.c_int,
                           C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double), _vp, _vp]),
    "hpcgx_kappa_grad": (C.c_int, [_vp, _vp, _vp, _vp]),
}

_lib = None


class HpcgxError(RuntimeError):
    pass


def lib():
    """Loads libhpcgx.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HpcgxError(f"{LIB_PATH} is not built: run `python -m hpcg_amd.build`")
        # A process may hold only one HIP runtime.  torch ships its own
        # libamdhip64 and reaches it under a file name the dynamic linker does
        # not match against an already-loaded /opt/rocm copy, so if this
        # library came first, torch would load a second runtime whose device
        # discovery then fails ("No HIP GPUs are available").  Loaded after
        # torch, libhpcgx.so binds to torch's runtime by its soname.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (rt, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = rt
            f.argtypes = args
        _lib = L
    return _lib


def colour_rows(nnz_row, cols, f2c=None):
    """hpcgx_colour_rows: (colour per row, breadth-first order, colours) of
    rows packed [n, 27], colours ordered by injected rows when f2c is given
    (host only)."""
    nnz_row = np.ascontiguousarray(nnz_row, np.uint8)
    cols = np.ascontiguousarray(cols, np.int32).reshape(-1)
    n = len(nnz_row)
    fa = None if f2c is None else np.ascontiguousarray(f2c, np.int32)
    fp = None if fa is None else fa.ctypes.data_as(C.c_void_p)
    col, order, nc = np.empty(n, np.int32), np.empty(n, np.int32), C.c_int(0)
    rc = lib().hpcgx_colour_rows(n, nnz_row, cols, fp, 0 if fa is None else len(fa), col, order, C.byref(nc))
    if rc != 0:
        raise HpcgxError(f"hpcgx_colour_rows: {ERRORS.get(rc, rc)}")
    return col, order, nc.value


def _ptr(v):
    """Device pointer of a DeviceVector, a torch tensor or an int."""
    if isinstance(v, DeviceVector):
        return v.ptr
    if hasattr(v, "data_ptr"):
        return v.data_ptr()
    return int(v)


class DeviceVector:
    """A device vector of one level, in the context's colour row order."""

    def __init__(self, ctx, level, ptr, npad, n):
        self.ctx, self.level, self.ptr, self.npad, self.n = ctx, level, ptr, npad, n

    def free(self):
        if self.ptr and self.ctx.h:
            self.ctx._call("hpcgx_vec_free", self.ctx.h, self.ptr)
        self.ptr = 0

    def upload(self, host):
        self.ctx.upload(self, host)
        return self

    def download(self):
        return self.ctx.download(self)


class Context:
    """One hpcgx context on a HIP device (the device side of OptimizeProblem)."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.hpcgx_create(device, C.byref(h))
        if rc != 0:
            raise HpcgxError(f"hpcgx_create failed: {ERRORS.get(rc, rc)}")
        self.h = h.value
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            lib().hpcgx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _call(self, name, *args):
        rc = getattr(lib(), name)(*args)
        if rc != 0:
            msg = lib().hpcgx_error_string(self.h)
            raise HpcgxError(f"{name}: {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")
        return rc

    # ---- problem
    def set_colour_order(self, order):
        self._call("hpcgx_set_colour_order", self.h, np.ascontiguousarray(order, np.int32))

    def generate(self, nx, ny, nz, levels=4, kappa=None):
        kp = None
        if kappa is not None:
            kappa = np.ascontiguousarray(kappa, np.float64)
            kp = kappa.ctypes.data_as(C.c_void_p)
        self._call("hpcgx_problem_generate", self.h, nx, ny, nz, levels, kp)
        return self

    def upload_level(self, level, dims, nnz_row, cols, vals, f2c=None):
        nnz_row = np.ascontiguousarray(nnz_row, np.uint8)
        cols = np.ascontiguousarray(cols, np.int32).reshape(-1)
        vals = np.ascontiguousarray(vals, np.float64).reshape(-1)
        f2c_arr = None if f2c is None else np.ascontiguousarray(f2c, np.int32)
        fp = None if f2c_arr is None else f2c_arr.ctypes.data_as(C.c_void_p)
        self._call("hpcgx_level_upload_packed", self.h, level, dims[0], dims[1], dims[2],
                   nnz_row, cols, vals, fp)

    def upload_rows(self, level, nnz_row, cols, vals, f2c=None, n_coarse=0):
        """General rows without a grid (hpcgx_level_upload_rows)."""
        nnz_row = np.ascontiguousarray(nnz_row, np.uint8)
        cols = np.ascontiguousarray(cols, np.int32).reshape(-1)
        vals = np.ascontiguousarray(vals, np.float64).reshape(-1)
        f2c_arr = None if f2c is None else np.ascontiguousarray(f2c, np.int32)
        fp = None if f2c_arr is None else f2c_arr.ctypes.data_as(C.c_void_p)
        self._call("hpcgx_level_upload_rows", self.h, level, len(nnz_row), nnz_row, cols, vals, fp,
                   0 if f2c_arr is None else (n_coarse or len(f2c_arr)))

    def colouring(self, level=0):
        """(colour of each natural row, number of colours, general layout?)"""
        nc, gen = C.c_int(0), C.c_int(0)
        col = np.zeros(self.info(level)["n"], np.int32)
        self._call("hpcgx_level_colouring", self.h, level, C.byref(nc), C.byref(gen),
                   col.ctypes.data_as(C.c_void_p))
        return col, nc.value, bool(gen.value)

    def info(self, level=0):
        out = np.zeros(8, np.int64)
        self._call("hpcgx_level_info", self.h, level, out)
        keys = ("n", "npad", "nnz", "nx", "ny", "nz", "slots", "levels")
        return {k: int(v) for k, v in zip(keys, out)}

    def export_level(self, level):
        n = self.info(level)["n"]
        nr = np.zeros(n, np.uint8)
        cols = np.zeros(n * 27, np.int32)
        vals = np.zeros(n * 27, np.float64)
        self._call("hpcgx_level_export_packed", self.h, level, nr, cols, vals)
        return nr, cols.reshape(n, 27), vals.reshape(n, 27)

    def set_diagonal(self, level, d):
        self._call("hpcgx_level_set_diagonal", self.h, level, np.ascontiguousarray(d, np.float64))

    def get_diagonal(self, level):
        d = np.zeros(self.info(level)["n"])
        self._call("hpcgx_level_get_diagonal", self.h, level, d)
        return d

    def prepare(self, precision):
        self._call("hpcgx_prepare_precision", self.h, precision)

    def set_column_format(self, fmt):
        """HPCGX_COLS_COMPACT (default), HPCGX_COLS_COMPACT_ALWAYS or HPCGX_COLS_EXPLICIT."""
        self._call("hpcgx_set_column_format", self.h, fmt)

    def column_forms(self, level=0):
        """Slices of a level per column form: {'delta', 'explicit', 'uniform'}
        (see column_reads for the form the kernels read)."""
        out = np.zeros(4, np.int64)
        self._call("hpcgx_level_column_forms", self.h, level, out)
        return {"delta": int(out[0]), "explicit": int(out[1]), "uniform": int(out[2])}

    def column_reads(self, level=0):
        """'compact' if the level's kernels read the compact copy, else 'explicit'."""
        out = np.zeros(4, np.int64)
        self._call("hpcgx_level_column_forms", self.h, level, out)
        return "compact" if out[3] else "explicit"

    def memory_bytes(self):
        return lib().hpcgx_memory_bytes(self.h)

    def set_stream(self, stream_ptr):
        self._call("hpcgx_set_stream", self.h, stream_ptr)

    def sync(self):
        self._call("hpcgx_sync", self.h)

    # ---- vectors
    def vec(self, level=0, host=None):
        p = C.c_void_p()
        self._call("hpcgx_vec_alloc", self.h, level, C.byref(p))
        inf = self.info(level)
        v = DeviceVector(self, level, p.value, inf["npad"], inf["n"])
        if host is not None:
            self.upload(v, host)
        return v

    def upload(self, v, host, level=None):
        level = v.level if level is None else level
        host = np.ascontiguousarray(host, np.float64)
        self._call("hpcgx_vec_upload", self.h, level, _ptr(v), host, host.size)

    def download(self, v, level=None):
        level = v.level if level is None else level
        out = np.empty(self.info(level)["n"])
        self._call("hpcgx_vec_download", self.h, level, _ptr(v), out, out.size)
        return out

    def from_natural(self, v, nat, level=None):
        """device natural-order buffer (torch tensor / pointer) -> device vector v"""
        level = v.level if level is None else level
        self._call("hpcgx_vec_from_natural", self.h, level, _ptr(v), _ptr(nat))

    def to_natural(self, v, nat, level=None):
        level = v.level if level is None else level
        self._call("hpcgx_vec_to_natural", self.h, level, _ptr(v), _ptr(nat))

    def fill(self, v, value, level=None):
        level = v.level if level is None else level
        self._call("hpcgx_vec_fill", self.h, level, _ptr(v), float(value))

    # ---- kernels
    def spmv(self, level, x, y):
        self._call("hpcgx_spmv", self.h, level, _ptr(x), _ptr(y))

    def symgs(self, level, r, x):
        self._call("hpcgx_symgs", self.h, level, _ptr(r), _ptr(x))

    def restrict(self, level, r, x, rc):
        self._call("hpcgx_restrict", self.h, level, _ptr(r), _ptr(x), _ptr(rc))

    def prolong(self, level, xc, x):
        self._call("hpcgx_prolong", self.h, level, _ptr(xc), _ptr(x))

    def mg(self, level, r, x):
        self._call("hpcgx_mg", self.h, level, _ptr(r), _ptr(x))

    def mg_mixed(self, precision, r, x):
        self._call("hpcgx_mg_mixed", self.h, precision, _ptr(r), _ptr(x))

    def dot(self, x, y, n=None):
        n = self.info(0)["n"] if n is None else n
        out = C.c_double()
        self._call("hpcgx_dot", self.h, n, _ptr(x), _ptr(y), C.byref(out))
        return out.value

    def dot_noalloc(self, x, y, n=None):
        """hpcgx_dot_noalloc: the dot with its partial sums in the library's
        static device scratch (allocates nothing)."""
        n = self.info(0)["n"] if n is None else n
        out = C.c_double()
        self._call("hpcgx_dot_noalloc", self.h, n, _ptr(x), _ptr(y), C.byref(out))
        return out.value

    def waxpby(self, alpha, x, beta, y, w, n=None):
        n = self.info(0)["n"] if n is None else n
        self._call("hpcgx_waxpby", self.h, n, alpha, _ptr(x), beta, _ptr(y), _ptr(w))

    def waxpby_dot(self, alpha, x, beta, y, w, n=None):
        n = self.info(0)["n"] if n is None else n
        out = C.c_double()
        self._call("hpcgx_waxpby_dot", self.h, n, alpha, _ptr(x), beta, _ptr(y), _ptr(w), C.byref(out))
        return out.value

    def cg(self, b, x, max_iter=50, tol=0.0, variant=HPCGX_CG_FP64, precond=True, times=None,
           history=False):
        it, nr, nr0 = C.c_int(0), C.c_double(0), C.c_double(0)
        hist = np.zeros(max_iter + 1) if history else None
        tp = None if times is None else times.ctypes.data_as(C.c_void_p)
        hp = None if hist is None else hist.ctypes.data_as(C.c_void_p)
        self._call("hpcgx_cg", self.h, variant, _ptr(b), _ptr(x), max_iter, tol, int(precond),
                   C.byref(it), C.byref(nr), C.byref(nr0), tp, hp)
        out = dict(niters=it.value, normr=nr.value, normr0=nr0.value)
        if history:
            out["hist"] = hist[: it.value + 1]
        return out

    def kappa_grad(self, lam, x, g):
        self._call("hpcgx_kappa_grad", self.h, _ptr(lam), _ptr(x), _ptr(g))
