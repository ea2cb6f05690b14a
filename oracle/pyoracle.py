"""ctypes access to the checkers: the C restatement (libhpcg_oracle.so) and
the reference's own code (oracle/_ref/libhpcg_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the measured or
shipped path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libhpcg_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhpcg_ref.so")
REF_OMP_SO = os.path.join(HERE, "_ref", "libhpcg_ref_omp.so")
REF_BIN = os.path.join(HERE, "_ref", "xhpcg_ref")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

_lib = None
_ref = {}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle oracle`")
        L = C.CDLL(ORACLE_SO)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp]
        L.orc_destroy.argtypes = [vp]
        L.orc_coarse.restype = vp
        L.orc_coarse.argtypes = [vp]
        L.orc_nrow.restype = C.c_int
        L.orc_nrow.argtypes = [vp]
        L.orc_nnz.restype = C.c_int64
        L.orc_nnz.argtypes = [vp]
        L.orc_dims.argtypes = [vp, _ip]
        for name, rt in (("orc_nnz_row", C.POINTER(C.c_uint8)), ("orc_cols", C.POINTER(C.c_int)),
                         ("orc_vals", C.POINTER(C.c_double)), ("orc_f2c", C.POINTER(C.c_int)),
                         ("orc_kappa", C.POINTER(C.c_double))):
            getattr(L, name).restype = rt
            getattr(L, name).argtypes = [vp]
        L.orc_rhs.argtypes = [vp, _dp, _dp, _dp]
        L.orc_get_diagonal.argtypes = [vp, _dp]
        L.orc_set_diagonal.argtypes = [vp, _dp]
        L.orc_set_colour_order.argtypes = [vp, _ip]
        L.orc_colour_greedy.argtypes = [vp, _ip, C.POINTER(C.c_int)]
        L.orc_spmv.argtypes = [vp, _dp, _dp]
        L.orc_symgs.argtypes = [vp, _dp, _dp, C.c_int]
        L.orc_restrict.argtypes = [vp, _dp, _dp, _dp]
        L.orc_prolong.argtypes = [vp, _dp, _dp]
        L.orc_mg.argtypes = [vp, _dp, _dp, C.c_int]
        L.orc_dot.restype = C.c_double
        L.orc_dot.argtypes = [C.c_int64, _dp, _dp]
        L.orc_waxpby.argtypes = [C.c_int64, C.c_double, _dp, C.c_double, _dp, _dp]
        L.orc_cg.restype = C.c_int
        L.orc_cg.argtypes = [vp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int,
                             C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double), vp]
        L.orc_mg_f32.argtypes = [vp, _fp, _fp]
        L.orc_mg_f16.argtypes = [vp, _fp, _fp]
        L.orc_plant_fault.argtypes = [C.c_int]
        L.orc_cg_mixed.restype = C.c_int
        L.orc_cg_mixed.argtypes = [vp, _dp, _dp, C.c_int, C.c_double, C.c_int,
                                   C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double), vp]
        L.orc_kappa_grad.argtypes = [vp, _dp, _dp, _dp]
        L.orc_create_rows.restype = vp
        L.orc_create_rows.argtypes = [C.c_int, _u8p, _ip, _dp]
        L.orc_link.restype = C.c_int
        L.orc_link.argtypes = [vp, vp, _ip]
        L.orc_colour_general.argtypes = [vp, _ip, C.POINTER(C.c_int)]
        L.orc_colour_general_order.argtypes = [vp, _ip, C.POINTER(C.c_int), _ip]
        L.orc_set_colouring.restype = C.c_int
        L.orc_set_colouring.argtypes = [vp, _ip, C.c_int]
        L.orc_level_colours.argtypes = [vp, _ip]
        L.orc_ncolours.restype = C.c_int
        L.orc_ncolours.argtypes = [vp]
        L.orc_f32_to_f16.restype = C.c_uint16
        L.orc_f32_to_f16.argtypes = [C.c_float]
        L.orc_f16_to_f32.restype = C.c_float
        L.orc_f16_to_f32.argtypes = [C.c_uint16]
        _lib = L
    return _lib


class Level:
    """One level of an oracle hierarchy (natural row order)."""

    def __init__(self, handle, owner):
        self.h = handle
        self._owner = owner
        L = lib()
        self.n = L.orc_nrow(handle)
        self.nnz = L.orc_nnz(handle)
        d = np.zeros(3, np.int32)
        L.orc_dims(handle, d)
        self.dims = tuple(int(v) for v in d)

    def arrays(self):
        """(nnz_row[n], cols[n,27], vals[n,27], f2c or None) -- copies."""
        L = lib()
        n = self.n
        nnz_row = np.ctypeslib.as_array(L.orc_nnz_row(self.h), (n,)).copy()
        cols = np.ctypeslib.as_array(L.orc_cols(self.h), (n * 27,)).reshape(n, 27).copy()
        vals = np.ctypeslib.as_array(L.orc_vals(self.h), (n * 27,)).reshape(n, 27).copy()
        c = L.orc_coarse(self.h)
        f2c = None
        if c:
            nc = L.orc_nrow(c)
            f2c = np.ctypeslib.as_array(L.orc_f2c(self.h), (nc,)).copy()
        return nnz_row, cols, vals, f2c

    def coarse(self):
        c = lib().orc_coarse(self.h)
        return Level(c, self._owner) if c else None

    def spmv(self, x):
        y = np.empty(self.n)
        lib().orc_spmv(self.h, np.ascontiguousarray(x, np.float64), y)
        return y

    def symgs(self, r, x, mode=2):
        """mode 0 natural (reference), 1 coloured, 2 coloured with the last
        colour's backward step elided (the device's ordering)."""
        x = np.array(x, np.float64, copy=True)
        lib().orc_symgs(self.h, np.ascontiguousarray(r, np.float64), x, int(mode))
        return x

    def restrict(self, r, Axf):
        rc = np.empty(self.coarse().n)
        lib().orc_restrict(self.h, np.ascontiguousarray(r, np.float64),
                           np.ascontiguousarray(Axf, np.float64), rc)
        return rc

    def prolong(self, xc, x):
        x = np.array(x, np.float64, copy=True)
        lib().orc_prolong(self.h, np.ascontiguousarray(xc, np.float64), x)
        return x

    def mg(self, r, mode=2):
        x = np.empty(self.n)
        lib().orc_mg(self.h, np.ascontiguousarray(r, np.float64), x, int(mode))
        return x

    def mg_f32(self, r):
        x = np.empty(self.n, np.float32)
        lib().orc_mg_f32(self.h, np.ascontiguousarray(r, np.float32), x)
        return x

    def mg_f16(self, r):
        x = np.empty(self.n, np.float32)
        lib().orc_mg_f16(self.h, np.ascontiguousarray(r, np.float32), x)
        return x

    def kappa_grad(self, lam, x):
        g = np.empty(self.n)
        lib().orc_kappa_grad(self.h, np.ascontiguousarray(lam, np.float64),
                             np.ascontiguousarray(x, np.float64), g)
        return g

    def diagonal(self):
        d = np.empty(self.n)
        lib().orc_get_diagonal(self.h, d)
        return d

    def set_diagonal(self, d):
        lib().orc_set_diagonal(self.h, np.ascontiguousarray(d, np.float64))

    def greedy_colours(self):
        """orc_colour_greedy (the reference's example, lower neighbours only)."""
        col = np.empty(self.n, np.int32)
        nc = C.c_int(0)
        lib().orc_colour_greedy(self.h, col, C.byref(nc))
        return col, nc.value

    def general_colours(self, with_order=False):
        """orc_colour_general: (colour per row, number of colours[, order])."""
        col = np.empty(self.n, np.int32)
        order = np.empty(self.n, np.int32)
        nc = C.c_int(0)
        lib().orc_colour_general_order(self.h, col, C.byref(nc), order)
        return (col, nc.value, order) if with_order else (col, nc.value)

    def set_colouring(self, colour):
        colour = np.ascontiguousarray(colour, np.int32)
        if lib().orc_set_colouring(self.h, colour, int(colour.max()) + 1) != 0:
            raise ValueError("colouring couples two rows of one colour")

    def ncolours(self):
        return lib().orc_ncolours(self.h)

    def colours(self):
        """The colouring the level is swept in now (after orc_link's ordering)."""
        col = np.empty(self.n, np.int32)
        lib().orc_level_colours(self.h, col)
        return col


class Problem(Level):
    """Oracle hierarchy: HPCG matrix (kappa None) or the kappa operator."""

    def __init__(self, nx, ny, nz, levels=4, kappa=None, _handle=None):
        L = lib()
        self._kappa = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
        kp = None if self._kappa is None else self._kappa.ctypes.data_as(C.c_void_p)
        h = _handle or L.orc_create(nx, ny, nz, levels, kp)
        if not h:
            raise ValueError("bad grid")
        super().__init__(h, self)
        self.levels = levels

    @classmethod
    def from_rows(cls, levels):
        """A hierarchy of general levels, fine to coarse: each entry is
        (nnz_row[n], cols[n,27], vals[n,27], f2c) with f2c the injection onto
        the next level (None on the coarsest).  Each level is coloured with
        orc_colour_general."""
        L = lib()
        hs = []
        for nr, cols, vals, _ in levels:
            nr = np.ascontiguousarray(nr, np.uint8)
            h = L.orc_create_rows(len(nr), nr, np.ascontiguousarray(cols, np.int32).reshape(-1),
                                  np.ascontiguousarray(vals, np.float64).reshape(-1))
            if not h:
                for g in hs:
                    L.orc_destroy(g)
                raise ValueError("rows rejected (empty/long row, column out of range or no diagonal)")
            hs.append(h)
        for k in range(len(hs) - 2, -1, -1):
            f2c = np.ascontiguousarray(levels[k][3], np.int32)
            if L.orc_link(hs[k], hs[k + 1], f2c) != 0:
                for g in hs[: k + 2]:
                    L.orc_destroy(g)
                raise ValueError("bad f2c")
        n = len(levels[0][0])
        return cls(n, 1, 1, levels=len(levels), _handle=hs[0])

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
        except Exception:
            pass

    def level(self, k):
        lv = self
        for _ in range(k):
            lv = lv.coarse()
        return lv

    def rhs(self):
        b, x, xe = np.empty(self.n), np.empty(self.n), np.empty(self.n)
        lib().orc_rhs(self.h, b, x, xe)
        return b, x, xe

    def set_colour_order(self, order):
        lib().orc_set_colour_order(self.h, np.ascontiguousarray(order, np.int32))

    def cg(self, b, x0=None, max_iter=50, tol=0.0, precond=True, mode=2, history=False):
        x = np.zeros(self.n) if x0 is None else np.array(x0, np.float64, copy=True)
        it, nr, nr0 = C.c_int(0), C.c_double(0), C.c_double(0)
        hist = np.zeros(max_iter + 1)
        lib().orc_cg(self.h, np.ascontiguousarray(b, np.float64), x, max_iter, tol, int(precond),
                     int(mode), C.byref(it), C.byref(nr), C.byref(nr0),
                     hist.ctypes.data_as(C.c_void_p))
        out = dict(x=x, niters=it.value, normr=nr.value, normr0=nr0.value)
        if history:
            out["hist"] = hist[: it.value + 1]
        return out

    def cg_mixed(self, b, x0=None, max_iter=500, tol=1e-9, mg_prec=32, history=False):
        x = np.zeros(self.n) if x0 is None else np.array(x0, np.float64, copy=True)
        it, nr, nr0 = C.c_int(0), C.c_double(0), C.c_double(0)
        hist = np.zeros(max_iter + 1)
        lib().orc_cg_mixed(self.h, np.ascontiguousarray(b, np.float64), x, max_iter, tol, mg_prec,
                           C.byref(it), C.byref(nr), C.byref(nr0), hist.ctypes.data_as(C.c_void_p))
        out = dict(x=x, niters=it.value, normr=nr.value, normr0=nr0.value)
        if history:
            out["hist"] = hist[: it.value + 1]
        return out


class planted_fault:
    """Context manager: the oracle's scaled fp16 V-cycle carries a known
    mistake (orc_plant_fault) for hierarchies built inside the block, so a
    test can show that its check catches it.  1: sigma inverted; 2: sigma
    dropped."""

    def __init__(self, which):
        self.which = int(which)

    def __enter__(self):
        lib().orc_plant_fault(self.which)
        return self

    def __exit__(self, *exc):
        lib().orc_plant_fault(0)
        return False


def dot(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return lib().orc_dot(x.size, x, y)


def waxpby(alpha, x, beta, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    w = np.empty_like(x)
    lib().orc_waxpby(x.size, alpha, x, beta, y, w)
    return w


def f32_to_f16_bits(v):
    return lib().orc_f32_to_f16(float(v))


def f16_bits_to_f32(h):
    return lib().orc_f16_to_f32(int(h))


# --------------------------------------------------------------------------
# the reference's own code (oracle/_ref/libhpcg_ref.so)

def ref_available(omp=False):
    return os.path.exists(REF_OMP_SO if omp else REF_SO)


def ref(omp=False):
    """The reference's own code: the serial no-FMA parity build, or (omp) the
    OpenMP build with the flags of setup/Make.GCC_OMP."""
    if omp not in _ref:
        path = REF_OMP_SO if omp else REF_SO
        if not os.path.exists(path):
            raise RuntimeError(f"reference build missing ({path}): run `make -C oracle ref`")
        R = C.CDLL(path)
        vp = C.c_void_p
        R.ref_create.restype = vp
        R.ref_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        R.ref_destroy.argtypes = [vp]
        R.ref_level_info.restype = C.c_int
        R.ref_level_info.argtypes = [vp, C.c_int, _i64p]
        R.ref_level_export.restype = C.c_int
        R.ref_level_export.argtypes = [vp, C.c_int, _u8p, _ip, _dp, _ip, vp]
        R.ref_rhs.argtypes = [vp, _dp, _dp, _dp]
        R.ref_set_diagonal.argtypes = [vp, C.c_int, _dp]
        R.ref_replace_rows.restype = C.c_int
        R.ref_replace_rows.argtypes = [vp, C.c_int, _u8p, _ip, _dp, vp]
        for f in ("ref_spmv", "ref_symgs", "ref_mg"):
            getattr(R, f).restype = C.c_int
            getattr(R, f).argtypes = [vp, C.c_int, _dp, _dp]
        R.ref_dot.restype = C.c_double
        R.ref_dot.argtypes = [C.c_int, _dp, _dp]
        R.ref_waxpby.restype = C.c_int
        R.ref_waxpby.argtypes = [C.c_int, C.c_double, _dp, C.c_double, _dp, _dp]
        R.ref_cg.restype = C.c_int
        R.ref_cg.argtypes = [vp, vp, _dp, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int),
                             C.POINTER(C.c_double), C.POINTER(C.c_double), vp]
        R.ref_timer.restype = C.c_double
        _ref[omp] = R
    return _ref[omp]


class RefProblem:
    """The reference's own problem (GenerateProblem + 3 coarse levels) and _ref kernels."""

    def __init__(self, nx, ny, nz, levels=4, omp=False):
        self.omp = omp
        self.h = ref(omp).ref_create(nx, ny, nz, levels)
        self.levels = levels
        self.info = []
        for l in range(levels):
            out = np.zeros(6, np.int64)
            ref(omp).ref_level_info(self.h, l, out)
            self.info.append(dict(nrow=int(out[0]), ncol=int(out[1]), nnz=int(out[2]),
                                  dims=(int(out[3]), int(out[4]), int(out[5]))))
        self.n = self.info[0]["nrow"]

    def __del__(self):
        try:
            ref(self.omp).ref_destroy(self.h)
        except Exception:
            pass

    def export(self, level):
        n = self.info[level]["nrow"]
        nnz_row = np.zeros(n, np.uint8)
        cols = np.zeros(n * 27, np.int32)
        vals = np.zeros(n * 27, np.float64)
        diag = np.zeros(n, np.int32)
        f2c = None
        ptr = None
        if level + 1 < self.levels:
            f2c = np.zeros(self.info[level + 1]["nrow"], np.int32)
            ptr = f2c.ctypes.data_as(C.c_void_p)
        ref(self.omp).ref_level_export(self.h, level, nnz_row, cols, vals, diag, ptr)
        return nnz_row, cols.reshape(n, 27), vals.reshape(n, 27), diag, f2c

    def rhs(self):
        b, x, xe = np.empty(self.n), np.empty(self.n), np.empty(self.n)
        ref(self.omp).ref_rhs(self.h, b, x, xe)
        return b, x, xe

    def set_diagonal(self, level, d):
        ref(self.omp).ref_set_diagonal(self.h, level, np.ascontiguousarray(d, np.float64))

    def replace_rows(self, level, nnz_row, cols, vals, f2c=None):
        """Overwrite a level's rows (same row count) and optionally its f2c."""
        f = None if f2c is None else np.ascontiguousarray(f2c, np.int32)
        rc = ref(self.omp).ref_replace_rows(
            self.h, level, np.ascontiguousarray(nnz_row, np.uint8),
            np.ascontiguousarray(cols, np.int32).reshape(-1), np.ascontiguousarray(vals, np.float64).reshape(-1),
            None if f is None else f.ctypes.data_as(C.c_void_p))
        if rc:
            raise ValueError(f"ref_replace_rows failed ({rc})")

    def spmv(self, x, level=0):
        y = np.empty(self.info[level]["nrow"])
        ref(self.omp).ref_spmv(self.h, level, np.array(x, np.float64), y)
        return y

    def symgs(self, r, x, level=0):
        x = np.array(x, np.float64, copy=True)
        ref(self.omp).ref_symgs(self.h, level, np.array(r, np.float64), x)
        return x

    def mg(self, r, level=0):
        x = np.zeros(self.info[level]["ncol"])
        ref(self.omp).ref_mg(self.h, level, np.array(r, np.float64), x)
        return x

    def cg(self, x0=None, max_iter=50, tol=0.0, precond=True, b=None):
        x = np.zeros(self.n) if x0 is None else np.array(x0, np.float64, copy=True)
        it, nr, nr0 = C.c_int(0), C.c_double(0), C.c_double(0)
        times = np.zeros(10)
        bb = None if b is None else np.ascontiguousarray(b, np.float64)
        bp = None if bb is None else bb.ctypes.data_as(C.c_void_p)
        ref(self.omp).ref_cg(self.h, bp, x, max_iter, tol, int(precond), C.byref(it), C.byref(nr),
                     C.byref(nr0), times.ctypes.data_as(C.c_void_p))
        return dict(x=x, niters=it.value, normr=nr.value, normr0=nr0.value, times=times)


def ref_dot(x, y):
    x = np.array(x, np.float64)
    y = np.array(y, np.float64)
    return ref().ref_dot(x.size, x, y)


def ref_waxpby(alpha, x, beta, y):
    x = np.array(x, np.float64)
    y = np.array(y, np.float64)
    w = np.empty_like(x)
    ref().ref_waxpby(x.size, alpha, x, beta, y, w)
    return w
