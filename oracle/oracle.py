"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the GPU library is compared with.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  Arrays are numpy, point-major (index i + m*(j + n*k)); operators
have shape (o, n, m, 27).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p
_i = ctypes.c_int
_u64 = ctypes.c_uint64


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        sigs = {
            "orc_hierarchy": (_i, [_i, _i, _i, _I]),
            "orc_gen_operator": (None, [_i, _u64, _i, _i, _i, _D]),
            "orc_gen_operator_planes": (None, [_i, _u64, _i, _i, _i, _i, _i, _D]),
            "orc_gen_rhs": (None, [_u64, _i, _i, _i, _D]),
            "orc_gs_lex": (None, [_i, _i, _i, _D, _D, _D, _i]),
            "orc_gs_colour8": (None, [_i, _i, _i, _D, _D, _D, _i]),
            "orc_gs_lex_f": (None, [_i, _i, _i, _F, _F, _F, _i]),
            "orc_gs_colour8_f": (None, [_i, _i, _i, _F, _F, _F, _i]),
            "orc_residual": (None, [_i, _i, _i, _D, _D, _D, _D]),
            "orc_l2_error_all": (ctypes.c_double, [_i, _i, _i, _D, _D, _D]),
            "orc_rhs_l2norm": (ctypes.c_double, [_i, _i, _i, _D]),
            "orc_restrict": (None, [_i, _i, _i, _D, _D, _D, _i, _i, _i, _D]),
            "orc_restrict_f": (None, [_i, _i, _i, _F, _F, _F, _i, _i, _i, _F]),
            "orc_interp": (None, [_i, _i, _i, _D, _i, _i, _i, _D]),
            "orc_interp_f": (None, [_i, _i, _i, _F, _i, _i, _i, _F]),
            "orc_rat": (None, [_i, _i, _i, _D, _i, _i, _i, _D]),
            "orc_lu_factor": (_i, [_i, _i, _i, _D, _D, _I]),
            "orc_lu_solve": (None, [_i, _D, _I, _D, _D]),
            "orc_mg_create": (_VP, [_i, _i, _i, _D]),
            "orc_mg_create_p": (_VP, [_i, _i, _i, _i, _D]),
            "orc_hierarchy_p": (_i, [_i, _i, _i, _i, _I]),
            "orc_gen_operator_p": (None, [_i, _u64, _i, _i, _i, _i, _D]),
            "orc_gs_lex_p": (None, [_i, _i, _i, _i, _D, _D, _D, _i]),
            "orc_gs_colour8_p": (None, [_i, _i, _i, _i, _D, _D, _D, _i]),
            "orc_gs_lex_p_f": (None, [_i, _i, _i, _i, _F, _F, _F, _i]),
            "orc_gs_colour8_p_f": (None, [_i, _i, _i, _i, _F, _F, _F, _i]),
            "orc_residual_p": (None, [_i, _i, _i, _i, _D, _D, _D, _D]),
            "orc_l2_error_all_p": (ctypes.c_double, [_i, _i, _i, _i, _D, _D, _D]),
            "orc_restrict_p": (None, [_i, _i, _i, _i, _D, _D, _D, _i, _i, _i, _D]),
            "orc_restrict_p_f": (None, [_i, _i, _i, _i, _F, _F, _F, _i, _i, _i, _F]),
            "orc_interp_p": (None, [_i, _i, _i, _i, _D, _i, _i, _i, _D]),
            "orc_interp_p_f": (None, [_i, _i, _i, _i, _F, _i, _i, _i, _F]),
            "orc_rat_p": (None, [_i, _i, _i, _i, _D, _i, _i, _i, _D]),
            "orc_lu_factor_p": (_i, [_i, _i, _i, _i, _D, _D, _I]),
            "orc_mg_destroy": (None, [_VP]),
            "orc_mg_levels": (_i, [_VP]),
            "orc_mg_set_smoother": (None, [_VP, _i]),
            "orc_mg_set_sweeps": (None, [_VP, _i, _i]),
            "orc_mg_enable_lu": (_i, [_VP]),
            "orc_mg_set_rhs": (None, [_VP, _D]),
            "orc_mg_set_x": (None, [_VP, _i, _D]),
            "orc_mg_get_x": (None, [_VP, _i, _D]),
            "orc_mg_get_r": (None, [_VP, _i, _D]),
            "orc_mg_get_A": (None, [_VP, _i, _D]),
            "orc_mg_gauss_seidel": (None, [_VP, _i, _i]),
            "orc_mg_restriction": (None, [_VP, _i]),
            "orc_mg_interpolation": (None, [_VP, _i]),
            "orc_mg_solve_exact": (None, [_VP, _i]),
            "orc_mg_v_cycle": (None, [_VP, _i]),
            "orc_mg_l2_error_all": (ctypes.c_double, [_VP, _i]),
            "orc_mg_rhs_l2norm": (ctypes.c_double, [_VP, _i]),
            "orc_mg_solve": (_i, [_VP, ctypes.c_double, _i, _D]),
            "orc_mg_solve_mixed": (_i, [_VP, ctypes.c_double, _i, _D]),
            "orc_mg_get_Af": (_i, [_VP, _i, _F]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    if a.dtype == np.float64:
        return a.ctypes.data_as(_D)
    if a.dtype == np.float32:
        return a.ctypes.data_as(_F)
    return a.ctypes.data_as(_I)


def hierarchy(m, n, o, prd=0):
    """Level sizes; prd holds the periodic axes (bit 0 x, bit 1 y, bit 2 z)."""
    d = np.zeros(48, dtype=np.int32)
    nl = lib().orc_hierarchy_p(m, n, o, prd, _p(d))
    return [tuple(int(v) for v in d[3 * l:3 * l + 3]) for l in range(nl)]


def gen_operator(family, seed, m, n, o, prd=0):
    A = np.zeros((o, n, m, 27), dtype=np.float64)
    if prd:
        lib().orc_gen_operator_p(family, seed, m, n, o, prd, _p(A))
    else:
        lib().orc_gen_operator(family, seed, m, n, o, _p(A))
    return A


def gen_operator_planes(family, seed, m, n, o, k0, k1):
    """Planes k0..k1-1 of gen_operator(family, seed, m, n, o)."""
    A = np.zeros((k1 - k0, n, m, 27), dtype=np.float64)
    lib().orc_gen_operator_planes(family, seed, m, n, o, k0, k1, _p(A))
    return A


def gen_rhs(seed, m, n, o):
    b = np.zeros((o, n, m), dtype=np.float64)
    lib().orc_gen_rhs(seed, m, n, o, _p(b))
    return b


def test_field(m, n, o, seed):
    """2u-1 with u = pgmg_u01(seed, 99, p): the field oracle/ref_driver.cc sets."""
    idx = np.arange(m * n * o, dtype=np.uint64)
    return (2.0 * _u01_np(seed, 99, idx) - 1.0).reshape(o, n, m)


_M64 = (1 << 64) - 1


def _mix64(z):
    z = (z + np.uint64(0x9e3779b97f4a7c15)) & np.uint64(_M64)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)) & np.uint64(_M64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)) & np.uint64(_M64)
    return z ^ (z >> np.uint64(31))


def _u01_np(seed, stream, idx):
    with np.errstate(over="ignore"):
        s = np.uint64(seed) ^ np.uint64((stream << 56) & _M64) ^ np.uint64((stream * 0x632be59bd9b4e019) & _M64)
        h0 = _mix64(np.array([s], dtype=np.uint64))[0]
        h = _mix64((np.uint64(h0) + idx.astype(np.uint64)) & np.uint64(_M64))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _arr(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def gs(A, x, r, mode=0, colour8=False, fp32=False, prd=0):
    o, n, m = x.shape
    dt = np.float32 if fp32 else np.float64
    A = _arr(A, dt); r = _arr(r, dt); x = _arr(x, dt).copy()
    if prd:
        f = {(False, False): lib().orc_gs_lex_p, (True, False): lib().orc_gs_colour8_p,
             (False, True): lib().orc_gs_lex_p_f, (True, True): lib().orc_gs_colour8_p_f}[(colour8, fp32)]
        f(m, n, o, prd, _p(A), _p(x), _p(r), mode)
        return x
    f = {(False, False): lib().orc_gs_lex, (True, False): lib().orc_gs_colour8,
         (False, True): lib().orc_gs_lex_f, (True, True): lib().orc_gs_colour8_f}[(colour8, fp32)]
    f(m, n, o, _p(A), _p(x), _p(r), mode)
    return x


def residual(A, x, r, prd=0):
    o, n, m = x.shape
    A = _arr(A, np.float64); x = _arr(x, np.float64); r = _arr(r, np.float64)
    out = np.zeros_like(x)
    if prd:
        lib().orc_residual_p(m, n, o, prd, _p(A), _p(x), _p(r), _p(out))
    else:
        lib().orc_residual(m, n, o, _p(A), _p(x), _p(r), _p(out))
    return out


def l2_error_all(A, x, r, prd=0):
    o, n, m = x.shape
    A = _arr(A, np.float64); x = _arr(x, np.float64); r = _arr(r, np.float64)
    if prd:
        return lib().orc_l2_error_all_p(m, n, o, prd, _p(A), _p(x), _p(r))
    return lib().orc_l2_error_all(m, n, o, _p(A), _p(x), _p(r))


def restrict(Af, xf, rf, coarse, fp32=False, prd=0):
    of, nf, mf = xf.shape
    mc, nc, oc = coarse
    dt = np.float32 if fp32 else np.float64
    Af = _arr(Af, dt); xf = _arr(xf, dt); rf = _arr(rf, dt)
    rc = np.zeros((oc, nc, mc), dtype=dt)
    if prd:
        (lib().orc_restrict_p_f if fp32 else lib().orc_restrict_p)(mf, nf, of, prd, _p(Af), _p(xf), _p(rf),
                                                                   mc, nc, oc, _p(rc))
    else:
        (lib().orc_restrict_f if fp32 else lib().orc_restrict)(mf, nf, of, _p(Af), _p(xf), _p(rf), mc, nc, oc, _p(rc))
    return rc


def interp(xc, xf, fp32=False, prd=0):
    oc, nc, mc = xc.shape
    of, nf, mf = xf.shape
    dt = np.float32 if fp32 else np.float64
    xc = _arr(xc, dt); out = _arr(xf, dt).copy()
    if prd:
        (lib().orc_interp_p_f if fp32 else lib().orc_interp_p)(mc, nc, oc, prd, _p(xc), mf, nf, of, _p(out))
    else:
        (lib().orc_interp_f if fp32 else lib().orc_interp)(mc, nc, oc, _p(xc), mf, nf, of, _p(out))
    return out


def rat(Af, coarse, prd=0):
    of, nf, mf = Af.shape[:3]
    mc, nc, oc = coarse
    Af = _arr(Af, np.float64)
    Ac = np.zeros((oc, nc, mc, 27), dtype=np.float64)
    if prd:
        lib().orc_rat_p(mf, nf, of, prd, _p(Af), mc, nc, oc, _p(Ac))
    else:
        lib().orc_rat(mf, nf, of, _p(Af), mc, nc, oc, _p(Ac))
    return Ac


def lu_factor(A, prd=0):
    o, n, m = A.shape[:3]
    N = m * n * o
    A = _arr(A, np.float64)
    LU = np.zeros((N, N), dtype=np.float64)  # column-major: LU[col, row]
    ipiv = np.zeros(N, dtype=np.int32)
    ok = lib().orc_lu_factor_p(m, n, o, prd, _p(A), _p(LU), _p(ipiv))
    return ok, LU, ipiv


def lu_solve(LU, ipiv, b):
    N = ipiv.shape[0]
    b = _arr(b, np.float64).reshape(-1)
    x = np.zeros(N, dtype=np.float64)
    lib().orc_lu_solve(N, _p(_arr(LU, np.float64)), _p(_arr(ipiv, np.int32)), _p(b), _p(x))
    return x


class Multigrid:
    """The oracle's hierarchy: mirrors multigrid<p_class> at one rank."""

    def __init__(self, A, colour8=False, lu=False, v_down=2, v_up=2, prd=0):
        self.dims = hierarchy(A.shape[2], A.shape[1], A.shape[0], prd)
        A = _arr(A, np.float64)
        self.h = lib().orc_mg_create_p(A.shape[2], A.shape[1], A.shape[0], prd, _p(A))
        lib().orc_mg_set_smoother(self.h, 1 if colour8 else 0)
        lib().orc_mg_set_sweeps(self.h, v_down, v_up)
        self.lu = bool(lib().orc_mg_enable_lu(self.h)) if lu else False

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mg_destroy(self.h)
            self.h = None

    @property
    def nlevels(self):
        return len(self.dims)

    def shape(self, l):
        m, n, o = self.dims[l]
        return (o, n, m)

    def set_rhs(self, b):
        lib().orc_mg_set_rhs(self.h, _p(_arr(b, np.float64)))

    def set_x(self, l, x):
        lib().orc_mg_set_x(self.h, l, _p(_arr(x, np.float64)))

    def get_x(self, l=0):
        x = np.zeros(self.shape(l))
        lib().orc_mg_get_x(self.h, l, _p(x))
        return x

    def get_r(self, l=0):
        r = np.zeros(self.shape(l))
        lib().orc_mg_get_r(self.h, l, _p(r))
        return r

    def get_A(self, l):
        A = np.zeros(self.shape(l) + (27,))
        lib().orc_mg_get_A(self.h, l, _p(A))
        return A

    def gauss_seidel(self, mode=0, level=0):
        lib().orc_mg_gauss_seidel(self.h, mode, level)

    def restriction(self, level):
        lib().orc_mg_restriction(self.h, level)

    def interpolation(self, level):
        lib().orc_mg_interpolation(self.h, level)

    def solve_exact(self, level):
        lib().orc_mg_solve_exact(self.h, level)

    def v_cycle(self, mode=0):
        lib().orc_mg_v_cycle(self.h, mode)

    def fmg_cycle(self, v_down=2, v_up=2, v_bottom=20):
        """multigrid::fmg_cycle (multigrid.cc:109-134) composed from the level
        operations: restrict to the bottom smoothing on the way, then for each
        j from the bottom up, down from level j again, v_bottom sweeps at the
        bottom and interpolate back to level j-1."""
        L = self.nlevels
        for i in range(L - 1):
            self.restriction(i + 1)
            for _ in range(v_down):
                self.gauss_seidel(0, i + 1)
        for j in range(L - 1, 0, -1):
            for i in range(j, L - 1):
                self.restriction(i + 1)
                for _ in range(v_down):
                    self.gauss_seidel(0, i + 1)
            for _ in range(v_bottom):
                self.gauss_seidel(0, L - 1)
            for i in range(L - 1, j - 1, -1):
                self.interpolation(i)
                for _ in range(v_up):
                    self.gauss_seidel(0, i - 1)

    def get_Af(self, l):
        """FP32 operator of level l made by the last mixed solve (None before one)."""
        A = np.zeros(self.shape(l) + (27,), dtype=np.float32)
        return A if lib().orc_mg_get_Af(self.h, l, _p(A)) else None

    def l2_error_all(self, level=0):
        return lib().orc_mg_l2_error_all(self.h, level)

    def rhs_l2norm(self, level=0):
        return lib().orc_mg_rhs_l2norm(self.h, level)

    def solve(self, tol, maxc=100, mixed=False):
        hist = np.zeros(maxc + 1)
        f = lib().orc_mg_solve_mixed if mixed else lib().orc_mg_solve
        c = f(self.h, tol, maxc, _p(hist))
        return c, hist[:c + 1].copy()
