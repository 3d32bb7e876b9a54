"""pgmg_amd - MI355X (gfx950) geometric multigrid, Python binding.

The product is the native library pgmg_amd/libpgmg_amd.so (hand-written HIP
kernels behind the C ABI of include/pgmg_capi.h).  This module is a thin
ctypes binding of that ABI for Python callers, tests and the benchmark; it
has no compute path of its own, and importing it fails loudly when the
library is not built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpgmg_amd.so")

PREC_FP64, PREC_MIXED = 0, 1
SMOOTH_EXACT, SMOOTH_COLOUR8 = 0, 1
BOTTOM_GS24, BOTTOM_LU = 0, 1
FAM_FACE7_UNIFORM, FAM_FACE7_LOGUNI, FAM_Q1_ANISO = 1, 2, 3
KC_GS, KC_RESIDUAL, KC_RESTRICT, KC_INTERP = 1, 2, 3, 4

ERR_NAMES = {1: "FILE", 2: "MEMORY", 3: "INTERNAL", 4: "MULTIGRID", 5: "MATH", 6: "SETUP", 7: "HIP",
             8: "ARG", 9: "UNSUPPORTED", 10: "ADMISSION", 11: "NOT_CONVERGED", 12: "NAN"}


class PgmgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pgmg error %d (%s): %s" % (code, ERR_NAMES.get(code, "?"), msg))
        self.code = code


class Dims(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("n", ctypes.c_int), ("o", ctypes.c_int),
                ("x_prd", ctypes.c_int), ("y_prd", ctypes.c_int), ("z_prd", ctypes.c_int)]


class Opts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("smoother", ctypes.c_int), ("bottom", ctypes.c_int),
                ("v_down", ctypes.c_int), ("v_up", ctypes.c_int), ("v_bottom", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("max_levels", ctypes.c_int)]


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p
_i = ctypes.c_int

_SIGS = {
    "pgmg_last_error": (ctypes.c_char_p, []),
    "pgmg_version": (_i, []),
    "pgmg_default_opts": (None, [ctypes.POINTER(Opts)]),
    "pgmg_hierarchy_shape": (_i, [ctypes.POINTER(Dims), _I, _I]),
    "pgmg_hier_footprint": (_i, [ctypes.POINTER(Dims), ctypes.POINTER(Opts), ctypes.POINTER(ctypes.c_uint64)]),
    "pgmg_hier_create": (_i, [ctypes.POINTER(Dims), ctypes.POINTER(Opts), ctypes.POINTER(_VP)]),
    "pgmg_hier_destroy": (_i, [_VP]),
    "pgmg_num_levels": (_i, [_VP, _I]),
    "pgmg_level_dims": (_i, [_VP, _i, _I]),
    "pgmg_set_operator27": (_i, [_VP, _D]),
    "pgmg_generate_problem": (_i, [_VP, _i, ctypes.c_uint64]),
    "pgmg_compute_rats": (_i, [_VP]),
    "pgmg_lu_active": (_i, [_VP, _I]),
    "pgmg_set_rhs": (_i, [_VP, _i, _D]),
    "pgmg_get_rhs": (_i, [_VP, _i, _D]),
    "pgmg_set_x": (_i, [_VP, _i, _D]),
    "pgmg_get_x": (_i, [_VP, _i, _D]),
    "pgmg_get_operator27": (_i, [_VP, _i, _D]),
    "pgmg_clear_x": (_i, [_VP, _i]),
    "pgmg_gauss_seidel": (_i, [_VP, _i, _i, _i]),
    "pgmg_restriction": (_i, [_VP, _i]),
    "pgmg_interpolation": (_i, [_VP, _i]),
    "pgmg_solve_exact": (_i, [_VP, _i, _i]),
    "pgmg_v_cycle": (_i, [_VP, _i]),
    "pgmg_set_sweeps": (_i, [_VP, _i, _i]),
    "pgmg_l2_error_all": (_i, [_VP, _i, _D]),
    "pgmg_rhs_l2norm": (_i, [_VP, _i, _D]),
    "pgmg_solve": (_i, [_VP, ctypes.c_double, _i, _I, _D]),
    "pgmg_residual": (_i, [_VP, _i, _D]),
    "pgmg_apply_operator": (_i, [_VP, _i, _D, _D]),
    "pgmg_copy_rhs_to_x": (_i, [_VP, _i]),
    "pgmg_rhs_x_dot": (_i, [_VP, _i, _D]),
    "pgmg_x_diff2": (_i, [_VP, _i, _D, _D]),
    "pgmg_x_diff2_noalloc": (_i, [_VP, _i, _D, _D]),
    "pgmg_synchronize": (_i, [_VP]),
    "pgmg_set_wait_limit": (_i, [_VP, _i]),
    "pgmg_gs_count": (_i, [_VP, ctypes.POINTER(ctypes.c_longlong)]),
    "pgmg_prof_enable": (_i, [_VP, _i, _i]),
    "pgmg_prof_read": (_i, [_VP, ctypes.POINTER(ctypes.c_longlong), _D, ctypes.POINTER(ctypes.c_longlong)]),
}

_LIB = None


def lib():
    """The loaded native library (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("pgmg_amd: native library %s is not built; run "
                               "`python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise PgmgError(rc, lib().pgmg_last_error().decode(errors="replace"))


def _dptr(a):
    return a.ctypes.data_as(_D)


def default_opts(**kw):
    o = Opts()
    lib().pgmg_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, int(v))
    return o


def _dims(m, n, o, prd):
    """prd: the periodic axes as a bit mask (1 = x, 2 = y, 4 = z)."""
    return Dims(m, n, o, prd & 1, (prd >> 1) & 1, (prd >> 2) & 1)


def hierarchy_shape(m, n, o, prd=0):
    d = _dims(m, n, o, prd)
    nl = ctypes.c_int()
    out = (ctypes.c_int * 48)()
    _check(lib().pgmg_hierarchy_shape(ctypes.byref(d), ctypes.byref(nl), out))
    return [(out[3 * l], out[3 * l + 1], out[3 * l + 2]) for l in range(nl.value)]


def footprint(m, n, o, prd=0, **kw):
    d = _dims(m, n, o, prd)
    b = ctypes.c_uint64()
    _check(lib().pgmg_hier_footprint(ctypes.byref(d), ctypes.byref(default_opts(**kw)), ctypes.byref(b)))
    return b.value


class Hierarchy:
    """A multigrid hierarchy resident on the GPU (one pgmg_hier handle).

    Method names follow the reference's multigrid class (multigrid.hh)."""

    def __init__(self, m, n, o, precision=PREC_FP64, smoother=SMOOTH_COLOUR8, bottom=BOTTOM_GS24,
                 v_down=2, v_up=2, use_graph=True, max_levels=0, prd=0):
        self.opts = default_opts(precision=precision, smoother=smoother, bottom=bottom,
                                 v_down=v_down, v_up=v_up, use_graph=use_graph, max_levels=max_levels)
        self.prd = prd
        d = _dims(m, n, o, prd)
        h = _VP()
        _check(lib().pgmg_hier_create(ctypes.byref(d), ctypes.byref(self.opts), ctypes.byref(h)))
        self.h = h
        self.dims = hierarchy_shape(m, n, o, prd)
        if max_levels > 0:
            self.dims = self.dims[:max_levels]

    def close(self):
        if getattr(self, "h", None):
            lib().pgmg_hier_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def nlevels(self):
        return len(self.dims)

    def shape(self, level):
        m, n, o = self.dims[level]
        return (o, n, m)

    def set_operator27(self, A):
        A = np.ascontiguousarray(A, dtype=np.float64)
        assert A.size == 27 * int(np.prod(self.shape(0)))
        _check(lib().pgmg_set_operator27(self.h, _dptr(A)))

    def generate_problem(self, family, seed):
        _check(lib().pgmg_generate_problem(self.h, family, seed))

    def compute_rats(self):
        _check(lib().pgmg_compute_rats(self.h))

    def lu_active(self):
        a = ctypes.c_int()
        _check(lib().pgmg_lu_active(self.h, ctypes.byref(a)))
        return bool(a.value)

    def set_rhs(self, b, level=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert b.size == int(np.prod(self.shape(level)))
        _check(lib().pgmg_set_rhs(self.h, level, _dptr(b)))

    def get_rhs(self, level=0):
        b = np.zeros(self.shape(level))
        _check(lib().pgmg_get_rhs(self.h, level, _dptr(b)))
        return b

    def set_x(self, x, level=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == int(np.prod(self.shape(level)))
        _check(lib().pgmg_set_x(self.h, level, _dptr(x)))

    def get_x(self, level=0):
        x = np.zeros(self.shape(level))
        _check(lib().pgmg_get_x(self.h, level, _dptr(x)))
        return x

    def get_operator27(self, level):
        A = np.zeros(self.shape(level) + (27,))
        _check(lib().pgmg_get_operator27(self.h, level, _dptr(A)))
        return A

    def clear_x(self, level=0):
        _check(lib().pgmg_clear_x(self.h, level))

    def gauss_seidel(self, mode=0, level=0, nsweeps=1):
        _check(lib().pgmg_gauss_seidel(self.h, mode, level, nsweeps))

    def restriction(self, level):
        _check(lib().pgmg_restriction(self.h, level))

    def interpolation(self, level):
        _check(lib().pgmg_interpolation(self.h, level))

    def solve_exact(self, level, smooth=False):
        _check(lib().pgmg_solve_exact(self.h, level, int(smooth)))

    def v_cycle(self, mode=0):
        _check(lib().pgmg_v_cycle(self.h, mode))

    def set_sweeps(self, v_down, v_up):
        """The public v_down / v_up of multigrid.hh:122-128, changed."""
        _check(lib().pgmg_set_sweeps(self.h, v_down, v_up))

    def l2_error_all(self, level=0):
        v = ctypes.c_double()
        _check(lib().pgmg_l2_error_all(self.h, level, ctypes.byref(v)))
        return v.value

    def rhs_l2norm(self, level=0):
        v = ctypes.c_double()
        _check(lib().pgmg_rhs_l2norm(self.h, level, ctypes.byref(v)))
        return v.value

    def solve(self, tol, max_cycles=100, check=True):
        """Returns (cycles, history of l2_error_all).  With check=False a
        solve that used up max_cycles returns normally."""
        hist = np.zeros(max_cycles + 1)
        c = ctypes.c_int()
        rc = lib().pgmg_solve(self.h, tol, max_cycles, ctypes.byref(c), _dptr(hist))
        if not (rc == 11 and not check):
            _check(rc)
        return c.value, hist[:c.value + 1].copy()

    def residual(self, level=0):
        out = np.zeros(self.shape(level))
        _check(lib().pgmg_residual(self.h, level, _dptr(out)))
        return out

    # operations on outside arrays (multigrid.hh:222-251)
    def apply_operator(self, v, level=0):
        """A v at a level (region::Ax)."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.size == int(np.prod(self.shape(level)))
        out = np.zeros(self.shape(level))
        _check(lib().pgmg_apply_operator(self.h, level, _dptr(v), _dptr(out)))
        return out

    def copy_rhs_to_x(self, level=0):
        _check(lib().pgmg_copy_rhs_to_x(self.h, level))

    def rhs_x_dot(self, level=0):
        v = ctypes.c_double()
        _check(lib().pgmg_rhs_x_dot(self.h, level, ctypes.byref(v)))
        return v.value

    def x_diff2(self, y, level=0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        assert y.size == int(np.prod(self.shape(level)))
        v = ctypes.c_double()
        _check(lib().pgmg_x_diff2(self.h, level, _dptr(y), ctypes.byref(v)))
        return v.value

    def x_diff2_noalloc(self, y, level=0):
        """x_diff2 through the allocation-free entry point (levels of up to
        about 8.3 M points)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        assert y.size == int(np.prod(self.shape(level)))
        v = ctypes.c_double()
        _check(lib().pgmg_x_diff2_noalloc(self.h, level, _dptr(y), ctypes.byref(v)))
        return v.value

    def synchronize(self):
        _check(lib().pgmg_synchronize(self.h))

    def set_wait_limit(self, polls):
        """Polls after which an exact sweep's wait between planes gives up."""
        _check(lib().pgmg_set_wait_limit(self.h, polls))

    def gs_count(self):
        v = ctypes.c_longlong()
        _check(lib().pgmg_gs_count(self.h, ctypes.byref(v)))
        return v.value

    def prof_enable(self, kernel_class, level=0):
        _check(lib().pgmg_prof_enable(self.h, kernel_class, level))

    def prof_read(self):
        """(launches, total milliseconds, grid points updated) since prof_enable."""
        n = ctypes.c_longlong()
        ms = ctypes.c_double()
        pts = ctypes.c_longlong()
        _check(lib().pgmg_prof_read(self.h, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(pts)))
        return n.value, ms.value, pts.value


# ---------------------------------------------------------------- batches

class Request(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("m", ctypes.c_int), ("n", ctypes.c_int), ("o", ctypes.c_int),
                ("family", ctypes.c_int), ("seed", ctypes.c_uint64), ("tol", ctypes.c_double),
                ("max_cycles", ctypes.c_int), ("x", _D)]


class Result(ctypes.Structure):
    _fields_ = [("admit_s", ctypes.c_double), ("done_s", ctypes.c_double), ("cycles", ctypes.c_int),
                ("status", ctypes.c_int), ("rel_residual", ctypes.c_double)]


_BATCH_SIGS = {
    "pgmg_batch_create": (_i, [ctypes.POINTER(Opts), ctypes.c_uint64, _i, ctypes.POINTER(_VP)]),
    "pgmg_batch_destroy": (_i, [_VP]),
    "pgmg_batch_admit": (_i, [_VP, ctypes.POINTER(Dims), _i, ctypes.c_uint64, ctypes.c_double, _i, _I]),
    "pgmg_batch_admit_operator": (_i, [_VP, ctypes.POINTER(Dims), _D, _D, ctypes.c_double, _i, _I]),
    "pgmg_batch_step": (_i, [_VP, _I, _I]),
    "pgmg_batch_slot_status": (_i, [_VP, _i, _I, _I, _D]),
    "pgmg_batch_get_x": (_i, [_VP, _i, _D]),
    "pgmg_batch_release": (_i, [_VP, _i]),
    "pgmg_batch_free_bytes": (_i, [_VP, ctypes.POINTER(ctypes.c_uint64)]),
    "pgmg_batch_tiles": (_i, [_VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_longlong)]),
    "pgmg_serve_run": (_i, [_VP, ctypes.POINTER(Request), _i, ctypes.POINTER(Result), _D]),
}
_SIGS.update(_BATCH_SIGS)

SLOT_FREE, SLOT_ACTIVE, SLOT_CONVERGED, SLOT_FAILED = 0, 1, 2, 3


class Batch:
    """Independent hierarchies of any sizes advanced together (pgmg_batch_*)."""

    TILE_CLASSES = ("gs_pair", "residual", "residual_top", "restrict", "interp", "add_x")

    def __init__(self, arena_bytes, max_slots=256, precision=PREC_MIXED, bottom=BOTTOM_GS24, v_down=2, v_up=2,
                 use_graph=True):
        self.opts = default_opts(precision=precision, smoother=SMOOTH_COLOUR8, bottom=bottom, v_down=v_down, v_up=v_up,
                                 use_graph=use_graph)
        h = _VP()
        _check(lib().pgmg_batch_create(ctypes.byref(self.opts), int(arena_bytes), max_slots, ctypes.byref(h)))
        self.h = h
        self.shapes = {}

    def close(self):
        if getattr(self, "h", None):
            lib().pgmg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def admit(self, m, n, o, family, seed, tol, max_cycles=100):
        """Returns the slot, or None when the batch has no room."""
        s = ctypes.c_int(-1)
        rc = lib().pgmg_batch_admit(self.h, ctypes.byref(Dims(m, n, o, 0, 0, 0)), family, seed, tol, max_cycles,
                                    ctypes.byref(s))
        if rc == 10:
            return None
        _check(rc)
        self.shapes[s.value] = (o, n, m)
        return s.value

    def admit_operator(self, A, b, tol, max_cycles=100):
        o, n, m = b.shape
        A = np.ascontiguousarray(A, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        s = ctypes.c_int(-1)
        rc = lib().pgmg_batch_admit_operator(self.h, ctypes.byref(Dims(m, n, o, 0, 0, 0)), _dptr(A), _dptr(b), tol,
                                             max_cycles, ctypes.byref(s))
        if rc == 10:
            return None
        _check(rc)
        self.shapes[s.value] = (o, n, m)
        return s.value

    def step(self):
        a, f = ctypes.c_int(), ctypes.c_int()
        _check(lib().pgmg_batch_step(self.h, ctypes.byref(a), ctypes.byref(f)))
        return a.value, f.value

    def run(self, max_steps=10000):
        """Steps until no slot is active; returns the number of steps."""
        steps = 0
        while steps < max_steps:
            a, f = self.step()
            steps += 1
            if a == 0 and not self._setting_up():
                break
        return steps

    def _setting_up(self):
        return any(self.status(s)[0] == SLOT_ACTIVE for s in self.shapes)

    def status(self, slot):
        st, c, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().pgmg_batch_slot_status(self.h, slot, ctypes.byref(st), ctypes.byref(c), ctypes.byref(r)))
        return st.value, c.value, r.value

    def get_x(self, slot):
        x = np.zeros(self.shapes[slot])
        _check(lib().pgmg_batch_get_x(self.h, slot, _dptr(x)))
        return x

    def release(self, slot):
        _check(lib().pgmg_batch_release(self.h, slot))
        self.shapes.pop(slot, None)

    def free_bytes(self):
        v = ctypes.c_uint64()
        _check(lib().pgmg_batch_free_bytes(self.h, ctypes.byref(v)))
        return v.value

    def tiles(self):
        """(work units per workgroup by launch class or None before the
        trial steps are done, number of graph replays of the cycle)."""
        c = (ctypes.c_int * len(self.TILE_CLASSES))()
        t, g = ctypes.c_int(), ctypes.c_longlong()
        _check(lib().pgmg_batch_tiles(self.h, c, ctypes.byref(t), ctypes.byref(g)))
        return (dict(zip(self.TILE_CLASSES, list(c))) if t.value else None), g.value

    def serve(self, requests, keep_x=True):
        """requests: list of dicts (arrival_s, m, n, o, family, seed, tol,
        max_cycles).  Returns (results as a list of dicts, wall seconds); with
        keep_x each result carries its solution as "x" (shape (o, n, m))."""
        n = len(requests)
        R = (Request * n)()
        xs = []
        for q, r in enumerate(requests):
            x = np.zeros((r["o"], r["n"], r["m"])) if keep_x else None
            xs.append(x)
            R[q] = Request(r["arrival_s"], r["m"], r["n"], r["o"], r["family"], r["seed"], r["tol"],
                           r.get("max_cycles", 100), _dptr(x) if keep_x else None)
        out = (Result * n)()
        wall = ctypes.c_double()
        _check(lib().pgmg_serve_run(self.h, R, n, out, ctypes.byref(wall)))
        res = [dict(admit_s=o_.admit_s, done_s=o_.done_s, cycles=o_.cycles, status=o_.status,
                    rel_residual=o_.rel_residual, x=x) for o_, x in zip(out, xs)]
        return res, wall.value
