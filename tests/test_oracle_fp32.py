"""Pins of the oracle's FP32 level functions and of its mixed-precision
defect correction.  CPU only.

The FP32 functions are the same source text as the FP64 ones, instantiated
with float (oracle/pgmg_oracle_level.inc), and the FP64 instantiation is
pinned bit for bit to the reference's own runs.  What the FP32 instantiation
adds is float arithmetic, so it is checked here against restatements written
independently of the oracle's C code, in numpy float32 scalars (every +, *
and / rounds to single precision, no contraction):

- Gauss-Seidel (lexicographic, colour-8, both modes, periodic axes): a
  point-by-point sweep in the order the reference visits points
  (region.cc:367-446 forward, 463-537 backward).
- Restriction: the weights come from the FP64 oracle's unit-impulse
  responses (pinned to the reference, even-grid quirk included); the
  residual and the accumulation in cube-scan order (region.cc:665-746) are
  restated in float32.
- Interpolation: the association orders of interpolate_full and
  interpolate_partial (region.cc:752-810), restated per fine point; and, on
  integer-valued inputs where every operation is exact in both precisions,
  bit-equality with the FP64 oracle.

Each restatement is also run with double accumulation and shown to differ
from the oracle, so a float function that silently computes in double would
fail these tests.

The defect correction (orc_mg_solve_mixed) is pinned by composing the same
level functions in Python after the V(v_down+L, v_up+L) schedule of
multigrid.cc:88-106; the composition, run on the FP64 level functions,
reproduces the FP64 oracle's V-cycle (pinned to the reference) bit for bit.
The mixed solution is checked against the exact discrete solution with the
bound |x - x*| <= |b - A x| / sigma_min(A)."""
import numpy as np
import pytest

from oracle import oracle as O

F32 = np.float32


def nb_index(i, j, k, m, n, o, prd):
    """Wrapped neighbour (or None outside a non-periodic axis) and whether it
    was reached across a periodic boundary."""
    g = False
    if not 0 <= i < m:
        if not prd & 1:
            return None, False
        i, g = i % m, True
    if not 0 <= j < n:
        if not prd & 2:
            return None, False
        j, g = j % n, True
    if not 0 <= k < o:
        if not prd & 4:
            return None, False
        k, g = k % o, True
    return (k, j, i), g


def py_sweep(A, x, r, order, dt, prd=0, acc64=False):
    """Gauss-Seidel over the points in `order`: the off-centre box sum from
    zero in box order, then (r - sum) * (1 / a_centre).  Neighbours across a
    periodic boundary are read from the copy taken at the start (the ghost
    strips of region.cc:1338-1365).  acc64: accumulate in double instead (a
    planted mistake)."""
    o, n, m = x.shape
    x = x.astype(dt).copy()
    ghost = x.copy()
    adt = np.float64 if acc64 else dt
    for (i, j, k) in order:
        ax = adt(0)
        for e in range(27):
            if e == 13:
                continue
            q, g = nb_index(i + e % 3 - 1, j + (e // 3) % 3 - 1, k + e // 9 - 1, m, n, o, prd)
            if q is None:
                continue
            v = ghost[q] if g else x[q]
            ax = adt(ax + adt(A[k, j, i, e]) * adt(v))
        x[k, j, i] = dt((adt(r[k, j, i]) - ax) * (adt(1) / adt(A[k, j, i, 13])))
    return x


def lex(m, n, o):
    return [(i, j, k) for k in range(o) for j in range(n) for i in range(m)]


def colour_order(m, n, o, reverse=False):
    cols = range(7, -1, -1) if reverse else range(8)
    return [(i, j, k) for c in cols for k in range(o) for j in range(n) for i in range(m)
            if (i & 1) + 2 * (j & 1) + 4 * (k & 1) == c]


def rand_problem(m, n, o, seed, prd=0):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 0, size=(o, n, m, 27))
    A[..., 13] = 30.0 + rng.uniform(0, 1, size=(o, n, m))
    for k in range(o):
        for j in range(n):
            for i in range(m):
                for e in range(27):
                    if nb_index(i + e % 3 - 1, j + (e // 3) % 3 - 1, k + e // 9 - 1, m, n, o, prd)[0] is None:
                        A[k, j, i, e] = 0.0
    x = rng.uniform(-1, 1, size=(o, n, m))
    r = rng.uniform(-1, 1, size=(o, n, m))
    return A.astype(F32), x.astype(F32), r.astype(F32)


GRIDS = [(5, 4, 3), (6, 5, 4), (7, 6, 5), (4, 7, 6)]


@pytest.mark.parametrize("dims", GRIDS)
@pytest.mark.parametrize("colour8", [False, True])
def test_f32_gauss_seidel_is_float_arithmetic(dims, colour8):
    m, n, o = dims
    A, x0, r = rand_problem(m, n, o, 100 + m)
    fwd = colour_order(m, n, o) if colour8 else lex(m, n, o)
    bwd = colour_order(m, n, o, True) if colour8 else lex(m, n, o)[::-1]
    x1 = py_sweep(A, x0, r, fwd, F32)
    got = O.gs(A, x0, r, 0, colour8=colour8, fp32=True)
    assert got.dtype == F32 and np.array_equal(got, x1)
    assert np.array_equal(O.gs(A, x0, r, 1, colour8=colour8, fp32=True), py_sweep(A, x1, r, bwd, F32))
    # the planted mistake (double accumulation) is seen
    assert not np.array_equal(py_sweep(A, x0, r, fwd, F32, acc64=True), got)


@pytest.mark.parametrize("dims,prd", [((6, 5, 4), 1), ((5, 6, 4), 2), ((6, 4, 6), 4), ((4, 4, 4), 7), ((5, 6, 7), 5)])
@pytest.mark.parametrize("colour8", [False, True])
def test_f32_gauss_seidel_periodic_is_float_arithmetic(dims, prd, colour8):
    m, n, o = dims
    A, x0, r = rand_problem(m, n, o, 200 + prd, prd)
    fwd = colour_order(m, n, o) if colour8 else lex(m, n, o)
    bwd = colour_order(m, n, o, True) if colour8 else lex(m, n, o)[::-1]
    x1 = py_sweep(A, x0, r, fwd, F32, prd)
    got = O.gs(A, x0, r, 0, colour8=colour8, fp32=True, prd=prd)
    assert np.array_equal(got, x1)
    # the reverse pass takes fresh ghosts (region.cc:457)
    assert np.array_equal(O.gs(A, x0, r, 1, colour8=colour8, fp32=True, prd=prd), py_sweep(A, x1, r, bwd, F32, prd))


def py_residual(A, x, r, dt, prd=0, acc64=False):
    """r - (box sum with the centre, in box order from zero)
    (region.cc:1833-1842)."""
    o, n, m = x.shape
    adt = np.float64 if acc64 else dt
    out = np.zeros((o, n, m), dtype=dt)
    for k in range(o):
        for j in range(n):
            for i in range(m):
                ax = adt(0)
                for e in range(27):
                    q, _ = nb_index(i + e % 3 - 1, j + (e // 3) % 3 - 1, k + e // 9 - 1, m, n, o, prd)
                    if q is not None:
                        ax = adt(ax + adt(A[k, j, i, e]) * adt(x[q]))
                out[k, j, i] = dt(adt(r[k, j, i]) - ax)
    return out


def restriction_weights(fine, coarse, prd):
    """W[c, f]: the weight of fine residual f in coarse point c, read from the
    FP64 oracle's response to unit residuals (x = 0, A = 0)."""
    m, n, o = fine
    N = m * n * o
    W = np.zeros((coarse[0] * coarse[1] * coarse[2], N))
    Z = np.zeros((o, n, m, 27))
    for f in range(N):
        e = np.zeros(N)
        e[f] = 1.0
        W[:, f] = O.restrict(Z, np.zeros((o, n, m)), e.reshape(o, n, m), coarse, prd=prd).reshape(-1)
    return W


def py_restrict(A, x, r, fine, coarse, prd, dt, acc64=False):
    """Coarse r: each fine residual times its weight, added to the coarse
    points in the scan order of region.cc:665-746 (cubes k, j, i; inside a
    cube local k, j, i)."""
    m, n, o = fine
    res = py_residual(A, x, r, dt, prd, acc64).reshape(-1)
    W = restriction_weights(fine, coarse, prd)

    def key(f):
        i, j, k = f % m, (f // m) % n, f // (m * n)
        return (k >> 1, j >> 1, i >> 1, k & 1, j & 1, i & 1)
    scan = sorted(range(m * n * o), key=key)
    rc = np.zeros(W.shape[0], dtype=dt)
    for f in scan:
        for c in np.nonzero(W[:, f])[0]:
            w = dt(W[c, f])
            assert float(w) == W[c, f]  # powers of two
            rc[c] = dt(rc[c] + w * res[f])
    return rc.reshape(coarse[2], coarse[1], coarse[0])


@pytest.mark.parametrize("fine,prd", [((7, 5, 9), 0), ((8, 6, 5), 0), ((6, 8, 8), 0), ((9, 7, 6), 0),
                                      ((8, 6, 5), 1), ((7, 8, 6), 2), ((6, 6, 8), 7), ((9, 6, 7), 5)])
def test_f32_restriction_is_float_arithmetic(fine, prd):
    m, n, o = fine
    coarse = O.hierarchy(m, n, o, prd)[1] if len(O.hierarchy(m, n, o, prd)) > 1 else \
        tuple((d + (1 if (prd >> a) & 1 else 2)) >> 1 for a, d in enumerate(fine))
    A, x, r = rand_problem(m, n, o, 300 + m + prd, prd)
    got = O.restrict(A, x, r, coarse, fp32=True, prd=prd)
    assert got.dtype == F32
    assert np.array_equal(got, py_restrict(A, x, r, fine, coarse, prd, F32))
    assert not np.array_equal(got, py_restrict(A, x, r, fine, coarse, prd, F32, acc64=True))


def py_interp_point(li, lj, lk, full, xs, ys, zs, v, dt):
    """Increment of fine point (local li, lj, lk) of a cube from the cube's
    eight coarse values v[0..7] (v0 + v1 (x) + v2 (y) + v4 (z) ...), in the
    association orders of interpolate_full (region.cc:764-771) and
    interpolate_partial (region.cc:795-809)."""
    h, q, e = dt(0.5), dt(0.25), dt(0.125)
    v0, v1, v2, v3, v4, v5, v6, v7 = v
    if full:
        t1 = dt(v0 + v1)
        t2 = dt(dt(t1 + v2) + v3)
        t3 = dt(v4 + v5)
        return {(0, 0, 0): v0,
                (1, 0, 0): dt(h * t1),
                (0, 1, 0): dt(h * dt(v0 + v2)),
                (1, 1, 0): dt(q * t2),
                (0, 0, 1): dt(h * dt(v0 + v4)),
                (1, 0, 1): dt(q * dt(t1 + t3)),
                (0, 1, 1): dt(q * dt(dt(dt(v0 + v2) + v4) + v6)),
                (1, 1, 1): dt(e * dt(dt(dt(t2 + t3) + v6) + v7))}[(li, lj, lk)]

    def half(a, b):
        return dt(h * dt(a + b))

    def quarter(a, b, c, d):
        return dt(q * dt(dt(dt(a + b) + c) + d))
    if (li, lj, lk) == (0, 0, 0):
        return v0
    if (li, lj, lk) == (1, 0, 0):
        return v1 if xs else half(v0, v1)
    if (li, lj, lk) == (0, 1, 0):
        return v2 if ys else half(v0, v2)
    if (li, lj, lk) == (0, 0, 1):
        return v4 if zs else half(v0, v4)
    if (li, lj, lk) == (1, 0, 1):
        return (v5 if xs else half(v4, v5)) if zs else (half(v1, v5) if xs else quarter(v0, v1, v4, v5))
    if (li, lj, lk) == (0, 1, 1):
        return (v6 if ys else half(v4, v6)) if zs else (half(v2, v6) if ys else quarter(v0, v2, v4, v6))
    t2 = (v3 if xs else half(v2, v3)) if ys else (half(v1, v3) if xs else quarter(v0, v1, v2, v3))
    if lk == 0:
        return t2
    t3 = (v7 if xs else half(v6, v7)) if ys else (half(v5, v7) if xs else quarter(v4, v5, v6, v7))
    return t3 if zs else dt(h * dt(t3 + t2))


def py_interp(xc, xf, prd, dt):
    """x_f += P x_c, fine point by fine point.  A cube is the last of an axis
    at c = (M-1)>>1; an even last cube of a non-periodic axis injects its
    upper point (the xs/ys/zs flags); the upper coarse neighbour of the last
    cube is clamped on a non-periodic axis and wraps on a periodic one."""
    oc, nc, mc = xc.shape
    of, nf, mf = xf.shape
    out = xf.astype(dt).copy()
    last = [(mf - 1) >> 1, (nf - 1) >> 1, (of - 1) >> 1]
    Mf, Mc = [mf, nf, of], [mc, nc, oc]
    for fk in range(of):
        for fj in range(nf):
            for fi in range(mf):
                c = [fi >> 1, fj >> 1, fk >> 1]
                loc = [fi & 1, fj & 1, fk & 1]
                up = []
                flags = []
                for a in range(3):
                    p = (prd >> a) & 1
                    up.append((c[a] + 1) % Mc[a] if p else min(c[a] + 1, Mc[a] - 1))
                    flags.append((not p) and c[a] == last[a] and Mf[a] % 2 == 0)
                full = all(c[a] < last[a] for a in range(3))
                v = []
                for dz in range(2):
                    for dy in range(2):
                        for dx in range(2):
                            v.append(dt(xc[up[2] if dz else c[2], up[1] if dy else c[1], up[0] if dx else c[0]]))
                inc = py_interp_point(loc[0], loc[1], loc[2], full, flags[0], flags[1], flags[2], v, dt)
                out[fk, fj, fi] = dt(out[fk, fj, fi] + inc)
    return out


@pytest.mark.parametrize("fine,prd", [((9, 7, 5), 0), ((8, 6, 10), 0), ((6, 9, 8), 0), ((10, 10, 10), 0),
                                      ((8, 6, 5), 1), ((9, 8, 6), 2), ((6, 6, 8), 7), ((9, 6, 7), 5)])
def test_f32_interpolation_is_float_arithmetic(fine, prd):
    m, n, o = fine
    mc, nc, oc = [(d + (1 if (prd >> a) & 1 else 2)) >> 1 for a, d in enumerate(fine)]
    rng = np.random.default_rng(400 + m + prd)
    xc = rng.uniform(-1, 1, size=(oc, nc, mc)).astype(F32)
    xf = rng.uniform(-1, 1, size=(o, n, m)).astype(F32)
    got = O.interp(xc, xf, fp32=True, prd=prd)
    assert got.dtype == F32
    assert np.array_equal(got, py_interp(xc, xf, prd, F32))
    # the same restatement in double matches the FP64 oracle (pinned to the
    # reference's interpolation bit for bit)
    assert np.array_equal(O.interp(xc.astype(np.float64), xf.astype(np.float64), prd=prd),
                          py_interp(xc.astype(np.float64), xf.astype(np.float64), prd, np.float64))
    # double arithmetic rounded once at the end differs from the float one
    assert not np.array_equal(got, O.interp(xc.astype(np.float64), xf.astype(np.float64), prd=prd).astype(F32))
    # integer inputs: every operation is exact, both precisions agree
    xci = rng.integers(-512, 513, size=(oc, nc, mc)).astype(F32)
    xfi = rng.integers(-512, 513, size=(o, n, m)).astype(F32)
    assert np.array_equal(O.interp(xci, xfi, fp32=True, prd=prd).astype(np.float64),
                          O.interp(xci.astype(np.float64), xfi.astype(np.float64), prd=prd))


# ---------------------------------------------------------------- mixed

def composed_cycle(As, xs, rs, dims, prd, v_down, v_up, colour8, lu, fp32):
    """One V(v_down+L, v_up+L) cycle (multigrid.cc:88-106) composed from the
    oracle's level functions; xs/rs are per-level lists, updated in place.
    Bottom: dense LU in FP64 if lu = (LU, ipiv), else 24 sweeps."""
    L = len(dims)
    dt = F32 if fp32 else np.float64

    def gs(l):
        xs[l] = O.gs(As[l], xs[l], rs[l], 0, colour8=colour8, fp32=fp32, prd=prd)
    for i in range(L - 1):
        for _ in range(v_down + i):
            gs(i)
        rs[i + 1] = O.restrict(As[i], xs[i], rs[i], dims[i + 1], fp32=fp32, prd=prd)
        xs[i + 1] = np.zeros_like(xs[i + 1])
    b = L - 1
    if lu is not None:
        xb = O.lu_solve(lu[0], lu[1], rs[b].astype(np.float64).reshape(-1))
        xs[b] = xb.reshape(rs[b].shape).astype(dt)
    else:
        for _ in range(24):
            gs(b)
    for i in range(L - 1, 0, -1):
        xs[i - 1] = O.interp(xs[i], xs[i - 1], fp32=fp32, prd=prd)
        for _ in range(v_up + i - 1):
            gs(i - 1)


CASES = [  # (family, seed, dims, prd, colour8, lu, v_down, v_up)
    (1, 31, (20, 17, 30), 0, False, False, 2, 2),
    (2, 32, (18, 13, 22), 0, True, True, 2, 2),
    (3, 33, (17, 17, 17), 0, False, True, 1, 3),
    (1, 34, (16, 16, 16), 7, True, False, 2, 2),
    (2, 35, (13, 10, 11), 5, False, False, 0, 2),
]


def build(case):
    fam, seed, (m, n, o), prd, colour8, lu, vd, vu = case
    A = O.gen_operator(fam, seed, m, n, o, prd)
    b = O.gen_rhs(seed, m, n, o)
    mg = O.Multigrid(A, colour8=colour8, lu=lu, v_down=vd, v_up=vu, prd=prd)
    return A, b, mg


@pytest.mark.parametrize("case", CASES)
def test_composition_reproduces_fp64_v_cycle(case):
    """The Python composition, on the FP64 level functions, is the FP64
    oracle's V-cycle (itself pinned to the reference's iterates)."""
    fam, seed, dims, prd, colour8, lu, vd, vu = case
    A, b, mg = build(case)
    mg.set_rhs(b)
    L = mg.nlevels
    As = [mg.get_A(l) for l in range(L)]
    xs = [np.zeros(mg.shape(l)) for l in range(L)]
    rs = [b] + [np.zeros(mg.shape(l)) for l in range(1, L)]
    fac = None
    if mg.lu:
        ok, LU, ipiv = O.lu_factor(As[-1], prd)
        assert ok
        fac = (LU, ipiv)
    for _ in range(2):
        mg.v_cycle()
        composed_cycle(As, xs, rs, mg.dims, prd, vd, vu, colour8, fac, False)
        assert np.array_equal(mg.get_x(0), xs[0])


@pytest.mark.parametrize("case", CASES)
def test_mixed_defect_correction_is_composition_of_pinned_parts(case):
    """Each defect-correction step of orc_mg_solve_mixed: r = b - A x (FP64
    residual), r32 = (float) r, one FP32 V-cycle from e = 0 on the FP64
    operators rounded to FP32, x += (double) e.  The FP32 operators are
    (float) of the FP64 Galerkin hierarchy at every level."""
    fam, seed, dims, prd, colour8, lu, vd, vu = case
    A, b, mg = build(case)
    mg.set_rhs(b)
    ncyc = 3
    c, hist = mg.solve(1e-30, ncyc, mixed=True)
    assert c == ncyc
    L = mg.nlevels
    A64 = [mg.get_A(l) for l in range(L)]
    A32 = [a.astype(F32) for a in A64]
    for l in range(L):
        assert np.array_equal(mg.get_Af(l), A32[l])
    fac = None
    if mg.lu:
        ok, LU, ipiv = O.lu_factor(A64[-1], prd)
        fac = (LU, ipiv)
    x = np.zeros(mg.shape(0))
    for cyc in range(ncyc):
        res = O.residual(A, x, b, prd)
        assert float(np.sum(res * res)) == pytest.approx(hist[cyc], rel=1e-13, abs=0)
        xs = [np.zeros(mg.shape(l), dtype=F32) for l in range(L)]
        rs = [res.astype(F32)] + [np.zeros(mg.shape(l), dtype=F32) for l in range(1, L)]
        composed_cycle(A32, xs, rs, mg.dims, prd, vd, vu, colour8, fac, True)
        x = x + xs[0].astype(np.float64)
    assert np.array_equal(mg.get_x(0), x)


def dense(A, prd=0):
    o, n, m = A.shape[:3]
    N = m * n * o
    D = np.zeros((N, N))
    for k in range(o):
        for j in range(n):
            for i in range(m):
                p = i + m * (j + n * k)
                for e in range(27):
                    q, _ = nb_index(i + e % 3 - 1, j + (e // 3) % 3 - 1, k + e // 9 - 1, m, n, o, prd)
                    if q is not None:
                        D[p, q[2] + m * (q[1] + n * q[0])] = A[k, j, i, e]
    return D


@pytest.mark.parametrize("fam,dims,colour8", [(1, (13, 11, 12), False), (3, (12, 12, 12), True),
                                              (2, (14, 9, 10), False)])
def test_mixed_solve_converges_to_the_fp64_solution(fam, dims, colour8):
    """The mixed solve stops at tol with its residual taken in FP64 from its
    own x, and lies within |b - A x| / sigma_min(A) of the exact discrete
    solution, as the FP64 solve does; so the two are within the sum of those
    bounds of each other."""
    m, n, o = dims
    A = O.gen_operator(fam, 41, m, n, o)
    b = O.gen_rhs(41, m, n, o)
    D = dense(A)
    xstar = np.linalg.solve(D, b.reshape(-1))
    smin = np.linalg.svd(D, compute_uv=False)[-1]
    tol = 1e-10
    out = {}
    for mixed in (False, True):
        mg = O.Multigrid(A, colour8=colour8)
        mg.set_rhs(b)
        c, hist = mg.solve(tol, 400, mixed=mixed)
        x = mg.get_x(0)
        e2 = O.l2_error_all(A, x, b)
        assert e2 == hist[-1]
        assert np.sqrt(e2 / np.sum(b * b)) < tol
        bound = np.sqrt(e2) / smin
        assert np.linalg.norm(x.reshape(-1) - xstar) <= bound * (1 + 1e-6)
        out[mixed] = (x, bound, c)
    (x64, b64, c64), (x32, b32, c32) = out[False], out[True]
    assert np.linalg.norm((x32 - x64).reshape(-1)) <= (b32 + b64) * (1 + 1e-6)
    # FP32 inner cycles cost at most a few extra defect-correction steps
    assert c32 <= c64 + 3
