"""GPU parity on grids with periodic axes, through the C ABI: the HIP library
against the oracle (bit for bit: Galerkin operators, both smoothers,
restriction, interpolation, V-cycle iterates, the LU bottom and the FP32
inner cycle) and against the real reference's periodic runs
(tests/golden/prd_*.npz: sweeps and interpolation bit for bit; the
restriction, Galerkin operators and cycles to rounding, since the reference
adds the contributions it carries through its ghost strips in another
order).  Sums of squares are reduced in a different order than a serial loop
and are compared to 1e-12 relative."""
import glob
import os
import sys

import numpy as np
import pytest

import pgmg_amd as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PFIX = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "prd_*.npz")))
NORM_RTOL = 1e-12
ROUND = 1e-13


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    return z, tuple(int(v) for v in z["meta"])


def maxrel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def setup(meta, generate=False, **kw):
    fam, seed, m, n, o, prd = meta
    A = O.gen_operator(fam, seed, m, n, o, prd)
    b = O.gen_rhs(seed, m, n, o)
    h = P.Hierarchy(m, n, o, prd=prd, **kw)
    if generate:
        h.generate_problem(fam, seed)
    else:
        h.set_operator27(A)
        h.set_rhs(b)
    h.compute_rats()
    return h, A, b


@pytest.mark.parametrize("name", PFIX)
def test_periodic_shape_generation_and_galerkin(name):
    z, meta = load(name)
    say("case", name)
    fam, seed, m, n, o, prd = meta
    assert P.hierarchy_shape(m, n, o, prd) == [tuple(int(v) for v in d) for d in z["dims"]]
    h, A, b = setup(meta, generate=True, smoother=P.SMOOTH_EXACT)
    assert np.array_equal(h.get_operator27(0), A)
    assert np.array_equal(h.get_rhs(0), b)
    mg = O.Multigrid(A, prd=prd)
    for l in range(1, h.nlevels):
        Ag = h.get_operator27(l)
        assert np.array_equal(Ag, mg.get_A(l)), "level %d" % l
        assert maxrel(Ag.reshape(-1), z["A%d" % l]) < ROUND
    h.close()


@pytest.mark.parametrize("name", PFIX)
def test_periodic_level_operations(name):
    z, meta = load(name)
    say("case", name)
    fam, seed, m, n, o, prd = meta
    h, A, b = setup(meta, smoother=P.SMOOTH_EXACT, use_graph=False)
    xt = O.test_field(m, n, o, seed + 1)
    h.set_x(xt)
    assert abs(h.l2_error_all() / z["l2_error_all_xt"][0] - 1) < NORM_RTOL
    h.gauss_seidel(0, 0)
    assert np.array_equal(h.get_x(0).reshape(-1), z["x_gs0"])
    h.gauss_seidel(1, 0)
    x1 = h.get_x(0)
    assert np.array_equal(x1.reshape(-1), z["x_gs1"])
    np.testing.assert_array_equal(h.residual(0), O.residual(A, x1, b, prd=prd))
    dims = h.dims
    h.restriction(1)
    rc = h.get_rhs(1)
    assert np.array_equal(rc, O.restrict(A, x1, b, dims[1], prd=prd))
    assert maxrel(rc.reshape(-1), z["r1_restrict"]) < ROUND
    assert not h.get_x(1).any()
    xc = O.test_field(*dims[1], seed + 2)
    h.set_x(xc, level=1)
    h.interpolation(1)
    assert np.array_equal(h.get_x(0).reshape(-1), z["x_interp"])
    v = O.test_field(m, n, o, seed + 3)
    D = h.apply_operator(v)
    # 0 - Av is exact, so the residual oracle with a zero source is -Av
    np.testing.assert_array_equal(D, -O.residual(A, v, np.zeros_like(b), prd=prd))
    h.close()


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("name", PFIX)
def test_periodic_v_cycles_exact_mode(name, graph):
    z, meta = load(name)
    say("case", name, "graph", graph)
    fam, seed, m, n, o, prd = meta
    h, A, b = setup(meta, smoother=P.SMOOTH_EXACT, use_graph=graph)
    mg = O.Multigrid(A, prd=prd)
    mg.set_rhs(b)
    hist = [h.l2_error_all()]
    for c in range(4):
        h.v_cycle()
        mg.v_cycle()
        assert np.array_equal(h.get_x(0), mg.get_x(0)), "cycle %d" % c
        hist.append(h.l2_error_all())
    np.testing.assert_allclose(hist, z["vcycle_res_gs24"], rtol=1e-10, atol=0)
    assert maxrel(h.get_x(0).reshape(-1), z["x_v4_gs24"]) < 1e-11
    h.close()


@pytest.mark.parametrize("name", PFIX)
def test_periodic_lu_bottom(name):
    z, meta = load(name)
    say("case", name)
    fam, seed, m, n, o, prd = meta
    h, A, b = setup(meta, smoother=P.SMOOTH_EXACT, bottom=P.BOTTOM_LU)
    mg = O.Multigrid(A, prd=prd, lu=True)
    assert h.lu_active() == mg.lu == bool(z["lu_enabled"][0])
    mg.set_rhs(b)
    hist = [h.l2_error_all()]
    for c in range(4):
        h.v_cycle()
        mg.v_cycle()
        hist.append(h.l2_error_all())
    assert np.array_equal(h.get_x(0), mg.get_x(0))
    np.testing.assert_allclose(hist, z["vcycle_res_lu"], rtol=1e-8, atol=0)
    h.close()


@pytest.mark.parametrize("name", PFIX)
def test_periodic_v_cycles_colour8(name):
    z, meta = load(name)
    say("case", name)
    fam, seed, m, n, o, prd = meta
    h, A, b = setup(meta, smoother=P.SMOOTH_COLOUR8)
    mg = O.Multigrid(A, colour8=True, prd=prd)
    mg.set_rhs(b)
    for c in range(3):
        h.v_cycle()
        mg.v_cycle()
    assert np.array_equal(h.get_x(0), mg.get_x(0))
    assert abs(h.l2_error_all() / mg.l2_error_all() - 1) < NORM_RTOL
    h.close()


@pytest.mark.parametrize("smoother", [P.SMOOTH_EXACT, P.SMOOTH_COLOUR8])
@pytest.mark.parametrize("meta", [(1, 5, 13, 10, 11, 5), (3, 7, 12, 18, 9, 3), (2, 9, 20, 17, 30, 6)])
def test_periodic_mixed_solve_matches_oracle(meta, smoother):
    say("case", meta, "smoother", smoother)
    fam, seed, m, n, o, prd = meta
    h, A, b = setup(meta, smoother=smoother, precision=P.PREC_MIXED)
    mg = O.Multigrid(A, colour8=smoother == P.SMOOTH_COLOUR8, prd=prd)
    mg.set_rhs(b)
    # the high-contrast f7log case converges slowly: both stop at 40 cycles
    c_o, hist_o = mg.solve(1e-10, 40, mixed=True)
    c_g, hist_g = h.solve(1e-10, 40, check=False)
    assert c_g == c_o
    if c_o < 40:
        assert np.sqrt(hist_g[-1] / h.rhs_l2norm()) < 1e-10
    np.testing.assert_allclose(hist_g, hist_o, rtol=NORM_RTOL, atol=0)
    assert np.array_equal(h.get_x(0), mg.get_x(0))
    h.close()


@pytest.mark.parametrize("prd", [1, 2, 4, 3, 7])
def test_periodic_64_cube_properties(prd):
    # full-size sweep on the device only: a fixed point stays fixed, and a
    # V-cycle of the f7 family with a zero-mean source reduces the residual
    say("case 64^3 prd", prd)
    m = n = o = 64
    h = P.Hierarchy(m, n, o, prd=prd, smoother=P.SMOOTH_COLOUR8, precision=P.PREC_FP64)
    h.generate_problem(1, 31)
    h.compute_rats()
    assert h.dims == O.hierarchy(m, n, o, prd)
    b = h.get_rhs(0)
    b -= b.mean()
    h.set_rhs(b)
    e0 = h.l2_error_all()
    for c in range(6):
        h.v_cycle()
    e6 = h.l2_error_all()
    assert e6 < 1e-6 * e0
    h.close()


@pytest.mark.parametrize("dims,prd", [((262, 6, 5), 1), ((256, 4, 4), 7), ((130, 7, 6), 3), ((18, 22, 14), 7)])
def test_periodic_row_kernels_colour8(dims, prd):
    # rows long enough for the segmented row kernels (FP64 beyond 128 points,
    # FP32 beyond 256), rows ending inside a lane, and a periodic x axis,
    # whose halo point m the row kernels read from the ghost copy
    say("case", dims, "prd", prd)
    m, n, o = dims
    meta = (1, 41, m, n, o, prd)
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        h, A, b = setup(meta, smoother=P.SMOOTH_COLOUR8, precision=prec)
        mg = O.Multigrid(A, colour8=True, prd=prd)
        mg.set_rhs(b)
        if prec == P.PREC_FP64:
            for c in range(3):
                h.v_cycle()
                mg.v_cycle()
        else:
            c_o, hist_o = mg.solve(1e-30, 3, mixed=True)
            c_g, hist_g = h.solve(1e-30, 3, check=False)
            assert c_g == c_o == 3
        assert np.array_equal(h.get_x(0), mg.get_x(0)), prec
        h.close()


def _random_cases(n=8, seed=2026):
    rng = np.random.default_rng(seed)
    cases = []
    while len(cases) < n:
        dims = tuple(int(v) for v in rng.integers(3, 30, size=3))
        prd = int(rng.integers(1, 8))
        if len(O.hierarchy(*dims, prd)) > 1:
            cases.append((dims, prd, int(rng.integers(1, 4))))
    return cases


@pytest.mark.parametrize("dims,prd,fam", _random_cases())
def test_periodic_random_shapes(dims, prd, fam):
    say("case", dims, "prd", prd, "family", fam)
    _periodic_case(dims, prd, fam)


# Periodic axes of size 2 on a level that restricts: its one cube is also
# the last, so its local 1 reaches coarse 0 twice, as that point's own share
# and through the wrap (two such axes on (261, 2, 2)).
SIZE_TWO = [((18, 29, 2), 4, 1), ((18, 29, 2), 7, 2), ((2, 24, 29), 1, 3), ((128, 128, 8), 7, 2),
            ((520, 4, 4), 6, 3)]


def test_size_two_cases_reach_a_restricting_level():
    for dims, prd, fam in SIZE_TWO:
        levels = O.hierarchy(*dims, prd)[:-1]
        assert any(prd >> a & 1 and l[a] == 2 for l in levels for a in range(3)), dims
    assert any(sum(prd >> a & 1 and l[a] == 2 for a in range(3)) == 2
               for dims, prd, fam in SIZE_TWO for l in O.hierarchy(*dims, prd)[:-1])


@pytest.mark.parametrize("dims,prd,fam", SIZE_TWO)
def test_periodic_axis_of_size_two(dims, prd, fam):
    say("case", dims, "prd", prd, "family", fam)
    _periodic_case(dims, prd, fam)


def _periodic_case(dims, prd, fam):
    """Both smoothers: Galerkin operators and two V-cycles bitwise against
    the oracle; mixed precision: two defect-correction cycles bitwise."""
    m, n, o = dims
    meta = (fam, 77, m, n, o, prd)
    for smoother in (P.SMOOTH_EXACT, P.SMOOTH_COLOUR8):
        h, A, b = setup(meta, smoother=smoother)
        mg = O.Multigrid(A, colour8=smoother == P.SMOOTH_COLOUR8, prd=prd)
        mg.set_rhs(b)
        for l in range(1, h.nlevels):
            assert np.array_equal(h.get_operator27(l), mg.get_A(l))
        for c in range(2):
            h.v_cycle()
            mg.v_cycle()
        assert np.array_equal(h.get_x(0), mg.get_x(0)), smoother
        h.close()
    h, A, b = setup(meta, smoother=P.SMOOTH_COLOUR8, precision=P.PREC_MIXED)
    mg = O.Multigrid(A, colour8=True, prd=prd)
    mg.set_rhs(b)
    c_o, _ = mg.solve(1e-30, 2, mixed=True)
    c_g, _ = h.solve(1e-30, 2, check=False)
    assert np.array_equal(h.get_x(0), mg.get_x(0))
    h.close()


@pytest.mark.parametrize("dims,prd", [((48, 40, 30), 5), ((36, 50, 33), 7)])
def test_periodic_exact_large_level(dims, prd):
    # a level above the one-workgroup size: exact sweeps on the plane-skewed
    # copies, with the ghosts refreshed on them before each sweep
    say("case", dims, "prd", prd)
    m, n, o = dims
    h, A, b = setup((2, 13, m, n, o, prd), smoother=P.SMOOTH_EXACT)
    mg = O.Multigrid(A, prd=prd)
    mg.set_rhs(b)
    for c in range(2):
        h.v_cycle()
        mg.v_cycle()
    assert np.array_equal(h.get_x(0), mg.get_x(0))
    h.close()


def test_periodic_mixed_exact_large_level():
    m, n, o, prd = 48, 40, 30, 5
    h, A, b = setup((1, 19, m, n, o, prd), smoother=P.SMOOTH_EXACT, precision=P.PREC_MIXED)
    mg = O.Multigrid(A, prd=prd)
    mg.set_rhs(b)
    c_o, _ = mg.solve(1e-30, 3, mixed=True)
    c_g, _ = h.solve(1e-30, 3, check=False)
    assert c_g == c_o == 3
    assert np.array_equal(h.get_x(0), mg.get_x(0))
    h.close()
