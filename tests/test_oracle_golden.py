"""The oracle against the golden vectors written by the real reference
(tests/golden/make_golden.py).  CPU only.

Bit-for-bit where the reference is deterministic (the 24-sweep bottom path);
the bottom-LU path is compared with a tolerance because the reference calls
MKL's blocked dgetrf/dgetrs, whose rounding differs from the oracle's
unblocked factorisation."""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                  if not os.path.basename(p).startswith(("solve_", "slices_", "summary_", "prd_", "fmg_", "diag_")))
SOLVES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "solve_*.npz")))


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def problem(z):
    fam, seed, m, n, o = [int(v) for v in z["meta"]]
    return fam, seed, m, n, o, O.gen_operator(fam, seed, m, n, o), O.gen_rhs(seed, m, n, o)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_fixture_set_present():
    assert len(FIXTURES) >= 7 and len(SOLVES) >= 3


@pytest.mark.parametrize("name", FIXTURES)
def test_inputs_match_what_reference_received(name):
    z = load(name)
    fam, seed, m, n, o, A, b = problem(z)
    assert sha(A) == str(z["sha_A0"])
    assert sha(b) == str(z["sha_b"])


@pytest.mark.parametrize("name", FIXTURES)
def test_hierarchy_shape(name):
    z = load(name)
    fam, seed, m, n, o = [int(v) for v in z["meta"]]
    assert [list(d) for d in O.hierarchy(m, n, o)] == z["dims"].tolist()


@pytest.mark.parametrize("name", FIXTURES)
def test_coarse_boxes_within_unit_offsets(name):
    """Every stencil box the reference built lies in [-1,1]^3 (the premise of
    the fixed 27-entry layout)."""
    z = load(name)
    for key in z.files:
        if key.startswith("S") and key.endswith("_range"):
            lo, hi = z[key]
            assert lo[0::2].min() >= -1 and hi[1::2].max() <= 2


@pytest.mark.parametrize("name", FIXTURES)
def test_galerkin_operators_bitwise(name):
    z = load(name)
    fam, seed, m, n, o, A, b = problem(z)
    mg = O.Multigrid(A)
    for l in range(1, mg.nlevels):
        assert np.array_equal(mg.get_A(l).reshape(-1), z["A%d" % l]), "level %d" % l


@pytest.mark.parametrize("name", FIXTURES)
def test_level_operations_bitwise(name):
    z = load(name)
    fam, seed, m, n, o, A, b = problem(z)
    mg = O.Multigrid(A)
    mg.set_rhs(b)
    assert mg.rhs_l2norm() == z["rhs_l2norm"][0]
    mg.set_x(0, O.test_field(m, n, o, seed + 1))
    assert mg.l2_error_all() == z["l2_error_all_xt"][0]
    mg.gauss_seidel(0, 0)
    assert np.array_equal(mg.get_x(0).reshape(-1), z["x_gs0"])
    mg.gauss_seidel(1, 0)
    assert np.array_equal(mg.get_x(0).reshape(-1), z["x_gs1"])
    if mg.nlevels > 1:
        mg.restriction(1)
        assert np.array_equal(mg.get_r(1).reshape(-1), z["r1_restrict"])
        assert np.array_equal(mg.get_x(1).reshape(-1), z["x1_after_restrict"])
        mc, nc, oc = mg.dims[1]
        mg.set_x(1, O.test_field(mc, nc, oc, seed + 2))
        mg.interpolation(1)
        assert np.array_equal(mg.get_x(0).reshape(-1), z["x_interp"])


@pytest.mark.parametrize("name", FIXTURES)
def test_v_cycles_bitwise(name):
    z = load(name)
    fam, seed, m, n, o, A, b = problem(z)
    mg = O.Multigrid(A)
    mg.set_rhs(b)
    hist = [mg.l2_error_all()]
    for c in range(4):
        mg.v_cycle()
        hist.append(mg.l2_error_all())
        if c == 0:
            assert np.array_equal(mg.get_x(0).reshape(-1), z["x_v1"])
    assert np.array_equal(np.array(hist), z["vcycle_res_gs24"])
    assert np.array_equal(mg.get_x(0).reshape(-1), z["x_v4_gs24"])


@pytest.mark.parametrize("name", [f for f in FIXTURES if "9x9x9" not in f])
def test_v_cycles_lu_bottom(name):
    z = load(name)
    fam, seed, m, n, o, A, b = problem(z)
    mg = O.Multigrid(A, lu=True)
    assert mg.lu == bool(z["lu_enabled"][0])
    mg.set_rhs(b)
    hist = [mg.l2_error_all()]
    for c in range(4):
        mg.v_cycle()
        hist.append(mg.l2_error_all())
    x = mg.get_x(0).reshape(-1)
    ref = z["x_v4_lu"]
    assert np.abs(x - ref).max() <= 1e-13 * np.abs(ref).max()
    np.testing.assert_allclose(np.array(hist), z["vcycle_res_lu"], rtol=1e-8)


@pytest.mark.parametrize("name", SOLVES)
def test_full_solve(name):
    z = load(name)
    info = json.loads(str(z["info"]))
    fam, seed, m, n, o, A, b = problem(z)
    mg = O.Multigrid(A, lu=name.endswith("_lu"))
    mg.set_rhs(b)
    c, hist = mg.solve(float(z["tol"]), 100)
    assert c == info["cycles"]
    x = mg.get_x(0).reshape(-1)
    if name.endswith("_lu"):
        np.testing.assert_allclose(hist, z["res"], rtol=1e-6)
        assert np.abs(x - z["x"]).max() <= 1e-13 * np.abs(z["x"]).max()
    else:
        assert np.array_equal(hist, z["res"])
        assert np.array_equal(x, z["x"])


def test_constructor_lu_falls_back():
    """The reference's constructor-time factorisation meets uninitialised
    coefficients and falls back to 24 sweeps in every fixture run, so its
    default solve uses the 24-sweep bottom."""
    for name in FIXTURES:
        assert int(load(name)["aexact_after_ctor"][0]) == 0


def test_full_size_summary_fixture():
    """The 257^3 reference run kept as a summary: ten cycles, each cutting
    the residual, below 1e-10 relative at the end; the source norm is the
    oracle's (same generator)."""
    z = load("summary_f7u_257_gs24")
    fam, seed, m, n, o = [int(v) for v in z["meta"]]
    res = z["res"]
    assert len(res) == 11 and np.all(np.diff(res) < 0)
    assert np.sqrt(res[-1] / z["rhs_l2norm"][0]) < float(z["tol"]) <= np.sqrt(res[-2] / z["rhs_l2norm"][0])
    b = O.gen_rhs(seed, m, n, o)
    np.testing.assert_allclose(float(np.sum(b * b)), float(z["rhs_l2norm"][0]), rtol=1e-12)
    assert len(str(z["sha_x"])) == 64 and len(z["sample_x"]) == len(z["sample_idx"])


@pytest.mark.parametrize("name", ["fmg_f7u_20x17x30", "fmg_q1_33x29x40", "fmg_prd_f7u_13x10x11_p5"])
def test_fmg_schedule_matches_reference(name):
    """The fmg_cycle schedule, as the oracle composes it, against the compiled
    reference's own fmg_cycle: the iterate after it (bit for bit on
    non-periodic grids; to 1e-14 with periodic axes, where the reference's
    restriction adds the wrapped contributions in another order) and after a
    further V-cycle."""
    z = load(name)
    fam, seed, m, n, o, prd = [int(v) for v in z["meta"]]
    A = O.gen_operator(fam, seed, m, n, o, prd)
    mg = O.Multigrid(A, prd=prd)
    mg.set_rhs(O.gen_rhs(seed, m, n, o))
    mg.fmg_cycle()
    x = mg.get_x(0).reshape(-1)
    x1 = mg.get_x(1).reshape(-1)
    l2 = mg.l2_error_all()
    mg.v_cycle()
    xv = mg.get_x(0).reshape(-1)
    if prd:
        scale = np.abs(z["x_fmg"]).max()
        assert np.abs(x - z["x_fmg"]).max() <= 1e-14 * scale
        assert np.abs(xv - z["x_fmg_v"]).max() <= 1e-14 * scale
        assert abs(l2 / z["l2_fmg"][0] - 1) < 1e-10
    else:
        assert np.array_equal(x, z["x_fmg"])
        assert np.array_equal(x1, z["x1_fmg"])
        assert np.array_equal(xv, z["x_fmg_v"])
        assert abs(l2 / z["l2_fmg"][0] - 1) < 1e-12
