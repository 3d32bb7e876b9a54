"""The drop-in C++ API (include/pgmg/) driven like the reference's own
programs (gs_parallel.cc, guide_test.cc, the fixture solve), built by
tests/cpp/Makefile against libpgmg_amd.so."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def run(args, timeout=300):
    p = subprocess.run([os.path.join(BIN, args[0])] + [str(a) for a in args[1:]],
                       capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def test_guide_matches_reference():
    """mg_guide's update rule against the sequence the reference's own
    mg_guide produced (oracle/ref_driver.cc `guide`)."""
    out = run(["guide_unit"])
    ref = open(os.path.join(GOLDEN, "guide_sequence.txt")).read()
    assert out == ref


def problem_simple_operator(m, n, o):
    A = np.zeros((o, n, m, 27))
    b = np.zeros((o, n, m))
    for k in range(o):
        for j in range(n):
            for i in range(m):
                inner = 0 < i < m - 1 and 0 < j < n - 1 and 0 < k < o - 1
                if inner:
                    for e in (4, 10, 12, 14, 16, 22):
                        A[k, j, i, e] = 1.0
                    A[k, j, i, 13] = -6.0
                else:
                    A[k, j, i, 13] = -0.375
                b[k, j, i] = 1.0 if (i, j, k) == (3, 3, 3) else (-1.0 if (i, j, k) == (12, 12, 12) else 0.0)
    return A, b


@pytest.mark.gpu
def test_lone_region_gauss_seidel_history(tmp_path):
    """gs_parallel.cc with the reference's problem_simple: every sweep's sum of
    squared residuals and the final iterate, against the oracle."""
    m = n = o = 24
    sweeps = 12
    xf = tmp_path / "x.bin"
    out = run(["gs_history", "simple", m, n, o, sweeps, xf]).splitlines()
    hist = [float(l.split()[1]) for l in out[:sweeps + 1]]
    A, b = problem_simple_operator(m, n, o)
    x = np.zeros((o, n, m))
    ref = [O.l2_error_all(A, x, b)]
    for s in range(sweeps):
        x = O.gs(A, x, b, 0)
        ref.append(O.l2_error_all(A, x, b))
    np.testing.assert_allclose(hist, ref, rtol=1e-12)
    assert np.array_equal(np.fromfile(xf).reshape(o, n, m), x)
    assert out[-1] == "gs_num %d" % sweeps


@pytest.mark.gpu
def test_multigrid_cycles_match_reference_fixture(tmp_path):
    z = np.load(os.path.join(GOLDEN, "solve_f7u_65_gs24.npz"))
    fam, seed, m, n, o = [int(v) for v in z["meta"]]
    xf = tmp_path / "x.bin"
    out = run(["mg_drivers", "cycles", fam, seed, m, n, o, xf]).splitlines()
    assert out[0] == "levels 4 65x65x65 33x33x33 17x17x17 9x9x9"
    res = [float(l.split()[2]) for l in out if l.startswith("res ")]
    sweeps = [int(l.split()[4]) for l in out if l.startswith("res ") and "sweeps" in l]
    np.testing.assert_allclose(res, z["res"], rtol=1e-12)
    assert sweeps == [42] * (len(res) - 1)
    assert np.array_equal(np.fromfile(xf), z["x"])


@pytest.mark.gpu
def test_solve_v_cycles_guide():
    out = run(["mg_drivers", "guide", 1, 5, 33, 33, 33, "/dev/null"]).splitlines()
    g = [l.split() for l in out if l.startswith("guide ")]
    assert len(g) == 3
    tol = (1.0 ** 2) * (np.finfo(float).eps ** 2) * 1e4 * 33 ** 3
    for row in g:
        assert row[3] == "0"
        assert float(row[5]) <= tol


@pytest.mark.gpu
@pytest.mark.parametrize("prd", [0, 5])
def test_fmg_cycle_against_oracle_composition(prd):
    """fmg_cycle (multigrid.cc:109-134) through the facade against the same
    schedule composed from the oracle's level operations (prd: periodic
    axes)."""
    m, n, o = 20, 17, 30
    fam = "1:%d" % prd if prd else 1
    out = run(["mg_drivers", "fmg", fam, 9, m, n, o, "/dev/null"]).splitlines()
    l2 = float([l for l in out if l.startswith("fmg ")][0].split()[1])
    A = O.gen_operator(1, 9, m, n, o, prd)
    b = O.gen_rhs(9, m, n, o)
    mg = O.Multigrid(A, prd=prd)
    mg.set_rhs(b)
    L = mg.nlevels
    for i in range(L - 1):
        mg.restriction(i + 1)
        for _ in range(2):
            mg.gauss_seidel(0, i + 1)
    for j in range(L - 1, 0, -1):
        for i in range(j, L - 1):
            mg.restriction(i + 1)
            for _ in range(2):
                mg.gauss_seidel(0, i + 1)
        for _ in range(20):
            mg.gauss_seidel(0, L - 1)
        for i in range(L - 1, j - 1, -1):
            mg.interpolation(i)
            for _ in range(2):
                mg.gauss_seidel(0, i - 1)
    assert abs(l2 / mg.l2_error_all() - 1) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["slices_f7u_20x17x30", "slices_q1_18x13x22"])
def test_slice_output_matches_reference_bytes(tmp_path, name):
    """output_x / output_r / output_residual after two exact-mode V-cycles,
    byte for byte against the files the reference's own output routines
    wrote for the same problem (region.cc:1562-1660)."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    fam, seed, m, n, o, k = [int(v) for v in z["meta"]]
    prefix = str(tmp_path / "s")
    out = run(["mg_drivers", "slices", fam, seed, m, n, o, prefix, k])
    assert "slices done" in out
    for field in ("x", "r", "res"):
        for lev in (0, 1):
            got = np.fromfile("%s_%s%d.bin" % (prefix, field, lev), dtype=np.uint8)
            assert np.array_equal(got, z["%s%d" % (field, lev)]), (field, lev)


@pytest.mark.gpu
def test_pcg_with_vcycle_preconditioner():
    """CG preconditioned by one V-cycle, driven through Ax and
    source_soln_dot as PGMG's CG users do; then the other outside-array
    helpers on the result."""
    out = run(["mg_drivers", "pcg", 1, 5, 33, 29, 40, "/dev/null"]).splitlines()
    hist = [float(l.split()[2]) for l in out if l.startswith("pcg ")]
    bb = hist[0]
    assert hist[-1] < 1e-20 * bb and len(hist) <= 16
    assert all(hist[q + 1] < hist[q] for q in range(len(hist) - 1))
    vals = dict((l.split()[0], l.split()[1:]) for l in out if not l.startswith(("pcg", "levels")))
    vals = {k: [float(t) for t in v if t != "bb"] for k, v in vals.items()}
    true_res, bb2 = vals["true"][0], vals["true"][1]
    assert bb2 == bb and true_res < 1e-18 * bb
    assert vals["diff_self"][0] == 0.0 and vals["diff_rhs"][0] == 0.0
    assert abs(vals["rhs_ratio"][0] - 2.25) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("fam,seed,dims,prd", [(1, 5, (13, 10, 11), 5), (3, 7, (12, 18, 9), 3)])
def test_multigrid_cycles_periodic(tmp_path, fam, seed, dims, prd):
    # a p_class with periodic axes through the drop-in API, against the oracle
    # (itself pinned to the reference's periodic runs)
    m, n, o = dims
    xf = tmp_path / "x.bin"
    out = run(["mg_drivers", "cycles", "%d:%d" % (fam, prd), seed, m, n, o, xf]).splitlines()
    assert out[0] == "levels %d " % len(O.hierarchy(m, n, o, prd)) + " ".join(
        "%dx%dx%d" % d for d in O.hierarchy(m, n, o, prd))
    res = [float(l.split()[2]) for l in out if l.startswith("res ")]
    mg = O.Multigrid(O.gen_operator(fam, seed, m, n, o, prd), prd=prd)
    mg.set_rhs(O.gen_rhs(seed, m, n, o))
    c, hist = mg.solve(1e-10, 100)
    np.testing.assert_allclose(res, hist, rtol=1e-12)
    assert np.array_equal(np.fromfile(xf).reshape(o, n, m), mg.get_x(0))
