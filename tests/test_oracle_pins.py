"""Pins of the oracle by what the mathematics fixes, independent of the
reference run: closed-form transfer operators, a brute-force dense Galerkin
product, an independent pure-Python Gauss-Seidel in lexicographic and
hyperplane order, fixed points, the Jacobi special case of the colour
smoother, and LAPACK's 2x2 known answer (lapack_test.cc).  CPU only."""
import random

import numpy as np
import pytest

from oracle import oracle as O


def P1(M):
    """1-D interpolation matrix (fine x coarse) of region.cc:752-810 with the
    even-size injection of the last point."""
    K = (M + 2) >> 1
    P = np.zeros((M, K))
    for f in range(M):
        c, l = f >> 1, f & 1
        last = (M - 1) >> 1
        if l == 0:
            P[f, c] = 1
        elif c == last and M % 2 == 0:
            P[f, c + 1] = 1
        else:
            P[f, c] = P[f, c + 1] = 0.5
    return P


def P3(m, n, o):
    return np.kron(P1(o), np.kron(P1(n), P1(m)))


def dense(A):
    o, n, m = A.shape[:3]
    N = m * n * o
    D = np.zeros((N, N))
    for k in range(o):
        for j in range(n):
            for i in range(m):
                p = i + m * (j + n * k)
                for e in range(27):
                    di, dj, dk = e % 3 - 1, (e // 3) % 3 - 1, e // 9 - 1
                    qi, qj, qk = i + di, j + dj, k + dk
                    if 0 <= qi < m and 0 <= qj < n and 0 <= qk < o:
                        D[p, qi + m * (qj + n * qk)] = A[k, j, i, e]
    return D


def rand_op(m, n, o, seed, sym=True):
    """Random diagonally dominant 27-point operator (zero outside the grid)."""
    rng = np.random.default_rng(seed)
    D = np.zeros((m * n * o, m * n * o))
    A = rng.uniform(-1, 0, size=(o, n, m, 27))
    A[..., 13] = 30.0 + rng.uniform(0, 1, size=(o, n, m))
    for k in range(o):
        for j in range(n):
            for i in range(m):
                for e in range(27):
                    di, dj, dk = e % 3 - 1, (e // 3) % 3 - 1, e // 9 - 1
                    if not (0 <= i + di < m and 0 <= j + dj < n and 0 <= k + dk < o):
                        A[k, j, i, e] = 0.0
    return A


@pytest.mark.parametrize("m,n,o,levels", [
    (65, 65, 65, [(65, 65, 65), (33, 33, 33), (17, 17, 17), (9, 9, 9)]),
    (20, 17, 30, [(20, 17, 30), (11, 9, 16), (6, 5, 9)]),
    (33, 65, 129, [(33, 65, 129), (17, 33, 65), (9, 17, 33), (5, 9, 17)]),
])
def test_hierarchy_table(m, n, o, levels):
    assert O.hierarchy(m, n, o) == levels


def test_hierarchy_depths():
    assert len(O.hierarchy(257, 257, 257)) == 6
    assert len(O.hierarchy(513, 513, 513)) == 7
    assert O.hierarchy(9, 9, 9) == [(9, 9, 9)]


@pytest.mark.parametrize("dims", [(9, 7, 5), (8, 6, 10), (5, 12, 7), (6, 6, 6)])
def test_interpolation_is_P(dims):
    m, n, o = dims
    (mc, nc, oc) = O.hierarchy(m, n, o)[1] if len(O.hierarchy(m, n, o)) > 1 else ((m + 2) >> 1, (n + 2) >> 1, (o + 2) >> 1)
    rng = np.random.default_rng(1)
    xc = rng.uniform(-1, 1, size=(oc, nc, mc))
    xf0 = rng.uniform(-1, 1, size=(o, n, m))
    xf = O.interp(xc, xf0)
    ref = xf0.reshape(-1) + P3(m, n, o) @ xc.reshape(-1)
    np.testing.assert_allclose(xf.reshape(-1), ref, rtol=0, atol=1e-15)


def test_interpolation_reproduces_linear_fields():
    """mg_test.cc:30-33: a field linear in the coarse index interpolates to
    the linear fine field (odd sizes: exact, every point)."""
    m, n, o = 17, 9, 13
    mc, nc, oc = (m + 2) >> 1, (n + 2) >> 1, (o + 2) >> 1
    J = np.arange(nc, dtype=float)
    xc = np.broadcast_to(J[None, :, None], (oc, nc, mc)).copy()
    xf = O.interp(xc, np.zeros((o, n, m)))
    j = np.arange(n, dtype=float)
    assert np.array_equal(xf, np.broadcast_to(j[None, :, None] / 2, (o, n, m)))
    xf1 = O.interp(np.ones((oc, nc, mc)), np.zeros((o, n, m)))
    assert np.array_equal(xf1, np.ones((o, n, m)))


@pytest.mark.parametrize("dims", [(9, 7, 5), (5, 11, 13)])
def test_restriction_is_adjoint_on_odd_grids(dims):
    """R = P^T/8 for odd sizes: <R u, v> = <u, P v>/8."""
    m, n, o = dims
    mc, nc, oc = (m + 2) >> 1, (n + 2) >> 1, (o + 2) >> 1
    rng = np.random.default_rng(2)
    u = rng.uniform(-1, 1, size=(o, n, m))
    Ru = O.restrict(np.zeros((o, n, m, 27)), np.zeros((o, n, m)), u, (mc, nc, oc))
    ref = P3(m, n, o).T @ u.reshape(-1) / 8
    np.testing.assert_allclose(Ru.reshape(-1), ref, rtol=0, atol=1e-15)


def test_restriction_even_boundary_quirk():
    """On even grids the reference's restrict_partial (region.cc:888-895)
    splits local offset (0,1,1) by the x injection flag, so R != P^T/8 only in
    boundary cubes; interior coarse points keep full weighting."""
    m = n = o = 8
    mc = (m + 2) >> 1
    rng = np.random.default_rng(3)
    u = rng.uniform(-1, 1, size=(o, n, m))
    Ru = O.restrict(np.zeros((o, n, m, 27)), np.zeros((o, n, m)), u, (mc, mc, mc))
    ref = (P3(m, n, o).T @ u.reshape(-1) / 8).reshape(mc, mc, mc)
    diff = np.abs(Ru - ref) > 1e-14
    assert diff.any()
    inner = diff[:mc - 2, :mc - 2, :mc - 2]
    assert not inner.any()
    # one unit residual in a boundary cube at local (0,1,1): cube (3,0,0)
    # holds fine (6,1,1); the reference sends it to coarse y = 1 only.
    e = np.zeros((o, n, m)); e[1, 1, 6] = 1.0
    Re = O.restrict(np.zeros((o, n, m, 27)), np.zeros((o, n, m)), e, (mc, mc, mc))
    nz = {tuple(int(v) for v in t): Re[tuple(t)] for t in np.argwhere(Re != 0)}
    assert nz == {(0, 1, 3): 0.0625, (1, 1, 3): 0.0625}


@pytest.mark.parametrize("dims", [(5, 5, 5), (6, 5, 4), (7, 4, 6), (4, 4, 4)])
def test_rat_is_dense_galerkin_product(dims):
    m, n, o = dims
    mc, nc, oc = (m + 2) >> 1, (n + 2) >> 1, (o + 2) >> 1
    A = rand_op(m, n, o, 4)
    Ac = O.rat(A, (mc, nc, oc))
    P = P3(m, n, o)
    ref = P.T @ dense(A) @ P / 8
    np.testing.assert_allclose(dense(Ac), ref, rtol=0, atol=1e-12)


def test_rat_preserves_symmetry_and_zero_row_sums():
    """Q1 Laplacian stencil (problem_simple.cc:18, x12: centre 32, edges -2,
    corners -1, faces 0) with zero row sums away from the boundary: the
    Galerkin operator stays symmetric and interior rows still sum to zero."""
    m = n = o = 9
    st = np.zeros(27)
    for e in range(27):
        d = [abs(e % 3 - 1), abs((e // 3) % 3 - 1), abs(e // 9 - 1)]
        s = sum(d)
        st[e] = {0: 32.0, 1: 0.0, 2: -2.0, 3: -1.0}[s]
    A = np.broadcast_to(st, (o, n, m, 27)).copy()
    for k in range(o):
        for j in range(n):
            for i in range(m):
                for e in range(27):
                    di, dj, dk = e % 3 - 1, (e // 3) % 3 - 1, e // 9 - 1
                    if not (0 <= i + di < m and 0 <= j + dj < n and 0 <= k + dk < o):
                        A[k, j, i, e] = 0.0
    Ac = O.rat(A, (5, 5, 5))
    D = dense(Ac)
    np.testing.assert_allclose(D, D.T, rtol=0, atol=1e-13)
    assert abs(Ac[2, 2, 2].sum()) < 1e-13


def py_gs(A, x, r, order):
    o, n, m = x.shape
    x = x.copy()
    for (i, j, k) in order:
        ax = 0.0
        for e in range(27):
            if e == 13:
                continue
            di, dj, dk = e % 3 - 1, (e // 3) % 3 - 1, e // 9 - 1
            qi, qj, qk = i + di, j + dj, k + dk
            if 0 <= qi < m and 0 <= qj < n and 0 <= qk < o:
                ax = ax + float(A[k, j, i, e]) * float(x[qk, qj, qi])
        x[k, j, i] = (float(r[k, j, i]) - ax) * (1.0 / float(A[k, j, i, 13]))
    return x


def test_gauss_seidel_lexicographic_and_wavefront_orders():
    """An independent pure-Python sweep (IEEE doubles, no contraction) in
    lexicographic order equals the oracle bit for bit, and so does the same
    sweep in hyperplane i+2j+4k order with points shuffled inside a plane."""
    m, n, o = 5, 4, 3
    A = rand_op(m, n, o, 5)
    rng = np.random.default_rng(6)
    x0 = rng.uniform(-1, 1, size=(o, n, m))
    r = rng.uniform(-1, 1, size=(o, n, m))
    lex = [(i, j, k) for k in range(o) for j in range(n) for i in range(m)]
    x_lex = py_gs(A, x0, r, lex)
    assert np.array_equal(x_lex, O.gs(A, x0, r, 0))
    planes = {}
    for (i, j, k) in lex:
        planes.setdefault(i + 2 * j + 4 * k, []).append((i, j, k))
    order = []
    rnd = random.Random(7)
    for t in sorted(planes):
        pts = planes[t][:]
        rnd.shuffle(pts)
        order += pts
    assert np.array_equal(py_gs(A, x0, r, order), x_lex)
    # reverse sweep (mode 1 = forward then backward)
    assert np.array_equal(O.gs(A, x0, r, 1), py_gs(A, x_lex, r, lex[::-1]))


def test_colour8_order():
    m, n, o = 6, 5, 4
    A = rand_op(m, n, o, 8)
    rng = np.random.default_rng(9)
    x0 = rng.uniform(-1, 1, size=(o, n, m))
    r = rng.uniform(-1, 1, size=(o, n, m))
    order = [(i, j, k) for c in range(8) for k in range(o) for j in range(n) for i in range(m)
             if (i & 1) + 2 * (j & 1) + 4 * (k & 1) == c]
    assert np.array_equal(O.gs(A, x0, r, 0, colour8=True), py_gs(A, x0, r, order))


def test_colour8_on_diagonal_operator_is_jacobi():
    m, n, o = 6, 5, 4
    A = np.zeros((o, n, m, 27))
    A[..., 13] = np.random.default_rng(10).uniform(1, 2, size=(o, n, m))
    r = np.random.default_rng(11).uniform(-1, 1, size=(o, n, m))
    x = O.gs(A, np.zeros((o, n, m)), r, 0, colour8=True)
    assert np.array_equal(x, r * (1.0 / A[..., 13]))


def test_exact_solution_is_fixed_point():
    m, n, o = 5, 4, 6
    A = rand_op(m, n, o, 12)
    b = np.random.default_rng(13).uniform(-1, 1, size=(o, n, m))
    xs = np.linalg.solve(dense(A), b.reshape(-1)).reshape(o, n, m)
    np.testing.assert_allclose(O.gs(A, xs, b, 1), xs, rtol=0, atol=1e-14)
    np.testing.assert_allclose(O.gs(A, xs, b, 0, colour8=True), xs, rtol=0, atol=1e-14)
    assert O.l2_error_all(A, xs, b) < 1e-26


def test_lu_known_answer_lapack_test():
    """lapack_test.cc: A = [[1,2],[1,1]], b = (3,7) -> x = (11,-4)."""
    A = np.zeros((1, 1, 2, 27))
    A[0, 0, 0, 13], A[0, 0, 0, 14] = 1.0, 2.0
    A[0, 0, 1, 12], A[0, 0, 1, 13] = 1.0, 1.0
    ok, LU, ipiv = O.lu_factor(A)
    assert ok
    assert np.array_equal(O.lu_solve(LU, ipiv, np.array([3.0, 7.0])), np.array([11.0, -4.0]))


def test_lu_matches_dense_solve_and_fallback_rule():
    m, n, o = 5, 4, 3
    A = rand_op(m, n, o, 14)
    b = np.random.default_rng(15).uniform(-1, 1, size=m * n * o)
    ok, LU, ipiv = O.lu_factor(A)
    assert ok
    np.testing.assert_allclose(O.lu_solve(LU, ipiv, b), np.linalg.solve(dense(A), b), rtol=1e-12, atol=1e-14)
    # singular (all zero) matrix: the reference falls back to Gauss-Seidel
    ok0, _, _ = O.lu_factor(np.zeros((o, n, m, 27)))
    assert not ok0


@pytest.mark.parametrize("family", [1, 2, 3])
def test_generator_planes_match_whole_grid(family):
    """The plane-range generator returns the same rows as the whole-grid one;
    the Q1 operator is symmetric with zero row sums away from the boundary
    (a sum of element stiffness matrices of constant tensors)."""
    m, n, o = 11, 9, 8
    A = O.gen_operator(family, 17, m, n, o)
    for k0, k1 in [(0, 3), (2, 8), (7, 8)]:
        assert np.array_equal(O.gen_operator_planes(family, 17, m, n, o, k0, k1), A[k0:k1])
    if family == 3:
        D = dense(A)
        np.testing.assert_allclose(D, D.T, rtol=0, atol=1e-12)
        interior = A[1:-1, 1:-1, 1:-1].sum(axis=-1)
        np.testing.assert_allclose(interior, 0.0, atol=1e-12)
