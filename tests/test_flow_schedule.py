"""Model check, on the CPU, of the schedule of the one-launch exact sweeps
(k_gs_flow in pgmg_amd/csrc/pgmg_kernels.hip).

Each k-plane is a workgroup that walks steps tau = 0, 1, ...; step tau
updates row u = tau - 8s (u = i + 2j) of plane k for every sweep s in
flight, once plane k-1 has finished row min(u+3, umax) of sweep s and plane
k+1 row min(u+3, umax) of sweep s-1 (progress counted in steps).  Here the
planes advance in random interleavings, one whole step at a time (no step
that may run at the same time as another writes a row that step reads), and
the result must equal nsw lexicographic sweeps of the oracle bit for bit;
no interleaving may leave every unfinished plane waiting."""
import random

import numpy as np
import pytest

from oracle import oracle as O


def needs(tau, nsw, umax):
    nb = na = 0
    for s in range(nsw):
        u = tau - 8 * s
        if 0 <= u <= umax:
            top = min(u + 3, umax)
            nb = max(nb, top + 8 * s + 1)
            if s > 0:
                na = max(na, top + 8 * (s - 1) + 1)
    return nb, na


def gs_point(A, x, r, i, j, k):
    o, n, m = x.shape
    ax = 0.0
    for e in range(27):
        if e == 13:
            continue
        qi, qj, qk = i + e % 3 - 1, j + (e // 3) % 3 - 1, k + e // 9 - 1
        if 0 <= qi < m and 0 <= qj < n and 0 <= qk < o:
            ax = ax + float(A[k, j, i, e]) * float(x[qk, qj, qi])
    x[k, j, i] = (float(r[k, j, i]) - ax) * (1.0 / float(A[k, j, i, 13]))


def simulate(A, x0, r, nsw, rng):
    o, n, m = x0.shape
    umax = (m - 1) + 2 * (n - 1)
    nsteps = umax + 1 + 8 * (nsw - 1)
    x = x0.copy()
    prog = [0] * o
    while True:
        ready = []
        for k in range(o):
            if prog[k] == nsteps:
                continue
            nb, na = needs(prog[k], nsw, umax)
            if (k == 0 or prog[k - 1] >= nb) and (k == o - 1 or prog[k + 1] >= na):
                ready.append(k)
        if not ready:
            assert all(p == nsteps for p in prog), "every unfinished plane waits: %s" % prog
            return x
        k = rng.choice(ready)
        tau = prog[k]
        for s in range(nsw):
            u = tau - 8 * s
            if not 0 <= u <= umax:
                continue
            for j in range(max(0, (u - (m - 1) + 1) >> 1), min(n - 1, u >> 1) + 1):
                gs_point(A, x, r, u - 2 * j, j, k)
        prog[k] = tau + 1


@pytest.mark.parametrize("dims,nsw,seed", [((9, 7, 6), 1, 1), ((9, 7, 6), 2, 2), ((13, 10, 11), 3, 3),
                                           ((6, 11, 5), 2, 4), ((5, 4, 12), 4, 5), ((2, 3, 9), 3, 6),
                                           ((17, 3, 4), 2, 7)])
def test_flow_schedule_equals_lexicographic_sweeps(dims, nsw, seed):
    m, n, o = dims
    A = O.gen_operator(1, 50 + seed, m, n, o)
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-1, 1, size=(o, n, m))
    r = rng.uniform(-1, 1, size=(o, n, m))
    ref = x0.copy()
    for _ in range(nsw):
        ref = O.gs(A, ref, r, 0)
    for trial in range(3):
        x = simulate(A, x0, r, nsw, random.Random(100 * seed + trial))
        assert np.array_equal(x, ref)


def test_flow_schedule_catches_a_short_wait():
    """The check is sharp: waiting for row u+2 of the plane below instead of
    u+3 lets some interleaving read a value too early."""
    m, n, o = 9, 7, 6
    A = O.gen_operator(1, 60, m, n, o)
    rng = np.random.default_rng(8)
    x0 = rng.uniform(-1, 1, size=(o, n, m))
    r = rng.uniform(-1, 1, size=(o, n, m))
    ref = O.gs(A, x0, r, 0)
    global needs
    orig = needs

    def short(tau, nsw, umax):
        nb, na = orig(tau, nsw, umax)
        return max(0, nb - 1), na
    needs = short
    try:
        bad = any(not np.array_equal(simulate(A, x0, r, 1, random.Random(t)), ref) for t in range(20))
    finally:
        needs = orig
    assert bad
