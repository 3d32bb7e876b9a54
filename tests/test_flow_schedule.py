"""Model check, on the CPU, of the schedule of the one-launch exact sweeps
(k_gs_flow in pgmg_amd/csrc/pgmg_kernels.hip).

Each k-plane is a workgroup that walks steps tau = 0, 1, ...; step tau
updates row u = tau - LAG*s (u = i + 2j) of plane k for every sweep s in
flight.  A step has four phases, and the planes interleave at phase
granularity in random orders:
  wait    until plane k-1 has published row min(u+3, umax) of sweep s as
          done, and plane k+1 row min(u+3, umax) of sweep s-1;
  read    every neighbour value the step's points use (one snapshot);
  publish that the previous step is done (its stores are in memory: the
          kernel knows it once this step's loads, issued after them, have
          completed);
  write   the new values.
The counts are published one step late, and published after the last step.
The result must equal nsw lexicographic sweeps of the oracle bit for bit, and
no interleaving may leave every unfinished plane waiting."""
import random

import numpy as np
import pytest

from oracle import oracle as O

LAG = 16  # rows between consecutive sweeps in flight (kFlowLag)


def needs(tau, nsw, umax, lag=LAG):
    nb = na = 0
    for s in range(nsw):
        u = tau - lag * s
        if 0 <= u <= umax:
            top = min(u + 3, umax)
            nb = max(nb, top + lag * s + 1)
            if s > 0:
                na = max(na, top + lag * (s - 1) + 1)
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


def gs_value(A, x, r, i, j, k):
    """gs_point's new value for (i, j, k), without writing it."""
    o, n, m = x.shape
    ax = 0.0
    for e in range(27):
        if e == 13:
            continue
        qi, qj, qk = i + e % 3 - 1, j + (e // 3) % 3 - 1, k + e // 9 - 1
        if 0 <= qi < m and 0 <= qj < n and 0 <= qk < o:
            ax = ax + float(A[k, j, i, e]) * float(x[qk, qj, qi])
    return (float(r[k, j, i]) - ax) * (1.0 / float(A[k, j, i, 13]))


def simulate(A, x0, r, nsw, rng, lag=LAG, need=None):
    need = need or needs
    o, n, m = x0.shape
    umax = (m - 1) + 2 * (n - 1)
    nsteps = umax + 1 + lag * (nsw - 1)
    x = x0.copy()
    pub = [0] * o           # published counts
    step = [0] * o          # step each plane is in
    phase = ["wait"] * o    # wait -> read -> publish -> write -> (next step)
    vals = [None] * o
    while True:
        ready = []
        for k in range(o):
            if phase[k] == "done":
                continue
            if phase[k] == "wait":
                nb, na = need(step[k], nsw, umax, lag)
                if not ((k == 0 or pub[k - 1] >= nb) and (k == o - 1 or pub[k + 1] >= na)):
                    continue
            ready.append(k)
        if not ready:
            assert all(p == "done" for p in phase), "every unfinished plane waits: %s %s" % (step, pub)
            return x
        # mostly the highest plane that may move: upper planes then read as
        # early as the waits allow, where a missing wait would show
        k = max(ready) if rng.random() < 0.7 else rng.choice(ready)
        tau = step[k]
        if phase[k] == "wait":
            phase[k] = "read"
        elif phase[k] == "read":
            v = []
            for s in range(nsw):
                u = tau - lag * s
                if 0 <= u <= umax:
                    for j in range(max(0, (u - (m - 1) + 1) >> 1), min(n - 1, u >> 1) + 1):
                        v.append((u - 2 * j, j, gs_value(A, x, r, u - 2 * j, j, k)))
            vals[k] = v
            phase[k] = "publish"
        elif phase[k] == "publish":
            pub[k] = tau
            phase[k] = "write"
        else:
            for (i, j, val) in vals[k]:
                x[k, j, i] = val
            step[k] = tau + 1
            if step[k] == nsteps:
                pub[k] = nsteps
                phase[k] = "done"
            else:
                phase[k] = "wait"


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
    def short(tau, nsw, umax, lag):
        nb, na = needs(tau, nsw, umax, lag)
        return max(0, nb - 1), na
    assert any(not np.array_equal(simulate(A, x0, r, 1, random.Random(t), need=short), ref) for t in range(40))


def test_flow_schedule_needs_the_sweep_lag():
    """With sweeps in flight only eight rows apart the late publication could
    leave two neighbouring planes each waiting for the other."""
    m, n, o = 9, 7, 6
    A = O.gen_operator(1, 61, m, n, o)
    rng = np.random.default_rng(9)
    x0 = rng.uniform(-1, 1, size=(o, n, m))
    r = rng.uniform(-1, 1, size=(o, n, m))
    stuck = 0
    for t in range(20):
        try:
            simulate(A, x0, r, 3, random.Random(t), lag=8)
        except AssertionError as e:
            stuck += "waits" in str(e)
    assert stuck > 0


def pipelined_hyperplanes(A, x0, r, nsw, lag=8):
    """The one-workgroup schedule: sweep s updates hyperplane t = i+2j+4k at
    step t + lag*s, every point of a step from the values before the step."""
    o, n, m = x0.shape
    tmax = (m - 1) + 2 * (n - 1) + 4 * (o - 1)
    x = x0.copy()
    for st in range(tmax + 1 + lag * (nsw - 1)):
        new = []
        for k in range(o):
            for j in range(n):
                for i in range(m):
                    d = st - (i + 2 * j + 4 * k)
                    if d >= 0 and d % lag == 0 and d // lag < nsw:
                        new.append((k, j, i, gs_value(A, x, r, i, j, k)))
        for k, j, i, v in new:
            x[k, j, i] = v
    return x


@pytest.mark.parametrize("dims,nsw", [((9, 9, 9), 4), ((5, 7, 6), 3), ((8, 3, 10), 5)])
def test_pipelined_hyperplanes_equal_lexicographic_sweeps(dims, nsw):
    m, n, o = dims
    A = O.gen_operator(3, 70, m, n, o)  # 27-point Q1: corner neighbours too
    rng = np.random.default_rng(nsw)
    x0 = rng.uniform(-1, 1, size=(o, n, m))
    r = rng.uniform(-1, 1, size=(o, n, m))
    ref = x0.copy()
    for _ in range(nsw):
        ref = O.gs(A, ref, r, 0)
    assert np.array_equal(pipelined_hyperplanes(A, x0, r, nsw), ref)
    # seven hyperplanes apart is too close: the corner neighbour (1, 1, 1),
    # seven planes later, would be read before its previous-sweep update
    assert not np.array_equal(pipelined_hyperplanes(A, x0, r, nsw, lag=7), ref)
