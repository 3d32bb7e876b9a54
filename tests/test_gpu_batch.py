"""Batches and serving on the GPU: every slot's result equals the lone
hierarchy's bit for bit (tolerance 0), alone or in any mix; admission by arena
space; the open-loop serving loop completes every request."""
import sys

import numpy as np
import pytest

import pgmg_amd as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PROBLEMS = [
    # (m, n, o, family, seed)
    (17, 13, 22, P.FAM_Q1_ANISO, 3),
    (33, 33, 33, P.FAM_FACE7_UNIFORM, 4),
    (20, 17, 30, P.FAM_FACE7_UNIFORM, 5),
    (40, 21, 9, P.FAM_Q1_ANISO, 6),
    (9, 9, 9, P.FAM_FACE7_UNIFORM, 7),
    (65, 33, 17, P.FAM_FACE7_LOGUNI, 8),
]


def say(*a):
    print(*a, flush=True)


def alone(m, n, o, fam, seed, precision, bottom, tol, maxc):
    h = P.Hierarchy(m, n, o, precision=precision, smoother=P.SMOOTH_COLOUR8, bottom=bottom)
    h.generate_problem(fam, seed)
    h.compute_rats()
    c, hist = h.solve(tol, maxc, check=False)
    x = h.get_x(0)
    say("alone", (m, n, o, fam, seed), "cycles", c, "lu_active", h.lu_active())
    h.close()
    return c, x


@pytest.mark.parametrize("precision", [P.PREC_MIXED, P.PREC_FP64])
@pytest.mark.parametrize("bottom", [P.BOTTOM_GS24, P.BOTTOM_LU])
def test_batch_equals_alone(precision, bottom):
    say("batch", precision, bottom)
    tol, maxc = 1e-8, 40
    ref = [alone(m, n, o, f, s, precision, bottom, tol, maxc) for (m, n, o, f, s) in PROBLEMS]
    with P.Batch(2 << 30, max_slots=16, precision=precision, bottom=bottom) as b:
        slots = [b.admit(m, n, o, f, s, tol, maxc) for (m, n, o, f, s) in PROBLEMS]
        assert None not in slots
        steps = b.run()
        stats = [b.status(s) for s in slots]
        say("steps", steps, "slots (state, cycles, rel):", stats, "alone cycles:", [r[0] for r in ref])
        for q, s in enumerate(slots):
            st, c, rel = stats[q]
            assert c == ref[q][0], (q, c, ref[q][0])
            assert np.array_equal(b.get_x(s), ref[q][1]), q
            b.release(s)


def test_alone_vs_mix_in_batch():
    """One request solved by itself in a batch and inside a mix of others
    admitted at different times: identical bits."""
    tol = 1e-8
    m, n, o, f, s = 24, 19, 31, P.FAM_Q1_ANISO, 99
    with P.Batch(1 << 30, max_slots=8) as b:
        sl = b.admit(m, n, o, f, s, tol)
        b.run()
        x_alone = b.get_x(sl)
        b.release(sl)
    with P.Batch(2 << 30, max_slots=8) as b:
        others = [b.admit(*p[:3], p[3], p[4], tol) for p in PROBLEMS[:3]]
        b.step()
        b.step()
        sl = b.admit(m, n, o, f, s, tol)  # joins while the others are mid-solve
        others += [b.admit(*p[:3], p[3], p[4], tol) for p in PROBLEMS[3:]]
        b.run()
        assert b.status(sl)[0] == P.SLOT_CONVERGED
        assert np.array_equal(b.get_x(sl), x_alone)


def test_host_operator_admission_matches_oracle_solution():
    m, n, o = 21, 18, 15
    A = O.gen_operator(P.FAM_FACE7_UNIFORM, 31, m, n, o)
    bvec = O.gen_rhs(31, m, n, o)
    mg = O.Multigrid(A, colour8=True)
    mg.set_rhs(bvec)
    c, hist = mg.solve(1e-9, 40, mixed=True)
    with P.Batch(1 << 30, max_slots=4) as b:
        s = b.admit_operator(A, bvec, 1e-9)
        b.run()
        st, cb, rel = b.status(s)
        assert st == P.SLOT_CONVERGED and cb == c
        assert np.array_equal(b.get_x(s), mg.get_x(0))


def test_admission_by_arena_space():
    fp = P.footprint(65, 65, 65, precision=P.PREC_MIXED)
    with P.Batch(int(fp * 2.5), max_slots=8) as b:
        s1 = b.admit(65, 65, 65, P.FAM_FACE7_UNIFORM, 1, 1e-8)
        s2 = b.admit(65, 65, 65, P.FAM_FACE7_UNIFORM, 2, 1e-8)
        assert s1 is not None and s2 is not None
        assert b.admit(65, 65, 65, P.FAM_FACE7_UNIFORM, 3, 1e-8) is None  # no room
        b.run()
        b.release(s1)
        s3 = b.admit(65, 65, 65, P.FAM_FACE7_UNIFORM, 3, 1e-8)
        assert s3 is not None
        b.run()
        assert b.status(s3)[0] == P.SLOT_CONVERGED
    with P.Batch(1 << 28, max_slots=1) as b:
        assert b.admit(9, 9, 9, 1, 1, 1e-8) is not None
        assert b.admit(9, 9, 9, 1, 2, 1e-8) is None  # no free slot


def test_serving_open_loop_small():
    rng = np.random.default_rng(5)
    reqs = []
    t = 0.0
    for q in range(24):
        t += rng.exponential(0.004)
        dims = [int(round(np.exp(rng.uniform(np.log(17), np.log(48))))) for _ in range(3)]
        reqs.append(dict(arrival_s=t, m=dims[0], n=dims[1], o=dims[2], family=P.FAM_Q1_ANISO, seed=1000 + q,
                         tol=1e-8, max_cycles=60))
    with P.Batch(4 << 30, max_slots=16) as b:
        res, wall = b.serve(reqs)
        assert b.free_bytes() == 4 << 30
    for r, q in zip(res, reqs):
        assert r["status"] == 0 and r["rel_residual"] < 1e-8
        assert r["admit_s"] >= q["arrival_s"] - 1e-9
        assert r["done_s"] > r["admit_s"]
    assert wall > 0


def test_serving_returns_each_solution_bitwise_as_alone():
    """Staggered requests of uneven sizes served by pgmg_serve_run, several
    on the GPU at once: each returned x equals the same request solved by a
    lone hierarchy bit for bit (tolerance 0), and one is also checked against
    the oracle's mixed-precision solve of the same problem."""
    rng = np.random.default_rng(11)
    reqs = []
    t = 0.0
    for q in range(20):
        t += rng.exponential(0.002)
        dims = [int(round(np.exp(rng.uniform(np.log(17), np.log(56))))) for _ in range(3)]
        fam = [P.FAM_Q1_ANISO, P.FAM_FACE7_UNIFORM][q % 2]
        reqs.append(dict(arrival_s=t, m=dims[0], n=dims[1], o=dims[2], family=fam, seed=2000 + q,
                         tol=1e-8, max_cycles=80))
    with P.Batch(4 << 30, max_slots=8) as b:
        res, wall = b.serve(reqs)
        assert b.free_bytes() == 4 << 30
    overlap = sum(1 for a in res for c in res if a is not c and a["admit_s"] < c["admit_s"] < a["done_s"])
    say("requests", len(reqs), "wall", wall, "admissions during another solve", overlap)
    assert overlap > 0
    for q, (r, rq) in enumerate(zip(res, reqs)):
        assert r["status"] == 0 and r["rel_residual"] < 1e-8, (q, r)
        c, x = alone(rq["m"], rq["n"], rq["o"], rq["family"], rq["seed"], P.PREC_MIXED, P.BOTTOM_GS24, 1e-8, 80)
        assert r["cycles"] == c, q
        assert np.array_equal(r["x"], x), q
    rq = reqs[1]
    A = O.gen_operator(rq["family"], rq["seed"], rq["m"], rq["n"], rq["o"])
    mg = O.Multigrid(A, colour8=True)
    mg.set_rhs(O.gen_rhs(rq["seed"], rq["m"], rq["n"], rq["o"]))
    c, hist = mg.solve(1e-8, 80, mixed=True)
    assert c == res[1]["cycles"]
    assert np.array_equal(res[1]["x"], mg.get_x(0))


@pytest.mark.parametrize("v_down,v_up", [(0, 2), (1, 3)])
def test_batch_mixed_sweep_counts_match_oracle(v_down, v_up):
    """Mixed precision with other sweep counts, v_down = 0 included (no
    level-0 down-sweep, so the FP32 correction must be cleared before the
    cycle): the batch against the oracle's defect correction."""
    m, n, o = 33, 29, 40
    A = O.gen_operator(P.FAM_FACE7_UNIFORM, 12, m, n, o)
    bvec = O.gen_rhs(12, m, n, o)
    mg = O.Multigrid(A, colour8=True, v_down=v_down, v_up=v_up)
    mg.set_rhs(bvec)
    c, hist = mg.solve(1e-30, 4, mixed=True)
    with P.Batch(1 << 30, max_slots=2, v_down=v_down, v_up=v_up) as b:
        s = b.admit_operator(A, bvec, 1e-30, max_cycles=4)
        b.run()
        assert b.status(s)[1] == c == 4
        assert np.array_equal(b.get_x(s), mg.get_x(0))


def test_serving_request_that_never_fits():
    """A request larger than the whole arena is reported (admission error)
    instead of waiting forever; the others are served."""
    fp = P.footprint(33, 33, 33, precision=P.PREC_MIXED)
    reqs = [dict(arrival_s=0.0, m=33, n=33, o=33, family=P.FAM_Q1_ANISO, seed=1, tol=1e-8, max_cycles=60),
            dict(arrival_s=0.0, m=129, n=129, o=129, family=P.FAM_Q1_ANISO, seed=2, tol=1e-8, max_cycles=60),
            dict(arrival_s=0.001, m=33, n=33, o=33, family=P.FAM_Q1_ANISO, seed=3, tol=1e-8, max_cycles=60)]
    with P.Batch(int(fp * 1.5), max_slots=4) as b:
        res, wall = b.serve(reqs)
        assert b.free_bytes() == int(fp * 1.5)
    assert res[0]["status"] == 0 and res[2]["status"] == 0
    assert res[1]["status"] == 10  # PGMG_ERR_ADMISSION


def test_batch_tiles_and_graph_keep_the_bits():
    """The trial steps of the tile shapes and the graph-captured cycles that
    follow give the bits of eager cycles at one unit per workgroup; a slot
    admitted after the graph is captured (the plan changes, the graph is
    dropped and captured again) still equals its lone solve."""
    tol, maxc = 1e-30, 30  # every slot runs its 30 cycles
    p0, p1 = PROBLEMS[1], PROBLEMS[3]
    ref0 = alone(*p0, P.PREC_MIXED, P.BOTTOM_GS24, tol, maxc)
    ref1 = alone(*p1, P.PREC_MIXED, P.BOTTOM_GS24, tol, maxc)
    with P.Batch(2 << 30, max_slots=4) as b:
        s0 = b.admit(*p0, tol, maxc)
        for _ in range(12):
            b.step()
        tiles, replays = b.tiles()
        say("tiles", tiles, "graph replays", replays)
        assert tiles is not None and set(tiles.values()) <= {1, 2, 4, 8}
        assert replays >= 1
        s1 = b.admit(*p1, tol, maxc)
        b.run()
        tiles2, replays2 = b.tiles()
        say("graph replays at the end", replays2)
        assert tiles2 == tiles and replays2 > replays
        for s, ref in ((s0, ref0), (s1, ref1)):
            st, c, rel = b.status(s)
            assert c == ref[0] == maxc, (s, c, ref[0])
            assert np.array_equal(b.get_x(s), ref[1]), s
    with P.Batch(1 << 30, max_slots=2, use_graph=False) as b:
        s0 = b.admit(*p0, tol, maxc)
        b.run()
        assert b.tiles()[1] == 0
        assert np.array_equal(b.get_x(s0), ref0[1])
