"""The two largest configurations at their full sizes, checked through
properties that need no oracle run at that size:
- 513^3 Q1 finite elements (27-point operator, per-cell anisotropic
  tensors): the mixed-precision and FP64 colour solves both reach 1e-10; the
  residual each reports is the l2 residual of the x it returns; the mixed
  iterate agrees with the FP64 one (defect correction converges to the FP64
  fixed point).
- 256 independent problems (axes from {33, 65, 129}, log-uniform face
  coefficients, the bench's set) in one batch, mixed precision to 1e-8: every
  slot converges, and one slot of each size equals its lone solve bit for bit
  (cycle count and x)."""
import numpy as np
import pytest

import pgmg_amd as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_q1_513_mixed_and_fp64():
    n, tol = 513, 1e-10
    xs = {}
    for prec, name in [(P.PREC_FP64, "fp64"), (P.PREC_MIXED, "mixed")]:
        say("513^3 Q1", name)
        h = P.Hierarchy(n, n, n, precision=prec, smoother=P.SMOOTH_COLOUR8)
        h.generate_problem(P.FAM_Q1_ANISO, 777)
        h.compute_rats()
        c, hist = h.solve(tol, 100)
        bn = h.rhs_l2norm()
        rel = (hist[-1] / bn) ** 0.5
        # the reported residual is that of the iterate left in x
        assert abs(h.l2_error_all() - hist[-1]) <= 1e-12 * hist[-1]
        say(name, "cycles", c, "rel", rel)
        assert rel < tol and c <= 20
        xs[name] = h.get_x(0)
        h.close()
    err = np.linalg.norm(xs["mixed"] - xs["fp64"]) / np.linalg.norm(xs["fp64"])
    say("mixed vs fp64", err)
    assert err < 1e-12


def _bench_sizes(count=256):
    rng = np.random.default_rng(2024)  # bench.py's batch_256
    return [tuple(int(v) for v in rng.choice([33, 65, 129], size=3)) for _ in range(count)]


def test_batch_256_full_size():
    tol, maxc = 1e-8, 1000
    sizes = _bench_sizes()
    need = sum(P.footprint(m, n, o, precision=P.PREC_MIXED) for (m, n, o) in sizes)
    with P.Batch(int(need * 1.05) + (1 << 30), max_slots=len(sizes)) as b:
        slots = [b.admit(m, n, o, P.FAM_FACE7_LOGUNI, 5000 + q, tol, maxc) for q, (m, n, o) in enumerate(sizes)]
        assert None not in slots
        steps = b.run(max_steps=maxc + 5)
        st = [b.status(s) for s in slots]
        say("steps", steps, "tiles", b.tiles())
        assert all(s[0] == P.SLOT_CONVERGED and s[2] < tol for s in st)
        # one slot of each size against its lone solve
        picks = {}
        for q, d in enumerate(sizes):
            picks.setdefault(max(d), q)
        for q in sorted(picks.values()):
            m, n, o = sizes[q]
            h = P.Hierarchy(m, n, o, precision=P.PREC_MIXED, smoother=P.SMOOTH_COLOUR8, bottom=P.BOTTOM_GS24)
            h.generate_problem(P.FAM_FACE7_LOGUNI, 5000 + q)
            h.compute_rats()
            c, hist = h.solve(tol, maxc, check=False)
            say("slot", q, (m, n, o), "cycles", st[q][1], "alone", c)
            assert c == st[q][1]
            assert np.array_equal(b.get_x(slots[q]), h.get_x(0))
            h.close()
