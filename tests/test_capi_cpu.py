"""The native library loads, exports every function include/*.h declares, and
its host-only entry points agree with the oracle.  No GPU calls."""
import ctypes
import os
import re

import pytest

import pgmg_amd
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for mt in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*[\s\*]+(pgmg_[A-Za-z0-9_]+)\s*\(", text, re.M):
            if "static" in mt.group(0) or "PGMG_HD" in mt.group(0):
                continue  # header-only inline helpers, not library exports
            names.add(mt.group(1))
    return sorted(names)


def test_library_exports_all_declared_symbols():
    assert os.path.exists(pgmg_amd.LIB_PATH), "build the library first"
    L = ctypes.CDLL(pgmg_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_defaults():
    assert pgmg_amd.lib().pgmg_version() == 2
    o = pgmg_amd.default_opts()
    assert (o.precision, o.smoother, o.bottom, o.v_down, o.v_up, o.v_bottom) == (0, 1, 0, 2, 2, 20)


@pytest.mark.parametrize("dims", [(65, 65, 65), (257, 257, 257), (513, 513, 513), (20, 17, 30), (33, 65, 129), (9, 9, 9)])
def test_hierarchy_shape_matches_oracle(dims):
    assert pgmg_amd.hierarchy_shape(*dims) == O.hierarchy(*dims)


@pytest.mark.parametrize("dims,prd", [((64, 64, 64), 7), ((33, 33, 33), 1), ((20, 17, 30), 6), ((13, 10, 11), 5),
                                      ((256, 256, 256), 7), ((100, 2, 300), 3)])
def test_periodic_hierarchy_shape_matches_oracle(dims, prd):
    assert pgmg_amd.hierarchy_shape(*dims, prd=prd) == O.hierarchy(*dims, prd)


def test_footprint_scales():
    f257 = pgmg_amd.footprint(257, 257, 257)
    f257m = pgmg_amd.footprint(257, 257, 257, precision=pgmg_amd.PREC_MIXED)
    assert 3.5e9 < f257 < 6e9
    assert f257m > f257 * 1.4
    assert pgmg_amd.footprint(513, 513, 513, precision=pgmg_amd.PREC_MIXED) < 80e9


def test_footprint_counts_exact_skewed_copies():
    # the exact smoother's plane-skewed copies: 29 arrays of
    # R * (m + 2n + 4) * (o + 2) elements per level whose x (halo included)
    # does not fit the one workgroup's 64 KB of LDS in FP64, or that has more
    # than 1024 points (FP64, plus FP32 in mixed precision)
    def skew_elems(m, n, o):
        R = min((m + 1) // 2 + 2, n + 2)
        return R * (m + 2 * n + 4) * (o + 2)
    big = [d for d in pgmg_amd.hierarchy_shape(257, 257, 257)
           if (d[0] + 2) * (d[1] + 2) * (d[2] + 2) * 8 > 65536 or d[0] * d[1] * d[2] > 1024]
    extra = sum(29 * skew_elems(*d) for d in big)
    for prec, word in [(pgmg_amd.PREC_FP64, 8), (pgmg_amd.PREC_MIXED, 12)]:
        base = pgmg_amd.footprint(257, 257, 257, precision=prec)
        exact = pgmg_amd.footprint(257, 257, 257, precision=prec, smoother=pgmg_amd.SMOOTH_EXACT)
        assert exact - base == extra * word
