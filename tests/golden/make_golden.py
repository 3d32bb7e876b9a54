"""Writes the golden vectors in tests/golden/ by running the real PGMG
reference (built by `make -f oracle/Makefile.ref` into oracle/_ref/).

The reference's own test programs print and assert nothing, so these vectors
are the pins for parity: coarse (Galerkin) operators at every level, one
forward and one symmetric Gauss-Seidel sweep, a restriction, an
interpolation, V-cycle residual histories and iterates with the default
24-sweep bottom and with the bottom LU enabled, on odd, even and non-cubic
grids of all three synthetic problem families, non-periodic and with
periodic axes.  Inputs are regenerated from
(family, seed, m, n, o) by include/pgmg_problems.h; the files keep SHA-256
digests of the operator and source the reference actually received.

Run from the repository root in a container that has /root/reference and
/opt/conda (MPICH, MKL):  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = os.path.join(ROOT, "oracle", "_ref", "pgmg_ref")
OUT = os.path.join(ROOT, "tests", "golden")

# (name, family, seed, m, n, o)
FIXTURES = [
    ("f7u_17x17x17", 1, 11, 17, 17, 17),
    ("f7u_20x17x30", 1, 12, 20, 17, 30),
    ("f7u_16x16x16", 1, 13, 16, 16, 16),
    ("f7u_9x9x9", 1, 17, 9, 9, 9),
    ("f7log_33x20x11", 2, 14, 33, 20, 11),
    ("q1_18x13x22", 3, 15, 18, 13, 22),
    ("q1_12x12x12", 3, 16, 12, 12, 12),
]
# Periodic grids: (name, family, seed, m, n, o, prd), prd the periodic axes
# as a bit mask (1 = x, 2 = y, 4 = z).  The same records as FIXTURES.
PFIXTURES = [
    ("prd_f7u_13x10x11_p5", 1, 5, 13, 10, 11, 5),
    ("prd_q1_12x18x9_p3", 3, 7, 12, 18, 9, 3),
    ("prd_f7log_20x17x30_p6", 2, 9, 20, 17, 30, 6),
    ("prd_q1_10x14x12_p7", 3, 4, 10, 14, 12, 7),
    ("prd_f7u_16x16x16_p7", 1, 5, 16, 16, 16, 7),
    # a periodic axis of size 2 on a restricting level (two on (261, 2, 2))
    ("prd_f7u_18x29x2_p4", 1, 6, 18, 29, 2, 4),
    ("prd_q1_2x24x29_p1", 3, 8, 2, 24, 29, 1),
    ("prd_f7log_520x4x4_p6", 2, 10, 520, 4, 4, 6),
]
# (name, family, seed, m, n, o, bottom, tol)
SOLVES = [
    ("solve_f7u_65_gs24", 1, 12345, 65, 65, 65, "gs24", 1e-10),
    ("solve_f7u_65_lu", 1, 12345, 65, 65, 65, "lu", 1e-10),
    ("solve_q1_33x29x40_gs24", 3, 77, 33, 29, 40, "gs24", 1e-10),
]


# Full-size solves kept as summaries (the iterate is too large to store):
# every cycle's sum of squared residuals, the SHA-256 of the final x and a
# few of its values.  (name, family, seed, m, n, o, bottom, tol)
SUMMARIES = [
    ("summary_f7u_257_gs24", 1, 12345, 257, 257, 257, "gs24", 1e-10),
]
SAMPLE_POINTS = [0, 1, 257 * 128 + 77, 257 * 257 * 128 + 257 * 128 + 128, 257 ** 3 - 1]


# One fmg_cycle from x = 0, then one V-cycle: (name, family, seed, m, n, o,
# prd).
FMGS = [
    ("fmg_f7u_20x17x30", 1, 9, 20, 17, 30, 0),
    ("fmg_q1_33x29x40", 3, 21, 33, 29, 40, 0),
    ("fmg_prd_f7u_13x10x11_p5", 1, 9, 13, 10, 11, 5),
]


# The reference's diagnostic_S / diagnostic_A / diagnostic_x of every level
# (x after two V-cycles): (name, family or "simple", seed, m, n, o, prd).
# "simple" is the reference's own problem_simple (boundary boxes 1x1x1).
DIAGS = [
    ("diag_f7u_13x10x11", "1", 5, 13, 10, 11, 0),
    ("diag_simple_14x12x10", "simple", 0, 14, 12, 10, 0),
    ("diag_simple_18x9x21", "simple", 0, 18, 9, 21, 0),
    ("diag_q1_18x13x22", "3", 15, 18, 13, 22, 0),
    ("diag_prd_f7u_13x10x11_p5", "1", 5, 13, 10, 11, 5),
    ("diag_prd_simple_12x10x14_p3", "simple", 0, 12, 10, 14, 3),
    ("diag_prd_q1_10x14x12_p7", "3", 4, 10, 14, 12, 7),
    # three levels (boxes only): coarse boxes from coarse boxes
    ("diag_simple_30x20x25", "simple", 0, 30, 20, 25, 0),
    ("diag_prd_simple_28x22x17_p5", "simple", 0, 28, 22, 17, 5),
]
DIAG_BOXES_ONLY = ("diag_simple_30x20x25", "diag_prd_simple_28x22x17_p5")


# z-slice files written by the reference's output_x / output_r /
# output_residual after two V-cycles: (name, family, seed, m, n, o, k).
SLICES = [
    ("slices_f7u_20x17x30", 1, 7, 20, 17, 30, 11),
    ("slices_q1_18x13x22", 3, 3, 18, 13, 22, 9),
]


def read_records(path):
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        name = data[pos:pos + 32].split(b"\0")[0].decode()
        dtype, count = struct.unpack_from("<iq", data, pos + 32)
        pos += 44
        nbytes = count * (8 if dtype == 0 else 4)
        arr = np.frombuffer(data[pos:pos + nbytes], dtype=np.float64 if dtype == 0 else np.int32).copy()
        pos += nbytes
        out[name] = arr
    return out


def run_text(args):
    """The reference's standard output, without its constructor's '#' lines."""
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = "/opt/conda/lib"
    p = subprocess.run([REF] + [str(a) for a in args], capture_output=True, text=True, env=env, timeout=600)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("reference run failed: %s" % args)
    return "".join(l + "\n" for l in p.stdout.splitlines() if not l.startswith("#"))


def run(args):
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = "/opt/conda/lib"
    p = subprocess.run([REF] + [str(a) for a in args], capture_output=True, text=True, env=env, timeout=1800)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("reference run failed: %s" % args)
    return json.loads(p.stdout.strip().splitlines()[-1])


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    if not os.path.exists(REF):
        raise SystemExit("build the reference first: make -f oracle/Makefile.ref")
    only = set(sys.argv[1:])  # optional: names of the fixtures to (re)write

    def want(name):
        return not only or name in only
    with tempfile.TemporaryDirectory() as tmpdir:
        write_all(want, tmpdir)


def write_all(want, tmpdir):
    tmp = os.path.join(tmpdir, "records.bin")
    for name, fam, seed, m, n, o, prd in [f + (0,) for f in FIXTURES] + PFIXTURES:
        if not want(name):
            continue
        run(["fixture", "%d:%d" % (fam, prd) if prd else fam, seed, m, n, o, tmp])
        rec = read_records(tmp)
        dims = rec["dims"].reshape(-1, 3)
        meta = [fam, seed, m, n, o] + ([prd] if prd else [])
        save = {"meta": np.array(meta, dtype=np.int64), "dims": dims,
                "aexact_after_ctor": rec["aexact_after_ctor"]}
        for k, v in rec.items():
            if k in ("dims", "aexact_after_ctor"):
                continue
            if k == "A0":
                save["sha_A0"] = np.array(sha(v.reshape(o, n, m, 27)))
                continue
            if k == "b":
                save["sha_b"] = np.array(sha(v.reshape(o, n, m)))
                continue
            if k.startswith("S"):
                # box bounds: keep only the min/max over the level (the
                # [-1,1]^3 claim) to stay small
                s = v.reshape(-1, 6)
                save[k + "_range"] = np.array([s.min(axis=0), s.max(axis=0)])
                continue
            save[k] = v
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **save)
        print(name, "levels", dims.tolist(), flush=True)
    for name, fam, seed, m, n, o, bottom, tol in SUMMARIES:
        if not want(name):
            continue
        info = run(["solve", fam, seed, m, n, o, bottom, tol, 100, tmp])
        rec = read_records(tmp)
        x = rec["x"]
        np.savez_compressed(os.path.join(OUT, name + ".npz"),
                            meta=np.array([fam, seed, m, n, o], dtype=np.int64), tol=np.array(tol),
                            res=rec["res"], sha_x=np.array(sha(x)), sample_idx=np.array(SAMPLE_POINTS, dtype=np.int64),
                            sample_x=x[SAMPLE_POINTS], rhs_l2norm=rec["rhs_l2norm"], info=np.array(json.dumps(info)))
        print(name, info, flush=True)
    for name, fam, seed, m, n, o, bottom, tol in SOLVES:
        if not want(name):
            continue
        info = run(["solve", fam, seed, m, n, o, bottom, tol, 100, tmp])
        rec = read_records(tmp)
        np.savez_compressed(os.path.join(OUT, name + ".npz"),
                            meta=np.array([fam, seed, m, n, o], dtype=np.int64),
                            tol=np.array(tol), res=rec["res"], x=rec["x"], rhs_l2norm=rec["rhs_l2norm"],
                            info=np.array(json.dumps(info)))
        print(name, info, flush=True)
    for name, fam, seed, m, n, o, k in SLICES:
        if not want(name):
            continue
        prefix = os.path.join(tmpdir, "slice")
        run(["slices", fam, seed, m, n, o, k, prefix])
        save = {"meta": np.array([fam, seed, m, n, o, k], dtype=np.int64)}
        for field in ("x", "r", "res"):
            for lev in (0, 1):
                path = "%s_%s%d.bin" % (prefix, field, lev)
                save["%s%d" % (field, lev)] = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
                os.remove(path)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **save)
        print(name, flush=True)
    for name, fam, seed, m, n, o, prd in DIAGS:
        if not want(name):
            continue
        famarg = "%s:%d" % (fam, prd) if prd else fam
        save = {"meta": np.array([0 if fam == "simple" else int(fam), seed, m, n, o, prd], dtype=np.int64)}
        lev = 0
        while True:
            txt = run_text(["diag", famarg, seed, m, n, o, lev, "S"])
            if not txt:
                break
            save["S%d" % lev] = np.frombuffer(txt.encode(), dtype=np.uint8)
            if name in DIAG_BOXES_ONLY:
                lev += 1
                continue
            save["A%d" % lev] = np.frombuffer(run_text(["diag", famarg, seed, m, n, o, lev, "A"]).encode(), dtype=np.uint8)
            if not prd:
                save["x%d" % lev] = np.frombuffer(run_text(["diag", famarg, seed, m, n, o, lev, "x"]).encode(),
                                                  dtype=np.uint8)
            lev += 1
        save["levels"] = np.array(lev)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **save)
        print(name, "levels", lev, flush=True)
    for name, fam, seed, m, n, o, prd in FMGS:
        if not want(name):
            continue
        run(["fmg", "%d:%d" % (fam, prd) if prd else fam, seed, m, n, o, tmp])
        rec = read_records(tmp)
        save = {"meta": np.array([fam, seed, m, n, o, prd], dtype=np.int64)}
        save.update(rec)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **save)
        print(name, flush=True)


if __name__ == "__main__":
    main()
