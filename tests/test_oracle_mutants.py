"""Planted mistakes: each oracle rule test must fail when the oracle's C
source carries the mistake it is meant to catch.  Every mutant is a copy of
oracle/raster_ref.c with one or two statements changed, built with the
oracle's own flags into a temporary directory, and the targeted tests run
against it in a child pytest (ORC_LIB names the mutant library); the child
must report failures.  Nothing here touches the product or the GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "raster_ref.c")
NOISE = os.path.join(ROOT, "oracle", "redwood_ref.c")
CFLAGS = ["-O1", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-w"]

# (name, [(original statement, mutated statement)], test file, -k expression)
MUTANTS = [
    # mip-mapped filtering (tests/test_oracle_mip.py)
    ("mip blend weights swapped",
     [("float x1 = g * t1[k];", "float x1 = f * t1[k];"),
      ("float x2 = f * t2[k];", "float x2 = g * t2[k];")], "test_oracle_mip.py", "lml"),
    ("mip level by floor(lambda)",
     [("float f = ceilf(lambda + 0.5f) - 1.0f;", "float f = floorf(lambda);")],
     "test_oracle_mip.py", "lmn"),
    ("lambda used for frac(lambda)", [("float f = lambda - fl;", "float f = lambda;")],
     "test_oracle_mip.py", "lml"),
    ("second mip level not advanced",
     [("sample_level(t, d1 + 1, layer, s, tc, lin, t2);", "sample_level(t, d1, layer, s, tc, lin, t2);")],
     "test_oracle_mip.py", "lml"),
    ("rho from the x axis only", [("float r2 = rx > ry ? rx : ry;", "float r2 = rx;")],
     "test_oracle_mip.py", "aniso"),
    ("rho as the smaller axis", [("float r2 = rx > ry ? rx : ry;", "float r2 = rx < ry ? rx : ry;")],
     "test_oracle_mip.py", "aniso"),
    ("v derivative terms dropped",
     [("rx = rx + rx2;", "rx = rx + 0.0f * rx2;"), ("ry = ry + ry2;", "ry = ry + 0.0f * ry2;")],
     "test_oracle_mip.py", "rot"),
    ("lambda = log2(rho^2)", [("lambda = lambda * 0.5f;", "lambda = lambda * 1.0f;")],
     "test_oracle_mip.py", "lml or lmn"),
    ("no magnification threshold for NEAREST_MIPMAP_*", [("    c = 0.5f;", "    c = 0.0f;")],
     "test_oracle_mip.py", "nml"),
    # coverage, snapping, depth, filtering (tests/test_oracle_rules.py)
    ("fill rule flipped", [("return dy < 0 || (dy == 0 && dx > 0);", "return dy > 0 || (dy == 0 && dx < 0);")],
     "test_oracle_rules.py", "watertight or tie_rule"),
    ("depth test LEQUAL", [("if (z < C->zbuf[pix]) {", "if (z <= C->zbuf[pix]) {")],
     "test_oracle_rules.py", "depth_less"),
    ("truncating snap", [("X[i] = (int32_t)rintf(xw * ORC_SUBPIX);", "X[i] = (int32_t)(xw * ORC_SUBPIX);")],
     "test_oracle_rules.py", "snapping"),
    ("far plane keeps outside vertices", [("case 1: return v->w - v->z;", "case 1: return v->w - v->z + v->w;")],
     "test_oracle_rules.py", "clipping"),
    ("bilinear half-texel offset dropped", [("float uu = u - 0.5f;", "float uu = u;")],
     "test_oracle_rules.py", "bilinear"),
    ("log2 series term off", [("s = s + (1.0f / 5.0f);", "s = s + (1.0f / 6.0f);")],
     "test_oracle_rules.py", "log2"),
]


def _build(tmp_path, edits):
    src = open(SRC).read()
    for old, new in edits:
        assert src.count(old) == 1, f"statement not found exactly once: {old!r}"
        src = src.replace(old, new)
    c = tmp_path / "raster_mut.c"
    c.write_text(src)
    so = tmp_path / "liborc_mut.so"
    subprocess.run(["gcc", *CFLAGS, "-shared", "-o", str(so), str(c), NOISE, "-lm"], check=True)
    return str(so)


@pytest.mark.parametrize("name,edits,target,expr", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(tmp_path, name, edits, target, expr):
    so = _build(tmp_path, edits)
    env = dict(os.environ, ORC_LIB=so)
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                          os.path.join(ROOT, "tests", target), "-k", expr],
                         env=env, capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert res.returncode == 1, (f"mutant '{name}' was not caught by {target} -k '{expr}' "
                                 f"(exit {res.returncode}):\n{res.stdout[-2000:]}")


def test_unmutated_copy_passes(tmp_path):
    # the harness itself: an unchanged copy passes the same selections
    so = _build(tmp_path, [])
    env = dict(os.environ, ORC_LIB=so)
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                          os.path.join(ROOT, "tests", "test_oracle_mip.py"),
                          os.path.join(ROOT, "tests", "test_oracle_rules.py")],
                         env=env, capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert res.returncode == 0, res.stdout[-2000:]
