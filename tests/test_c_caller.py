"""The drop-in boundary from C: tests/c_caller/drop_in.c is written against
the reference's headers (include/alm.h, approx/approx.h, spmv/spmv.h) and
links with -lacdalm in place of the reference's objects.  Its checks follow
the reference's test mains (TEST_SPMV spmv/spmv.c:261-386, TEST_APPROX
approx/approx.c:710-798, TEST_ALM alm.c:366-395) with assertions."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "acd-alm_amd", "lib")
SRC = os.path.join(ROOT, "tests", "c_caller", "drop_in.c")


def build(tmp_path):
    exe = str(tmp_path / "drop_in")
    cmd = ["gcc", "-std=gnu11", "-O2", "-ffp-contract=off", "-Wall", "-Werror",
           "-I" + os.path.join(ROOT, "include"), SRC, "-L" + LIBDIR, "-lacdalm",
           "-lm", "-Wl,-rpath," + LIBDIR, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_caller_compiles_and_links(tmp_path):
    """The reference's prototypes, compiled as C and resolved against the
    library (no GPU needed)."""
    assert os.path.exists(os.path.join(LIBDIR, "libacdalm.so"))
    build(tmp_path)


@pytest.mark.gpu
@pytest.mark.parametrize("part", ["spmv", "approx", "alm"])
def test_c_caller_runs(tmp_path, part):
    exe = build(tmp_path)
    print(f"[c caller] {part}", flush=True)
    r = subprocess.run([exe, part], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "drop_in: ok" in r.stdout


def test_c_caller_passes_with_the_reference_build(tmp_path):
    """The same program linked with the reference compiled from its sources
    (oracle/_ref, built by the oracle's Makefile where the reference is
    present): the checks hold for the reference itself, so they test the
    drop-in property and not an accident of this library."""
    ref = os.path.join(ROOT, "oracle", "_ref", "libacdref_approx.so")
    if not os.path.exists(ref):
        pytest.skip("reference build absent (oracle/_ref)")
    exe = str(tmp_path / "drop_in_ref")
    cmd = ["gcc", "-std=gnu11", "-O2", "-ffp-contract=off", "-Wall", "-Werror",
           "-I" + os.path.join(ROOT, "include"), SRC, ref, "-lm",
           "-Wl,-rpath," + os.path.dirname(ref), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "drop_in: ok" in r.stdout
