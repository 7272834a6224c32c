"""CPU builds with AddressSanitizer and UndefinedBehaviorSanitizer (gcc, no
GPU code; tests/sanitize/Makefile): the host library's file formats and
orbital builders (tests/sanitize/host_io.c), and the plain C caller of the
drop-in symbols (tests/c_caller/cpsd_caller.c) linked against the host
library and the oracle, on the reference's Example0 inputs.  Each runs with
no sanitizer report and leak-free; a planted out-of-bounds read
(tests/sanitize/selfcheck.c) shows the instrumentation is live."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from example_inputs import GOLDEN, example0_oracle_state
import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
REPORTS = ("AddressSanitizer", "LeakSanitizer", "runtime error")


@pytest.fixture(scope="module")
def built():
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not installed")
    r = subprocess.run(["make", "-B", "-s", "-C", SAN], capture_output=True, text=True, timeout=300)
    if r.returncode != 0 and "sanitize" in r.stderr and "cannot find" in r.stderr:
        pytest.skip("gcc has no sanitizer runtime here: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr[-3000:]
    return SAN


def run(args, **kw):
    return subprocess.run(args, capture_output=True, text=True, env=ENV, timeout=300, **kw)


def clean(r):
    assert r.returncode == 0, r.stderr[-4000:]
    for rep in REPORTS:
        assert rep not in r.stderr, r.stderr[-4000:]


def test_instrumentation_is_live(built):
    r = run([os.path.join(built, "san_selfcheck")])
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stderr or "runtime error" in r.stderr


def test_host_io_under_sanitizers(built, tmp_path):
    r = run([os.path.join(built, "san_host_io"), str(tmp_path), os.path.join(O.DATA, "psp", "O00.dat")])
    clean(r)
    assert "host io ok" in r.stdout


def test_c_caller_over_oracle_under_sanitizers(built):
    """two calls of Example0's 10 steps through the drop-in's prototypes; the
    energies are the oracle's (the same C code at another optimisation level,
    no contraction: 1e-10 Ha)"""
    d = os.path.join(GOLDEN, "example0")
    r = run([os.path.join(built, "san_caller"), os.path.join(d, "cpsdv3.d"), os.path.join(d, "c2.ion00"),
             os.path.join(O.DATA, "psp"), os.path.join(O.DATA, "PSI0.DAT"), "2"])
    clean(r)
    calls = [l.split() for l in r.stdout.splitlines() if l.startswith("call")]
    assert len(calls) == 2
    ctl, st = example0_oracle_state()
    for row in calls:
        vals = np.array([float(v) for v in row[2:]])
        Er, der, _ = st.step(ctl["inner"], ctl["dt"], ctl["fmass"])
        assert np.max(np.abs(vals[:10] - Er)) < 1e-10
        assert abs(vals[10] - der) < 1e-10
    fft = [l.split() for l in r.stdout.splitlines() if l.startswith("fft roundtrip")]
    assert fft and float(fft[0][2]) < 1e-13
