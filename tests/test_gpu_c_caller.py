"""A plain C program (tests/c_caller/cpsd_caller.c, built by
__graft_entry__.build()) calls the drop-in symbols cpsd(), rc_fft3f_ and
cr_fft3b_ with its own prototypes on the reference's Example0 inputs, as the
reference's driver would after a relink; its energies are checked against
the oracle (1e-9 Ha), its FFT round trip against N^3 times the identity."""
import os
import subprocess
import sys

import numpy as np
import pytest

from example_inputs import GOLDEN, example0_oracle_state
import oracle_lib as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "c_caller", "cpsd_caller")


def test_c_caller_example0(gpu):
    print("c caller example0", flush=True)
    sys.stdout.flush()
    assert os.path.exists(EXE), "build the C caller first (__graft_entry__.build())"
    d = os.path.join(GOLDEN, "example0")
    r = subprocess.run([EXE, os.path.join(d, "cpsdv3.d"), os.path.join(d, "c2.ion00"),
                        os.path.join(O.DATA, "psp"), os.path.join(O.DATA, "PSI0.DAT"), "3"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [l.split() for l in r.stdout.splitlines()]
    calls = [l for l in lines if l and l[0] == "call"]
    assert len(calls) == 3
    ctl, st = example0_oracle_state()
    for row in calls:
        vals = np.array([float(v) for v in row[2:]])
        Er, der, dcr = st.step(ctl["inner"], ctl["dt"], ctl["fmass"])
        assert np.max(np.abs(vals[:10] - Er)) < 1e-9
        assert abs(vals[10] - der) < 1e-9
        assert abs(vals[11] - dcr) < 1e-9 * max(1.0, abs(dcr))
    fft = [l for l in lines if l[:2] == ["fft", "roundtrip"]]
    assert fft and float(fft[0][2]) < 1e-13
