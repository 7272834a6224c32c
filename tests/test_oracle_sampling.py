"""The sampled CPU-baseline estimator (oracle/sample_timing.c).

bench.py reports the serial restatement's seconds per SD step from a bounded
sample of each component; these checks keep that estimate consistent with a
full oracle step at a size where the full step is cheap to time.
"""
import time

import numpy as np
import pytest

import oracle_lib as ol
from example_inputs import example0_oracle_state


def test_parts_add_up():
    p = ol.step_time_sampled(16, 4, 2, 8, lag_iters=2, inner=1, sample=1)
    parts = np.array([p[k] for k in ol.SAMPLE_PARTS[:-1]])
    assert np.all(parts >= 0.0)
    assert p["total"] > 0.0
    assert abs(parts.sum() - p["total"]) <= 1e-12 * max(1.0, p["total"])


def test_estimate_tracks_full_step():
    """Sampled estimate vs a timed full oracle step on the Example0 cell (C2,
    32^3, 4 orbitals): the estimator times the same loops, so the two agree
    to well inside 2x."""
    _, st = example0_oracle_state()
    st.step(1)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        st.step(1)
    full = (time.perf_counter() - t0) / reps
    nproj = int(sum(st.lmmax[k] for k in st.s.katm))
    est = ol.step_time_sampled(st.s.nfft, int(st.ne[0]), st.s.nion, nproj,
                               lag_iters=3, inner=1, sample=int(st.ne[0]))["total"]
    if full < 2e-3:
        pytest.skip("full step too short to time reliably (%.2g s)" % full)
    assert 0.5 < est / full < 2.0, (est, full)
