"""The colour pass's filtered Kern-Frenkel test (signed squares against a
band, q's side in its body frame) against the reference arithmetic, on 2^20
pairs whose patch cosines sit within 1e-15 to 1e-9 of cosmax on the p side,
the q side or both (tests/gpu_checks/kf_filter_check.hip, built by
__graft_entry__.build()).  Every pair the filter decides must match the exact
decision, and some pairs must fall inside the band, so the fallback is
exercised."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "patchyparticles_amd", "bin", "kf_filter_check")


@pytest.mark.gpu
def test_filtered_pair_test_at_the_thresholds():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built: run __graft_entry__.build()")
    res = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    m = re.search(r"pairs (\d+) certain (\d+) band (\d+) mismatches (\d+)", res.stdout)
    assert m, res.stdout + res.stderr
    pairs, certain, band, bad = map(int, m.groups())
    assert bad == 0, res.stdout
    assert band > 0 and certain + band == pairs, res.stdout
    # three quarters of the pairs are placed at a threshold, about half of
    # those inside the band (|cos - cosmax| below ~5e-13)
    assert certain > 0.4 * pairs, res.stdout
    assert res.returncode == 0, res.stdout + res.stderr
