"""The device's double-precision sin, cos, exp, log and sqrt, called as the
sampler calls them, against the host's C library on the ranges the sampler
feeds them (tests/gpu_checks/devmath_check.hip, built by
__graft_entry__.build()): the trial rotation angle (MC.c:321-325), the
Marsaglia axis factor (utils.c:115-131), |d| of the pair test (MC.c:244),
exp(-dE/T) beyond the acceptance table (MC.c:332) and the volume move's
ln V / exp (MC.c:474-499).

sqrt is correctly rounded on both sides, so it must agree exactly.  For sin,
cos, exp and log neither side promises correct rounding (the host's libm and
the device's are each within 1 ulp of the true value), so the two may differ
by up to 2 ulp; at 1e-16 relative that changes a proposal, not the law it is
drawn from."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "patchyparticles_amd", "bin", "devmath_check")
MAX_ULP = 2


@pytest.mark.gpu
def test_device_math_against_host_libm():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built: run __graft_entry__.build()")
    res = subprocess.run([EXE, str(MAX_ULP)], capture_output=True, text=True, timeout=120)
    rows = re.findall(r"fn (\w+) n (\d+) exact (\d+) maxulp (\d+)", res.stdout)
    assert {r[0] for r in rows} == {"sin", "cos", "sqrt", "exp", "log"}, res.stdout + res.stderr
    for name, n, exact, worst in rows:
        n, exact, worst = int(n), int(exact), int(worst)
        assert n == 1 << 20, res.stdout
        assert worst <= (0 if name == "sqrt" else MAX_ULP), res.stdout
        # most results are bit-identical to the host's
        assert exact > 0.9 * n, res.stdout
    assert res.returncode == 0, res.stdout + res.stderr
