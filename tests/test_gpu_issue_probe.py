"""The VALU issue-rate probe (tests/cpp/valu_issue_probe.hip) runs on the GPU
and reports rates in the range a CU can issue; the fp32 add issues faster
than the fp64 FMA (the cost model behind the ES kernel's arithmetic choices).
"""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build",
                     "valu_issue_probe")


def test_valu_issue_probe_rates(gpu):
    assert os.path.exists(PROBE), "build the library first (python -m __graft_entry__)"
    print("valu issue probe", flush=True)
    out = subprocess.run([PROBE], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rates = {}
    for line in out.stdout.splitlines():
        r = json.loads(line)
        rates[(r["instruction"], r["waves_per_cu"])] = r["winst_per_cu_clock"]
    assert len(rates) == 26
    assert all(0.3 < v < 2.5 for v in rates.values()), rates
    assert rates[("v_add_f32", 32)] > 1.2 * rates[("v_fma_f64", 32)], rates
