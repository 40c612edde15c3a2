"""Both ULV factor kernels against the oracle: the per-reflector kernel (k_ulv_factor)
and the blocked MFMA kernel (k_ulv_factor_mf).  The library picks one per precision;
HSS_FACTOR_KERNEL forces either on every level it fits, so each runs in a process of its
own (factor_kernel_check.py) over the golden fixtures, C1 and an n = 4096 log-kernel matrix,
Float64 (parity bounds as test_gpu_parity) and Float32 (refined to 1e-10)."""
import json
import os
import pathlib
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = pathlib.Path(__file__).resolve().parent


@pytest.mark.parametrize("kernel,name", [("vector", "k_ulv_factor"), ("mfma", "k_ulv_factor_mf")])
def test_factor_kernel_parity(kernel, name):
    env = dict(os.environ, HSS_FACTOR_KERNEL=kernel)
    print(f"factor kernel {kernel}", flush=True)
    p = subprocess.run([sys.executable, str(HERE / "factor_kernel_check.py")], env=env, capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    out = json.loads(p.stdout.strip().splitlines()[-1])
    assert out["ok"]
    # every case ran the kernel asked for (golden cases with nodes wider than the blocked
    # kernel's limit may mix in the per-reflector kernel; C1 and the log matrix fit)
    for case in ("c1", "log4096", "c1/fp32", "log4096/fp32"):
        assert name in out["kernels"][case], out["kernels"]
    if kernel == "vector":
        assert all(ks == ["k_ulv_factor"] for ks in out["kernels"].values() if ks), out["kernels"]
