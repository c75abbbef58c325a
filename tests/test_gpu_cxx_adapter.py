"""Host C++ through include/adm_cxx.hpp: the demo program (tests/cxx/,
built by adamantine_amd/build.py) runs the plain 32x16x8 Q1 workload on the
device through the C-ABI; its vmult and its 20 forward-Euler steps agree
with the oracle at the parity bars, and with the Python binding of the
same library."""
import os
import subprocess

import numpy as np
import pytest

from adamantine_amd.problem import FORWARD_EULER
from adamantine_amd.workloads import plain_small
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
DEMO = os.path.join(os.path.dirname(__file__), "cxx", "adapter_demo")


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_cxx_adapter_demo(tmp_path):
    import torch
    from adamantine_amd.thermal import ThermalOperatorDevice
    assert os.path.exists(DEMO), "build the library first (it builds the demo)"
    out = str(tmp_path / "demo.bin")
    print("C++ adapter demo: 32x16x8 Q1, vmult + 20 FE steps", flush=True)
    r = subprocess.run([DEMO, out], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout.strip(), r.stderr.strip(), flush=True)
    assert r.returncode == 0
    raw = open(out, "rb").read()
    n = int(np.frombuffer(raw[:8], np.int64)[0])
    vals = np.frombuffer(raw[8:], np.float64)
    v, T = vals[:n], vals[n:2 * n]
    p = plain_small()
    assert n == p.n_dofs
    o = Oracle(p)
    T0 = np.full(n, 300.0)
    o.set_time(1e-4)
    assert rel(v, o.vmult(T0)) < 1e-12
    op = ThermalOperatorDevice(p)
    op.set_time(1e-4)
    dst = op.initialize_dof_vector()
    op.vmult(dst, torch.tensor(T0, device="cuda"))
    torch.cuda.synchronize()
    assert rel(v, dst.cpu().numpy()) < 1e-15
    want, t = Oracle(p).evolve(FORWARD_EULER, 0.0, 1e-6, 20, T0.copy())
    assert rel(T, want) < 1e-10
    assert T.max() > 300.0
