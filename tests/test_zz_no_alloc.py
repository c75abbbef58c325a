"""adm_vmult_noalloc, the entry point that allocates nothing (the plane
kernel's patch buffer in the library's static scratch array), against
adm_vmult on the same inputs: two calls in a row, the second on a smaller
mesh; a mesh whose patch buffer does not fit is refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from adamantine_amd.workloads import plain_large, random_temperature

pytestmark = pytest.mark.gpu

VMULT_TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _vmult(op, fn, T, t):
    import torch
    from adamantine_amd import thermal as TH
    op.set_time(t)
    src = torch.tensor(T, device="cuda")
    dst = op.initialize_dof_vector(7.0)  # stale contents must not matter
    TH.check(getattr(TH.load_library(), fn)(
        op.handle, TH._dev_ptr(src, op.m()), TH._dev_ptr(dst, op.m()), 0,
        TH._stream_ptr()))
    torch.cuda.synchronize()
    return dst.cpu().numpy()


@pytest.mark.parametrize("mode", ["plane", "line"])
def test_noalloc_matches_vmult(mode):
    from adamantine_amd.thermal import (LAUNCH_LINE_COLORED, LAUNCH_PLANE,
                                        ThermalOperatorDevice)
    launch = LAUNCH_PLANE if mode == "plane" else LAUNCH_LINE_COLORED
    for n_cells, seed in (((48, 24, 12), 11), ((20, 10, 6), 12)):
        print("no-alloc vmult:", mode, n_cells, "seed", seed, flush=True)
        p = plain_large(n_cells=n_cells)
        T = random_temperature(p, seed=seed, low=300.0, high=2000.0)
        ref_op = ThermalOperatorDevice(p)
        op = ThermalOperatorDevice(p)
        ref_op.set_launch_config(launch, 4)
        op.set_launch_config(launch, 4)
        for t in (0.0, 2e-4):
            want = _vmult(ref_op, "adm_vmult", T, t)
            got = _vmult(op, "adm_vmult_noalloc", T, t)
            assert rel(got, want) < VMULT_TOL, (n_cells, t, rel(got, want))
        lq0, pw0 = ref_op.qp_state()
        lq1, pw1 = op.qp_state()
        np.testing.assert_array_equal(pw1, pw0)
        np.testing.assert_array_equal(lq1, lq0)


def test_noalloc_refuses_a_mesh_too_large():
    import torch
    from adamantine_amd import thermal as TH
    from adamantine_amd.thermal import ThermalOperatorDevice
    # 23 x 27 patches x 32 layers x 315 nodes = 6.26 million doubles
    p = plain_large(n_cells=(160, 80, 32))
    op = ThermalOperatorDevice(p)
    src = op.initialize_dof_vector(300.0)
    dst = op.initialize_dof_vector(0.0)
    rc = TH.load_library().adm_vmult_noalloc(
        op.handle, TH._dev_ptr(src, op.m()), TH._dev_ptr(dst, op.m()), 0,
        TH._stream_ptr())
    assert rc == 3  # ADM_ERR_UNSUPPORTED
    assert b"scratch" in TH.load_library().adm_last_error()
    torch.cuda.synchronize()
    assert float(dst.abs().max()) == 0.0  # nothing was launched
