"""shci_spmv_noalloc (the allocation-free product) against the existing
entry point it varies, shci_spmv with SHCI_DEVICE_PTRS, through the C ABI:
two builds in a row, the second over a smaller det list (the scratch keeps
the first call's contents), both layouts, FP64 and FP32 stores.  The CSR
product has no atomics, so its results are bit-identical; the segment
product adds member sums with atomics (FP64 1e-13, FP32 1e-6 relative to the
vector norm).  The context's device memory must not change across the call,
and a time_sym build is refused."""
import ctypes as C

import numpy as np
import pytest

from shci_amd import _abi
from shci_amd import integrals as si


def _dets(n, n_up, n_dn, h1, eri, eps_list):
    from shci_amd.host import HostSystem

    host = HostSystem.from_arrays(n, n_up, n_dn, h1, eri, 0.0)
    coefs = np.ones((1, 1))
    for eps in eps_list:
        host.select(coefs, eps)
        coefs = np.full((1, host.n_dets), 1.0 / np.sqrt(host.n_dets))
    return host.dets()


def _products(ham, x, fp32):
    import torch

    lib, h = ham.ctx.lib, ham.ctx.handle
    flags = _abi.SHCI_FP32 if fp32 else 0
    xd = torch.tensor(x, dtype=torch.float64, device="cuda")
    y_ref = torch.empty_like(xd)
    y_na = torch.full_like(xd, float("nan"))
    torch.cuda.synchronize()
    ham.ctx.check(lib.shci_spmv(h, C.c_void_p(xd.data_ptr()), C.c_void_p(y_ref.data_ptr()),
                                flags | _abi.SHCI_DEVICE_PTRS), "shci_spmv")
    before = ham.ctx.memory()
    ham.ctx.check(lib.shci_spmv_noalloc(h, C.c_void_p(xd.data_ptr()), C.c_void_p(y_na.data_ptr()), flags),
                  "shci_spmv_noalloc")
    assert ham.ctx.memory() == before  # nothing allocated through the context
    return y_ref.cpu().numpy(), y_na.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
def test_noalloc_matches_device_spmv(layout):
    from shci_amd.hamiltonian import Hamiltonian

    n, nu, nd = 16, 4, 4
    _, h1, eri, _ = si.synthetic(n, seed=31)
    up, dn = _dets(n, nu, nd, h1, eri, [1e-3, 2e-4, 5e-5])
    sizes = [len(up), len(up) // 3]  # the second call on a smaller det list
    assert sizes[1] > 1000
    ham = Hamiltonian(n, h1, eri, 0.0, store="both", layout=layout)
    for call, m in enumerate(sizes):
        ham.clear()
        ham.update(up[:m], dn[:m])
        x = np.random.default_rng(100 + call).uniform(-1, 1, m)
        for fp32 in (False, True):
            y_ref, y_na = _products(ham, x, fp32)
            print(f"{layout} call {call}: {m} dets, fp32={fp32}, max |diff| {np.max(np.abs(y_na - y_ref)):.2e}",
                  flush=True)
            assert np.all(np.isfinite(y_na))
            if layout == "csr":
                np.testing.assert_array_equal(y_na, y_ref)
            else:
                tol = 1e-6 if fp32 else 1e-13
                assert np.linalg.norm(y_na - y_ref) <= tol * np.linalg.norm(y_ref)
    ham.clear()


@pytest.mark.gpu
def test_noalloc_refuses_time_sym():
    import torch

    from shci_amd.hamiltonian import Hamiltonian

    n, ne = 10, 3
    _, h1, eri, _ = si.synthetic(n, seed=5)
    up, dn = _dets(n, ne, ne, h1, eri, [1e-2])
    keep = up <= dn  # canonical form of the time-reversal pairs
    up, dn = up[keep], dn[keep]
    ham = Hamiltonian(n, h1, eri, 0.0, store="fp64", time_sym=True)
    ham.update(up, dn)
    xd = torch.zeros(len(up), dtype=torch.float64, device="cuda")
    yd = torch.zeros_like(xd)
    rc = ham.ctx.lib.shci_spmv_noalloc(ham.ctx.handle, C.c_void_p(xd.data_ptr()), C.c_void_p(yd.data_ptr()), 0)
    assert rc == -3 and b"time_sym" in ham.ctx.lib.shci_last_error(ham.ctx.handle)
    ham.clear()
