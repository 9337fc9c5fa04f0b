"""Heat-bath determinant selection and the deterministic PT2 correction on the
device.

``Selector`` is one sweep of the variational loop (src/solver/solver.h:435-495
with ChemSystem::find_connected_dets, src/chem/chem_system.cc:289-402) through
``shci_select``; ``GpuSystem`` keeps the det list and eps_tried_prev with the
interface of ``shci_amd.host.HostSystem`` (``select``, ``dets``, ``reset``,
``n_dets``), so the same variational loop runs with either.  ``pt_dtm`` is
the deterministic PT2 correction (src/solver/solver.h:637-764)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .hamiltonian import Hamiltonian, ShciError

_u64p = C.POINTER(C.c_uint64)
_dp = C.POINTER(C.c_double)


class Selector:
    """Selection over the integrals of ``ham`` (its device context).
    ``time_sym`` (default: ``ham.time_sym``) applies the reference's time_sym
    rules: canonical dets (up <= dn), sqrt2 on eps_min of a det with up != dn,
    canonical candidates (solver.h:447-455)."""

    def __init__(self, ham: Hamiltonian, time_sym=None):
        self.ham = ham
        self.ctx = ham.ctx
        self.time_sym = getattr(ham, "time_sym", False) if time_sym is None else bool(time_sym)
        self.last_ms = 0.0
        self.last_phases_ms = {}

    def select(self, up, dn, coefs, eps_var, eps_tried_prev):
        """One sweep: returns the new dets (up, dn), (up, dn) ascending.
        ``eps_tried_prev`` (float64, one per det) is updated in place."""
        import time

        up = np.ascontiguousarray(up, dtype=np.uint64)
        dn = np.ascontiguousarray(dn, dtype=np.uint64)
        coefs = np.ascontiguousarray(np.atleast_2d(coefs), dtype=np.float64)
        if coefs.shape[1] != len(up) or len(dn) != len(up):
            raise ShciError("coefs / dets length mismatch")
        if not (isinstance(eps_tried_prev, np.ndarray) and eps_tried_prev.dtype == np.float64
                and eps_tried_prev.flags.c_contiguous and len(eps_tried_prev) == len(up)):
            raise ShciError("eps_tried_prev must be a contiguous float64 array with one entry per det")
        self.ham.ensure_integrals()
        lib = self.ctx.lib
        n_new = C.c_uint64()
        t0 = time.perf_counter()
        self.ctx.check(lib.shci_select(self.ctx.handle, up.ctypes.data_as(_u64p), dn.ctypes.data_as(_u64p), len(up),
                                       coefs.ctypes.data_as(_dp), coefs.shape[0], float(eps_var),
                                       eps_tried_prev.ctypes.data_as(_dp), C.byref(n_new),
                                       _abi.SHCI_TIME_SYM if self.time_sym else 0), "shci_select")
        m = n_new.value
        nu = np.zeros(m, dtype=np.uint64)
        nd = np.zeros(m, dtype=np.uint64)
        self.ctx.check(lib.shci_select_result(self.ctx.handle, nu.ctypes.data_as(_u64p), nd.ctypes.data_as(_u64p), m),
                       "shci_select_result")
        self.last_ms = (time.perf_counter() - t0) * 1e3
        ph = (C.c_float * 4)()
        self.ctx.check(lib.shci_select_timings(self.ctx.handle, ph), "shci_select_timings")
        self.last_phases_ms = dict(setup=ph[0], walk=ph[1], unique=ph[2], output=ph[3])
        return nu, nd


class GpuSystem:
    """The variational det list with selection on the device: HostSystem's
    ``select(coefs, eps_var)``, ``dets()``, ``reset()`` and ``n_dets``."""

    def __init__(self, ham: Hamiltonian, hf_up: int, hf_dn: int):
        self.selector = Selector(ham)  # time_sym follows the Hamiltonian's
        self.hf_up, self.hf_dn = int(hf_up), int(hf_dn)
        self.reset()

    def reset(self):
        self._up = np.array([self.hf_up], dtype=np.uint64)
        self._dn = np.array([self.hf_dn], dtype=np.uint64)
        self._eps = np.full(1, 1.0e100)

    @property
    def n_dets(self) -> int:
        return len(self._up)

    def dets(self):
        return self._up.copy(), self._dn.copy()

    def set_dets(self, up, dn):
        self._up = np.ascontiguousarray(up, dtype=np.uint64).copy()
        self._dn = np.ascontiguousarray(dn, dtype=np.uint64).copy()
        self._eps = np.full(len(self._up), 1.0e100)

    def select(self, coefs, eps_var) -> int:
        """One selection sweep; appends the new dets, returns how many."""
        nu, nd = self.selector.select(self._up, self._dn, coefs, eps_var, self._eps)
        self._up = np.concatenate([self._up, nu])
        self._dn = np.concatenate([self._dn, nd])
        self._eps = np.concatenate([self._eps, np.full(len(nu), 1.0e100)])
        return len(nu)


def pt_dtm(ham: Hamiltonian, up, dn, coefs, eps_pt_dtm, energy_var, n_batches=0):
    """Deterministic PT2 correction of one state over the integrals of
    ``ham`` (Solver::get_energy_pt_dtm, src/solver/solver.h:637-764): returns
    (E_pt2, number of PT dets, device ms).  ``coefs``: that state's
    coefficients; ``energy_var``: its variational energy."""
    import time

    up = np.ascontiguousarray(up, dtype=np.uint64)
    dn = np.ascontiguousarray(dn, dtype=np.uint64)
    c = np.ascontiguousarray(coefs, dtype=np.float64).ravel()
    if len(c) != len(up) or len(dn) != len(up):
        raise ShciError("coefs / dets length mismatch")
    ham.ensure_integrals()
    ctx = ham.ctx
    e = C.c_double()
    m = C.c_uint64()
    t0 = time.perf_counter()
    ctx.check(ctx.lib.shci_pt_dtm(ctx.handle, up.ctypes.data_as(_u64p), dn.ctypes.data_as(_u64p), len(up),
                                  c.ctypes.data_as(_dp), float(eps_pt_dtm), float(energy_var), int(n_batches),
                                  C.byref(e), C.byref(m)), "shci_pt_dtm")
    return e.value, m.value, (time.perf_counter() - t0) * 1e3
