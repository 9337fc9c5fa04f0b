"""ctypes bindings of libshci_host.so: integrals (FCIDUMP semantics of
src/chem/integrals.cc) and the CPU heat-bath selection
(src/chem/chem_system.cc:74-402, src/solver/solver.h:435-495), which stays on
the CPU and feeds determinant lists to the GPU path."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native

_u64p = C.POINTER(C.c_uint64)
_dp = C.POINTER(C.c_double)


def _lib():
    return _native.host_lib()


class HostSystem:
    """Integrals + the variational determinant list + selection state."""

    def __init__(self, handle, n_up, n_dn, n_states=1):
        lib = _lib()
        if not handle:
            raise RuntimeError("shci_host: " + lib.shci_host_last_error().decode())
        self._h = handle
        self.n_up, self.n_dn, self.n_states = n_up, n_dn, n_states
        n = C.c_uint()
        core = C.c_double()
        up = C.c_uint64()
        dn = C.c_uint64()
        sz = C.c_size_t()
        ehf = C.c_double()
        lib.shci_host_system_info(self._h, C.byref(n), C.byref(core), C.byref(up), C.byref(dn), C.byref(sz),
                                  C.byref(ehf))
        self.n_orbs, self.core, self.hf_up, self.hf_dn = n.value, core.value, up.value, dn.value
        self.eri_size, self.energy_hf = sz.value, ehf.value
        lib.shci_host_reset_dets(self._h)

    @classmethod
    def from_arrays(cls, n, n_up, n_dn, h1, eri, core=0.0, n_states=1, hf=None):
        h1 = np.ascontiguousarray(h1, dtype=np.float64)
        eri = np.ascontiguousarray(eri, dtype=np.float64)
        hf_up, hf_dn = hf if hf is not None else ((1 << n_up) - 1, (1 << n_dn) - 1)
        h = _lib().shci_host_system_from_arrays(n, n_up, n_dn, h1.ctypes.data_as(_dp), eri.ctypes.data_as(_dp),
                                                core, hf_up, hf_dn)
        return cls(h, n_up, n_dn, n_states)

    @classmethod
    def from_fcidump(cls, path, n_up, n_dn, occs_up=(), occs_dn=(), reorder=-1, n_states=1):
        ou = (C.c_uint * max(1, len(occs_up)))(*occs_up)
        od = (C.c_uint * max(1, len(occs_dn)))(*occs_dn)
        h = _lib().shci_host_system_from_fcidump(str(path).encode(), n_up, n_dn, ou, len(occs_up), od,
                                                 len(occs_dn), reorder)
        return cls(h, n_up, n_dn, n_states)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib().shci_host_system_free(self._h)
            self._h = None

    def integrals(self):
        n = self.n_orbs
        h1 = np.zeros((n, n))
        eri = np.zeros(self.eri_size)
        order = (C.c_uint * n)()
        _lib().shci_host_system_integrals(self._h, h1.ctypes.data_as(_dp), eri.ctypes.data_as(_dp), order)
        return h1, eri, np.array(order[:], dtype=np.int64)

    @property
    def n_dets(self) -> int:
        return int(_lib().shci_host_n_dets(self._h))

    def dets(self):
        n = self.n_dets
        up = np.zeros(n, dtype=np.uint64)
        dn = np.zeros(n, dtype=np.uint64)
        _lib().shci_host_get_dets(self._h, up.ctypes.data_as(_u64p), dn.ctypes.data_as(_u64p))
        return up, dn

    def set_dets(self, up, dn):
        up = np.ascontiguousarray(up, dtype=np.uint64)
        dn = np.ascontiguousarray(dn, dtype=np.uint64)
        _lib().shci_host_set_dets(self._h, up.ctypes.data_as(_u64p), dn.ctypes.data_as(_u64p), len(up))

    def reset(self):
        _lib().shci_host_reset_dets(self._h)

    def run_variation(self, eps_vars, schedule, target_error=5e-5, n_states=1, mixed=False, gpu_select=False,
                      device=0, verbose=False, wf_files=False, force_var=False):
        """The reference's variational loop in C++ (Solver::run_all_variations,
        solver.h:338-397) through the drop-in adapter: GPU build and Davidson,
        selection by the host queue walk or on the device.  ``wf_files``: load
        wf_eps1_<eps>.dat from the working directory instead of running an
        eps_var whose file exists (unless ``force_var``), and save it after
        running one (solver.h:352-375).  Returns the final variational
        energies; the det list stays here (``dets()``)."""
        ev = np.ascontiguousarray(eps_vars, dtype=np.float64)
        sc = np.ascontiguousarray(schedule, dtype=np.float64)
        out = np.zeros(n_states)
        mode = (1 if mixed else 0) | (2 if gpu_select else 0) | (4 if wf_files else 0) | (8 if force_var else 0)
        rc = _lib().shci_host_run_variation(self._h, ev.ctypes.data_as(_dp), len(ev), sc.ctypes.data_as(_dp),
                                            len(sc), target_error, n_states, mode, device, int(verbose),
                                            out.ctypes.data_as(_dp))
        if rc != 0:
            raise RuntimeError("shci_host: " + _lib().shci_host_last_error().decode())
        return out

    def select(self, coefs, eps_var) -> int:
        """One selection sweep; appends new dets, returns how many were added."""
        coefs = np.atleast_2d(np.ascontiguousarray(coefs, dtype=np.float64))
        r = _lib().shci_host_select(self._h, coefs.ctypes.data_as(_dp), coefs.shape[0], eps_var)
        if r < 0:
            raise RuntimeError("shci_host: " + _lib().shci_host_last_error().decode())
        return int(r)


def _orb_order_arg(n_orbs, orb_order):
    if orb_order is None:
        return None, C.cast(None, C.POINTER(C.c_uint))
    o = np.ascontiguousarray(orb_order, dtype=np.uint32)
    if o.shape != (n_orbs,):
        raise ValueError(f"orb_order must have {n_orbs} entries")
    return o, o.ctypes.data_as(C.POINTER(C.c_uint))


def dump_1rdm(path, rdm1, orb_order=None):
    """RDM::dump_1rdm (rdm.cc:202-221): ``1rdm.csv`` from a spin-summed n x n
    1-RDM (``Hamiltonian.rdm``), orbitals in the original order given by
    ``orb_order`` (``HostSystem.integrals()[2]``; identity if None)."""
    r1 = np.ascontiguousarray(rdm1, dtype=np.float64)
    n = r1.shape[0]
    keep, op = _orb_order_arg(n, orb_order)
    if _lib().shci_host_dump_1rdm(str(path).encode(), n, op, r1.ctypes.data_as(_dp)) != 0:
        raise RuntimeError("shci_host_dump_1rdm: " + _lib().shci_host_io_last_error().decode())


def dump_2rdm(path, rdm2, n_orbs, orb_order=None, csv=True):
    """RDM::dump_2rdm (rdm.cc:271-331): ``2rdm.csv`` (csv=True) or the
    Dice-convention ``spatialRDM.txt`` (csv=False) from the packed 2-RDM of
    ``Hamiltonian.rdm``."""
    r2 = np.ascontiguousarray(rdm2, dtype=np.float64)
    nn = n_orbs * n_orbs
    if r2.shape != (nn * (nn + 1) // 2,):
        raise ValueError("rdm2 must be the packed 2-RDM of n_orbs orbitals")
    keep, op = _orb_order_arg(n_orbs, orb_order)
    if _lib().shci_host_dump_2rdm(str(path).encode(), n_orbs, op, r2.ctypes.data_as(_dp), 1 if csv else 0) != 0:
        raise RuntimeError("shci_host_dump_2rdm: " + _lib().shci_host_io_last_error().decode())


def wf_filename(eps_var) -> str:
    """Solver::get_wf_filename (solver.h:1415-1417)."""
    return "wf_eps1_%#.2e.dat" % eps_var


def save_wf(path, n_up, n_dn, up, dn, coefs, energy_hf, energy_var, time_sym=False):
    """Solver::save_variation_result (solver.h:1217-1224): the reference's
    wavefunction file (BaseSystem::serialize, base_system.h:125-133) of dets
    (up, dn) and coefs (n_states x n_dets)."""
    up = np.ascontiguousarray(up, dtype=np.uint64)
    dn = np.ascontiguousarray(dn, dtype=np.uint64)
    c = np.ascontiguousarray(np.atleast_2d(coefs), dtype=np.float64)
    ev = np.ascontiguousarray(np.atleast_1d(energy_var), dtype=np.float64)
    if len(dn) != len(up) or c.shape[1] != len(up) or len(ev) != c.shape[0]:
        raise ValueError("dets, coefs and energy_var disagree in size")
    rc = _lib().shci_host_save_wf(str(path).encode(), n_up, n_dn, up.ctypes.data_as(_u64p), dn.ctypes.data_as(_u64p),
                                  len(up), c.ctypes.data_as(_dp), c.shape[0], float(energy_hf),
                                  ev.ctypes.data_as(_dp), 1 if time_sym else 0)
    if rc != 0:
        raise RuntimeError("shci_host_save_wf: " + _lib().shci_host_wf_last_error().decode())


def load_wf(path) -> dict:
    """Solver::load_variation_result (solver.h:1179-1215): the fields of a
    wavefunction file as a dict (n_up, n_dn, up, dn, coefs, energy_hf,
    energy_var, time_sym)."""
    lib = _lib()
    nu, nd, ns = C.c_uint(), C.c_uint(), C.c_uint()
    n = C.c_size_t()
    ts = C.c_int()
    p = str(path).encode()
    if lib.shci_host_wf_info(p, C.byref(nu), C.byref(nd), C.byref(n), C.byref(ns), C.byref(ts)) != 0:
        raise RuntimeError("shci_host_wf_info: " + lib.shci_host_wf_last_error().decode())
    up = np.zeros(n.value, dtype=np.uint64)
    dn = np.zeros(n.value, dtype=np.uint64)
    coefs = np.zeros((ns.value, n.value))
    ev = np.zeros(ns.value)
    ehf = C.c_double()
    if lib.shci_host_load_wf(p, up.ctypes.data_as(_u64p), dn.ctypes.data_as(_u64p), n.value,
                             coefs.ctypes.data_as(_dp), ns.value, C.byref(ehf), ev.ctypes.data_as(_dp)) != 0:
        raise RuntimeError("shci_host_load_wf: " + lib.shci_host_wf_last_error().decode())
    return dict(n_up=nu.value, n_dn=nd.value, up=up, dn=dn, coefs=coefs, energy_hf=ehf.value, energy_var=ev,
                time_sym=bool(ts.value))
