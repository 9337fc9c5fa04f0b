"""CPU oracle for the variational hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only as the checker (or the timed CPU
baseline).  The product path (``shci_amd``) never imports it and fails loudly
when its HIP library is missing.

``oracle/oracle.cc`` restates the reference algorithms; each function cites the
reference file:line it follows.  How each part is pinned (every check is
independent of the oracle's own formulas):

=====================================  ===================================================
oracle part (reference)                pinned by
=====================================  ===================================================
diff / permutation_factor              tests/test_oracle_dets.py: half_det_test.cc vectors;
(src/det/half_det.cc:120-206)          brute-force creation/annihilation sign on all pairs
                                       of strings with n <= 8
Slater-Condon elements                 tests/test_oracle_elements.py: Jordan-Wigner
(src/chem/chem_system.cc:409-609)      second-quantised H (dense, <= 12 spin orbitals)
                                       element by element; dense FCI eigenvalue
FCIDUMP read / keys / reorder          tests/test_integrals.py: synthetic round trip;
(src/chem/integrals.cc:37-422)         examples/C HF energy and orbital energies (log)
Hamiltonian::update (connections)      tests/test_oracle_hamiltonian.py: brute-force O(N^2)
(src/solver/hamiltonian.h:87-598)      pattern; examples/C nnz and absingles counts (log);
                                       incremental == from-scratch
SparseMatrix::mul                      dense product of the symmetric expansion
(src/solver/sparse_matrix.cc:28-56)
Davidson::diagonalize                  DavidsonTest.HilbertSystem known answer
(src/solver/davidson.cc:8-196)          (src/solver/davidson_test.cc:26-50); scipy eigh;
                                       examples/C per-iteration "Davidson #k" trajectory
RDM 1/2 (src/chem/rdm.cc)              trace / partial-trace identities; energy identity
                                       rdm.cc:731-745 vs the eigenvalue; finite differences
time_sym element                       tests/test_oracle_time_sym.py: equals P^T H P of the
(src/base_system.h:86-101)             full space's dense H (first principles); its spectrum
                                       is the spin-flip-even part of H's
selection sweep                        tests/test_oracle_selection.py: examples/C per-
(src/solver/solver.h:435-495,          iteration det counts of the reference's log with
src/chem/chem_system.cc:115-402)       this selection driving the loop; hand-built
                                       cases (a det whose only connections are known);
                                       the criteria's limits (eps -> 0 gives every det
                                       within two excitations; eps_tried_prev skips)
wavefunction file (wf_format.py;       tests/test_wf_files.py: the published varint
src/solver/solver.h:1179-1224,         examples; a file assembled byte by byte from the
src/base_system.h:125-133)             field list.  No file written by the reference
                                       exists in its checkout and its serializer (hps)
                                       is absent: beyond hps's published encoding the
                                       bytes are parity unpinned
=====================================  ===================================================

The reference itself cannot be compiled here (its submodules fgpl, hps, eigen,
json and googletest are empty), so there is no ``oracle/_ref`` build.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
_lib = None

_u64p = C.POINTER(C.c_uint64)
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
    lib = C.CDLL(str(_LIB_PATH))
    lib.orc_last_error.restype = C.c_char_p
    lib.orc_system_new.restype = C.c_void_p
    lib.orc_system_new.argtypes = [C.c_uint, C.c_uint, C.c_uint, _dp, _dp, C.c_double]
    lib.orc_system_from_fcidump.restype = C.c_void_p
    lib.orc_system_from_fcidump.argtypes = [C.c_char_p, C.c_uint, C.c_uint, C.POINTER(C.c_uint), C.c_uint,
                                            C.POINTER(C.c_uint), C.c_uint, C.c_int]
    lib.orc_system_free.argtypes = [C.c_void_p]
    lib.orc_system_info.argtypes = [C.c_void_p, C.POINTER(C.c_uint), _dp, _u64p, _u64p, C.POINTER(C.c_size_t)]
    lib.orc_system_integrals.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(C.c_uint)]
    lib.orc_system_set_integrals.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
    lib.orc_diff.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(C.c_int)]
    lib.orc_elem.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, _dp]
    lib.orc_ham_update.argtypes = [C.c_void_p, _u64p, _u64p, C.c_size_t]
    lib.orc_ham_clear.argtypes = [C.c_void_p]
    lib.orc_ham_update_existing.argtypes = [C.c_void_p]
    lib.orc_ham_stats.argtypes = [C.c_void_p, C.POINTER(C.c_size_t), _dp]
    lib.orc_ham_export.argtypes = [C.c_void_p, _i64p, _i64p, _dp]
    lib.orc_ham_diag.argtypes = [C.c_void_p, _dp]
    lib.orc_ham_mul.argtypes = [C.c_void_p, _dp, _dp]
    lib.orc_ham_time_mul.restype = C.c_double
    lib.orc_ham_time_mul.argtypes = [C.c_void_p, _dp, C.c_int]
    lib.orc_davidson.argtypes = [C.c_void_p, C.c_uint, _dp, C.c_double, _dp, _dp, C.POINTER(C.c_int), _dp,
                                 C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.orc_sm_new.restype = C.c_void_p
    lib.orc_sm_new.argtypes = [C.c_size_t]
    lib.orc_sm_free.argtypes = [C.c_void_p]
    lib.orc_sm_append.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_double]
    lib.orc_sm_davidson.argtypes = [C.c_void_p, C.c_uint, _dp, C.c_double, _dp, _dp, C.POINTER(C.c_int), _dp,
                                    C.c_int, C.POINTER(C.c_int)]
    lib.orc_rdm.argtypes = [C.c_void_p, C.c_uint, _dp, _dp, _dp]
    lib.orc_select.argtypes = [C.c_void_p, _u64p, _u64p, C.c_size_t, _dp, C.c_uint, C.c_double, _dp,
                               C.POINTER(C.c_size_t), C.c_int]
    lib.orc_select_result.argtypes = [C.c_void_p, _u64p, _u64p]
    lib.orc_elem_time_sym.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, _dp]
    lib.orc_pt_dtm.argtypes = [C.c_void_p, _u64p, _u64p, C.c_size_t, _dp, C.c_double, C.c_double, _dp,
                               C.POINTER(C.c_size_t)]
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + _load().orc_last_error().decode())


def diff(left: int, right: int):
    """``HalfDet::diff``: returns (n_diffs, left_only, right_only, permutation_factor)."""
    lib = _load()
    lo = (C.c_uint * 2)()
    ro = (C.c_uint * 2)()
    pf = C.c_int()
    nd = lib.orc_diff(C.c_uint64(left), C.c_uint64(right), lo, ro, C.byref(pf))
    return nd, (lo[0], lo[1]), (ro[0], ro[1]), pf.value


class System:
    """Integrals + determinant list + the reference-layout Hamiltonian."""

    def __init__(self, handle, n_up, n_dn):
        if not handle:
            raise RuntimeError("oracle: " + _load().orc_last_error().decode())
        self._h = handle
        self.n_up, self.n_dn = n_up, n_dn
        lib = _load()
        n = C.c_uint()
        core = C.c_double()
        up = C.c_uint64()
        dn = C.c_uint64()
        sz = C.c_size_t()
        lib.orc_system_info(self._h, C.byref(n), C.byref(core), C.byref(up), C.byref(dn), C.byref(sz))
        self.n_orbs, self.core, self.hf_up, self.hf_dn, self.eri_size = n.value, core.value, up.value, dn.value, sz.value
        self.n_dets = 0

    @classmethod
    def from_arrays(cls, n, n_up, n_dn, h1, eri, core=0.0):
        h1 = np.ascontiguousarray(h1, dtype=np.float64)
        eri = np.ascontiguousarray(eri, dtype=np.float64)
        h = _load().orc_system_new(n, n_up, n_dn, _dptr(h1), _dptr(eri), core)
        s = cls(h, n_up, n_dn)
        s.hf_up, s.hf_dn = (1 << n_up) - 1, (1 << n_dn) - 1
        return s

    @classmethod
    def from_fcidump(cls, path, n_up, n_dn, occs_up=(), occs_dn=(), reorder=-1):
        ou = (C.c_uint * max(1, len(occs_up)))(*occs_up)
        od = (C.c_uint * max(1, len(occs_dn)))(*occs_dn)
        h = _load().orc_system_from_fcidump(str(path).encode(), n_up, n_dn, ou, len(occs_up), od, len(occs_dn),
                                            reorder)
        return cls(h, n_up, n_dn)

    def __del__(self):
        if getattr(self, "_h", None):
            _load().orc_system_free(self._h)
            self._h = None

    def integrals(self):
        n = self.n_orbs
        h1 = np.zeros((n, n))
        eri = np.zeros(self.eri_size)
        order = (C.c_uint * n)()
        _load().orc_system_integrals(self._h, _dptr(h1), _dptr(eri), order)
        return h1, eri, np.array(order[:], dtype=np.int64)

    def set_integrals(self, h1, eri, core=0.0):
        h1 = np.ascontiguousarray(h1, dtype=np.float64)
        eri = np.ascontiguousarray(eri, dtype=np.float64)
        _load().orc_system_set_integrals(self._h, _dptr(h1), _dptr(eri), core)

    def elem(self, up_i, dn_i, up_j, dn_j, n_excite=-1):
        out = C.c_double()
        _check(_load().orc_elem(self._h, C.c_uint64(up_i), C.c_uint64(dn_i), C.c_uint64(up_j),
                                C.c_uint64(dn_j), n_excite, C.byref(out)))
        return out.value

    def update(self, up, dn):
        """``Hamiltonian::update`` on an append-only det list."""
        up = np.ascontiguousarray(up, dtype=np.uint64)
        dn = np.ascontiguousarray(dn, dtype=np.uint64)
        _check(_load().orc_ham_update(self._h, up.ctypes.data_as(_u64p), dn.ctypes.data_as(_u64p), len(up)))
        self.n_dets = len(up)

    def clear(self):
        _load().orc_ham_clear(self._h)
        self.n_dets = 0

    def update_existing_elems(self):
        _check(_load().orc_ham_update_existing(self._h))

    def stats(self):
        st = (C.c_size_t * 8)()
        tm = (C.c_double * 4)()
        _load().orc_ham_stats(self._h, st, tm)
        return dict(n_stored=st[0], nnz=st[1], singles=st[2], n_unique_alpha=st[3], n_unique_beta=st[4],
                    n_dets=st[5], t_abdet=tm[0], t_abm1=tm[1], t_absingles=tm[2], t_matrix=tm[3])

    def export_upper(self):
        """Stored upper-triangle rows: (row_ptr, cols, vals) in the reference's append order."""
        st = self.stats()
        n, m = st["n_dets"], st["n_stored"]
        rp = np.zeros(n + 1, dtype=np.int64)
        cols = np.zeros(m, dtype=np.int64)
        vals = np.zeros(m)
        _load().orc_ham_export(self._h, rp.ctypes.data_as(_i64p), cols.ctypes.data_as(_i64p), _dptr(vals))
        return rp, cols, vals

    def diag(self):
        d = np.zeros(self.n_dets)
        _load().orc_ham_diag(self._h, _dptr(d))
        return d

    def mul(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        _load().orc_ham_mul(self._h, _dptr(x), _dptr(y))
        return y

    def time_mul(self, x, reps=1):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return _load().orc_ham_time_mul(self._h, _dptr(x), reps)

    def davidson(self, init, tol, n_states=None):
        init = np.atleast_2d(np.ascontiguousarray(init, dtype=np.float64))
        k = init.shape[0] if n_states is None else n_states
        n = init.shape[1]
        ev = np.zeros(k)
        evec = np.zeros((k, n))
        conv = C.c_int()
        traj = np.zeros(64)
        nt = C.c_int()
        ns = C.c_int()
        _check(_load().orc_davidson(self._h, k, _dptr(init), tol, _dptr(ev), _dptr(evec), C.byref(conv),
                                    _dptr(traj), 64, C.byref(nt), C.byref(ns)))
        return ev, evec, bool(conv.value), traj[: nt.value].copy(), ns.value

    def rdm(self, coefs, want_2rdm=True):
        coefs = np.atleast_2d(np.ascontiguousarray(coefs, dtype=np.float64))
        n = self.n_orbs
        r1 = np.zeros((n, n))
        r2 = np.zeros(n * n * (n * n + 1) // 2) if want_2rdm else None
        _check(_load().orc_rdm(self._h, coefs.shape[0], _dptr(coefs), _dptr(r1),
                               _dptr(r2) if want_2rdm else C.cast(None, _dp)))
        return r1, r2

    def elem_time_sym(self, up_i, dn_i, up_j, dn_j):
        """H_ts between time-reversal-symmetrised states (base_system.h:86-101)."""
        out = C.c_double()
        _check(_load().orc_elem_time_sym(self._h, up_i, dn_i, up_j, dn_j, C.byref(out)))
        return out.value

    def select(self, up, dn, coefs, eps_var, eps_tried_prev, time_sym=False):
        """One selection sweep over the dets (up, dn): the new dets, (up, dn)
        ascending; ``eps_tried_prev`` (float64, one per det) updated in place."""
        up = np.ascontiguousarray(up, dtype=np.uint64)
        dn = np.ascontiguousarray(dn, dtype=np.uint64)
        coefs = np.atleast_2d(np.ascontiguousarray(coefs, dtype=np.float64))
        assert eps_tried_prev.dtype == np.float64 and eps_tried_prev.flags.c_contiguous
        m = C.c_size_t()
        _check(_load().orc_select(self._h, up.ctypes.data_as(_u64p), dn.ctypes.data_as(_u64p), len(up),
                                  _dptr(coefs), coefs.shape[0], eps_var, _dptr(eps_tried_prev), C.byref(m),
                                  int(time_sym)))
        nu = np.zeros(m.value, dtype=np.uint64)
        nd = np.zeros(m.value, dtype=np.uint64)
        _check(_load().orc_select_result(self._h, nu.ctypes.data_as(_u64p), nd.ctypes.data_as(_u64p)))
        return nu, nd


    def pt_dtm(self, up, dn, coefs, eps_pt_dtm, energy_var):
        """Deterministic PT2 correction (one batch) of the wavefunction
        (up, dn, coefs of one state): (E_pt2, number of PT dets)."""
        up = np.ascontiguousarray(up, dtype=np.uint64)
        dn = np.ascontiguousarray(dn, dtype=np.uint64)
        c = np.ascontiguousarray(coefs, dtype=np.float64).ravel()
        e = C.c_double()
        m = C.c_size_t()
        _check(_load().orc_pt_dtm(self._h, up.ctypes.data_as(_u64p), dn.ctypes.data_as(_u64p), len(up), _dptr(c),
                                  eps_pt_dtm, energy_var, C.byref(e), C.byref(m)))
        return e.value, m.value


def sparse_davidson(dim, entries, init, tol):
    """Davidson on a directly filled upper-triangle SparseMatrix (DavidsonTest)."""
    lib = _load()
    m = lib.orc_sm_new(dim)
    try:
        for i, j, v in entries:
            lib.orc_sm_append(m, i, j, v)
        init = np.atleast_2d(np.ascontiguousarray(init, dtype=np.float64))
        k, n = init.shape
        ev = np.zeros(k)
        evec = np.zeros((k, n))
        conv = C.c_int()
        traj = np.zeros(64)
        nt = C.c_int()
        _check(lib.orc_sm_davidson(m, k, _dptr(init), tol, _dptr(ev), _dptr(evec), C.byref(conv), _dptr(traj), 64,
                                   C.byref(nt)))
        return ev, evec, bool(conv.value), traj[: nt.value].copy()
    finally:
        lib.orc_sm_free(m)


def upper_digest(rp, cols, vals):
    """Order-free digest of the (row, column, FP64 value bits) triples of an
    upper-triangle export (row_ptr, cols, vals): equal digests and equal
    row pointers stand for the same pattern, signs and bit-identical values
    in any within-row order; a wrong column, a dropped entry or a value one
    ulp off changes it.  For exports too large to sort."""
    rows = np.repeat(np.arange(len(rp) - 1, dtype=np.uint64), np.diff(rp).astype(np.int64))
    with np.errstate(over="ignore"):
        h = rows * np.uint64(0x9E3779B97F4A7C15)
        h ^= np.asarray(cols).astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
        h ^= np.ascontiguousarray(vals, dtype=np.float64).view(np.uint64) * np.uint64(0x165667B19E3779F9)
        h ^= h >> np.uint64(31)
        h *= np.uint64(0xD6E8FEB86659FD93)
        h ^= h >> np.uint64(32)
        return int(np.bitwise_xor.reduce(h)), int(h.sum(dtype=np.uint64))


def unpack_2rdm(packed, n):
    """Full ``D[p, q, r, s]`` from the reference's packed layout (rdm.cc:233-241, 683-685)."""
    a = (np.arange(n)[:, None] * n + np.arange(n)[None, :])  # a[p, s] = p*n + s
    A = a[:, None, None, :]          # index p*n + s
    B = a[None, :, :, None]          # index q*n + r  (b[q, r])
    hi = np.maximum(A, B)
    lo = np.minimum(A, B)
    return packed[hi * (hi + 1) // 2 + lo]
