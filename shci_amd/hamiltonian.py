"""Python mirror of the reference's variational call surface, backed by the
gfx950 C-ABI (include/shci_hip.h):

* :class:`Hamiltonian`  -- ``Hamiltonian<S>`` (src/solver/hamiltonian.h:14-79):
  ``matrix``, ``update``, ``update_existing_elems``, ``clear``.
* :class:`SparseMatrix` -- the consumer view of ``SparseMatrix``
  (src/solver/sparse_matrix.h:14-50): ``get_diag``, ``count_n_elems``,
  ``count_n_rows``, ``mul`` (real and complex), ``get_row``,
  ``get_connections``, ``clear``.
* :class:`Davidson`     -- ``Davidson`` (src/solver/davidson.h:6-29):
  ``diagonalize``, ``get_lowest_eigenvalues``, ``get_lowest_eigenvectors``,
  ``converged``.

Errors from the library raise ``RuntimeError`` (the reference throws
``std::runtime_error``); non-convergence is reported through
``Davidson.converged`` as in the reference.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._native import hip_lib

__all__ = ["Context", "Hamiltonian", "SparseMatrix", "Davidson", "ShciError"]


class ShciError(RuntimeError):
    pass


_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


def _vp(a):
    return C.c_void_p(a.ctypes.data)


class Context:
    """One device context (stream + device buffers) per process and device."""

    _by_device: dict = {}

    def __init__(self, device: int = 0):
        self.lib = hip_lib()
        h = C.c_void_p()
        rc = self.lib.shci_ctx_create(C.byref(h), device)
        self._h = h
        if rc != 0:
            raise ShciError(f"shci_ctx_create failed ({rc}): {self.lib.shci_last_error(h).decode()}")
        self.device = device

    @classmethod
    def get(cls, device: int = 0) -> "Context":
        if device not in cls._by_device:
            cls._by_device[device] = cls(device)
        return cls._by_device[device]

    @property
    def handle(self):
        return self._h

    def check(self, rc: int, what: str = ""):
        if rc != 0:
            raise ShciError(f"{what} failed ({rc}): {self.lib.shci_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self.lib.shci_ctx_destroy(self._h)
            self._h = None

    def memory(self) -> int:
        b = C.c_uint64()
        self.check(self.lib.shci_memory(self._h, C.byref(b)), "shci_memory")
        return b.value


class SparseMatrix:
    """Device-resident Hamiltonian (full-symmetric CSR + FP64 diagonal)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self._built = False

    # -- container queries (sparse_matrix.h:14-22) --
    def counts(self):
        n = C.c_uint64()
        up = C.c_uint64()
        el = C.c_uint64()
        self.ctx.check(self.ctx.lib.shci_get_counts(self.ctx.handle, C.byref(n), C.byref(up), C.byref(el)),
                       "shci_get_counts")
        return n.value, up.value, el.value

    def count_n_rows(self) -> int:
        if not self._built:
            return 0
        n = C.c_uint64()
        self.ctx.check(self.ctx.lib.shci_get_counts(self.ctx.handle, C.byref(n), None, None), "shci_get_counts")
        return n.value

    def count_n_elems(self) -> int:
        """``2 * stored_upper - n`` (sparse_matrix.cc:13-26)."""
        return self.counts()[2]

    def get_diag(self, i=None):
        n = self.count_n_rows()
        d = np.zeros(n)
        self.ctx.check(self.ctx.lib.shci_get_diag(self.ctx.handle, d.ctypes.data_as(_dp)), "shci_get_diag")
        return d if i is None else float(d[i])

    # -- products (sparse_matrix.cc:28-85) --
    def mul(self, vec, fp32: bool = False):
        vec = np.asarray(vec)
        if np.iscomplexobj(vec):
            return self.mul(vec.real, fp32) + 1j * self.mul(vec.imag, fp32)
        x = np.ascontiguousarray(vec, dtype=np.float64)
        y = np.zeros_like(x)
        flags = _abi.SHCI_FP32 if fp32 else 0
        self.ctx.check(self.ctx.lib.shci_spmv(self.ctx.handle, _vp(x), _vp(y), flags), "shci_spmv")
        return y

    def mul_block(self, X, fp32: bool = True):
        X = np.ascontiguousarray(np.atleast_2d(X), dtype=np.float64)
        Y = np.zeros_like(X)
        flags = _abi.SHCI_FP32 if fp32 else 0
        self.ctx.check(self.ctx.lib.shci_spmm(self.ctx.handle, _vp(X), _vp(Y), X.shape[0], flags), "shci_spmm")
        return Y

    def mul_device(self, x_ptr: int, y_ptr: int, fp32: bool = False):
        """y = H x for float64 device pointers (e.g. torch tensors' data_ptr())."""
        flags = _abi.SHCI_DEVICE_PTRS | (_abi.SHCI_FP32 if fp32 else 0)
        self.ctx.check(self.ctx.lib.shci_spmv(self.ctx.handle, C.c_void_p(x_ptr), C.c_void_p(y_ptr), flags),
                       "shci_spmv")

    # -- structure export (sparse_matrix.h:44-48) --
    def export_upper(self):
        """Stored upper triangle in caller indices: (row_ptr, cols, vals); each
        row holds its diagonal first, then columns j > i."""
        n, up, _ = self.counts()
        rp = np.zeros(n + 1, dtype=np.int64)
        cols = np.zeros(up, dtype=np.int64)
        vals = np.zeros(up)
        self.ctx.check(self.ctx.lib.shci_export_upper(self.ctx.handle, rp.ctypes.data_as(_i64p),
                                                      cols.ctypes.data_as(_i64p), vals.ctypes.data_as(_dp)),
                       "shci_export_upper")
        return rp, cols, vals

    def get_row(self, i: int):
        rp, cols, vals = self.export_upper()
        return cols[rp[i]:rp[i + 1]].copy(), vals[rp[i]:rp[i + 1]].copy()

    def get_connections(self):
        rp, cols, _ = self.export_upper()
        return [cols[rp[i]:rp[i + 1]].tolist() for i in range(len(rp) - 1)]

    def clear(self):
        self._built = False

    # -- timing helpers (bench / profiling) --
    def time_spmv(self, fp32: bool = False, reps: int = 10, f64_vectors: bool = False):
        """Average time (ms) and algorithmic bytes of a device-resident SpMV.
        ``fp32`` selects the FP32 store; ``f64_vectors`` pairs it with FP64
        vectors (the mixed Davidson's shift and energy re-evaluation)."""
        ms = C.c_float()
        by = C.c_double()
        flags = (_abi.SHCI_FP32 if fp32 else 0) | (_abi.SHCI_F64_VECTORS if f64_vectors else 0)
        self.ctx.check(self.ctx.lib.shci_time_spmv(self.ctx.handle, flags, reps, C.byref(ms), C.byref(by)),
                       "shci_time_spmv")
        return ms.value, by.value

    def time_spmm(self, k: int = 4, reps: int = 10):
        """Average time (ms) and algorithmic bytes of an FP32 block product with k vectors."""
        ms = C.c_float()
        by = C.c_double()
        self.ctx.check(self.ctx.lib.shci_time_spmm(self.ctx.handle, k, reps, C.byref(ms), C.byref(by)),
                       "shci_time_spmm")
        return ms.value, by.value

    def capture_ritz(self, max_vectors: int):
        """Keep the ground-state Ritz vector behind each trajectory entry of
        the following Davidson calls (0 stops and frees)."""
        self.ctx.check(self.ctx.lib.shci_capture_ritz(self.ctx.handle, int(max_vectors)), "shci_capture_ritz")

    def ritz_energies(self, cap: int = 4096):
        """FP64 energies <w|H|w>/<w|w> of the captured Ritz vectors of the
        latest Davidson call; needs the FP64 store over the same dets."""
        e = np.zeros(cap)
        m = C.c_uint32()
        self.ctx.check(self.ctx.lib.shci_ritz_energies(self.ctx.handle, e.ctypes.data_as(_dp), cap, C.byref(m)),
                       "shci_ritz_energies")
        return e[: min(cap, m.value)].copy()

    def layout_stats(self):
        st = (C.c_uint64 * 14)()
        self.ctx.check(self.ctx.lib.shci_layout_stats(self.ctx.handle, st), "shci_layout_stats")
        names = ["n_dets", "nnz", "nnz_alloc", "segments_layout", "segments", "memberships", "large_segments",
                 "value_slots", "mixed_runs", "alpha_groups", "beta_groups", "lanes", "tiles", "tile_dets"]
        return dict(zip(names, list(st)))

    def segment_ranges(self):
        """Per (size class, type) range of the segment layout: a list of 27
        dicts (class, type, segments, memberships, slots); class 0..8 = m > 1024,
        > 256, > 64, 33-64, 17-32, 9-16, 5-8, 3-4, 2; type C, A, B."""
        out = (C.c_uint64 * 81)()
        self.ctx.check(self.ctx.lib.shci_segment_ranges(self.ctx.handle, out), "shci_segment_ranges")
        kept = (C.c_uint64 * 27)()
        self.ctx.check(self.ctx.lib.shci_segment_kept(self.ctx.handle, kept), "shci_segment_kept")
        return [{"class": r // 3, "type": "CAB"[r % 3], "segments": out[3 * r], "memberships": out[3 * r + 1],
                 "slots": out[3 * r + 2], "kept_pairs": kept[r]} for r in range(27)]

    def build_timings(self):
        t = (C.c_float * 8)()
        self.ctx.check(self.ctx.lib.shci_get_build_timings(self.ctx.handle, t), "shci_get_build_timings")
        names = ["group", "singles", "diag", "count", "scan_alloc", "fill", "compact", "total"]
        return dict(zip(names, list(t)))


class Hamiltonian:
    """``Hamiltonian<ChemSystem>`` over the device.  ``store`` selects the value
    precision(s) kept: ``"fp64"``, ``"fp32"`` or ``"both"``.  ``layout``:
    ``"csr"`` (full-symmetric CSR; supports row export) or ``"segments"``
    (dets grouped by shared cores, each connected pair stored once as a value
    slot without column index; faster SpMV, less memory).  ``time_sym``: the
    dets (up <= dn) stand for time-reversal-symmetrised states and every
    vector is in that basis (ChemSystem time_sym, base_system.h:86-122)."""

    def __init__(self, n_orb, h1, eri, core=0.0, device=0, store="fp64", keep_zeros=False, layout="csr",
                 time_sym=False):
        self.ctx = Context.get(device)
        self.matrix = SparseMatrix(self.ctx)
        self.n_dets = 0
        self.n_dets_prev = 0
        self.store = store
        self.keep_zeros = keep_zeros
        if layout not in ("csr", "segments"):
            raise ShciError(f"unknown layout {layout!r}")
        self.layout = layout
        self.time_sym = bool(time_sym)
        self.set_integrals(n_orb, h1, eri, core)

    def set_integrals(self, n_orb, h1, eri, core=0.0):
        self.n_orb = int(n_orb)
        self._h1 = np.ascontiguousarray(h1, dtype=np.float64)
        self._eri = np.ascontiguousarray(eri, dtype=np.float64)
        npair = self.n_orb * (self.n_orb + 1) // 2
        if self._h1.shape != (self.n_orb, self.n_orb) or self._eri.size != npair * (npair + 1) // 2:
            raise ShciError("integral shapes do not match n_orb")
        self.core = float(core)
        self.ctx.check(self.ctx.lib.shci_set_integrals(self.ctx.handle, self.n_orb, self._h1.ctypes.data_as(_dp),
                                                       self._eri.ctypes.data_as(_dp), self.core),
                       "shci_set_integrals")
        self.ctx.integrals_owner = self

    def ensure_integrals(self):
        """Makes these integrals the context's current ones (contexts are
        shared per device; the selection reads the context's integrals)."""
        if getattr(self.ctx, "integrals_owner", None) is not self:
            self.set_integrals(self.n_orb, self._h1, self._eri, self.core)

    def _flags(self):
        f = {"fp64": _abi.SHCI_STORE_FP64, "fp32": _abi.SHCI_STORE_FP32,
             "both": _abi.SHCI_STORE_FP64 | _abi.SHCI_STORE_FP32}[self.store]
        if self.keep_zeros:
            f |= _abi.SHCI_KEEP_ZEROS
        if self.layout == "segments":
            f |= _abi.SHCI_LAYOUT_SEGMENTS
        if self.time_sym:
            f |= _abi.SHCI_TIME_SYM
        return f

    def update(self, up, dn):
        """hamiltonian.h:87-112: ``up``/``dn`` is the whole (append-only) det list."""
        up = np.ascontiguousarray(up, dtype=np.uint64)
        dn = np.ascontiguousarray(dn, dtype=np.uint64)
        self.n_dets_prev = self.n_dets
        self.n_dets = len(up)
        if self.n_dets == self.n_dets_prev and self.matrix._built:
            return
        self.ctx.check(self.ctx.lib.shci_build_hamiltonian(self.ctx.handle, _vp(up), _vp(dn), self.n_dets,
                                                           self.n_dets_prev, self._flags()),
                       "shci_build_hamiltonian")
        self.matrix._built = True

    def rebuild(self):
        """Rebuild from the device-resident det list (timed builds)."""
        self.ctx.check(self.ctx.lib.shci_rebuild_hamiltonian(self.ctx.handle, self._flags()),
                       "shci_rebuild_hamiltonian")

    def update_existing_elems(self):
        """hamiltonian.h:114-126 with the integrals last passed to set_integrals."""
        self.ctx.check(self.ctx.lib.shci_update_existing_elems(self.ctx.handle), "shci_update_existing_elems")

    def clear(self):
        """hamiltonian.h:128-155: drops the dets and the matrix (device memory released)."""
        self.n_dets = self.n_dets_prev = 0
        self.matrix.clear()
        self.ctx.check(self.ctx.lib.shci_clear(self.ctx.handle), "shci_clear")

    def rdm(self, coefs):
        """(1-RDM n x n, packed 2-RDM) of ``coefs`` (n_states x n_dets)."""
        coefs = np.ascontiguousarray(np.atleast_2d(coefs), dtype=np.float64)
        n = self.n_orb
        r1 = np.zeros((n, n))
        r2 = np.zeros(n * n * (n * n + 1) // 2)
        self.ctx.check(self.ctx.lib.shci_rdm(self.ctx.handle, coefs.shape[0], coefs.ctypes.data_as(_dp),
                                             r1.ctypes.data_as(_dp), r2.ctypes.data_as(_dp)), "shci_rdm")
        return r1, r2


class Davidson:
    """davidson.h:6-29.  ``mode``: ``"fp64"`` (reference algorithm) or ``"mixed"``."""

    def __init__(self, n_states: int = 1):
        self.n_states = n_states
        self.lowest_eigenvalues = np.zeros(n_states)
        self.lowest_eigenvectors = None
        self.converged = False
        self.trajectory = np.zeros(0)
        self.trajectory_t = np.zeros(0)

    def diagonalize(self, matrix: SparseMatrix, initial_vectors, target_error, verbose=False, mode="fp64",
                    max_calls=1):
        init = np.ascontiguousarray(np.atleast_2d(initial_vectors), dtype=np.float64)
        k, n = init.shape
        ns = min(k, self.n_states) if k >= 1 else self.n_states
        evals = np.zeros(ns)
        evecs = np.zeros((ns, n))
        conv = C.c_int()
        cap = 64 * max(1, max_calls) + 8
        te = np.zeros(cap)
        tt = np.zeros(cap)
        nt = C.c_int()
        m = _abi.SHCI_DAVIDSON_MIXED if mode == "mixed" else _abi.SHCI_DAVIDSON_FP64
        ctx = matrix.ctx
        ctx.check(ctx.lib.shci_davidson(ctx.handle, ns, init.ctypes.data_as(_dp), float(target_error), m,
                                        int(max_calls), evals.ctypes.data_as(_dp), evecs.ctypes.data_as(_dp),
                                        C.byref(conv), te.ctypes.data_as(_dp), tt.ctypes.data_as(_dp), cap,
                                        C.byref(nt)), "shci_davidson")
        self.lowest_eigenvalues = evals
        self.lowest_eigenvectors = evecs
        self.converged = bool(conv.value)
        self.trajectory = te[: nt.value].copy()
        self.trajectory_t = tt[: nt.value].copy()
        if verbose:
            for i, e in enumerate(self.trajectory):
                print(f"Davidson #{i}: {e:.10f}")
        return evals

    def get_lowest_eigenvalues(self):
        return self.lowest_eigenvalues

    def get_lowest_eigenvectors(self):
        return self.lowest_eigenvectors
