/* C-ABI of libshci_hip.so: the MI355X (gfx950) variational hot path of SHCI.
 *
 * The reference ("Arrow", QMC-Cornell/shci) computes this path on the CPU in
 *   Hamiltonian<S>::update          src/solver/hamiltonian.h:87-598
 *   SparseMatrix::mul               src/solver/sparse_matrix.cc:28-56
 *   Davidson::diagonalize           src/solver/davidson.cc:8-196
 *   RDM::get_1rdm / get_2rdm        src/chem/rdm.cc:12-131, 374-395
 * the step before it, the heat-bath selection sweep
 *   Solver<S>::run_variation loop   src/solver/solver.h:435-495
 *   ChemSystem::find_connected_dets src/chem/chem_system.cc:289-402
 * and the step after it, the deterministic PT2 correction
 *   Solver<S>::get_energy_pt_dtm    src/solver/solver.h:637-764
 * Each entry point below names the reference interface it replaces.  All
 * arguments are plain pointers and sizes; host pointers unless a flag says
 * otherwise.  Every call returns SHCI_OK (0) or a negative status; the message
 * is available from shci_last_error().  One context per process and device,
 * calls serialised by the caller (like the reference's master thread).
 *
 * Determinants: one uint64 per spin (bit k = orbital k), n_orb <= 64.
 * Integrals:    h1 dense n x n (symmetric, the value Integrals::get_1b returns);
 *               eri packed by the reference key combine4(p,q,r,s) =
 *               combine2(combine2(p,q), combine2(r,s)) (src/chem/integrals.cc:410-422),
 *               one chemists' (pq|rs) per 8-fold class, length P(P+1)/2, P = n(n+1)/2.
 */
#ifndef SHCI_HIP_H_
#define SHCI_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct shci_ctx shci_ctx;

enum {
  SHCI_OK = 0,
  SHCI_ERR_HIP = -1,         /* a HIP runtime call failed */
  SHCI_ERR_OOM = -2,         /* device allocation failed */
  SHCI_ERR_UNSUPPORTED = -3, /* n_orb > 64, time_sym, ... */
  SHCI_ERR_ARG = -4,         /* bad argument */
  SHCI_ERR_STATE = -5        /* call order (e.g. SpMV before a build) */
};

/* Build / SpMV flags. */
enum {
  SHCI_STORE_FP64 = 1u << 0,   /* keep FP64 off-diagonal values (FP64 path) */
  SHCI_STORE_FP32 = 1u << 1,   /* keep FP32 off-diagonal values (mixed path) */
  SHCI_KEEP_ZEROS = 1u << 2,   /* keep |H_ij| < 1e-12 connections (exact RDM adjoint) */
  SHCI_DEVICE_PTRS = 1u << 3,  /* vector/det arguments are device pointers */
  SHCI_FP32 = 1u << 4,         /* SpMV in FP32 storage / FP32 row accumulation */
  SHCI_LAYOUT_SEGMENTS = 1u << 5, /* build: segment layout instead of the CSR: dets
                                    grouped by shared (alpha-1, beta-1), (alpha-2) and
                                    (beta-2) cores, each connected pair stored once as a
                                    value slot of its owning segment, no column indices;
                                    faster SpMV, less memory; no row export */
  SHCI_F64_VECTORS = 1u << 6,  /* shci_time_spmv: FP64 vectors with the store SHCI_FP32
                                  selects (FP32 values and FP64 accumulation: the
                                  mixed Davidson's shift and energy re-evaluation) */
  SHCI_TIME_SYM = 1u << 7      /* build: time-reversal symmetry (ChemSystem time_sym,
                                  base_system.h:86-122).  Each det (host pointers, up <= dn,
                                  n_up == n_dn, the reference's canonical form, solver.h:452)
                                  stands for the state (|up,dn> + |dn,up>)/sqrt2, or |up,up>.
                                  Every later call works in that basis: vectors and
                                  coefficients of n_dets entries, H_ts elements of
                                  base_system.h:86-101 (products, diagonal, Davidson);
                                  shci_rdm unpacks like unpack_time_sym.  The matrix is built
                                  over the dets and their reverses; shci_get_counts answers
                                  n_dets only (SHCI_ERR_UNSUPPORTED when an element count
                                  is asked for); shci_export_upper is unsupported. */
};

/* Davidson modes. */
enum {
  SHCI_DAVIDSON_FP64 = 0,  /* reference algorithm, FP64 everywhere */
  SHCI_DAVIDSON_MIXED = 1  /* FP32 CSR + vectors; FP64 diag, Gram-Schmidt, Rayleigh-Ritz, refinement */
};

/* Context lifetime.  No counterpart in the reference (it owns no device). */
int shci_ctx_create(shci_ctx** ctx, int device);
int shci_ctx_destroy(shci_ctx* ctx);
const char* shci_last_error(const shci_ctx* ctx);

/* Integrals as read by ChemSystem::get_hamiltonian_elem through
 * Integrals::get_1b / get_2b (src/chem/integrals.cc:390-408). */
int shci_set_integrals(shci_ctx* ctx, uint32_t n_orb, const double* h1, const double* eri_packed,
                       double energy_core);

/* Hamiltonian<S>::update (src/solver/hamiltonian.h:87-112).  dets up[i], dn[i]
 * in the caller's order; the matrix is the full variational H over them
 * (diagonal + every pair with excitation <= 2 and |H_ij| >= 1e-12,
 * hamiltonian.h:440,466,506,527).  n_dets_prev follows the reference's
 * append-only contract: n_dets == n_dets_prev returns at once (hamiltonian.h:91).
 * flags: SHCI_STORE_FP64 and/or SHCI_STORE_FP32, SHCI_KEEP_ZEROS, SHCI_DEVICE_PTRS. */
int shci_build_hamiltonian(shci_ctx* ctx, const uint64_t* up, const uint64_t* dn, uint64_t n_dets,
                           uint64_t n_dets_prev, uint32_t flags);

/* Rebuild from the determinants already resident on the device (same list as
 * the last shci_build_hamiltonian); used to time device-resident builds. */
int shci_rebuild_hamiltonian(shci_ctx* ctx, uint32_t flags);

/* Hamiltonian<S>::update_existing_elems (hamiltonian.h:114-126,
 * sparse_matrix.cc:131-141): recompute every stored value with the current
 * integrals on the fixed connection pattern. */
int shci_update_existing_elems(shci_ctx* ctx);

/* Hamiltonian<S>::clear / SparseMatrix::clear (hamiltonian.h:128-155,
 * sparse_matrix.cc:93-100): release the determinants, the matrix and all
 * work buffers (the integrals stay); their device memory, cached blocks and
 * pooled chunks included, goes back to the driver. */
int shci_clear(shci_ctx* ctx);

/* SparseMatrix::count_n_elems (sparse_matrix.cc:13-26) and row count:
 * n_stored_upper counts the diagonal + pairs i<j (the reference's storage);
 * n_elems = 2 * n_stored_upper - n_dets (its printed "nonzero elems").  Any
 * pointer may be NULL.  Under SHCI_TIME_SYM only n_dets is answered. */
int shci_get_counts(shci_ctx* ctx, uint64_t* n_dets, uint64_t* n_stored_upper, uint64_t* n_elems);

/* SparseMatrix::get_diag() (sparse_matrix.h:16): FP64 diagonal, caller order. */
int shci_get_diag(shci_ctx* ctx, double* diag);

/* SparseMatrix::mul (sparse_matrix.cc:28-56): y = H x, caller order.
 * FP64 by default; with SHCI_FP32 the FP32 store is used (values FP32, row
 * accumulation FP32, diagonal FP64).  SHCI_DEVICE_PTRS: x, y on the device. */
int shci_spmv(shci_ctx* ctx, const double* x, double* y, uint32_t flags);

/* y = H x like shci_spmv with SHCI_DEVICE_PTRS, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  x and y are device pointers (caller order, FP64); flags:
 * SHCI_FP32 selects the FP32 store as in shci_spmv.  The working vectors (x
 * and y in internal order and the beta-order copies of the A segments) live
 * in a fixed device array of 256 MiB per GPU, held until the process ends, so
 * n_dets is at most 8,388,608 (SHCI_ERR_UNSUPPORTED beyond; also under
 * SHCI_TIME_SYM).  Every call on one GPU uses that GPU's one array: calls that
 * overlap on a GPU (two contexts, two threads) race.  The matrix must be
 * built; the call returns when y is written. */
int shci_spmv_noalloc(shci_ctx* ctx, const double* x, double* y, uint32_t flags);

/* Block product Y = H X for k vectors (k <= 8), X, Y row-major k x n, caller
 * order.  The reference declares a multi-vector mul (sparse_matrix.h:28-32)
 * but never defines it; this is the block form used by the multi-root path. */
int shci_spmm(shci_ctx* ctx, const double* X, double* Y, uint32_t k, uint32_t flags);

/* Davidson::diagonalize (src/solver/davidson.cc:8-196) on the built H.
 * init: n_states x n_dets (caller order); evals[n_states], evecs n_states x n.
 * tol: eigenvalue-change tolerance (davidson.cc:174-185).  mode: SHCI_DAVIDSON_*.
 * max_calls: the reference performs at most 2*5*n_states - n_states
 * iterations per call and its caller re-invokes with the Ritz vectors
 * (solver.h:430-530); max_calls > 1 repeats that warm restart here until
 * converged.  traj_e/traj_t (optional, capacity traj_cap) receive the
 * ground-state estimate and elapsed seconds after every Rayleigh-Ritz step;
 * in MIXED mode the last entry is the FP64 re-evaluated energy.  MIXED mode
 * also accepts a call's convergence only if every Ritz vector's residual
 * |Hc - Ec| (FP64 vectors) is at most max(10 sqrt(tol), 1e-4); otherwise
 * the call counts as unconverged and is warm-restarted while calls remain
 * (*converged = 0 if none remain). */
int shci_davidson(shci_ctx* ctx, uint32_t n_states, const double* init, double tol, uint32_t mode,
                  uint32_t max_calls, double* evals, double* evecs, int* converged, double* traj_e,
                  double* traj_t, int traj_cap, int* n_traj);

/* RDM::get_1rdm (rdm.cc:12-131) over every single excitation of the det
 * list, whether or not the matrix stores the pair (|H| < 1e-12 is cut
 * there), and RDM::get_2rdm(const SparseMatrix&) (rdm.cc:374-395) over the
 * built connections; FP64 accumulation, states summed (rdm.cc:227,652).
 * coefs: n_states x n_dets (caller order).
 * rdm1: n x n; rdm2: packed n^2(n^2+1)/2 in the reference layout
 * (index a = p*n+s, b = q*n+r, a >= b, rdm.cc:644-657).  Either may be NULL.
 * This is the adjoint of E(h, v): dE/dh = rdm1, dE/d(ps|qr) = D_pqrs / 2. */
int shci_rdm(shci_ctx* ctx, uint32_t n_states, const double* coefs, double* rdm1, double* rdm2);

/* SparseMatrix::get_row / get_connections (sparse_matrix.h:44-48): the stored
 * upper triangle (diagonal + j > i) in caller indices.  row_ptr[n+1];
 * cols/vals sized n_stored_upper (from shci_get_counts).  Any may be NULL
 * except row_ptr. */
int shci_export_upper(shci_ctx* ctx, int64_t* row_ptr, int64_t* cols, double* vals);

/* One heat-bath selection sweep of the variational loop (solver.h:435-495,
 * the body of the loop over dets with ChemSystem::find_connected_dets,
 * chem_system.cc:289-402; queues of chem_system.cc:115-287) on the device.
 * up/dn: the n_dets variational dets (host, caller order; orbitals < n_orb,
 * equal electron counts); coefs: n_states x n_dets (host); eps_tried_prev:
 * n_dets values, in/out (1e100 for dets never tried; the sweep sets
 * 1.000000001 * eps_min for every det it walks, solver.h:466).  For each det
 * with eps_min = eps_var / max_state |c| below its eps_tried_prev: singles
 * with S >= eps_min and |H_ij| >= eps_min, doubles with
 * eps_min <= |H| < eps_tried_prev, not in the variational space.  *n_new
 * receives the number of distinct new dets; shci_select_result copies them
 * out, (up, dn) ascending (the reference appends them in the order of its
 * hash set).  flags: SHCI_TIME_SYM applies the reference's time_sym rules
 * (solver.h:447-455): the dets are canonical (up <= dn), eps_min is multiplied
 * by sqrt2 for a det with up != dn, and candidates are tested against the
 * space and returned in their canonical form. */
int shci_select(shci_ctx* ctx, const uint64_t* up, const uint64_t* dn, uint64_t n_dets, const double* coefs,
                uint32_t n_states, double eps_var, double* eps_tried_prev, uint64_t* n_new, uint32_t flags);
/* The last sweep's new dets: n must equal its *n_new. */
int shci_select_result(shci_ctx* ctx, uint64_t* new_up, uint64_t* new_dn, uint64_t n);
/* Deterministic PT2 correction of one state (Solver::get_energy_pt_dtm,
 * solver.h:637-764) on the device.  up/dn, coefs: the n_dets variational dets
 * (host, caller order) and that state's coefficients.  For every det i, the
 * dets a that find_connected_dets yields with eps_min = eps_pt_dtm / |c_i|
 * and no upper cut (chem_system.cc:289-402), outside the variational space
 * and with |H_ai c_i| >= eps_pt_dtm, add H_ai c_i to hc_sums[a]; then
 * *energy_pt = sum_a hc_sums[a]^2 / (energy_var - H_aa) and *n_pt_dets the
 * number of such a.  Each hc_sums[a] adds its terms in increasing i.
 * n_batches: partitions of the a by hash (0 = as few as device memory
 * allows); the result does not depend on it beyond the summation order. */
int shci_pt_dtm(shci_ctx* ctx, const uint64_t* up, const uint64_t* dn, uint64_t n_dets, const double* coefs,
                double eps_pt_dtm, double energy_var, uint32_t n_batches, double* energy_pt, uint64_t* n_pt_dets);

/* Phases of the last sweep (no reference counterpart; diagnostics), HIP
 * events on the context's stream, ms[4]: upload, checks, queues and the
 * variational-space hash set; the walk (one append pass); any repeated or
 * chunked passes, sort and unique; unpack and download. */
int shci_select_timings(shci_ctx* ctx, float* ms);

/* Timing of the most recent build phases and a timed SpMV loop, measured with
 * HIP events on the context's stream (milliseconds).  phases[8]: sort/group,
 * singles, diagonal, count, scan+alloc, fill, compact, total. */
int shci_get_build_timings(shci_ctx* ctx, float* phases);
/* Average kernel time of `reps` SpMVs (device-resident vectors) and the
 * algorithmic bytes one SpMV moves.  flags: SHCI_FP32 selects the FP32 store,
 * SHCI_F64_VECTORS FP64 vectors with it. */
int shci_time_spmv(shci_ctx* ctx, uint32_t flags, int reps, float* ms_avg, double* bytes);

/* Average kernel time of `reps` FP32 block products with k = 4 device-resident
 * vectors (the multi-root product of shci_spmm) and the algorithmic bytes of
 * one product (values read once for the 4 vectors).  No reference
 * counterpart (the reference declares a multi-vector mul but never defines
 * it, sparse_matrix.h:28-32). */
int shci_time_spmm(shci_ctx* ctx, uint32_t k, int reps, float* ms_avg, double* bytes);

/* Ritz-vector capture (no reference counterpart; measurement).  With
 * max_vectors > 0, every later shci_davidson call keeps, in device memory, the
 * ground-state Ritz vector behind each of its trajectory entries (the start
 * vector's and every Rayleigh-Ritz step's; not the final re-evaluation of
 * MIXED mode), at most max_vectors of them, widened to FP64; each call starts
 * afresh.  0 stops the capture and frees the vectors.  The capture's widening
 * copy is queued after each entry's timestamp. */
int shci_capture_ritz(shci_ctx* ctx, uint32_t max_vectors);
/* FP64 energies <w|H|w> / <w|w> of the captured vectors with the FP64 store
 * (the matrix rebuilt with SHCI_STORE_FP64 over the same dets): e[k] for
 * k < min(cap, captured); *n_captured receives the number captured.  This is
 * how a MIXED run's time to a given FP64 accuracy is measured: the
 * trajectory's timestamps, the energies from here. */
int shci_ritz_energies(shci_ctx* ctx, double* e, uint32_t cap, uint32_t* n_captured);

/* Layout statistics (no reference counterpart; diagnostics): stats[14] =
 * n_dets, stored off-diagonal entries (both directions), allocated entries
 * (both directions), segment layout?, segments, segment memberships,
 * segments of more than 64 members, value slots, mixed-double runs of the CSR
 * search, alpha groups, beta groups, CSR SpMV lanes per row, tiles of the
 * segments with at most 64 members, distinct dets summed over the tiles. */
int shci_layout_stats(shci_ctx* ctx, uint64_t* stats);

/* Segment-layout ranges (no reference counterpart; diagnostics): for each of
 * the 27 (size class, type) ranges r = 3 * class + type (classes m > 1024,
 * > 256, > 64, 33-64, 17-32, 9-16, 5-8, 3-4, 2; types C, A, B),
 * out[3 * r] = segments, out[3 * r + 1] = memberships, out[3 * r + 2] =
 * value slots.  SHCI_ERR_STATE unless the segment layout is built. */
int shci_segment_ranges(shci_ctx* ctx, uint64_t* out);
/* Segment-layout kept pairs (no reference counterpart; diagnostics): out[r]
 * = pairs of range r (as above) that its segments own and keep (|H| >= 1e-12
 * or SHCI_KEEP_ZEROS), as counted by the last fill; the range's other value
 * slots hold zeros. */
int shci_segment_kept(shci_ctx* ctx, uint64_t* out);

/* Device memory held by the context (bytes): its buffers, its cached blocks
 * and its pooled physical chunks. */
int shci_memory(shci_ctx* ctx, uint64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* SHCI_HIP_H_ */
