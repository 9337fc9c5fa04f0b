/* C API of libshci_host.so: the CPU side of the variational stage that stays
 * on the host (integrals and heat-bath selection) plus the variational driver
 * that calls the GPU path through include/shci_hip.h.
 *
 *   integrals           src/chem/integrals.cc:37-422 (FCIDUMP read, HF det, reorder)
 *   heat-bath selection src/chem/chem_system.cc:74-402, src/solver/solver.h:435-495
 *   variational loop    src/solver/solver.h:338-539 (run_all_variations / run_variation)
 *   RDM files           src/chem/rdm.cc:202-331
 *   wavefunction files  src/solver/solver.h:1179-1224, src/base_system.h:125-133
 */
#ifndef SHCI_HOST_H_
#define SHCI_HOST_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* shci_host_last_error(void);

/* Integrals::load (integrals.cc:17-35) without the binary cache.
 * reorder: -1 reference default (!explicit occupations), 0 off, 1 on. */
void* shci_host_system_from_fcidump(const char* path, unsigned n_up, unsigned n_dn, const unsigned* occs_up,
                                    unsigned n_occs_up, const unsigned* occs_dn, unsigned n_occs_dn, int reorder);
void* shci_host_system_from_arrays(unsigned n_orbs, unsigned n_up, unsigned n_dn, const double* h1_dense,
                                   const double* eri_packed, double core, uint64_t hf_up, uint64_t hf_dn);
void shci_host_system_free(void* sys);
int shci_host_system_info(void* sys, unsigned* n_orbs, double* core, uint64_t* hf_up, uint64_t* hf_dn,
                          size_t* eri_size, double* e_hf);
int shci_host_system_integrals(void* sys, double* h1_dense, double* eri_packed, unsigned* orb_order);

/* Variational determinant list (append-only, starts at the HF det). */
int shci_host_reset_dets(void* sys);
size_t shci_host_n_dets(void* sys);
int shci_host_get_dets(void* sys, uint64_t* up, uint64_t* dn);
int shci_host_set_dets(void* sys, const uint64_t* up, const uint64_t* dn, size_t n);

/* One selection sweep of run_variation (solver.h:435-495); coefs n_states x n_dets.
 * Returns the number of new determinants, or -1 on error. */
long long shci_host_select(void* sys, const double* coefs, unsigned n_states, double eps_var);

/* Solver::run_all_variations (solver.h:338-397) with the Hamiltonian build and
 * Davidson on the GPU (include/shci_hip.h).  eps_vars / schedule as in the
 * reference config; mode bit 0: mixed Davidson (else FP64); bit 1: selection
 * on the device (shci_select, else the host queue walk); bit 2: wavefunction
 * files in the working directory, as the reference keeps them: each eps_var
 * first tries to load wf_eps1_<eps_var %#.2e>.dat and skips its iterations
 * when that succeeds, and saves the file after running them
 * (solver.h:352-375); bit 3: force_var, never load (still save).
 * energies[n_states] receives the final variational energies; the det list
 * stays in sys. */
int shci_host_run_variation(void* sys, const double* eps_vars, unsigned n_eps, const double* schedule,
                            unsigned n_sched, double target_error, unsigned n_states, int mode, int device,
                            int verbose, double* energies);

/* RDM file outputs (src/chem/rdm.cc), from shci_rdm's arrays: rdm1 n x n
 * row-major (spin-summed), rdm2 packed a = p n + s, b = q n + r, a >= b.
 * orb_order[i] = original (FCIDUMP) index of internal orbital i, as from
 * shci_host_system_integrals; NULL = identity.  Return 0, or -1 with
 * shci_host_io_last_error().
 *   shci_host_dump_1rdm  RDM::dump_1rdm (rdm.cc:202-221): "p,r,1rdm", p <= r,
 *                        |x| >= 1e-9, "%d,%d,%#.15g"
 *   shci_host_dump_2rdm  RDM::dump_2rdm (rdm.cc:271-331): csv != 0 ->
 *                        "p,q,r,s,2rdm" (p <= q, |x| >= 1e-9, %#.15g); csv == 0
 *                        -> spatialRDM.txt (n, then "p q s r value" in the
 *                        original order for |x| > 1e-6, %10.8g) */
int shci_host_dump_1rdm(const char* path, unsigned n_orbs, const unsigned* orb_order, const double* rdm1);
int shci_host_dump_2rdm(const char* path, unsigned n_orbs, const unsigned* orb_order, const double* rdm2, int csv);
const char* shci_host_io_last_error(void);

/* Wavefunction files wf_eps1_<eps_var>.dat (Solver::save_variation_result /
 * load_variation_result, solver.h:1179-1224): the fields of
 * BaseSystem::serialize (base_system.h:125-133) in the reference's hps
 * encoding (unsigned integers as base-128 varints, low group first; doubles
 * as 8 little-endian bytes; vectors as a varint size and the elements; a det
 * as its up and dn half-dets of 2 uint64 chunks each).  Dets of n_orb <= 64:
 * chunk 1 is written as 0, and a file with a nonzero chunk 1 is rejected.
 * coefs: n_states x n_dets row-major.  shci_host_wf_info reads the counts
 * (any pointer may be NULL) for sizing the buffers of shci_host_load_wf,
 * which rejects counts that differ from the file's.  Return 0, or -1 with
 * shci_host_wf_last_error(). */
int shci_host_save_wf(const char* path, unsigned n_up, unsigned n_dn, const uint64_t* up, const uint64_t* dn,
                      size_t n_dets, const double* coefs, unsigned n_states, double energy_hf,
                      const double* energy_var, int time_sym);
int shci_host_wf_info(const char* path, unsigned* n_up, unsigned* n_dn, size_t* n_dets, unsigned* n_states,
                      int* time_sym);
int shci_host_load_wf(const char* path, uint64_t* up, uint64_t* dn, size_t n_dets, double* coefs, unsigned n_states,
                      double* energy_hf, double* energy_var);
const char* shci_host_wf_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SHCI_HOST_H_ */
