// The reference's variational loop (src/solver/solver.h:338-539) with the
// heat-bath selection (host queue walk, or the device through the adapter)
// feeding the GPU Hamiltonian build and Davidson through the drop-in adapter
// (include/shci_gpu/shci_gpu.h).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <exception>
#include <stdexcept>
#include <vector>

#include "host_state.h"
#include "shci_gpu/shci_gpu.h"
#include "shci_host.h"

namespace shci_host {

// The system type handed to Hamiltonian<S>::update.
struct VarSystem {
  const Integrals* ints;
  const std::vector<Det>* dets;
};

}  // namespace shci_host

namespace shci_gpu {
template <>
struct SystemTraits<shci_host::VarSystem> {
  static size_t n_dets(const shci_host::VarSystem& s) { return s.dets->size(); }
  static void dets(const shci_host::VarSystem& s, std::vector<uint64_t>& up, std::vector<uint64_t>& dn) {
    up.resize(s.dets->size());
    dn.resize(s.dets->size());
    for (size_t i = 0; i < s.dets->size(); i++) up[i] = (*s.dets)[i].up, dn[i] = (*s.dets)[i].dn;
  }
  static void integrals(const shci_host::VarSystem& s, unsigned& n_orb, std::vector<double>& h1,
                        std::vector<double>& eri, double& core) {
    const auto& I = *s.ints;
    n_orb = I.n_orbs;
    h1.resize((size_t)n_orb * n_orb);
    for (unsigned p = 0; p < n_orb; p++)
      for (unsigned q = 0; q < n_orb; q++) h1[(size_t)p * n_orb + q] = I.get_1b(p, q);
    eri = I.eri;
    core = I.energy_core;
  }
  static bool time_sym(const shci_host::VarSystem&) { return false; }
};
}  // namespace shci_gpu

using namespace shci_host;

extern "C" int shci_host_run_variation(void* h, const double* eps_vars, unsigned n_eps, const double* schedule,
                                       unsigned n_sched, double target_error, unsigned n_states, int mode,
                                       int device, int verbose, double* energies) {
  auto* s = static_cast<HostSystem*>(h);
  try {
    shci_ctx* ctx = shci_gpu::default_context(device);
    const bool mixed = (mode & 1) != 0, gpu_select = (mode & 2) != 0;
    const uint32_t store = mixed ? (SHCI_STORE_FP32 | SHCI_STORE_FP64) : SHCI_STORE_FP64;
    shci_gpu::Hamiltonian<VarSystem> hamiltonian(ctx, store);
    shci_gpu::Selector<VarSystem> selector(ctx);
    std::vector<uint64_t> new_up, new_dn;
    VarSystem sys{&s->ints, &s->dets};
    s->dets.assign(1, Det{s->ints.hf_up, s->ints.hf_dn});
    s->eps_tried.assign(1, 1.0e100);
    std::vector<std::vector<double>> coefs(n_states, std::vector<double>(1, 1e-16));
    coefs[0][0] = 1.0;
    std::vector<double> energy_var(n_states, 0.0);
    double eps_prev = 1e100;
    unsigned it_sched = 0;
    size_t iteration_global = 0;
    for (unsigned e = 0; e < n_eps; e++) {
      const double eps_var = eps_vars[e];
      std::vector<std::pair<double, bool>> plan;
      while (it_sched < n_sched && schedule[it_sched] >= eps_prev) it_sched++;
      while (it_sched < n_sched && schedule[it_sched] > eps_var) plan.push_back({schedule[it_sched++], false});
      plan.push_back({eps_var, true});
      for (const auto& step : plan) {
        const double eps = step.first;
        const bool until = step.second;
        shci_gpu::Davidson davidson(n_states, mixed ? SHCI_DAVIDSON_MIXED : SHCI_DAVIDSON_FP64);
        size_t n_dets = s->dets.size();
        bool dets_converged = false, converged = false;
        while (!converged) {
          if (!dets_converged) {
            size_t n_new;
            if (gpu_select) {
              n_new = selector.select(sys, coefs, eps, s->eps_tried, new_up, new_dn);
              for (size_t i = 0; i < n_new; i++) s->dets.push_back(Det{new_up[i], new_dn[i]});
            } else {
              n_new = s->sel->select(s->dets, coefs, eps, s->eps_tried);
            }
            for (auto& c : coefs) c.resize(s->dets.size(), 1e-16);
            if (verbose) printf("Number of dets / new dets: %zu / %zu\n", s->dets.size(), n_new);
            hamiltonian.update(sys);
          }
          const double tol = until ? target_error / 500000 : target_error / 50;
          davidson.diagonalize(hamiltonian.matrix, coefs, tol, verbose != 0);
          energy_var = davidson.get_lowest_eigenvalues();
          coefs = davidson.get_lowest_eigenvectors();
          iteration_global++;
          if (verbose) printf("Iteration %zu eps1= %#.2e ndets= %zu energy= %.8f\n", iteration_global, eps,
                              s->dets.size(), energy_var[0]);
          if (s->dets.size() < n_dets * 1.001) dets_converged = true;
          if (dets_converged && davidson.converged) converged = true;
          n_dets = s->dets.size();
          if (!until) break;
        }
      }
      eps_prev = eps_var;
    }
    for (unsigned k = 0; k < n_states; k++) energies[k] = energy_var[k];
    return 0;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return -1;
  }
}
