// The reference's variational loop (src/solver/solver.h:338-539) with the
// heat-bath selection (host queue walk, or the device through the adapter)
// feeding the GPU Hamiltonian build and Davidson through the drop-in adapter
// (include/shci_gpu/shci_gpu.h).
#include <algorithm>
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

namespace {

// Solver::load_variation_result (solver.h:1179-1215): false when the file
// cannot be opened (the variation then runs); a file that opens but does not
// parse, or holds another system's shape, is an error.
bool load_wf(const char* name, HostSystem& s, unsigned n_states, std::vector<std::vector<double>>& coefs,
             std::vector<double>& energy_var) {
  if (FILE* f = std::fopen(name, "rb")) {
    std::fclose(f);
  } else {
    return false;
  }
  unsigned n_up, n_dn, ns;
  size_t n;
  int ts;
  if (shci_host_wf_info(name, &n_up, &n_dn, &n, &ns, &ts)) throw std::runtime_error(shci_host_wf_last_error());
  if (n_up != s.n_up || n_dn != s.n_dn || ns != n_states || ts || n == 0)
    throw std::runtime_error(std::string(name) + ": holds another system (electrons, states, time_sym or no dets)");
  std::vector<uint64_t> up(n), dn(n);
  std::vector<double> c((size_t)ns * n), ev(ns);
  double e_hf;
  if (shci_host_load_wf(name, up.data(), dn.data(), n, c.data(), ns, &e_hf, ev.data()))
    throw std::runtime_error(shci_host_wf_last_error());
  s.dets.resize(n);
  for (size_t i = 0; i < n; i++) s.dets[i] = Det{up[i], dn[i]};
  s.eps_tried.assign(n, 1.0e100);
  coefs.assign(ns, std::vector<double>(n));
  for (unsigned k = 0; k < ns; k++) std::copy(c.begin() + (size_t)k * n, c.begin() + (size_t)(k + 1) * n, coefs[k].begin());
  energy_var = ev;
  return true;
}

// Solver::save_variation_result (solver.h:1217-1224).
void save_wf(const char* name, const HostSystem& s, const std::vector<std::vector<double>>& coefs,
             const std::vector<double>& energy_var) {
  const size_t n = s.dets.size();
  std::vector<uint64_t> up(n), dn(n);
  for (size_t i = 0; i < n; i++) up[i] = s.dets[i].up, dn[i] = s.dets[i].dn;
  std::vector<double> c(coefs.size() * n);
  for (const auto& ck : coefs)
    if (ck.size() != n) throw std::runtime_error("save_wf: coefficients and dets differ in length");
  for (size_t k = 0; k < coefs.size(); k++) std::copy(coefs[k].begin(), coefs[k].end(), c.begin() + k * n);
  if (shci_host_save_wf(name, s.n_up, s.n_dn, up.data(), dn.data(), n, c.data(), (unsigned)coefs.size(),
                        s.sel->hf_energy(), energy_var.data(), 0))
    throw std::runtime_error(shci_host_wf_last_error());
}

}  // namespace

extern "C" int shci_host_run_variation(void* h, const double* eps_vars, unsigned n_eps, const double* schedule,
                                       unsigned n_sched, double target_error, unsigned n_states, int mode,
                                       int device, int verbose, double* energies) {
  auto* s = static_cast<HostSystem*>(h);
  try {
    shci_ctx* ctx = shci_gpu::default_context(device);
    const bool mixed = (mode & 1) != 0, gpu_select = (mode & 2) != 0;
    const bool wf_files = (mode & 4) != 0, force_var = (mode & 8) != 0;
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
      char wf_name[64];
      std::snprintf(wf_name, sizeof wf_name, "wf_eps1_%#.2e.dat", eps_var);
      if (wf_files && !force_var && load_wf(wf_name, *s, n_states, coefs, energy_var)) {
        // solver.h:374-377: the loaded list restarts the selection bookkeeping.
        hamiltonian.clear();
        if (verbose) printf("Loaded %zu dets from: %s\n", s->dets.size(), wf_name);
        eps_prev = eps_var;
        continue;
      }
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
      if (wf_files) save_wf(wf_name, *s, coefs, energy_var);
      eps_prev = eps_var;
    }
    for (unsigned k = 0; k < n_states; k++) energies[k] = energy_var[k];
    return 0;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return -1;
  }
}
