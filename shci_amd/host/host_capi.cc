// C API of the host library (libshci_host.so): integrals + CPU heat-bath
// selection state, for the Python side and the C++ variational driver.
// Declared in include/shci_host.h.
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_state.h"

using namespace shci_host;

namespace shci_host {
thread_local std::string g_host_err;
}  // namespace shci_host

extern "C" {

const char* shci_host_last_error() { return g_host_err.c_str(); }

void* shci_host_system_from_fcidump(const char* path, unsigned n_up, unsigned n_dn,
                                    const unsigned* occs_up, unsigned n_occs_up,
                                    const unsigned* occs_dn, unsigned n_occs_dn, int reorder) {
  try {
    auto* s = new HostSystem();
    s->n_up = n_up;
    s->n_dn = n_dn;
    s->ints.load_fcidump(path, n_up, n_dn, std::vector<unsigned>(occs_up, occs_up + n_occs_up),
                         std::vector<unsigned>(occs_dn, occs_dn + n_occs_dn), reorder);
    s->sel = new Selector(s->ints, n_up, n_dn);
    return s;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return nullptr;
  }
}

void* shci_host_system_from_arrays(unsigned n_orbs, unsigned n_up, unsigned n_dn,
                                   const double* h1_dense, const double* eri_packed, double core,
                                   uint64_t hf_up, uint64_t hf_dn) {
  try {
    if (n_orbs > 64) throw std::runtime_error("n_orbs > 64 unsupported");
    auto* s = new HostSystem();
    s->n_up = n_up;
    s->n_dn = n_dn;
    s->ints.set_arrays(n_orbs, h1_dense, eri_packed, core);
    s->ints.hf_up = hf_up;
    s->ints.hf_dn = hf_dn;
    s->sel = new Selector(s->ints, n_up, n_dn);
    return s;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return nullptr;
  }
}

void shci_host_system_free(void* h) { delete static_cast<HostSystem*>(h); }

int shci_host_system_info(void* h, unsigned* n_orbs, double* core, uint64_t* hf_up, uint64_t* hf_dn,
                          size_t* eri_size, double* e_hf) {
  auto* s = static_cast<HostSystem*>(h);
  *n_orbs = s->ints.n_orbs;
  *core = s->ints.energy_core;
  *hf_up = s->ints.hf_up;
  *hf_dn = s->ints.hf_dn;
  *eri_size = s->ints.eri.size();
  *e_hf = s->sel->hf_energy();
  return 0;
}

int shci_host_system_integrals(void* h, double* h1_dense, double* eri_packed, unsigned* orb_order) {
  auto* s = static_cast<HostSystem*>(h);
  const unsigned n = s->ints.n_orbs;
  for (unsigned p = 0; p < n; p++)
    for (unsigned q = 0; q < n; q++) h1_dense[(size_t)p * n + q] = s->ints.get_1b(p, q);
  std::memcpy(eri_packed, s->ints.eri.data(), s->ints.eri.size() * sizeof(double));
  if (orb_order) std::memcpy(orb_order, s->ints.orb_order.data(), n * sizeof(unsigned));
  return 0;
}

// Reset the variational space to the HF determinant.
int shci_host_reset_dets(void* h) {
  auto* s = static_cast<HostSystem*>(h);
  s->dets.assign(1, Det{s->ints.hf_up, s->ints.hf_dn});
  s->eps_tried.assign(1, 1.0e100);
  return 0;
}

size_t shci_host_n_dets(void* h) { return static_cast<HostSystem*>(h)->dets.size(); }

int shci_host_get_dets(void* h, uint64_t* up, uint64_t* dn) {
  auto* s = static_cast<HostSystem*>(h);
  for (size_t i = 0; i < s->dets.size(); i++) up[i] = s->dets[i].up, dn[i] = s->dets[i].dn;
  return 0;
}

int shci_host_set_dets(void* h, const uint64_t* up, const uint64_t* dn, size_t n) {
  auto* s = static_cast<HostSystem*>(h);
  s->dets.resize(n);
  for (size_t i = 0; i < n; i++) s->dets[i] = Det{up[i], dn[i]};
  s->eps_tried.assign(n, 1.0e100);
  return 0;
}

// One selection sweep with coefficients coefs[n_states][n_dets] (row-major).
// Returns the number of new determinants (appended), or -1 on error.
long long shci_host_select(void* h, const double* coefs, unsigned n_states, double eps_var) {
  auto* s = static_cast<HostSystem*>(h);
  try {
    const size_t n = s->dets.size();
    std::vector<std::vector<double>> c(n_states, std::vector<double>(n));
    for (unsigned k = 0; k < n_states; k++) std::memcpy(c[k].data(), coefs + (size_t)k * n, n * sizeof(double));
    return (long long)s->sel->select(s->dets, c, eps_var, s->eps_tried);
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return -1;
  }
}

}  // extern "C"
