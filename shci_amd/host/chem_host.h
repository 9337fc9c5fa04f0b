// Host-side chemistry system for the GPU variational path: integrals (FCIDUMP
// semantics of src/chem/integrals.cc), Slater-Condon elements needed by the
// CPU-side heat-bath selection, and the selection itself
// (src/chem/chem_system.cc:74-402, src/solver/solver.h:413-539).  Selection
// and PT2 stay on the CPU; the Hamiltonian build and Davidson run on the GPU
// through the C-ABI in include/shci_hip.h.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace shci_host {

// Integral keys, src/chem/integrals.cc:410-422.
inline size_t combine2(size_t a, size_t b) {
  return a > b ? (a * (a + 1)) / 2 + b : (b * (b + 1)) / 2 + a;
}
inline size_t combine4(size_t a, size_t b, size_t c, size_t d) {
  return combine2(combine2(a, b), combine2(c, d));
}

struct Det {
  uint64_t up, dn;
  bool operator==(const Det& o) const { return up == o.up && dn == o.dn; }
  bool operator<(const Det& o) const { return up < o.up || (up == o.up && dn < o.dn); }
};

// One-/two-electron integrals in the reference's keyed layout: h1 by combine2,
// eri by combine4 (the 8-fold unique chemists' integrals), core energy.
struct Integrals {
  unsigned n_orbs = 0;
  unsigned n_elecs = 0;
  double energy_core = 0.0;
  std::vector<double> h1;
  std::vector<double> eri;
  std::vector<unsigned> orb_order;  // new index -> FCIDUMP orbital (0-based)
  uint64_t hf_up = 0, hf_dn = 0;

  void resize(unsigned n);
  double get_1b(unsigned p, unsigned q) const { return h1[combine2(p, q)]; }
  double get_2b(unsigned p, unsigned q, unsigned r, unsigned s) const {
    return eri[combine4(p, q, r, s)];
  }
  // Reads an FCIDUMP with the reference's semantics (|x| < 1e-9 dropped,
  // duplicates keep the larger |x|), builds the HF det (lowest-filled or the
  // explicit occupations) and optionally reorders orbitals by HF orbital energy.
  // reorder: -1 = reference default (!explicit occupations), 0 = off, 1 = on.
  void load_fcidump(const std::string& path, unsigned n_up, unsigned n_dn,
                    const std::vector<unsigned>& occs_up, const std::vector<unsigned>& occs_dn,
                    int reorder);
  void set_arrays(unsigned n, const double* h1_dense, const double* eri_packed, double core);
};

// Heat-bath selection state for one system (src/chem/chem_system.cc:74-196).
class Selector {
 public:
  Selector(const Integrals& ints, unsigned n_up, unsigned n_dn);
  // One selection sweep of run_variation (src/solver/solver.h:435-495): for
  // every det i with eps_min = eps_var / max_state |c_i| below eps_tried_prev[i],
  // generate connected dets above eps_min that are not yet in the variational
  // space.  New dets are appended to `dets` (sorted by (up, dn) within the sweep
  // so the order is reproducible) and eps_tried_prev is updated.  Returns the
  // number of new dets.
  size_t select(std::vector<Det>& dets, const std::vector<std::vector<double>>& coefs,
                double eps_var, std::vector<double>& eps_tried_prev);
  double hf_energy() const;
  double elem_single(const Det& di, const Det& dj) const;  // H_ij for a single excitation

 private:
  const Integrals& ints_;
  unsigned n_up_, n_dn_, n_orbs_;
  struct Sr {
    double S;
    unsigned r;
  };
  struct Hrs {
    double H;
    unsigned r, s;
  };
  std::vector<std::vector<Sr>> singles_queue_;
  std::vector<std::vector<Hrs>> hci_queue_;
  double max_singles_ = 0.0, max_hci_ = 0.0;
  void setup_singles_queue();
  void setup_hci_queue();
  template <class F>
  void find_connected(const Det& det, double eps_max, double eps_min, F&& handler) const;
};

}  // namespace shci_host
