// Drop-in C++ adapter: the reference's Hamiltonian<S> / SparseMatrix /
// Davidson call surface (src/solver/hamiltonian.h:14-79,
// src/solver/sparse_matrix.h:12-60, src/solver/davidson.h:6-29) and the
// selection sweep of its variational loop (src/solver/solver.h:435-495)
// implemented over the C-ABI of include/shci_hip.h.  Header-only; link
// libshci_hip.so.
//
// A system type S plugs in through shci_gpu::SystemTraits<S>, which must
// provide
//   static size_t n_dets(const S&);
//   static void dets(const S&, std::vector<uint64_t>& up, std::vector<uint64_t>& dn);
//   static void integrals(const S&, unsigned& n_orb, std::vector<double>& h1_dense,
//                         std::vector<double>& eri_packed, double& core);
//   static bool time_sym(const S&);
// (for the reference's ChemSystem: get_n_dets(), dets[i].up/.dn chunk 0 and
// Integrals::get_1b/get_2b over the combine4 keys, src/chem/integrals.h:42-44).
// Errors throw std::runtime_error like the reference.
#pragma once

#include <algorithm>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "shci_hip.h"

namespace shci_gpu {

inline void check(shci_ctx* ctx, int rc, const char* what) {
  if (rc != SHCI_OK) throw std::runtime_error(std::string(what) + ": " + shci_last_error(ctx));
}

// One device context per process (the reference's one-rank-per-process model).
inline shci_ctx* default_context(int device = 0) {
  static shci_ctx* ctx = nullptr;
  if (!ctx) {
    if (shci_ctx_create(&ctx, device) != SHCI_OK) throw std::runtime_error("shci_ctx_create failed");
  }
  return ctx;
}

template <class S>
struct SystemTraits;

// get_row() view (src/solver/sparse_vector.h:8-45 read interface).
class SparseVector {
 public:
  size_t size() const { return indices_.size(); }
  size_t get_index(size_t i) const { return indices_.at(i); }
  double get_value(size_t i) const { return values_.at(i); }
  std::vector<size_t> get_connections() const { return indices_; }
  std::vector<size_t> indices_;
  std::vector<double> values_;
};

class SparseMatrix {
 public:
  explicit SparseMatrix(shci_ctx* ctx = nullptr) : ctx_(ctx ? ctx : default_context()) {}

  // The diagonal is copied to the host once per build or update and served
  // from there: get_diag(i) is O(1), as in the reference (sparse_matrix.h:14),
  // whose callers read it element by element inside loops.
  double get_diag(size_t i) const { return diag_cache().at(i); }
  std::vector<double> get_diag() const { return diag_cache(); }
  size_t count_n_elems() const {
    uint64_t n, up, el;
    check(ctx_, shci_get_counts(ctx_, &n, &up, &el), "shci_get_counts");
    return el;
  }
  size_t count_n_rows() const {
    if (!built_) return 0;
    uint64_t n;
    check(ctx_, shci_get_counts(ctx_, &n, nullptr, nullptr), "shci_get_counts");
    return n;
  }
  std::vector<double> mul(const std::vector<double>& vec) const {
    std::vector<double> y(vec.size());
    check(ctx_, shci_spmv(ctx_, vec.data(), y.data(), 0), "shci_spmv");
    return y;
  }
  std::vector<std::complex<double>> mul(const std::vector<std::complex<double>>& vec) const {
    const size_t n = vec.size();
    std::vector<double> re(n), im(n);
    for (size_t i = 0; i < n; i++) re[i] = vec[i].real(), im[i] = vec[i].imag();
    const auto yr = mul(re), yi = mul(im);
    std::vector<std::complex<double>> y(n);
    for (size_t i = 0; i < n; i++) y[i] = {yr[i], yi[i]};
    return y;
  }
  SparseVector get_row(size_t i) const {
    export_cache();
    SparseVector r;
    for (int64_t k = rp_[i]; k < rp_[i + 1]; k++) {
      r.indices_.push_back((size_t)cols_[k]);
      r.values_.push_back(vals_[k]);
    }
    return r;
  }
  std::vector<std::vector<size_t>> get_connections() const {
    export_cache();
    std::vector<std::vector<size_t>> c(rp_.size() - 1);
    for (size_t i = 0; i + 1 < rp_.size(); i++)
      for (int64_t k = rp_[i]; k < rp_[i + 1]; k++) c[i].push_back((size_t)cols_[k]);
    return c;
  }
  void clear() {
    built_ = false;
    rp_.clear();
    cols_.clear();
    vals_.clear();
    diag_.clear();
    diag_ok_ = false;
  }
  shci_ctx* context() const { return ctx_; }
  // Called after every build and update_existing_elems: drops the host copies.
  void mark_built() {
    built_ = true;
    rp_.clear();
    diag_ok_ = false;
  }

 private:
  const std::vector<double>& diag_cache() const {
    if (!diag_ok_) {
      diag_.assign(count_n_rows(), 0.0);
      if (!diag_.empty()) check(ctx_, shci_get_diag(ctx_, diag_.data()), "shci_get_diag");
      diag_ok_ = true;
    }
    return diag_;
  }
  void export_cache() const {
    if (!rp_.empty()) return;
    uint64_t n, up, el;
    check(ctx_, shci_get_counts(ctx_, &n, &up, &el), "shci_get_counts");
    rp_.resize(n + 1);
    cols_.resize(up);
    vals_.resize(up);
    check(ctx_, shci_export_upper(ctx_, rp_.data(), cols_.data(), vals_.data()), "shci_export_upper");
  }
  shci_ctx* ctx_;
  bool built_ = false;
  mutable std::vector<int64_t> rp_, cols_;
  mutable std::vector<double> vals_;
  mutable std::vector<double> diag_;
  mutable bool diag_ok_ = false;
};

// hamiltonian.h:14-79.  store_flags: SHCI_STORE_FP64 and/or SHCI_STORE_FP32.
template <class S>
class Hamiltonian {
 public:
  explicit Hamiltonian(shci_ctx* ctx = nullptr, uint32_t store_flags = SHCI_STORE_FP64)
      : matrix(ctx), flags_(store_flags) {}

  SparseMatrix matrix;

  void update(const S& system) {
    using T = SystemTraits<S>;
    n_dets_prev_ = n_dets_;
    n_dets_ = T::n_dets(system);
    if (n_dets_prev_ == n_dets_) return;  // hamiltonian.h:91
    upload_integrals(system);
    std::vector<uint64_t> up, dn;
    T::dets(system, up, dn);
    // time_sym: the dets are the reference's canonical symmetrised states
    // (up <= dn); vectors stay in that basis (base_system.h:86-122)
    const uint32_t f = flags_ | (T::time_sym(system) ? SHCI_TIME_SYM : 0u);
    check(matrix.context(), shci_build_hamiltonian(matrix.context(), up.data(), dn.data(), n_dets_, n_dets_prev_, f),
          "shci_build_hamiltonian");
    matrix.mark_built();
  }

  void update_existing_elems(const S& system) {
    upload_integrals(system, true);
    check(matrix.context(), shci_update_existing_elems(matrix.context()), "shci_update_existing_elems");
    matrix.mark_built();
  }

  void clear() {
    n_dets_ = n_dets_prev_ = 0;
    ints_set_ = false;
    matrix.clear();
  }

 private:
  void upload_integrals(const S& system, bool force = false) {
    if (ints_set_ && !force) return;
    unsigned n_orb;
    std::vector<double> h1, eri;
    double core;
    SystemTraits<S>::integrals(system, n_orb, h1, eri, core);
    check(matrix.context(), shci_set_integrals(matrix.context(), n_orb, h1.data(), eri.data(), core),
          "shci_set_integrals");
    ints_set_ = true;
  }
  size_t n_dets_ = 0, n_dets_prev_ = 0;
  uint32_t flags_;
  bool ints_set_ = false;
};

// One heat-bath selection sweep of the variational loop on the device
// (solver.h:435-495, ChemSystem::find_connected_dets chem_system.cc:289-402):
// the new dets come back (up, dn) ascending in new_up / new_dn (the caller
// appends them, as the reference appends its hash set's dets); eps_tried_prev
// (one per det, 1e100 for dets never tried) is updated and grown with 1e100
// for the new dets.  coefs: n_states vectors of the current det count.
template <class S>
class Selector {
 public:
  explicit Selector(shci_ctx* ctx = nullptr) : ctx_(ctx ? ctx : default_context()) {}

  size_t select(const S& system, const std::vector<std::vector<double>>& coefs, double eps_var,
                std::vector<double>& eps_tried_prev, std::vector<uint64_t>& new_up, std::vector<uint64_t>& new_dn) {
    using T = SystemTraits<S>;
    if (!ints_set_) {
      unsigned n_orb;
      std::vector<double> h1, eri;
      double core;
      T::integrals(system, n_orb, h1, eri, core);
      check(ctx_, shci_set_integrals(ctx_, n_orb, h1.data(), eri.data(), core), "shci_set_integrals");
      ints_set_ = true;
    }
    std::vector<uint64_t> up, dn;
    T::dets(system, up, dn);
    const size_t n = up.size();
    eps_tried_prev.resize(n, 1.0e100);
    std::vector<double> c(coefs.size() * n, 0.0);
    for (size_t s = 0; s < coefs.size(); s++)
      std::copy(coefs[s].begin(), coefs[s].begin() + std::min(n, coefs[s].size()), c.begin() + s * n);
    uint64_t m = 0;
    check(ctx_,
          shci_select(ctx_, up.data(), dn.data(), n, c.data(), (uint32_t)coefs.size(), eps_var,
                      eps_tried_prev.data(), &m, T::time_sym(system) ? SHCI_TIME_SYM : 0u),
          "shci_select");
    new_up.resize(m);
    new_dn.resize(m);
    check(ctx_, shci_select_result(ctx_, new_up.data(), new_dn.data(), m), "shci_select_result");
    eps_tried_prev.resize(n + m, 1.0e100);
    return m;
  }

 private:
  shci_ctx* ctx_;
  bool ints_set_ = false;
};

// davidson.h:6-29.  mode: SHCI_DAVIDSON_FP64 (reference algorithm) or SHCI_DAVIDSON_MIXED.
class Davidson {
 public:
  explicit Davidson(unsigned n_states, uint32_t mode = SHCI_DAVIDSON_FP64)
      : n_states_(n_states), mode_(mode), lowest_eigenvalues_(n_states), lowest_eigenvectors_(n_states) {}

  void diagonalize(const SparseMatrix& matrix, const std::vector<std::vector<double>>& initial_vectors,
                   const double target_error, const bool verbose = false) {
    const size_t dim = initial_vectors[0].size();
    const unsigned ns = (unsigned)std::min<size_t>(std::min<size_t>(dim, initial_vectors.size()), n_states_);
    std::vector<double> init(ns * dim), evals(ns), evecs(ns * dim);
    for (unsigned s = 0; s < ns; s++) std::copy(initial_vectors[s].begin(), initial_vectors[s].end(), init.begin() + s * dim);
    std::vector<double> te(64), tt(64);
    int nt = 0, conv = 0;
    check(matrix.context(),
          shci_davidson(matrix.context(), ns, init.data(), target_error, mode_, 1, evals.data(), evecs.data(), &conv,
                        te.data(), tt.data(), 64, &nt),
          "shci_davidson");
    if (verbose)
      for (int k = 0; k < nt; k++) printf("Davidson #%d:  %.10f\n", k, te[k]);
    converged = conv != 0;
    for (unsigned s = 0; s < ns; s++) {
      lowest_eigenvalues_[s] = evals[s];
      lowest_eigenvectors_[s].assign(evecs.begin() + s * dim, evecs.begin() + (s + 1) * dim);
    }
    for (unsigned s = ns; s < initial_vectors.size() && s < n_states_; s++) lowest_eigenvectors_[s] = initial_vectors[s];
  }
  std::vector<double> get_lowest_eigenvalues() const { return lowest_eigenvalues_; }
  std::vector<std::vector<double>> get_lowest_eigenvectors() const { return lowest_eigenvectors_; }
  bool converged = false;

 private:
  unsigned n_states_;
  uint32_t mode_;
  std::vector<double> lowest_eigenvalues_;
  std::vector<std::vector<double>> lowest_eigenvectors_;
};

}  // namespace shci_gpu
