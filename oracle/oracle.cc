// CPU oracle: a plain C++17/OpenMP restatement of the reference SHCI ("Arrow")
// variational hot path.  THIS FILE IS TEST INFRASTRUCTURE.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load liboracle.so,
// and only as the checker (or the timed CPU baseline); the product path never
// links, loads or calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference repository root).  Determinants are restricted to n_orb <= 64
// (one uint64 per spin), which covers every workload this build targets; the
// reference's second chunk (src/det/half_det.h:10) is therefore always zero.
//
// Pinning (what checks each part) is listed in the header of oracle/__init__.py.

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

namespace orc {

constexpr double EPS = 1.0e-12;  // src/util.h:50

// ---------------------------------------------------------------------------
// Integral keys: src/chem/integrals.cc:410-422
// ---------------------------------------------------------------------------
static inline size_t combine2(size_t a, size_t b) {
  return a > b ? (a * (a + 1)) / 2 + b : (b * (b + 1)) / 2 + a;
}
static inline size_t combine4(size_t a, size_t b, size_t c, size_t d) {
  return combine2(combine2(a, b), combine2(c, d));
}

struct Integrals {
  unsigned n_orbs = 0;
  unsigned n_elecs = 0;
  double energy_core = 0.0;
  std::vector<double> h1;   // keyed by combine2 (missing key -> 0)
  std::vector<double> eri;  // keyed by combine4 (missing key -> 0)
  std::vector<unsigned> orb_order, orb_order_inv;
  uint64_t hf_up = 0, hf_dn = 0;

  void resize(unsigned n) {
    n_orbs = n;
    const size_t np = (size_t)n * (n + 1) / 2;
    h1.assign(np, 0.0);
    eri.assign(np * (np + 1) / 2, 0.0);
  }
  // src/chem/integrals.cc:390-393
  double get_1b(unsigned p, unsigned q) const { return h1[combine2(p, q)]; }
  // src/chem/integrals.cc:395-408 (no Dooh/Coov lz filter: C1/D2h-type groups only)
  double get_2b(unsigned p, unsigned q, unsigned r, unsigned s) const {
    return eri[combine4(p, q, r, s)];
  }
};

// Reducer used on duplicate keys: keep the larger |x| (src/chem/integrals.cc:104-111).
static inline void set_max_abs(std::vector<double>& store, size_t key, double x) {
  double& a = store[key];
  if (a == 0.0) {
    a = x;
  } else if (std::abs(a) < std::abs(x)) {
    a = x;
  }
}

struct RawIntegral {
  double H;
  unsigned p, q, r, s;
};

// FCIDUMP reader: src/chem/integrals.cc:37-114.  Header tokens split on
// whitespace, ',', '=' (the reference's regex also splits on an inner '-').
static void read_fcidump(const std::string& path, Integrals& ints, std::vector<RawIntegral>& raw) {
  std::ifstream f(path);
  if (!f.good()) throw std::runtime_error("cannot open FCIDUMP");
  std::string line;
  bool in_orbsym = false, done = false;
  unsigned n_sym_read = 0;
  unsigned n_orbs = 0, n_elecs = 0;
  while (!done && std::getline(f, line)) {
    std::string tok;
    std::vector<std::string> words;
    for (char c : line) {
      if (c == ' ' || c == '\t' || c == ',' || c == '=' || c == '\r') {
        if (!tok.empty()) words.push_back(tok), tok.clear();
      } else {
        tok.push_back(c);
      }
    }
    if (!tok.empty()) words.push_back(tok);
    for (size_t w = 0; w < words.size(); w++) {
      const std::string& m = words[w];
      if (m == "NORB" || m == "NORBS") {
        n_orbs = (unsigned)std::stoul(words.at(++w));
      } else if (m == "NELEC") {
        n_elecs = (unsigned)std::stoul(words.at(++w));
      } else if (m == "ORBSYM") {
        in_orbsym = true;
      } else if (m == "&END" || m == "/" || m == "END") {
        done = true;
        break;
      } else if (in_orbsym) {
        bool numeric = !m.empty() && (std::isdigit((unsigned char)m[0]) || m[0] == '-');
        if (numeric) {
          n_sym_read++;
          if (n_sym_read == n_orbs) in_orbsym = false;
        } else {
          in_orbsym = false;
        }
      }
    }
  }
  if (n_orbs == 0) throw std::runtime_error("FCIDUMP: NORB missing");
  ints.resize(n_orbs);
  ints.n_elecs = n_elecs;
  double x;
  unsigned p, q, r, s;
  while (f >> x >> p >> q >> r >> s) {
    if (std::abs(x) < 1.0e-9) continue;  // integrals.cc:89
    raw.push_back({x, p, q, r, s});
  }
}

static void key_integrals(Integrals& ints, const std::vector<RawIntegral>& raw,
                          const std::vector<unsigned>& order_inv) {
  // src/chem/integrals.cc:94-113 and 352-381 (re-keyed through orb_order_inv)
  std::fill(ints.h1.begin(), ints.h1.end(), 0.0);
  std::fill(ints.eri.begin(), ints.eri.end(), 0.0);
  for (const auto& it : raw) {
    if (it.p == 0 && it.q == 0 && it.r == 0 && it.s == 0) {
      ints.energy_core = it.H;
    } else if (it.r == 0 && it.s == 0) {
      set_max_abs(ints.h1, combine2(order_inv[it.p - 1], order_inv[it.q - 1]), it.H);
    } else {
      set_max_abs(ints.eri,
                  combine4(order_inv[it.p - 1], order_inv[it.q - 1], order_inv[it.r - 1],
                           order_inv[it.s - 1]),
                  it.H);
    }
  }
}

// Orbital energies for the reorder: src/chem/integrals.cc:228-255
static std::vector<double> orb_energies(const Integrals& ints) {
  std::vector<double> e(ints.n_orbs);
  std::vector<unsigned> up, dn;
  for (unsigned i = 0; i < 64; i++) {
    if (ints.hf_up >> i & 1) up.push_back(i);
    if (ints.hf_dn >> i & 1) dn.push_back(i);
  }
  for (unsigned i = 0; i < ints.n_orbs; i++) {
    e[i] = ints.get_1b(i, i);
    double direct = 0, exchange = 0;
    for (unsigned orb : up) {
      if (orb == i) {
        direct += ints.get_2b(i, i, orb, orb);
      } else {
        direct += 2 * ints.get_2b(i, i, orb, orb);
        exchange -= ints.get_2b(i, orb, orb, i);
      }
    }
    for (unsigned orb : dn) {
      if (orb == i) {
        direct += ints.get_2b(i, i, orb, orb);
      } else {
        direct += 2 * ints.get_2b(i, i, orb, orb);
        exchange -= ints.get_2b(i, orb, orb, i);
      }
    }
    e[i] += 0.5 * (direct + exchange);
  }
  return e;
}

// Integrals::load minus the cache: read_fcidump, generate_det_hf (lowest-filled
// or explicit occupations, src/chem/integrals.cc:155-226), reorder_orbs (257-384).
static void load_integrals(const std::string& path, unsigned n_up, unsigned n_dn,
                           const std::vector<unsigned>& occs_up, const std::vector<unsigned>& occs_dn,
                           int reorder_flag, Integrals& ints) {
  std::vector<RawIntegral> raw;
  read_fcidump(path, ints, raw);
  std::vector<unsigned> ident(ints.n_orbs);
  std::iota(ident.begin(), ident.end(), 0u);
  key_integrals(ints, raw, ident);
  bool explicit_orbs = false;
  ints.hf_up = ints.hf_dn = 0;
  if (!occs_up.empty() || !occs_dn.empty()) {
    for (unsigned o : occs_up) ints.hf_up |= 1ull << o;
    for (unsigned o : occs_dn) ints.hf_dn |= 1ull << o;
    explicit_orbs = true;
  } else {
    for (unsigned i = 0; i < n_up; i++) ints.hf_up |= 1ull << i;
    for (unsigned i = 0; i < n_dn; i++) ints.hf_dn |= 1ull << i;
  }
  const auto e = orb_energies(ints);
  ints.orb_order = ident;
  const bool reorder = reorder_flag < 0 ? !explicit_orbs : reorder_flag != 0;
  if (reorder) {
    std::stable_sort(ints.orb_order.begin(), ints.orb_order.end(),
                     [&](unsigned a, unsigned b) { return e[a] < e[b] - EPS; });
  }
  ints.orb_order_inv.assign(ints.n_orbs, 0);
  for (unsigned i = 0; i < ints.n_orbs; i++) ints.orb_order_inv[ints.orb_order[i]] = i;
  uint64_t up = 0, dn = 0;
  for (unsigned i = 0; i < ints.n_orbs; i++) {
    if (ints.hf_up >> ints.orb_order[i] & 1) up |= 1ull << i;
    if (ints.hf_dn >> ints.orb_order[i] & 1) dn |= 1ull << i;
  }
  ints.hf_up = up;
  ints.hf_dn = dn;
  key_integrals(ints, raw, ints.orb_order_inv);
}

// ---------------------------------------------------------------------------
// HalfDet::diff for one 64-bit chunk: src/det/half_det.cc:120-206
// ---------------------------------------------------------------------------
struct DiffResult {
  unsigned left_only[2] = {0, 0};
  unsigned right_only[2] = {0, 0};
  unsigned n_diffs = 0;
  int permutation_factor = 1;
};

static inline DiffResult diff(uint64_t left, uint64_t right) {
  DiffResult res;
  if (left == right) return res;
  unsigned n_left_only = 0, n_right_only = 0, helper = 0;
  uint64_t lo = left & ~right;
  while (lo) {
    if (n_left_only >= 2) {
      res.n_diffs = 3;
      return res;
    }
    const unsigned tz = __builtin_ctzll(lo);
    const uint64_t bit = 1ull << tz;
    lo &= ~bit;
    res.left_only[n_left_only++] = tz;
    helper += __builtin_popcountll(left & (bit - 1));
  }
  uint64_t ro = right & ~left;
  while (ro) {
    if (n_right_only >= 2) {
      res.n_diffs = 3;
      return res;
    }
    const unsigned tz = __builtin_ctzll(ro);
    const uint64_t bit = 1ull << tz;
    ro &= ~bit;
    res.right_only[n_right_only++] = tz;
    helper += __builtin_popcountll(right & (bit - 1));
  }
  res.n_diffs = n_left_only;
  if (helper & 1) res.permutation_factor = -1;
  return res;
}

static inline std::vector<unsigned> occ(uint64_t h) {
  std::vector<unsigned> r;
  while (h) {
    const unsigned tz = __builtin_ctzll(h);
    h &= h - 1;
    r.push_back(tz);
  }
  return r;
}

struct Det {
  uint64_t up, dn;
  bool operator==(const Det& o) const { return up == o.up && dn == o.dn; }
};
struct DetHash {
  size_t operator()(const Det& d) const {
    return std::hash<uint64_t>()(d.up) * 1000003u ^ std::hash<uint64_t>()(d.dn);
  }
};

// ---------------------------------------------------------------------------
// Slater-Condon elements: src/chem/chem_system.cc:409-609
// ---------------------------------------------------------------------------
struct Chem {
  const Integrals* ints;

  // chem_system.cc:444-457
  double one_body_diag(const Det& d) const {
    double e = 0.0;
    for (unsigned o : occ(d.up)) e += ints->get_1b(o, o);
    if (d.up == d.dn) {
      e *= 2;
    } else {
      for (unsigned o : occ(d.dn)) e += ints->get_1b(o, o);
    }
    return e;
  }
  // chem_system.cc:495-542 (diag_helper is empty during variation)
  double two_body_diag(const Det& d) const {
    const auto ou = occ(d.up), od = occ(d.dn);
    double direct = 0.0, exchange = 0.0;
    for (size_t i = 0; i < ou.size(); i++)
      for (size_t j = i + 1; j < ou.size(); j++) {
        direct += ints->get_2b(ou[i], ou[i], ou[j], ou[j]);
        exchange -= ints->get_2b(ou[i], ou[j], ou[j], ou[i]);
      }
    if (d.up == d.dn) {
      direct *= 2;
      exchange *= 2;
    } else {
      for (size_t i = 0; i < od.size(); i++)
        for (size_t j = i + 1; j < od.size(); j++) {
          direct += ints->get_2b(od[i], od[i], od[j], od[j]);
          exchange -= ints->get_2b(od[i], od[j], od[j], od[i]);
        }
    }
    for (size_t i = 0; i < ou.size(); i++)
      for (size_t j = 0; j < od.size(); j++) direct += ints->get_2b(ou[i], ou[i], od[j], od[j]);
    return direct + exchange;
  }
  // chem_system.cc:544-550
  double one_body_single(const DiffResult& du, const DiffResult& dd) const {
    const DiffResult& d = du.n_diffs == 1 ? du : dd;
    return d.permutation_factor * ints->get_1b(d.left_only[0], d.right_only[0]);
  }
  // chem_system.cc:552-577
  double two_body_single(const Det& di, const DiffResult& du, const DiffResult& dd) const {
    const bool is_up = du.n_diffs == 1;
    const DiffResult& d = is_up ? du : dd;
    const unsigned oi = d.left_only[0], oj = d.right_only[0];
    const uint64_t same = is_up ? di.up : di.dn;
    uint64_t oppo = is_up ? di.dn : di.up;
    double e = 0.0;
    for (unsigned o : occ(same)) {
      if (o == oi || o == oj) continue;
      e -= ints->get_2b(oi, o, o, oj);
      const double direct = ints->get_2b(oi, oj, o, o);
      if (oppo >> o & 1) {
        oppo &= ~(1ull << o);
        e += 2 * direct;
      } else {
        e += direct;
      }
    }
    for (unsigned o : occ(oppo)) e += ints->get_2b(oi, oj, o, o);
    e *= d.permutation_factor;
    return e;
  }
  // chem_system.cc:579-609
  double two_body_double(const DiffResult& du, const DiffResult& dd) const {
    double e = 0.0;
    if (du.n_diffs == 0) {
      if (dd.n_diffs != 2) return 0.0;
      e = ints->get_2b(dd.left_only[0], dd.right_only[0], dd.left_only[1], dd.right_only[1]) -
          ints->get_2b(dd.left_only[0], dd.right_only[1], dd.left_only[1], dd.right_only[0]);
      e *= dd.permutation_factor;
    } else if (dd.n_diffs == 0) {
      if (du.n_diffs != 2) return 0.0;
      e = ints->get_2b(du.left_only[0], du.right_only[0], du.left_only[1], du.right_only[1]) -
          ints->get_2b(du.left_only[0], du.right_only[1], du.left_only[1], du.right_only[0]);
      e *= du.permutation_factor;
    } else {
      if (du.n_diffs != 1 || dd.n_diffs != 1) return 0.0;
      e = ints->get_2b(du.left_only[0], du.right_only[0], dd.left_only[0], dd.right_only[0]);
      e *= du.permutation_factor * dd.permutation_factor;
    }
    return e;
  }
  // base_system.h:86-101: the element between time-reversal-symmetrised
  // states (dets with up != dn stand for (|up,dn> + |dn,up>) / sqrt2).
  double elem_time_sym(const Det& di, const Det& dj, int n_excite) const {
    constexpr double SQRT2 = 1.4142135623730951;  // src/util.h:58
    double h = elem(di, dj, n_excite);
    if (di.up == di.dn) {
      if (dj.up != dj.dn) h *= SQRT2;
    } else {
      if (dj.up == dj.dn) {
        h *= SQRT2;
      } else {
        const Det di_rev{di.dn, di.up};
        h += elem(di_rev, dj, -1);
      }
    }
    return h;
  }
  // chem_system.cc:409-442
  double elem(const Det& di, const Det& dj, int n_excite) const {
    DiffResult du, dd;
    if (n_excite < 0) {
      du = diff(di.up, dj.up);
      if (du.n_diffs > 2) return 0.0;
      dd = diff(di.dn, dj.dn);
      n_excite = du.n_diffs + dd.n_diffs;
      if (n_excite > 2) return 0.0;
    } else if (n_excite > 0) {
      du = diff(di.up, dj.up);
      if (du.n_diffs < (unsigned)n_excite) dd = diff(di.dn, dj.dn);
    }
    if (n_excite == 0) return one_body_diag(di) + two_body_diag(di) + ints->energy_core;
    if (n_excite == 1) return one_body_single(du, dd) + two_body_single(di, du, dd);
    if (n_excite == 2) return two_body_double(du, dd);
    throw std::runtime_error("Calling hamiltonian with >2 excitation");
  }
};

// ---------------------------------------------------------------------------
// SparseMatrix: src/solver/sparse_matrix.{h,cc}, src/solver/sparse_vector.h
// ---------------------------------------------------------------------------
struct SparseMatrix {
  std::vector<std::vector<size_t>> idx;
  std::vector<std::vector<double>> val;
  std::vector<double> diag;

  void set_dim(size_t dim) {  // sparse_matrix.cc:87-91
    idx.resize(dim);
    val.resize(dim);
    diag.resize(dim, 0.0);
  }
  void append_elem(size_t i, size_t j, double v) {  // sparse_matrix.cc:8-11
    idx[i].push_back(j);
    val[i].push_back(v);
    if (i == j) diag[i] = v;
  }
  size_t n_stored() const {
    size_t s = 0;
    for (const auto& r : idx) s += r.size();
    return s;
  }
  size_t count_n_elems() const { return n_stored() * 2 - idx.size(); }  // sparse_matrix.cc:13-26
  // sparse_matrix.cc:28-56 (one rank: the allreduce is the identity)
  std::vector<double> mul(const std::vector<double>& vec) const {
    const size_t dim = idx.size();
    std::vector<double> res(dim, 0.0);
#pragma omp parallel for schedule(static, 1)
    for (size_t i = 0; i < dim; i++) {
      double diff_i = 0.0;
      for (size_t k = 0; k < idx[i].size(); k++) {
        const size_t j = idx[i][k];
        const double h = val[i][k];
        diff_i += h * vec[j];
        if (i != j) {
          const double diff_j = h * vec[i];
#pragma omp atomic
          res[j] += diff_j;
        }
      }
#pragma omp atomic
      res[i] += diff_i;
    }
    return res;
  }
};

// ---------------------------------------------------------------------------
// Hamiltonian<ChemSystem>: src/solver/hamiltonian.h:87-598 (time_sym = false)
// ---------------------------------------------------------------------------
template <class T>
static void sort_by_first(std::vector<T>& v1, std::vector<T>& v2) {  // src/util.h:77-91
  std::vector<std::pair<T, T>> v;
  v.reserve(v1.size());
  for (size_t i = 0; i < v1.size(); i++) v.emplace_back(v1[i], v2[i]);
  std::sort(v.begin(), v.end(), [](const std::pair<T, T>& a, const std::pair<T, T>& b) {
    return a.first < b.first;
  });
  for (size_t i = 0; i < v1.size(); i++) v1[i] = v[i].first, v2[i] = v[i].second;
}

struct Hamiltonian {
  SparseMatrix matrix;
  size_t n_dets = 0, n_dets_prev = 0;
  unsigned n_up = 0, n_dn = 0;
  bool sort_by_det_id = false;
  bool has_double_excitation = true;
  std::vector<uint64_t> unique_alphas, unique_betas;
  std::unordered_map<uint64_t, size_t> alpha_to_id, beta_to_id;
  std::unordered_map<uint64_t, std::pair<std::vector<size_t>, std::vector<size_t>>> abm1_to_ab_ids;
  std::vector<std::vector<size_t>> alpha_id_to_single_ids, beta_id_to_single_ids;
  std::vector<std::vector<size_t>> alpha_id_to_beta_ids, alpha_id_to_det_ids;
  std::vector<std::vector<size_t>> beta_id_to_alpha_ids, beta_id_to_det_ids;
  // Phase statistics, printed by the reference (hamiltonian.h:406-412, 593-596).
  size_t singles_cnt = 0;
  double t_abdet = 0, t_abm1 = 0, t_absingles = 0, t_matrix = 0;

  // hamiltonian.h:87-112
  void update(const Chem& sys, const std::vector<Det>& dets) {
    n_dets_prev = n_dets;
    n_dets = dets.size();
    if (n_dets_prev == n_dets) return;
    sort_by_det_id = (n_dets < n_dets_prev * 1.15);
    double t0 = omp_get_wtime();
    update_abdet(dets);
    double t1 = omp_get_wtime();
    t_abdet = t1 - t0;
    if (has_double_excitation) {
      update_abm1(dets);
      double t2 = omp_get_wtime();
      t_abm1 = t2 - t1;
      update_absingles(dets);
      abm1_to_ab_ids.clear();
      t_absingles = omp_get_wtime() - t2;
      t1 = omp_get_wtime();
    }
    update_matrix(sys, dets);
    t_matrix = omp_get_wtime() - t1;
    alpha_id_to_single_ids.clear();
    alpha_id_to_single_ids.shrink_to_fit();
    beta_id_to_single_ids.clear();
    beta_id_to_single_ids.shrink_to_fit();
  }

  // hamiltonian.h:157-254
  void update_abdet(const std::vector<Det>& dets) {
    std::unordered_set<size_t> updated_alphas, updated_betas;
    for (size_t i = n_dets_prev; i < n_dets; i++) {
      const uint64_t alpha = dets[i].up, beta = dets[i].dn;
      size_t alpha_id, beta_id;
      auto ita = alpha_to_id.find(alpha);
      if (ita == alpha_to_id.end()) {
        alpha_id = alpha_to_id.size();
        alpha_to_id[alpha] = alpha_id;
        unique_alphas.push_back(alpha);
        alpha_id_to_beta_ids.resize(alpha_id + 1);
        alpha_id_to_det_ids.resize(alpha_id + 1);
      } else {
        alpha_id = ita->second;
      }
      updated_alphas.insert(alpha_id);
      auto itb = beta_to_id.find(beta);
      if (itb == beta_to_id.end()) {
        beta_id = beta_to_id.size();
        beta_to_id[beta] = beta_id;
        unique_betas.push_back(beta);
        beta_id_to_alpha_ids.resize(beta_id + 1);
        beta_id_to_det_ids.resize(beta_id + 1);
      } else {
        beta_id = itb->second;
      }
      updated_betas.insert(beta_id);
      alpha_id_to_beta_ids[alpha_id].push_back(beta_id);
      alpha_id_to_det_ids[alpha_id].push_back(i);
      beta_id_to_alpha_ids[beta_id].push_back(alpha_id);
      beta_id_to_det_ids[beta_id].push_back(i);
    }
    if (sort_by_det_id) {
      for (size_t a : updated_alphas) sort_by_first(alpha_id_to_det_ids[a], alpha_id_to_beta_ids[a]);
      for (size_t b : updated_betas) sort_by_first(beta_id_to_det_ids[b], beta_id_to_alpha_ids[b]);
    } else {
      for (size_t a : updated_alphas) sort_by_first(alpha_id_to_beta_ids[a], alpha_id_to_det_ids[a]);
      for (size_t b : updated_betas) sort_by_first(beta_id_to_alpha_ids[b], beta_id_to_det_ids[b]);
    }
  }

  // hamiltonian.h:256-305
  void update_abm1(const std::vector<Det>& dets) {
    std::unordered_set<size_t> updated_alphas, updated_betas;
    for (size_t i = n_dets_prev; i < n_dets; i++) {
      const size_t alpha_id = alpha_to_id[dets[i].up];
      if (!updated_alphas.count(alpha_id)) {
        for (unsigned e : occ(dets[i].up)) {
          abm1_to_ab_ids[dets[i].up & ~(1ull << e)].first.push_back(alpha_id);
        }
        updated_alphas.insert(alpha_id);
      }
      const size_t beta_id = beta_to_id[dets[i].dn];
      if (!updated_betas.count(beta_id)) {
        for (unsigned e : occ(dets[i].dn)) {
          abm1_to_ab_ids[dets[i].dn & ~(1ull << e)].second.push_back(beta_id);
        }
        updated_betas.insert(beta_id);
      }
    }
  }

  // hamiltonian.h:307-413
  void update_absingles(const std::vector<Det>& dets) {
    std::unordered_set<size_t> updated_alphas, updated_betas;
    alpha_id_to_single_ids.resize(alpha_to_id.size());
    beta_id_to_single_ids.resize(beta_to_id.size());
    for (size_t i = n_dets_prev; i < n_dets; i++) {
      updated_alphas.insert(alpha_to_id[dets[i].up]);
      updated_betas.insert(beta_to_id[dets[i].dn]);
    }
    const size_t n_ua = alpha_to_id.size(), n_ub = beta_to_id.size();
    std::vector<omp_lock_t> locks(std::max(n_ua, n_ub));
    for (auto& l : locks) omp_init_lock(&l);
    const auto pass = [&](const std::vector<uint64_t>& uniq, bool is_alpha,
                          std::vector<std::vector<size_t>>& singles,
                          const std::unordered_set<size_t>& updated) {
      const size_t n_u = uniq.size();
#pragma omp parallel for schedule(static, 1)
      for (size_t id = 0; id < n_u; id++) {
        for (unsigned e : occ(uniq[id])) {
          const uint64_t m1 = uniq[id] & ~(1ull << e);
          auto it = abm1_to_ab_ids.find(m1);
          if (it == abm1_to_ab_ids.end()) continue;
          const auto& lst = is_alpha ? it->second.first : it->second.second;
          for (size_t other : lst) {
            if (other == id) continue;
            if (id > other && updated.count(id) && updated.count(other)) continue;
            omp_set_lock(&locks[id]);
            singles[id].push_back(other);
            omp_unset_lock(&locks[id]);
            omp_set_lock(&locks[other]);
            singles[other].push_back(id);
            omp_unset_lock(&locks[other]);
          }
        }
      }
    };
    pass(unique_alphas, true, alpha_id_to_single_ids, updated_alphas);
    pass(unique_betas, false, beta_id_to_single_ids, updated_betas);
    for (auto& l : locks) omp_destroy_lock(&l);
    size_t cnt = 0;
#pragma omp parallel for schedule(static, 1) reduction(+ : cnt)
    for (size_t a = 0; a < n_ua; a++) {
      std::sort(alpha_id_to_single_ids[a].begin(), alpha_id_to_single_ids[a].end());
      cnt += alpha_id_to_single_ids[a].size();
    }
#pragma omp parallel for schedule(static, 1) reduction(+ : cnt)
    for (size_t b = 0; b < n_ub; b++) {
      std::sort(beta_id_to_single_ids[b].begin(), beta_id_to_single_ids[b].end());
      cnt += beta_id_to_single_ids[b].size();
    }
    singles_cnt = cnt;
  }

  // hamiltonian.h:415-598
  void update_matrix(const Chem& sys, const std::vector<Det>& dets) {
    matrix.set_dim(n_dets);
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t det_id = 0; det_id < n_dets; det_id++) {
      const Det& det = dets[det_id];
      const bool is_new = det_id >= n_dets_prev;
      if (is_new) matrix.append_elem(det_id, det_id, sys.elem(det, det, 0));
      const size_t start_id = is_new ? det_id + 1 : n_dets_prev;
      // Same beta: single or double alpha excitations (431-442).
      const size_t beta_id = beta_to_id.at(det.dn);
      for (size_t j : beta_id_to_det_ids[beta_id]) {
        if (j < start_id) continue;
        const double H = sys.elem(det, dets[j], -1);
        if (std::abs(H) < EPS) continue;
        matrix.append_elem(det_id, j, H);
      }
      // Same alpha: single or double beta excitations (456-468).
      const size_t alpha_id = alpha_to_id.at(det.up);
      for (size_t j : alpha_id_to_det_ids[alpha_id]) {
        if (j < start_id) continue;
        const double H = sys.elem(det, dets[j], -1);
        if (std::abs(H) < EPS) continue;
        matrix.append_elem(det_id, j, H);
      }
      // Mixed alpha-single x beta-single doubles (482-532).
      if (!has_double_excitation) continue;
      const auto& alpha_singles = alpha_id_to_single_ids[alpha_id];
      const auto& beta_singles = beta_id_to_single_ids[beta_id];
      for (size_t alpha_single : alpha_singles) {
        if (alpha_id_to_beta_ids.size() <= alpha_single) continue;
        const auto& rel_beta = alpha_id_to_beta_ids[alpha_single];
        const auto& rel_det = alpha_id_to_det_ids[alpha_single];
        const size_t n_rel = rel_beta.size();
        if (sort_by_det_id) {
          const size_t s0 = std::lower_bound(rel_det.begin(), rel_det.end(), start_id) - rel_det.begin();
          for (size_t r = s0; r < n_rel; r++) {
            if (std::binary_search(beta_singles.begin(), beta_singles.end(), rel_beta[r])) {
              const size_t j = rel_det[r];
              const double H = sys.elem(det, dets[j], 2);
              if (std::abs(H) < EPS) continue;
              matrix.append_elem(det_id, j, H);
            }
          }
        } else {
          size_t ptr = 0;
          for (size_t beta_single : beta_singles) {
            while (ptr < n_rel && rel_beta[ptr] < beta_single) ptr++;
            if (ptr == n_rel) break;
            if (rel_beta[ptr] == beta_single) {
              const size_t j = rel_det[ptr];
              ptr++;
              if (j < start_id) continue;
              const double H = sys.elem(det, dets[j], 2);
              if (std::abs(H) < EPS) continue;
              matrix.append_elem(det_id, j, H);
            }
          }
        }
      }
    }
  }

  // hamiltonian.h:114-126 + sparse_matrix.cc:131-141
  void update_existing_elems(const Chem& sys, const std::vector<Det>& dets) {
    const size_t dim = matrix.idx.size();
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t i = 0; i < dim; i++) {
      for (size_t k = 0; k < matrix.idx[i].size(); k++) {
        const size_t j = matrix.idx[i][k];
        const double v = sys.elem(dets[i], dets[j], -1);
        matrix.val[i][k] = v;
        if (i == j) matrix.diag[i] = v;
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Dense symmetric eigensolver for the Rayleigh-Ritz step (stands in for
// Eigen::SelfAdjointEigenSolver at src/solver/davidson.cc:145-148).  Cyclic
// Jacobi; eigenvalues ascending, eigenvectors as columns.
// ---------------------------------------------------------------------------
static void jacobi_eigh(int n, const std::vector<double>& a_in, std::vector<double>& w,
                        std::vector<double>& v) {
  std::vector<double> a = a_in;  // row-major n x n, symmetric
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) off += a[(size_t)p * n + q] * a[(size_t)p * n + q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; p++) {
      for (int q = p + 1; q < n; q++) {
        const double apq = a[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::abs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * akp - s * akq;
          a[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * apk - s * aqk;
          a[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = c * vkp - s * vkq;
          v[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
    }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(),
            [&](int x, int y) { return a[(size_t)x * n + x] < a[(size_t)y * n + y]; });
  w.resize(n);
  std::vector<double> vs((size_t)n * n);
  for (int j = 0; j < n; j++) {
    w[j] = a[(size_t)ord[j] * n + ord[j]];
    for (int i = 0; i < n; i++) vs[(size_t)i * n + j] = v[(size_t)i * n + ord[j]];
  }
  v.swap(vs);
}

static double dot_omp(const std::vector<double>& a, const std::vector<double>& b) {
  // src/util.cc:53-79 (per-thread partial sums)
  const int nt = omp_get_max_threads();
  std::vector<double> part(nt, 0.0);
  const size_t n = a.size();
#pragma omp parallel for
  for (size_t i = 0; i < n; i++) part[omp_get_thread_num()] += a[i] * b[i];
  double s = 0.0;
  for (int t = 0; t < nt; t++) s += part[t];
  return s;
}

// ---------------------------------------------------------------------------
// Davidson::diagonalize: src/solver/davidson.cc:8-196
// ---------------------------------------------------------------------------
struct DavidsonResult {
  std::vector<double> evals;
  std::vector<std::vector<double>> evecs;
  bool converged = false;
  std::vector<double> trajectory;  // ground-state estimate after each SpMV batch ("Davidson #k")
  size_t n_spmv = 0;
};

static DavidsonResult davidson(const SparseMatrix& m, const std::vector<std::vector<double>>& init,
                               double tol, unsigned n_states_req) {
  DavidsonResult R;
  const size_t N_STORE = 5;
  const size_t dim = init[0].size();
  const unsigned n_states = (unsigned)std::min<size_t>(dim, std::min<size_t>(init.size(), n_states_req));
  R.evals.assign(n_states_req, 0.0);
  R.evecs.assign(n_states_req, std::vector<double>(dim));
  if (dim == 1) {
    R.evals[0] = m.diag[0];
    R.evecs[0].assign(1, 1.0);
    R.converged = true;
    return R;
  }
  const size_t n_store = std::min(dim, N_STORE);
  std::vector<double> lowest(n_states), lowest_prev(n_states, 0.0);
  std::vector<std::vector<double>> v(n_states * n_store, std::vector<double>(dim)), Hv(n_states * n_store);
  std::vector<std::vector<double>> w(n_states), Hw(n_states);
  for (unsigned s = 0; s < n_states; s++) {
    double norm = std::sqrt(dot_omp(init[s], init[s]));
    for (size_t j = 0; j < dim; j++) v[s][j] = init[s][j] / norm;
    if (s > 0) {
      for (unsigned k = 0; k < s; k++) {
        const double ip = dot_omp(v[s], v[k]);
        for (size_t j = 0; j < dim; j++) v[s][j] -= ip * v[k][j];
      }
      norm = std::sqrt(dot_omp(v[s], v[s]));
      for (size_t j = 0; j < dim; j++) v[s][j] /= norm;
    }
  }
  const size_t K = n_states * n_store;
  std::vector<double> hk(K * K, 0.0), ek(K * K, 0.0);
  auto H = [&](size_t i, size_t j) -> double& { return hk[i * K + j]; };
  bool converged = false;
  size_t n_converged = 0;
  for (unsigned s = 0; s < n_states; s++) {
    Hv[s] = m.mul(v[s]);
    R.n_spmv++;
    lowest[s] = dot_omp(v[s], Hv[s]);
    H(s, s) = lowest[s];
    w[s] = v[s];
    Hw[s] = Hv[s];
  }
  R.trajectory.push_back(lowest[0]);
  lowest_prev = lowest;
  size_t i_state_precond = 0;
  for (size_t it = n_states; it < n_store * n_states * 2; it++) {
    const size_t it_circ = it % K;
    if (it >= K) {
      if (it_circ < n_states - 1) continue;
      if (it_circ == n_states - 1) {
        for (unsigned s = 0; s < n_states; s++) v[s] = w[s], Hv[s] = Hw[s];
        for (unsigned s = 0; s < n_states; s++) {
          lowest[s] = dot_omp(v[s], Hv[s]);
          H(s, s) = lowest[s];
          for (unsigned k = s + 1; k < n_states; k++) {
            const double e = dot_omp(v[s], Hv[k]);
            H(s, k) = e;
            H(k, s) = e;
          }
        }
        continue;
      }
    }
    const double lam = lowest[i_state_precond];
#pragma omp parallel for
    for (size_t j = 0; j < dim; j++) {
      const double d = lam - m.diag[j];
      v[it_circ][j] = std::abs(d) < 1.0e-8 ? 0.0 : (Hw[i_state_precond][j] - lam * w[i_state_precond][j]) / d;
    }
    for (size_t i = 0; i < it_circ; i++) {
      const double ip = dot_omp(v[it_circ], v[i]);
#pragma omp parallel for
      for (size_t j = 0; j < dim; j++) v[it_circ][j] -= ip * v[i][j];
    }
    const double norm = std::sqrt(dot_omp(v[it_circ], v[it_circ]));
    if (norm < 1e-12) {
      converged = true;
      break;
    }
#pragma omp parallel for
    for (size_t j = 0; j < dim; j++) v[it_circ][j] /= norm;
    Hv[it_circ] = m.mul(v[it_circ]);
    R.n_spmv++;
    for (size_t i = 0; i <= it_circ; i++) H(it_circ, i) = dot_omp(v[i], Hv[it_circ]);
    if (i_state_precond + 1 == n_states) {
      const int ks = (int)it_circ + 1;
      std::vector<double> sub((size_t)ks * ks);
      for (int i = 0; i < ks; i++)
        for (int j = 0; j < ks; j++) sub[(size_t)i * ks + j] = i >= j ? H(i, j) : H(j, i);  // lower triangle
      std::vector<double> ev, evec;
      jacobi_eigh(ks, sub, ev, evec);
      for (unsigned s = 0; s < n_states; s++) {
        lowest[s] = ev[s];
        const double factor = evec[s] < 0 ? -1.0 : 1.0;  // evec(0, s)
        for (int i = 0; i < ks; i++) ek[(size_t)i * K + s] = evec[(size_t)i * ks + s] * factor;
        std::vector<double> ws(dim), hws(dim);
#pragma omp parallel for
        for (size_t j = 0; j < dim; j++) {
          double wj = 0.0, hwj = 0.0;
          for (int i = 0; i < ks; i++) {
            wj += v[i][j] * ek[(size_t)i * K + s];
            hwj += Hv[i][j] * ek[(size_t)i * K + s];
          }
          ws[j] = wj;
          hws[j] = hwj;
        }
        w[s].swap(ws);
        Hw[s].swap(hws);
      }
      R.trajectory.push_back(lowest[0]);
      for (unsigned s = (unsigned)n_converged; s < n_states; s++) {
        if (std::abs(lowest[s] - lowest_prev[s]) > tol) break;
        n_converged++;
      }
      if (n_converged == n_states) converged = true;
      if (!converged) lowest_prev = lowest;
      if (converged) break;
    }
    i_state_precond++;
    if (i_state_precond >= n_states) i_state_precond = n_converged;
  }
  for (unsigned s = 0; s < n_states; s++) {
    R.evals[s] = lowest[s];
    R.evecs[s] = w[s];
  }
  for (unsigned s = n_states; s < init.size() && s < n_states_req; s++) R.evecs[s] = init[s];
  R.converged = converged;
  return R;
}

// ---------------------------------------------------------------------------
// RDMs: src/chem/rdm.cc (time_sym = false)
// ---------------------------------------------------------------------------
struct RDM {
  unsigned n_orbs, n_up, n_dn, n_states;
  const std::vector<Det>& dets;
  const std::vector<std::vector<double>>& coefs;
  std::vector<double> one_rdm;  // n x n row-major
  std::vector<double> two_rdm;  // packed, rdm.cc:233-241

  RDM(unsigned n, unsigned nu, unsigned nd, const std::vector<Det>& d,
      const std::vector<std::vector<double>>& c)
      : n_orbs(n), n_up(nu), n_dn(nd), n_states((unsigned)c.size()), dets(d), coefs(c) {}

  // rdm.cc:223-231
  void write_in_1rdm(unsigned p, unsigned q, double factor, size_t i, size_t j) {
    double value = 0.0;
    for (unsigned s = 0; s < n_states; s++) value += coefs[s][i] * coefs[s][j];
    value *= factor;
#pragma omp atomic
    one_rdm[(size_t)p * n_orbs + q] += value;
  }
  // rdm.cc:12-131 (non-time_sym branch; orb_sym filter omitted: C1 only,
  // callers with symmetry pass all orbitals in one irrep)
  void get_1rdm(const std::vector<unsigned>& orb_sym) {
    one_rdm.assign((size_t)n_orbs * n_orbs, 0.0);
    std::unordered_map<Det, size_t, DetHash> det2ind;
    for (size_t i = 0; i < dets.size(); i++) det2ind[dets[i]] = i;
#pragma omp parallel for schedule(dynamic, 10)
    for (size_t i = 0; i < dets.size(); i++) {
      const Det& d = dets[i];
      const auto ou = occ(d.up), od = occ(d.dn);
      for (unsigned e = 0; e < n_up; e++) {
        const unsigned p = ou[e];
        for (unsigned r = 0; r < n_orbs; r++) {
          if (orb_sym[p] != orb_sym[r]) continue;
          if (p != r && (d.up >> r & 1)) continue;
          Det nd = d;
          nd.up = (nd.up & ~(1ull << p)) | (1ull << r);
          auto it = det2ind.find(nd);
          if (it == det2ind.end()) continue;
          write_in_1rdm(p, r, diff(d.up, nd.up).permutation_factor, i, it->second);
        }
      }
      for (unsigned e = 0; e < n_dn; e++) {
        const unsigned p = od[e];
        for (unsigned r = 0; r < n_orbs; r++) {
          if (orb_sym[p] != orb_sym[r]) continue;
          if (p != r && (d.dn >> r & 1)) continue;
          Det nd = d;
          nd.dn = (nd.dn & ~(1ull << p)) | (1ull << r);
          auto it = det2ind.find(nd);
          if (it == det2ind.end()) continue;
          write_in_1rdm(p, r, diff(d.dn, nd.dn).permutation_factor, i, it->second);
        }
      }
    }
  }
  // rdm.cc:644-657
  void write_in_2rdm(unsigned p, unsigned q, unsigned r, unsigned s, double factor, size_t i,
                     size_t j) {
    const size_t a = (size_t)p * n_orbs + s, b = (size_t)q * n_orbs + r;
    if (a >= b) {
      double value = 0.0;
      for (unsigned st = 0; st < n_states; st++) value += coefs[st][i] * coefs[st][j];
      value *= factor;
#pragma omp atomic
      two_rdm[(a * (a + 1)) / 2 + b] += value;
    }
  }
  // rdm.cc:243-269
  static int permfac_ccaa(uint64_t h, unsigned p, unsigned q, unsigned r, unsigned s) {
    unsigned c = 0;
    const unsigned ann[2] = {s, r}, cre[2] = {q, p};
    for (unsigned o : ann) {
      c += __builtin_popcountll(h & ((1ull << o) - 1));
      h &= ~(1ull << o);
    }
    for (unsigned o : cre) {
      c += __builtin_popcountll(h & ((1ull << o) - 1));
      h |= 1ull << o;
    }
    return c % 2 == 0 ? 1 : -1;
  }
  // rdm.cc:435-642, with connected_det = dets[j], this_det = dets[i], tr_factor = 1.
  void elements(const Det& cd, size_t j, const Det& td, size_t i) {
    const double tf = 1.0;
    const auto ou = occ(td.up), od = occ(td.dn);
    if (td.up == cd.up) {
      if (td.dn == cd.dn) {
        for (unsigned a = 0; a < n_up; a++)
          for (unsigned b = a + 1; b < n_up; b++) {
            const unsigned s = ou[a], r = ou[b];
            write_in_2rdm(s, r, r, s, tf, i, j);
            write_in_2rdm(s, r, s, r, -tf, i, j);
            write_in_2rdm(r, s, s, r, tf, i, j);
            write_in_2rdm(r, s, r, s, -tf, i, j);
          }
        for (unsigned a = 0; a < n_dn; a++)
          for (unsigned b = a + 1; b < n_dn; b++) {
            const unsigned s = od[a], r = od[b];
            write_in_2rdm(s, r, r, s, tf, i, j);
            write_in_2rdm(s, r, s, r, -tf, i, j);
            write_in_2rdm(r, s, s, r, tf, i, j);
            write_in_2rdm(r, s, r, s, -tf, i, j);
          }
        for (unsigned a = 0; a < n_up; a++)
          for (unsigned b = 0; b < n_dn; b++) {
            const unsigned s = ou[a], r = od[b];
            write_in_2rdm(s, r, r, s, tf, i, j);
            write_in_2rdm(r, s, s, r, tf, i, j);
          }
      } else {
        const DiffResult dcb = diff(cd.dn, td.dn);
        if (dcb.n_diffs == 1) {
          const unsigned b1 = dcb.right_only[0], b2 = dcb.left_only[0];
          for (unsigned e = 0; e < n_dn; e++) {
            const unsigned p = od[e];
            if (p != b1 && p != b2) {
              double sf = tf * permfac_ccaa(td.dn, p, b2, b1, p);
              write_in_2rdm(p, b2, b1, p, sf, i, j);
              write_in_2rdm(b2, p, b1, p, -sf, i, j);
              write_in_2rdm(p, b2, p, b1, -sf, i, j);
              write_in_2rdm(b2, p, p, b1, sf, i, j);
              sf = tf * permfac_ccaa(cd.dn, p, b1, b2, p);
              write_in_2rdm(p, b1, b2, p, sf, i, j);
              write_in_2rdm(b1, p, b2, p, -sf, i, j);
              write_in_2rdm(p, b1, p, b2, -sf, i, j);
              write_in_2rdm(b1, p, p, b2, sf, i, j);
            }
          }
          for (unsigned e = 0; e < n_up; e++) {
            const unsigned p = ou[e];
            double sf = tf * dcb.permutation_factor;
            write_in_2rdm(p, b2, b1, p, sf, i, j);
            write_in_2rdm(b2, p, p, b1, sf, i, j);
            sf = tf * diff(td.dn, cd.dn).permutation_factor;
            write_in_2rdm(p, b1, b2, p, sf, i, j);
            write_in_2rdm(b1, p, p, b2, sf, i, j);
          }
        } else if (dcb.n_diffs == 2) {
          const unsigned b1 = dcb.right_only[0], b2 = dcb.right_only[1];
          const unsigned b3 = dcb.left_only[0], b4 = dcb.left_only[1];
          double sf = tf * dcb.permutation_factor;
          write_in_2rdm(b3, b4, b2, b1, sf, i, j);
          write_in_2rdm(b3, b4, b1, b2, -sf, i, j);
          write_in_2rdm(b4, b3, b2, b1, -sf, i, j);
          write_in_2rdm(b4, b3, b1, b2, sf, i, j);
          sf = tf * diff(td.dn, cd.dn).permutation_factor;
          write_in_2rdm(b1, b2, b4, b3, sf, i, j);
          write_in_2rdm(b1, b2, b3, b4, -sf, i, j);
          write_in_2rdm(b2, b1, b4, b3, -sf, i, j);
          write_in_2rdm(b2, b1, b3, b4, sf, i, j);
        }
      }
      return;
    }
    const DiffResult dca = diff(cd.up, td.up);
    if (dca.n_diffs == 1) {
      if (td.dn == cd.dn) {
        const unsigned a1 = dca.right_only[0], a2 = dca.left_only[0];
        for (unsigned e = 0; e < n_up; e++) {
          const unsigned p = ou[e];
          if (p != a1 && p != a2) {
            double sf = tf * permfac_ccaa(td.up, p, a2, a1, p);
            write_in_2rdm(p, a2, a1, p, sf, i, j);
            write_in_2rdm(p, a2, p, a1, -sf, i, j);
            write_in_2rdm(a2, p, a1, p, -sf, i, j);
            write_in_2rdm(a2, p, p, a1, sf, i, j);
            sf = tf * permfac_ccaa(cd.up, p, a1, a2, p);
            write_in_2rdm(p, a1, a2, p, sf, i, j);
            write_in_2rdm(p, a1, p, a2, -sf, i, j);
            write_in_2rdm(a1, p, a2, p, -sf, i, j);
            write_in_2rdm(a1, p, p, a2, sf, i, j);
          }
        }
        for (unsigned e = 0; e < n_dn; e++) {
          const unsigned p = od[e];
          double sf = tf * dca.permutation_factor;
          write_in_2rdm(a2, p, p, a1, sf, i, j);
          write_in_2rdm(p, a2, a1, p, sf, i, j);
          sf = tf * diff(td.up, cd.up).permutation_factor;
          write_in_2rdm(a1, p, p, a2, sf, i, j);
          write_in_2rdm(p, a1, a2, p, sf, i, j);
        }
      } else {
        const DiffResult dcb = diff(cd.dn, td.dn);
        if (dcb.n_diffs == 1) {
          const unsigned a1 = dca.right_only[0], a2 = dca.left_only[0];
          const unsigned b1 = dcb.right_only[0], b2 = dcb.left_only[0];
          double sf = tf * dca.permutation_factor * dcb.permutation_factor;
          write_in_2rdm(a2, b2, b1, a1, sf, i, j);
          write_in_2rdm(b2, a2, a1, b1, sf, i, j);
          sf = tf * diff(td.up, cd.up).permutation_factor * diff(td.dn, cd.dn).permutation_factor;
          write_in_2rdm(a1, b1, b2, a2, sf, i, j);
          write_in_2rdm(b1, a1, a2, b2, sf, i, j);
        }
      }
    } else if (dca.n_diffs == 2) {
      if (td.dn == cd.dn) {
        const unsigned a1 = dca.right_only[0], a2 = dca.right_only[1];
        const unsigned a3 = dca.left_only[0], a4 = dca.left_only[1];
        double sf = tf * dca.permutation_factor;
        write_in_2rdm(a3, a4, a2, a1, sf, i, j);
        write_in_2rdm(a3, a4, a1, a2, -sf, i, j);
        write_in_2rdm(a4, a3, a2, a1, -sf, i, j);
        write_in_2rdm(a4, a3, a1, a2, sf, i, j);
        sf = tf * diff(td.up, cd.up).permutation_factor;
        write_in_2rdm(a1, a2, a4, a3, sf, i, j);
        write_in_2rdm(a1, a2, a3, a4, -sf, i, j);
        write_in_2rdm(a2, a1, a4, a3, -sf, i, j);
        write_in_2rdm(a2, a1, a3, a4, sf, i, j);
      }
    }
  }
  // rdm.cc:374-395 over the stored upper-triangle rows.
  void get_2rdm(const SparseMatrix& m) {
    const size_t n = n_orbs;
    two_rdm.assign(n * n * (n * n + 1) / 2, 0.0);
#pragma omp parallel for schedule(dynamic, 10)
    for (size_t i = 0; i < m.idx.size(); i++) {
      for (size_t k = 0; k < m.idx[i].size(); k++) {
        const size_t j = m.idx[i][k];
        elements(dets[j], j, dets[i], i);
      }
    }
  }
  double two_rdm_elem(unsigned p, unsigned q, unsigned r, unsigned s) const {  // rdm.cc:683-685
    const size_t a = (size_t)p * n_orbs + s, b = (size_t)q * n_orbs + r;
    return a > b ? two_rdm[(a * (a + 1)) / 2 + b] : two_rdm[(b * (b + 1)) / 2 + a];
  }
};

}  // namespace orc

// ===========================================================================
// C API (ctypes).  All handles are opaque.  Return 0 on success, -1 on error
// (message via orc_last_error).
// ===========================================================================
using namespace orc;

static thread_local std::string g_err;


// ---------------------------------------------------------------------------
// One heat-bath selection sweep: src/solver/solver.h:435-495 with
// ChemSystem::find_connected_dets (src/chem/chem_system.cc:289-402; second
// rejection and active space off, all orbitals in one irrep), restated
// without the sorted queues: every single and double excitation of a walked
// det is tested against the queue criteria directly.
//   single p -> r: S(p, r) > 0 and S >= eps_min (singles queue,
//     chem_system.cc:237-259), |H_ij| >= eps_min (solver.h:456-459);
//   double: |H| > 0 and eps_min <= |H| < eps_tried_prev, |H| the hci-queue
//     value (chem_system.cc:115-196,261-287): same spin |(pr|qs) - (ps|qr)|,
//     opposite spin |(pr|qs)| (p, r up; q, s down);
//   never a det of the variational space.
// time_sym (solver.h:447-455): eps_min times sqrt2 for a det with up != dn,
// candidates tested and kept in their canonical form (up <= dn), the single's
// element taken with the candidate as found.
// ---------------------------------------------------------------------------
static double selection_S(const Integrals& I, unsigned p, unsigned r, unsigned n_elecs) {
  const unsigned n = I.n_orbs;
  std::vector<double> el;
  for (unsigned o = 0; o < n; o++) {
    const double di = I.get_2b(p, r, o, o);
    el.push_back(di);
    if (o != p && o != r) el.push_back(di - I.get_2b(p, o, o, r));
  }
  std::sort(el.begin(), el.end());
  double smin = I.get_1b(p, r), smax = smin;
  for (unsigned i = 0; i < std::min(n_elecs - 1, 2 * n - 2); i++) {
    smin += el[i];
    smax += el[2 * n - 3 - i];
  }
  return std::max(std::abs(smin), std::abs(smax));
}

static std::vector<Det> select_sweep(const Integrals& I, const Chem& chem, const std::vector<Det>& dets,
                                     const double* coefs, unsigned n_states, double eps_var, double* eps_tried,
                                     bool time_sym) {
  const unsigned n = I.n_orbs;
  const size_t nd = dets.size();
  const unsigned ne = (unsigned)(__builtin_popcountll(dets[0].up) + __builtin_popcountll(dets[0].dn));
  std::vector<double> S((size_t)n * n, 0.0);
  for (unsigned p = 0; p < n; p++)
    for (unsigned r = 0; r < n; r++)
      if (p != r) S[p * n + r] = selection_S(I, p, r, ne);
  std::unordered_set<Det, DetHash> var(dets.begin(), dets.end());
  const uint64_t all = n < 64 ? (1ull << n) - 1 : ~0ull;
  std::vector<std::vector<Det>> found(omp_get_max_threads());
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t i = 0; i < nd; i++) {
    auto& out = found[omp_get_thread_num()];
    double mc = coefs[i];
    for (unsigned s = 1; s < n_states; s++)
      if (std::abs(coefs[s * nd + i]) > std::abs(mc)) mc = coefs[s * nd + i];
    const Det& d = dets[i];
    double emin = eps_var / std::abs(mc);
    if (time_sym && d.up != d.dn) emin *= 1.4142135623730951;  // Util::SQRT2
    const double emax = eps_tried[i];
    if (emin >= emax) continue;
    auto keep = [&](const Det& c, bool single) {
      const Det reg = (time_sym && c.up > c.dn) ? Det{c.dn, c.up} : c;
      if (var.count(reg)) return;
      if (single && std::abs(chem.elem(d, c, 1)) < emin) return;
      out.push_back(reg);
    };
    for (int spin = 0; spin < 2; spin++) {
      const uint64_t str = spin ? d.dn : d.up;
      const auto oc = occ(str), vi = occ(all & ~str);
      for (unsigned p : oc)
        for (unsigned r : vi) {
          const double s = S[p * n + r];
          if (s == 0.0 || s < emin) continue;
          const uint64_t t = (str & ~(1ull << p)) | (1ull << r);
          keep(spin ? Det{d.up, t} : Det{t, d.dn}, true);
        }
      for (size_t a = 0; a < oc.size(); a++)
        for (size_t b = a + 1; b < oc.size(); b++)
          for (size_t x = 0; x < vi.size(); x++)
            for (size_t y = x + 1; y < vi.size(); y++) {
              const unsigned p = oc[a], q = oc[b], r = vi[x], s = vi[y];
              const double H = std::abs(I.get_2b(p, r, q, s) - I.get_2b(p, s, q, r));
              if (H == 0.0 || H < emin || H >= emax) continue;
              const uint64_t t = (str & ~(1ull << p) & ~(1ull << q)) | (1ull << r) | (1ull << s);
              keep(spin ? Det{d.up, t} : Det{t, d.dn}, false);
            }
    }
    const auto ou = occ(d.up), vu = occ(all & ~d.up), od = occ(d.dn), vd = occ(all & ~d.dn);
    for (unsigned p : ou)
      for (unsigned r : vu)
        for (unsigned q : od)
          for (unsigned s : vd) {
            const double H = std::abs(I.get_2b(p, r, q, s));
            if (H == 0.0 || H < emin || H >= emax) continue;
            keep(Det{(d.up & ~(1ull << p)) | (1ull << r), (d.dn & ~(1ull << q)) | (1ull << s)}, false);
          }
    eps_tried[i] = 1.000000001 * emin;
  }
  std::vector<Det> res;
  for (auto& v : found) res.insert(res.end(), v.begin(), v.end());
  auto lt = [](const Det& a, const Det& b) { return a.up < b.up || (a.up == b.up && a.dn < b.dn); };
  std::sort(res.begin(), res.end(), lt);
  res.erase(std::unique(res.begin(), res.end()), res.end());
  return res;
}


// ---------------------------------------------------------------------------
// Deterministic PT2 correction, one batch: Solver::get_energy_pt_dtm
// (src/solver/solver.h:637-764).  For every variational det i, the dets a
// that find_connected_dets yields with eps_min = eps_pt_dtm / |c_i| and no
// upper cut (chem_system.cc:289-402: singles with S(p, r) >= eps_min, doubles
// with queue |H| >= eps_min), restated without the queues; a outside the
// variational space with |H_ai c_i| >= eps_pt_dtm adds H_ai c_i to
// hc_sums[a] (solver.h:710-720, summed in increasing i here).  Then
// E_pt2 = sum_a hc_sums[a]^2 / (E_var - H_aa) (solver.h:728-733), summed over
// a in (up, dn) order.
// ---------------------------------------------------------------------------
static double pt_dtm(const Integrals& I, const Chem& chem, const std::vector<Det>& dets, const double* coefs,
                     double eps_pt, double energy_var, size_t* n_pt_dets) {
  const unsigned n = I.n_orbs;
  const size_t nd = dets.size();
  const unsigned ne = (unsigned)(__builtin_popcountll(dets[0].up) + __builtin_popcountll(dets[0].dn));
  std::vector<double> S((size_t)n * n, 0.0);
  for (unsigned p = 0; p < n; p++)
    for (unsigned r = 0; r < n; r++)
      if (p != r) S[p * n + r] = selection_S(I, p, r, ne);
  std::unordered_set<Det, DetHash> var(dets.begin(), dets.end());
  const uint64_t all = n < 64 ? (1ull << n) - 1 : ~0ull;
  struct Term {
    Det a;
    size_t i;
    double hc;
  };
  std::vector<std::vector<Term>> found(omp_get_max_threads());
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t i = 0; i < nd; i++) {
    auto& out = found[omp_get_thread_num()];
    const double c = coefs[i];
    const double emin = eps_pt / std::abs(c);
    if (!(emin < 1e300)) continue;  // c == 0: nothing passes
    const Det& d = dets[i];
    auto add = [&](const Det& a, int n_excite) {
      if (var.count(a)) return;
      const double hc = chem.elem(d, a, n_excite) * c;
      if (std::abs(hc) < eps_pt) return;
      out.push_back(Term{a, i, hc});
    };
    for (int spin = 0; spin < 2; spin++) {
      const uint64_t str = spin ? d.dn : d.up;
      const auto oc = occ(str), vi = occ(all & ~str);
      for (unsigned p : oc)
        for (unsigned r : vi) {
          const double s = S[p * n + r];
          if (s == 0.0 || s < emin) continue;
          const uint64_t t = (str & ~(1ull << p)) | (1ull << r);
          add(spin ? Det{d.up, t} : Det{t, d.dn}, 1);
        }
      for (size_t a = 0; a < oc.size(); a++)
        for (size_t b = a + 1; b < oc.size(); b++)
          for (size_t x = 0; x < vi.size(); x++)
            for (size_t y = x + 1; y < vi.size(); y++) {
              const unsigned p = oc[a], q = oc[b], r = vi[x], s = vi[y];
              const double H = std::abs(I.get_2b(p, r, q, s) - I.get_2b(p, s, q, r));
              if (H == 0.0 || H < emin) continue;
              const uint64_t t = (str & ~(1ull << p) & ~(1ull << q)) | (1ull << r) | (1ull << s);
              add(spin ? Det{d.up, t} : Det{t, d.dn}, 2);
            }
    }
    const auto ou = occ(d.up), vu = occ(all & ~d.up), od = occ(d.dn), vd = occ(all & ~d.dn);
    for (unsigned p : ou)
      for (unsigned r : vu)
        for (unsigned q : od)
          for (unsigned s : vd) {
            const double H = std::abs(I.get_2b(p, r, q, s));
            if (H == 0.0 || H < emin) continue;
            add(Det{(d.up & ~(1ull << p)) | (1ull << r), (d.dn & ~(1ull << q)) | (1ull << s)}, 2);
          }
  }
  std::vector<Term> all_terms;
  for (auto& v : found) all_terms.insert(all_terms.end(), v.begin(), v.end());
  std::sort(all_terms.begin(), all_terms.end(), [](const Term& x, const Term& y) {
    if (x.a.up != y.a.up) return x.a.up < y.a.up;
    if (x.a.dn != y.a.dn) return x.a.dn < y.a.dn;
    return x.i < y.i;
  });
  // The per-det sums add their terms in increasing i (the order the device
  // keeps, so each contribution is bit-identical); the total over the PT
  // dets is a compensated (Neumaier) sum, so that a comparison of totals over
  // millions of same-sign contributions measures the other side's rounding,
  // not this one's (a plain running sum drifts by ~1e-12 relative there).
  double e = 0.0, comp = 0.0;
  size_t np = 0;
  for (size_t k = 0; k < all_terms.size();) {
    size_t k2 = k;
    double s = 0.0;
    while (k2 < all_terms.size() && all_terms[k2].a == all_terms[k].a) s += all_terms[k2++].hc;
    const double haa = chem.elem(all_terms[k].a, all_terms[k].a, 0);
    const double t = s * s / (energy_var - haa);
    const double u = e + t;
    comp += std::fabs(e) >= std::fabs(t) ? (e - u) + t : (t - u) + e;
    e = u;
    np++;
    k = k2;
  }
  *n_pt_dets = np;
  return e + comp;
}

struct OrcSystem {
  Integrals ints;
  Chem chem;
  unsigned n_up = 0, n_dn = 0;
  std::vector<Det> dets;
  Hamiltonian ham;
  std::vector<Det> selected;  // the last selection sweep's new dets
  OrcSystem() { chem.ints = &ints; }
};

#define ORC_TRY(body)        \
  try {                      \
    body;                    \
    return 0;                \
  } catch (const std::exception& e) { \
    g_err = e.what();        \
    return -1;               \
  }

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// Integrals from arrays: h1 dense n*n (symmetric), eri packed by combine4.
void* orc_system_new(unsigned n_orbs, unsigned n_up, unsigned n_dn, const double* h1,
                     const double* eri_packed, double core) {
  auto* s = new OrcSystem();
  s->ints.resize(n_orbs);
  s->n_up = n_up;
  s->n_dn = n_dn;
  s->ints.energy_core = core;
  for (unsigned p = 0; p < n_orbs; p++)
    for (unsigned q = 0; q <= p; q++) s->ints.h1[combine2(p, q)] = h1[(size_t)p * n_orbs + q];
  std::memcpy(s->ints.eri.data(), eri_packed, s->ints.eri.size() * sizeof(double));
  s->ham.n_up = n_up;
  s->ham.n_dn = n_dn;
  return s;
}

// Integrals from an FCIDUMP file (reference semantics incl. reorder).
// reorder: -1 = reference default (!explicit occupations), 0 = off, 1 = on.
void* orc_system_from_fcidump(const char* path, unsigned n_up, unsigned n_dn, const unsigned* occs_up,
                              unsigned n_occs_up, const unsigned* occs_dn, unsigned n_occs_dn,
                              int reorder) {
  try {
    auto* s = new OrcSystem();
    std::vector<unsigned> ou(occs_up, occs_up + n_occs_up), od(occs_dn, occs_dn + n_occs_dn);
    load_integrals(path, n_up, n_dn, ou, od, reorder, s->ints);
    if (s->ints.n_orbs > 64) throw std::runtime_error("oracle supports n_orbs <= 64");
    s->n_up = n_up;
    s->n_dn = n_dn;
    s->ham.n_up = n_up;
    s->ham.n_dn = n_dn;
    return s;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void orc_system_free(void* h) { delete static_cast<OrcSystem*>(h); }

int orc_system_info(void* h, unsigned* n_orbs, double* core, uint64_t* hf_up, uint64_t* hf_dn,
                    size_t* eri_size) {
  auto* s = static_cast<OrcSystem*>(h);
  *n_orbs = s->ints.n_orbs;
  *core = s->ints.energy_core;
  *hf_up = s->ints.hf_up;
  *hf_dn = s->ints.hf_dn;
  *eri_size = s->ints.eri.size();
  return 0;
}

int orc_system_integrals(void* h, double* h1_dense, double* eri_packed, unsigned* orb_order) {
  auto* s = static_cast<OrcSystem*>(h);
  const unsigned n = s->ints.n_orbs;
  for (unsigned p = 0; p < n; p++)
    for (unsigned q = 0; q < n; q++) h1_dense[(size_t)p * n + q] = s->ints.get_1b(p, q);
  std::memcpy(eri_packed, s->ints.eri.data(), s->ints.eri.size() * sizeof(double));
  if (orb_order)
    for (unsigned i = 0; i < n; i++) orb_order[i] = s->ints.orb_order.empty() ? i : s->ints.orb_order[i];
  return 0;
}

int orc_diff(uint64_t left, uint64_t right, unsigned* left_only, unsigned* right_only, int* pf) {
  const DiffResult d = diff(left, right);
  left_only[0] = d.left_only[0];
  left_only[1] = d.left_only[1];
  right_only[0] = d.right_only[0];
  right_only[1] = d.right_only[1];
  *pf = d.permutation_factor;
  return (int)d.n_diffs;
}

int orc_elem(void* h, uint64_t up_i, uint64_t dn_i, uint64_t up_j, uint64_t dn_j, int n_excite,
             double* out) {
  auto* s = static_cast<OrcSystem*>(h);
  ORC_TRY(*out = s->chem.elem(Det{up_i, dn_i}, Det{up_j, dn_j}, n_excite));
}

// Append-only determinant update + Hamiltonian::update (reference semantics:
// the first n_prev dets must equal the previous call's list).
int orc_elem_time_sym(void* h, uint64_t up_i, uint64_t dn_i, uint64_t up_j, uint64_t dn_j, double* out) {
  auto* s = static_cast<OrcSystem*>(h);
  *out = s->chem.elem_time_sym(Det{up_i, dn_i}, Det{up_j, dn_j}, -1);
  return 0;
}

int orc_ham_update(void* h, const uint64_t* up, const uint64_t* dn, size_t n_dets) {
  auto* s = static_cast<OrcSystem*>(h);
  ORC_TRY({
    if (n_dets < s->dets.size()) throw std::runtime_error("det list may only grow");
    for (size_t i = 0; i < s->dets.size(); i++)
      if (s->dets[i].up != up[i] || s->dets[i].dn != dn[i])
        throw std::runtime_error("det prefix changed");
    for (size_t i = s->dets.size(); i < n_dets; i++) s->dets.push_back(Det{up[i], dn[i]});
    s->ham.update(s->chem, s->dets);
  });
}

int orc_ham_clear(void* h) {
  auto* s = static_cast<OrcSystem*>(h);
  unsigned nu = s->n_up, nd = s->n_dn;
  s->ham = Hamiltonian();
  s->ham.n_up = nu;
  s->ham.n_dn = nd;
  s->dets.clear();
  return 0;
}

int orc_ham_update_existing(void* h) {
  auto* s = static_cast<OrcSystem*>(h);
  ORC_TRY(s->ham.update_existing_elems(s->chem, s->dets));
}

// Replace the integrals (same n_orbs), e.g. before update_existing_elems.
int orc_system_set_integrals(void* h, const double* h1, const double* eri_packed, double core) {
  auto* s = static_cast<OrcSystem*>(h);
  const unsigned n = s->ints.n_orbs;
  s->ints.energy_core = core;
  for (unsigned p = 0; p < n; p++)
    for (unsigned q = 0; q <= p; q++) s->ints.h1[combine2(p, q)] = h1[(size_t)p * n + q];
  std::memcpy(s->ints.eri.data(), eri_packed, s->ints.eri.size() * sizeof(double));
  return 0;
}

// stats[0..7] = n_stored, count_n_elems, singles_cnt, n_unique_alphas,
// n_unique_betas, n_dets; times[0..3] = abdet, abm1, absingles, matrix (s).
int orc_ham_stats(void* h, size_t* stats, double* times) {
  auto* s = static_cast<OrcSystem*>(h);
  stats[0] = s->ham.matrix.n_stored();
  stats[1] = s->ham.matrix.idx.empty() ? 0 : s->ham.matrix.count_n_elems();
  stats[2] = s->ham.singles_cnt;
  stats[3] = s->ham.unique_alphas.size();
  stats[4] = s->ham.unique_betas.size();
  stats[5] = s->dets.size();
  times[0] = s->ham.t_abdet;
  times[1] = s->ham.t_abm1;
  times[2] = s->ham.t_absingles;
  times[3] = s->ham.t_matrix;
  return 0;
}

// Export stored upper-triangle rows (row order, in append order within a row).
int orc_ham_export(void* h, int64_t* row_ptr, int64_t* cols, double* vals) {
  auto* s = static_cast<OrcSystem*>(h);
  const auto& m = s->ham.matrix;
  int64_t k = 0;
  for (size_t i = 0; i < m.idx.size(); i++) {
    row_ptr[i] = k;
    for (size_t t = 0; t < m.idx[i].size(); t++, k++) {
      cols[k] = (int64_t)m.idx[i][t];
      vals[k] = m.val[i][t];
    }
  }
  row_ptr[m.idx.size()] = k;
  return 0;
}

int orc_ham_diag(void* h, double* d) {
  auto* s = static_cast<OrcSystem*>(h);
  std::memcpy(d, s->ham.matrix.diag.data(), s->ham.matrix.diag.size() * sizeof(double));
  return 0;
}

int orc_ham_mul(void* h, const double* x, double* y) {
  auto* s = static_cast<OrcSystem*>(h);
  const size_t n = s->ham.matrix.idx.size();
  std::vector<double> xv(x, x + n);
  const auto r = s->ham.matrix.mul(xv);
  std::memcpy(y, r.data(), n * sizeof(double));
  return 0;
}

// Time `reps` SpMVs (the reference layout and atomic scatter); returns seconds per SpMV.
double orc_ham_time_mul(void* h, const double* x, int reps) {
  auto* s = static_cast<OrcSystem*>(h);
  const size_t n = s->ham.matrix.idx.size();
  std::vector<double> xv(x, x + n);
  const double t0 = omp_get_wtime();
  double sink = 0.0;
  for (int r = 0; r < reps; r++) sink += s->ham.matrix.mul(xv)[0];
  const double t = (omp_get_wtime() - t0) / reps;
  return sink == sink ? t : t;
}

// Davidson over the system's Hamiltonian matrix.  init: n_states x n_dets.
// traj receives up to traj_cap ground-state estimates (Davidson #0, #1, ...).
int orc_davidson(void* h, unsigned n_states, const double* init, double tol, double* evals,
                 double* evecs, int* converged, double* traj, int traj_cap, int* n_traj,
                 int* n_spmv) {
  auto* s = static_cast<OrcSystem*>(h);
  ORC_TRY({
    const size_t n = s->ham.matrix.idx.size();
    std::vector<std::vector<double>> iv(n_states, std::vector<double>(n));
    for (unsigned k = 0; k < n_states; k++)
      std::memcpy(iv[k].data(), init + (size_t)k * n, n * sizeof(double));
    const auto R = davidson(s->ham.matrix, iv, tol, n_states);
    for (unsigned k = 0; k < n_states; k++) {
      evals[k] = R.evals[k];
      std::memcpy(evecs + (size_t)k * n, R.evecs[k].data(), n * sizeof(double));
    }
    *converged = R.converged ? 1 : 0;
    int nt = 0;
    for (double e : R.trajectory)
      if (nt < traj_cap) traj[nt++] = e;
    *n_traj = nt;
    *n_spmv = (int)R.n_spmv;
  });
}

// A bare SparseMatrix (for the reference's DavidsonTest, src/solver/davidson_test.cc).
void* orc_sm_new(size_t dim) {
  auto* m = new SparseMatrix();
  m->set_dim(dim);
  return m;
}
void orc_sm_free(void* m) { delete static_cast<SparseMatrix*>(m); }
void orc_sm_append(void* m, size_t i, size_t j, double v) {
  static_cast<SparseMatrix*>(m)->append_elem(i, j, v);
}
int orc_sm_davidson(void* mh, unsigned n_states, const double* init, double tol, double* evals,
                    double* evecs, int* converged, double* traj, int traj_cap, int* n_traj) {
  auto* m = static_cast<SparseMatrix*>(mh);
  ORC_TRY({
    const size_t n = m->idx.size();
    std::vector<std::vector<double>> iv(n_states, std::vector<double>(n));
    for (unsigned k = 0; k < n_states; k++)
      std::memcpy(iv[k].data(), init + (size_t)k * n, n * sizeof(double));
    const auto R = davidson(*m, iv, tol, n_states);
    for (unsigned k = 0; k < n_states; k++) {
      evals[k] = R.evals[k];
      std::memcpy(evecs + (size_t)k * n, R.evecs[k].data(), n * sizeof(double));
    }
    *converged = R.converged ? 1 : 0;
    int nt = 0;
    for (double e : R.trajectory)
      if (nt < traj_cap) traj[nt++] = e;
    *n_traj = nt;
  });
}

// 1-RDM (n x n) and 2-RDM (packed, n^2(n^2+1)/2) of coefs (n_states x n_dets)
// over the system's det list and stored Hamiltonian connections.
int orc_select(void* h, const uint64_t* up, const uint64_t* dn, size_t n_dets, const double* coefs,
               unsigned n_states, double eps_var, double* eps_tried, size_t* n_new, int time_sym) {
  try {
    auto* s = static_cast<OrcSystem*>(h);
    std::vector<Det> dets(n_dets);
    for (size_t i = 0; i < n_dets; i++) dets[i] = Det{up[i], dn[i]};
    s->selected = select_sweep(s->ints, s->chem, dets, coefs, n_states, eps_var, eps_tried, time_sym != 0);
    *n_new = s->selected.size();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int orc_select_result(void* h, uint64_t* up, uint64_t* dn) {
  auto* s = static_cast<OrcSystem*>(h);
  for (size_t i = 0; i < s->selected.size(); i++) {
    up[i] = s->selected[i].up;
    dn[i] = s->selected[i].dn;
  }
  return 0;
}

int orc_pt_dtm(void* h, const uint64_t* up, const uint64_t* dn, size_t n_dets, const double* coefs, double eps_pt,
               double energy_var, double* energy_pt, size_t* n_pt_dets) {
  try {
    auto* s = static_cast<OrcSystem*>(h);
    std::vector<Det> dets(n_dets);
    for (size_t i = 0; i < n_dets; i++) dets[i] = Det{up[i], dn[i]};
    *energy_pt = pt_dtm(s->ints, s->chem, dets, coefs, eps_pt, energy_var, n_pt_dets);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int orc_rdm(void* h, unsigned n_states, const double* coefs, double* rdm1, double* rdm2) {
  auto* s = static_cast<OrcSystem*>(h);
  ORC_TRY({
    const size_t n = s->dets.size();
    std::vector<std::vector<double>> c(n_states, std::vector<double>(n));
    for (unsigned k = 0; k < n_states; k++) std::memcpy(c[k].data(), coefs + (size_t)k * n, n * sizeof(double));
    RDM r(s->ints.n_orbs, s->n_up, s->n_dn, s->dets, c);
    std::vector<unsigned> sym(s->ints.n_orbs, 1);
    if (rdm1) {
      r.get_1rdm(sym);
      std::memcpy(rdm1, r.one_rdm.data(), r.one_rdm.size() * sizeof(double));
    }
    if (rdm2) {
      r.get_2rdm(s->ham.matrix);
      std::memcpy(rdm2, r.two_rdm.data(), r.two_rdm.size() * sizeof(double));
    }
  });
}

}  // extern "C"
