// Host-side integrals + heat-bath selection.  See chem_host.h.
#include "chem_host.h"

#include <omp.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <numeric>
#include <stdexcept>

namespace shci_host {

namespace {

constexpr double kEps = 1.0e-12;  // src/util.h:50

inline uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

// Open-addressing set of determinants; concurrent lookups are safe once built.
class DetSet {
 public:
  void reserve(size_t n) {
    size_t cap = 1024;
    while (cap < n * 2) cap <<= 1;
    if (cap <= keys_.size()) return;
    std::vector<Det> old;
    old.swap(keys_);
    keys_.assign(cap, Det{~0ull, ~0ull});
    mask_ = cap - 1;
    size_ = 0;
    for (const Det& d : old)
      if (!(d.up == ~0ull && d.dn == ~0ull)) insert(d);
  }
  bool has(const Det& d) const {
    if (keys_.empty()) return false;
    size_t h = mix64(d.up * 0x9e3779b97f4a7c15ULL ^ d.dn) & mask_;
    while (true) {
      const Det& k = keys_[h];
      if (k == d) return true;
      if (k.up == ~0ull && k.dn == ~0ull) return false;
      h = (h + 1) & mask_;
    }
  }
  void insert(const Det& d) {
    if ((size_ + 1) * 2 > keys_.size()) reserve(size_ + 1);
    size_t h = mix64(d.up * 0x9e3779b97f4a7c15ULL ^ d.dn) & mask_;
    while (true) {
      Det& k = keys_[h];
      if (k == d) return;
      if (k.up == ~0ull && k.dn == ~0ull) {
        k = d;
        size_++;
        return;
      }
      h = (h + 1) & mask_;
    }
  }
  size_t size() const { return size_; }

 private:
  std::vector<Det> keys_;
  size_t mask_ = 0;
  size_t size_ = 0;
};

struct RawIntegral {
  double H;
  unsigned p, q, r, s;
};

inline void keep_max_abs(double& a, double x) {
  if (a == 0.0 || std::abs(a) < std::abs(x)) a = x;
}

}  // namespace

void Integrals::resize(unsigned n) {
  n_orbs = n;
  const size_t np = (size_t)n * (n + 1) / 2;
  h1.assign(np, 0.0);
  eri.assign(np * (np + 1) / 2, 0.0);
  orb_order.resize(n);
  std::iota(orb_order.begin(), orb_order.end(), 0u);
}

void Integrals::set_arrays(unsigned n, const double* h1_dense, const double* eri_packed, double core) {
  resize(n);
  energy_core = core;
  for (unsigned p = 0; p < n; p++)
    for (unsigned q = 0; q <= p; q++) h1[combine2(p, q)] = h1_dense[(size_t)p * n + q];
  std::memcpy(eri.data(), eri_packed, eri.size() * sizeof(double));
}

// src/chem/integrals.cc:37-114 (read), 155-226 (HF det), 228-384 (reorder).
void Integrals::load_fcidump(const std::string& path, unsigned n_up, unsigned n_dn,
                             const std::vector<unsigned>& occs_up,
                             const std::vector<unsigned>& occs_dn, int reorder) {
  std::ifstream f(path);
  if (!f.good()) throw std::runtime_error("cannot open FCIDUMP: " + path);
  std::string line;
  unsigned norb = 0, nelec = 0, n_sym = 0;
  bool in_sym = false, done = false;
  while (!done && std::getline(f, line)) {
    std::vector<std::string> words;
    std::string tok;
    for (char c : line) {
      if (std::isspace((unsigned char)c) || c == ',' || c == '=') {
        if (!tok.empty()) words.push_back(tok), tok.clear();
      } else {
        tok.push_back(c);
      }
    }
    if (!tok.empty()) words.push_back(tok);
    for (size_t w = 0; w < words.size(); w++) {
      const std::string& m = words[w];
      if (m == "NORB" || m == "NORBS") {
        norb = (unsigned)std::stoul(words.at(++w));
      } else if (m == "NELEC") {
        nelec = (unsigned)std::stoul(words.at(++w));
      } else if (m == "ORBSYM") {
        in_sym = true;
      } else if (m == "&END" || m == "/" || m == "END") {
        done = true;
        break;
      } else if (in_sym) {
        if (!m.empty() && (std::isdigit((unsigned char)m[0]) || m[0] == '-')) {
          if (++n_sym == norb) in_sym = false;
        } else {
          in_sym = false;
        }
      }
    }
  }
  if (norb == 0) throw std::runtime_error("FCIDUMP without NORB");
  if (norb > 64) throw std::runtime_error("GPU path supports n_orbs <= 64");
  std::vector<RawIntegral> raw;
  double x;
  unsigned p, q, r, s;
  while (f >> x >> p >> q >> r >> s) {
    if (std::abs(x) < 1.0e-9) continue;
    raw.push_back({x, p, q, r, s});
  }
  resize(norb);
  n_elecs = nelec;
  auto key_all = [&](const std::vector<unsigned>& inv) {
    std::fill(h1.begin(), h1.end(), 0.0);
    std::fill(eri.begin(), eri.end(), 0.0);
    for (const auto& it : raw) {
      if (it.p == 0 && it.q == 0 && it.r == 0 && it.s == 0) {
        energy_core = it.H;
      } else if (it.r == 0 && it.s == 0) {
        keep_max_abs(h1[combine2(inv[it.p - 1], inv[it.q - 1])], it.H);
      } else {
        keep_max_abs(eri[combine4(inv[it.p - 1], inv[it.q - 1], inv[it.r - 1], inv[it.s - 1])], it.H);
      }
    }
  };
  std::vector<unsigned> ident(norb);
  std::iota(ident.begin(), ident.end(), 0u);
  key_all(ident);
  hf_up = hf_dn = 0;
  const bool explicit_orbs = !occs_up.empty() || !occs_dn.empty();
  if (explicit_orbs) {
    for (unsigned o : occs_up) hf_up |= 1ull << o;
    for (unsigned o : occs_dn) hf_dn |= 1ull << o;
  } else {
    for (unsigned i = 0; i < n_up; i++) hf_up |= 1ull << i;
    for (unsigned i = 0; i < n_dn; i++) hf_dn |= 1ull << i;
  }
  const bool do_reorder = reorder < 0 ? !explicit_orbs : reorder != 0;
  if (!do_reorder) return;
  std::vector<double> e(norb);
  for (unsigned i = 0; i < norb; i++) {
    e[i] = get_1b(i, i);
    double direct = 0, exchange = 0;
    for (uint64_t h : {hf_up, hf_dn}) {
      for (unsigned o = 0; o < norb; o++) {
        if (!(h >> o & 1)) continue;
        if (o == i) {
          direct += get_2b(i, i, o, o);
        } else {
          direct += 2 * get_2b(i, i, o, o);
          exchange -= get_2b(i, o, o, i);
        }
      }
    }
    e[i] += 0.5 * (direct + exchange);
  }
  std::stable_sort(orb_order.begin(), orb_order.end(),
                   [&](unsigned a, unsigned b) { return e[a] < e[b] - kEps; });
  std::vector<unsigned> inv(norb);
  for (unsigned i = 0; i < norb; i++) inv[orb_order[i]] = i;
  uint64_t up = 0, dn = 0;
  for (unsigned i = 0; i < norb; i++) {
    if (hf_up >> orb_order[i] & 1) up |= 1ull << i;
    if (hf_dn >> orb_order[i] & 1) dn |= 1ull << i;
  }
  hf_up = up;
  hf_dn = dn;
  key_all(inv);
}

Selector::Selector(const Integrals& ints, unsigned n_up, unsigned n_dn)
    : ints_(ints), n_up_(n_up), n_dn_(n_dn), n_orbs_(ints.n_orbs) {
  setup_singles_queue();
  setup_hci_queue();
}

// src/chem/chem_system.cc:237-259 (all orbitals in one irrep).
void Selector::setup_singles_queue() {
  const unsigned n = n_orbs_;
  const unsigned n_elecs = n_up_ + n_dn_;
  singles_queue_.assign(n, {});
  max_singles_ = 0.0;
  for (unsigned p = 0; p < n; p++) {
    for (unsigned r = 0; r < n; r++) {
      if (p == r) continue;
      std::vector<double> el;
      el.reserve(2 * n);
      for (unsigned o = 0; o < n; o++) {
        const double ex = ints_.get_2b(p, o, o, r);
        const double di = ints_.get_2b(p, r, o, o);
        if (o == p || o == r) {
          el.push_back(di);
        } else {
          el.push_back(di);
          el.push_back(di - ex);
        }
      }
      std::sort(el.begin(), el.end());
      double smin = ints_.get_1b(p, r), smax = smin;
      for (unsigned i = 0; i < std::min(n_elecs - 1, 2 * n - 2); i++) {
        smin += el[i];
        smax += el[2 * n - 3 - i];
      }
      const double S = std::max(std::abs(smin), std::abs(smax));
      if (S == 0.0) continue;
      singles_queue_[p].push_back({S, r});
    }
    std::sort(singles_queue_[p].begin(), singles_queue_[p].end(),
              [](const Sr& a, const Sr& b) { return a.S > b.S; });
    if (!singles_queue_[p].empty()) max_singles_ = std::max(max_singles_, singles_queue_[p].front().S);
  }
}

// src/chem/chem_system.cc:115-196 and 261-287 (all orbitals in one irrep).
void Selector::setup_hci_queue() {
  const unsigned n = n_orbs_;
  hci_queue_.assign(combine2(n, 2 * n), {});
  max_hci_ = 0.0;
#pragma omp parallel for schedule(dynamic, 1) reduction(max : max_hci_)
  for (unsigned p = 0; p < n; p++) {
    // Same spin: |(pr|qs) - (ps|qr)|, r < s.
    for (unsigned q = p + 1; q < n; q++) {
      auto& lst = hci_queue_[combine2(p, q)];
      for (unsigned r = 0; r < n; r++)
        for (unsigned s = r + 1; s < n; s++) {
          if (p == r || q == s || p == s || q == r) continue;
          const double H = std::abs(ints_.get_2b(p, r, q, s) - ints_.get_2b(p, s, q, r));
          if (H == 0.0) continue;
          lst.push_back({H, r, s});
        }
      std::sort(lst.begin(), lst.end(), [](const Hrs& a, const Hrs& b) { return a.H > b.H; });
      if (!lst.empty()) max_hci_ = std::max(max_hci_, lst.front().H);
    }
    // Opposite spin: |(pr|q's')|, q' = q - n >= p, s' = s - n.
    for (unsigned q = n + p; q < 2 * n; q++) {
      auto& lst = hci_queue_[combine2(p, q)];
      const unsigned qq = q - n;
      for (unsigned r = 0; r < n; r++)
        for (unsigned s = 0; s < n; s++) {
          if (p == r || qq == s) continue;  // get_hci_queue_elem, 263 (offset indices)
          const double H = std::abs(ints_.get_2b(p, r, qq, s));
          if (H == 0.0) continue;
          lst.push_back({H, r, s + n});
        }
      std::sort(lst.begin(), lst.end(), [](const Hrs& a, const Hrs& b) { return a.H > b.H; });
      if (!lst.empty()) max_hci_ = std::max(max_hci_, lst.front().H);
    }
  }
}

double Selector::hf_energy() const {
  double e1 = 0, direct = 0, exchange = 0;
  std::vector<unsigned> ou, od;
  for (unsigned o = 0; o < 64; o++) {
    if (ints_.hf_up >> o & 1) ou.push_back(o);
    if (ints_.hf_dn >> o & 1) od.push_back(o);
  }
  for (unsigned o : ou) e1 += ints_.get_1b(o, o);
  for (unsigned o : od) e1 += ints_.get_1b(o, o);
  auto same = [&](const std::vector<unsigned>& oc) {
    for (size_t i = 0; i < oc.size(); i++)
      for (size_t j = i + 1; j < oc.size(); j++) {
        direct += ints_.get_2b(oc[i], oc[i], oc[j], oc[j]);
        exchange -= ints_.get_2b(oc[i], oc[j], oc[j], oc[i]);
      }
  };
  same(ou);
  same(od);
  for (unsigned a : ou)
    for (unsigned b : od) direct += ints_.get_2b(a, a, b, b);
  return e1 + direct + exchange + ints_.energy_core;
}

// Single-excitation element, src/chem/chem_system.cc:544-577 (used by the
// selection's single filter, src/solver/solver.h:456-459).
double Selector::elem_single(const Det& di, const Det& dj) const {
  const bool is_up = di.up != dj.up;
  const uint64_t a = is_up ? di.up : di.dn, b = is_up ? dj.up : dj.dn;
  const unsigned oi = __builtin_ctzll(a & ~b), oj = __builtin_ctzll(b & ~a);
  const unsigned cnt = __builtin_popcountll(a & ((1ull << oi) - 1)) + __builtin_popcountll(b & ((1ull << oj) - 1));
  const int pf = (cnt & 1) ? -1 : 1;
  const double one = pf * ints_.get_1b(oi, oj);
  uint64_t oppo = is_up ? di.dn : di.up;
  double e = 0.0;
  for (uint64_t m = a; m; m &= m - 1) {
    const unsigned o = __builtin_ctzll(m);
    if (o == oi || o == oj) continue;
    e -= ints_.get_2b(oi, o, o, oj);
    const double direct = ints_.get_2b(oi, oj, o, o);
    if (oppo >> o & 1) {
      oppo &= ~(1ull << o);
      e += 2 * direct;
    } else {
      e += direct;
    }
  }
  for (uint64_t m = oppo; m; m &= m - 1) e += ints_.get_2b(oi, oj, __builtin_ctzll(m), __builtin_ctzll(m));
  e *= pf;
  return one + e;
}

// src/chem/chem_system.cc:289-402 (second_rejection and active space off).
template <class F>
void Selector::find_connected(const Det& det, double eps_max, double eps_min, F&& handler) const {
  if (eps_max < eps_min) return;
  const unsigned n = n_orbs_;
  unsigned occ_up[64], occ_dn[64];
  unsigned nu = 0, nd = 0;
  for (uint64_t m = det.up; m; m &= m - 1) occ_up[nu++] = __builtin_ctzll(m);
  for (uint64_t m = det.dn; m; m &= m - 1) occ_dn[nd++] = __builtin_ctzll(m);
  const unsigned n_elecs = nu + nd;
  if (eps_min <= max_singles_) {
    for (unsigned pid = 0; pid < n_elecs; pid++) {
      const unsigned p = pid < nu ? occ_up[pid] : occ_dn[pid - nu];
      for (const Sr& sr : singles_queue_[p]) {
        if (sr.S < eps_min) break;
        const unsigned r = sr.r;
        Det c = det;
        if (pid < nu) {
          if (det.up >> r & 1) continue;
          c.up = (c.up & ~(1ull << p)) | (1ull << r);
        } else {
          if (det.dn >> r & 1) continue;
          c.dn = (c.dn & ~(1ull << p)) | (1ull << r);
        }
        handler(c, 1);
      }
    }
  }
  if (eps_min > max_hci_) return;
  for (unsigned pid = 0; pid < n_elecs; pid++) {
    for (unsigned qid = pid + 1; qid < n_elecs; qid++) {
      const unsigned p = pid < nu ? occ_up[pid] : occ_dn[pid - nu] + n;
      const unsigned q = qid < nu ? occ_up[qid] : occ_dn[qid - nu] + n;
      unsigned p2 = p, q2 = q;
      if (p >= n && q >= n) {
        p2 -= n;
        q2 -= n;
      } else if (p < n && q >= n && p > q - n) {
        p2 = q - n;
        q2 = p + n;
      }
      for (const Hrs& hrs : hci_queue_[combine2(p2, q2)]) {
        const double H = hrs.H;
        if (H < eps_min) break;
        if (H >= eps_max) continue;
        unsigned r = hrs.r, s = hrs.s;
        if (p >= n && q >= n) {
          r += n;
          s += n;
        } else if (p < n && q >= n && p > q - n) {
          const unsigned tmp_r = s - n;
          s = r + n;
          r = tmp_r;
        }
        const bool occ_r = r < n ? (det.up >> r & 1) : (det.dn >> (r - n) & 1);
        if (occ_r) continue;
        const bool occ_s = s < n ? (det.up >> s & 1) : (det.dn >> (s - n) & 1);
        if (occ_s) continue;
        Det c = det;
        if (p < n) c.up &= ~(1ull << p); else c.dn &= ~(1ull << (p - n));
        if (q < n) c.up &= ~(1ull << q); else c.dn &= ~(1ull << (q - n));
        if (r < n) c.up |= 1ull << r; else c.dn |= 1ull << (r - n);
        if (s < n) c.up |= 1ull << s; else c.dn |= 1ull << (s - n);
        handler(c, 2);
      }
    }
  }
}

size_t Selector::select(std::vector<Det>& dets, const std::vector<std::vector<double>>& coefs,
                        double eps_var, std::vector<double>& eps_tried_prev) {
  const size_t n_dets = dets.size();
  eps_tried_prev.resize(n_dets, 1.0e100);
  DetSet var;
  var.reserve(n_dets);
  for (const Det& d : dets) var.insert(d);
  const int nt = omp_get_max_threads();
  std::vector<std::vector<Det>> found(nt);
  auto compact = [](std::vector<Det>& v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  };
#pragma omp parallel
  {
    const int tid = omp_get_thread_num();
    auto& mine = found[tid];
    size_t next_compact = 1u << 22;
#pragma omp for schedule(dynamic, 64)
    for (size_t i = 0; i < n_dets; i++) {
      double max_coef = coefs[0][i];
      for (size_t s = 1; s < coefs.size(); s++)
        if (std::abs(coefs[s][i]) > std::abs(max_coef)) max_coef = coefs[s][i];
      const double eps_min = eps_var / std::abs(max_coef);
      if (eps_min >= eps_tried_prev[i]) continue;
      const Det& det = dets[i];
      find_connected(det, eps_tried_prev[i], eps_min, [&](const Det& c, int n_excite) {
        if (var.has(c)) return;
        if (n_excite == 1 && std::abs(elem_single(det, c)) < eps_min) return;
        mine.push_back(c);
      });
      eps_tried_prev[i] = 1.000000001 * eps_min;
      if (mine.size() > next_compact) {
        compact(mine);
        next_compact = std::max(next_compact, mine.size() * 2);
      }
    }
    compact(mine);
  }
  std::vector<Det> all;
  size_t tot = 0;
  for (auto& v : found) tot += v.size();
  all.reserve(tot);
  for (auto& v : found) {
    all.insert(all.end(), v.begin(), v.end());
    std::vector<Det>().swap(v);
  }
  compact(all);
  dets.insert(dets.end(), all.begin(), all.end());
  eps_tried_prev.resize(dets.size(), 1.0e100);
  return all.size();
}

}  // namespace shci_host
