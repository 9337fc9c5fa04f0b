// Wavefunction files of the reference, wf_eps1_<eps_var>.dat
// (Solver::save_variation_result / load_variation_result,
// src/solver/solver.h:1179-1224): the serialized BaseSystem, its fields in the
// order of BaseSystem::serialize (src/base_system.h:125-133):
//
//   n_up, n_dn          unsigned     -> varint
//   dets                vector<Det>  -> varint count, then per det up, dn;
//                                       a HalfDet is its 2 uint64 chunks
//                                       (half_det.h:58-66), orbitals 0-63 in
//                                       chunk 0, each a varint
//   coefs               vector<vector<double>> -> varint n_states, then per
//                                       state a varint length and the doubles
//   energy_hf           double       -> 8 bytes
//   energy_var          vector<double>
//   time_sym            bool         -> varint (one byte, 0 or 1)
//
// The encoding is that of the hps serializer the reference links (an absent
// submodule): unsigned integers as base-128 varints, low 7-bit group first,
// bit 7 set on every byte but the last; floating-point numbers as their 8
// little-endian IEEE-754 bytes; a vector as its size (varint) followed by its
// elements.  Dets here hold n_orb <= 64 orbitals: a file whose chunk 1 is not
// zero is rejected.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "shci_host.h"

namespace {

thread_local std::string g_wf_err;

int wf_fail(const std::string& m) {
  g_wf_err = m;
  return -1;
}

void put_varint(std::string& b, uint64_t v) {
  while (v >= 128) {
    b.push_back((char)((v & 127) | 128));
    v >>= 7;
  }
  b.push_back((char)v);
}

void put_double(std::string& b, double x) {
  unsigned char raw[8];
  std::memcpy(raw, &x, 8);  // x86-64 and gfx hosts are little-endian
  b.append(reinterpret_cast<const char*>(raw), 8);
}

struct Reader {
  const std::string& b;
  size_t pos = 0;
  std::string err;

  bool varint(uint64_t& v) {
    v = 0;
    for (int k = 0; k < 10; k++) {
      if (pos >= b.size()) return fail("truncated varint");
      const uint8_t c = (uint8_t)b[pos++];
      if (k == 9 && c > 1) return fail("varint exceeds 64 bits");
      v |= (uint64_t)(c & 127) << (7 * k);
      if (!(c & 128)) return true;
    }
    return fail("varint longer than 10 bytes");
  }
  bool u32(unsigned& v) {
    uint64_t x;
    if (!varint(x)) return false;
    if (x > 0xffffffffull) return fail("unsigned field exceeds 32 bits");
    v = (unsigned)x;
    return true;
  }
  bool f64(double& x) {
    if (b.size() - pos < 8) return fail("truncated double");
    std::memcpy(&x, b.data() + pos, 8);
    pos += 8;
    return true;
  }
  bool fail(const char* m) {
    err = m;
    return false;
  }
};

struct Wf {
  unsigned n_up = 0, n_dn = 0;
  std::vector<uint64_t> up, dn;
  std::vector<std::vector<double>> coefs;
  double energy_hf = 0.0;
  std::vector<double> energy_var;
  bool time_sym = false;
};

// Parses a whole file into w (counts only when `counts_only`: the dets and
// coefficients are skipped, not stored).
int parse(const char* path, Wf& w, bool counts_only, size_t& n_dets) {
  if (!path) return wf_fail("null path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return wf_fail(std::string("cannot open ") + path);
  std::string buf;
  char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) buf.append(chunk, got);
  const bool read_err = std::ferror(f) != 0;
  std::fclose(f);
  if (read_err) return wf_fail(std::string("read error on ") + path);
  Reader r{buf};
  auto bad = [&](const char* what) { return wf_fail(std::string(path) + ": " + what + ": " + r.err); };
  uint64_t nd;
  if (!r.u32(w.n_up) || !r.u32(w.n_dn)) return bad("n_up/n_dn");
  if (!r.varint(nd)) return bad("det count");
  if (nd > (buf.size() - r.pos) / 4) return wf_fail(std::string(path) + ": det count larger than the file");
  n_dets = (size_t)nd;
  if (!counts_only) {
    w.up.resize(n_dets);
    w.dn.resize(n_dets);
  }
  for (size_t i = 0; i < n_dets; i++) {
    uint64_t c[4];
    for (auto& x : c)
      if (!r.varint(x)) return bad("det");
    if (c[1] != 0 || c[3] != 0)
      return wf_fail(std::string(path) + ": det " + std::to_string(i) + " occupies an orbital >= 64");
    if (!counts_only) w.up[i] = c[0], w.dn[i] = c[2];
  }
  uint64_t ns;
  if (!r.varint(ns)) return bad("coefs");
  if (ns > 64) return wf_fail(std::string(path) + ": more than 64 states");
  w.coefs.resize(ns);
  for (auto& cs : w.coefs) {
    uint64_t len;
    if (!r.varint(len)) return bad("coefs");
    if (len != n_dets) return wf_fail(std::string(path) + ": a state's coefficient count differs from the det count");
    if (buf.size() - r.pos < 8 * len) return wf_fail(std::string(path) + ": truncated coefficients");
    if (!counts_only) {
      cs.resize(len);
      std::memcpy(cs.data(), buf.data() + r.pos, 8 * len);
    }
    r.pos += 8 * len;
  }
  if (!r.f64(w.energy_hf)) return bad("energy_hf");
  uint64_t ne;
  if (!r.varint(ne)) return bad("energy_var");
  if (ne != ns) return wf_fail(std::string(path) + ": energy_var and coefs hold different state counts");
  w.energy_var.resize(ne);
  for (auto& e : w.energy_var)
    if (!r.f64(e)) return bad("energy_var");
  uint64_t ts;
  if (!r.varint(ts)) return bad("time_sym");
  if (ts > 1) return wf_fail(std::string(path) + ": time_sym is not 0 or 1");
  w.time_sym = ts != 0;
  if (r.pos != buf.size()) return wf_fail(std::string(path) + ": trailing bytes after time_sym");
  return 0;
}

}  // namespace

extern "C" {

const char* shci_host_wf_last_error(void) { return g_wf_err.c_str(); }

int shci_host_save_wf(const char* path, unsigned n_up, unsigned n_dn, const uint64_t* up, const uint64_t* dn,
                      size_t n_dets, const double* coefs, unsigned n_states, double energy_hf,
                      const double* energy_var, int time_sym) {
  if (!path || (n_dets && (!up || !dn || (n_states && !coefs))) || (n_states && !energy_var))
    return wf_fail("null argument");
  std::string b;
  b.reserve(32 + n_dets * (24 + 8 * (size_t)n_states));
  put_varint(b, n_up);
  put_varint(b, n_dn);
  put_varint(b, n_dets);
  for (size_t i = 0; i < n_dets; i++) {
    put_varint(b, up[i]);
    put_varint(b, 0);  // chunk 1: orbitals 64-127
    put_varint(b, dn[i]);
    put_varint(b, 0);
  }
  put_varint(b, n_states);
  for (unsigned s = 0; s < n_states; s++) {
    put_varint(b, n_dets);
    for (size_t i = 0; i < n_dets; i++) put_double(b, coefs[(size_t)s * n_dets + i]);
  }
  put_double(b, energy_hf);
  put_varint(b, n_states);
  for (unsigned s = 0; s < n_states; s++) put_double(b, energy_var[s]);
  put_varint(b, time_sym ? 1 : 0);
  FILE* f = std::fopen(path, "wb");
  if (!f) return wf_fail(std::string("cannot create ") + path);
  const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
  if (std::fclose(f) != 0 || !ok) return wf_fail(std::string("write error on ") + path);
  return 0;
}

int shci_host_wf_info(const char* path, unsigned* n_up, unsigned* n_dn, size_t* n_dets, unsigned* n_states,
                      int* time_sym) {
  Wf w;
  size_t nd = 0;
  if (parse(path, w, true, nd)) return -1;
  if (n_up) *n_up = w.n_up;
  if (n_dn) *n_dn = w.n_dn;
  if (n_dets) *n_dets = nd;
  if (n_states) *n_states = (unsigned)w.coefs.size();
  if (time_sym) *time_sym = w.time_sym ? 1 : 0;
  return 0;
}

int shci_host_load_wf(const char* path, uint64_t* up, uint64_t* dn, size_t n_dets, double* coefs, unsigned n_states,
                      double* energy_hf, double* energy_var) {
  Wf w;
  size_t nd = 0;
  if (parse(path, w, false, nd)) return -1;
  if (nd != n_dets || w.coefs.size() != n_states)
    return wf_fail(std::string(path) + ": holds " + std::to_string(nd) + " dets and " +
                   std::to_string(w.coefs.size()) + " states, the buffers " + std::to_string(n_dets) + " and " +
                   std::to_string(n_states));
  if (n_dets && (!up || !dn || (n_states && !coefs))) return wf_fail("null argument");
  for (size_t i = 0; i < nd; i++) up[i] = w.up[i], dn[i] = w.dn[i];
  for (unsigned s = 0; s < n_states; s++)
    if (nd) std::memcpy(coefs + (size_t)s * nd, w.coefs[s].data(), 8 * nd);
  if (energy_hf) *energy_hf = w.energy_hf;
  if (energy_var)
    for (unsigned s = 0; s < n_states; s++) energy_var[s] = w.energy_var[s];
  return 0;
}

}  // extern "C"
