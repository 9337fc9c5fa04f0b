// RDM file outputs of the reference (src/chem/rdm.cc): 1rdm.csv, 2rdm.csv and
// the Dice-convention spatialRDM.txt, written from the GPU RDM of shci_rdm
// (1-RDM n x n row-major, spin-summed; 2-RDM packed a = p n + s, b = q n + r,
// a >= b).  Orbitals are printed in the original (FCIDUMP) order through
// orb_order (internal orbital i = original orbital orb_order[i],
// integrals.cc:258-283).  The loops, thresholds and printf formats are the
// reference's, so the files are byte-compatible with its downstream readers.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "shci_host.h"

namespace {

thread_local std::string g_io_err;

int io_fail(const std::string& m) {
  g_io_err = m;
  return -1;
}

// Element D_pqrs of the packed 2-RDM (RDM::combine4_2rdm, rdm.cc:233-241).
double rdm2_elem(const double* rdm2, size_t n, unsigned p, unsigned q, unsigned r, unsigned s) {
  const size_t a = (size_t)p * n + s, b = (size_t)q * n + r;
  return a > b ? rdm2[a * (a + 1) / 2 + b] : rdm2[b * (b + 1) / 2 + a];
}

std::vector<unsigned> order_or_identity(unsigned n, const unsigned* orb_order) {
  std::vector<unsigned> o(n);
  for (unsigned i = 0; i < n; i++) o[i] = orb_order ? orb_order[i] : i;
  return o;
}

}  // namespace

extern "C" {

const char* shci_host_io_last_error(void) { return g_io_err.c_str(); }

// RDM::dump_1rdm (rdm.cc:202-221).
int shci_host_dump_1rdm(const char* path, unsigned n_orbs, const unsigned* orb_order, const double* rdm1) {
  if (!path || !rdm1) return io_fail("null argument");
  const std::vector<unsigned> ord = order_or_identity(n_orbs, orb_order);
  FILE* f = std::fopen(path, "w");
  if (!f) return io_fail(std::string("cannot open ") + path);
  std::fprintf(f, "p,r,1rdm\n");
  for (unsigned p = 0; p < n_orbs; p++) {
    for (unsigned r = p; r < n_orbs; r++) {
      const double v = rdm1[(size_t)p * n_orbs + r];
      if (std::abs(v) < 1e-9) continue;
      std::fprintf(f, "%d,%d,%#.15g\n", (int)ord[p], (int)ord[r], v);
    }
  }
  return std::fclose(f) == 0 ? 0 : io_fail("write failed");
}

// RDM::dump_2rdm (rdm.cc:271-331): csv -> 2rdm.csv, else spatialRDM.txt.
int shci_host_dump_2rdm(const char* path, unsigned n_orbs, const unsigned* orb_order, const double* rdm2, int csv) {
  if (!path || !rdm2) return io_fail("null argument");
  const size_t n = n_orbs;
  const std::vector<unsigned> ord = order_or_identity(n_orbs, orb_order);
  std::vector<unsigned> inv(n_orbs);
  for (unsigned i = 0; i < n_orbs; i++) {
    if (ord[i] >= n_orbs) return io_fail("orb_order is not a permutation");
    inv[ord[i]] = i;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) return io_fail(std::string("cannot open ") + path);
  if (csv) {
    std::fprintf(f, "p,q,r,s,2rdm\n");
    for (unsigned p = 0; p < n_orbs; p++)
      for (unsigned q = p; q < n_orbs; q++)
        for (unsigned s = 0; s < n_orbs; s++)
          for (unsigned r = 0; r < n_orbs; r++) {
            if (p == q && s > r) continue;
            const double v = rdm2_elem(rdm2, n, p, q, r, s);
            if (std::abs(v) < 1.0e-9) continue;
            std::fprintf(f, "%d,%d,%d,%d,%#.15g\n", (int)ord[p], (int)ord[q], (int)ord[r], (int)ord[s], v);
          }
  } else {
    std::fprintf(f, "%d\n", (int)n_orbs);
    for (unsigned p = 0; p < n_orbs; p++)
      for (unsigned q = 0; q < n_orbs; q++)
        for (unsigned s = 0; s < n_orbs; s++)  // r and s switched (Dice convention)
          for (unsigned r = 0; r < n_orbs; r++) {
            const double v = rdm2_elem(rdm2, n, inv[p], inv[q], inv[r], inv[s]);
            if (std::abs(v) > 1.e-6) std::fprintf(f, "%3d   %3d   %3d   %3d   %10.8g\n", (int)p, (int)q, (int)s, (int)r, v);
          }
  }
  return std::fclose(f) == 0 ? 0 : io_fail("write failed");
}

}  // extern "C"
