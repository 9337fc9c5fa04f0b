// State behind the opaque handle of include/shci_host.h.
#pragma once

#include <string>
#include <vector>

#include "chem_host.h"

namespace shci_host {

extern thread_local std::string g_host_err;

struct HostSystem {
  Integrals ints;
  unsigned n_up = 0, n_dn = 0;
  Selector* sel = nullptr;
  std::vector<Det> dets;
  std::vector<double> eps_tried;
  ~HostSystem() { delete sel; }
};

}  // namespace shci_host
