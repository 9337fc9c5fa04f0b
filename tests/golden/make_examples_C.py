"""Writes the examples/C fixtures from the reference's own example run:

* ``C_FCIDUMP``         -- copy of examples/C/FCIDUMP (22 orbitals, D2h, 4 up / 2 dn)
* ``examples_C.json``   -- per-iteration data printed in examples/C/out
  (dets, "Number of nonzero elems", "Full size of absingles", "Outer size of
  a/b singles", the "Davidson #k" energies), the deterministic PT2 stage
  ("Number of dtm pt dets", "PT dtm correction (eps1= ..., eps_pt_dtm= ...)",
  the "Variational energy" it starts from) and examples/C/result.json.

Run in the development container (it reads /root/reference); the outputs are
committed so the tests do not need the reference at run time.
"""
import json
import re
import shutil
from pathlib import Path

REF = Path("/root/reference/examples/C")
OUT = Path(__file__).resolve().parent


def main():
    shutil.copyfile(REF / "FCIDUMP", OUT / "C_FCIDUMP")
    iters = []
    cur = None
    for line in (REF / "out").read_text().splitlines():
        m = re.match(r"Number of dets / new dets: ([\d,]+) / ([\d,]+)", line)
        if m:
            cur = {"n_dets": int(m.group(1).replace(",", "")), "n_new": int(m.group(2).replace(",", "")),
                   "davidson": []}
            iters.append(cur)
            continue
        m = re.match(r"Outer size of a/b singles: ([\d,]+) / ([\d,]+)", line)
        if m and cur is not None:
            cur["n_unique_alpha"] = int(m.group(1).replace(",", ""))
            cur["n_unique_beta"] = int(m.group(2).replace(",", ""))
        m = re.match(r"Full size of absingles: ([\d,]+)", line)
        if m and cur is not None:
            cur["singles"] = int(m.group(1).replace(",", ""))
        m = re.match(r"Number of nonzero elems: ([\d,]+)", line)
        if m and cur is not None:
            cur["nnz"] = int(m.group(1).replace(",", ""))
        m = re.match(r"Davidson #(\d+):\s+(-?[\d.]+)", line)
        if m and cur is not None:
            cur["davidson"].append(float(m.group(2)))
        m = re.match(r"Iteration (\d+) eps1= ([\d.e+-]+) ndets= ([\d,]+) energy= (-?[\d.]+)", line)
        if m and cur is not None:
            cur["eps"] = float(m.group(2))
            cur["energy"] = float(m.group(4))
    pt_dtm = []
    var_energy = None
    n_pt = None
    for line in (REF / "out").read_text().splitlines():
        m = re.match(r"Variational energy: (-?[\d.]+) Ha", line)
        if m:
            var_energy = float(m.group(1))
        m = re.match(r"Number of dtm pt dets: ([\d,]+)", line)
        if m:
            n_pt = int(m.group(1).replace(",", ""))
        m = re.match(r"PT dtm correction \(eps1= ([\d.e+-]+), eps_pt_dtm= ([\d.e+-]+)\): (-?[\d.]+)", line)
        if m:
            pt_dtm.append({"eps_var": float(m.group(1)), "eps_pt_dtm": float(m.group(2)),
                           "correction": float(m.group(3)), "n_pt_dets": n_pt, "energy_var": var_energy})
    result = json.loads((REF / "result.json").read_text())
    data = {"n_up": 4, "n_dn": 2, "energy_hf": result["energy_hf"], "energy_var": result["energy_var"],
            "iterations": iters, "pt_dtm": pt_dtm}
    (OUT / "examples_C.json").write_text(json.dumps(data, indent=1))


if __name__ == "__main__":
    main()
