"""Writes tests/golden/fcc_binary_vacancy_basis.json from the reference's
FCC_binary_vacancy fixture files (prim.json, basis_sets/bset.default/
basis.json and bspecs.json, formation_energy_eci.json): the prototype
clusters, multiplicities and the cluster-function terms parsed from the
basis.json function strings.  Run in a container that holds the reference
(read as data; nothing of the reference is executed)."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from casmcode_clexmonte_amd.system import parse_casm_function_latex  # noqa: E402

REF = "/root/reference/tests/unit/clexmonte/data/FCC_binary_vacancy"


def main():
    prim = json.load(open(os.path.join(REF, "prim.json")))
    basis = json.load(open(os.path.join(REF, "basis_sets/bset.default/basis.json")))
    eci = json.load(open(os.path.join(REF, "formation_energy_eci.json")))
    orbits = []
    for o in basis["orbits"]:
        funcs = []
        for cf in o["cluster_functions"]:
            key = [k for k in cf if k.startswith("\\Phi")][0]
            terms = parse_casm_function_latex(cf[key])
            funcs.append({"linear_function_index": cf["linear_function_index"],
                          "terms": [[c, sorted([[int(s), int(f)] for s, f in d.items()])] for c, d in terms]})
        orbits.append({"mult": o["mult"], "prototype_sites": o["prototype"]["sites"],
                       "max_length": o["prototype"]["max_length"], "functions": funcs})
    eci_by_lfi = {}
    for o in eci["orbits"]:
        for cf in o["cluster_functions"]:
            if "eci" in cf:
                eci_by_lfi[cf["linear_function_index"]] = cf["eci"]
    out = {"source": "tests/unit/clexmonte/data/FCC_binary_vacancy (reference fixture)",
           "prim": prim, "max_length": {"2": 4.1}, "site_basis": "occupation",
           "site_functions": [[0.0, 1.0, 0.0], [0.0, 0.0, 1.0]],
           "corr_size": 9, "nlist_size": 19,
           "orbits": orbits, "eci": eci_by_lfi}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fcc_binary_vacancy_basis.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
