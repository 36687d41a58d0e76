"""Writes tests/golden/casm_clex_fixtures.json from two of the reference's
fixture directories, read as data (nothing of the reference is executed):

* tests/unit/clexmonte/data/Clex_ZrO_Occ: prim.json, the composition axes of
  system.json, formation_energy_eci.json,
  run_params_complete.json (canonical series) and run_params_sgc_complete.json
  (semi-grand canonical series);
* tests/unit/clexmonte/data/FCC_binary_vacancy: prim.json,
  formation_energy_eci.json and formation_energy_sparse_eci.json (a list of
  [linear_function_index, eci] pairs, kept as is).

The ECI files keep only what a cluster expansion needs (bspecs, site
functions, and per orbit: mult, prototype sites and lengths, and the cluster
functions with their latex, linear_function_index and eci).  Run in a
container that holds the reference."""
import json
import os

REF = "/root/reference/tests/unit/clexmonte/data"


def _load(*parts):
    with open(os.path.join(REF, *parts)) as f:
        return json.load(f)


def compact_eci(d):
    orbits = []
    for o in d["orbits"]:
        p = o["prototype"]
        orbits.append({
            "linear_orbit_index": o["linear_orbit_index"] if "linear_orbit_index" in o else len(orbits),
            "mult": o["mult"],
            "prototype": {"sites": p["sites"], "max_length": p.get("max_length", 0.0),
                          "min_length": p.get("min_length", 0.0)},
            "cluster_functions": o["cluster_functions"],
        })
    return {"bspecs": d["bspecs"], "site_functions": d.get("site_functions"), "orbits": orbits}


def main():
    out = {
        "source": "reference fixtures tests/unit/clexmonte/data/{Clex_ZrO_Occ,FCC_binary_vacancy}",
        "Clex_ZrO_Occ": {
            "prim": _load("Clex_ZrO_Occ", "prim.json"),
            "composition_axes": _load("Clex_ZrO_Occ", "system.json")["composition_axes"],
            "formation_energy_eci": compact_eci(_load("Clex_ZrO_Occ", "formation_energy_eci.json")),
            "run_params_complete": _load("Clex_ZrO_Occ", "run_params_complete.json"),
            "run_params_sgc_complete": _load("Clex_ZrO_Occ", "run_params_sgc_complete.json"),
        },
        "FCC_binary_vacancy": {
            "prim": _load("FCC_binary_vacancy", "prim.json"),
            "formation_energy_eci": compact_eci(_load("FCC_binary_vacancy", "formation_energy_eci.json")),
            "formation_energy_sparse_eci": _load("FCC_binary_vacancy", "formation_energy_sparse_eci.json"),
        },
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "casm_clex_fixtures.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=None, separators=(",", ":"))
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
