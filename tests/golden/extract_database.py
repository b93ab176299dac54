"""Writes the species, collision-integral and Chung tables this package reads, from the
reference's database files.

Run once where the reference tree is mounted, with the destination directory as its argument
(peregrine_amd/data); the outputs are committed data:
- species.yaml: the species of the H/O/N mixtures (H2, H, O, O2, OH, H2O, HO2, H2O2, N2, AR)
  with their molecular data as the reference compiles them (NIST-JANAF thermo tables,
  Lennard-Jones and polar transport parameters, critical constants), each entry keeping
  its source note;
- integrals.yaml: the Stockmayer collision integrals Omega(2,2)* and A* of Monchick and Mason,
  J. Chem. Phys. 35, 1676 (1961), against T* and delta*;
- chung.yaml: the dense-gas coefficient tables of Chung et al., Ind. Eng. Chem. Res. 27, 671
  (1988).
"""

import sys
from pathlib import Path

import yaml

SRC = Path("/root/reference/peregrine/simulator/mixture/database")
KEEP = "H2 H O O2 OH H2O HO2 H2O2 N2 AR".split()


def main(dst):
    dst = Path(dst)
    species = yaml.safe_load(open(SRC / "species.yaml"))
    sub = {name: species[name] for name in KEEP if name in species}
    with open(dst / "species.yaml", "w") as f:
        f.write("# Molecular data of the H/O/N species: NIST-JANAF thermo tables (J/kmol/K, J/kmol),\n"
                "# Lennard-Jones and polar transport parameters (SI), critical constants; extracted by\n"
                "# tests/golden/extract_database.py.\n")
        yaml.safe_dump(sub, f, sort_keys=False, default_flow_style=None, width=100)
    ints = yaml.safe_load(open(SRC / "integrals.yaml"))
    with open(dst / "integrals.yaml", "w") as f:
        f.write("# Stockmayer (12-6-3) collision integrals: omega22 = Omega(2,2)*, astar = A*, rows T*,\n"
                "# columns delta*. Monchick and Mason, J. Chem. Phys. 35, 1676 (1961).\n")
        yaml.safe_dump({k: ints[k] for k in ("delta", "tstar22", "omega22", "tstar", "astar")}, f,
                       sort_keys=False, default_flow_style=None, width=100)
    chung = yaml.safe_load(open(SRC / "chung.yaml"))
    with open(dst / "chung.yaml", "w") as f:
        f.write("# Chung, Ajlan, Lee and Starling, Ind. Eng. Chem. Res. 27, 671 (1988): rows (a0, a1, a2, a3)\n"
                "# of E_i = a0 + a1 w + a2 mu_r^4 + a3 kappa, viscosity E1..E10, conductivity B1..B7.\n")
        yaml.safe_dump(chung, f, sort_keys=False, default_flow_style=None, width=100)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else Path(__file__).resolve().parent)
