"""Writes h2o2.yaml, the H/O-only subset of GRI-Mech 3.0, from the reference's GRI30.yaml.

Run once where the reference tree is mounted, with the destination directory as its argument
(peregrine_amd/data/mechanisms); the output is committed data. The
species kept are H2, H, O, O2, OH, H2O, HO2, H2O2 and N2 (the bath, last); a
reaction is kept when every species on both sides is among them, and its
third-body efficiencies are restricted to them.
"""

import sys
from pathlib import Path

import yaml

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
from peregrine_amd.mixture.parser import Equation  # noqa: E402

SRC = "/root/reference/peregrine/simulator/mixture/database/mechanisms/GRI30.yaml"
KEEP = "H2 H O O2 OH H2O HO2 H2O2 N2".split()


def species_of(equation):
    """Every species an equation names on either side (its third bodies aside)."""
    eq = Equation(equation)
    return list(eq.reactants) + list(eq.products)


def main(dst):
    d = yaml.safe_load(open(SRC))
    rx = []
    for r in d["reactions"]:
        if all(x in KEEP for x in species_of(r["equation"])):
            r = dict(r)
            if "efficiencies" in r:
                r["efficiencies"] = {k: v for k, v in r["efficiencies"].items() if k in KEEP}
            rx.append(r)
    sp = sorted((s for s in d["species"] if s["name"] in KEEP), key=lambda s: KEEP.index(s["name"]))
    out = {
        "description": "H2/O2 with N2 as the bath: the H/O-only reactions and species of GRI-Mech 3.0 "
        "(Smith et al., http://combustion.berkeley.edu/gri-mech/version30/text30.html), third-body "
        "efficiencies restricted to these species.",
        "units": d["units"],
        "phases": [{"name": "h2o2", "thermo": "ideal-gas", "elements": ["O", "H", "N"], "species": KEEP,
                    "kinetics": "gas", "state": {"T": 300.0, "P": "1 atm"}}],
        "species": sp,
        "reactions": rx,
    }
    with open(dst, "w") as f:
        yaml.safe_dump(out, f, sort_keys=False, default_flow_style=None, width=100)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "h2o2.yaml")
