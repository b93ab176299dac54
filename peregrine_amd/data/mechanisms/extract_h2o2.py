"""Writes h2o2.yaml, the H/O-only subset of GRI-Mech 3.0, from the reference's GRI30.yaml.

Run once where the reference tree is mounted; the output is committed data. The
species kept are H2, H, O, O2, OH, H2O, HO2, H2O2 and N2 (the bath, last); a
reaction is kept when every species on both sides is among them, and its
third-body efficiencies are restricted to them.
"""

import re
import sys

import yaml

SRC = "/root/reference/peregrine/simulator/mixture/database/mechanisms/GRI30.yaml"
KEEP = "H2 H O O2 OH H2O HO2 H2O2 N2".split()


def names(side):
    side = re.sub(r"\(\s*\+[^)]+\)", "", side)
    out = []
    for t in (t.strip() for t in side.split("+")):
        if t and t != "M":
            m = re.match(r"^(\d+\.?\d*)\s+(.+)$", t)
            out.append(m.group(2) if m else t)
    return out


def main(dst):
    d = yaml.safe_load(open(SRC))
    rx = []
    for r in d["reactions"]:
        lhs, rhs = re.split(r"<=>|=>", r["equation"])
        if all(x in KEEP for x in names(lhs) + names(rhs)):
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
