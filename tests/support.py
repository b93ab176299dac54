"""Shared test helpers: cases, seeded states, and the oracle run on host arrays.

Arrays here are numpy in the device layout the oracle takes: (ncomp, nk, nj, ni),
C-contiguous. ``to_host`` / ``from_host`` convert to and from the reference's
host indexing [i, j, k, comp].
"""

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

from peregrine_amd import grid  # noqa: E402
from peregrine_amd.config import Config  # noqa: E402

# constants only: a calorically perfect air and a thermally perfect one (NASA-7 fit of air)
AIR_CPG = {"Air": {"MW": 28.96, "cp0": 1005.0, "mu0": 1.8591191080521142e-05, "kappa0": 0.02625394405190068}}
# constant transport of the Taylor-Green case at Re 1600 (mu = rho0 V0 L / 1600, L = 1 m, V0 Mach 0.1
# at 300 K and 1 atm; kappa = mu cp / 0.71)
_R_AIR = 8314.46261815324 / 28.96
TGV_MU = (101325.0 / (_R_AIR * 300.0)) * 0.1 * np.sqrt(1005.0 / (1005.0 - _R_AIR) * _R_AIR * 300.0) / 1600.0
TGV_KAPPA = TGV_MU * 1005.0 / 0.71
AIR_NASA7 = {
    # 0.79 N2 + 0.21 O2 by mole: the molar average of the two GRI-Mech 3.0 NASA-7 fits
    "Air": {
        "MW": 28.85064,
        "NASA7": {
            "T-range": [300.0, 1000.0, 3500.0],
            "low": [3.40027067, 0.000483195742, -1.06301196e-06, 2.42372488e-09, -1.2502517e-12, -1029.93907, 3.88890578],
            "high": [3.00137855, 0.00148695006, -6.0826904e-07, 1.23755417e-10, -9.88622096e-15, -957.586304, 5.86979569],
        },
        "mu0": 1.8591191080521142e-05,
        "kappa0": 0.02625394405190068,
    }
}
# two ideal species with different weights and heats: exercises the mixture sums
BINARY_CPG = {
    "A": {"MW": 28.0, "cp0": 1040.0, "mu0": 1.8e-5, "kappa0": 0.026, "lewis": 1.0},
    "B": {"MW": 44.0, "cp0": 846.0, "mu0": 1.5e-5, "kappa0": 0.017, "lewis": 1.3},
}

H2_AIR_Y = {"H2": 0.02852, "O2": 0.22636}  # stoichiometric in air, N2 the rest


def config(gas="air-cpg", recon="constant", primary="kepaec", limiter=None, chem=False, integrator="tvd-rk3", dt=1e-7,
           trans=None, alpha=None):
    """A case config; trans (a transport model) makes it Navier-Stokes with alpha damping."""
    c = Config()
    c["simulation"]["simulator"] = "navier-stokes" if trans else "euler"
    m = c["mixture"]
    m["trans"] = trans
    if alpha is not None:
        c["diff-flux"]["values"] = {"alpha": alpha}
    if gas == "air-cpg":
        m["species"], m["eos"] = AIR_CPG, "cpg"
    elif gas == "air-tpg":
        m["species"], m["eos"], m["T-range"] = AIR_NASA7, "tpg", (300.0, 3500.0)
    elif gas == "binary":
        m["species"], m["eos"] = BINARY_CPG, "cpg"
    elif gas in ("o2n2-srk", "o2n2-pr"):
        # the transcritical pair: library O2 and N2 (JANAF thermo, critical constants), a cubic real gas
        m["species"], m["eos"], m["T-range"] = ["O2", "N2"], "real-gas" if gas == "o2n2-srk" else "peng-robinson", (
            100.0, 1000.0)
    elif gas == "h2o2":
        m["species"], m["eos"], m["T-range"] = "h2o2.yaml", "tpg", (300.0, 3500.0)
        if chem:
            m["chemistry"] = "finite-rate"
            c["chem-integration"]["kind"] = "coupled"
    else:
        raise ValueError(gas)
    a = c["adv-flux"]
    a["recon"], a["primary"], a["limiter"] = recon, primary, limiter
    c["time-integration"]["integrator"] = integrator
    c["simulation"]["dt"] = dt
    c.validate()
    return c


def to_dev(a):
    """host [i, j, k, comp] -> device layout (comp, k, j, i), contiguous"""
    return np.ascontiguousarray(np.asarray(a, float).transpose(3, 2, 1, 0))


def to_host(a):
    return a.transpose(3, 2, 1, 0)


class HostBlock:
    """One periodic cube block's arrays on the host, in the device layout."""

    def __init__(self, ncell, ng, ne, nregs, lengths=(1.0, 1.0, 1.0), jostle=0.0, seed=0, nstate=8):
        self.ncell, self.ng, self.ne = tuple(ncell), ng, ne
        nodes = [np.linspace(0.0, L, n + 1) for L, n in zip(lengths, ncell)]
        inner = np.stack(np.meshgrid(*nodes, indexing="ij"), axis=-1)
        if jostle:
            rng = np.random.default_rng(seed)
            d = np.array([L / n for L, n in zip(lengths, ncell)])
            pert = jostle * d * (rng.random(inner.shape) - 0.5)
            # keep the periodic end planes matched
            for a in range(3):
                pert_a = np.moveaxis(pert, a, 0)
                pert_a[-1] = pert_a[0]
            inner = inner + pert
        self.nodes = grid.halo_nodes(inner, ng, (True, True, True), lengths)
        self.cshape = tuple(n + 2 * ng for n in ncell)
        self.regs = [np.zeros((ne,) + self.cshape[::-1]) for _ in range(nregs)]
        self.state = np.zeros((nstate,) + self.cshape[::-1])
        self.S = [to_dev(grid.areavecs(self.nodes, a)) for a in range(3)]
        self.F = [np.zeros((ne,) + s.shape[1:]) for s in self.S]
        self.detJinv = to_dev(grid.detJinv(self.nodes, ng)[..., None])
        self.xcc = grid.cellcenters(self.nodes)
        # the viscous terms' geometry and arrays (Navier-Stokes)
        self.xccd = to_dev(self.xcc)
        self.Jinv = to_dev(grid.Jinv(self.nodes).reshape(self.cshape + (9,)))
        self.xf = [to_dev(grid.facecenters(self.nodes, a)) for a in range(3)]
        ns = ne - 4
        self.trstate = np.zeros((2 + ns,) + self.cshape[::-1])
        self.hsp = np.zeros((ns,) + self.cshape[::-1])
        self.grads = np.zeros((3 * (ne - 1),) + self.cshape[::-1])

    def interior(self):
        g = self.ng
        return np.s_[:, g:-g, g:-g, g:-g]


def oracle_block(hb, integrator, chemistry, viscous=False, alpha=1.0, walls=()):
    """walls: the faces (0 i-low, 1 i-high, 2 j-low, ...) that are adiabatic slip walls, or a dict
    {face: (kind, (u, v, w), T)} with kind 0 slp-adia .. 5 mov-isot."""
    from oracle import oracle as O

    b = O.Block()
    b.ncell = (O.C.c_int * 3)(*hb.ncell)
    b.ng = hb.ng
    for r, a in enumerate(hb.regs):
        b.regs[r] = O.arr(a)
    b.nregs = len(hb.regs)
    b.state = O.arr(hb.state)
    b.iF, b.jF, b.kF = (O.arr(f) for f in hb.F)
    b.iS, b.jS, b.kS = (O.arr(s) for s in hb.S)
    b.detJinv = O.arr(hb.detJinv)
    b.integrator = O.INTEGRATOR[integrator]
    b.chemistry = int(chemistry)
    b.viscous, b.alpha = int(viscous), alpha
    b.trstate, b.hsp, b.grads = O.arr(hb.trstate), O.arr(hb.hsp), O.arr(hb.grads)
    b.Jinv, b.xcc = O.arr(hb.Jinv), O.arr(hb.xccd)
    for a in range(3):
        b.xf[a] = O.arr(hb.xf[a])
    spec = walls if isinstance(walls, dict) else {f: (0, (0.0, 0.0, 0.0), 0.0) for f in walls}
    for f in range(6):
        kind, wall, T = spec.get(f, (-1, (0.0, 0.0, 0.0), 0.0))
        b.bc[f] = kind + 1
        for d in range(3):
            b.bcwall[f][d] = wall[d]
        b.bcT[f] = T
    b.nstate = hb.state.shape[0]
    return b


def h2_prims(xcc, ncell, seed=20240611, nspots=8, lengths=(0.01, 0.01, 0.01), species=None):
    """The reacting case's seeded start: 1 atm, at rest, stoichiometric H2-air, T 1000 K plus
    Gaussian hot spots (sigma 0.05 L) clipped to 1500 K."""
    rng = np.random.default_rng(seed)
    L = np.array(lengths)
    centres = rng.random((nspots, 3)) * L
    sigma = 0.05 * L[0]
    x = xcc
    T = np.full(x.shape[:3], 1000.0)
    bump = np.zeros(x.shape[:3])
    for c in centres:
        # the nearest periodic image of each centre
        d = x - c
        d -= L * np.round(d / L)
        bump += np.exp(-(d**2).sum(axis=-1) / (2 * sigma**2))
    T = np.minimum(T + 500.0 * bump, 1500.0)
    names = species or ["H2", "H", "O", "O2", "OH", "H2O", "HO2", "H2O2", "N2"]
    q = np.zeros(x.shape[:3] + (4 + len(names),))
    q[..., 0] = 101325.0
    q[..., 4] = T
    for n, name in enumerate(names[:-1]):
        q[..., 5 + n] = H2_AIR_Y.get(name, 0.0)
    return q


def smooth_prims(xcc, ne, seed=3, ns=1):
    """A smooth, fully three-dimensional state: every flux row well away from zero."""
    rng = np.random.default_rng(seed)
    x, y, z = xcc[..., 0], xcc[..., 1], xcc[..., 2]
    ph = rng.uniform(0.0, 2.0 * np.pi, 8)
    tp = 2.0 * np.pi
    q = np.zeros(xcc.shape[:3] + (ne,))
    q[..., 0] = 101325.0 * (1 + 0.2 * np.sin(tp * x + ph[0]) * np.cos(tp * y + ph[1]))
    q[..., 1] = 60.0 * np.cos(tp * x + ph[2]) + 20.0 * np.sin(tp * z + ph[3])
    q[..., 2] = 15.0 * np.sin(tp * y + ph[4]) - 25.0
    q[..., 3] = -10.0 * np.cos(tp * z + ph[5]) + 5.0 * np.sin(tp * x)
    q[..., 4] = 300.0 * (1 + 0.3 * np.cos(tp * x + ph[6]) * np.sin(tp * z + ph[7]))
    for n in range(ne - 5):
        q[..., 5 + n] = (0.5 + 0.3 * np.sin(tp * (x + y) + ph[n % 8])) / max(ns - 1, 1)
    return q


def close(got, want, within, floor=0.0):
    """|got - want| <= within * max|want| per component row (the last axis on host layout,
    the first on device layout: pass component-first arrays)."""
    got, want = np.asarray(got), np.asarray(want)
    scale = np.abs(want).reshape(want.shape[0], -1).max(axis=1)
    scale = np.maximum(scale, floor)
    err = np.abs(got - want).reshape(want.shape[0], -1).max(axis=1)
    return bool((err <= within * scale).all()), err / np.where(scale > 0, scale, 1.0)


def halo_eos(oc, hb, ne):
    """The oracle's halo exchange of Q (self-periodic) and its EOS over the halo slabs."""
    from oracle import oracle as O

    ng = hb.ng
    for a in range(3):
        O.lib.pgo_fill_periodic(O.arr(hb.regs[0]), ne, O.ints(hb.ncell), ng, ng, a)
    for a in range(3):
        for lo in (True, False):
            start, ext = [ng] * 3, list(hb.ncell)
            start[a] = 0 if lo else ng + hb.ncell[a]
            ext[a] = ng
            O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.box(start, ext))


def reacting_flow(x, n, L):
    """Hot spots in H2-air with radicals, plus a swirling velocity and a pressure wave, so every
    conserved row, momentum included, carries more than rounding."""
    q = h2_prims(x, (n,) * 3, lengths=L)
    q[..., 6] = 1e-4
    k = 2 * np.pi / L[0]
    X, Y, Z = x[..., 0] * k, x[..., 1] * k, x[..., 2] * k
    q[..., 0] *= 1.0 + 0.05 * np.sin(X) * np.cos(Z)
    q[..., 1] = 40.0 * np.sin(X) * np.cos(Y)
    q[..., 2] = -40.0 * np.cos(X) * np.sin(Y)
    q[..., 3] = 20.0 * np.sin(Z)
    return q


# --- interface orientations (tests/communication/test_orientation.py:13-60) ------------------

def orientations():
    """The 24 signed axis permutations a right-handed grid can meet another in: digit m names the
    neighbour's axis ours runs along, 1 2 3 forward and 4 5 6 backward."""
    import itertools

    out = []
    for axes in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            P = np.zeros((3, 3))
            for m, (a, s) in enumerate(zip(axes, signs)):
                P[m, a] = s
            if np.linalg.det(P) > 0:
                out.append("".join(str(a + 1 + (3 if s < 0 else 0)) for a, s in zip(axes, signs)))
    assert len(out) == 24
    return out


def signedperm(S):
    return tuple((int(d) - 1) % 3 for d in S), tuple(int(d) > 3 for d in S)


def reorient(a, S):
    """An array (i, j, k, ...) in the first block's layout as a block met with digits S holds it."""
    axes, flips = signedperm(S)
    for m, f in enumerate(flips):
        if f:
            a = np.flip(a, axis=m)
    return np.moveaxis(a, (0, 1, 2), axes)


class ReorientedPair:
    """A cube lattice whose block 1 the partitioner relabels so that block 0 meets it with digits S."""

    def __init__(self, S, **kwargs):
        from peregrine_amd.mesher import CubeMesher

        self.S, self.cube = S, CubeMesher(**kwargs)

    def __getattr__(self, name):
        return getattr(self.cube, name)

    def fill(self, mb):
        from peregrine_amd.partition import Conditioner

        self.cube.fill(mb)
        axes, flips = signedperm(self.S)
        # the partitioner's form: which old axis each new one holds, and which new ones run backwards
        perm = [axes.index(m) for m in range(3)]
        Conditioner().reorient(mb, mb.blocks[1], perm, [flips[perm[m]] for m in range(3)])


def reorient_back(a, S):
    """The inverse of reorient: a block met with digits S's array in the first block's layout."""
    axes, flips = signedperm(S)
    a = np.moveaxis(a, axes, (0, 1, 2))
    for m, f in enumerate(flips):
        if f:
            a = np.flip(a, axis=m)
    return a


FIVE_CPG = {
    # five calorically perfect species: a species count the main library is not compiled for
    "A": {"MW": 28.0, "cp0": 1040.0, "mu0": 1.8e-5, "kappa0": 0.026, "lewis": 1.0},
    "B": {"MW": 44.0, "cp0": 846.0, "mu0": 1.5e-5, "kappa0": 0.017, "lewis": 1.3},
    "C": {"MW": 4.0, "cp0": 5193.0, "mu0": 2.0e-5, "kappa0": 0.15, "lewis": 0.4},
    "D": {"MW": 32.0, "cp0": 918.0, "mu0": 2.0e-5, "kappa0": 0.026, "lewis": 1.1},
    "E": {"MW": 18.0, "cp0": 1864.0, "mu0": 1.0e-5, "kappa0": 0.018, "lewis": 0.9},
}


def trimmed_mechanism(directory, keep=21):
    """The H2/O2 mechanism file with its first `keep` reactions only (the falloff reaction the 21st),
    written into directory: a mechanism whose structure the main library is not compiled for."""
    import yaml

    from peregrine_amd.mixture.mixture import DATA

    with open(DATA / "mechanisms" / "h2o2.yaml") as f:
        doc = yaml.safe_load(f)
    doc["reactions"] = doc["reactions"][:keep]
    path = Path(directory) / f"h2o2-first{keep}.yaml"
    with open(path, "w") as f:
        yaml.safe_dump(doc, f, sort_keys=False)
    return str(path)

