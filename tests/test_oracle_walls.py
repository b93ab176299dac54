"""The oracle's adiabatic slip wall (euler/kernels/bcs/slp-adia-wall.mako, bcs/wall.mako and the
Navier-Stokes equivalents) against what it must do: the mirror halo, no mass, species or energy
through the wall, a fluid at rest stays at rest, and a slip wall carries no shear."""

import ctypes

import numpy as np
import pytest

import support as sp
from oracle import oracle as O
from peregrine_amd.simulator import get_simulator


def _setup(gas="binary", trans=None, recon="muscl", primary="hllc", limiter="koren", ncell=(6, 8, 5),
           lengths=(0.6, 0.8, 0.5), jostle=0.0):
    cfg = sp.config(gas, recon=recon, primary=primary, limiter=limiter, trans=trans)
    sim = get_simulator(cfg)
    oc = O.OracleCase(sim.mixture, recon, limiter, primary)
    hb = sp.HostBlock(ncell, sim.ng, sim.ne, 3, lengths=lengths, jostle=jostle, seed=3)
    return sim, oc, hb


def _fill(oc, hb, q):
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), hb.cshape))


def test_the_mirror_halo_reflects_the_normal_momentum():
    sim, oc, hb = _setup()
    q = sp.smooth_prims(hb.xcc, sim.ne, ns=2)
    _fill(oc, hb, q)
    g = hb.ng
    O.lib.pgo_wall_mirror(oc.ref, 1, 0, 0, O.dbls((0.0, 0.0, 0.0)), 0.0, O.ints(hb.ncell), g, 8, O.arr(hb.regs[0]),
                          O.arr(hb.state), O.arr(hb.S[1]))
    Q = hb.regs[0]
    for h in range(g):
        halo, mirror = g - 1 - h, g + h
        # j is the second array axis from the end: (comp, k, j, i)
        assert np.array_equal(Q[0][g:-g, halo, g:-g], Q[0][g:-g, mirror, g:-g])
        assert np.array_equal(Q[1][g:-g, halo, g:-g], Q[1][g:-g, mirror, g:-g])
        # the unit normal is S / |S|: one rounding from (0, 1, 0)
        assert np.abs(Q[2][g:-g, halo, g:-g] + Q[2][g:-g, mirror, g:-g]).max() <= 4e-16 * np.abs(Q[2]).max()
        assert np.array_equal(Q[4:][:, g:-g, halo, g:-g], Q[4:][:, g:-g, mirror, g:-g])


@pytest.mark.parametrize("viscous", [False, True])
def test_walls_pass_no_mass_species_or_energy(viscous):
    """Walls at both j faces, periodic in i and k: the volume sums of mass, species, energy and the
    i and k momenta change by nothing over a stage (the walls' faces carry only pressure and
    normal stress along j)."""
    sim, oc, hb = _setup(trans="constant" if viscous else None)
    q = sp.smooth_prims(hb.xcc, sim.ne, ns=2)
    _fill(oc, hb, q)
    b = sp.oracle_block(hb, "rk1", False, viscous=viscous, walls=(2, 3))
    Q0 = hb.regs[0].copy()
    O.lib.pgo_stage(oc.ref, ctypes.byref(b), 0, 1e-6)
    vol = 1.0 / hb.detJinv[hb.interior()]
    change = ((hb.regs[0] - Q0)[hb.interior()] * vol).reshape(sim.ne, -1).sum(axis=1)
    scale = (np.abs(Q0[hb.interior()]) * vol).reshape(sim.ne, -1).sum(axis=1)
    for row in (0, 1, 3, 4, 5):
        assert abs(change[row]) <= 1e-14 * scale[row], (row, change[row] / scale[row])
    # the j momentum does change: the walls push
    assert abs(change[2]) > 1e-10 * scale[2]


@pytest.mark.parametrize("viscous", [False, True])
def test_a_fluid_at_rest_between_walls_stays_at_rest(viscous):
    sim, oc, hb = _setup(trans="constant" if viscous else None)
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0], q[..., 4], q[..., 5] = 101325.0, 300.0, 0.3
    _fill(oc, hb, q)
    b = sp.oracle_block(hb, "tvd-rk3", False, viscous=viscous, walls=(0, 1, 2, 3, 4, 5))
    Q0 = hb.regs[0].copy()
    for _ in range(3):
        O.lib.pgo_step(oc.ref, ctypes.byref(b), 1e-5)
    # the flux function's own p S at the cell's inner faces and the wall's p S meet to rounding:
    # the fluid keeps still to a velocity of 1e-12 of the sound speed
    Q, st = hb.regs[0][hb.interior()], hb.state[hb.interior()]
    speed = np.sqrt(Q[1] ** 2 + Q[2] ** 2 + Q[3] ** 2) / Q[0]
    print("largest speed / c", (speed / st[7]).max())
    assert (speed <= 1e-12 * st[7]).all()
    assert np.abs(Q[0] / Q0[hb.interior()][0] - 1).max() < 1e-14


def test_a_slip_wall_carries_no_shear():
    """A shear u(y) along a wall at j: the wall face's flux has no tangential stress; its normal
    momentum flux is the pressure's (no normal velocity)."""
    sim, oc, hb = _setup(trans="constant")
    x = hb.xcc
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0], q[..., 4], q[..., 5] = 101325.0, 300.0, 0.3
    q[..., 1] = 20.0 + 50.0 * x[..., 1]
    _fill(oc, hb, q)
    g = hb.ng
    F = np.zeros_like(hb.F[1])
    O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.trstate), O.arr(hb.hsp),
                        O.box((0, 0, 0), hb.cshape))
    O.lib.pgo_wall_flux(oc.ref, 1, 0, 0, O.dbls((0.0, 0.0, 0.0)), 0.0, O.ints(hb.ncell), g, 1, O.arr(hb.regs[0]),
                        O.arr(hb.state), O.arr(hb.trstate), O.arr(hb.xccd), O.arr(hb.xf[1]), O.arr(hb.S[1]), O.arr(F))
    wall = F[:, g:-g, g, g:-g]
    S = hb.S[1][:, g:-g, g, g:-g]
    p = hb.state[0][g:-g, g, g:-g]
    assert np.array_equal(wall[1], np.zeros_like(wall[1]))
    assert np.abs(wall[2] - p * S[1]).max() <= 1e-12 * np.abs(p * S[1]).max()
    assert np.array_equal(wall[0], np.zeros_like(wall[0])) and np.array_equal(wall[4], np.zeros_like(wall[4]))


# --- no-slip, moving and isothermal walls (navstokes/kernels/bcs/*.mako, euler slp-isot-wall.mako)

WALL = O.dbls((0.0, 0.0, 0.0))


def _mirror(oc, hb, kind, wall=(0.0, 0.0, 0.0), Tw=0.0, axis=1, side=0):
    O.lib.pgo_wall_mirror(oc.ref, axis, side, kind, O.dbls(wall), Tw, O.ints(hb.ncell), hb.ng, hb.state.shape[0],
                          O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.S[axis]))


def _wallflux(oc, hb, kind, wall=(0.0, 0.0, 0.0), Tw=0.0, axis=1, side=0):
    F = np.zeros_like(hb.F[axis])
    O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.trstate), O.arr(hb.hsp),
                        O.box((0, 0, 0), hb.cshape))
    O.lib.pgo_wall_flux(oc.ref, axis, side, kind, O.dbls(wall), Tw, O.ints(hb.ncell), hb.ng, 1, O.arr(hb.regs[0]),
                        O.arr(hb.state), O.arr(hb.trstate), O.arr(hb.xccd), O.arr(hb.xf[axis]), O.arr(hb.S[axis]),
                        O.arr(F))
    return F


@pytest.mark.parametrize("wall", [(0.0, 0.0, 0.0), (12.0, -3.0, 5.0)], ids=["still", "moving"])
def test_a_held_wall_halo_reflects_the_velocity_about_the_walls(wall):
    sim, oc, hb = _setup(trans="constant")
    q = sp.smooth_prims(hb.xcc, sim.ne, ns=2)
    _fill(oc, hb, q)
    g = hb.ng
    st0 = hb.state.copy()
    _mirror(oc, hb, 2 if wall == (0.0, 0.0, 0.0) else 4, wall)
    Q = hb.regs[0]
    for h in range(g):
        halo, mirror = g - 1 - h, g + h
        rho = Q[0][g:-g, mirror, g:-g]
        assert np.array_equal(Q[0][g:-g, halo, g:-g], rho)
        ke, keh = 0.0, 0.0
        for d in range(3):
            u, uh = Q[1 + d][g:-g, mirror, g:-g] / rho, Q[1 + d][g:-g, halo, g:-g] / rho
            # the two cells' mean velocity is the wall's, to rounding
            assert np.abs(0.5 * (u + uh) - wall[d]).max() <= 1e-13 * max(1.0, np.abs(u).max())
            ke, keh = ke + u * u, keh + uh * uh
        # the internal energy carried over: E_h - 0.5 rho |u_h|^2 = E - 0.5 rho |u|^2
        eint = Q[4][g:-g, mirror, g:-g] - 0.5 * rho * ke
        assert np.abs(Q[4][g:-g, halo, g:-g] - 0.5 * rho * keh - eint).max() <= 1e-12 * np.abs(eint).max()
        assert np.array_equal(hb.state[:, g:-g, halo, g:-g], st0[:, g:-g, mirror, g:-g])


@pytest.mark.parametrize("kind", [1, 3], ids=["slp-isot", "no-slp-isot"])
def test_an_isothermal_halo_holds_the_wall_temperature(kind):
    sim, oc, hb = _setup(trans="constant")
    q = sp.smooth_prims(hb.xcc, sim.ne, ns=2)
    _fill(oc, hb, q)
    g = hb.ng
    st0 = hb.state.copy()
    _mirror(oc, hb, kind, Tw=420.0)
    for h in range(g):
        halo, mirror = g - 1 - h, g + h
        assert (hb.state[1][g:-g, halo, g:-g] == 420.0).all()
        assert np.array_equal(hb.state[0][g:-g, halo, g:-g], st0[0][g:-g, mirror, g:-g])
        Y = hb.regs[0][5] / hb.regs[0][0]
        assert np.abs(Y[g:-g, halo, g:-g] - Y[g:-g, mirror, g:-g]).max() <= 1e-15


def test_a_moving_wall_drags_by_the_newtonian_stress():
    """A wall moving at U along x under fluid at rest: the wall face's x-momentum flux is
    -mu |S| U outward / dn and its energy flux the wall's work U times that; no heat."""
    sim, oc, hb = _setup(trans="constant")
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0], q[..., 4], q[..., 5] = 101325.0, 300.0, 0.3
    _fill(oc, hb, q)
    U = 7.5
    F = _wallflux(oc, hb, 4, (U, 0.0, 0.0))
    g = hb.ng
    wall = F[:, g:-g, g, g:-g]
    S = hb.S[1][:, g:-g, g, g:-g]
    mu = hb.trstate[0][g:-g, g, g:-g]
    dn = np.abs(sp.to_host(hb.xccd)[g:-g, g, g:-g, 1] - sp.to_host(hb.xf[1])[g:-g, g, g:-g, 1]).T
    Sabs = np.sqrt((S**2).sum(axis=0))
    want = -mu * Sabs * (-1.0) * U / dn
    assert np.abs(wall[1] - want).max() <= 1e-12 * np.abs(want).max()
    assert np.abs(wall[4] - U * want).max() <= 1e-12 * np.abs(U * want).max()
    assert wall[0].max() == 0.0 and np.abs(wall[3]).max() <= 1e-12 * np.abs(want).max()


def test_an_isothermal_wall_conducts_by_fourier_with_the_walls_kappa():
    """Fluid at rest at 300 K under a 420 K wall: the wall's energy flux is
    -kappa(T_w) |S| outward (T_w - T) / dn, kappa the mixture's at the wall's T."""
    sim, oc, hb = _setup(gas="h2o2", trans="kinetic-theory")
    q = np.zeros(hb.cshape + (sim.ne,))
    # air: O2 carried, N2 the rest
    q[..., 0], q[..., 4], q[..., 5 + sim.mixture.spnames.index("O2")] = 101325.0, 300.0, 0.23
    _fill(oc, hb, q)
    Tw = 420.0
    F = _wallflux(oc, hb, 3, Tw=Tw)
    g = hb.ng
    # kappa at the wall: the transport of the same mixture at p and T_w
    hw = sp.HostBlock(hb.ncell, hb.ng, sim.ne, 1, lengths=(0.6, 0.8, 0.5), nstate=hb.state.shape[0])
    qw = q.copy()
    qw[..., 4] = Tw
    _fill(oc, hw, qw)
    O.lib.pgo_transport(oc.ref, O.arr(hw.regs[0]), O.arr(hw.state), O.arr(hw.trstate), O.arr(hw.hsp),
                        O.box((0, 0, 0), hw.cshape))
    kappa = hw.trstate[1][g:-g, g, g:-g]
    S = hb.S[1][:, g:-g, g, g:-g]
    dn = np.abs(sp.to_host(hb.xccd)[g:-g, g, g:-g, 1] - sp.to_host(hb.xf[1])[g:-g, g, g:-g, 1]).T
    want = -kappa * (-1.0) * (Tw - 300.0) / dn * np.sqrt((S**2).sum(axis=0))
    got = F[4][g:-g, g, g:-g]
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # the hot wall below heats the fluid: energy crosses the low face along +j, into the block
    assert want.min() > 0.0


@pytest.mark.parametrize("kinds", [(2, 2), (3, 3)], ids=["no-slp-adia", "no-slp-isot"])
def test_a_box_of_still_held_walls_keeps_its_mass_and_an_adiabatic_one_its_energy(kinds):
    """Every face a still no-slip wall: mass is conserved through a stage; with adiabatic walls the
    energy too (no work, no heat), with isothermal walls at the fluid's mean T nearly so."""
    sim, oc, hb = _setup(trans="constant")
    q = sp.smooth_prims(hb.xcc, sim.ne, ns=2)
    q[..., 1:4] *= 0.1
    _fill(oc, hb, q)
    Tw = float(q[..., 4].mean())
    walls = {f: (kinds[0], (0.0, 0.0, 0.0), Tw) for f in range(6)}
    b = sp.oracle_block(hb, "rk1", False, viscous=True, walls=walls)
    Q0 = hb.regs[0].copy()
    O.lib.pgo_stage(oc.ref, ctypes.byref(b), 0, 1e-6)
    vol = 1.0 / hb.detJinv[hb.interior()]
    change = ((hb.regs[0] - Q0)[hb.interior()] * vol).reshape(sim.ne, -1).sum(axis=1)
    scale = (np.abs(Q0[hb.interior()]) * vol).reshape(sim.ne, -1).sum(axis=1)
    assert abs(change[0]) <= 1e-14 * scale[0] and abs(change[5]) <= 1e-14 * scale[5]
    if kinds[0] == 2:
        assert abs(change[4]) <= 1e-14 * scale[4]
