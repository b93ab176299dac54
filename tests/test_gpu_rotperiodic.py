"""Rotational periodics on the device, after the reference's own tests
(tests/communication/test_xchg.py:142-213): a 90-degree sector of an annulus joined end to end by
rotational periodics. The momentum that comes across is turned (a constant vector exactly as the
face's rotation turns it, scalars carried as they are), radial and azimuthal fields land at the
halo's own centres (the halo nodes being the far end's, turned), the gradients of a viscous case
turn as vectors and the velocity gradient as a tensor; and the sector steps as the first quarter of
the full annulus, whose quarters meet as interior interfaces."""

import numpy as np
import pytest

import support as sp

pytestmark = pytest.mark.gpu

RECONS = {"constant": ("kepaec", None), "muscl": ("hllc", "koren")}


def _annulus(recon="muscl", trans=None, mbdims=(1, 1, 2), sweep=90.0, blkdims=(4, 4, 4), dt=1e-7, fused=True):
    import peregrine_amd as pg

    primary, limiter = RECONS[recon]
    cfg = sp.config("air-cpg", recon=recon, primary=primary, limiter=limiter, trans=trans, dt=dt)
    cfg["bcvals"]["wall"] = {"kind": "slp-adia-wall"}
    cfg["backend-mi355x"]["fused-stage"] = fused
    mesh = pg.mesher.AnnulusMesher(mbdims=list(mbdims), blkdims=list(blkdims), p1=[0, 0, 0], p2=[1, 0, 0],
                                   p3=[0, 0.5, 0], sweep=sweep, thickness=0.5, periodic=True, defaultbc="wall")
    return pg.get_solver(cfg, mesh, pg.get_backend("mi355x", cfg))


def radial(x):
    y, z = x[..., 1], x[..., 2]
    return np.stack([np.zeros_like(y), y, z], axis=-1) / np.hypot(y, z)[..., None]


def azimuthal(x):
    y, z = x[..., 1], x[..., 2]
    return np.stack([np.zeros_like(y), -z, y], axis=-1) / np.hypot(y, z)[..., None]


@pytest.mark.parametrize("recon", list(RECONS))
def test_a_rotational_periodic_turns_the_momentum(gpu, recon):
    mb = _annulus(recon)
    ng = mb.ng
    a, b = mb.blocks
    for blk in mb.blocks:
        Q = blk.Q.get()
        Q[..., 0], Q[..., 1:4], Q[..., 4] = 1.0, (1.0, 2.0, 3.0), 5.0
        blk.Q.set(Q)
    mb.exchange(0)
    mb.synchronize()
    A, B = a.Q.get(), b.Q.get()
    # block 0's far end comes back through its low face turned by -90 about +x, (y, z) -> (z, -y);
    # block 1's through its high face by +90
    for g in range(ng):
        assert np.allclose(A[ng, ng, g, 1:4], (1.0, 3.0, -2.0), rtol=1e-15, atol=1e-15)
        assert np.allclose(B[ng, ng, -1 - g, 1:4], (1.0, -3.0, 2.0), rtol=1e-15, atol=1e-15)
    assert np.allclose(A[ng, ng, 0, 1:4], a.faces[5].rot @ [1.0, 2.0, 3.0], rtol=1e-15, atol=1e-15)
    assert A[ng, ng, 0, 0] == 1.0 and A[ng, ng, 0, 4] == 5.0
    # the interior interface between the blocks is a plain copy
    assert np.array_equal(A[ng:-ng, ng:-ng, -ng:], B[ng:-ng, ng:-ng, ng:2 * ng])


@pytest.mark.parametrize("path", ["exchange", "halo-pass"])
@pytest.mark.parametrize("field", [radial, azimuthal])
def test_a_rotational_periodic_lands_a_field_at_the_halos_centres(gpu, field, path):
    """A wrong sign or angle in the turn would land the field somewhere other than the halo's own
    centres; the halo pass (the fused stage's Q copy with the EOS over the halo) the same."""
    mb = _annulus("muscl")
    ng = mb.ng
    for blk in mb.blocks:
        Q = blk.Q.get()
        Q[..., 0], Q[..., 4] = 1.0, 5.0
        Q[..., 1:4] = field(blk.xcc.get())
        blk.Q.set(Q)
    if path == "exchange":
        mb.exchange(0)
    else:
        mb._halo_pass(0)
    mb.synchronize()
    checked = 0
    for blk in mb.blocks:
        for nface, f in blk.faces.items():
            if not f.rotperiodic:
                continue
            ks = np.s_[:ng] if f.is_low else np.s_[-ng:]
            halo = blk.Q.get()[ng:-ng, ng:-ng, ks]
            centres = blk.xcc.get()[ng:-ng, ng:-ng, ks]
            assert np.allclose(halo[..., 1:4], field(centres), rtol=0, atol=1e-14), (blk.nblki, nface)
            assert (halo[..., 0] == 1.0).all() and (halo[..., 4] == 5.0).all()
            if path == "halo-pass":
                st = blk.state.get()[ng:-ng, ng:-ng, ks]
                assert np.isfinite(st).all() and (st[..., 1] > 0).all()
            checked += 1
    assert checked == 2


def test_a_rotational_periodic_turns_the_gradients(gpu):
    """The gradient exchange of a viscous case: every row's gradient turns as a vector, the velocity
    gradient (rows u, v, w) as a tensor, R G R^T."""
    mb = _annulus("muscl", trans="constant")
    ng, ngrad = mb.ng, mb.sim.ngrad
    rng = np.random.default_rng(5)
    for blk in mb.blocks:
        blk.grads.set(rng.standard_normal(blk.grads.shape))
    src = [blk.grads.get() for blk in mb.blocks]
    mb._exchange(lambda b: b.grads, 3 * ngrad, 1, 0, 1)
    mb.synchronize()
    a = mb.blocks[0]
    R = a.faces[5].rot
    # block 0's low halo plane is block 1's last interior plane, turned
    got = a.grads.get()[ng:-ng, ng:-ng, ng - 1].reshape(-1, ngrad, 3)
    G = src[1][ng:-ng, ng:-ng, -ng - 1].reshape(-1, ngrad, 3)
    want = G @ R.T
    want[:, :3, :] = np.einsum("ab,nbc,dc->nad", R, G[:, :3, :], R)
    assert np.allclose(got, want, rtol=0, atol=1e-14 * np.abs(G).max())


def _swirl(x):
    """An axisymmetric swirl about the x axis: a quarter turn maps it onto itself."""
    X, y, z = x[..., 0], x[..., 1], x[..., 2]
    r = np.hypot(y, z)
    ur, ut, ux = 5.0 * np.cos(2 * np.pi * X) * (r - 0.5), 60.0 * r, 30.0 + 10.0 * np.sin(2 * np.pi * X)
    q = np.zeros(x.shape[:3] + (5,))
    q[..., 0] = 101325.0 * (1.0 + 0.05 * np.sin(2 * np.pi * X) * r)
    q[..., 1] = ux
    q[..., 2] = (ur * y - ut * z) / r
    q[..., 3] = (ur * z + ut * y) / r
    q[..., 4] = 300.0 * (1.0 + 0.1 * np.cos(2 * np.pi * X) * (1.5 - r))
    return q


@pytest.mark.parametrize("trans", [None, "constant"], ids=["euler", "navier-stokes"])
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-kernel"])
def test_a_sector_steps_as_the_quarter_of_the_full_annulus(gpu, trans, fused):
    """MUSCL-Koren + HLLC, tvd-rk3, 4 steps of an axisymmetric swirl: the 90-degree sector joined to
    itself by rotational periodics against the first of four quarter blocks of the full turn, within
    1e-12 of each component's largest value (the same geometry; the turn rounds)."""
    runs = []
    for mbdims, sweep in (((1, 1, 1), 90.0), ((1, 1, 4), 360.0)):
        mb = _annulus("muscl", trans=trans, mbdims=mbdims, sweep=sweep, blkdims=(9, 7, 7), dt=1e-5, fused=fused)
        assert mb.fused == fused
        mb.set_prims([_swirl(blk.xcc.get()) for blk in mb.blocks])
        for _ in range(4):
            mb.step(1e-5)
        g = mb.ng
        runs.append(mb.blocks[0].Q.get()[g:-g, g:-g, g:-g])
    want, got = runs[1], runs[0]
    scale = np.abs(want).reshape(-1, 5).max(axis=0)
    err = (np.abs(got - want).reshape(-1, 5).max(axis=0) / scale).max()
    print(f"sector vs full turn ({trans or 'euler'}, {'fused' if fused else 'per-kernel'}): {err:.3e}")
    assert err <= 1e-12
