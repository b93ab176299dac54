"""Steps replayed from one captured HIP graph against the same steps launched eagerly: to the bit
after every step, with eager work of the program's own (CFL reductions read back on the host,
arrays read back) between the replays, for the reacting case (coupled and sub-stepped chemistry,
fused and per-kernel), Taylor-Green by RK4 (per-kernel), a viscous walled channel and a
rotationally periodic sector."""

import numpy as np
import pytest

import support as sp

pytestmark = pytest.mark.gpu


def _h2o2(graphs, substep=False, fused=True):
    import peregrine_amd as pg

    cfg = sp.config("h2o2", recon="muscl", primary="hllc", limiter="koren", chem=True, dt=1e-8)
    cfg["backend-mi355x"]["fused-stage"] = fused
    if substep:
        ci = cfg["chem-integration"]
        ci["kind"] = "sub-stepped"
        ci["max-substeps"], ci["bulk-substeps"], ci["stiff-bins"], ci["equilibrium-tolerance"] = 50, 2, 8, 1e-6
    cfg["backend-mi355x"]["graphs"] = graphs
    n, L = 16, (0.01,) * 3
    mb = pg.get_solver(cfg, pg.mesher.CubeMesher(blkdims=[n + 1] * 3, lengths=L, periodic=[True] * 3),
                       pg.get_backend("mi355x", cfg))
    mb.set_prims([sp.reacting_flow(mb.blocks[0].xcc.get(), n, L)])
    return mb


def _channel(graphs):
    import peregrine_amd as pg

    cfg = sp.config("binary", recon="muscl", primary="hllc", limiter="koren", trans="constant", dt=2e-6)
    cfg["bcvals"]["wall2"] = {"kind": "no-slp-isot-wall", "T": 350.0}
    cfg["bcvals"]["wall3"] = {"kind": "mov-adia-wall", "u": 20.0, "v": 0.0, "w": 0.0}
    cfg["backend-mi355x"]["graphs"] = graphs
    mesh = pg.mesher.CubeMesher(blkdims=[17, 13, 9], lengths=[0.4, 0.3, 0.2], periodic=[True, False, True],
                                bcnames={3: "wall2", 4: "wall3"})
    mb = pg.get_solver(cfg, mesh, pg.get_backend("mi355x", cfg))
    mb.set_prims([sp.smooth_prims(mb.blocks[0].xcc.get(), 6, ns=2)])
    return mb


def _tgv(graphs):
    """Taylor-Green at 16^3, KEEP + Navier-Stokes, RK4: the per-kernel sequence (RK4 has no fused
    stage)."""
    import peregrine_amd as pg

    cfg = sp.config("air-tpg", recon="constant", primary="kepaec", trans="constant", integrator="rk4", dt=1e-4)
    cfg["backend-mi355x"]["graphs"] = graphs
    n = 16
    mesh = pg.mesher.CubeMesher(blkdims=[n + 1] * 3, lengths=[2 * np.pi] * 3, periodic=[True] * 3)
    mb = pg.get_solver(cfg, mesh, pg.get_backend("mi355x", cfg))
    x = mb.blocks[0].xcc.get()
    q = np.zeros(x.shape[:3] + (5,))
    q[..., 0] = 101325.0 * (1 + 0.01 * np.cos(2 * x[..., 0]) * np.cos(2 * x[..., 1]))
    q[..., 1] = 35.0 * np.sin(x[..., 0]) * np.cos(x[..., 1]) * np.cos(x[..., 2])
    q[..., 2] = -35.0 * np.cos(x[..., 0]) * np.sin(x[..., 1]) * np.cos(x[..., 2])
    q[..., 4] = 300.0
    mb.set_prims([q])
    return mb


def _sector(graphs):
    import peregrine_amd as pg
    from test_gpu_rotperiodic import _swirl

    cfg = sp.config("air-cpg", recon="muscl", primary="hllc", limiter="koren", trans="constant", dt=1e-5)
    cfg["bcvals"]["wall"] = {"kind": "slp-adia-wall"}
    cfg["backend-mi355x"]["graphs"] = graphs
    mesh = pg.mesher.AnnulusMesher(mbdims=[1, 1, 2], blkdims=[9, 7, 5], p1=[0, 0, 0], p2=[1, 0, 0], p3=[0, 0.5, 0],
                                   sweep=90, thickness=0.5, periodic=True, defaultbc="wall")
    mb = pg.get_solver(cfg, mesh, pg.get_backend("mi355x", cfg))
    mb.set_prims([_swirl(blk.xcc.get()) for blk in mb.blocks])
    return mb


CASES = {
    "h2o2-coupled": (lambda g: _h2o2(g), 1e-8),
    "h2o2-per-kernel": (lambda g: _h2o2(g, fused=False), 1e-8),
    "tgv-rk4": (_tgv, 1e-4),
    "h2o2-substepped": (lambda g: _h2o2(g, substep=True), 1e-8),
    "channel": (_channel, 2e-6),
    "sector": (_sector, 1e-5),
}


@pytest.mark.parametrize("case", list(CASES))
def test_graph_replays_are_the_eager_steps(gpu, case):
    make, dt = CASES[case]
    graphed, eager = make(True), make(False)
    for n in range(5):
        graphed.step(dt)
        eager.step(dt)
        # the program's own eager work between replays: reductions read back, arrays read back
        assert graphed.cflmax() == eager.cflmax()
        for a, b in zip(graphed.blocks, eager.blocks):
            assert np.array_equal(a.Q.get(), b.Q.get()), (case, n)
            assert np.array_equal(a.state.get(), b.state.get()), (case, n)
    # the first step follows from_prims (its halos ready) and launches; the rest replay one graph
    assert graphed.replays == 4 and eager.replays == 0
    assert graphed.nrt == eager.nrt == 5


def test_an_uneven_batch_steps_as_its_cases_one_by_one(gpu):
    """Three periodic H2/O2 boxes of different sizes in one solver, against each box stepped by a
    solver of its own, eagerly: to the bit. The batch runs each phase as one launch over the three
    (pga_batch_*), with the autotuner's tiles and replayed from one graph per step, with tiles forced
    to the narrowest candidates, and block at a time with the blocks overlapping on side streams."""
    import peregrine_amd as pg

    dims = [(16, 12, 20), (24, 16, 12), (12, 20, 16)]
    L = (0.01,) * 3

    def solver(dlist, graphs, batched="auto", tiles=None):
        cfg = sp.config("h2o2", recon="muscl", primary="hllc", limiter="koren", chem=True, dt=1e-8)
        cfg["backend-mi355x"]["graphs"] = graphs
        cfg["backend-mi355x"]["batched"] = batched
        mb = pg.get_solver(cfg, pg.mesher.BatchMesher(dlist, lengths=L), pg.get_backend("mi355x", cfg))
        if tiles is not None:
            mb.tiles = tiles
        mb.set_prims([sp.h2_prims(b.xcc.get(), b.ncell, lengths=L) for b in mb.blocks])
        return mb

    narrow = {c: {"stage": (16, 2), "fluxk": 16} for c in ("narrow", "small", "medium", "wide")}
    together = solver(dims, True)
    forced = solver(dims, False, tiles=narrow)
    # block at a time, launched eagerly, the blocks' stages overlapping on side streams
    overlapped = solver(dims, False, batched=False)
    alone = [solver([d], False) for d in dims]
    assert together.batched and forced.batched and not overlapped.batched and not alone[0].batched
    for _ in range(4):
        for mb in (together, forced, overlapped, *alone):
            mb.step(1e-8)
    print("tiles", together.tiles)
    assert together.replays == 3 and overlapped._streams is not None
    assert set(together.tiles) == {"narrow"}
    for blocks in zip(together.blocks, forced.blocks, overlapped.blocks, alone):
        want = blocks[-1].blocks[0]
        for blk in blocks[:-1]:
            assert np.array_equal(blk.Q.get(), want.Q.get())
            assert np.array_equal(blk.state.get(), want.state.get())


def test_a_step_captured_and_replayed_through_the_c_abi(gpu):
    """A reacting step's launches captured on a stream of their own by pga_graph_begin/end and replayed
    three times by pga_graph_launch: to the bit of three steps launched eagerly."""
    import ctypes

    import torch

    import peregrine_amd as pg
    from peregrine_amd import native

    L = (0.01,) * 3

    def solver():
        cfg = sp.config("h2o2", recon="muscl", primary="hllc", limiter="koren", chem=True, dt=1e-8)
        cfg["backend-mi355x"]["graphs"] = False
        mb = pg.get_solver(cfg, pg.mesher.CubeMesher(blkdims=[17] * 3, lengths=L, periodic=[True] * 3),
                           pg.get_backend("mi355x", cfg))
        mb.set_prims([sp.h2_prims(b.xcc.get(), b.ncell, lengths=L) for b in mb.blocks])
        return mb

    eager, replayed = solver(), solver()
    for mb in (eager, replayed):
        mb.step(1e-8)
    lib = native.lib()
    stream = torch.cuda.Stream()
    graph = ctypes.c_void_p()
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        native.check(lib.pga_graph_begin(ctypes.c_void_p(stream.cuda_stream)), "pga_graph_begin")
        replayed._fused_step()
        native.check(lib.pga_graph_end(ctypes.c_void_p(stream.cuda_stream), ctypes.byref(graph)), "pga_graph_end")
    try:
        for _ in range(3):
            native.check(lib.pga_graph_launch(graph, ctypes.c_void_p(stream.cuda_stream)), "pga_graph_launch")
            eager.step(1e-8)
        stream.synchronize()
    finally:
        lib.pga_graph_destroy(graph)
    assert np.array_equal(replayed.blocks[0].Q.get(), eager.blocks[0].Q.get())
    assert np.array_equal(replayed.blocks[0].state.get(), eager.blocks[0].state.get())


def test_allfinite_finds_a_planted_nan(gpu):
    import torch

    from peregrine_amd import native

    lib = native.lib()
    a = torch.ones((5, 6, 7, 9), dtype=torch.float64, device="cuda")
    out = torch.full((2,), 7, dtype=torch.int32, device="cuda")
    view = native.view(a.data_ptr(), 9, 7, 6, 5)
    box = native.box((1, 1, 1), (7, 5, 4))
    native.check(lib.pga_allfinite(view, 5, box, out.data_ptr(), None), "pga_allfinite")
    assert out.tolist() == [1, 7]
    a[3, 2, 3, 4] = float("nan")  # component 3 of cell (4, 3, 2)
    native.check(lib.pga_allfinite(view, 5, box, out.data_ptr(), None), "pga_allfinite")
    assert out.tolist() == [0, 7]
    # outside the box, or past the components checked, it is not seen
    a[3, 2, 3, 4] = 1.0
    a[4, 0, 0, 0] = float("inf")
    native.check(lib.pga_allfinite(view, 4, box, out.data_ptr(), None), "pga_allfinite")
    assert out.tolist() == [1, 7]


@pytest.mark.skip(reason="forking the blocks' side streams inside a captured HIP graph crashed the host process "
                         "(segmentation fault) in the graph's replay on a 2x2x1 lattice of the reacting case; no cause "
                         "is shown, so under graph replay the solver keeps every block on the launch stream")
def test_side_streams_forked_inside_a_captured_graph(gpu):
    """A 2x2x1 lattice of the reacting case, each block's sweeps and stage on one of four side streams
    forked from the launch stream after the halo pass and joined before the next, captured into the
    step's graph and replayed three times: to the bit of the same steps launched eagerly."""
    import peregrine_amd as pg
    from peregrine_amd.solver import Solver

    class Forked(Solver):
        def _block_streams(self):
            torch = self.backend.torch
            if self._streams is None:
                self._streams = [torch.cuda.Stream(device=self.backend.device) for _ in range(4)]
            main = torch.cuda.current_stream(self.backend.device)
            for s in self._streams:
                s.wait_stream(main)
            for n, blk in enumerate(self.blocks):
                yield blk, torch.cuda.stream(self._streams[n % 4])
            for s in self._streams:
                main.wait_stream(s)

    def lattice(graphs):
        cfg = sp.config("h2o2", recon="muscl", primary="hllc", limiter="koren", chem=True, dt=1e-8)
        cfg["backend-mi355x"]["graphs"] = graphs
        cfg["backend-mi355x"]["batched"] = False
        n, L = 16, (0.01,) * 3
        mesh = pg.mesher.CubeMesher(mbdims=(2, 2, 1), blkdims=[n // 2 + 1, n // 2 + 1, n + 1], lengths=L,
                                    periodic=[True] * 3)
        mb = pg.get_solver(cfg, mesh, pg.get_backend("mi355x", cfg), cls=Forked)
        mb.set_prims([sp.reacting_flow(b.xcc.get(), n, L) for b in mb.blocks])
        return mb

    eager, replayed = lattice(False), lattice(True)
    for _ in range(5):
        eager.step(1e-8)
        replayed.step(1e-8)
    assert replayed.replays == 3
    for a, b in zip(eager.blocks, replayed.blocks):
        assert np.array_equal(a.Q.get(), b.Q.get())
