"""The solver plugins in a run on the device (plugins/solver/report.py:20-60,
plugins/solver/nancheck.py:10-44, docs/plugins.md): report prints each due step's time, step size
and CFL numbers (dt times the device reduction) and the cost at the end; nancheck stops a run whose
solution turned non-finite and writes where, a file per block."""

import numpy as np
import pytest

import support as sp

pytestmark = pytest.mark.gpu


def _solver(plugins, niter):
    import peregrine_amd as pg

    cfg = sp.config("air-cpg", recon="muscl", primary="hllc", limiter="koren", dt=1e-6)
    cfg["simulation"]["niter"] = niter
    for name, sect in plugins.items():
        cfg["plugins"][name] = sect
    mesh = pg.mesher.CubeMesher(blkdims=[13] * 3, lengths=[1.0] * 3, periodic=[True] * 3)
    mb = pg.get_solver(cfg, mesh, pg.get_backend("mi355x", cfg))
    mb.set_prims([sp.smooth_prims(mb.blocks[0].xcc.get(), 5)])
    return mb


def test_report_prints_the_steps_cfl_and_the_cost(gpu, capsys):
    mb = _solver({"report": {"niter-out": 2}}, 4)
    capsys.readouterr()
    mb.run()
    out = capsys.readouterr().out
    assert out.count("MAX CFL") == 2 and out.count("nrt: ") == 2 and "nrt: 4" in out
    assert "acoustic" in out and "convective" in out and "Seconds/Iteration/Cell" in out
    # the printed CFL is the step size times the reduction over the cells
    cfl = mb.cflmax()
    assert f"MAX CFL       : {cfl['combined'] * 1e-6:.3f}" in out


def test_nancheck_stops_the_run_and_says_where(gpu, tmp_path, monkeypatch):
    from peregrine_amd.plugins import NanCheckError

    monkeypatch.chdir(tmp_path)
    mb = _solver({"nancheck": {}}, 3)
    ng = mb.ng
    blk = mb.blocks[0]
    Q = blk.Q.get()
    Q[ng + 3, ng + 4, ng + 5, 4] = np.nan
    blk.Q.set(Q)
    with pytest.raises(NanCheckError, match="non-finite state at step 1"):
        mb.run()
    assert mb.nrt == 1
    log = (tmp_path / "nans_0.log").read_text().splitlines()
    assert log[0] == "Non-finite state, block 0" and len(log) > 1
    x, y, z = blk.xcc.get()[ng + 3, ng + 4, ng + 5]
    assert f"x = {x} y = {y} z = {z}" in log


def test_trace_samples_the_primitives_at_the_nearest_cells(gpu, tmp_path, monkeypatch):
    """A point (its own cadence, every step) and a line of 5 points (the section's, every second
    step) on the reacting gas: a header naming the primitives, then a row per point per due step,
    the time, the nearest cell's centre and its p, u, v, w, T and carried Y; a second run appends."""
    import peregrine_amd as pg

    monkeypatch.chdir(tmp_path)
    cfg = sp.config("h2o2", recon="muscl", primary="hllc", limiter="koren", chem=True, dt=1e-8)
    cfg["simulation"]["niter"] = 4
    cfg["plugins"]["trace"] = {"niter-out": 2, "traces": {
        "probe": {"type": "point", "p0": [0.0031, 0.0052, 0.0077], "file": "probe.csv", "niter-out": 1},
        "line": {"type": "line", "p0": [0.0, 0.001, 0.002], "p1": [0.01, 0.009, 0.002], "n": 5, "file": "line.csv"}}}
    n, L = 12, (0.01,) * 3
    mb = pg.get_solver(cfg, pg.mesher.CubeMesher(blkdims=[n + 1] * 3, lengths=L, periodic=[True] * 3),
                       pg.get_backend("mi355x", cfg))
    mb.set_prims([sp.reacting_flow(mb.blocks[0].xcc.get(), n, L)])
    mb.run()
    probe = (tmp_path / "probe.csv").read_text().splitlines()
    line = (tmp_path / "line.csv").read_text().splitlines()
    names = ["p", "u", "v", "w", "T", *mb.sim.mixture.spnames[:-1]]
    assert probe[0] == "time, x, y, z, " + ", ".join(names) == line[0]
    assert len(probe) == 1 + 4 and len(line) == 1 + 2 * 5
    # the last probe row: the cell nearest the point, its primitives now
    ng, blk = mb.ng, mb.blocks[0]
    xcc = blk.xcc.get()[blk.interior]
    d = ((xcc - [0.0031, 0.0052, 0.0077]) ** 2).sum(axis=-1)
    i, j, k = np.unravel_index(d.argmin(), d.shape)
    Q, st = blk.Q.get()[ng + i, ng + j, ng + k], blk.state.get()[ng + i, ng + j, ng + k]
    want = [mb.tme.get(), *xcc[i, j, k], st[0], *(Q[1:4] / Q[0]), st[1], *(Q[5:] / Q[0])]
    got = [float(v) for v in probe[-1].split(",")]
    assert np.allclose(got, want, rtol=1e-8, atol=0)
    # a second run appends to the files it finds
    mb.run()
    assert len((tmp_path / "probe.csv").read_text().splitlines()) == 1 + 8


def _reacting(tmp_path, writer=None):
    import peregrine_amd as pg

    cfg = sp.config("h2o2", recon="muscl", primary="hllc", limiter="koren", chem=True, dt=1e-8)
    if writer is not None:
        cfg["plugins"]["writer"] = writer
    n, L = 12, (0.01,) * 3
    mesh = pg.mesher.CubeMesher(blkdims=[n + 1] * 3, lengths=L, periodic=[True] * 3)
    mb = pg.get_solver(cfg, mesh, pg.get_backend("mi355x", cfg))
    mb.set_prims([sp.reacting_flow(mb.blocks[0].xcc.get(), n, L)])
    return cfg, mesh, mb


def _interior(mb):
    g = mb.ng
    return [blk.Q.get()[g:-g, g:-g, g:-g] for blk in mb.blocks]


def _scaled(a, b):
    return (np.abs(a - b).reshape(-1, a.shape[-1]).max(axis=0) / np.abs(b).reshape(-1, b.shape[-1]).max(axis=0)).max()


def test_a_restart_continues_as_the_case_would_have(gpu, tmp_path):
    """After tests/io/test_restart.py:79-104: three steps, a double-precision result, two more;
    the case started again from the result holds the step, the time and the solution (to the
    primitives' round trip, Q -> p, T -> Q, in which e(T)'s ln T polynomials magnify a last-bit
    change of T about 1e5-fold: 1e-10) and its next two steps are the first case's (1e-9)."""
    import peregrine_amd as pg
    from peregrine_amd.restart import RestartReader, RestartWriter

    cfg, mesh, mb = _reacting(tmp_path)
    for _ in range(3):
        mb.step(1e-8)
    path = RestartWriter(mb, precision="double", cfg=cfg).write(mb, mb.nrt, mb.tme.get(), tmp_path)
    assert path.name == "q.00000003.npz"
    written, tme = _interior(mb), mb.tme.get()
    for _ in range(2):
        mb.step(1e-8)
    expect = _interior(mb)
    reader = RestartReader(path)
    assert reader.nrt == 3 and reader.privars == ["p", "u", "v", "w", "T", *mb.sim.mixture.spnames[:-1]]
    again = pg.get_solver(reader.cfg, mesh, pg.get_backend("mi355x", reader.cfg), state=reader)
    assert again.nrt == 3 and again.tme.get() == tme
    for a, b in zip(_interior(again), written):
        assert _scaled(a, b) < 1e-10
    for _ in range(2):
        again.step(1e-8)
    for a, b in zip(_interior(again), expect):
        assert _scaled(a, b) < 1e-9
    assert again.tme.get() == mb.tme.get()
    # the restarted case's steps after its first replay the captured graph, as the first case's do
    assert again.replays == 1 and mb.replays == 4


def test_the_writer_plugins_result_restarts_the_case(gpu, tmp_path):
    """The plugin writes single precision unless told otherwise, each variable a [k, j, i] array of
    a block's interior; the species sum to one; its result restarts the case to that rounding."""
    import peregrine_amd as pg
    from peregrine_amd.restart import RestartReader

    cfg, mesh, mb = _reacting(tmp_path, writer={"dir": str(tmp_path / "out")})
    for _ in range(3):
        mb.step(1e-8)
    mb.plugins[0](mb)
    written = _interior(mb)
    path = tmp_path / "out" / "q.00000003.npz"
    with np.load(path) as f:
        assert f["results_000000/p"].dtype == np.float32 and f["results_000000/p"].shape == (12, 12, 12)
        total = sum(f[f"results_000000/{s}"].astype(np.float64) for s in mb.sim.mixture.spnames)
        assert np.abs(total - 1.0).max() < 1e-6
    again = pg.get_solver(cfg, mesh, pg.get_backend("mi355x", cfg), state=RestartReader(path))
    assert again.nrt == 3
    for a, b in zip(_interior(again), written):
        assert _scaled(a, b) < 2e-6
