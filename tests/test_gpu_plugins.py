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
