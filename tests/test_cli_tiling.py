"""The job interface's tiling keys and plan report on the host
(solver/src/Run.jl:316-404, 650-781; solver/src/Tiles.jl; Advise.jl:219-339):
`out_of_core`, `vram_budget_bytes` and `vram_headroom` decide the driver;
PLAN's out_of_core block carries what the wizard reads (fits, store_bytes,
the tile plan, or the advice when nothing fits); the reach scan and tile
plan meet what they mean; `tau_max: "auto"` is the reference horizon in core.
"""
import json
import math

import numpy as np
import pytest

from test_cli import cli, config


@pytest.fixture(scope="module")
def tl(pg):
    from peregrine_desktop_amd import run, tiling
    return run, tiling


def setup(run, tmp_path, **kw):
    cfg, path = config(tmp_path, **kw)
    m, _, _ = run._model_for(cfg)
    g = run._setup(cfg)[-1]
    sp = run.solver_params(cfg, g)
    return cfg, path, m, g, sp


def test_decompose_honours_the_out_of_core_keys(tl, tmp_path):
    run, tiling = tl
    cfg, _, m, g, sp = setup(run, tmp_path)
    whole = g.ncells * 11                      # Jacobi pair + warm-start store
    asked = []

    def budget(headroom):
        asked.append(headroom)
        return int(20 * whole * headroom)

    def dec(**kw):
        c = dict(cfg, **kw)
        return tiling.decompose(c, g, m, sp.p, sp.ctl, sp.nctl, sp.warm, budget)

    assert dec()["mode"] == "incore" and asked[-1] == 0.85
    assert dec(vram_headroom=0.5)["budget"] == 10 * whole and asked[-1] == 0.5
    assert dec(vram_headroom=0.04)["mode"] == "ooc"
    for on in (True, "true", "always"):
        assert dec(out_of_core=on)["mode"] == "ooc"
    n = len(asked)
    d = dec(vram_budget_bytes=whole // 2)
    assert d["mode"] == "ooc" and d["budget"] == whole // 2 and len(asked) == n   # a pinned budget asks nothing
    for off in (False, "false", "never"):
        assert dec(vram_budget_bytes=whole // 2, out_of_core=off)["mode"] == "incore"
    assert tiling.decompose(cfg, g, m, sp.p, sp.ctl, sp.nctl, sp.warm, lambda h: None)["mode"] == "cpu"


def test_plan_reports_the_tiling(tl, tmp_path):
    run, tiling = tl
    cfg, path, m, g, sp = setup(run, tmp_path, vram_budget_bytes=600_000)
    r = cli(path, "plan")
    assert r.returncode == 0, r.stderr
    p = json.loads(r.stdout.split("PLAN ", 1)[1])
    ooc = p["out_of_core"]
    assert ooc["mode"] == "ooc" and ooc["fits"] is True and ooc["budget_bytes"] == 600_000
    for k in ("halo_cells", "tile_cells", "loaded_cells", "tiles_per_round", "resident_bytes", "amplification",
              "reach_cm", "reach_cells", "store_bytes", "io_bytes_per_round"):
        assert k in ooc, k
    assert ooc["store_bytes"] == g.ncells * 7
    assert ooc["resident_bytes"] <= 600_000 and ooc["amplification"] >= 1.0
    # in the wizard's arithmetic the store joins the card bytes
    assert p["bytes_total"] + ooc["store_bytes"] > p["bytes_total"]
    rec = p["recommend"]
    assert rec["tau_auto"] and rec["tau_applied"] == rec["tau_max"] and rec["tau_options"]
    assert all(o["tau"] <= 0.5 for o in rec["tau_options"])


def test_plan_advises_when_nothing_fits(tl, tmp_path):
    run, _ = tl
    cfg, path, m, g, sp = setup(run, tmp_path, vram_budget_bytes=200_000)
    r = cli(path, "plan")
    assert r.returncode == 0, r.stderr
    ooc = json.loads(r.stdout.split("PLAN ", 1)[1])["out_of_core"]
    assert ooc["mode"] == "ooc" and ooc["fits"] is False
    col = g.ncells // (g.n[0] * g.n[1])
    assert ooc["column_bytes"] == col * 7
    # the largest halo whose (1 + 2h)^2 columns fit the budget
    h = ooc["max_halo_cells"]
    assert (1 + 2 * h) ** 2 * col * 7 <= 200_000 < (3 + 2 * h) ** 2 * col * 7
    assert ooc["min_window_cells"] == 1 + 2 * max(ooc["halo_cells"])
    assert isinstance(ooc["warm_start_would_help"], bool)


def test_solve_takes_the_tiled_driver_when_asked(tl, tmp_path):
    """out_of_core = true selects the tiled backend before any device work
    (here there is no device, so the solve then fails loudly)."""
    run, _ = tl
    cfg, path, m, g, sp = setup(run, tmp_path, out_of_core=True, vram_budget_bytes=1 << 30)
    r = cli(path, "solve")
    lines = [json.loads(ln.split(" ", 1)[1]) for ln in r.stdout.splitlines() if ln.startswith("PROGRESS ")]
    be = [ln for ln in lines if ln.get("phase") == "backend"]
    assert be and be[0]["backend"] == "hip_tiled"


def test_reach_of_a_double_integrator(pg, tl):
    """Without drag the farthest a plain-scheme backup reaches is a coast from
    the velocity box's corner: sqrt(2) vmax dt (any push that lengthens the
    step ends outside the box and is rejected). The scan samples commands and
    directions, so it lands just inside that supremum."""
    _, tiling = tl
    Z = (0.0,) * 9
    m = pg.Model((100.0, 0, 0, 0, 100.0, 0, 0, 0, 5.0), Z, (0, 0, 0), Z, Z, (0, 0, 0), (5, 5, 0.15), 0.0)
    g = pg.Grid6.build((20, 20, 8, 7, 7, 5), (0, 0, -math.pi, -120, -120, -2), (300, 300, math.pi, 120, 120, 2))
    p = pg.Params.plain(dt=0.1, nsub=4)
    dxy, dh = tiling.reach_extent(g, m, p, pg.control_set(2), nangle=16, nsample=64)
    assert dxy == pytest.approx(math.sqrt(2) * 120 * 0.1, rel=1e-2)
    assert dh == pytest.approx(2.0 * 0.1, rel=1e-2)        # a coast at the largest spin, likewise
    hx, hy = tiling.halo_cells(g, dxy, margin=2)
    assert hx == math.ceil(dxy / float(g.step[0])) + 4


def test_plan_tiles_is_the_largest_tile_that_fits(pg, tl):
    _, tiling = tl
    g = pg.Grid6.build((60, 40, 8, 5, 5, 5), (0, 0, -math.pi, -1, -1, -1), (100, 100, math.pi, 1, 1, 1))
    budget = 3_000_000
    tp = tiling.plan_tiles(g, budget, 3, 3, warm=True)
    assert tp["bytes"] <= budget
    b_next = tiling.tile_bytes(g, min(tp["wx"] + 2, 60), min(tp["wy"] + 2, 40), 3, 3, True)[0]
    assert b_next > budget
    assert tp["ntiles"] == -(-60 // tp["wx"]) * -(-40 // tp["wy"])
    assert tiling.plan_tiles(g, 1000, 3, 3) is None


def test_tau_in_core_is_the_reference(pg, tl):
    _, tiling = tl
    rec = tiling.recommend_tau("incore", None, None, pg.Params(), None, 0, 0, True, 2, 16, 128, 7.0, 1, {}, 1.0)
    assert rec["tau_max"] == 0.5 and rec["options"] == [] and rec["cost"] == 0.0
