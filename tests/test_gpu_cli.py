"""The solve.jl job interface end to end on the device: `solve` mode on a
small field with a square robot and one obstacle, under the reference's
default scheme (Kuhn simplex lookups, RK2, per-cell CFL horizon with a
5-rung ladder, adaptive probes, a rotating slice of 8 over the level-4
octahedron lattice, 2 rounds of pattern search, the Int8 warm start) and
the escape pass.

Parity: every chunk file equals the oracle's table, solved with the same
sweep count and encoded by the card writer, byte for byte. Contract: the
card passes the checks a card reader makes (TABLE_FORMAT.md sections 2-8):
sizes and digests follow from the manifest, the target cell reads zero,
velocity axes behave as field frame, escape codes never decode as a route,
and the honing block re-derives from MODEL.JSON's own matrices.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from conftest import nthreads, oracle_problem
from test_cli import GOLD, closed_loop, config

pytestmark = pytest.mark.gpu

ITERS, ESC_ITERS = 30, 8


@pytest.fixture(scope="module")
def solved(pg, tmp_path_factory):
    from peregrine_desktop_amd import run
    tmp = tmp_path_factory.mktemp("cli")
    cfg, path = config(tmp, iterations=ITERS, tolerance=-1.0, escape_iterations=ESC_ITERS, escape_tolerance=-1.0)
    print(f"cli solve: {path}", flush=True)
    assert run.main([str(path), "solve"]) == 0
    return cfg, run


def oracle_tables(pg, orc, cfg, run):
    """The oracle's tables for the same config, through the CLI's own host
    set-up (grid, occupancy, parameters, seeds)."""
    from peregrine_desktop_amd.configs import Problem
    m, _, _ = run._model_for(cfg)
    bounds, polys, robot, gcfg, gi, g = run._setup(cfg)
    names, states, _ = run.load_targets(cfg["targets"])
    sp = run.solver_params(cfg, g)
    occ = pg.build_occupancy(g, polys, sp.margin, robot, gi.hsub, sp.clearance, bounds=bounds,
                             wall_clearance_cm=gi.wall_clearance)
    cap = np.float32(sp.cap)
    out = []
    for st in states:
        seeds = pg.target_cells(g, st, sp.ttol)
        op = oracle_problem(orc, Problem(0, g, m, sp.p, sp.ctl, [], occ, seeds))
        V = np.full(g.ncells, cap, np.float32)
        V[seeds] = 0
        wpol = np.zeros(3 * g.ncells, np.int8)
        for it in range(ITERS):
            W = np.empty_like(V)
            if sp.warm:
                op.sweep_warm(V, W, wpol, phase=it, nthreads=nthreads())
            else:
                op.sweep(V, W, phase=it, nthreads=nthreads())
            W[seeds] = 0
            V = W
        for it in range(ESC_ITERS):
            W = np.empty_like(V)
            op.sweep(V, W, phase=it, nthreads=nthreads(), esc=1)
            V = W
        out.append(V)
    return g, sp, out


def test_cli_tables_equal_the_oracle(pg, orc, solved):
    cfg, run = solved
    from peregrine_desktop_amd import card
    man = json.load(open(os.path.join(cfg["out_dir"], "MANIFEST.JSON")))
    g, sp, tables = oracle_tables(pg, orc, cfg, run)
    enc = man["encoding"]
    for t, V in zip(man["targets"], tables):
        assert t["iterations"] == ITERS and t["escape_iterations"] == ESC_ITERS
        want, reached, escaped = card.encode(V, enc["dtype"], enc["scale"], sp.cap, enc["escape_scale"])
        got = b"".join(open(os.path.join(cfg["out_dir"], t["file_pattern"] % c), "rb").read()
                       for c in range(t["n_chunks"]))
        assert got == want, f"target {t['name']}: {sum(a != b for a, b in zip(got, want))} bytes differ"
        assert t["reached_frac"] == reached / V.size and t["escape_frac"] == escaped / V.size
        assert escaped > 0


def test_cli_card_is_the_card_the_reference_checker_passed(solved):
    """tests/golden/cli/card_verdict.json holds the reference's verifier
    (wizard/verify_tables.py, run unchanged on the host-written card of this
    same job) passing, and that card's digests: the device card has them."""
    cfg, run = solved
    verdict = json.loads((GOLD / "card_verdict.json").read_text())
    assert verdict["exit_status"] == 0 and any("PASS" in ln for ln in verdict["stdout"])
    man = json.load(open(os.path.join(cfg["out_dir"], "MANIFEST.JSON")))
    assert [t["sha256"] for t in man["targets"]] == [t["sha256"] for t in verdict["targets"]]
    assert man["grid"]["n"] == verdict["grid_n"]


def test_forced_tiling_writes_the_same_card(pg, solved, tmp_path):
    """out_of_core = "always" with a device budget and tiles of 4 x 4 cells
    over the whole heading axis: the tiled driver (windows in pinned host
    memory) writes every chunk byte for byte as the in-HBM solve did, and
    the manifest records the tiling."""
    cfg, run = solved
    cfg2, path2 = config(tmp_path, iterations=ITERS, tolerance=-1.0, escape_iterations=ESC_ITERS,
                         escape_tolerance=-1.0, out_of_core="always", vram_budget_bytes=1 << 30,
                         tile_cells=[4, 4, 8], tau_max=0.5)
    print(f"cli tiled solve: {path2}", flush=True)
    assert run.main([str(path2), "solve"]) == 0
    man = json.load(open(os.path.join(cfg["out_dir"], "MANIFEST.JSON")))
    man2 = json.load(open(os.path.join(cfg2["out_dir"], "MANIFEST.JSON")))
    assert man["solver"]["driver"] == "whole_grid" and man2["solver"]["driver"] == "tiled"
    til = man2["solver"]["tiling"]
    assert til["tiles_per_round"] > 1 and til["tile_cells"][:2] == [4, 4]
    for t, t2 in zip(man["targets"], man2["targets"]):
        assert t["sha256"] == t2["sha256"] and t["n_chunks"] == t2["n_chunks"]
        for c in range(t["n_chunks"]):
            a = open(os.path.join(cfg["out_dir"], t["file_pattern"] % c), "rb").read()
            b = open(os.path.join(cfg2["out_dir"], t2["file_pattern"] % c), "rb").read()
            assert a == b, f"target {t['name']} chunk {c} differs"


class Card:
    """A card reader over the manifest (TABLE_FORMAT.md sections 3-6)."""

    def __init__(self, root):
        self.root = root
        self.m = json.load(open(os.path.join(root, "MANIFEST.JSON")))
        g, e = self.m["grid"], self.m["encoding"]
        self.n, self.lo = g["n"], g["min"]
        self.step = [(g["max"][k] - g["min"][k]) / (self.n[k] if g["wrap"][k] else self.n[k] - 1) for k in range(6)]
        self.e = e

    def subs(self, s):
        out = []
        for k in range(6):
            f = (s[k] - self.lo[k]) / self.step[k]
            i = int(round(f))
            out.append(i % self.n[k] if self.m["grid"]["wrap"][k] else min(max(i, 0), self.n[k] - 1))
        return out

    def raw(self, t, subs):
        idx = 0
        for k in range(6):
            idx = idx * self.n[k] + subs[k]
        e = self.e
        c, off = idx >> e["chunk_shift"], (idx & (e["chunk_elements"] - 1)) * e["elem_bytes"]
        with open(os.path.join(self.root, self.m["targets"][t]["file_pattern"] % c), "rb") as fh:
            fh.seek(off)
            return int.from_bytes(fh.read(e["elem_bytes"]), "little")

    def cell(self, t, s):
        r, e = self.raw(t, self.subs(s)), self.e
        if r == e["unreachable"] or (e["escape_base"] and r >= e["escape_base"]):
            return math.inf
        return r * e["scale"]

    def escape(self, t, s):
        r, e = self.raw(t, self.subs(s)), self.e
        if e["escape_base"] and e["escape_base"] <= r < e["unreachable"]:
            return e["escape_scale"] * (r - e["escape_base"]) ** 2
        return None


def test_cli_card_contract(solved):
    cfg, run = solved
    root = cfg["out_dir"]
    c = Card(root)
    m = c.m
    total = m["grid"]["total_cells"]
    assert total == math.prod(m["grid"]["n"])
    fb, wc = m["grid"]["field_bounds"], m["solver"]["wall_clearance_cm"]
    inset = [m["grid"]["min"][0] - fb[0], m["grid"]["min"][1] - fb[1], fb[2] - m["grid"]["max"][0],
             fb[3] - m["grid"]["max"][1]]
    assert min(inset) >= 0
    e = m["encoding"]
    assert e["escape_base"] + e["escape_codes"] == e["unreachable"]
    assert e["escape_scale"] * (e["escape_codes"] - 1) ** 2 >= m["solver"]["value_cap_s"] - 1e-6
    for t in m["targets"]:
        assert t["bytes"] == total * e["elem_bytes"] and t["n_chunks"] == -(-total // e["chunk_elements"])
        h = hashlib.sha256()
        for k in range(t["n_chunks"]):
            h.update(open(os.path.join(root, t["file_pattern"] % k), "rb").read())
        assert h.hexdigest() == t["sha256"]
        assert c.cell(t["index"], t["state"]) < 0.25
    # field-frame velocity: moving toward the target beats rest beats moving away
    st = list(m["targets"][0]["state"])
    probe = list(st)
    probe[0] = min(probe[0] + 100.0, m["grid"]["max"][0])
    speed = min(100.0, m["grid"]["max"][3])
    sign = -1.0 if probe[0] > st[0] else 1.0
    vs = {}
    for label, vx in (("toward", sign * speed), ("rest", 0.0), ("away", -sign * speed)):
        s = list(probe)
        s[3], s[4] = vx, 0.0
        vs[label] = c.cell(0, s)
    if all(math.isfinite(v) for v in vs.values()):
        assert vs["toward"] <= vs["rest"] <= vs["away"], vs
    # an escape cell never also decodes as a route
    for d in range(0, 200, 7):
        s = list(st)
        s[0] = min(s[0] + d, m["grid"]["max"][0])
        x = c.escape(0, s)
        if x is not None:
            assert not math.isfinite(c.cell(0, s)) and x >= 0
    # MODEL.JSON: the honing loop re-derived from the card's own matrices
    mj = json.load(open(os.path.join(root, "MODEL.JSON")))
    assert "A_u" in mj["nonzero_blocks"] and mj["control"]["saturation"]["knee"] > 0
    A_u = np.array(mj["A_u"])
    A = np.array(mj["A_s"])[:, 3:]
    S = np.array(mj["A_sgn"])[:, 3:]
    eps = mj["csign"]["coulomb_eps"][3:]
    h = mj["honing"]
    ev = np.sort_complex(np.linalg.eigvals(closed_loop(h, A_u, A, S, eps)))
    want = np.sort_complex(np.array([-p for p in h["poles"]["third"]] + [-h["omega"]] * 6, complex))
    assert np.allclose(ev, want, atol=2e-3 * max(1.0, h["omega"]))
    assert mj["regression_sha256"] == hashlib.sha256(open(GOLD / "regression.toml", "rb").read()).hexdigest()
