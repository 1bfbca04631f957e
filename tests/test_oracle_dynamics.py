"""Pins of the oracle's drivetrain dynamics against the reference.

The reference's calibration tool restates the solver's model in numpy
(calibration/fit_drivetrain.py) and writes the drivetrain_fit.toml that the
solver's `load_model` reads (solver/src/Pipeline.jl:406-440). Its outputs,
committed under tests/golden/dynamics/ by tests/golden/make_dynamics_vectors.py,
pin what the oracle's own restatement of PeregrineSolver.jl:100-279 cannot
pin by itself:

  * the TOML block layout through `load_model`, and every term of
    a = B.u_sat + A.v + q.w^2 + S.csign(v) + D.|v|v + c at nonzero (u, v, w)
    (`traction_gain` + `eval_terms`, fit_drivetrain.py:559-580,1160-1197);
  * forward Euler with the body-frame w x v terms and the R(h) position
    update, over a 6 s free-drift trajectory (`synth_drive`, :1321-1388).

The free-motion invariants need no fixture: with every model block zero the
robot coasts, so its field-frame velocity and spin are constant and its
position advances by v_field * t. `cell_update` rotates the cell's field
velocity into the body frame (PeregrineSolver.jl:1234-1236) and `lookahead`
rotates the successor's back with the new heading (:1141-1146); over a table
that is affine in position, field velocity and spin the 64-corner
interpolant is exact, so one backup reads off the successor's coordinates.
"""
import ctypes
import json
import math
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden" / "dynamics"


@pytest.fixture(scope="module")
def run(pg):
    from peregrine_desktop_amd import run
    return run


def body_accel(orc, mc, u, v, dtype=np.float64):
    out = np.empty(4, dtype)
    u = np.asarray(u, dtype)
    v = np.asarray(v, dtype)
    f = orc.lib().orc_body_accel_f64 if dtype == np.float64 else orc.lib().orc_body_accel_f32
    f(ctypes.byref(mc), u.ctypes.data, v.ctypes.data, out.ctypes.data)
    return out


def test_golden_toml_loads_into_the_model_blocks(run):
    """load_model keeps the tool's row = response, column = regressor layout."""
    m, _ = run.load_model(str(GOLD / "drivetrain_fit.toml"))
    a = json.loads((GOLD / "drift.json").read_text())["model"]
    assert m.A == tuple(np.float32(np.array(a["A"])).astype(float).ravel())
    assert m.S == tuple(np.float32(np.array(a["S"])).astype(float).ravel())
    assert m.D == tuple(np.float32(np.array(a["D"])).astype(float).ravel())
    assert m.q == tuple(np.float32(a["q"]).astype(float))
    assert m.knee == pytest.approx(0.55, rel=1e-7)


def test_acceleration_matches_the_calibration_tool(run, orc):
    """Every term of the model at 96 seeded (u, v, w), to FP64 rounding."""
    m, _ = run.load_model(str(GOLD / "drivetrain_fit.toml"))
    mc = m.c_struct()
    ref = json.loads((GOLD / "accel.json").read_text())
    U = np.array(ref["u_raw"])
    Vb = np.array(ref["v"])
    want = np.array(ref["accel"])
    gains = np.array(ref["gain"])
    worst = 0.0
    for u, v, a, gn in zip(U, Vb, want, gains):
        got = body_accel(orc, mc, u, v)
        assert got[3] == pytest.approx(gn, rel=1e-12, abs=1e-12)
        scale = np.abs(a).max() + 1.0
        worst = max(worst, float(np.abs(got[:3] - a).max() / scale))
        # the Float32 path evaluates the same terms in float
        g32 = body_accel(orc, mc, u, v, np.float32)
        assert np.abs(g32[:3].astype(np.float64) - a).max() / scale < 5e-6
    assert worst < 1e-12, worst


def test_free_drift_matches_the_calibration_tool(pg, orc):
    """6 s of forward Euler at 1/2000 s (B = 0: no command), 150 Hz samples
    of field position, heading and field velocity, to FP64 rounding."""
    d = json.loads((GOLD / "drift.json").read_text())
    md = d["model"]
    Z = (0.0,) * 9
    m = pg.Model(Z, tuple(np.ravel(md["A"])), tuple(md["q"]), tuple(np.ravel(md["S"])),
                 tuple(np.ravel(md["D"])), tuple(md["c"]), tuple(md["coulomb_eps"]), 0.0)
    mc = m.c_struct()
    rows = np.array(d["rows"])
    fine = float(d["fine_step_s"])
    L = orc.lib()
    s = np.zeros(6)
    nxt = np.empty(6)
    u = np.zeros(3)
    done = 0
    scale = np.abs(rows[:, 1:]).max(axis=0) + 1.0
    worst = np.zeros(5)
    for r in rows:
        while done < int(r[0]):
            L.orc_step_state_f64(ctypes.byref(mc), s.ctypes.data, u.ctypes.data, fine, 1, 0, nxt.ctypes.data)
            s[:] = nxt
            done += 1
        ch, sh = math.cos(s[2]), math.sin(s[2])
        got = np.array([s[0], s[1], s[2], ch * s[3] - sh * s[4], sh * s[3] + ch * s[4]])
        worst = np.maximum(worst, np.abs(got - r[1:]) / scale)
    assert done > 11000 and abs(s[2]) > 3.0      # the drift turned half a revolution
    assert worst.max() < 1e-10, worst


def zero_model(pg):
    Z = (0.0,) * 9
    return pg.Model(Z, Z, (0, 0, 0), Z, Z, (0, 0, 0), (5, 5, 0.15), 0.0)


@pytest.mark.parametrize("rk2", [0, 1])
def test_free_motion_conserves_field_velocity(pg, orc, rk2):
    """No forces: the body velocity turns against the spin, so the field
    velocity stays put (to the integrator's order) and the heading advances
    by w t."""
    mc = zero_model(pg).c_struct()
    L = orc.lib()
    h0, w = 0.7, 2.5
    vf = np.array([80.0, -35.0])
    c, s = math.cos(h0), math.sin(h0)
    vb = np.array([c * vf[0] + s * vf[1], -s * vf[0] + c * vf[1]])
    errs = []
    for nsub in (64, 128):
        s0 = np.array([10.0, 20.0, h0, vb[0], vb[1], w])
        out = np.empty(6)
        L.orc_step_state_f64(ctypes.byref(mc), s0.ctypes.data, np.zeros(3).ctypes.data, 0.4, nsub, rk2,
                             out.ctypes.data)
        assert out[2] == pytest.approx(h0 + w * 0.4, rel=1e-12)
        assert out[5] == w
        ch, sh = math.cos(out[2]), math.sin(out[2])
        vf1 = np.array([ch * out[3] - sh * out[4], sh * out[3] + ch * out[4]])
        errs.append(float(np.abs(vf1 - vf).max()))
        # Euler's rotation (I - dJ) grows |v| by sqrt(1 + d^2) per substep,
        # d = w h: |v| (w t)^2 / (2 nsub) in all; a reversed w x v term
        # would turn the field velocity by 2 w t instead
        assert errs[-1] < 1.2 * np.hypot(*vf) * (w * 0.4) ** 2 / (2 * nsub)
    if not rk2:
        assert errs[1] < 0.6 * errs[0]           # first order: halves with the step


@pytest.fixture(scope="module")
def affine_setup(pg):
    g = pg.Grid6.build((9, 9, 8, 7, 7, 5), (0, 0, -math.pi, -120, -120, -4), (200, 200, math.pi, 120, 120, 4))
    ii = np.indices(g.n).reshape(6, -1)
    coords = [g.axisvalue(k, ii[k]).astype(np.float64) for k in range(6)]
    return g, coords


@pytest.mark.parametrize("iw,nsub", [(2, 4), (3, 64), (1, 64)])
def test_backup_reads_the_free_successor(pg, orc, affine_setup, iw, nsub):
    """One FP64 backup over V = C + s_k returns tau + C + s'_k for the free
    successor s': position + v_field * tau, the same field velocity, the
    same spin. Exact when w = 0; within the Euler error otherwise."""
    g, coords = affine_setup
    mc = zero_model(pg).c_struct()
    dt = 0.05
    p = pg.Params(dt=dt, nsub=nsub, checks=3, adaptive_checks=False, nearest=False, cap=1.0e4, ntau=1,
                  cfl=0.0, hmax=0.0, ncoarse=0, rounds=0, rk2=False, vclamp=False, simplex=False)
    ctl = pg.ctl_soa(np.array([[0, 0, 0], [1, 0, 0]], np.float32))
    occ = np.zeros(g.n[0] * g.n[1] * g.n[2], np.uint8)
    cell = (4, 4, 5, 4, 2, iw)                    # h = pi/4, v_field = (40, -40)
    idx = int(g.flatten(*cell))
    st = [float(coords[k][idx]) for k in range(6)]
    tau = float(np.float32(dt))
    want = {0: st[0] + st[3] * tau, 1: st[1] + st[4] * tau, 3: st[3], 4: st[4], 5: st[5]}
    C = 500.0
    L = orc.lib()
    best = ctypes.c_int32()
    for k, wk in want.items():
        V = (C + coords[k]).astype(np.float64)
        got = L.orc_cell_update_f64(idx, V.ctypes.data, occ.ctypes.data, ctypes.byref(g.c()), ctypes.byref(mc),
                                    ctl.ctypes.data, 2, ctypes.byref(p.c()), 0, 0, ctypes.byref(best))
        err = abs(got - (tau + C + wk))
        tol = 1e-9 if st[5] == 0.0 else (0.02 if k < 3 else 0.05)
        assert err < tol, (k, got - tau - C, wk)
        assert best.value == 0                    # equal successors: coast wins the tie
