"""Write the drivetrain-model golden vectors in tests/golden/dynamics/.

The solver's model, a = B.u_sat + A.v + q.w^2 + S.csign(v) + D.|v|v + c in
the body frame (solver/src/PeregrineSolver.jl:100-171), is also restated in
numpy by the reference's calibration tool, which fits it and writes the
drivetrain_fit.toml the solver loads (calibration/fit_drivetrain.py). This
script runs that tool's own functions -- it needs the reference tree, so it
runs only where that tree is, never at test time -- and commits what they
produce as data:

  * drivetrain_fit.toml  -- a dense synthetic fit (every block nonzero,
    omega^2 on, knee 0.55) serialised by `_pack_basis` + `write_toml`
    (fit_drivetrain.py:971-1058,1132-1157), the file layout the solver's
    `load_model` reads (Pipeline.jl:406-440);
  * accel.json  -- body-frame accelerations at 96 seeded nonzero (u, v, w)
    from `traction_gain` (:559-580) and `eval_terms` (:1160-1197) dotted
    with each row's coefficients;
  * drift.json  -- a free-drift trajectory from `synth_drive` (:1321-1388)
    with B = 0 and no jitter or noise: forward Euler at 1/2000 s through
    A, q, S, D, c, the w x v terms and the R(h) position update, sampled
    at 150 Hz (field-frame position, heading and velocity).

Coefficients, knee and Coulomb widths are float32-representable, so the
Float32 model the solver holds is the same model the float64 numpy code
evaluates. Usage:  python3 tests/golden/make_dynamics_vectors.py [REF_ROOT]
"""
from __future__ import annotations

import json
import sys
from pathlib import Path
from types import SimpleNamespace

import numpy as np

HERE = Path(__file__).resolve().parent
OUT = HERE / "dynamics"


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def fit_results(fd, coef_rows, names):
    """Stand-ins for the tool's FitResult records: only the fields
    `_pack_basis` reads, with the statistics set to neutral values."""
    out = []
    for resp, coef in zip(fd.RESPONSES, coef_rows):
        k = len(names)
        out.append(SimpleNamespace(
            name=resp, names=list(names), coef=np.asarray(coef, dtype=np.float64),
            stderr_hac=np.zeros(k), stderr=np.zeros(k), tstat=np.zeros(k), vif=np.ones(k),
            r2=1.0, r2_adj=1.0, cv_r2=1.0, rmse=0.0, resid_std=0.0, response_std=1.0,
            block_resid_rms=[0.0], max_vif=1.0, n=1000))
    return out


def main(ref_root: str) -> None:
    sys.path.insert(0, str(Path(ref_root) / "calibration"))
    import fit_drivetrain as fd

    eps = tuple(float(e) for e in f32(fd.COULOMB_EPS))
    # The tool reads its module constant at call time; the float32 widths
    # make numpy evaluate exactly the Float32 model the solver loads.
    fd.COULOMB_EPS = eps
    knee = float(f32(0.55))

    names = (list(fd.MECANUM) + list(fd.STATES) + ["omega_sq"]
             + ["sgn_" + s for s in fd.STATES] + ["absv_" + s for s in fd.STATES] + ["const"])
    B = f32([[318.44, 6.12, -4.87], [-5.33, 241.07, 3.94], [0.128, -0.061, 13.706]])
    A = f32([[-2.13, -0.41, -8.72], [0.37, -4.96, 2.3], [0.039, 0.078, -4.67]])
    q = f32([-3.5, 2.1, 0.12])
    S = f32([[-11.6, -3.37, 3.29], [2.07, -6.33, 1.0], [-0.56, -1.17, -0.28]])
    D = f32([[-0.021, 0.003, 0.729], [0.001, -0.049, 0.871], [-0.0004, -0.0006, -0.312]])
    c = f32([0.9, -0.6, 0.05])
    rows = [np.concatenate([B[r], A[r], [q[r]], S[r], D[r], [c[r]]]) for r in range(3)]
    res = fit_results(fd, rows, names)

    OUT.mkdir(parents=True, exist_ok=True)
    fd.write_toml({"units": {"distance": "cm", "angle": "rad"},
                   "mecanum_basis": fd._pack_basis(res, fd.MECANUM, knee)},
                  str(OUT / "drivetrain_fit.toml"))

    rng = np.random.default_rng(20261204)
    n = 96
    u_raw = rng.uniform(-1.0, 1.0, (n, 3))
    u_raw[:8] *= 0.02                      # inside the knee's linear region
    v = np.column_stack([rng.uniform(-160, 160, n), rng.uniform(-160, 160, n), rng.uniform(-5, 5, n)])
    v[8:24] *= 0.02                        # inside the Coulomb ramps
    u_raw, v = f32(u_raw), f32(v)
    gain = fd.traction_gain(u_raw, knee)
    X = fd.eval_terms(u_raw * gain[:, None], v, names, eps=eps)
    acc = np.column_stack([X @ rows[r] for r in range(3)])
    (OUT / "accel.json").write_text(json.dumps({
        "source": "calibration/fit_drivetrain.py traction_gain + eval_terms, coefficient rows of drivetrain_fit.toml",
        "knee": knee, "coulomb_eps": list(eps),
        "u_raw": u_raw.tolist(), "v": v.tolist(), "gain": gain.tolist(), "accel": acc.tolist()}))

    # A constant push in every axis drives the drift out of the Coulomb
    # ramps: ~20 cm/s and ~0.5 rad/s, so w x v and R(h) both matter.
    c_drift = f32([60.0, -45.0, 2.5])
    rows_d = fd.synth_drive(np.zeros((3, 3)), A, q, c_drift, S=S, D=D, dur=6.0, hz=150.0, jitter=0.0,
                            pos_noise=0.0, vel_noise=0.0, seed=7)
    t_ms = rows_d[:, 0]
    step = np.rint(t_ms / 1000.0 * 2000.0).astype(np.int64)
    keep = np.arange(0, len(step), 3)
    (OUT / "drift.json").write_text(json.dumps({
        "source": "calibration/fit_drivetrain.py synth_drive(B=0, jitter=0, pos_noise=0, vel_noise=0, seed=7)",
        "fine_step_s": 1.0 / 2000.0,
        "model": {"A": A.tolist(), "q": q.tolist(), "S": S.tolist(), "D": D.tolist(),
                  "c": c_drift.tolist(), "coulomb_eps": list(eps)},
        "columns": ["fine_steps", "x", "y", "h", "vfx", "vfy"],
        "rows": [[int(step[i])] + [float(x) for x in rows_d[i, 5:10]] for i in keep]}))
    print(f"wrote {OUT}/drivetrain_fit.toml, accel.json ({n} states), drift.json ({keep.size} samples)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
