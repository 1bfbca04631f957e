"""Adjoint dL/dtheta on the device against the CPU FP64 adjoint and against
central finite differences of L itself.

L(theta) is the mean FP64 V over seeded start states (x, y, heading uniform,
velocity and spin zero). The device and the CPU oracle run the same adjoint
in FP64 on the same table, so they agree to summation order (atomics on the
device): the bar is 1e-8 relative. Finite differences perturb theta, solve
the FP64 table again to 1e-12 s and difference L; the step is 1e-3 of the
parameter (the model is stored in float, so the step actually taken is the
difference of the rounded values): the bar is 1e-3 relative.
"""
import numpy as np
import pytest

from conftest import nthreads, oracle_problem

pytestmark = pytest.mark.gpu


def fp64_L(pg, prob, model, starts, tol=1e-12):
    with pg.HipSolver(prob.grid, model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        its, l = s.solve_f64(50000, tol)
        assert l < tol, (its, l)
        L, grad, it = s.adjoint(starts)
        return L, grad, it, s.download_f64()


def test_device_adjoint_matches_cpu_fp64_adjoint(pg, orc):
    prob = pg.configs.problem(5, shape=(8, 8, 6, 5, 5, 5))
    starts = pg.configs.adjoint_starts(prob, 512)
    L, grad, it, V = fp64_L(pg, prob, prob.model, starts)
    assert it > 0
    op = oracle_problem(orc, prob)
    g_cpu, L_cpu, it_cpu, _ = op.adjoint_f64(V, prob.seeds, starts, nthreads=nthreads())
    assert it_cpu > 0
    assert L == pytest.approx(L_cpu, rel=1e-12)
    scale = np.max(np.abs(g_cpu))
    assert np.max(np.abs(grad - g_cpu)) <= 1e-8 * scale, (grad, g_cpu)


NAMES = ([f"B{r}{c}" for r in (1, 2, 3) for c in (1, 2, 3)] + [f"A{r}{c}" for r in (1, 2, 3) for c in (1, 2, 3)]
         + ["q1", "q2", "q3", "S11", "S22", "S33", "D11", "D22", "D33", "c1", "c2", "c3"])


def test_device_adjoint_against_central_differences(pg):
    """All 30 parameters, h = 1e-3 of the parameter (1e-5 for a zero one)."""
    prob = pg.configs.problem(5, shape=(12, 12, 8, 5, 5, 7))
    starts = pg.configs.adjoint_starts(prob, 1024)
    L, grad, it, _ = fp64_L(pg, prob, prob.model, starts)
    th = prob.model.theta()
    for j in range(30):
        h = 1e-3 * max(abs(th[j]), 1e-2)
        tp, tm = th.copy(), th.copy()
        tp[j] += h
        tm[j] -= h
        # the step the float model actually takes
        h_eff = float(np.float32(tp[j])) - float(np.float32(tm[j]))
        Lp = fp64_L(pg, prob, prob.model.with_theta(tp), starts)[0]
        Lm = fp64_L(pg, prob, prob.model.with_theta(tm), starts)[0]
        fd = (Lp - Lm) / h_eff
        print(f"theta[{j}] {NAMES[j]} {th[j]:+.4g}: adjoint {grad[j]:+.6e} fd {fd:+.6e}", flush=True)
        assert abs(grad[j] - fd) <= 1e-3 * max(abs(fd), 1e-3 * np.max(np.abs(grad))), (j, grad[j], fd)


def test_autograd_function(pg):
    import torch
    from importlib import import_module
    ag = import_module(pg.__name__ + ".autograd")
    prob = pg.configs.problem(5, shape=(8, 8, 6, 5, 5, 5))
    starts = pg.configs.adjoint_starts(prob, 256)
    theta = torch.tensor(prob.model.theta(), dtype=torch.float64, device="cuda", requires_grad=True)
    L = ag.value_loss(theta, prob, starts)
    (2.0 * L).backward()
    L_ref, g_ref, _, _ = fp64_L(pg, prob, prob.model, starts)
    assert float(L) == pytest.approx(L_ref, rel=1e-12)
    assert np.allclose(theta.grad.cpu().numpy(), 2.0 * g_ref, rtol=1e-10, atol=0)
