"""Active boxes against computing every box.

A Jacobi sweep skips a box when nothing its backups can read, and none of
its own cells, changed in the previous sweep: min(V, TV) would leave it as it
is. The bar is bit equality with the same solver computing every box, after
every sweep: the FP32 table, the 16-bit mirror where there is one, the
sup-norm and the summed decrease, also across the restarts of the set (an
upload, a kernel change, the polish of a 16-bit store).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPE = (24, 24, 16, 9, 9, 11)


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def forms(pg):
    return {
        "lds": (pg.abi.PG_VSTORE_F32, pg.abi.PG_KERNEL_LDS),
        "fast32": (pg.abi.PG_VSTORE_F32, pg.abi.PG_KERNEL_FAST),
        "u16": (pg.abi.PG_VSTORE_U16_TILE, None),
        "f16": (pg.abi.PG_VSTORE_F16_TILE, None),
    }


def pair(pg, prob, vstore, kernel):
    """Two solvers of one problem: `a` with the active set, `b` without."""
    a = pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ, vstore=vstore)
    b = pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ, vstore=vstore)
    b.set_active(False)
    for s in (a, b):
        if kernel is not None:
            s.set_kernel(kernel)
        s.fill_seed(prob.seeds)
    return a, b


def same_sweep(a, b, it, mirror):
    ra = a.sweep(it, it % 2 == 1)
    rb = b.sweep(it, it % 2 == 1)
    x, y = a.download(), b.download()
    bad = np.nonzero(bits(x) != bits(y))[0]
    assert bad.size == 0, f"sweep {it}: {bad.size} cells differ, first {bad[:5].tolist()}"
    if mirror:
        assert np.array_equal(bits(a.download_mirror()), bits(b.download_mirror())), f"sweep {it}: mirror differs"
    assert ra.linf == rb.linf and ra.sum_decrease == rb.sum_decrease
    assert b.last_sweep_boxes()[0] == b.last_sweep_boxes()[1]
    return ra, a.last_sweep_boxes()


@pytest.mark.parametrize("form", ["lds", "fast32", "u16", "f16"])
def test_active_sweeps_equal_full_sweeps(pg, form):
    """A whole solve to ||dV||_inf < 1e-3 on a reduced problem-2 grid with
    obstacles and 125 controls, compared sweep by sweep."""
    prob = pg.configs.problem(2, shape=SHAPE)
    vstore, kernel = forms(pg)[form]
    a, b = pair(pg, prob, vstore, kernel)
    with a, b:
        done = []
        for it in range(400):
            r, (c, t) = same_sweep(a, b, it, vstore != pg.abi.PG_VSTORE_F32)
            done.append(c)
            if r.linf < 1e-3:
                break
        assert r.linf < 1e-3
        print(f"{form}: {len(done)} sweeps, boxes computed {sum(done) / (len(done) * t):.3f} of all")
        assert done[0] == t            # the first sweep computes every box
        assert min(done[1:]) < t // 2  # and later sweeps skip


def test_restarts(pg):
    """An upload, a switch to the global kernel and back, and the change of
    kernel each restart the set; results stay equal throughout."""
    prob = pg.configs.problem(2, shape=SHAPE)
    a, b = pair(pg, prob, pg.abi.PG_VSTORE_F32, pg.abi.PG_KERNEL_LDS)
    rng = np.random.default_rng(7)
    with a, b:
        it = 0
        for _ in range(40):
            same_sweep(a, b, it, False)
            it += 1
        V = a.download()
        reached = np.nonzero(V < prob.params.cap)[0]
        pick = rng.choice(reached, size=reached.size // 100, replace=False)
        V[pick] *= np.float32(0.9)
        a.upload(V)
        b.upload(V)
        _, (c, t) = same_sweep(a, b, it, False)
        assert c == t
        it += 1
        for _ in range(5):
            same_sweep(a, b, it, False)
            it += 1
        for s in (a, b):
            s.set_kernel(pg.abi.PG_KERNEL_GLOBAL)
        for _ in range(2):
            same_sweep(a, b, it, False)
            it += 1
        for s in (a, b):
            s.set_kernel(pg.abi.PG_KERNEL_LDS)
        _, (c, t) = same_sweep(a, b, it, False)
        assert c == t
        it += 1
        for _ in range(5):
            _, (c, t) = same_sweep(a, b, it, False)
            it += 1
        assert c < t


def test_polish_after_active_16bit(pg):
    """30 sweeps of the 16-bit store, then 3 exact polish sweeps."""
    prob = pg.configs.problem(2, shape=SHAPE)
    a, b = pair(pg, prob, pg.abi.PG_VSTORE_U16_TILE, None)
    with a, b:
        for it in range(30):
            same_sweep(a, b, it, True)
        ra = a.polish(3)
        rb = b.polish(3)
        assert ra == rb
        assert np.array_equal(bits(a.download()), bits(b.download()))
