"""Pins of the host pre-steps that both the HIP path and the oracle consume:
`feasible_bounds` and `build_occupancy` (solver/src/Pipeline.jl:29-136,
148-399). A mistake here would reach both sides of every parity test, so
each check states the geometric property directly:

  * the reference's wall self-test (solver/src/SelfTest.jl:876-963): a
    60 x 20 cm chassis reaches 30 cm along its long axis and 10 across, the
    table spans the union over headings, no free cell hangs the footprint
    through the wall at any angle its heading bin covers, the low edge is
    reachable, a point robot is inset by the clearance alone, and a chassis
    too big for the field is refused;
  * obstacles for a point robot: a cell is blocked iff its node lies inside
    the rectangle or closer than the clearance to it (closed-form distance
    to an axis-aligned box);
  * obstacles for a footprint: a cell is blocked iff at some angle of its
    heading bin the rotated rectangle overlaps the obstacle, decided by the
    separating-axis theorem for convex polygons (a different algorithm from
    the reference's vertex-containment plus edge-crossing test).
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def geo(pg):
    from peregrine_desktop_amd import occupancy, run
    return occupancy, run


CHASSIS = np.array([[-30.0, 30.0, 30.0, -30.0], [-10.0, -10.0, 10.0, 10.0]])


def test_walls_self_test(pg, geo):
    occm, run = geo
    field = (0.0, 0.0, 200.0, 200.0)
    clr = 4.0
    r0 = occm.footprint_reach(CHASSIS, 0.0)
    r90 = occm.footprint_reach(CHASSIS, math.pi / 2)
    assert r0[0] == pytest.approx(30.0, abs=1e-9) and r0[2] == pytest.approx(10.0, abs=1e-9)
    assert r90[0] == pytest.approx(10.0, abs=1e-9) and r90[2] == pytest.approx(30.0, abs=1e-9)

    nh, hsub = 16, 3
    fb = run.feasible_bounds(field, CHASSIS, clr, nh, hsub)
    bins = occm.heading_bin_angles(nh, hsub)
    best = min(max(occm.footprint_reach(CHASSIS, th)[0] for th in angs) for angs in bins)
    assert fb["xlo"] == pytest.approx(field[0] + clr + best, abs=1e-9)
    assert fb["xlo"] < field[0] + clr + 30.0

    g = pg.Grid6.build((41, 41, nh, 3, 3, 3), (fb["xlo"], fb["ylo"], -math.pi, -50, -50, -3),
                       (fb["xhi"], fb["yhi"], math.pi, 50, 50, 3))
    occ = occm.build_occupancy(g, [], 0.0, CHASSIS, hsub, clr, bounds=field, wall_clearance_cm=clr)
    occ = occ.reshape(g.n[0], g.n[1], nh).astype(bool)
    worst = -math.inf
    nfree = 0
    for k in range(nh):
        thc = float(g.axisvalue(2, k))
        half = float(np.float32(g.step[2])) / 2
        angs = [thc - half + 2 * half * s / (hsub - 1) for s in range(hsub)]
        for i in range(g.n[0]):
            px = float(g.axisvalue(0, i))
            for j in range(g.n[1]):
                if occ[i, j, k]:
                    continue
                nfree += 1
                py = float(g.axisvalue(1, j))
                for th in angs:
                    # the rotated corners themselves, not footprint_reach
                    c, s = math.cos(th), math.sin(th)
                    bx = px + c * CHASSIS[0] - s * CHASSIS[1]
                    by = py + s * CHASSIS[0] + c * CHASSIS[1]
                    worst = max(worst, field[0] + clr - bx.min(), bx.max() - (field[2] - clr),
                                field[1] + clr - by.min(), by.max() - (field[3] - clr))
    assert nfree > 0
    assert worst <= 1.0e-3, worst
    assert not occ[0].all()                      # the low x edge is free at some heading

    fbp = run.feasible_bounds(field, None, clr, nh, hsub)
    assert fbp["xlo"] == pytest.approx(field[0] + clr, abs=1e-9)
    assert fbp["xhi"] == pytest.approx(field[2] - clr, abs=1e-9)
    with pytest.raises(run.SolveError):
        run.feasible_bounds((0.0, 0.0, 40.0, 40.0), CHASSIS, clr, nh, hsub)


def box_distance(px, py, x0, x1, y0, y1):
    dx = np.maximum(np.maximum(x0 - px, 0.0), px - x1)
    dy = np.maximum(np.maximum(y0 - py, 0.0), py - y1)
    return np.hypot(dx, dy)


@pytest.mark.parametrize("clearance", [0.0, 7.5])
def test_point_robot_obstacles_are_the_rectangle_and_its_clearance(pg, geo, clearance):
    occm, _ = geo
    g = pg.Grid6.build((37, 29, 4, 2, 2, 2), (0, 0, -math.pi, -1, -1, -1), (180, 140, math.pi, 1, 1, 1))
    rects = [(61.3, 47.9, 33.1, 21.7), (128.2, 96.6, 18.3, 40.9)]
    polys = [occm.rect(*r) for r in rects]
    occ = occm.build_occupancy(g, polys, 0.0, None, 3, clearance).reshape(g.n[0], g.n[1], g.n[2]).astype(bool)
    xs = g.axisvalue(0, np.arange(g.n[0])).astype(np.float64)
    ys = g.axisvalue(1, np.arange(g.n[1])).astype(np.float64)
    PX, PY = np.meshgrid(xs, ys, indexing="ij")
    want = np.zeros(PX.shape, bool)
    for cx, cy, w, h in rects:
        d = box_distance(PX, PY, cx - w / 2, cx + w / 2, cy - h / 2, cy + h / 2)
        inside = (np.abs(PX - cx) < w / 2) & (np.abs(PY - cy) < h / 2)
        want |= inside | (d < clearance)
    for k in range(g.n[2]):                      # a point has no heading
        assert np.array_equal(occ[:, :, k], want)
    assert want.sum() > 20


def sat_overlap(P, Q):
    """Separating-axis test for two convex polygons (2xN, any orientation):
    they overlap iff no edge normal of either separates their projections."""
    for poly in (P, Q):
        n = poly.shape[1]
        for i in range(n):
            ex = poly[0, (i + 1) % n] - poly[0, i]
            ey = poly[1, (i + 1) % n] - poly[1, i]
            ax, ay = -ey, ex
            pp = ax * P[0] + ay * P[1]
            qq = ax * Q[0] + ay * Q[1]
            if pp.max() < qq.min() or qq.max() < pp.min():
                return False
    return True


def test_footprint_obstacles_match_separating_axes(pg, geo):
    occm, _ = geo
    nh, hsub = 8, 3
    g = pg.Grid6.build((23, 19, nh, 2, 2, 2), (0, 0, -math.pi, -1, -1, -1), (220, 180, math.pi, 1, 1, 1))
    robot = np.array([[-21.0, 17.0, 17.0, -21.0], [-9.0, -9.0, 9.0, 9.0]])
    obst = occm.rect(104.7, 91.3, 38.9, 25.3)
    occ = occm.build_occupancy(g, [obst], 0.0, robot, hsub, 0.0).reshape(g.n[0], g.n[1], nh).astype(bool)
    nblocked = 0
    mism = []
    for k in range(nh):
        thc = float(g.axisvalue(2, k))
        half = float(np.float32(g.step[2])) / 2
        angs = [thc - half + 2 * half * s / (hsub - 1) for s in range(hsub)]
        for i in range(g.n[0]):
            px = float(g.axisvalue(0, i))
            for j in range(g.n[1]):
                py = float(g.axisvalue(1, j))
                hit = False
                for th in angs:
                    c, s = math.cos(th), math.sin(th)
                    body = np.array([px + c * robot[0] - s * robot[1], py + s * robot[0] + c * robot[1]])
                    hit |= sat_overlap(body, obst)
                nblocked += hit
                if hit != occ[i, j, k]:
                    mism.append((i, j, k))
    assert nblocked > 100
    assert not mism, mism[:10]
