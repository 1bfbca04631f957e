"""Configuration-space occupancy on the (x, y, h) grid.

Host pre-step of the solve, mirroring `build_occupancy` and its geometry
helpers (solver/src/Pipeline.jl:29-136, 148-399): per heading bin, the union
over `substeps` angles spanning the bin of "the footprint at that pose is
within `clearance_cm` of an obstacle or outside the wall box", then an
optional per-slice dilation by `margin_cm`. Result is one byte per cell in
the kernel's (ix*Ny + iy)*Nh + ih order.
"""
from __future__ import annotations

import math

import numpy as np

from .model import Grid6


def inpoly(px: float, py: float, poly: np.ndarray) -> bool:
    """Ray-casting point in polygon (Pipeline.jl:29-43). poly is 2xN."""
    n = poly.shape[1]
    inside = False
    j = n - 1
    eps = np.finfo(np.float64).eps
    for i in range(n):
        xi, yi = poly[0, i], poly[1, i]
        xj, yj = poly[0, j], poly[1, j]
        if ((yi > py) != (yj > py)) and (px < (xj - xi) * (py - yi) / (yj - yi + eps) + xi):
            inside = not inside
        j = i
    return inside


def inpoly_grid(px: np.ndarray, py: np.ndarray, poly: np.ndarray) -> np.ndarray:
    """Vectorised `inpoly` over arrays of points (identical arithmetic)."""
    n = poly.shape[1]
    inside = np.zeros(np.broadcast(px, py).shape, dtype=bool)
    j = n - 1
    eps = np.finfo(np.float64).eps
    for i in range(n):
        xi, yi = poly[0, i], poly[1, i]
        xj, yj = poly[0, j], poly[1, j]
        crosses = (yi > py) != (yj > py)
        xint = (xj - xi) * (py - yi) / (yj - yi + eps) + xi
        inside ^= crosses & (px < xint)
        j = i
    return inside


def seg_cross(p1, p2, p3, p4) -> bool:
    def d(a, b, c):
        return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])
    d1 = d(p3, p4, p1); d2 = d(p3, p4, p2)
    d3 = d(p1, p2, p3); d4 = d(p1, p2, p4)
    return ((d1 > 0) != (d2 > 0)) and ((d3 > 0) != (d4 > 0))


def polys_overlap(A: np.ndarray, B: np.ndarray) -> bool:
    na, nb = A.shape[1], B.shape[1]
    for i in range(na):
        if inpoly(A[0, i], A[1, i], B):
            return True
    for j in range(nb):
        if inpoly(B[0, j], B[1, j], A):
            return True
    for i in range(na):
        a1 = (A[0, i], A[1, i]); a2 = (A[0, (i + 1) % na], A[1, (i + 1) % na])
        for j in range(nb):
            b1 = (B[0, j], B[1, j]); b2 = (B[0, (j + 1) % nb], B[1, (j + 1) % nb])
            if seg_cross(a1, a2, b1, b2):
                return True
    return False


def pt_seg_dist(p, a, b) -> float:
    abx = b[0] - a[0]; aby = b[1] - a[1]
    L2 = abx * abx + aby * aby
    t = 0.0 if L2 <= 0 else min(max(((p[0] - a[0]) * abx + (p[1] - a[1]) * aby) / L2, 0.0), 1.0)
    dx = p[0] - (a[0] + t * abx); dy = p[1] - (a[1] + t * aby)
    return math.sqrt(dx * dx + dy * dy)


def seg_seg_dist(a1, a2, b1, b2) -> float:
    if seg_cross(a1, a2, b1, b2):
        return 0.0
    return min(pt_seg_dist(a1, b1, b2), pt_seg_dist(a2, b1, b2),
               pt_seg_dist(b1, a1, a2), pt_seg_dist(b2, a1, a2))


def polys_too_close(A: np.ndarray, B: np.ndarray, d: float) -> bool:
    if polys_overlap(A, B):
        return True
    if d <= 0:
        return False
    na, nb = A.shape[1], B.shape[1]
    for i in range(na):
        a1 = (A[0, i], A[1, i]); a2 = (A[0, (i + 1) % na], A[1, (i + 1) % na])
        for j in range(nb):
            b1 = (B[0, j], B[1, j]); b2 = (B[0, (j + 1) % nb], B[1, (j + 1) % nb])
            if seg_seg_dist(a1, a2, b1, b2) < d:
                return True
    return False


def point_too_close(px: float, py: float, B: np.ndarray, d: float) -> bool:
    if inpoly(px, py, B):
        return True
    if d <= 0:
        return False
    nb = B.shape[1]
    for j in range(nb):
        b1 = (B[0, j], B[1, j]); b2 = (B[0, (j + 1) % nb], B[1, (j + 1) % nb])
        if pt_seg_dist((px, py), b1, b2) < d:
            return True
    return False


def heading_bin_angles(nh: int, substeps: int):
    step = 2 * math.pi / nh
    half = step / 2
    out = []
    for k in range(nh):
        c = -math.pi + k * step
        if substeps <= 1:
            out.append([c])
        else:
            out.append([c - half + 2 * half * (s - 1) / (substeps - 1) for s in range(1, substeps + 1)])
    return out


def footprint_reach(robot, theta: float):
    if robot is None or robot.shape[1] < 3:
        return (0.0, 0.0, 0.0, 0.0)
    c, s = math.cos(theta), math.sin(theta)
    xlo = xhi = ylo = yhi = -math.inf
    for i in range(robot.shape[1]):
        rx = c * robot[0, i] - s * robot[1, i]
        ry = s * robot[0, i] + c * robot[1, i]
        xlo = max(xlo, -rx); xhi = max(xhi, rx)
        ylo = max(ylo, -ry); yhi = max(yhi, ry)
    return (xlo, xhi, ylo, yhi)


def wall_box(bounds, robot, clearance: float, theta: float):
    rxlo, rxhi, rylo, ryhi = footprint_reach(robot, theta)
    d = float(clearance)
    return (bounds[0] + d + rxlo, bounds[2] - d - rxhi, bounds[1] + d + rylo, bounds[3] - d - ryhi)


def build_occupancy(g: Grid6, polys, margin_cm: float = 0.0, robot=None, substeps: int = 3,
                    clearance_cm: float = 0.0, bounds=None, wall_clearance_cm=None) -> np.ndarray:
    """Occupancy bytes, (ix*Ny + iy)*Nh + ih order (Pipeline.jl:287-399)."""
    if wall_clearance_cm is None:
        wall_clearance_cm = clearance_cm
    nx, ny, nh = g.n[0], g.n[1], g.n[2]
    occ = np.zeros((nx, ny, nh), dtype=bool)
    point_robot = robot is None or robot.shape[1] < 3
    rad = 0.0 if point_robot else float(np.max(np.sqrt(robot[0] ** 2 + robot[1] ** 2)))
    tol = 1.0e-3
    xs = g.axisvalue(0, np.arange(nx)).astype(np.float64)
    ys = g.axisvalue(1, np.arange(ny)).astype(np.float64)
    for k in range(nh):
        thc = float(g.axisvalue(2, k))
        half = float(np.float32(g.step[2])) / 2
        if substeps <= 1:
            angles = [thc]
        else:
            angles = [thc - half + 2 * half * (s - 1) / (substeps - 1) for s in range(1, substeps + 1)]
        if bounds is not None:
            for th in angles:
                bxlo, bxhi, bylo, byhi = wall_box(bounds, robot, wall_clearance_cm, th)
                occ[(xs < bxlo - tol) | (xs > bxhi + tol), :, k] = True
                occ[:, (ys < bylo - tol) | (ys > byhi + tol), k] = True
        for th in angles:
            Rr = None
            if not point_robot:
                R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
                Rr = R @ robot
            for P in polys:
                pad = rad + margin_cm + clearance_cm
                x0 = P[0].min() - pad; x1 = P[0].max() + pad
                y0 = P[1].min() - pad; y1 = P[1].max() + pad
                i0 = max(0, math.floor((x0 - float(g.lo[0])) / float(g.step[0])))
                i1 = min(nx - 1, math.ceil((x1 - float(g.lo[0])) / float(g.step[0])))
                j0 = max(0, math.floor((y0 - float(g.lo[1])) / float(g.step[1])))
                j1 = min(ny - 1, math.ceil((y1 - float(g.lo[1])) / float(g.step[1])))
                if i1 < i0 or j1 < j0:
                    continue
                if point_robot and clearance_cm <= 0:
                    PX, PY = np.meshgrid(xs[i0:i1 + 1], ys[j0:j1 + 1], indexing="ij")
                    occ[i0:i1 + 1, j0:j1 + 1, k] |= inpoly_grid(PX, PY, P)
                    continue
                for i in range(i0, i1 + 1):
                    for j in range(j0, j1 + 1):
                        if occ[i, j, k]:
                            continue
                        px, py = xs[i], ys[j]
                        if point_robot:
                            occ[i, j, k] = point_too_close(px, py, P, clearance_cm)
                        else:
                            body = Rr + np.array([[px], [py]])
                            occ[i, j, k] = polys_too_close(body, P, clearance_cm)
    if margin_cm > 0:
        rx = max(0, math.ceil(margin_cm / float(g.step[0])))
        ry = max(0, math.ceil(margin_cm / float(g.step[1])))
        if rx > 0 or ry > 0:
            src = occ.copy()
            for k in range(nh):
                ii, jj = np.nonzero(src[:, :, k])
                for di in range(-rx, rx + 1):
                    for dj in range(-ry, ry + 1):
                        a = ii + di; b = jj + dj
                        ok = (a >= 0) & (a < nx) & (b >= 0) & (b < ny)
                        occ[a[ok], b[ok], k] = True
    return np.ascontiguousarray(occ.reshape(-1).astype(np.uint8))


def rect(cx: float, cy: float, w: float, h: float) -> np.ndarray:
    """Axis-aligned rectangle as a 2x4 polygon (counter-clockwise)."""
    return np.array([[cx - w / 2, cx + w / 2, cx + w / 2, cx - w / 2],
                     [cy - h / 2, cy - h / 2, cy + h / 2, cy + h / 2]], dtype=np.float64)
