"""Problem description types, mirroring the reference's Julia interface.

`Grid6`, `Model`, `Params`, `control_set` and `target_cells` keep the names,
argument meaning and error behaviour of solver/src/PeregrineSolver.jl and
solver/src/Pipeline.jl, so host code written against the reference reads the
same here. Values are held in the exact float32 bits the reference computes.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

F32 = np.float32


@dataclass(frozen=True)
class Grid6:
    """Uniform 6-D grid over (x, y, h, vx, vy, w), heading periodic.

    Mirrors `Grid6` (PeregrineSolver.jl:47-60): the heading axis spans
    [-pi, pi) with step Float32(2pi/n); every other axis spans lo..hi
    inclusive with step Float32((hi - lo)/(n - 1)), the quotient taken in
    Float64 as Julia does for Real bounds.
    """

    n: tuple
    lo: tuple
    step: tuple

    @staticmethod
    def build(n, lo, hi) -> "Grid6":
        n = tuple(int(v) for v in n)
        if len(n) != 6 or any(v < 2 for v in n):
            raise ValueError(f"every axis needs at least 2 samples, got {n}")
        st = []
        l3 = []
        for k in range(6):
            if k == 2:
                st.append(F32(2.0 * math.pi / n[k]))
                l3.append(F32(-math.pi))
            else:
                st.append(F32((float(hi[k]) - float(lo[k])) / (n[k] - 1)))
                l3.append(F32(lo[k]))
        return Grid6(n, tuple(l3), tuple(st))

    @property
    def ncells(self) -> int:
        return int(np.prod(np.array(self.n, dtype=np.int64)))

    def axisvalue(self, k: int, i) -> np.ndarray:
        """g.lo[k] + Float32(i) * g.step[k] (PeregrineSolver.jl:65), in float32."""
        i = np.asarray(i)
        return F32(self.lo[k]) + i.astype(F32) * F32(self.step[k])

    def flatten(self, i1, i2, i3, i4, i5, i6):
        n = [np.int64(v) for v in self.n]
        return ((((np.int64(i1) * n[1] + i2) * n[2] + i3) * n[3] + i4) * n[4] + i5) * n[5] + i6

    def unflatten(self, idx):
        idx = np.asarray(idx, dtype=np.int64)
        out = []
        r = idx
        for k in range(5, 0, -1):
            out.append(r % self.n[k])
            r = r // self.n[k]
        out.append(r)
        return tuple(reversed(out))

    def hi(self, k: int) -> float:
        return float(self.lo[k]) + float(self.step[k]) * (self.n[k] - 1)

    def c(self) -> abi.PgGrid:
        g = abi.PgGrid()
        for k in range(6):
            g.n[k] = self.n[k]
            g.lo[k] = float(self.lo[k])
            g.step[k] = float(self.step[k])
        return g


@dataclass(frozen=True)
class Model:
    """Body-frame drivetrain regression (PeregrineSolver.jl:100-109).

    a = B*u_sat + A*v + q*w^2 + S*csign(v) + D*|v|v + c, row-major 3x3 blocks.
    knee <= 0 or NaN disables traction saturation.
    """

    B: tuple
    A: tuple
    q: tuple
    S: tuple
    D: tuple
    c: tuple
    eps: tuple = (5.0, 5.0, 0.15)
    knee: float = 0.0

    def __post_init__(self):
        for name, size in (("B", 9), ("A", 9), ("q", 3), ("S", 9), ("D", 9), ("c", 3), ("eps", 3)):
            v = tuple(float(F32(x)) for x in getattr(self, name))
            if len(v) != size:
                raise ValueError(f"Model.{name} needs {size} values, got {len(v)}")
            object.__setattr__(self, name, v)
        object.__setattr__(self, "knee", float(F32(self.knee)))

    def c_struct(self) -> abi.PgModel:
        m = abi.PgModel()
        for name in ("B", "A", "q", "S", "D", "c", "eps"):
            arr = getattr(m, name)
            for i, v in enumerate(getattr(self, name)):
                arr[i] = v
        m.knee = self.knee
        return m

    def theta(self) -> np.ndarray:
        """The 30 differentiable drivetrain parameters: B(9), A(9), q(3),
        diag S(3), diag D(3), c(3)."""
        S = np.array(self.S, dtype=np.float64).reshape(3, 3)
        D = np.array(self.D, dtype=np.float64).reshape(3, 3)
        return np.concatenate([np.array(self.B), np.array(self.A), np.array(self.q),
                               np.diag(S), np.diag(D), np.array(self.c)]).astype(np.float64)

    def with_theta(self, th) -> "Model":
        th = np.asarray(th, dtype=np.float64)
        S = np.array(self.S, dtype=np.float64).reshape(3, 3).copy()
        D = np.array(self.D, dtype=np.float64).reshape(3, 3).copy()
        S[np.diag_indices(3)] = th[21:24]
        D[np.diag_indices(3)] = th[24:27]
        return Model(tuple(th[0:9]), tuple(th[9:18]), tuple(th[18:21]), tuple(S.ravel()),
                     tuple(D.ravel()), tuple(th[27:30]), self.eps, self.knee)


@dataclass(frozen=True)
class Params:
    """Backup parameters (PeregrineSolver.jl:897-927).

    The defaults are the reference constructor's defaults; `Params.plain()`
    is the benchmark's plain scheme (solver/bench/compare.jl:50-53) with the
    substep bound solver_params derives from dt/substeps (Run.jl:276).
    """

    dt: float = 0.05
    nsub: int = 4
    checks: int = 3
    adaptive_checks: bool = True
    nearest: bool = False
    cap: float = 60.0
    ntau: int = 1
    tau_ratio: float = 2.0
    cfl: float = 0.0
    tau_min: float = 0.004
    tau_max: float = 0.5
    hmax: float = 0.0125
    ncoarse: int = 0
    rounds: int = 0
    delta0: float = 0.35
    rk2: bool = True
    vclamp: bool = False
    simplex: bool = True

    @staticmethod
    def plain(dt: float = 0.05, nsub: int = 4, checks: int = 3, cap: float = 60.0) -> "Params":
        dt32 = F32(dt)
        hmax = float(F32(dt32 / F32(max(nsub, 1))))
        return Params(dt=float(dt32), nsub=nsub, checks=checks, adaptive_checks=False,
                      nearest=False, cap=cap, ntau=1, tau_ratio=2.0, cfl=0.0, tau_min=0.004,
                      tau_max=0.5, hmax=hmax, ncoarse=0, rounds=0, delta0=0.35, rk2=False,
                      vclamp=False, simplex=False)

    def c(self) -> abi.PgParams:
        p = abi.PgParams()
        for name, typ in abi.PgParams._fields_:
            v = getattr(self, name)
            setattr(p, name, float(F32(v)) if typ is ctypes.c_float else int(v))
        return p

    def substeps(self) -> int:
        """substeps_for(dt) of the plain scheme (PeregrineSolver.jl:985-989)."""
        if F32(self.hmax) <= 0:
            return int(self.nsub)
        q = F32(self.dt) / F32(self.hmax)
        q = min(q, F32(1.0e4))
        return max(1, min(int(q) + 1, 64))


# ---------------------------------------------------------------------------
# Control sets
# ---------------------------------------------------------------------------

def control_set(level: int) -> np.ndarray:
    """Octahedron-surface controls, coast first (PeregrineSolver.jl:318-346).

    Returns an (nctl, 3) float32 array sorted by (L1 norm, squared L2 norm),
    stable in generation order, as Julia's `sort!` is.
    """
    if level < 1:
        raise ValueError("control level must be >= 1")
    seen = {(F32(0), F32(0), F32(0))}
    out = [(F32(0), F32(0), F32(0))]
    L = level
    for sd in (-1, 1):
        for ss in (-1, 1):
            for st in (-1, 1):
                for i in range(L + 1):
                    for j in range(L - i + 1):
                        k = L - i - j
                        a = F32(i) / F32(L)
                        b = F32(j) / F32(L)
                        c = F32(k) / F32(L)
                        p = (F32(0) if a == 0 else F32(sd) * a,
                             F32(0) if b == 0 else F32(ss) * b,
                             F32(0) if c == 0 else F32(st) * c)
                        if p not in seen:
                            seen.add(p)
                            out.append(p)
    key = [(F32(abs(p[0]) + abs(p[1]) + abs(p[2])), F32(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]))
           for p in out]
    order = sorted(range(len(out)), key=lambda i: key[i])
    return np.array([out[i] for i in order], dtype=F32)


def cube_controls(levels) -> np.ndarray:
    """Cube control set over `levels` per axis: coast first, then the rest in
    lexicographic order with u1 slowest. {-1,0,1} gives 27, {-1,-1/2,0,1/2,1}
    gives 125."""
    lv = [F32(v) for v in levels]
    rest = []
    for a in lv:
        for b in lv:
            for c in lv:
                if a == 0 and b == 0 and c == 0:
                    continue
                rest.append((a, b, c))
    return np.array([(F32(0), F32(0), F32(0))] + rest, dtype=F32)


def ctl_soa(ctl: np.ndarray) -> np.ndarray:
    """(nctl, 3) -> the reference's SoA vector [u1...; u2...; u3...] (Run.jl:283-285)."""
    ctl = np.asarray(ctl, dtype=F32)
    return np.ascontiguousarray(np.concatenate([ctl[:, 0], ctl[:, 1], ctl[:, 2]]))


# ---------------------------------------------------------------------------
# Seeds
# ---------------------------------------------------------------------------

def target_cells(g: Grid6, s, tol) -> np.ndarray:
    """Cells within `tol` of target `s` on every axis, heading wrapped, else
    the single nearest cell (Pipeline.jl:504-536). 0-based, sorted, unique."""
    s = [F32(v) for v in s]
    tol = [F32(v) for v in tol]
    ranges = []
    for k in range(6):
        c = (s[k] - F32(g.lo[k])) / F32(g.step[k])
        if k == 2:
            n3 = F32(g.n[2])
            c = c - np.floor(c / n3) * n3
        r = tol[k] / F32(g.step[k])
        lo = int(math.ceil(float(F32(c - r))))
        hi = int(math.floor(float(F32(c + r))))
        if k != 2:
            lo = max(0, lo)
            hi = min(g.n[k] - 1, hi)
        ranges.append(range(lo, hi + 1))
    idx = []
    import itertools
    for i1, i2, i3, i4, i5, i6 in itertools.product(*ranges):
        idx.append(int(g.flatten(i1, i2, i3 % g.n[2], i4, i5, i6)))
    if not idx:
        near = []
        for k in range(6):
            c = (s[k] - F32(g.lo[k])) / F32(g.step[k])
            r = int(np.rint(np.float64(c)))
            near.append(r % g.n[2] if k == 2 else min(max(r, 0), g.n[k] - 1))
        idx.append(int(g.flatten(*near)))
    return np.unique(np.array(idx, dtype=np.int64))


def disc_target_cells(g: Grid6, s, radius: float, dh: float) -> np.ndarray:
    """Seeds for a disc target: grid nodes within `radius` of (s[0], s[1]) in
    the plane, with wrapped heading error < dh and zero velocity, else the
    reference's nearest-cell fallback. Velocity axes must carry 0 as a node."""
    xs = g.axisvalue(0, np.arange(g.n[0])).astype(np.float64)
    ys = g.axisvalue(1, np.arange(g.n[1])).astype(np.float64)
    hs = g.axisvalue(2, np.arange(g.n[2])).astype(np.float64)
    zero = []
    for k in (3, 4, 5):
        vals = g.axisvalue(k, np.arange(g.n[k]))
        hit = np.nonzero(vals == 0)[0]
        if hit.size != 1:
            raise ValueError(f"axis {k} has no grid node at zero velocity")
        zero.append(int(hit[0]))
    out = []
    for i in range(g.n[0]):
        for j in range(g.n[1]):
            if math.hypot(xs[i] - s[0], ys[j] - s[1]) > radius:
                continue
            for k in range(g.n[2]):
                d = (hs[k] - s[2] + math.pi) % (2 * math.pi) - math.pi
                if abs(d) < dh:
                    out.append(int(g.flatten(i, j, k, *zero)))
    if not out:
        return target_cells(g, s, [0.0] * 6)
    return np.unique(np.array(out, dtype=np.int64))
