#!/usr/bin/env python3
"""Command-line entry point, the drop-in for solver/solve.jl (solve.jl:1-61).

    python3 peregrine-desktop_amd/solve.py <config.json> [plan|solve|honing [run_dir]]
    python3 peregrine-desktop_amd/solve.py --self-test

stdout carries `PLAN {json}`, `PROGRESS {json}`, `HONING {json}` and
`WROTE <path>` lines; stderr carries `ERROR <message>` and up to eight
traceback lines. Exit code 0 is success, 1 an error, 2 a usage error or a
missing config file. The output tree is out_dir/TABLES/TxxCyyyy.BIN,
out_dir/MANIFEST.JSON and out_dir/MODEL.JSON.
"""
import importlib.util
import sys
from pathlib import Path


def _package():
    name = "peregrine_desktop_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg = Path(__file__).resolve().parent
    spec = importlib.util.spec_from_file_location(name, pkg / "__init__.py", submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def self_test() -> int:
    """A small solve on device 0 checked against what any correct solve
    satisfies: seeds hold zero, every value lies in [0, cap], one more sweep
    from the converged table changes it by less than the tolerance, and the
    escape pass leaves every real route untouched while writing escape
    times in [dt, cap) only where there was none."""
    import numpy as np
    pg = _package()
    prob = pg.configs.problem(2, shape=(16, 16, 12, 7, 7, 9))
    cap = np.float32(prob.params.cap)
    fails = []
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        its, _, linf = s.solve(500, -1.0, 1e-3)
        V = s.download()
        if not np.all(V[prob.seeds] == 0):
            fails.append("a seed cell is not zero")
        if not (np.all(V >= 0) and np.all(V <= cap)):
            fails.append("a value lies outside [0, cap]")
        if s.sweep(its).linf >= 1e-3:
            fails.append("the stopped table is not a fixed point to 1e-3")
        V = s.download()
        s.solve_escape(200, -1.0, 1e-3)
        E = s.download()
    routes = V < cap
    if not np.array_equal(E[routes], V[routes]):
        fails.append("the escape pass changed a real route")
    esc = -E[E < 0]
    if esc.size == 0 or not (np.all(esc >= np.float32(prob.params.dt)) and np.all(esc < cap)):
        fails.append("escape times are missing or outside [dt, cap)")
    for f in fails:
        print(f"FAIL {f}", file=sys.stderr)
    print(f"self-test: {its} sweeps, {'PASS' if not fails else 'FAIL'}", flush=True)
    return 0 if not fails else 1


if __name__ == "__main__":
    if sys.argv[1:2] == ["--self-test"]:
        sys.exit(self_test())
    _package()
    from peregrine_desktop_amd import run
    sys.exit(run.main())
