"""Run the reference's card checker on the CLI fixture's card and record it.

The card is the one tests/test_gpu_cli.py has the device write (three
targets on the fixture field, the reference's default scheme, 30 value and
8 escape sweeps each), written here on the host from the oracle's tables by
the same card writer and MANIFEST/MODEL.JSON writers the CLI uses. The
reference's verifier (wizard/verify_tables.py) then runs on it unchanged, as
a subprocess, in a container that holds the reference tree; it never ships
with the tests. What it printed, its exit status and each target's SHA-256
go to tests/golden/cli/card_verdict.json; tests/test_gpu_cli.py asserts the
device's card has the same digests, so the verdict covers the device card.

    python3 tests/golden/make_card_verdict.py [REF_ROOT]
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from conftest import load_package  # noqa: E402

pg = load_package()
from peregrine_desktop_amd import card, run  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from test_cli import config  # noqa: E402
from test_gpu_cli import ESC_ITERS, ITERS, oracle_tables  # noqa: E402


def main(ref_root: str) -> None:
    orc.lib()
    tmp = Path(tempfile.mkdtemp())
    cfg, _ = config(tmp, iterations=ITERS, tolerance=-1.0, escape_iterations=ESC_ITERS, escape_tolerance=-1.0)
    print("oracle tables ...", flush=True)
    g, sp, tables = oracle_tables(pg, orc, cfg, run)
    m, _, zeroed = run._model_for(cfg)
    bounds, polys, robot, gcfg, gi, g = run._setup(cfg)
    names, states, _ = run.load_targets(cfg["targets"])
    dtype = "u16"
    scale = card.DTYPES[dtype][2]
    chunk = 1 << 23
    escape_scale = card.escape_scale_for(dtype, sp.cap)
    out = Path(cfg["out_dir"])
    tables_dir = out / "TABLES"
    os.makedirs(tables_dir, exist_ok=True)
    entries = []
    for ti0, (st, V) in enumerate(zip(states, tables)):
        info = card.write_table(str(tables_dir), ti0, V, dtype, scale, chunk, sp.cap, escape_scale)
        entries.append(run.card_entry(ti0, names[ti0], st, info, ITERS, 0.0, ESC_ITERS, 0.0))
    run.write_card_files(str(out), cfg, g, gi, bounds, sp, dtype, scale, chunk, True, ESC_ITERS, -1.0, escape_scale,
                         robot, False, entries, cfg["regression"], cfg["field"], m, zeroed)
    checker = Path(ref_root) / "wizard" / "verify_tables.py"
    r = subprocess.run([sys.executable, str(checker), str(out)], capture_output=True, text=True, timeout=600)
    print(r.stdout[-3000:], r.stderr[-2000:], f"exit {r.returncode}", sep="\n")
    verdict = {
        "checker": "wizard/verify_tables.py, unchanged, run as a subprocess on the card",
        "card": f"CLI fixture (tests/golden/cli), default scheme, {ITERS} value and {ESC_ITERS} escape sweeps per "
                "target, tables from the oracle, written by card.write_table and run.write_card_files",
        "exit_status": r.returncode,
        "stdout": r.stdout.splitlines(),
        "stderr": r.stderr.splitlines(),
        "targets": [{"name": e["name"], "sha256": e["sha256"], "n_chunks": e["n_chunks"], "bytes": e["bytes"]}
                    for e in entries],
        "grid_n": list(g.n),
    }
    (HERE / "cli" / "card_verdict.json").write_text(json.dumps(verdict, indent=1) + "\n")
    print(f"wrote {HERE / 'cli' / 'card_verdict.json'}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
