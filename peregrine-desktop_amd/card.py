"""The SD-card table writer: a solved V table -> chunked on-card files.

Restates the reference's encoding and output step (solver/src/Pipeline.jl:
620-840, docs/TABLE_FORMAT.md sections 1-7): the four on-card dtypes with the
unreachable sentinel and the square-law escape band, 8.3-named chunk files of
2^chunk_shift values, and the SHA-256 of the logical table in index order.
Vectorised with numpy over one chunk at a time, so a table is written without
holding its encoding whole.
"""
from __future__ import annotations

import hashlib
import json
import math
import os
from dataclasses import dataclass
from datetime import datetime, timezone

import numpy as np

# dtype -> (bytes per value, sentinel, default scale, first escape code);
# Pipeline.jl:620-625. Float types carry the escape in the sign and have no
# band; their sentinel is NaN.
DTYPES = {
    "u8": (1, 0xFF, 0.025, 0x00E0),
    "u16": (2, 0xFFFF, 0.001, 0xFC00),
    "f16": (2, None, 1.0, 0),
    "f32": (4, None, 1.0, 0),
}


def escape_codes(dtype: str) -> int:
    """Raw codes of the escape band of an integer dtype (Pipeline.jl:628-630)."""
    nbytes, sentinel, _, base = DTYPES[dtype]
    return 0 if base == 0 else sentinel - base


def escape_scale_for(dtype: str, cap: float) -> float:
    """Escape seconds per code squared (Pipeline.jl:651-653): the band spans
    0..cap in escape_codes - 1 squared steps."""
    n = escape_codes(dtype)
    return 0.0 if n <= 1 else float(cap) / float(n - 1) ** 2


def encode(V: np.ndarray, dtype: str, scale: float, cap: float, escape_scale: float = 0.0):
    """Seconds -> on-card bytes, and the counts of reached and escaped cells
    (Pipeline.jl:685-775). A finite negative value is an escape time (-v);
    a value >= cap, infinite or NaN is unreachable; a real value that would
    land in the escape band saturates to the sentinel."""
    V = np.asarray(V, dtype=np.float32)
    cap32 = np.float32(cap)
    esc = (V < 0) & np.isfinite(V)
    unr = ~esc & ((V >= cap32) | np.isinf(V) | np.isnan(V))
    reached = int(np.count_nonzero(~esc & ~unr))
    escaped = int(np.count_nonzero(esc))
    nbytes, sentinel, _, base = DTYPES[dtype]
    if dtype == "f32":
        out = V.copy()
        out[unr] = np.float32(np.nan)
        return out.astype("<f4").tobytes(), reached, escaped
    if dtype == "f16":
        out = V.astype(np.float16)
        out[unr] = np.float16(np.nan)
        return out.astype("<f2").tobytes(), reached, escaped
    hasband = escape_scale > 0.0 and base != 0
    ebase = base if hasband else sentinel
    ecodes = max(escape_codes(dtype), 1)
    einv = 1.0 / escape_scale if escape_scale > 0.0 else 0.0
    out = np.empty(V.shape, dtype=np.uint32)
    # real routes: round(v / scale) to nearest, ties to even; saturate at the band
    vd = V.astype(np.float64)
    r = np.rint(vd / scale)
    real = ~esc & ~unr
    rr = np.where(real, r, 0).astype(np.uint64)
    out[:] = np.where(rr >= ebase, sentinel, rr).astype(np.uint32)
    out[unr] = sentinel
    if escaped:
        e = -vd[esc]
        code = np.rint(np.sqrt(np.maximum(e, 0.0) * einv)).astype(np.uint64)
        out[esc] = (ebase + np.minimum(code, ecodes - 1)).astype(np.uint32)
    typ = "<u1" if dtype == "u8" else "<u2"
    return out.astype(typ).tobytes(), reached, escaped


def decode(raw: np.ndarray, dtype: str, scale: float, escape_scale: float = 0.0):
    """On-card raw values -> (seconds, escape seconds), as a reader does
    (TABLE_FORMAT.md section 6): unreachable and escape cells are +inf
    seconds; escape seconds are +inf except on escape cells."""
    nbytes, sentinel, _, base = DTYPES[dtype]
    if dtype in ("f16", "f32"):
        x = np.asarray(raw).astype(np.float64)
        secs = np.where(np.isnan(x) | (x < 0), np.inf, x)
        esc = np.where((x < 0) & ~np.isnan(x), -x, np.inf)
        return secs, esc
    x = np.asarray(raw).astype(np.int64)
    hasband = escape_scale > 0.0 and base != 0
    unr = (x == sentinel) | (hasband & (x >= base))
    secs = np.where(unr, np.inf, x * scale)
    esc = np.where(hasband & (x >= base) & (x != sentinel), escape_scale * (x - base).astype(np.float64) ** 2,
                   np.inf)
    return secs, esc


def chunk_of(idx: int, chunk_shift: int, elem_bytes: int):
    """Flat index -> (chunk number, byte offset in that chunk) (section 5)."""
    return idx >> chunk_shift, (idx & ((1 << chunk_shift) - 1)) * elem_bytes


def chunk_name(target: int, chunk: int) -> str:
    return f"T{target:02d}C{chunk:04d}.BIN"


@dataclass
class TableInfo:
    n_chunks: int
    bytes: int
    sha256: str
    reached: int
    escaped: int
    cells: int


def write_table(tables_dir: str, target: int, V: np.ndarray, dtype: str, scale: float, chunk_elements: int,
                cap: float, escape_scale: float = 0.0) -> TableInfo:
    """One target's table as chunk files (Pipeline.jl:783-822); the SHA-256
    covers the concatenated logical table, independent of the chunking."""
    if chunk_elements <= 0 or chunk_elements & (chunk_elements - 1):
        raise ValueError("chunk_elements must be a power of two")
    os.makedirs(tables_dir, exist_ok=True)
    V = np.asarray(V, dtype=np.float32).reshape(-1)
    cells = V.size
    nchunks = -(-cells // chunk_elements)
    if nchunks > 10000:
        raise ValueError(f"{nchunks} chunks exceed the 8.3 name's 4 digits")
    h = hashlib.sha256()
    reached = escaped = total = 0
    for c in range(nchunks):
        part, r, e = encode(V[c * chunk_elements:(c + 1) * chunk_elements], dtype, scale, cap, escape_scale)
        with open(os.path.join(tables_dir, chunk_name(target, c)), "wb") as fh:
            fh.write(part)
        h.update(part)
        reached += r
        escaped += e
        total += len(part)
    return TableInfo(nchunks, total, h.hexdigest(), reached, escaped, cells)


def manifest(grid, targets: list, infos: list, dtype: str, scale: float, chunk_elements: int, cap: float,
             escape_scale: float, field_bounds=None, solver: dict | None = None, model_file: str = "MODEL.JSON",
             generator: str = "peregrine-desktop_amd") -> dict:
    """MANIFEST.JSON (TABLE_FORMAT.md section 2). `targets` holds
    (name, state[6]) pairs; `infos` the matching TableInfo."""
    n = [int(v) for v in grid.n]
    lo = [float(v) for v in grid.lo]
    hi = [float(np.float32(grid.lo[k]) + np.float32(grid.step[k]) * np.float32(n[k] - 1)) for k in range(6)]
    # the heading axis is periodic: its span is the full turn
    hi[2] = float(np.float32(grid.lo[2]) + np.float32(grid.step[2]) * np.float32(n[2]))
    nbytes = DTYPES[dtype][0]
    sentinel = DTYPES[dtype][1]
    ebase = DTYPES[dtype][3] if escape_scale > 0 else 0
    total = math.prod(n)
    out = {
        "schema_version": 1,
        "generated_utc": datetime.now(timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"),
        "generator": generator,
        "model_file": model_file,
        "grid": {
            "axes": ["x", "y", "h", "vx", "vy", "w"],
            "n": n, "min": lo, "max": hi,
            "wrap": [False, False, True, False, False, False],
            "units": ["cm", "cm", "rad", "cm/s", "cm/s", "rad/s"],
            "frame": "field",
            "total_cells": total,
            "index_formula": "((((ix*Ny+iy)*Nh+ih)*Nvx+ivx)*Nvy+ivy)*Nw+iw",
        },
        "encoding": {
            "dtype": dtype, "elem_bytes": nbytes, "scale": scale, "unit": "seconds",
            "unreachable": sentinel if sentinel is not None else "NaN",
            "escape_base": ebase, "escape_codes": escape_codes(dtype) if ebase else 0,
            "escape_scale": escape_scale, "escape_unit": "seconds",
            "byte_order": "little", "order": "row_major_c",
            "chunk_elements": chunk_elements, "chunk_shift": int(chunk_elements).bit_length() - 1,
        },
        "targets": [],
    }
    if field_bounds is not None:
        out["grid"]["field_bounds"] = [float(v) for v in field_bounds]
    for i, ((name, state), info) in enumerate(zip(targets, infos)):
        unreached = info.cells - info.reached
        out["targets"].append({
            "index": i, "name": name, "state": [float(v) for v in state],
            "file_pattern": f"TABLES/T{i:02d}C%04d.BIN",
            "n_chunks": info.n_chunks, "bytes": info.bytes, "sha256": info.sha256,
            "reached_frac": info.reached / info.cells,
            "escape_frac": info.escaped / info.cells,
            "escape_of_unreached_frac": (info.escaped / unreached) if unreached else 0.0,
        })
    if solver:
        out["solver"] = solver
    return out


def write_manifest(root: str, m: dict) -> None:
    with open(os.path.join(root, "MANIFEST.JSON"), "w") as fh:
        json.dump(m, fh, indent=2)
