"""The on-card table writer against docs/TABLE_FORMAT.md's worked example and
the encoding rules of Pipeline.jl:620-840 (CPU only)."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import load_package

pg = load_package()
from importlib import import_module  # noqa: E402

card = import_module(pg.__name__ + ".card")


def test_worked_example_index_chunk_offset():
    """TABLE_FORMAT.md section 7: n = [46,46,24,13,13,13], the subscripts
    (11, 25, 16, 8, 5, 8) fold to 28,035,145, which is chunk 3, byte offset
    5,738,642 of T00C0003.BIN at chunk_shift 23 and 2-byte values."""
    n = [46, 46, 24, 13, 13, 13]
    sub = [11, 25, 16, 8, 5, 8]
    idx = 0
    for k in range(6):
        idx = idx * n[k] + sub[k]
    assert idx == 28035145
    assert card.chunk_of(idx, 23, 2) == (3, 5738642)
    assert card.chunk_name(0, 3) == "T00C0003.BIN"


def test_worked_example_decode():
    """12345 is 12.345 s; 64612 is unreachable with a 0.57 s escape; 65535 is
    unreachable with no way out (escape_scale 5.7448e-05)."""
    secs, esc = card.decode(np.array([12345, 64612, 65535], np.uint16), "u16", 0.001, 5.7448e-05)
    assert secs[0] == pytest.approx(12.345)
    assert np.isinf(secs[1]) and np.isinf(secs[2])
    assert esc[1] == pytest.approx(0.57448, rel=1e-4)
    assert np.isinf(esc[0]) and np.isinf(esc[2])


def test_escape_scale_spans_the_cap():
    """The manifest's escape_scale for u16 and a 60 s cap: 60 / 1022^2."""
    assert card.escape_codes("u16") == 1023
    assert card.escape_codes("u8") == 31
    assert card.escape_codes("f16") == 0
    assert card.escape_scale_for("u16", 60.0) == pytest.approx(5.7448e-05, rel=1e-4)
    assert card.escape_scale_for("f32", 60.0) == 0.0


@pytest.mark.parametrize("dtype", ["u8", "u16", "f16", "f32"])
def test_encode_rules(dtype):
    cap = 60.0
    es = card.escape_scale_for(dtype, cap)
    V = np.array([0.0, 1.2345, 5.0, 59.999, 60.0, 61.0, np.inf, np.nan, -0.25, -3.0], np.float32)
    raw, reached, escaped = card.encode(V, dtype, card.DTYPES[dtype][2], cap, es)
    # reached counts real routes before saturation (Pipeline.jl:751-758)
    assert reached == 4 and escaped == 2
    nb = card.DTYPES[dtype][0]
    typ = {"u8": "<u1", "u16": "<u2", "f16": "<f2", "f32": "<f4"}[dtype]
    x = np.frombuffer(raw, dtype=typ)
    assert x.size * nb == len(raw) == V.size * nb
    secs, esc = card.decode(x, dtype, card.DTYPES[dtype][2], es)
    scale = card.DTYPES[dtype][2]
    if dtype in ("u8", "u16"):
        # real routes: round to the nearest code; u8 saturates 5 s and up at
        # the band (code 224 = 5.6 s): 59.999 s is past it and saturates
        assert secs[0] == 0.0
        # the float value is divided in double, as Julia's Float32 / Float64 does
        assert secs[1] == pytest.approx(np.rint(float(np.float32(1.2345)) / scale) * scale)
        if dtype == "u16":
            assert secs[2] == pytest.approx(5.0) and secs[3] == pytest.approx(59.999)
        else:
            assert np.isinf(secs[3])
        # escape cells: unreachable, with the square-law time
        assert np.isinf(secs[8]) and np.isinf(secs[9])
        code = np.rint(np.sqrt(0.25 / es))
        assert esc[8] == pytest.approx(es * code ** 2)
    else:
        assert secs[1] == pytest.approx(1.2345, rel=1e-3)
        assert esc[8] == pytest.approx(0.25, rel=1e-3) and esc[9] == pytest.approx(3.0, rel=1e-3)
    assert np.all(np.isinf(secs[4:8]))
    assert np.all(np.isinf(esc[:8]))


def test_u16_ties_round_to_even():
    """round(UInt32, v / scale) rounds half to even: 0.0005 s / 0.001 -> 0
    and 0.0015 / 0.001 -> 2 when the quotient is exactly half-way."""
    V = np.array([0.5, 1.5, 2.5], np.float32)
    raw, _, _ = card.encode(V, "u16", 1.0, 60.0, 0.0)
    assert list(np.frombuffer(raw, "<u2")) == [0, 2, 2]


def test_write_table_chunks_and_sha(tmp_path):
    rng = np.random.default_rng(0)
    V = rng.uniform(0, 70, 1000).astype(np.float32)
    V[::7] = -rng.uniform(0, 10, V[::7].size)
    es = card.escape_scale_for("u16", 60.0)
    info = card.write_table(str(tmp_path / "TABLES"), 2, V, "u16", 0.001, 256, 60.0, es)
    assert info.n_chunks == 4 and info.bytes == 2000 and info.cells == 1000
    files = sorted(os.listdir(tmp_path / "TABLES"))
    assert files == [f"T02C{c:04d}.BIN" for c in range(4)]
    data = b"".join((tmp_path / "TABLES" / f).read_bytes() for f in files)
    whole, reached, escaped = card.encode(V, "u16", 0.001, 60.0, es)
    assert data == whole
    assert info.sha256 == hashlib.sha256(whole).hexdigest()
    assert (reached, escaped) == (info.reached, info.escaped)
    # the hash is of the logical stream: a different chunking gives the same
    other = card.write_table(str(tmp_path / "T2"), 2, V, "u16", 0.001, 64, 60.0, es)
    assert other.sha256 == info.sha256 and other.n_chunks == 16


def test_manifest_fields(tmp_path):
    prob = pg.configs.problem(1, shape=(8, 8, 6, 5, 5, 5))
    V = np.full(prob.grid.ncells, 60.0, np.float32)
    V[prob.seeds] = 0.0
    es = card.escape_scale_for("u16", 60.0)
    info = card.write_table(str(tmp_path / "TABLES"), 0, V, "u16", 0.001, 1 << 12, 60.0, es)
    m = card.manifest(prob.grid, [("center", (182.88, 182.88, 0, 0, 0, 0))], [info], "u16", 0.001, 1 << 12, 60.0,
                      es, field_bounds=(0, 0, 365.76, 365.76))
    card.write_manifest(str(tmp_path), m)
    m2 = json.loads((tmp_path / "MANIFEST.JSON").read_text())
    assert m2["grid"]["total_cells"] == prob.grid.ncells
    assert m2["encoding"]["chunk_shift"] == 12 and m2["encoding"]["escape_base"] == 0xFC00
    t = m2["targets"][0]
    assert t["n_chunks"] == -(-prob.grid.ncells // (1 << 12)) and t["bytes"] == 2 * prob.grid.ncells
    assert t["reached_frac"] == pytest.approx(prob.seeds.size / prob.grid.ncells)
    # steps recovered as a reader does (section 3): (max - min) / (n - 1), heading / n
    g = m2["grid"]
    for k in range(6):
        div = g["n"][k] if g["wrap"][k] else g["n"][k] - 1
        assert (g["max"][k] - g["min"][k]) / div == pytest.approx(prob.grid.step[k], rel=1e-6)
