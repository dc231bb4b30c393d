"""The committed tuning table (fastprox_amd/tuning.txt, written by
fastprox_amd/autotune.py and read by csrc/autotune.cpp): every line parses
to a kernel the library builds, every tuned class has an entry for every
size, and the small-batch lines (max_images 1 and 8) precede the full-batch
line of their size, which the first-match lookup relies on."""

from __future__ import annotations

import pathlib

import pytest

from fastprox_amd import autotune

TABLE = pathlib.Path(autotune.__file__).with_name("tuning.txt")


def _rows():
    rows = []
    for ln in TABLE.read_text().splitlines():
        body = ln.split("#")[0].split()
        if not body:
            continue
        assert len(body) in (11, 12), ln
        v = [float(x) for x in body]
        rows.append({"k": int(v[0]), "tsize": int(v[1]), "iso": int(v[2]), "sw": int(v[3]), "max_px": int(v[4]),
                     "kind": int(v[5]), "TT": int(v[6]), "RY": int(v[7]), "tTT": int(v[8]), "tRY": int(v[9]),
                     "frac": v[10], "max_images": int(v[11]) if len(v) == 12 else 0})
    return rows


def test_every_line_names_a_built_kernel():
    for r in _rows():
        if r["k"] == 2:
            assert r["kind"] == 0 and r["TT"] in (2, 4, 6, 8) and 1 <= r["RY"] <= 4096
            if r["frac"] > 0:
                assert r["tTT"] in (2, 4, 6, 8) and r["tTT"] >= r["TT"] and 0 < r["frac"] < 1
        else:
            assert r["kind"] == 3 and 1 <= r["RY"] <= 4096 and r["frac"] == 0
            # several iterations per launch: float32 volumes without pixel weights only
            assert r["TT"] == 1 or (r["TT"] in (2, 3) and r["tsize"] == 4 and r["sw"] == 0)


@pytest.mark.parametrize("cls", autotune.CLASSES, ids=lambda c: f"k{c[0]}-b{c[1]}-iso{int(c[2])}-sw{int(c[3])}")
def test_each_class_covers_its_sizes_small_batches_first(cls):
    k, tb, iso, sw = cls
    rows = [r for r in _rows() if (r["k"], r["tsize"], bool(r["iso"]), bool(r["sw"])) == (k, tb, iso, sw)]
    sizes = autotune.SIZES_2D if k == 2 else autotune.SIZES_3D
    assert sorted({r["max_px"] for r in rows}) == sorted(s ** k for s in sizes)
    full = [r for r in rows if r["max_images"] == 0]
    assert [r["max_px"] for r in full] == sorted(s ** k for s in sizes)
    for i, r in enumerate(rows):
        if r["max_images"]:
            assert r["max_images"] in autotune.SMALL_BATCHES
            # its size's full-batch line comes later
            assert any(q["max_px"] == r["max_px"] and q["max_images"] == 0 for q in rows[i + 1:])
