"""The shipped tuning table (tuning_gfx950.txt, written by tags_sweep_autotune
on an MI355X and read by bench.py and the sweeps through the autotuner's
cache): every entry parses as the library's cache reader expects, its values
are ones tags_sweep_set_tuning accepts, keys are unique, and the groups of
samples chosen for large systems fit the device memory bound the autotuner
applies when it tries them. No GPU needed (the variant counts come from the
library's host-side tables)."""
import os
import re

import pytest

import tags_pkg

T = tags_pkg.load()
TABLE = os.path.join(tags_pkg.PKG_DIR, "tuning_gfx950.txt")
KEY_RE = re.compile(r"^(\S+) cu=(\d+) form=(hem|mhem) plan=(\d+) prec=(mixed|fp64) N=(\d+) m=(\d+)$")
FIELDS = ("big", "small", "width", "chunks", "rows", "inflight", "batch", "nodal_above",
          "gmres", "ms")
HBM_BYTES = 288e9  # MI355X


def _entries():
    out = []
    for line in open(TABLE):
        line = line.rstrip("\n")
        if not line.strip() or line.startswith("#"):
            continue
        key, _, vals = line.partition(" | ")
        out.append((key, vals))
    return out


def test_table_parses_and_keys_are_unique():
    ents = _entries()
    assert len(ents) >= 8
    keys = [k for k, _ in ents]
    assert len(keys) == len(set(keys))
    for key, vals in ents:
        assert KEY_RE.match(key), key
        names = [kv.split("=")[0] for kv in vals.split()]
        assert tuple(names) == FIELDS, (key, names)


@pytest.mark.parametrize("entry", _entries(), ids=lambda e: e[0].split(" ", 2)[2])
def test_values_are_accepted(entry):
    key, vals = entry
    m = KEY_RE.match(key)
    prec = T.PREC_MIXED if m.group(5) == "mixed" else T.PREC_FP64
    N = int(m.group(6))
    v = {kv.split("=")[0]: float(kv.split("=")[1]) for kv in vals.split()}
    nvar = T.lib().tags_gemm_variants(prec)
    assert nvar == (34 if prec == T.PREC_MIXED else 16)
    for f in ("big", "small"):
        assert v[f] == int(v[f]) and -1 <= v[f] < nvar, (f, v[f])
    assert v["width"] in (-1, 0, 128, 256)
    assert v["chunks"] == -1 or 1 <= v["chunks"] <= 64
    assert v["rows"] in (-1, 0, 1, 2)
    assert 0 <= v["inflight"] <= 4
    assert v["batch"] == -1 or 1 <= v["batch"] <= 64
    assert v["gmres"] == 0 or 1 <= v["gmres"] <= 32
    if prec == T.PREC_FP64:
        assert v["nodal_above"] == -1 and v["gmres"] == 0
    assert v["ms"] > 0
    if v["batch"] > 1 and N >= 8192:
        # groups in flight within 80 % of the device's memory: W (16 bytes per
        # entry) and, in the mixed form, W32 (8 bytes) per sample
        per = N * N * (24.0 if prec == T.PREC_MIXED else 16.0)
        k = v["inflight"] if v["inflight"] > 0 else 4
        assert v["batch"] * k * per <= 0.8 * HBM_BYTES
