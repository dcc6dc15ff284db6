"""The fused batch's tuning table (gdhip/tuning.py), host side: the key of a
shell shape, lookups with the default for absent shapes, merging a tuned
batch into a table, and the committed table's coverage of the uneven-shell
workload's shapes."""
import json
import os


class _Shape:
    def __init__(self, precision, lmax, nr):
        self.precision, self.lmax, self.mmax, self.nr = precision, lmax, lmax, nr
        self.nlat, self.nlon = 3 * (lmax + 1) // 2, 3 * (lmax + 1)


def test_key_and_default(tmp_path):
    from gdhip import tuning

    s = _Shape("fp32", 63, 49)
    assert tuning.shape_key(s) == "fp32 63 63 96 192 49"
    path = str(tmp_path / "t.json")
    assert tuning.load(path) == {}
    assert tuning.synth_tiles([s], path) == [0]


def test_record_merges_and_sorts(tmp_path):
    from gdhip import tuning

    path = str(tmp_path / "t.json")
    a, b = _Shape("fp64", 31, 33), _Shape("fp32", 127, 97)
    tuning.record([a], [1], path, note="first")
    tuning.record([b, a], [2, 0], path)
    doc = json.load(open(path))
    assert doc["note"] == "first"
    assert list(doc["tiles"]) == sorted(doc["tiles"])
    assert tuning.synth_tiles([a, b, _Shape("bf16x3", 31, 33)], path) == [0, 2, 0]


def test_committed_table_covers_the_uneven_shell_shapes():
    """Every (lmax, nr) of the 64-shell workload has an FP64 and an FP32 entry,
    each a tile the precision has (tile 2 only with FP32 operands)."""
    from gdhip import tuning

    assert os.path.exists(tuning.TABLE)
    table = tuning.load()
    for prec in ("fp64", "fp32"):
        for L in (31, 47, 63, 95, 127):
            for nr in (33, 49, 65, 97):
                t = table.get(tuning.shape_key(_Shape(prec, L, nr)))
                assert t in ((0, 1) if prec == "fp64" else (0, 1, 2)), (prec, L, nr, t)
