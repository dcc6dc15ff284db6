"""Tuning table of the fused shell batch: the Legendre-synthesis tile of each
shell shape, as measured by the batch autotuner (gd_batch_tune) on MI355X
(gfx950).  Every tile gives the same bits, so the table only moves time.

The table (synth_tiles_gfx950.json) maps "precision lmax mmax nlat nlon nr"
to a tile; `record` merges a tuned batch into it.
"""
import json
import os

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "synth_tiles_gfx950.json")


def shape_key(step):
    return "%s %d %d %d %d %d" % (step.precision, step.lmax, step.mmax, step.nlat, step.nlon, step.nr)


def load(path=TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get("tiles", {})


def synth_tiles(steps, path=TABLE):
    """Each step's tile from the table; 0 (the step's own synthesis form)
    where its shape is absent."""
    table = load(path)
    return [int(table.get(shape_key(s), 0)) for s in steps]


def record(steps, tiles, path=TABLE, note=None):
    """Adds the tuned tiles of `steps` to the table at `path`."""
    doc = {"tiles": {}}
    if os.path.exists(path):
        with open(path) as f:
            doc = json.load(f)
    doc.setdefault("tiles", {})
    for s, t in zip(steps, tiles):
        doc["tiles"][shape_key(s)] = int(t)
    if note:
        doc["note"] = note
    doc["tiles"] = dict(sorted(doc["tiles"].items()))
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")


def tune_shapes(shapes, prec, reps=5, path=TABLE):
    """Runs the batch autotuner over one shell of each (lmax, nr) in `shapes`
    (grid 3(L+1)/2 x 3(L+1)) and records the chosen tiles at `path`."""
    from .fields import seeded_state
    from .step import CNAB2Step, StepBatch

    steps = []
    for k, (L, nr) in enumerate(shapes):
        s = CNAB2Step(L, 3 * (L + 1) // 2, 3 * (L + 1), nr, precision=prec)
        s.set_state_complex(seeded_state(700 + k, L, nr, amp=1e-3))
        steps.append(s)
    b = StepBatch(steps, mode="fused", tiles=[0] * len(steps))
    tiles = b.tune(reps)
    record(steps, tiles, path, note="gd_batch_tune on MI355X (gfx950), %d repetitions" % reps)
    out = {"prec": prec, "tiles": tiles, "assignment_ms": b.tune_ms}
    b.close()
    for s in steps:
        s.close()
    return out


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="tune the fused batch's synthesis tiles of the uneven-shell shapes")
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--prec", nargs="+", default=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--reps", type=int, default=10)
    a = ap.parse_args()
    shapes = [(L, nr) for L in (31, 47, 63, 95, 127) for nr in (33, 49, 65, 97)]
    for p in a.prec:
        print(json.dumps(tune_shapes(shapes, p, a.reps, a.out)), flush=True)
