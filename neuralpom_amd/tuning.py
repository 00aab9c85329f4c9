"""Launch shapes per workload (pom_batch_tune), from the table the tuner
writes (``autotune.py`` -> ``tuning.json`` beside this file).

An ensemble takes the entry whose workload is nearest to its own: same
precision first, then the closest total of 2-D points (log scale), then the
closest member count.  Without a table (or an entry for the precision) it
keeps the library's defaults."""
from __future__ import annotations

import ctypes as C
import json
import math
from pathlib import Path

from . import _lib as L

TABLE = Path(__file__).resolve().parent / "tuning.json"
KEYS = ("map2_blocks_per_group", "map2_min_groups", "flat2_blocks_per_group", "flat2_min_groups",
        "flat3_blocks_per_group", "flat3_min_groups")
DEFAULT = dict(zip(KEYS, (4, 768, 1, 768, 1, 768)))

_cache: dict = {}


def load_table(path: Path = TABLE):
    key = str(path)
    if key not in _cache:
        _cache[key] = json.loads(path.read_text()) if path.exists() else {"entries": []}
    return _cache[key]


def workload_key(shapes, dtype_name: str):
    return {"dtype": dtype_name, "members": len(shapes), "points2d": int(sum(im * jm for im, jm in shapes))}


def select(shapes, dtype_name: str, table=None):
    """(entry name, shapes dict) for a workload; ("default", DEFAULT) without a match."""
    table = load_table() if table is None else table
    key = workload_key(shapes, dtype_name)
    cands = [e for e in table.get("entries", []) if e["key"]["dtype"] == dtype_name]
    if not cands:
        return "default", dict(DEFAULT)

    def dist(e):
        k = e["key"]
        return (abs(math.log(k["points2d"]) - math.log(max(key["points2d"], 1))),
                abs(k["members"] - key["members"]))
    best = min(cands, key=dist)
    return best["workload"], {k: int(best["tuning"][k]) for k in KEYS}


def apply(lib, handle, tuning: dict):
    t = L.Tuning(*[int(tuning[k]) for k in KEYS])
    L.check(lib.pom_batch_tune(handle, C.byref(t)), "pom_batch_tune")


def current(lib, handle) -> dict:
    t = L.Tuning()
    L.check(lib.pom_batch_get_tuning(handle, C.byref(t)), "pom_batch_get_tuning")
    return {k: int(getattr(t, k)) for k in KEYS}
