"""Launch-shape tables (neuralpom_amd.tuning): selection of the nearest
tuned workload, the defaults without a match, and the committed table's
entries within the range pom_batch_tune accepts."""
import math

from neuralpom_amd import _lib as L
from neuralpom_amd import tuning as TU

TABLE = {"entries": [
    {"workload": "small", "key": {"dtype": "f32", "members": 1, "points2d": 7380},
     "tuning": dict(TU.DEFAULT, flat2_blocks_per_group=2)},
    {"workload": "big", "key": {"dtype": "f32", "members": 64, "points2d": 780936},
     "tuning": dict(TU.DEFAULT, map2_blocks_per_group=8)},
    {"workload": "big64", "key": {"dtype": "f64", "members": 64, "points2d": 780936},
     "tuning": dict(TU.DEFAULT, flat3_min_groups=1536)},
]}


def test_select_nearest_by_points_then_members():
    assert TU.select([(90, 82)], "f32", TABLE) == ("small", TABLE["entries"][0]["tuning"])
    assert TU.select([(90, 82)] * 60, "f32", TABLE)[0] == "big"
    assert TU.select([(100, 100)] * 3, "f32", TABLE)[0] == "small"
    assert TU.select([(354, 322)] * 16, "f32", TABLE)[0] == "big"
    # the precision is matched first
    assert TU.select([(90, 82)], "f64", TABLE)[0] == "big64"


def test_select_defaults_without_entries():
    assert TU.select([(90, 82)], "f32", {"entries": []}) == ("default", TU.DEFAULT)
    name, shapes = TU.select([(90, 82)], "f64", {"entries": TABLE["entries"][:2]})
    assert name == "default" and shapes == TU.DEFAULT


def test_keys_match_the_c_struct():
    assert tuple(f for f, _ in L.Tuning._fields_) == TU.KEYS


def test_committed_table_is_valid():
    table = TU.load_table()
    assert table["entries"], "tuning.json has no entries"
    for e in table["entries"]:
        assert set(e["tuning"]) == set(TU.KEYS), e["workload"]
        assert all(1 <= int(v) <= (1 << 20) for v in e["tuning"].values()), e["workload"]
        assert e["key"]["dtype"] in ("f32", "f64")
        # the chosen shapes are never slower than the defaults they were measured against
        assert e["ms_per_step"] <= e["default_ms"] + 1e-9
        assert math.isfinite(e["ms_per_step"])
