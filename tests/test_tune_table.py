"""The committed tuning table (pmdd_amd/tune/gfx950.json, written by
python -m pmdd_amd.tune) holds, per objective count, launch shapes the engine
accepts (pmdd_ctx_set_tuning's ranges)."""
import json
import os

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pmdd_amd",
                     "tune", "gfx950.json")


def test_table_shapes_are_in_range():
    t = json.load(open(TABLE))
    assert t["device"] == "gfx950"
    for d, cfg in t["tuning"].items():
        assert 1 <= int(d) <= 8
        assert len(cfg["filter_threads"]) == 3
        assert all(n in (128, 256, 512) for n in cfg["filter_threads"])
        assert 256 <= cfg["cross_workgroups"] <= 16384
        assert 1 <= cfg["cross_split_max"] <= 16
        assert 256 <= cfg["refill_waves"] <= 32768
        m = t["measured"][d]
        # the chosen shapes are measured beside the built-in ones
        assert m["tuned_ms"]["kernel"] > 0 and m["builtin_ms"]["kernel"] > 0
