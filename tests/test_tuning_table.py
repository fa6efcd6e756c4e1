"""The committed tuning table and the autotuner's host logic (no GPU): the table
parses, every entry names only known knobs with values the library accepts, the
uneven-sweep entry is found by the workload's own signature, and apply() sets
exactly the defaults overlaid with the entry."""
import json

from mcluminescence_amd import autotune
from mcluminescence_amd import workloads as W

ACCEPTED = {
    "holes_per_cell": lambda v: 0.25 <= v <= 64.0,
    "persist_frac": lambda v: 0.0 < v <= 4.0,
    "wave_search_max": lambda v: 0.0 <= v <= 64.0,
    "lookahead_max": lambda v: 0.0 <= v <= 65535.0,
    "waves_per_cluster": lambda v: v in (0.0, 1.0, 2.0, 4.0, 8.0),
}


class _Ctx:
    def __init__(self):
        self.knobs = {}

    def tune_set(self, key, value):
        assert key in ACCEPTED and ACCEPTED[key](float(value)), (key, value)
        self.knobs[key] = float(value)


def test_defaults_and_space_are_accepted_values():
    assert set(autotune.DEFAULTS) == set(ACCEPTED)
    for k, v in autotune.DEFAULTS.items():
        assert ACCEPTED[k](v)
    assert set(autotune.SPACE) == set(ACCEPTED)
    for k, vals in autotune.SPACE.items():
        assert all(ACCEPTED[k](v) for v in vals)


def test_committed_table_is_valid_and_keyed_by_signature():
    table = json.loads(autotune.TABLE.read_text())
    assert table, "the committed table is empty"
    for sig, entry in table.items():
        assert "|s" in sig and "|tl" in sig, sig
        for k, v in entry.items():
            if k == "ms":
                assert v > 0
            else:
                assert k in ACCEPTED and ACCEPTED[k](float(v)), (sig, k, v)
    assert autotune.signature(W.uneven_sweep()) in table


def test_apply_overlays_the_entry_on_the_defaults():
    runs = W.uneven_sweep()
    table = autotune.load_table()
    ctx = _Ctx()
    knobs = autotune.apply(ctx, runs, table)
    entry = table[autotune.signature(runs)]
    for k, v in autotune.DEFAULTS.items():
        assert ctx.knobs[k] == float(entry.get(k, v)) == knobs[k]
    # an unknown workload keeps the defaults
    ctx2 = _Ctx()
    autotune.apply(ctx2, W.tl_glow_16x500(sims=3), {})
    assert ctx2.knobs == autotune.DEFAULTS


def test_signature_buckets_clusters_by_electrons():
    runs = W.iso_decay_4096x400(sims=100)          # 400 electrons: class 2 (257-512)
    sig = autotune.signature(runs)
    assert sig.startswith("c2:128|") and sig.endswith("|s12|tl0")
    assert autotune.signature([]) == "empty"
