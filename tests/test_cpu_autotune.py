"""The autotuner's selection rules (gmapf_amd/autotune.py select) on synthetic
measurement tables: the p99 gate, the batched/plain ratio gate, re-measurement
of the best few candidates before the gates, and block-shape rows competing
with the default shape, and the p99 frame deciding between equally fast
points."""
from gmapf_amd.autotune import _entry, select, shape_options


def row(T, Q, B, solves, p99, plain, spw=0):
    return dict(T=T, Q=Q, budget=B, spw=spw, solves_per_s=solves, p50_ms=p99 * 0.9, p99_ms=p99,
                plain_solves_per_s=plain, ratio=solves / plain)


def confirm_from(overrides=None, calls=None):
    overrides = overrides or {}

    def confirm(r):
        key = (r["T"], r["Q"], r["budget"], r["spw"])
        if calls is not None:
            calls.append(key)
        return overrides.get(key, (r["solves_per_s"], r["p99_ms"]))
    return confirm


def test_target_needs_ratio_fastest_does_not():
    table = [row(4, 16, 128, 10e6, 0.5, 1.2e6),   # ratio 8.3
             row(8, 32, 128, 12e6, 0.6, 3.0e6),   # faster, ratio 4
             row(2, 16, 64, 4e6, 0.3, 1.0e6)]
    target, fastest = select(table, 4.0, 5.0, confirm_from())
    assert (target["T"], target["Q"]) == (4, 16)
    assert (fastest["T"], fastest["Q"]) == (8, 32)


def test_p99_gate_uses_the_remeasured_figures():
    table = [row(4, 16, 256, 11e6, 3.9, 1.2e6),   # passes in the table, fails when re-measured
             row(4, 16, 128, 10e6, 0.5, 1.2e6),
             row(2, 16, 128, 9e6, 0.5, 1.2e6),
             row(4, 32, 512, 13e6, 4.5, 1.2e6)]   # over the p99 gate: never a candidate
    calls = []
    target, fastest = select(table, 4.0, 5.0, confirm_from({(4, 16, 256, 0): (11e6, 4.2)}, calls), top=3)
    assert target["budget"] == 128 and fastest["budget"] == 128
    assert (4, 32, 512, 0) not in calls
    assert len(calls) == len(set(calls)) == 3  # each candidate re-measured once, shared by both picks


def test_ratio_gate_uses_the_remeasured_solves():
    table = [row(4, 16, 128, 6.6e6, 0.5, 1.2e6),  # ratio 5.5 in the table, 4.6 re-measured
             row(4, 16, 96, 6.1e6, 0.4, 1.2e6)]   # ratio 5.08
    target, fastest = select(table, 4.0, 5.0, confirm_from({(4, 16, 128, 0): (5.5e6, 0.5)}))
    assert target["budget"] == 96
    assert fastest["budget"] == 96 and fastest["solves_confirm"] == 6.1e6


def test_only_the_best_few_are_remeasured():
    table = [row(4, 16, b, s, 0.5, 1.0e6) for b, s in ((64, 7e6), (96, 8e6), (128, 9e6), (192, 10e6))]
    calls = []
    select(table, 4.0, 5.0, confirm_from(calls=calls), top=2)
    assert sorted(c[2] for c in calls) == [128, 192]


def test_nothing_within_the_frame_gate():
    table = [row(64, 256, 512, 3e6, 9.0, 1e6), row(32, 128, 256, 2e6, 6.0, 1e6)]
    target, fastest = select(table, 4.0, 5.0, confirm_from())
    assert target is None
    assert fastest["p99_ms"] == 6.0  # the lowest p99


def test_block_shape_rows_compete_and_are_reported():
    table = [row(4, 16, 128, 10e6, 0.5, 1.2e6), row(4, 16, 128, 10.5e6, 0.55, 1.2e6, spw=8)]
    target, fastest = select(table, 4.0, 5.0, confirm_from())
    assert target["spw"] == 8 and _entry(target)["spw"] == 8
    assert shape_options(4) == [8, 4] and shape_options(2) == [16, 8] and shape_options(64) == []


def test_equal_rates_go_to_the_lower_p99():
    # Re-measured within 1% of each other: the shorter p99 frame wins over the
    # (slightly) faster point, for both picks.
    table = [row(4, 16, 192, 11.854e6, 0.50, 1.25e6),
             row(4, 16, 128, 11.840e6, 0.40, 1.25e6),
             row(2, 16, 128, 11.52e6, 0.30, 1.25e6)]   # 2.8% slower: not a tie
    confirm = confirm_from({(4, 16, 192, 0): (11.854e6, 0.53)})
    target, fastest = select(table, 4.0, 5.0, confirm)
    assert target["budget"] == 128 and target["T"] == 4
    assert fastest["budget"] == 128 and fastest["T"] == 4


def test_a_clear_rate_lead_beats_a_lower_p99():
    table = [row(4, 16, 192, 12.0e6, 0.53, 1.25e6),    # 1.4% faster
             row(4, 16, 128, 11.84e6, 0.40, 1.25e6)]
    target, fastest = select(table, 4.0, 5.0, confirm_from())
    assert target["budget"] == 192 and fastest["budget"] == 192
    target, _ = select(table, 4.0, 5.0, confirm_from(), tie=0.02)
    assert target["budget"] == 128
