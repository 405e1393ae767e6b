"""Host logic of the launch-shape tuner (no GPU): candidate generation, the
tie-break, and the shipped MI355X table."""
import math

import numpy as np

from emtrace import autotune


def test_candidates_fit_lds_and_granules():
    for p, n, batch in ((16384, 65536, False), (1000, 8192, False), (8188, 713262, True),
                        (7, 130, False)):
        cands = autotune.candidates(p, n, 256, batch=batch)
        assert cands
        for S, block, tile in cands:
            assert S in (1, 2, 4, 8, 16, 32) and block in (256, 512)
            assert tile % max(256, S * autotune.GROUP) == 0
            assert tile * 16 + autotune.LDS_STATIC <= autotune.LDS_BYTES
            if batch:
                assert S <= 8


def test_pick_prefers_fewer_lanes_among_ties():
    rep = [(1, 256, 512, 0, 10.5), (2, 256, 512, 0, 10.04), (4, 512, 512, 0, 10.0),
           (8, 512, 512, 0, 9.99)]
    assert autotune._pick(rep)[:3] == (2, 256, 512)
    rep = [(1, 256, 512, 0, 12.0), (4, 512, 512, 0, 10.0), (8, 512, 512, 0, float("inf"))]
    assert autotune._pick(rep)[:3] == (4, 512, 512)


def test_unroll_stage_times_the_fastest_shape_at_each_unroll():
    """Stage 2 runs the stage-1 winner at U = 8, 4, 2; an explicit unroll is
    kept only when it beats the default by more than the tie margin."""
    cost = {(1, 256, 512): 10.0, (2, 256, 512): 8.0, (4, 512, 512): 9.0}
    gain = {0: 1.0, 8: 1.0, 4: 0.95, 2: 0.90}
    calls = []

    def run(launch):
        calls.append((launch.lanes_per_line, launch.block, launch.tile, launch.unroll))
        return cost[(launch.lanes_per_line, launch.block, launch.tile)] * gain[launch.unroll]

    best, report = autotune._tune(list(cost), run)
    assert calls[:3] == [(1, 256, 512, 0), (2, 256, 512, 0), (4, 512, 512, 0)]
    assert calls[3:] == [(2, 256, 512, 8), (2, 256, 512, 4), (2, 256, 512, 2)]
    assert best[:4] == (2, 256, 512, 2) and len(report) == 6
    gain.update({4: 1.0, 2: 1.005})
    best, _ = autotune._tune(list(cost), run)
    assert best[:4] == (2, 256, 512, 0)


def test_old_cache_entries_read_as_default_unroll(tmp_path, monkeypatch):
    monkeypatch.setenv("EMTRACE_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune._store("k", [4, 512, 4096, 1.5])
    assert autotune._lookup("k") == [4, 512, 4096, 0, 1.5]
    launch = autotune._launch(autotune._lookup("k"))
    assert (launch.lanes_per_line, launch.tile, launch.block, launch.unroll) == (4, 4096, 512, 0)


def test_shipped_shapes_are_candidates_of_their_problem():
    for key, (S, block, tile, U, _) in autotune.SHIPPED.items():
        assert U in (0, 2, 4, 8)
        kind, arch, cus, *rest = key.split(":")
        assert arch == "gfx950" and cus == "256"
        if kind == "plain":
            p, n, stop = int(rest[0]), int(rest[1]), rest[3] == "1"
            cands = autotune.candidates(p, n, 256, stop=stop)
        else:
            p, n = int(rest[1]), int(rest[2])
            cands = autotune.candidates(p, n, 256, batch=True)
            cands += [(s, b, min(t, int(math.ceil(p / 256) * 256))) for s, b, t in cands]
        assert (S, block, tile) in cands, key


def test_stop_candidates_include_wide_groups_and_small_blocks():
    full = autotune.candidates(16384, 65536, 256)
    stop = autotune.candidates(16384, 65536, 256, stop=True)
    assert max(c[0] for c in full) <= 8  # enough waves already at S = 4
    assert (32, 64, 512) in stop and min(c[1] for c in stop) == 64


def test_register_resident_trace_takes_no_launch_shape():
    """Line ends without checkpoints over 4,096..16,384 charges take the
    register-resident trace, whose launch is fixed: the tuner does not time
    shapes for it (the library ignores a given one)."""
    from emtrace import api, rules
    assert not api.takes_launch_shape(4096, rules.angle())
    assert not api.takes_launch_shape(16384, rules.angle())
    assert api.takes_launch_shape(4095, rules.angle())
    assert api.takes_launch_shape(16385, rules.angle())
    assert api.takes_launch_shape(16384, rules.angle(stop=False))
    assert api.takes_launch_shape(16384, rules.angle(), ckpt=True)
    assert api.takes_launch_shape(16384, rules.plain())
    charges = np.zeros((8192, 4), np.float32)
    seeds = np.zeros((64, 3), np.float32)
    launch, report = autotune.tune_plain(charges, seeds, rules.angle())
    assert (launch.lanes_per_line, launch.tile, launch.block, launch.unroll) == (0, 0, 0, 0)
    assert "no launch shape" in report[0][0]
