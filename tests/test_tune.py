"""The counting settings table on the host (no device needed): size
classes, the MI355X defaults and the candidate set."""
from __future__ import annotations

import gpu_bitruss_amd as gb


def test_size_classes():
    assert gb.tune_class(0) == 0 and gb.tune_class(3) == 0
    assert gb.tune_class(4) == 1 and gb.tune_class(15) == 1 and gb.tune_class(16) == 2
    assert gb.tune_class(2_000_000) == 10 and gb.tune_class(58_883_633) == 12
    assert gb.tune_class(1 << 62) == 15
    # a batch union is classed by its edges per graph
    assert gb.tune_class(86_700_000, 4096) == gb.tune_class(86_700_000 // 4096) == 7


def test_defaults_are_candidates():
    cands = set(gb.tune_candidates())
    assert len(cands) == 48
    for c in range(16):
        cfg = gb.tune_default(c)
        assert cfg in cands
        wave, block, threads, hub = cfg
        assert wave < block and threads in (256, 512) and hub in (256, 512)
