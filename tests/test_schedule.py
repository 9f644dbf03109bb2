"""The built-in K-split rules of the contractions (host only, no GPU)."""
from __future__ import annotations

import ctypes

import pytest


def _rule(rule, grid, chunks, slots, max_chunks=1 << 30):
    from summit_amd import _native
    out = ctypes.c_int32()
    _native.call("gpe_schedule_split_rule", rule, grid, chunks, max_chunks, slots, ctypes.byref(out))
    return out.value


def test_phase1_rule_at_the_headline_and_small_shapes():
    # N=100,000 -> 782 workgroups per split (two probe-column groups of 391 row blocks),
    # 512 resident slots, M=500,000 -> 3,907 chunks: add splits while the grid holds
    # fewer than 16 rounds and each split keeps >= 128 chunks -> 11
    assert _rule(1, 782, 3907, 512) == 11
    # N=5,000, M=10,000: 79 chunks, too few to split
    assert _rule(1, 40, 79, 512) == 1
    # at most 16 splits, and never more chunks per split than allowed
    assert _rule(1, 1, 100000, 512) == 16
    assert _rule(1, 782, 3907, 512, max_chunks=100) == 40


def test_fill_rule():
    # 782 workgroups per split in rounds of 512 slots: the smallest split count
    # whose last round is >= 90% full wins
    fills = {s: (782 * s) / (-(-782 * s // 512) * 512) for s in range(1, 17)}
    first = next(s for s in range(1, 17) if fills[s] >= 0.9)
    assert _rule(0, 782, 3907, 512) == first
    assert _rule(0, 512, 3907, 512) == 1


def test_rule_arguments_are_checked():
    from summit_amd import _native
    with pytest.raises(_native.GpeError, match="split rule"):
        _rule(2, 10, 10, 10)
