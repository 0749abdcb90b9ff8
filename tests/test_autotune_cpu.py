"""Autotuner host logic (no GPU): size buckets, launch weights, and that a
tuned table is applied to a mesh of the matching size only."""
import json

from solids4foam_amd import autotune


class FakeMesh:
    def __init__(self, n):
        self.nCells = n
        self.tiles = None

    def set_tiles(self, cap, mn):
        self.tiles = (cap, mn)


def test_bucket_is_the_nearest_power_of_two():
    assert autotune.bucket(8_000_000) == str(2 ** 23)
    assert autotune.bucket(4096) == "4096"
    assert autotune.bucket(5000) == "4096" and autotune.bucket(6000) == "8192"


def test_weights_count_launches_per_outer_iteration():
    w = autotune.weights(7.0)
    assert w["residual"] == w["grad"] == 1.0
    assert w["spmv"] == w["pcg_update"] == 7.0 and w["vcycle"] == 8.0 and w["finalize"] == 23.0


def test_apply_uses_the_matching_bucket(tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"tiles": {"4096": {"max_tiles": 65536, "min_slices": 2, "cells": 4096}}}))
    table = autotune.load_table(str(p))
    m = FakeMesh(4100)
    assert autotune.apply(m, table) == (65536, 2) and m.tiles == (65536, 2)
    m2 = FakeMesh(40000)
    assert autotune.apply(m2, table) is None and m2.tiles is None
