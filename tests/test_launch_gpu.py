"""Launch shape and HIP graphs: groups of four transport rounds replayed from a captured
graph, the lanes per CU of each round and the refill block change no bit of a result.
Graph replays are checked one by one against eager runs, with eager work and host
synchronisation between them, and with other work items (seeds) reusing the graph."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def slab():
    from pyradmc_amd import workloads

    grid, xs, src, _ = workloads.slab_phantom()
    return grid, xs, src


def _engine(grid, xs, **launch):
    from pyradmc_amd import HipEngine

    eng = HipEngine(grid, xs)
    if launch:
        eng.set_launch(**launch)
    return eng


def _same(a, b) -> None:
    np.testing.assert_array_equal(a.dose, b.dose)
    np.testing.assert_array_equal(a.dose_sigma, b.dose_sigma)
    for name in ("energy_emitted", "energy_deposited", "energy_escaped", "energy_unscored"):
        assert getattr(a, name) == getattr(b, name), name


def test_launch_shape_and_graphs_change_no_bit(slab):
    grid, xs, src = slab
    eager = _engine(grid, xs, use_graphs=False).run(src, 300_000, 4, seed=11)
    assert eager.counters["graph_launches"] == 0
    for launch in ({}, {"round_lanes_per_cu": 1024}, {"round_lanes_per_cu": 4096, "refill_block": 256},
                   {"refill_block": 128, "use_graphs": False}, {"round_lanes_per_cu": 1024, "push_chunk": 256},
                   {"refill_block": 512, "push_chunk": 128, "use_graphs": False}):
        got = _engine(grid, xs, **launch).run(src, 300_000, 4, seed=11)
        _same(eager, got)
        assert (got.counters["graph_launches"] > 0) == launch.get("use_graphs", True), launch


def test_graph_replays_match_eager_runs(slab):
    grid, xs, src = slab
    eng = _engine(grid, xs)
    ref = _engine(grid, xs, use_graphs=False)
    expected = {seed: ref.run(src, 200_000, 2, seed=seed) for seed in (1, 2)}
    for seed in (1, 2, 1, 1, 2):  # replays of one captured graph across runs and seeds
        got = eng.run(src, 200_000, 2, seed=seed)
        assert got.counters["graph_launches"] > 0
        _same(expected[seed], got)
        _same(expected[seed], ref.run(src, 200_000, 2, seed=seed))  # eager work between replays


def test_profiled_runs_launch_one_by_one_between_graph_replays(slab):
    grid, xs, src = slab
    eng = _engine(grid, xs)
    plain = eng.run(src, 200_000, 2, seed=3)
    assert plain.counters["graph_launches"] > 0
    eng.set_profiling(True)
    timed = eng.run(src, 200_000, 2, seed=3)
    eng.set_profiling(False)
    _same(plain, timed)
    assert timed.counters["graph_launches"] == 0
    assert timed.counters["photon_kernel_ms"] > 0.0 and timed.counters["electron_kernel_ms"] > 0.0
    again = eng.run(src, 200_000, 2, seed=3)
    assert again.counters["graph_launches"] > 0
    _same(plain, again)


def test_launch_arguments_rejected(slab):
    grid, xs, _ = slab
    eng = _engine(grid, xs)
    for bad in ({"round_lanes_per_cu": 100}, {"round_lanes_per_cu": 16384}, {"refill_block": 96},
                {"refill_block": 2048}, {"push_chunk": 32}, {"push_chunk": 100}):
        with pytest.raises(ValueError):
            eng.set_launch(**bad)


def test_autotune_picks_a_shape_that_changes_no_bit(slab, tmp_path, monkeypatch):
    from pyradmc_amd import autotune

    monkeypatch.setenv("PYRADMC_AMD_CACHE", str(tmp_path / "autotune.json"))
    grid, xs, src = slab
    eng = _engine(grid, xs)
    candidates = {"round_lanes_per_cu": (1024, 2048), "refill_block": (64, 128), "chunk_size": (None, 100_000)}
    config = eng.autotune(src, 200_000, 2, seed=1, repeats=1, candidates=candidates)
    assert (tmp_path / "autotune.json").is_file()
    assert eng.launch["round_lanes_per_cu"] == config.round_lanes_per_cu
    again = _engine(grid, xs)
    assert again.autotune(src, 200_000, 2, seed=1) == config  # from the cache
    _same(_engine(grid, xs, use_graphs=False).run(src, 200_000, 2, seed=9), eng.run(src, 200_000, 2, seed=9))


def test_dij_autotune_picks_a_shape_that_changes_no_bit(tmp_path, monkeypatch):
    from pyradmc_amd import BeamletGridSource, VoxelGrid
    from pyradmc_amd.data.analytic import AnalyticCrossSections

    monkeypatch.setenv("PYRADMC_AMD_CACHE", str(tmp_path / "autotune.json"))
    grid = VoxelGrid.uniform_water((12, 12, 20), (0.5, 0.5, 0.5))
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    src = BeamletGridSource(4.0, -0.1, (1.0, 5.0), (1.0, 5.0), 3, 3)
    eng = _engine(grid, xs)
    candidates = {"round_lanes_per_cu": (512, 1024), "push_chunk": (64, 256), "chunk_size": (None, 50_000)}
    config = eng.autotune_dij(src, 20_000, 4, seed=1, repeats=1, candidates=candidates, truncation=1e-3,
                              value_dtype=np.float32)
    assert "dij|" in (tmp_path / "autotune.json").read_text()
    assert eng.launch["round_lanes_per_cu"] == config.round_lanes_per_cu
    tuned = eng.run_dij(src, 20_000, 4, seed=5, truncation=1e-3)
    plain = _engine(grid, xs, use_graphs=False).run_dij(src, 20_000, 4, seed=5, truncation=1e-3)
    for name in ("indptr", "indices", "dose", "sigma"):
        np.testing.assert_array_equal(getattr(tuned, name), getattr(plain, name))
