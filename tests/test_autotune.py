"""The autotuner's search and cache logic, on a fake engine and clock (no GPU)."""

from __future__ import annotations

from types import SimpleNamespace

from pyradmc_amd import autotune


class _FakeEngine:
    def __init__(self):
        self.chunk_size = None
        self.launch = {}
        self.calls = 0
        self.device_name = "fake"
        self._info = {"cu_count": 256}
        self.grid = SimpleNamespace(shape=(8, 8, 8))

    def set_launch(self, lanes, refill, push, graphs):
        self.launch = {"round_lanes_per_cu": lanes, "refill_block": refill, "push_chunk": push, "use_graphs": graphs}


def _cost(engine) -> float:
    """Separable cost with its minimum at lanes 3072, refill 128, chunk 2 Mi, graphs on."""
    lanes, refill, graphs = engine.launch["round_lanes_per_cu"], engine.launch["refill_block"], engine.launch["use_graphs"]
    chunk = engine.chunk_size or (8 << 20)
    return (1.0 + abs(lanes - 3072) / 1024 + abs(refill - 128) / 64 + abs(chunk - (2 << 20)) / (1 << 20)
            + (0.0 if graphs else 0.5))


def test_coordinate_descent_finds_the_separable_minimum():
    eng = _FakeEngine()
    clock = [0.0]

    def work():
        eng.calls += 1
        clock[0] += _cost(eng)

    candidates = {"round_lanes_per_cu": (1024, 2048, 3072, 4096), "refill_block": (64, 128, 256),
                  "chunk_size": (None, 4 << 20, 2 << 20, 1 << 20), "use_graphs": (True, False)}
    config, seconds, trials = autotune.tune(eng, work, candidates=candidates, repeats=2, timer=lambda: clock[0])
    assert (config.round_lanes_per_cu, config.refill_block, config.chunk_size, config.use_graphs) == (3072, 128, 2 << 20, True)
    assert seconds == 1.0
    assert eng.launch == {"round_lanes_per_cu": 3072, "refill_block": 128, "push_chunk": 64, "use_graphs": True}
    assert eng.chunk_size == 2 << 20
    configs = [tuple(sorted(c.items())) for c, _ in trials]
    assert len(configs) == len(set(configs))  # each configuration timed once
    assert eng.calls == 3 * len(trials)       # one warm run and two timed runs each


def test_cache_round_trip_applies_the_stored_configuration(tmp_path):
    eng = _FakeEngine()
    path = tmp_path / "autotune.json"
    cfg = autotune.LaunchConfig(chunk_size=1 << 20, round_lanes_per_cu=1536, refill_block=256, push_chunk=128,
                                use_graphs=False)
    autotune.store("k1", cfg, 0.5, path)
    autotune.store("k2", autotune.LaunchConfig(), 0.7, path)
    assert autotune.load(eng, "missing", path) is None
    got = autotune.load(eng, "k1", path)
    assert got == cfg
    assert eng.chunk_size == 1 << 20
    assert eng.launch == {"round_lanes_per_cu": 1536, "refill_block": 256, "push_chunk": 128, "use_graphs": False}


def test_workload_key_separates_what_the_shape_depends_on():
    eng = _FakeEngine()
    src = SimpleNamespace(max_energy=6.0)
    base = autotune.workload_key(eng, src, 10_000_000, 20)
    assert autotune.workload_key(eng, src, 12_000_000, 20) == base  # same power of two
    assert autotune.workload_key(eng, src, 40_000_000, 20) != base
    assert autotune.workload_key(eng, src, 10_000_000, 10) != base
    assert autotune.workload_key(eng, src, 10_000_000, 20, transport_electrons=False) != base


def test_second_pass_follows_interacting_knobs():
    """Lanes' best value depends on the refill block: a second sweep finds it."""
    eng = _FakeEngine()
    clock = [0.0]

    def cost():
        lanes, refill = eng.launch["round_lanes_per_cu"], eng.launch["refill_block"]
        best_lanes = 2048 if refill == 64 else 1024
        return 1.0 + abs(lanes - best_lanes) / 1024 + (1.5 if refill == 64 else 0.0)

    def work():
        clock[0] += cost()

    candidates = {"round_lanes_per_cu": (1024, 2048, 4096), "refill_block": (64, 256)}
    one, _, _ = autotune.tune(eng, work, candidates=candidates, repeats=1, timer=lambda: clock[0], passes=1)
    assert (one.round_lanes_per_cu, one.refill_block) == (2048, 256)
    two, t, _ = autotune.tune(eng, work, candidates=candidates, repeats=1, timer=lambda: clock[0], passes=2)
    assert (two.round_lanes_per_cu, two.refill_block) == (1024, 256) and t == 1.0


def test_failing_candidate_is_skipped():
    """A shape whose run raises (a queue overflow, say) is not a candidate."""
    eng = _FakeEngine()
    clock = [0.0]

    def work():
        if eng.launch["push_chunk"] == 256:
            raise RuntimeError("secondary queue overflow")
        clock[0] += 2.0 if eng.launch["push_chunk"] == 64 else 1.0

    config, t, trials = autotune.tune(eng, work, candidates={"push_chunk": (64, 128, 256)}, repeats=1,
                                      timer=lambda: clock[0])
    assert config.push_chunk == 128 and t == 1.0
    assert eng.launch["push_chunk"] == 128


def test_dij_tuning_times_run_dij_and_caches_under_its_own_key(tmp_path, monkeypatch):
    monkeypatch.setenv("PYRADMC_AMD_CACHE", str(tmp_path / "autotune.json"))
    eng = _FakeEngine()
    clock = [0.0]
    seen = []

    def run_dij(source, counts, n_batches, seed, **kw):
        seen.append((counts, n_batches, seed, kw))
        clock[0] += 1.0 + abs(eng.launch["round_lanes_per_cu"] - 1024) / 1024

    eng.run_dij = run_dij
    src = SimpleNamespace(max_energy=6.0, n_beamlets=225)
    candidates = {"round_lanes_per_cu": (512, 1024, 2048)}
    config = autotune.tune_dij(eng, src, 20_000, 10, seed=3, repeats=1, candidates=candidates,
                               timer=lambda: clock[0], truncation=1e-4)
    assert config.round_lanes_per_cu == 1024
    assert seen and all(s == (20_000, 10, 3, {"truncation": 1e-4}) for s in seen)
    cache = (tmp_path / "autotune.json").read_text()
    assert '"dij|' in cache and "|c225" in cache
    calls = len(seen)
    eng.launch = {}
    again = autotune.tune_dij(eng, src, 20_000, 10, seed=3, candidates=candidates, timer=lambda: clock[0],
                              truncation=1e-4)
    assert again == config and len(seen) == calls  # from the cache, nothing timed
    assert eng.launch["round_lanes_per_cu"] == 1024
