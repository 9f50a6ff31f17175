"""Autotuner for the transport launch shape.

Knobs, none of which changes a result (each particle carries its own random stream and
the tallies are integer sums, so dose and Dij are bit-identical under any of them):

* ``chunk_size`` - histories generated and resident at once (the queues hold
  ``chunk_size * queue_factor`` particles); the reference's own knob
  (pyradmc/backends/warp/engine.py:175-186 records its chunk sweep);
* ``round_lanes_per_cu`` - lanes per CU in each round's grid (persistent lanes; the
  resident share runs first, the rest as earlier lanes retire);
* ``refill_block`` - queue indices a wave claims with one atomic;
* ``push_chunk`` - output queue slots a wave reserves with one atomic;
* ``use_graphs`` - replay each group of four rounds from a captured HIP graph.

``tune`` runs a coordinate descent over candidate values, timing a caller-supplied
piece of work (the median of ``repeats`` runs per candidate), applies the best and
stores it in a JSON cache keyed by device and workload, so later engines on the same
workload pick it up with ``load``. ``tune_run`` times an open-field ``run``,
``tune_dij`` a whole ``run_dij``.
"""

from __future__ import annotations

import json
import math
import os
import statistics
import time
from dataclasses import asdict, dataclass
from pathlib import Path

DEFAULT_CANDIDATES = {
    "round_lanes_per_cu": (512, 768, 1024, 1536, 2048, 3072),
    "refill_block": (64, 128, 256, 512, 1024),
    "push_chunk": (64, 128, 256),
    "chunk_size": (None, 4 << 20, 2 << 20),
    "use_graphs": (True, False),
}


@dataclass
class LaunchConfig:
    chunk_size: int | None = None  # None: the engine sizes chunks from free memory
    round_lanes_per_cu: int = 2048
    refill_block: int = 64
    push_chunk: int = 64
    use_graphs: bool = True


def cache_path() -> Path:
    root = os.environ.get("PYRADMC_AMD_CACHE")
    return Path(root) if root else Path.home() / ".cache" / "pyradmc_amd" / "autotune.json"


def apply(engine, config: LaunchConfig) -> None:
    engine.chunk_size = config.chunk_size
    engine.set_launch(config.round_lanes_per_cu, config.refill_block, config.push_chunk, config.use_graphs)


def workload_key(engine, source, n_histories: int, n_batches: int, transport_electrons: bool = True) -> str:
    """Device, grid, source type and energy, history count (to a power of two) and
    batches: the things the best launch shape depends on."""
    grid = engine.grid
    scale = max(int(n_histories), 1).bit_length()
    parts = [engine.device_name, str(engine._info.get("cu_count")), "x".join(str(int(v)) for v in grid.shape),
             type(source).__name__, f"{float(source.max_energy):.4g}", f"e{int(bool(transport_electrons))}",
             f"n2^{scale}", f"b{int(n_batches)}"]
    return "|".join(parts)


def _read_cache(path: Path) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def load(engine, key: str, path: Path | None = None) -> LaunchConfig | None:
    """The cached configuration for ``key``, applied to the engine; None if absent."""
    entry = _read_cache(path or cache_path()).get(key)
    if not entry:
        return None
    config = LaunchConfig(**entry["config"])
    apply(engine, config)
    return config


def store(key: str, config: LaunchConfig, seconds: float, path: Path | None = None) -> None:
    path = path or cache_path()
    data = _read_cache(path)
    data[key] = {"config": asdict(config), "seconds": seconds}
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(".tmp")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
    os.replace(tmp, path)


def tune(engine, work, candidates: dict | None = None, repeats: int = 3, start: LaunchConfig | None = None,
         timer=time.perf_counter, log=None, passes: int = 2) -> tuple[LaunchConfig, float, list]:
    """Coordinate descent over ``candidates`` (knob -> values, in the order given):
    each knob in turn takes the value with the lowest median time of ``work()`` over
    ``repeats`` runs, the other knobs held at their best so far; up to ``passes``
    sweeps over the knobs (stopping early when a sweep changes nothing, since the
    knobs interact). Applies the best configuration to ``engine`` and returns
    (config, seconds, trials)."""
    candidates = DEFAULT_CANDIDATES if candidates is None else candidates
    best = LaunchConfig(**asdict(start)) if start else LaunchConfig()
    measured: dict = {}
    trials = []

    def measure(config: LaunchConfig) -> float:
        key = tuple(sorted(asdict(config).items()))
        if key not in measured:
            try:
                apply(engine, config)
                work()  # warm: allocations, code objects, graph capture
                times = []
                for _ in range(repeats):
                    t0 = timer()
                    work()
                    times.append(timer() - t0)
                measured[key] = statistics.median(times)
            except RuntimeError as exc:  # e.g. queue overflow at this shape: not a candidate
                measured[key] = math.inf
                if log:
                    log(f"autotune: {asdict(config)} failed: {exc}")
            trials.append((asdict(config), measured[key]))
            if log:
                log(f"autotune: {asdict(config)} -> {measured[key] * 1e3:.3f} ms")
        return measured[key]

    best_time = measure(best)
    for _ in range(max(passes, 1)):
        before = best
        for knob, values in candidates.items():
            for value in values:
                trial = LaunchConfig(**{**asdict(best), knob: value})
                t = measure(trial)
                if t < best_time:
                    best, best_time = trial, t
        if best == before:
            break
    apply(engine, best)
    return best, best_time, trials


def tune_run(engine, source, n_histories: int, n_batches: int, seed: int = 1, repeats: int = 3,
             candidates: dict | None = None, use_cache: bool = True, log=None, **run_kw) -> LaunchConfig:
    """Tune on ``engine.run(source, n_histories, n_batches, seed, **run_kw)``; a cached
    result for the same workload is applied without timing anything."""
    key = workload_key(engine, source, n_histories, n_batches, run_kw.get("transport_electrons", True))
    if use_cache:
        cached = load(engine, key)
        if cached is not None:
            return cached
    config, seconds, _ = tune(engine, lambda: engine.run(source, n_histories, n_batches, seed, **run_kw),
                              candidates=candidates, repeats=repeats, log=log)
    if use_cache:
        store(key, config, seconds)
    return config


def tune_dij(engine, source, n_histories_per_beamlet, n_batches: int, seed: int = 1, repeats: int = 3,
             candidates: dict | None = None, use_cache: bool = True, log=None, timer=time.perf_counter,
             passes: int = 2, **dij_kw) -> LaunchConfig:
    """Tune on ``engine.run_dij(source, n_histories_per_beamlet, n_batches, seed, **dij_kw)``
    (transport, reduction and the CSC on the host, timed whole); cached like ``tune_run``
    under a key that also names the beamlet count."""
    counts = n_histories_per_beamlet
    total = int(sum(counts)) if hasattr(counts, "__len__") else int(counts) * int(source.n_beamlets)
    key = "dij|" + workload_key(engine, source, total, n_batches, dij_kw.get("transport_electrons", True)) \
        + f"|c{int(source.n_beamlets)}"
    if use_cache:
        cached = load(engine, key)
        if cached is not None:
            return cached
    config, seconds, _ = tune(engine, lambda: engine.run_dij(source, counts, n_batches, seed, **dij_kw),
                              candidates=candidates, repeats=repeats, log=log, timer=timer, passes=passes)
    if use_cache:
        store(key, config, seconds)
    return config
