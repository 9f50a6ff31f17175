"""Phase-space sources: IAEA phase-space files and in-memory particle populations.

A phase space is a recorded population of particles (photons, electrons, positrons) at
arbitrary positions, directions and statistical weights, typically scored on a plane
below a treatment head. ``HipEngine.run`` transports it by one of two routes:

* host-sampled (``PhaseSpaceSource``, ``InMemoryPhaseSpaceSource``, as the reference's
  ``WarpEngine`` does): each block of histories draws records with replacement on the
  host and the columns go to the device with every transport call;
* device-resident (``DevicePhaseSpace`` around either): the population is copied to
  device memory once and the generator kernel draws each history's record itself,
  which removes the host sampling that otherwise bounds the run's rate.

Either way the generator sorts the records into the photon and electron queues by kind
and books ``weight * (E + 2 m_e c^2 [positrons])`` as emitted energy. A phase space has
no beamlets, so ``run_dij`` refuses it (pyradmc/geometry/phasespace.py:419-422).

File format (IAEA(NDS)-0484, Capote et al., Vienna 2006), as pyradmc reads it
(pyradmc/geometry/phasespace.py:1-28): an ASCII ``.IAEAheader`` of ``$KEYWORD:``
sections with ``//`` comments, and a binary ``.IAEAphsp`` of fixed-length records::

    int8 typ | float32 E | float32 x y z u v weight (stored ones only) | Nf float32 | Ni int32

``|typ|`` is the particle code (1 photon, 2 electron, 3 positron, 4 neutron, 5 proton)
and its sign is the sign of the third direction cosine ``w``, which is never stored;
``E < 0`` marks the first particle of a new primary history. A coordinate whose
``$RECORD_CONTENTS`` flag is 0 is absent from the record and constant over the file
(``$RECORD_CONSTANT``, in x, y, z, u, v, weight order). ``w = sign * sqrt(1 - u^2 - v^2)``;
``u^2 + v^2 > 1`` (rounding in the writer) renormalises ``(u, v)`` and sets ``w = 0``.

Record sampling is chunk-invariant: history ``h`` of a run with seed ``s`` draws record
``floor(U_h * n)`` of the ``n`` transportable records, where ``U_h`` is the ``h``-th
double of numpy's ``Generator(PCG64(s))`` (the stream advanced to the block's first
history, pyradmc/geometry/phasespace.py:367-382), so the records drawn do not depend on
how the run is cut into blocks, chunks or calls. Files are memory-mapped: construction
scans only the type and energy fields, and sampling gathers the drawn records.
"""

from __future__ import annotations

import itertools
import mmap
import warnings
from collections.abc import Iterator
from dataclasses import dataclass, field
from pathlib import Path
from typing import NamedTuple

import numpy as np

__all__ = [
    "DevicePhaseSpace",
    "IAEA_ELECTRON",
    "IAEA_PHOTON",
    "IAEA_POSITRON",
    "IAEAHeader",
    "IAEAPhaseSpace",
    "InMemoryPhaseSpaceSource",
    "PhaseSpaceSource",
    "PhspRecord",
    "read_iaea_header",
    "write_iaea_phase_space",
]

IAEA_PHOTON = 1
IAEA_ELECTRON = 2
IAEA_POSITRON = 3
_TRANSPORTED = (IAEA_PHOTON, IAEA_ELECTRON, IAEA_POSITRON)

# The six per-particle values that RECORD_CONTENTS can store or make constant, in file
# order. The seven flags of RECORD_CONTENTS are x, y, z, u, v, w, weight: the sixth (w)
# must be 1 and names no stored float.
COORDINATES = ("x", "y", "z", "u", "v", "weight")
_FLAG_OF_COORDINATE = (0, 1, 2, 3, 4, 6)
_W_FLAG = 5

_HEADER_SUFFIX = ".IAEAheader"
_DATA_SUFFIX = ".IAEAphsp"


class PhspRecord(NamedTuple):
    """One decoded particle (pyradmc/geometry/phasespace.py:79-99): ``energy`` > 0,
    ``(u, v, w)`` a unit vector, ``particle_type`` the IAEA code."""

    particle_type: int
    energy: float
    x: float
    y: float
    z: float
    u: float
    v: float
    w: float
    weight: float
    new_history: bool
    extra_floats: tuple[float, ...]
    extra_ints: tuple[int, ...]


@dataclass(frozen=True)
class IAEAHeader:
    """A parsed ``.IAEAheader`` (pyradmc/geometry/phasespace.py:102-113).

    ``byte_order`` is numpy's prefix ("<" for ``$BYTE_ORDER`` 1234, ">" for 4321);
    ``stored[i]`` says whether ``COORDINATES[i]`` is in every record; ``constants`` holds
    the value of each coordinate that is not."""

    byte_order: str
    record_length: int
    n_particles: int
    n_extra_floats: int
    n_extra_ints: int
    stored: tuple[bool, ...]
    constants: dict[str, float] = field(default_factory=dict)

    def record_dtype(self) -> np.dtype:
        """The numpy dtype of one record: ``typ``, ``e``, the stored coordinates, then
        ``extra_floats`` and ``extra_ints`` as sub-arrays when present."""
        o = self.byte_order
        names, formats = ["typ", "e"], [np.int8, o + "f4"]
        for name, kept in zip(COORDINATES, self.stored):
            if kept:
                names.append(name)
                formats.append(o + "f4")
        if self.n_extra_floats:
            names.append("extra_floats")
            formats.append((o + "f4", (self.n_extra_floats,)))
        if self.n_extra_ints:
            names.append("extra_ints")
            formats.append((o + "i4", (self.n_extra_ints,)))
        dt = np.dtype({"names": names, "formats": formats})
        if dt.itemsize != self.record_length:
            raise ValueError(f"record layout is {dt.itemsize} bytes, header says {self.record_length}")
        return dt


def _paths(path) -> tuple[Path, Path]:
    """(``.IAEAheader``, ``.IAEAphsp``) for a stem or either file's path."""
    p = Path(path)
    stem = p.with_suffix("") if p.suffix in (_HEADER_SUFFIX, _DATA_SUFFIX) else p
    return stem.with_suffix(_HEADER_SUFFIX), stem.with_suffix(_DATA_SUFFIX)


def _sections(text: str) -> dict[str, list[str]]:
    """``$KEYWORD:`` -> its value lines, with ``//`` comments and blank lines dropped."""
    out: dict[str, list[str]] = {}
    current = None
    for line in text.splitlines():
        s = line.strip()
        if s.startswith("$") and s.endswith(":"):
            current = out.setdefault(s[1:-1], [])
            continue
        value = s.split("//", 1)[0].strip()
        if current is not None and value:
            current.append(value)
    return out


def read_iaea_header(path) -> IAEAHeader:
    """Parse and validate a header against its own record layout and the data file
    (pyradmc/geometry/phasespace.py:138-246).

    Raises ``ValueError`` for: fewer than nine RECORD_CONTENTS entries, a w flag of 0,
    a RECORD_CONSTANT count that does not match the unstored coordinates, a BYTE_ORDER
    other than 1234/4321, a RECORD_LENGTH that disagrees with the flags, a data file
    whose size is not a whole number of records, or no particle count at all. A
    PARTICLES entry that disagrees with the data file only warns: the records on disk
    are what can be read."""
    header_path, data_path = _paths(path)
    sec = _sections(header_path.read_text())

    contents = sec.get("RECORD_CONTENTS", [])
    if len(contents) < 9:
        raise ValueError(f"RECORD_CONTENTS needs 7 flags plus Nf and Ni, got {len(contents)} entries")
    flags = [int(t) for t in contents[:7]]
    if flags[_W_FLAG] != 1:
        raise ValueError("RECORD_CONTENTS must store w (flag 6 = 1): w is rebuilt from u, v and the type sign")
    n_f, n_i = int(contents[7]), int(contents[8])
    stored = tuple(bool(flags[k]) for k in _FLAG_OF_COORDINATE)

    values = [float(t) for t in sec.get("RECORD_CONSTANT", [])]
    missing = [name for name, kept in zip(COORDINATES, stored) if not kept]
    if len(values) != len(missing):
        raise ValueError(f"RECORD_CONSTANT has {len(values)} values but {len(missing)} coordinates are constant")
    constants = dict(zip(missing, values))

    order_token = (sec.get("BYTE_ORDER") or ["1234"])[0]
    orders = {"1234": "<", "4321": ">"}
    if order_token not in orders:
        raise ValueError(f"unknown BYTE_ORDER {order_token!r}; expected 1234 or 4321")

    length = 5 + 4 * (sum(stored) + n_f + n_i)
    declared_length = sec.get("RECORD_LENGTH")
    if declared_length and int(declared_length[0]) != length:
        raise ValueError(f"RECORD_LENGTH {declared_length[0]} disagrees with the {length} bytes "
                         "implied by RECORD_CONTENTS")

    declared = int(sec["PARTICLES"][0]) if sec.get("PARTICLES") else None
    if data_path.exists():
        size = data_path.stat().st_size
        if size % length:
            raise ValueError(f"{data_path.name} size {size} is not a multiple of record length {length}")
        n = size // length
        if declared is not None and declared != n:
            warnings.warn(f"PARTICLES header says {declared} but {data_path.name} holds {n} records; "
                          "using the on-disk count", stacklevel=2)
    elif declared is None:
        raise ValueError("cannot determine particle count: no PARTICLES and no phsp file")
    else:
        n = declared
    return IAEAHeader(orders[order_token], length, n, n_f, n_i, stored, constants)


def _directions(u: np.ndarray, v: np.ndarray, negative: np.ndarray) -> tuple:
    """(u, v, w) for float32 ``u, v`` and the type-sign mask, in float32: ``w`` from the
    float64 ``1 - u^2 - v^2``, and a degenerate ``u^2 + v^2 > 1`` scaled back onto the
    unit circle with ``w = 0`` (pyradmc/geometry/phasespace.py:633-643)."""
    rr = u.astype(np.float64) ** 2 + v.astype(np.float64) ** 2
    sign = np.where(negative, -1.0, 1.0)
    w = np.where(rr <= 1.0, sign * np.sqrt(np.maximum(1.0 - rr, 0.0)), 0.0).astype(np.float32)
    over = rr > 1.0
    if np.any(over):
        scale = np.sqrt(rr[over]).astype(np.float32)
        u = u.copy()
        v = v.copy()
        u[over] /= scale
        v[over] /= scale
    return u, v, w


def _columns(recs: np.ndarray, header: IAEAHeader) -> dict[str, np.ndarray]:
    """Upload columns (float32 geometry, IAEA codes) of gathered records."""
    n = recs.shape[0]
    names = recs.dtype.names

    def coord(name):
        if name in names:
            return recs[name].astype(np.float32)
        return np.full(n, header.constants[name], dtype=np.float32)

    typ = recs["typ"].astype(np.int32)
    u, v, w = _directions(coord("u"), coord("v"), typ < 0)
    return {"particle_type": np.abs(typ), "energy": np.abs(recs["e"].astype(np.float32)),
            "x": coord("x"), "y": coord("y"), "z": coord("z"), "ux": u, "uy": v, "uz": w,
            "weight": coord("weight")}


class IAEAPhaseSpace:
    """Every record of a file in file order, for inspection and conversion
    (pyradmc/geometry/phasespace.py:323-364). Iterate inside its ``with`` block; each
    iteration starts at the first record. ``len`` is the particle count."""

    def __init__(self, path) -> None:
        self.header = read_iaea_header(path)
        self._data_path = _paths(path)[1]
        self._dtype = self.header.record_dtype()
        self._map = None
        self._file = None

    def __enter__(self) -> "IAEAPhaseSpace":
        self._file = self._data_path.open("rb")
        if self.header.n_particles:
            self._map = mmap.mmap(self._file.fileno(), 0, access=mmap.ACCESS_READ)
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def close(self) -> None:
        if self._map is not None:
            self._map.close()
            self._map = None
        if self._file is not None:
            self._file.close()
            self._file = None

    def __len__(self) -> int:
        return self.header.n_particles

    def __iter__(self) -> Iterator[PhspRecord]:
        if self._file is None:
            raise RuntimeError("iterate an IAEAPhaseSpace inside its `with` block")
        h = self.header
        block = 65536
        for start in range(0, h.n_particles, block):
            count = min(block, h.n_particles - start)
            recs = np.frombuffer(self._map, dtype=self._dtype, count=count,
                                 offset=start * h.record_length).copy()
            for r in recs:
                yield _record(r, h)

    def columns(self, start: int = 0, stop: int | None = None) -> dict[str, np.ndarray]:
        """Records ``[start, stop)`` as upload columns plus ``new_history`` (bool)."""
        if self._file is None:
            raise RuntimeError("read an IAEAPhaseSpace inside its `with` block")
        h = self.header
        stop = h.n_particles if stop is None else min(int(stop), h.n_particles)
        start = max(0, int(start))
        count = max(0, stop - start)
        if count == 0:
            recs = np.zeros(0, dtype=self._dtype)
        else:
            recs = np.frombuffer(self._map, dtype=self._dtype, count=count,
                                 offset=start * h.record_length)
        out = _columns(recs, h)
        out["new_history"] = recs["e"] < 0
        return out


def _record(r, h: IAEAHeader) -> PhspRecord:
    """One structured record as a ``PhspRecord``, direction rebuilt in float64
    (pyradmc/geometry/phasespace.py:249-256, 286-320)."""
    names = r.dtype.names
    val = {c: float(r[c]) if c in names else float(h.constants[c]) for c in COORDINATES}
    typ = int(r["typ"])
    e = float(r["e"])
    u, v = val["u"], val["v"]
    rr = u * u + v * v
    if rr <= 1.0:
        w = (-1.0 if typ < 0 else 1.0) * (1.0 - rr) ** 0.5
    else:
        s = rr ** 0.5
        u, v, w = u / s, v / s, 0.0
    ef = tuple(float(x) for x in r["extra_floats"]) if "extra_floats" in names else ()
    ei = tuple(int(x) for x in r["extra_ints"]) if "extra_ints" in names else ()
    return PhspRecord(abs(typ), abs(e), val["x"], val["y"], val["z"], u, v, w, val["weight"],
                      e < 0.0, ef, ei)


def _draw(seed: int, first_history: int, n: int, population: int) -> np.ndarray:
    """Population ranks drawn by histories ``[first_history, first_history + n)``:
    ``floor(U * population)`` of the stream's doubles from position ``first_history``,
    clamped below ``population`` (pyradmc/geometry/phasespace.py:367-382)."""
    bits = np.random.PCG64(int(seed))
    bits.advance(int(first_history))
    k = (np.random.Generator(bits).random(int(n)) * population).astype(np.int64)
    np.minimum(k, population - 1, out=k)
    return k


class _ReuseWarning:
    """Warns once when a run draws more histories than the population holds: records
    then repeat, and the batch sigmas miss the phase space's own (latent) variance
    (pyradmc/geometry/phasespace.py:385-418)."""

    def __init__(self, population: int) -> None:
        self.population = population
        self.drawn = 0
        self.warned = False

    def note(self, up_to: int) -> None:
        self.drawn = max(self.drawn, int(up_to))
        if not self.warned and self.drawn > self.population:
            self.warned = True
            warnings.warn(
                f"drawing more histories ({self.drawn}) than the phase space holds particles "
                f"({self.population}): records repeat, and the per-batch sigmas do not include "
                "the finite-phase-space latent variance", stacklevel=3)


class PhaseSpaceSource:
    """A source sampling records of an IAEA phase-space file with replacement
    (pyradmc/geometry/phasespace.py:421-659).

    Records of unsupported particles (neutrons, protons) leave the sampled population
    with a warning, or make construction raise with ``skip_unsupported=False``.
    ``len`` is the number of transportable records; ``max_energy`` sizes the tables."""

    def __init__(self, path, *, skip_unsupported: bool = True) -> None:
        self.header = read_iaea_header(path)
        self._data_path = _paths(path)[1]
        n = self.header.n_particles
        if n == 0:
            raise ValueError(f"phase-space file {path} contains no particles")
        self._dtype = self.header.record_dtype()
        self._file = self._data_path.open("rb")
        self._map = mmap.mmap(self._file.fileno(), 0, access=mmap.ACCESS_READ)
        scan = np.dtype({"names": ["typ", "e"], "formats": [np.int8, self.header.byte_order + "f4"],
                         "offsets": [0, 1], "itemsize": self.header.record_length})
        view = np.frombuffer(self._map, dtype=scan, count=n)
        code = np.abs(view["typ"].astype(np.int16))
        self._max_energy = float(np.max(np.abs(view["e"].astype(np.float64))))
        del view
        bad = np.flatnonzero(~np.isin(code, _TRANSPORTED))
        if bad.size:
            if not skip_unsupported:
                raise ValueError(
                    f"unsupported particle type {int(code[bad[0]])} in {path}; this engine transports "
                    "only photons, electrons and positrons (pass skip_unsupported=True to drop them)")
            warnings.warn(f"{bad.size} of {n} records in {self._data_path.name} are unsupported particle "
                          "types (not photon/electron/positron); dropping them from the sampled population",
                          stacklevel=2)
        # Rank k of the transportable records sits at file position k + (number of
        # unsupported records before it) = k + #{j : bad[j] - j <= k}.
        self._skip_key = (bad - np.arange(bad.size)).astype(np.int64)
        self._n_valid = n - int(bad.size)
        if self._n_valid == 0:
            raise ValueError(f"phase-space file {path} has no transportable particles")
        self._reuse = _ReuseWarning(self._n_valid)

    @property
    def max_energy(self) -> float:
        """The highest record energy in the file, MeV (transportable or not)."""
        return self._max_energy

    def __len__(self) -> int:
        return self._n_valid

    def positions(self, ranks: np.ndarray) -> np.ndarray:
        """File positions of the given ranks among the transportable records."""
        ranks = np.asarray(ranks, dtype=np.int64)
        if self._skip_key.size == 0:
            return ranks
        return ranks + np.searchsorted(self._skip_key, ranks, side="right")

    def sample_batch(self, seed: int, history_offset: int, n: int) -> dict[str, np.ndarray]:
        """Upload columns of histories ``[history_offset, history_offset + n)``:
        ``particle_type`` (IAEA codes, int32), ``energy``, ``x y z``, ``ux uy uz`` and
        ``weight`` (float32)."""
        if self._map is None:
            raise RuntimeError("the phase-space source is closed")
        self._reuse.note(history_offset + n)
        pos = self.positions(_draw(seed, history_offset, n, self._n_valid))
        recs = np.frombuffer(self._map, dtype=self._dtype, count=self.header.n_particles)[pos]
        return _columns(recs, self.header)

    def population_blocks(self, block: int = 1 << 22) -> Iterator[dict[str, np.ndarray]]:
        """The transportable records in file order as upload columns, ``block`` file
        records at a time (what a device-resident copy of the population uploads)."""
        if self._map is None:
            raise RuntimeError("the phase-space source is closed")
        h = self.header
        for start in range(0, h.n_particles, int(block)):
            count = min(int(block), h.n_particles - start)
            recs = np.frombuffer(self._map, dtype=self._dtype, count=count, offset=start * h.record_length)
            if self._skip_key.size:
                recs = recs[np.isin(np.abs(recs["typ"].astype(np.int16)), _TRANSPORTED)]
            else:
                recs = recs.copy()
            if recs.size:
                yield _columns(recs, h)

    def close(self) -> None:
        """Release the map and the file (idempotent)."""
        if getattr(self, "_map", None) is not None:
            self._map.close()
            self._map = None
        if getattr(self, "_file", None) is not None:
            self._file.close()
            self._file = None

    def __enter__(self) -> "PhaseSpaceSource":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:
        self.close()


class InMemoryPhaseSpaceSource:
    """A phase space held as columns, e.g. a treatment-head pre-solve's exit plane
    (pyradmc/geometry/phasespace.py:662-769): the sampling contract of
    ``PhaseSpaceSource`` over arrays instead of a file.

    Raises ``ValueError`` for an empty population, columns of another length than
    ``particle_type``, codes other than 1/2/3, a non-positive energy, a direction off
    the unit sphere by 1e-5 or more in ``|u|^2``, or a negative or non-finite weight."""

    def __init__(self, particle_type, energy, x, y, z, ux, uy, uz, weight) -> None:
        code = np.asarray(particle_type, dtype=np.int32).reshape(-1)
        if code.size == 0:
            raise ValueError("an in-memory phase space needs at least one particle")
        cols = {name: np.asarray(a, dtype=np.float64) for name, a in
                (("energy", energy), ("x", x), ("y", y), ("z", z), ("ux", ux), ("uy", uy),
                 ("uz", uz), ("weight", weight))}
        for name, a in cols.items():
            if a.shape != code.shape:
                raise ValueError(f"column {name!r} length {a.shape} != particle_type length {code.shape}")
        if not np.all(np.isin(code, _TRANSPORTED)):
            raise ValueError("particle_type must use the supported IAEA codes 1/2/3")
        if not np.all(cols["energy"] > 0.0):
            raise ValueError("every particle needs a positive energy")
        norm2 = cols["ux"] ** 2 + cols["uy"] ** 2 + cols["uz"] ** 2
        if not np.all(np.abs(norm2 - 1.0) < 1.0e-5):
            raise ValueError("directions must be unit vectors")
        w = cols["weight"]
        if not np.all(np.isfinite(w)) or np.any(w < 0.0):
            raise ValueError("weights must be finite and non-negative")
        self._code = code
        self._cols = cols
        self._reuse = _ReuseWarning(int(code.size))

    def __len__(self) -> int:
        return int(self._code.size)

    @property
    def max_energy(self) -> float:
        return float(self._cols["energy"].max())

    def columns(self) -> dict[str, np.ndarray]:
        """A copy of the population: ``particle_type`` plus the float64 columns."""
        out = {"particle_type": self._code.copy()}
        out.update({k: v.copy() for k, v in self._cols.items()})
        return out

    def sample_batch(self, seed: int, history_offset: int, n: int) -> dict[str, np.ndarray]:
        """Upload columns of histories ``[history_offset, history_offset + n)``."""
        self._reuse.note(history_offset + n)
        k = _draw(seed, history_offset, n, len(self))
        out = {"particle_type": self._code[k]}
        out.update({name: a[k].astype(np.float32) for name, a in self._cols.items()})
        return out

    def population_blocks(self, block: int = 1 << 22) -> Iterator[dict[str, np.ndarray]]:
        """The population in order as upload columns, ``block`` rows at a time."""
        for start in range(0, len(self), int(block)):
            sl = slice(start, start + int(block))
            out = {"particle_type": self._code[sl]}
            out.update({name: a[sl].astype(np.float32) for name, a in self._cols.items()})
            yield out


_tokens = itertools.count(1)


class DevicePhaseSpace:
    """A phase space whose transportable records live in device memory: ``HipEngine.run``
    uploads the population once per engine (36 bytes a record) and the generator kernel
    draws each history's record itself, so no record is sampled, decoded or copied on
    the host per run. Plays the role of pyradmc's ``DevicePhaseSpace``
    (pyradmc/backends/warp/presolve.py:710-742), here for any phase-space source.

    Sampling law: history ``h`` of seed ``s`` draws row ``floor(U * n)`` of the ``n``
    records, ``U`` from the first 64 bits of the history's own random stream. This is
    another stream than the host route's PCG64 draws, so a resident run and a
    host-sampled run of the same seed draw different records: both estimate the same
    dose and agree statistically, never bit for bit (as in the reference, whose device
    route draws with its own salted stream). The population must not change while an
    engine holds it; build a new ``DevicePhaseSpace`` for new records."""

    def __init__(self, source, *, block: int = 1 << 22) -> None:
        if not callable(getattr(source, "population_blocks", None)):
            raise TypeError(f"{type(source).__name__} has no population_blocks(); "
                            "expected a PhaseSpaceSource or an InMemoryPhaseSpaceSource")
        if block < 1:
            raise ValueError("block must be positive")
        self.source = source
        self.block = int(block)
        self.token = next(_tokens)
        self._n = len(source)
        self._max_energy = float(source.max_energy)
        self._reuse = _ReuseWarning(self._n)

    @property
    def max_energy(self) -> float:
        return self._max_energy

    def __len__(self) -> int:
        return self._n

    def population_blocks(self) -> Iterator[dict[str, np.ndarray]]:
        return self.source.population_blocks(self.block)

    def note_histories(self, up_to: int) -> None:
        """Warn once when a run draws more histories than the population holds."""
        self._reuse.note(up_to)

    def to_phase_space(self) -> InMemoryPhaseSpaceSource:
        """The population as an in-memory source (float32 values, in order)."""
        blocks = list(self.population_blocks())
        cols = {k: np.concatenate([b[k] for b in blocks]) for k in blocks[0]}
        return InMemoryPhaseSpaceSource(**cols)


def write_iaea_phase_space(path, columns: dict, *, new_history=None, byte_order: int = 1234,
                           constants: dict | None = None, original_histories: int | None = None) -> Path:
    """Write ``columns`` (``particle_type`` codes, ``energy``, ``x y z``, ``ux uy uz``,
    ``weight``) as an IAEA phase space; returns the data file's path.

    ``constants`` maps coordinates of ``COORDINATES`` to a value that every particle
    shares: they are left out of the records and written to ``$RECORD_CONSTANT``.
    ``new_history`` (bool per particle, default all True) sets the energy sign; the sign
    of ``uz`` sets the type sign. No extra floats or ints are written."""
    if byte_order not in (1234, 4321):
        raise ValueError("byte_order must be 1234 or 4321")
    constants = dict(constants or {})
    unknown = set(constants) - set(COORDINATES)
    if unknown:
        raise ValueError(f"constants for unknown coordinates {sorted(unknown)}")
    code = np.asarray(columns["particle_type"], dtype=np.int64).reshape(-1)
    n = code.size
    if np.any((code < 1) | (code > 127)):
        raise ValueError("particle_type codes must be in 1..127")
    source_name = {"x": "x", "y": "y", "z": "z", "u": "ux", "v": "uy", "weight": "weight"}
    stored = tuple(name not in constants for name in COORDINATES)
    header = IAEAHeader("<" if byte_order == 1234 else ">", 5 + 4 * sum(stored), n, 0, 0, stored,
                        {k: float(constants[k]) for k in COORDINATES if k in constants})
    recs = np.zeros(n, dtype=header.record_dtype())
    uz = np.asarray(columns["uz"], dtype=np.float64).reshape(-1)
    recs["typ"] = np.where(uz < 0, -code, code).astype(np.int8)
    fresh = np.ones(n, bool) if new_history is None else np.asarray(new_history, bool).reshape(-1)
    e = np.abs(np.asarray(columns["energy"], dtype=np.float32).reshape(-1))
    recs["e"] = np.where(fresh, -e, e)
    for name, kept in zip(COORDINATES, stored):
        if kept:
            recs[name] = np.asarray(columns[source_name[name]], dtype=np.float32).reshape(-1)
    header_path, data_path = _paths(path)
    data_path.parent.mkdir(parents=True, exist_ok=True)
    recs.tofile(data_path)
    flags = [int(k) for k in stored[:5]] + [1, int(stored[5])]
    labels = ("X", "Y", "Z", "U", "V", "W", "Weight")
    lines = ["$IAEA_INDEX:", "1000", "", "$TITLE:", "pyradmc_amd phase space", "",
             "$FILE_TYPE:", "0", "", "$CHECKSUM:", str(header.record_length * n), "",
             "$RECORD_CONTENTS:"]
    lines += [f"    {f}     // {lab} is stored ?" for f, lab in zip(flags, labels)]
    lines += ["    0     // Extra floats stored ?", "    0     // Extra longs stored ?", "",
              "$RECORD_CONSTANT:"]
    lines += [f"    {header.constants[k]!r}     // {k} constant" for k in COORDINATES if k in header.constants]
    lines += ["", "$RECORD_LENGTH:", str(header.record_length), "", "$BYTE_ORDER:", str(byte_order), "",
              "$ORIG_HISTORIES:", str(int(original_histories) if original_histories is not None
                                     else int(np.count_nonzero(fresh))), "",
              "$PARTICLES:", str(n), ""]
    for label, code_value in (("PHOTONS", 1), ("ELECTRONS", 2), ("POSITRONS", 3)):
        lines += [f"${label}:", str(int(np.count_nonzero(code == code_value))), ""]
    header_path.write_text("\n".join(lines))
    return data_path
