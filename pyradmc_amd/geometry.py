"""Voxel grid, photon spectrum, and the primary sources the engine generates in-kernel.

Drop-in restatements of the reference's plugin types (``VoxelGrid``,
``pyradmc/geometry/grid.py:226-337``; ``Spectrum``, ``pyradmc/geometry/spectrum.py:
35-143``; the five beam sources, ``pyradmc/geometry/source.py:276-726``): same
constructor arguments, validation, and attributes, so user code written against the
reference builds them unchanged. The engine reads their attributes and generates
primaries on the device; ``emit`` is not needed by the engine, it is kept so a
source stays usable by host code that samples one primary at a time.

:class:`SpectralParallelBeamSource` is an addition: a parallel rectangular field
with energies drawn from a :class:`Spectrum` (four uniforms: two for the energy, then
x and y). The reference has no source that is both parallel and spectral.
"""

from __future__ import annotations

import math
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np

from pyradmc_amd.data.materials import N_REGISTRY, WATER

# ----------------------------------------------------------------------------- grid


@dataclass(frozen=True)
class VoxelGrid:
    """Axis-aligned voxels, half-open (lower faces inside), density g/cm^3, material index."""

    shape: tuple[int, int, int]
    spacing: tuple[float, float, float]
    density: np.ndarray
    material: np.ndarray
    origin: tuple[float, float, float] = (0.0, 0.0, 0.0)

    def __post_init__(self) -> None:
        object.__setattr__(self, "shape", tuple(int(n) for n in self.shape))
        object.__setattr__(self, "spacing", tuple(float(s) for s in self.spacing))
        object.__setattr__(self, "origin", tuple(float(o) for o in self.origin))
        if any(n < 1 for n in self.shape):
            raise ValueError(f"empty grid shape {self.shape}")
        if any(s <= 0.0 for s in self.spacing):
            raise ValueError(f"non-positive spacing {self.spacing}")
        if self.density.shape != self.shape:
            raise ValueError(f"density shape {self.density.shape} != grid shape {self.shape}")
        if self.material.shape != self.shape:
            raise ValueError(f"material shape {self.material.shape} != grid shape {self.shape}")
        if np.any(self.density <= 0.0):
            raise ValueError("non-positive voxel density; vacuum is not supported")
        if np.any((self.material < 0) | (self.material >= N_REGISTRY)):
            raise ValueError("material index outside the registry")

    @classmethod
    def uniform_water(cls, shape, spacing, origin=(0.0, 0.0, 0.0)) -> "VoxelGrid":
        return cls(
            shape=tuple(shape),
            spacing=tuple(spacing),
            origin=tuple(origin),
            density=np.ones(tuple(shape), dtype=np.float64),
            material=np.full(tuple(shape), WATER, dtype=np.int32),
        )

    @property
    def voxel_volume(self) -> float:
        return self.spacing[0] * self.spacing[1] * self.spacing[2]

    @property
    def upper_corner(self) -> tuple[float, float, float]:
        return tuple(self.origin[a] + self.shape[a] * self.spacing[a] for a in range(3))

    @property
    def n_voxels(self) -> int:
        return self.shape[0] * self.shape[1] * self.shape[2]

    def contains(self, x: float, y: float, z: float) -> bool:
        hi = self.upper_corner
        lo = self.origin
        return lo[0] <= x < hi[0] and lo[1] <= y < hi[1] and lo[2] <= z < hi[2]

    def max_density_by_material(self) -> tuple[tuple[int, float], ...]:
        pairs = []
        for material in np.unique(self.material):
            rho_max = float(self.density[self.material == material].max())
            pairs.append((int(material), rho_max))
        return tuple(pairs)


# ------------------------------------------------------------------------- spectrum


class Spectrum:
    """Photon-number histogram sampled by CDF inversion (two uniforms per draw)."""

    def __init__(self, edges, weights, convention: str = "number") -> None:
        e = np.asarray(edges, dtype=np.float64)
        w = np.asarray(weights, dtype=np.float64)
        if e.ndim != 1 or e.size < 2:
            raise ValueError(f"a spectrum needs at least two edges, got shape {e.shape}")
        if not np.all(np.isfinite(e)) or not np.all(np.isfinite(w)):
            raise ValueError("spectrum edges and weights must be finite")
        if e[0] <= 0.0:
            raise ValueError(f"the first edge must be positive, got {e[0]} MeV")
        if not np.all(np.diff(e) > 0.0):
            raise ValueError("spectrum edges must be strictly increasing")
        if w.shape != (e.size - 1,):
            raise ValueError(f"need one weight per bin: {e.size - 1} bins but {w.size} weights")
        if np.any(w < 0.0):
            raise ValueError("spectrum weights must be non-negative")
        if convention not in ("number", "energy_fluence"):
            raise ValueError(f"unknown convention {convention!r}")
        if convention == "energy_fluence":
            w = w / (0.5 * (e[:-1] + e[1:]))
        total = float(w.sum())
        if total <= 0.0:
            raise ValueError("spectrum weights must include at least one positive value")
        self._edges = e
        self._p = w / total
        self._cdf = np.cumsum(self._p)
        for a in (self._edges, self._p, self._cdf):
            a.flags.writeable = False

    @property
    def max_energy(self) -> float:
        return float(self._edges[-1])

    @property
    def edges(self) -> np.ndarray:
        return self._edges

    @property
    def bin_probabilities(self) -> np.ndarray:
        return self._p

    @property
    def cdf(self) -> np.ndarray:
        return self._cdf

    @property
    def mean_energy(self) -> float:
        return float(np.sum(self._p * 0.5 * (self._edges[:-1] + self._edges[1:])))

    def sample_energies(self, u_bin, u_within) -> np.ndarray:
        k = np.minimum(np.searchsorted(self._cdf, u_bin), self._p.size - 1)
        lo = self._edges[k]
        return np.asarray(lo + (self._edges[k + 1] - lo) * u_within, dtype=np.float64)


def bremsstrahlung_6mv_like(n_bins: int = 60, e_lo: float = 0.1, e_hi: float = 6.0,
                            temperature: float = 1.8) -> Spectrum:
    """The "6 MV-like" spectrum of the heterogeneous-phantom and Dij workloads:
    ``n_bins`` equal bins on [0.1, 6] MeV, photon-number weight ``E exp(-E/1.8)`` at
    bin midpoints."""
    edges = np.linspace(e_lo, e_hi, n_bins + 1)
    mid = 0.5 * (edges[:-1] + edges[1:])
    return Spectrum(edges, mid * np.exp(-mid / temperature), convention="number")


# -------------------------------------------------------------------------- sources


class Source:
    """Open-field source base (``run``).

    The built-in types have in-kernel generators. Any other source (a subclass of a
    built-in type too) is host pre-sampled, as in the reference
    (``pyradmc/geometry/source.py:115-146``): it implements
    ``sample_batch(seed, history_offset, n)`` returning ``particle_type`` (int32 IAEA
    codes: 1 photon, 2 electron, 3 positron) and float32 columns ``energy, x, y, z,
    ux, uy, uz, weight`` for histories ``history_offset .. history_offset + n - 1``.
    History ``h`` then transports on the random stream keyed by ``h``.
    """

    @property
    def max_energy(self) -> float:  # pragma: no cover - interface
        raise NotImplementedError


class BeamletSource:
    """Beamlet-resolved source base (``run_dij``).

    Without an in-kernel generator a beamlet source is host pre-sampled through
    ``sample_beamlet_batch(seed, history_offset, n, beamlet)`` (columns as for
    ``Source.sample_batch``); ``history_offset`` is the Dij key of the batch's first
    history of that beamlet (``r`` correlated, ``offset_j + r`` independent).
    """

    @property
    def max_energy(self) -> float:  # pragma: no cover - interface
        raise NotImplementedError

    @property
    def n_beamlets(self) -> int:  # pragma: no cover - interface
        raise NotImplementedError


def _normalized(v, name: str) -> tuple[float, float, float]:
    norm = math.sqrt(sum(float(c) * float(c) for c in v))
    if norm == 0.0:
        raise ValueError(f"zero {name} vector")
    return (float(v[0]) / norm, float(v[1]) / norm, float(v[2]) / norm)


@dataclass(frozen=True)
class PencilBeamSource(Source):
    energy: float
    position: tuple[float, float, float]
    direction: tuple[float, float, float]

    def __post_init__(self) -> None:
        if self.energy <= 0.0:
            raise ValueError(f"non-positive energy {self.energy} MeV")
        object.__setattr__(self, "direction", _normalized(self.direction, "direction"))
        object.__setattr__(self, "position", tuple(float(c) for c in self.position))

    @property
    def max_energy(self) -> float:
        return float(self.energy)


@dataclass(frozen=True)
class ParallelBeamSource(Source):
    energy: float
    z: float
    x_range: tuple[float, float]
    y_range: tuple[float, float]

    def __post_init__(self) -> None:
        if self.energy <= 0.0:
            raise ValueError(f"non-positive energy {self.energy} MeV")
        if self.x_range[1] <= self.x_range[0] or self.y_range[1] <= self.y_range[0]:
            raise ValueError("empty field")

    @property
    def max_energy(self) -> float:
        return float(self.energy)


class SpectralParallelBeamSource(Source):
    """Parallel +z field over ``x_range`` x ``y_range`` at plane ``z``, energies from a
    :class:`Spectrum`; four uniforms per primary (bin, within-bin, x, y)."""

    def __init__(self, spectrum: Spectrum, z: float, x_range, y_range) -> None:
        if x_range[1] <= x_range[0] or y_range[1] <= y_range[0]:
            raise ValueError("empty field")
        self.spectrum = spectrum
        self.z = float(z)
        self.x_range = (float(x_range[0]), float(x_range[1]))
        self.y_range = (float(y_range[0]), float(y_range[1]))

    @property
    def max_energy(self) -> float:
        return self.spectrum.max_energy


@dataclass(frozen=True)
class BeamletGridSource(BeamletSource):
    """Parallel beamlet lattice along +z, x-major beamlet index ``jx * n_y + jy``."""

    energy: float
    z: float
    x_range: tuple[float, float]
    y_range: tuple[float, float]
    n_x: int
    n_y: int

    def __post_init__(self) -> None:
        if self.energy <= 0.0:
            raise ValueError(f"non-positive energy {self.energy} MeV")
        if self.x_range[1] <= self.x_range[0] or self.y_range[1] <= self.y_range[0]:
            raise ValueError("empty field")
        if self.n_x < 1 or self.n_y < 1:
            raise ValueError(f"lattice must be at least 1x1, got {self.n_x}x{self.n_y}")

    @property
    def max_energy(self) -> float:
        return float(self.energy)

    @property
    def n_beamlets(self) -> int:
        return self.n_x * self.n_y

    @staticmethod
    def _edge(bounds, i: int, n: int) -> float:
        lo, hi = bounds
        if i == 0:
            return lo
        if i == n:
            return hi
        return lo + (hi - lo) * (i / n)

    def beamlet_bounds(self, beamlet: int) -> tuple[float, float, float, float]:
        if not 0 <= beamlet < self.n_beamlets:
            raise IndexError(f"beamlet {beamlet} outside lattice of {self.n_beamlets}")
        jx, jy = divmod(beamlet, self.n_y)
        return (
            self._edge(self.x_range, jx, self.n_x),
            self._edge(self.x_range, jx + 1, self.n_x),
            self._edge(self.y_range, jy, self.n_y),
            self._edge(self.y_range, jy + 1, self.n_y),
        )


class _Fan:
    """Focal spot and rectangular apertures (engine frame, cm)."""

    def __init__(self, focal_point, centers, width_u, width_v, u_axis, v_axis) -> None:
        centers = tuple(tuple(float(c) for c in center) for center in centers)
        if len(centers) == 0:
            raise ValueError("a spectral source needs at least one aperture centre")
        if width_u < 0.0 or width_v < 0.0:
            raise ValueError(f"aperture widths must be non-negative, got ({width_u}, {width_v})")
        u = _normalized(u_axis, "u_axis")
        v = _normalized(v_axis, "v_axis")
        cross = (u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0])
        if math.sqrt(sum(c * c for c in cross)) < 1.0e-12:
            raise ValueError("aperture axes are parallel; they must span a plane")
        focal = tuple(float(c) for c in focal_point)
        for center in centers:
            if center == focal:
                raise ValueError(f"aperture centre {center} coincides with the focal point")
        self.focal_point = focal
        self.centers = centers
        self.width_u = float(width_u)
        self.width_v = float(width_v)
        self.u_axis = u
        self.v_axis = v


class SpectralBeamSource(Source):
    """Divergent polyenergetic open field through one aperture."""

    def __init__(self, spectrum: Spectrum, focal_point, center, width_u, width_v,
                 u_axis=(1.0, 0.0, 0.0), v_axis=(0.0, 1.0, 0.0)) -> None:
        self._spectrum = spectrum
        self._fan = _Fan(focal_point, (center,), width_u, width_v, u_axis, v_axis)

    @property
    def max_energy(self) -> float:
        return self._spectrum.max_energy

    @property
    def spectrum(self) -> Spectrum:
        return self._spectrum

    @property
    def focal_point(self):
        return self._fan.focal_point

    @property
    def center(self):
        return self._fan.centers[0]

    @property
    def centers(self):
        return self._fan.centers

    @property
    def width_u(self) -> float:
        return self._fan.width_u

    @property
    def width_v(self) -> float:
        return self._fan.width_v

    @property
    def u_axis(self):
        return self._fan.u_axis

    @property
    def v_axis(self):
        return self._fan.v_axis


class SpectralBeamletSource(BeamletSource):
    """Divergent polyenergetic beamlet fan; column order is the caller's centre order."""

    def __init__(self, spectrum: Spectrum, focal_point, centers: Sequence, width_u, width_v,
                 u_axis=(1.0, 0.0, 0.0), v_axis=(0.0, 1.0, 0.0)) -> None:
        self._spectrum = spectrum
        self._fan = _Fan(focal_point, tuple(centers), width_u, width_v, u_axis, v_axis)

    @property
    def max_energy(self) -> float:
        return self._spectrum.max_energy

    @property
    def n_beamlets(self) -> int:
        return len(self._fan.centers)

    @property
    def spectrum(self) -> Spectrum:
        return self._spectrum

    @property
    def focal_point(self):
        return self._fan.focal_point

    @property
    def centers(self):
        return self._fan.centers

    @property
    def width_u(self) -> float:
        return self._fan.width_u

    @property
    def width_v(self) -> float:
        return self._fan.width_v

    @property
    def u_axis(self):
        return self._fan.u_axis

    @property
    def v_axis(self):
        return self._fan.v_axis


# Floor of the Box-Muller uniform, so an exact zero draw gives a finite radius
# (pyradmc/geometry/source.py:731).
BOX_MULLER_FLOOR = 1.0e-12


class _SpotFan:
    """A divergent fan from a 2D-Gaussian focal spot (pyradmc/geometry/source.py:
    735-835): a primary starts on an aperture rectangle and travels along the line from
    a spot point ``focal + s_u u + s_v v`` (independent Gaussians of sigmas
    ``sigma_u, sigma_v`` by Box-Muller) through its start point. Six uniforms per
    history in fixed order: two for the energy, two for the offset in the rectangle, two
    for the spot (drawn at zero sigma too), so a zero-sigma source starts on the fan's
    own lines."""

    def __init__(self, fan: _Fan, sigma_u: float, sigma_v: float) -> None:
        if sigma_u < 0.0 or sigma_v < 0.0:
            raise ValueError(f"spot sigma must be non-negative, got ({sigma_u}, {sigma_v})")
        self.fan = fan
        self.sigma_u = float(sigma_u)
        self.sigma_v = float(sigma_v)

    def sample(self, spectrum: Spectrum, aperture: int, seed: int, history_offset: int, n: int) -> dict:
        """Host-sampled columns of histories ``[history_offset, history_offset + n)``:
        draws ``6h .. 6h + 5`` of ``Generator(PCG64(seed))`` for history ``h``, whatever
        the chunking and the aperture (pyradmc/geometry/source.py:795-835)."""
        bits = np.random.PCG64(int(seed))
        bits.advance(6 * int(history_offset))
        u = np.random.Generator(bits).random((int(n), 6))
        f = self.fan
        energy = spectrum.sample_energies(u[:, 0], u[:, 1])
        du = (u[:, 2] - 0.5) * f.width_u
        dv = (u[:, 3] - 0.5) * f.width_v
        radius = np.sqrt(-2.0 * np.log(np.maximum(u[:, 4], BOX_MULLER_FLOOR)))
        angle = 2.0 * np.pi * u[:, 5]
        su, sv = self.sigma_u * radius * np.cos(angle), self.sigma_v * radius * np.sin(angle)
        ua, va = np.asarray(f.u_axis), np.asarray(f.v_axis)
        start = np.asarray(f.centers[aperture]) + du[:, None] * ua + dv[:, None] * va
        spot = np.asarray(f.focal_point) + su[:, None] * ua + sv[:, None] * va
        d = start - spot
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        return {"particle_type": np.ones(int(n), np.int32), "energy": energy.astype(np.float32),
                "x": start[:, 0].astype(np.float32), "y": start[:, 1].astype(np.float32),
                "z": start[:, 2].astype(np.float32), "ux": d[:, 0].astype(np.float32),
                "uy": d[:, 1].astype(np.float32), "uz": d[:, 2].astype(np.float32),
                "weight": np.ones(int(n), np.float32)}


class GaussianSpotBeamSource(Source):
    """Photons born on a plane rectangle, aimed from a 2D-Gaussian focal spot
    (pyradmc/geometry/source.py:838-881): the head-input source. Generated in-kernel by
    the engine (six uniforms of the history's stream); ``sample_batch`` gives the same
    distribution on the host (for the head pre-solve or a host-side wrapper)."""

    def __init__(self, spectrum: Spectrum, focal_point, center, width_u, width_v, sigma_u, sigma_v,
                 u_axis=(1.0, 0.0, 0.0), v_axis=(0.0, 1.0, 0.0)) -> None:
        self._spectrum = spectrum
        self._spot = _SpotFan(_Fan(focal_point, (center,), width_u, width_v, u_axis, v_axis), sigma_u, sigma_v)

    @property
    def max_energy(self) -> float:
        return self._spectrum.max_energy

    @property
    def spectrum(self) -> Spectrum:
        return self._spectrum

    @property
    def focal_point(self):
        return self._spot.fan.focal_point

    @property
    def center(self):
        return self._spot.fan.centers[0]

    @property
    def centers(self):
        return self._spot.fan.centers

    @property
    def width_u(self) -> float:
        return self._spot.fan.width_u

    @property
    def width_v(self) -> float:
        return self._spot.fan.width_v

    @property
    def u_axis(self):
        return self._spot.fan.u_axis

    @property
    def v_axis(self):
        return self._spot.fan.v_axis

    @property
    def sigma_u(self) -> float:
        return self._spot.sigma_u

    @property
    def sigma_v(self) -> float:
        return self._spot.sigma_v

    def sample_batch(self, seed: int, history_offset: int, n: int) -> dict:
        return self._spot.sample(self._spectrum, 0, seed, history_offset, n)


class GaussianSpotBeamletSource(BeamletSource):
    """One plane rectangle per beamlet, aimed from a 2D-Gaussian focal spot
    (pyradmc/geometry/source.py:884-944); the draws never see the beamlet, so
    correlated Dij sampling replays energy, offset and spot point in every column."""

    def __init__(self, spectrum: Spectrum, focal_point, centers: Sequence, width_u, width_v, sigma_u, sigma_v,
                 u_axis=(1.0, 0.0, 0.0), v_axis=(0.0, 1.0, 0.0)) -> None:
        self._spectrum = spectrum
        self._spot = _SpotFan(_Fan(focal_point, tuple(centers), width_u, width_v, u_axis, v_axis), sigma_u,
                              sigma_v)

    @property
    def max_energy(self) -> float:
        return self._spectrum.max_energy

    @property
    def n_beamlets(self) -> int:
        return len(self._spot.fan.centers)

    @property
    def spectrum(self) -> Spectrum:
        return self._spectrum

    @property
    def focal_point(self):
        return self._spot.fan.focal_point

    @property
    def centers(self):
        return self._spot.fan.centers

    @property
    def width_u(self) -> float:
        return self._spot.fan.width_u

    @property
    def width_v(self) -> float:
        return self._spot.fan.width_v

    @property
    def u_axis(self):
        return self._spot.fan.u_axis

    @property
    def v_axis(self):
        return self._spot.fan.v_axis

    @property
    def sigma_u(self) -> float:
        return self._spot.sigma_u

    @property
    def sigma_v(self) -> float:
        return self._spot.sigma_v

    def sample_beamlet_batch(self, seed: int, history_offset: int, n: int, beamlet: int) -> dict:
        if not 0 <= beamlet < self.n_beamlets:
            raise IndexError(f"beamlet {beamlet} outside fan of {self.n_beamlets}")
        return self._spot.sample(self._spectrum, beamlet, seed, history_offset, n)
