"""Tabulated, material-resolved cross sections: the multi-material data source.

Restates the reference's compiled-table source (``TabulatedData``,
``pyradmc/data/tabulated/model.py:18-48``; ``TabulatedCrossSections``,
``pyradmc/data/tabulated/source.py:55-241``) over the same container fields: one
geometric photon grid and one geometric electron grid shared by every material, and
per-material rows in registry order. Queries interpolate linearly in value over
log E with flat clamping at the grid edges, the basis the flattened kernel tables use
(``pyradmc/data/tables.py:83-133``). Quantities restricted at the delta-ray cut are
tabulated at one cut, and a query at another raises.

The derived quantities the interface defines on top of the queries follow the
reference's interface (``pyradmc/data/interface.py:239-409``):

* the elastic screening from the material's elemental composition (Moliere);
* the restricted range by the trapezoid rule over 1/S on a 4096-node geometric grid
  from the cut to 64 MeV, and its inverse by interpolation in log E;
* the coherent cumulative: the compiled form factor resampled onto the kernel's
  momentum-transfer grid, or the flat (Thomson) one, ``x^2 / 2``.

The source answers in bulk (whole energy arrays per material), which is what
:func:`pyradmc_amd.data.tables.build_cross_section_tables` flattens; the scalar query
surface of the reference interface is there too. The EPICS parsers and the
precompiler that fill a ``TabulatedData`` from the evaluated libraries are not part
of this package (they need the EPDL/EEDL files); any ``TabulatedData`` - compiled
elsewhere, or synthetic as in the tests - drives the same transport.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from pyradmc_amd.data.analytic import RANGE_GRID_E_MAX, RANGE_GRID_POINTS, moliere_screening
from pyradmc_amd.data.materials import MATERIALS, REGISTRY_NAMES, WATER

COMPTON, PHOTOELECTRIC, PAIR, RAYLEIGH = 0, 1, 2, 3
PHOTON_PROCESSES = (COMPTON, PHOTOELECTRIC, PAIR, RAYLEIGH)
ELECTRON_QUANTITIES = ("restricted_stopping", "radiative_stopping", "moller", "csda_range",
                       "scattering_power")


@dataclass(frozen=True)
class TabulatedData:
    """Material-resolved data on two shared log-energy grids (the reference's compiled
    container, field for field). ``mu_over_rho`` maps a photon process (0 Compton,
    1 photoelectric, 2 pair, 3 Rayleigh) to ``[n_materials][n_photon]`` in cm^2/g;
    the electron arrays are ``[n_materials][n_electron]`` (MeV cm^2/g, cm^2/g, g/cm^2,
    rad^2 cm^2/g), the restricted ones at ``delta_cut``. ``materials`` names the rows,
    a prefix of the registry."""

    photon_energies: np.ndarray
    mu_over_rho: dict
    electron_energies: np.ndarray
    restricted_stopping: np.ndarray
    radiative_stopping: np.ndarray
    moller: np.ndarray
    csda_range: np.ndarray
    scattering_power: np.ndarray
    delta_cut: float
    materials: tuple
    provenance: str
    coherent_x: np.ndarray | None = None
    coherent_cumulative: np.ndarray | None = None


def _geometric(energies: np.ndarray) -> tuple[float, float, int]:
    """(log of the first node, inverse log spacing, node count) of a geometric grid."""
    n = int(energies.shape[0])
    if n < 2:
        raise ValueError("energy grid needs at least two nodes")
    log_e = np.log(energies)
    step = np.diff(log_e)
    if not np.allclose(step, step[0], rtol=1e-6):
        raise ValueError("energy grid is not geometric (uniform in log energy)")
    return float(log_e[0]), float((n - 1) / (log_e[-1] - log_e[0])), n


def _loglinear(row: np.ndarray, log_min: float, inv_dlog: float, n: int, energy) -> np.ndarray:
    """Linear in value over log E, clamped flat outside the grid (array queries)."""
    t = (np.log(np.asarray(energy, dtype=np.float64)) - log_min) * inv_dlog
    t = np.clip(t, 0.0, float(n - 1))
    i = np.minimum(t.astype(np.int64), n - 2)
    w = t - i
    return row[i] * (1.0 - w) + row[i + 1] * w


class TabulatedCrossSections:
    """Interpolating source over a :class:`TabulatedData`; drop-in for the reference's
    ``TabulatedCrossSections(data, geometry_densities)``. ``geometry_densities`` are the
    ``(material, maximum density)`` pairs present in the geometry; the Woodcock
    majorant is taken over them."""

    def __init__(self, data: TabulatedData, geometry_densities: tuple[tuple[int, float], ...] = ((WATER, 1.0),)):
        names = tuple(data.materials)
        if names != REGISTRY_NAMES[: len(names)]:
            raise ValueError(f"compiled materials {names} do not match the registry {REGISTRY_NAMES}")
        for material, density in geometry_densities:
            if not 0 <= material < len(names):
                raise ValueError(f"material index {material} not in the compiled table")
            if density <= 0.0:
                raise ValueError(f"non-positive density {density} for material {material}")
        self._data = data
        self._geometry_densities = tuple((int(m), float(d)) for m, d in geometry_densities)
        self._pe = np.asarray(data.photon_energies, dtype=np.float64)
        self._ee = np.asarray(data.electron_energies, dtype=np.float64)
        self._p_grid = _geometric(self._pe)
        self._e_grid = _geometric(self._ee)
        self._log_ee = np.log(self._ee)
        self._mu = {int(p): np.asarray(a, dtype=np.float64) for p, a in data.mu_over_rho.items()}
        self._electron = {q: np.asarray(getattr(data, q), dtype=np.float64) for q in ELECTRON_QUANTITIES}
        n_mat = len(names)
        for p, a in self._mu.items():
            if p not in PHOTON_PROCESSES or a.shape != (n_mat, self._pe.size):
                raise ValueError(f"photon table {p} has shape {a.shape}, expected {(n_mat, self._pe.size)}")
        for q, a in self._electron.items():
            if a.shape != (n_mat, self._ee.size):
                raise ValueError(f"electron table {q} has shape {a.shape}, expected {(n_mat, self._ee.size)}")
        self._range_grids: dict[tuple[int, float], tuple[np.ndarray, np.ndarray]] = {}

    # -- identity ------------------------------------------------------------------

    @property
    def n_materials(self) -> int:
        return len(self._data.materials)

    @property
    def geometry_densities(self) -> tuple[tuple[int, float], ...]:
        return self._geometry_densities

    @property
    def provenance(self) -> str:
        return self._data.provenance

    @property
    def delta_cut(self) -> float:
        return float(self._data.delta_cut)

    def _check_material(self, material: int) -> None:
        if not 0 <= material < self.n_materials:
            raise ValueError(f"material index {material} not in the compiled table")

    def _check_cut(self, delta_cut: float) -> None:
        if not math.isclose(delta_cut, self._data.delta_cut, rel_tol=1e-9, abs_tol=0.0):
            raise ValueError(
                f"tables were compiled at delta_cut={self._data.delta_cut} MeV but the transport "
                f"asked for {delta_cut} MeV; compile them at the run's ECUT")

    # -- bulk evaluation (what the table builder uses) -----------------------------

    def mu_channels(self, energy: np.ndarray, material: int) -> tuple[np.ndarray, ...]:
        """(compton, photo, pair, rayleigh) in cm^2/g; a channel the data lacks is 0."""
        self._check_material(material)
        e = np.asarray(energy, dtype=np.float64)
        out = []
        for p in PHOTON_PROCESSES:
            table = self._mu.get(p)
            out.append(np.zeros_like(e) if table is None else _loglinear(table[material], *self._p_grid, e))
        return tuple(out)

    def mu_total_bulk(self, energy: np.ndarray, material: int) -> np.ndarray:
        c = self.mu_channels(energy, material)
        return c[0] + c[1] + c[2] + c[3]

    def majorant_bulk(self, energy: np.ndarray) -> np.ndarray:
        e = np.asarray(energy, dtype=np.float64)
        return np.max([d * self.mu_total_bulk(e, m) for m, d in self._geometry_densities], axis=0)

    def electron_bulk(self, energy: np.ndarray, material: int, ecut: float) -> dict[str, np.ndarray]:
        self._check_material(material)
        self._check_cut(ecut)
        e = np.asarray(energy, dtype=np.float64)
        look = {q: _loglinear(self._electron[q][material], *self._e_grid, e) for q in ELECTRON_QUANTITIES}
        return {
            "stopping_restricted": look["restricted_stopping"],
            "stopping_radiative": look["radiative_stopping"],
            "moller": look["moller"],
            "csda_range": look["csda_range"],
            "scattering_power": look["scattering_power"],
            "log_eta": np.log(moliere_screening(MATERIALS[material].composition, e)),
        }

    def restricted_grid(self, material: int, delta_cut: float) -> tuple[np.ndarray, np.ndarray]:
        """(log E nodes, restricted range) from the cut to 64 MeV: trapezoid over 1/S."""
        key = (int(material), float(delta_cut))
        grid = self._range_grids.get(key)
        if grid is None:
            energies = np.geomspace(delta_cut, RANGE_GRID_E_MAX, RANGE_GRID_POINTS)
            inv_s = 1.0 / self.electron_bulk(energies, material, delta_cut)["stopping_restricted"]
            cumulative = np.concatenate(([0.0], np.cumsum(np.diff(energies) * 0.5 * (inv_s[1:] + inv_s[:-1]))))
            grid = (np.log(energies), cumulative)
            self._range_grids[key] = grid
        return grid

    def restricted_range_bulk(self, energy: np.ndarray, material: int, delta_cut: float) -> np.ndarray:
        log_e, cumulative = self.restricted_grid(material, delta_cut)
        e = np.clip(np.asarray(energy, dtype=np.float64), delta_cut, RANGE_GRID_E_MAX)
        return np.interp(np.log(e), log_e, cumulative)

    def coherent_cumulative(self, x_grid: np.ndarray, material: int) -> np.ndarray:
        self._check_material(material)
        x = np.asarray(x_grid, dtype=np.float64)
        if self._data.coherent_x is None or self._data.coherent_cumulative is None:
            return 0.5 * x * x
        return np.interp(x, np.asarray(self._data.coherent_x, dtype=np.float64),
                         np.asarray(self._data.coherent_cumulative, dtype=np.float64)[material])

    # -- scalar host API -------------------------------------------------------------

    def mu_over_rho(self, energy: float, material: int, process: int) -> float:
        if energy <= 0.0:
            raise ValueError(f"non-positive photon energy {energy} MeV")
        if process not in PHOTON_PROCESSES:
            raise ValueError(f"unknown photon process {process}")
        return float(self.mu_channels(np.array([energy]), material)[process][0])

    def mu_over_rho_total(self, energy: float, material: int) -> float:
        return float(self.mu_total_bulk(np.array([energy]), material)[0])

    def majorant(self, energy: float) -> float:
        return float(self.majorant_bulk(np.array([energy]))[0])

    def _electron_scalar(self, name: str, energy: float, material: int) -> float:
        if energy <= 0.0:
            raise ValueError(f"non-positive electron energy {energy} MeV")
        self._check_material(material)
        return float(_loglinear(self._electron[name][material], *self._e_grid, np.array([energy]))[0])

    def restricted_stopping_power(self, energy: float, material: int, delta_cut: float) -> float:
        self._check_cut(delta_cut)
        return self._electron_scalar("restricted_stopping", energy, material)

    def radiative_stopping_power(self, energy: float, material: int) -> float:
        return self._electron_scalar("radiative_stopping", energy, material)

    def moller_cross_section(self, energy: float, material: int, delta_cut: float) -> float:
        self._check_cut(delta_cut)
        return self._electron_scalar("moller", energy, material)

    def csda_range(self, energy: float, material: int) -> float:
        return self._electron_scalar("csda_range", energy, material)

    def scattering_power(self, energy: float, material: int, delta_cut: float) -> float:
        self._check_cut(delta_cut)
        return self._electron_scalar("scattering_power", energy, material)

    def elastic_screening(self, energy: float, material: int) -> float:
        self._check_material(material)
        return float(moliere_screening(MATERIALS[material].composition, np.array([energy]))[0])

    def restricted_range(self, energy: float, material: int, delta_cut: float) -> float:
        self._check_material(material)
        return float(self.restricted_range_bulk(np.array([energy]), material, delta_cut)[0])

    def energy_after_mass_path(self, energy: float, material: int, delta_cut: float, mass_path: float) -> float:
        """r^-1(r(E) - mass_path), clamped to [delta_cut, energy]."""
        log_e, cumulative = self.restricted_grid(material, delta_cut)
        remaining = self.restricted_range(energy, material, delta_cut) - mass_path
        if remaining <= 0.0:
            return float(delta_cut)
        e_end = float(np.exp(np.interp(remaining, cumulative, log_e)))
        return min(max(e_end, delta_cut), energy)

    # -- flattening -------------------------------------------------------------------

    def build_tables(self, ecut: float, pcut: float, e_max: float, n_points: int | None = None):
        from pyradmc_amd.data.tables import TABLE_POINTS, build_cross_section_tables

        return build_cross_section_tables(self, ecut=ecut, pcut=pcut, e_max=e_max,
                                          n_points=TABLE_POINTS if n_points is None else n_points)
