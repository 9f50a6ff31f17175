"""Flattened cross-section tables: the kernel-side face of a data source.

Same grids, node counts and field names as the reference's ``CrossSectionTables``
(``pyradmc/data/tables.py:139-362``), so an engine can consume either: two
``n_points`` log-energy grids (photon ``[PCUT/2, e_max]``, electron
``[ECUT/2, e_max]``), a majorant with ``1 + 1e-6`` headroom, the restricted range on
the electron grid and its inverse on one shared log-range grid, and the flat
(Thomson) coherent cumulative. Sources that expose bulk array evaluation (this
package's :class:`~pyradmc_amd.data.analytic.AnalyticCrossSections` and
:class:`~pyradmc_amd.data.tabulated.TabulatedCrossSections`) flatten in one
vectorized pass per material; any other object with the scalar query API is flattened
point by point.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from pyradmc_amd.constants import (
    COHERENT_POINTS,
    MAJORANT_HEADROOM,
    RAYLEIGH_MOMENTUM_TRANSFER_PER_MEV,
    TABLE_POINTS,
)

RANGE_R_MIN_FRACTION = 1.0e-6
COHERENT_X_MIN_FRACTION = 1.0e-7

ELECTRON_FIELDS = (
    "stopping_restricted",
    "stopping_radiative",
    "moller",
    "csda_range",
    "scattering_power",
)


@dataclass(frozen=True)
class CrossSectionTables:
    """Host float64 tables; the engine casts them to float32 on upload."""

    n_points: int
    ecut: float
    pcut: float
    e_max: float
    photon_log_e_min: float
    photon_inv_dlog: float
    electron_log_e_min: float
    electron_inv_dlog: float
    mu_compton: np.ndarray
    mu_photo: np.ndarray
    mu_pair: np.ndarray
    mu_rayleigh: np.ndarray
    majorant: np.ndarray
    stopping_restricted: np.ndarray
    stopping_radiative: np.ndarray
    moller: np.ndarray
    csda_range: np.ndarray
    scattering_power: np.ndarray
    log_eta: np.ndarray
    restricted_range: np.ndarray
    energy_of_restricted_range: np.ndarray
    range_log_r_min: float
    range_inv_dlog: float
    coherent_x: np.ndarray
    coherent_cumulative: np.ndarray
    n_coherent: int


def _grid_metadata(energies: np.ndarray, n_points: int) -> tuple[float, float]:
    log_min = float(np.log(energies[0]))
    log_max = float(np.log(energies[-1]))
    return log_min, (n_points - 1) / (log_max - log_min)


def build_cross_section_tables(
    source, ecut: float, pcut: float, e_max: float, n_points: int = TABLE_POINTS
) -> CrossSectionTables:
    """Flatten ``source`` onto the photon, electron and range grids."""
    if pcut <= 0.0 or ecut <= 0.0:
        raise ValueError(f"non-positive cutoff: pcut={pcut}, ecut={ecut}")
    if e_max <= max(pcut, ecut):
        raise ValueError(f"e_max={e_max} MeV does not cover the transport range")
    if n_points < 2:
        raise ValueError(f"need at least two grid nodes, got {n_points}")

    n_mat = source.n_materials
    photon_energies = np.geomspace(0.5 * pcut, e_max, n_points)
    electron_energies = np.geomspace(0.5 * ecut, e_max, n_points)
    bulk = hasattr(source, "mu_channels")

    channels = np.empty((4, n_mat, n_points))
    for m in range(n_mat):
        if bulk:
            channels[:, m, :] = np.stack(source.mu_channels(photon_energies, m))
        else:
            for j, e in enumerate(photon_energies):
                for p in range(4):
                    channels[p, m, j] = source.mu_over_rho(float(e), m, p)
    if bulk:
        majorant = source.majorant_bulk(photon_energies) * MAJORANT_HEADROOM
    else:
        majorant = np.array([source.majorant(float(e)) for e in photon_energies]) * MAJORANT_HEADROOM

    electron = {name: np.empty((n_mat, n_points)) for name in ELECTRON_FIELDS}
    log_eta = np.empty((n_mat, n_points))
    restricted_range = np.empty((n_mat, n_points))
    for m in range(n_mat):
        if bulk:
            values = source.electron_bulk(electron_energies, m, ecut)
            for name in ELECTRON_FIELDS:
                electron[name][m] = values[name]
            log_eta[m] = values["log_eta"]
            restricted_range[m] = source.restricted_range_bulk(electron_energies, m, ecut)
        else:
            for j, e in enumerate(electron_energies):
                e = float(e)
                electron["stopping_restricted"][m, j] = source.restricted_stopping_power(e, m, ecut)
                electron["stopping_radiative"][m, j] = source.radiative_stopping_power(e, m)
                electron["moller"][m, j] = source.moller_cross_section(e, m, ecut)
                electron["csda_range"][m, j] = source.csda_range(e, m)
                electron["scattering_power"][m, j] = source.scattering_power(e, m, ecut)
                log_eta[m, j] = np.log(source.elastic_screening(e, m))
                restricted_range[m, j] = source.restricted_range(e, m, ecut)

    # Inverse of the restricted range on one shared log-range grid, clamped into
    # each material's own [ecut, e_top].
    e_top = float(electron_energies[-1])
    r_top = float(restricted_range[:, -1].max())
    range_nodes = np.geomspace(r_top * RANGE_R_MIN_FRACTION, r_top, n_points)
    energy_of_rr = np.empty((n_mat, n_points))
    for m in range(n_mat):
        r_top_m = float(restricted_range[m, -1])
        mass_path = r_top_m - range_nodes
        if bulk:
            start = float(source.restricted_range_bulk(np.array([e_top]), m, ecut)[0])
            remaining = start - mass_path
            log_e, cumulative = source.restricted_grid(m, ecut)
            e_end = np.exp(np.interp(remaining, cumulative, log_e))
            e_end = np.minimum(np.maximum(e_end, ecut), e_top)
            energy_of_rr[m] = np.where(remaining <= 0.0, ecut, e_end)
        else:
            for j in range(n_points):
                energy_of_rr[m, j] = source.energy_after_mass_path(e_top, m, ecut, float(mass_path[j]))

    p_min, p_inv = _grid_metadata(photon_energies, n_points)
    e_min, e_inv = _grid_metadata(electron_energies, n_points)
    range_log_r_min = float(np.log(range_nodes[0]))
    range_inv_dlog = (n_points - 1) / float(np.log(range_nodes[-1]) - np.log(range_nodes[0]))

    x_max = RAYLEIGH_MOMENTUM_TRANSFER_PER_MEV * e_max
    coherent_x = np.geomspace(x_max * COHERENT_X_MIN_FRACTION, x_max, COHERENT_POINTS)
    if hasattr(source, "coherent_cumulative"):
        coherent = np.stack([source.coherent_cumulative(coherent_x, m) for m in range(n_mat)])
    else:
        coherent = np.stack([0.5 * coherent_x**2 for _ in range(n_mat)])

    return CrossSectionTables(
        n_points=n_points,
        ecut=ecut,
        pcut=pcut,
        e_max=e_max,
        photon_log_e_min=p_min,
        photon_inv_dlog=p_inv,
        electron_log_e_min=e_min,
        electron_inv_dlog=e_inv,
        mu_compton=channels[0],
        mu_photo=channels[1],
        mu_pair=channels[2],
        mu_rayleigh=channels[3],
        majorant=majorant,
        log_eta=log_eta,
        restricted_range=restricted_range,
        energy_of_restricted_range=energy_of_rr,
        range_log_r_min=range_log_r_min,
        range_inv_dlog=range_inv_dlog,
        coherent_x=coherent_x,
        coherent_cumulative=coherent,
        n_coherent=COHERENT_POINTS,
        **electron,
    )
