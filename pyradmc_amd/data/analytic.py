"""Closed-form water cross-sections and electron stopping data, evaluated in bulk.

The host-side data source of the transport path: the same closed forms the
reference's ``AnalyticCrossSections`` answers through its scalar query API
(``pyradmc/data/analytic.py:117-321``, ``pyradmc/data/berger_seltzer.py:47-240``,
``pyradmc/data/goudsmit_saunderson.py:140-399``), written here as array functions so
a whole 2048-node table flattens in one vectorized pass instead of a Python loop per
node. The scalar methods on :class:`AnalyticCrossSections` keep the reference's
host-API surface (``CrossSectionSource``, ``pyradmc/data/interface.py:80-443``) for
drop-in use; they evaluate the same array functions at one point.

Channels: Klein-Nishina Compton on free electrons; ``E^-3`` photoelectric anchored at
50 keV; a pair channel calibrated by least squares so the total matches the NIST
XCOM water totals; no Rayleigh (zero column). Electrons: Berger-Seltzer restricted
collision stopping with the Sternheimer density effect; closed-form restricted Moller
cross-section; a log-quadratic through three ESTAR radiative anchors; CSDA range by
trapezoid over ``1/S_total``; the Class-II transport-moment scattering power of the
Moliere-screened Rutherford law.

Units: MeV, cm^2/g, MeV cm^2/g, g/cm^2, rad^2 cm^2/g.
"""

from __future__ import annotations

import math

import numpy as np

from pyradmc_amd.constants import ELECTRON_MASS_MEV
from pyradmc_amd.data.materials import (
    AVOGADRO,
    STANDARD_ATOMIC_WEIGHT,
    WATER,
    WATER_DATA,
    MaterialData,
)

CLASSICAL_ELECTRON_RADIUS_CM = 2.817_940_3262e-13
TWO_PI_RE2_MC2 = 2.0 * math.pi * CLASSICAL_ELECTRON_RADIUS_CM**2 * ELECTRON_MASS_MEV
PAIR_THRESHOLD_MEV = 2.0 * ELECTRON_MASS_MEV

# The elastic-scattering (Moliere) layer keeps its own, shorter rest-mass constant,
# exactly as the reference GS module does (goudsmit_saunderson.py:104-108); the two
# values differ at the 1e-10 level and both are reproduced where they are used.
_GS_ELECTRON_MASS_MEV = 0.51099895
_HBAR_C_MEV_FM = 197.3269804
_BOHR_RADIUS_FM = 5.29177210903e4
_FINE_STRUCTURE = 7.2973525693e-3
_THOMAS_FERMI_MEV = _HBAR_C_MEV_FM / (0.885 * _BOHR_RADIUS_FM)

_PHOTOELECTRIC_ANCHOR_ENERGY_MEV = 0.050
_PHOTOELECTRIC_ANCHOR_WATER = 0.030

# NIST XCOM total mass attenuation of liquid water (coherent included), cm^2/g:
# the calibration anchors of the pair channel (Hubbell & Seltzer, NISTIR 5632).
_WATER_TOTAL_ANCHORS = np.array(
    [
        (2.0, 0.04942),
        (3.0, 0.03969),
        (4.0, 0.03403),
        (5.0, 0.03031),
        (6.0, 0.02770),
        (8.0, 0.02429),
        (10.0, 0.02219),
        (15.0, 0.01941),
        (20.0, 0.01813),
    ]
)

RANGE_GRID_E_MAX = 64.0
RANGE_GRID_POINTS = 4096


# --------------------------------------------------------------------------- photons


def klein_nishina_total_cm2(energy: np.ndarray) -> np.ndarray:
    """Klein-Nishina total cross-section per free electron (cm^2)."""
    alpha = np.asarray(energy, dtype=np.float64) / ELECTRON_MASS_MEV
    one_plus_2a = 1.0 + 2.0 * alpha
    log_term = np.log(one_plus_2a)
    bracket = (
        (1.0 + alpha) / alpha**2 * (2.0 * (1.0 + alpha) / one_plus_2a - log_term / alpha)
        + log_term / (2.0 * alpha)
        - (1.0 + 3.0 * alpha) / one_plus_2a**2
    )
    return 2.0 * math.pi * CLASSICAL_ELECTRON_RADIUS_CM**2 * bracket


def photoelectric_water(energy: np.ndarray) -> np.ndarray:
    """tau/rho of water, ``0.030 (0.05/E)^3`` cm^2/g."""
    e = np.asarray(energy, dtype=np.float64)
    return _PHOTOELECTRIC_ANCHOR_WATER * (_PHOTOELECTRIC_ANCHOR_ENERGY_MEV / e) ** 3


def fit_pair_coefficients() -> tuple[float, float, float]:
    """Least-squares ``c0 + c1 L + c2 L^2`` so that KN + photo + pair hits NIST totals."""
    energies = _WATER_TOTAL_ANCHORS[:, 0]
    totals = _WATER_TOTAL_ANCHORS[:, 1]
    kn = WATER_DATA.electrons_per_gram * klein_nishina_total_cm2(energies)
    tau = photoelectric_water(energies)
    shape = (1.0 - PAIR_THRESHOLD_MEV / energies) ** 3
    target = (totals - kn - tau) / shape
    log_e = np.log(energies / PAIR_THRESHOLD_MEV)
    design = np.stack([np.ones_like(log_e), log_e, log_e**2], axis=1)
    coeffs, *_ = np.linalg.lstsq(design, target, rcond=None)
    return float(coeffs[0]), float(coeffs[1]), float(coeffs[2])


def pair_water(energy: np.ndarray, coeffs: tuple[float, float, float]) -> np.ndarray:
    """kappa/rho of water: threshold cube times the calibrated log-quadratic, >= 0."""
    e = np.asarray(energy, dtype=np.float64)
    out = np.zeros_like(e)
    above = e > PAIR_THRESHOLD_MEV
    ea = e[above]
    shape = (1.0 - PAIR_THRESHOLD_MEV / ea) ** 3
    log_e = np.log(ea / PAIR_THRESHOLD_MEV)
    c0, c1, c2 = coeffs
    out[above] = np.maximum(0.0, shape * (c0 + c1 * log_e + c2 * log_e**2))
    return out


# ------------------------------------------------------------------------- electrons


def density_effect(tau: np.ndarray, material: MaterialData) -> np.ndarray:
    """Sternheimer density-effect correction, piecewise in ``x = log10(p/mc)``."""
    a, m, x0, x1, cbar, delta0 = material.sternheimer
    x = 0.5 * np.log10(tau * (tau + 2.0))
    high = 4.6052 * x - cbar
    mid = 4.6052 * x - cbar + a * np.power(np.maximum(x1 - x, 0.0), m)
    low = delta0 * np.power(10.0, 2.0 * (x - x0))
    return np.where(x >= x1, high, np.where(x >= x0, mid, low))


def restricted_collision_stopping(
    energy: np.ndarray, material: MaterialData, delta_cut: float | np.ndarray
) -> np.ndarray:
    """Berger-Seltzer restricted collision stopping power (MeV cm^2/g)."""
    e = np.asarray(energy, dtype=np.float64)
    tau = e / ELECTRON_MASS_MEV
    beta_sq = tau * (tau + 2.0) / (tau + 1.0) ** 2
    w = np.minimum(np.asarray(delta_cut, dtype=np.float64) / e, 0.5)
    g_minus = (
        -1.0
        - beta_sq
        + np.log(4.0 * (1.0 - w) * w)
        + 1.0 / (1.0 - w)
        + ((tau * w) ** 2 / 2.0 + (2.0 * tau + 1.0) * np.log(1.0 - w)) / (tau + 1.0) ** 2
    )
    log_term = np.log(
        tau**2 * (tau + 2.0) / (2.0 * (material.mean_excitation_mev / ELECTRON_MASS_MEV) ** 2)
    )
    prefactor = TWO_PI_RE2_MC2 * material.electrons_per_gram / beta_sq
    return prefactor * (log_term + g_minus - density_effect(tau, material))


def restricted_moller_cross_section(
    energy: np.ndarray, material: MaterialData, delta_cut: float
) -> np.ndarray:
    """Closed-form restricted Moller cross-section (cm^2/g); zero at or below 2 cut."""
    e = np.asarray(energy, dtype=np.float64)
    out = np.zeros_like(e)
    live = e > 2.0 * delta_cut
    el = e[live]
    tau = el / ELECTRON_MASS_MEV
    beta_sq = tau * (tau + 2.0) / (tau + 1.0) ** 2
    w = delta_cut / el
    bracket = (
        (tau / (tau + 1.0)) ** 2 * (0.5 - w)
        + 1.0 / w
        - 1.0 / (1.0 - w)
        + (2.0 * tau + 1.0) / (tau + 1.0) ** 2 * np.log(w / (1.0 - w))
    )
    out[live] = TWO_PI_RE2_MC2 * material.electrons_per_gram / (beta_sq * el) * bracket
    return out


def radiative_fit_coefficients(material: MaterialData) -> tuple[float, float, float]:
    """Exact log-quadratic through the material's three ESTAR radiative anchors."""
    anchors = material.radiative_anchors
    log_e = np.log([e for e, _ in anchors])
    log_s = np.log([s for _, s in anchors])
    design = np.stack([np.ones_like(log_e), log_e, log_e**2], axis=1)
    c = np.linalg.solve(design, log_s)
    return float(c[0]), float(c[1]), float(c[2])


def radiative_stopping(energy: np.ndarray, coeffs: tuple[float, float, float]) -> np.ndarray:
    log_e = np.log(np.asarray(energy, dtype=np.float64))
    c0, c1, c2 = coeffs
    return np.exp(c0 + c1 * log_e + c2 * log_e**2)


def csda_range_table(material: MaterialData) -> tuple[np.ndarray, np.ndarray]:
    """``(log E, range)``: trapezoid of ``1/S_total`` on 600 log nodes, 1 keV-30 MeV."""
    coeffs = radiative_fit_coefficients(material)
    energies = np.geomspace(1.0e-3, 30.0, 600)
    total = restricted_collision_stopping(energies, material, energies / 2.0) + radiative_stopping(
        energies, coeffs
    )
    inverse_total = 1.0 / total
    steps = np.diff(energies) * 0.5 * (inverse_total[1:] + inverse_total[:-1])
    return np.log(energies), np.concatenate(([0.0], np.cumsum(steps)))


def restricted_range_grid(material: MaterialData, delta_cut: float) -> tuple[np.ndarray, np.ndarray]:
    """``(log E, r(E))`` for the restricted-collision range from ``delta_cut`` to 64 MeV."""
    energies = np.geomspace(delta_cut, RANGE_GRID_E_MAX, RANGE_GRID_POINTS)
    inv_s = 1.0 / restricted_collision_stopping(energies, material, delta_cut)
    cumulative = np.concatenate(
        ([0.0], np.cumsum(np.diff(energies) * 0.5 * (inv_s[1:] + inv_s[:-1])))
    )
    return np.log(energies), cumulative


# ------------------------------------------------------ elastic scattering (Moliere)


def moliere_screening(composition: tuple[tuple[int, float], ...], energy: np.ndarray) -> np.ndarray:
    """Moliere screening parameter eta; compound = weighted geometric mean."""
    e = np.asarray(energy, dtype=np.float64)
    m = _GS_ELECTRON_MASS_MEV
    total_energy = e + m
    momentum = np.sqrt(e * (e + 2.0 * m))
    beta = momentum / total_energy
    weight_sum = 0.0
    log_sum = np.zeros_like(e)
    for z, fraction in composition:
        weight = fraction * z * (z + 1) / STANDARD_ATOMIC_WEIGHT[z]
        chi_0 = _THOMAS_FERMI_MEV * z ** (1.0 / 3.0) / momentum
        correction = 1.13 + 3.76 * (_FINE_STRUCTURE * z / beta) ** 2
        weight_sum += weight
        log_sum = log_sum + weight * np.log(0.25 * chi_0 * chi_0 * correction)
    return np.exp(log_sum / weight_sum)


def first_transport_moment(eta: np.ndarray) -> np.ndarray:
    """``G_1 = 2 eta (1+eta) ln((1+eta)/eta) - 2 eta`` of screened Rutherford."""
    eta = np.asarray(eta, dtype=np.float64)
    return 2.0 * eta * (1.0 + eta) * np.log((1.0 + eta) / eta) - 2.0 * eta


def _moller_transport_antiderivative(u: np.ndarray, tau: np.ndarray) -> np.ndarray:
    c = tau + 2.0
    denominator = tau * u * u - c
    root = np.sqrt(tau * c)
    return (
        2.0 * tau * c / (tau + 1.0) ** 2 * (u - 1.0) / denominator
        + 2.0 * root / (tau + 1.0) ** 2 * np.arctanh(np.sqrt(tau / c) * u)
        + 2.0 / (c * (u + 1.0))
        - np.log(1.0 - u) / c
        + (5.0 * tau * tau + 12.0 * tau + 5.0) / ((tau + 1.0) ** 2 * c) * np.log(u + 1.0)
        - 2.0 * (2.0 * tau + 1.0) / (tau + 1.0) ** 2 * np.log(u)
        + c / u
        - c / (2.0 * u * u)
    )


def hard_moller_transport_cross_section(energy: np.ndarray, delta_cut: float) -> np.ndarray:
    """First transport cross-section of explicit Moller events (cm^2 per electron)."""
    e = np.asarray(energy, dtype=np.float64)
    out = np.zeros_like(e)
    live = e > 2.0 * delta_cut
    el = e[live]
    m = _GS_ELECTRON_MASS_MEV
    tau = el / m
    beta_sq = tau * (tau + 2.0) / (tau + 1.0) ** 2

    def primary_cosine(transfer: np.ndarray) -> np.ndarray:
        remaining = el - transfer
        return np.sqrt(remaining * (el + 2.0 * m) / (el * (remaining + 2.0 * m)))

    u_cut = primary_cosine(np.full_like(el, delta_cut))
    u_half = primary_cosine(0.5 * el)
    dimensionless = _moller_transport_antiderivative(u_cut, tau) - _moller_transport_antiderivative(
        u_half, tau
    )
    prefactor = 2.0 * math.pi * CLASSICAL_ELECTRON_RADIUS_CM**2 * m / (beta_sq * el)
    out[live] = np.maximum(0.0, prefactor * dimensionless)
    return out


def _unit_screened_rutherford_transport(
    composition: tuple[tuple[int, float], ...], energy: np.ndarray
) -> np.ndarray:
    e = np.asarray(energy, dtype=np.float64)
    eta = moliere_screening(composition, e)
    m = _GS_ELECTRON_MASS_MEV
    tau = e / m
    pv = m * tau * (tau + 2.0) / (tau + 1.0)
    sigma_el = math.pi * CLASSICAL_ELECTRON_RADIUS_CM**2 * (m / pv) ** 2 / (eta * (1.0 + eta))
    return sigma_el * first_transport_moment(eta)


def transport_moment_scattering_power(
    composition: tuple[tuple[int, float], ...], energy: np.ndarray, delta_cut: float
) -> np.ndarray:
    """Class-II mass scattering power ``2 sum w N_A/A [Z^2 s_tr + Z (s_tr - s_M,hard)]``."""
    sigma_transport = _unit_screened_rutherford_transport(composition, energy)
    nuclear_atoms = sum(
        fraction * AVOGADRO / STANDARD_ATOMIC_WEIGHT[z] * z * z for z, fraction in composition
    )
    nuclear = 2.0 * nuclear_atoms * sigma_transport
    hard = hard_moller_transport_cross_section(energy, delta_cut)
    soft = np.maximum(0.0, sigma_transport - hard)
    electrons_per_gram = sum(
        fraction * AVOGADRO / STANDARD_ATOMIC_WEIGHT[z] * z for z, fraction in composition
    )
    return nuclear + 2.0 * electrons_per_gram * soft


# ------------------------------------------------------------------------- the source


class AnalyticCrossSections:
    """Analytic water data source; drop-in for the reference's ``AnalyticCrossSections``.

    ``geometry_densities`` are the ``(material, max density)`` pairs the Woodcock
    majorant is taken over (``pyradmc/data/analytic.py:117-165``). Water only.
    """

    def __init__(self, geometry_densities: tuple[tuple[int, float], ...] = ((WATER, 1.0),)) -> None:
        for material, density in geometry_densities:
            if not 0 <= material < self.n_materials:
                raise ValueError(
                    f"material index {material} beyond the analytic source (water only)"
                )
            if density <= 0.0:
                raise ValueError(f"non-positive density {density} for material {material}")
        self._geometry_densities = tuple((int(m), float(d)) for m, d in geometry_densities)
        self._pair_coeffs = fit_pair_coefficients()
        self._radiative_coeffs = radiative_fit_coefficients(WATER_DATA)
        self._range_log_e, self._range_values = csda_range_table(WATER_DATA)
        self._restricted_grids: dict[float, tuple[np.ndarray, np.ndarray]] = {}

    # -- identity ------------------------------------------------------------------

    @property
    def n_materials(self) -> int:
        return 1

    @property
    def geometry_densities(self) -> tuple[tuple[int, float], ...]:
        return self._geometry_densities

    @property
    def provenance(self) -> str:
        return (
            "analytic closed-form water (Klein-Nishina without binding, power-law "
            "photoelectric, fitted pair, Berger-Seltzer ICRU-37 stopping)"
        )

    # -- bulk evaluation (what the table builder uses) -----------------------------

    def _check_material(self, material: int) -> None:
        if not 0 <= material < self.n_materials:
            raise ValueError(f"material index {material} beyond the analytic source (water only)")

    def mu_channels(self, energy: np.ndarray, material: int) -> tuple[np.ndarray, ...]:
        """``(compton, photo, pair, rayleigh)`` mass coefficients at each energy."""
        self._check_material(material)
        e = np.asarray(energy, dtype=np.float64)
        compton = WATER_DATA.electrons_per_gram * klein_nishina_total_cm2(e)
        photo = photoelectric_water(e)
        pair = pair_water(e, self._pair_coeffs)
        return compton, photo, pair, np.zeros_like(e)

    def mu_total_bulk(self, energy: np.ndarray, material: int) -> np.ndarray:
        compton, photo, pair, _ = self.mu_channels(energy, material)
        return compton + photo + pair

    def majorant_bulk(self, energy: np.ndarray) -> np.ndarray:
        e = np.asarray(energy, dtype=np.float64)
        result = None
        for material, density in self._geometry_densities:
            value = density * self.mu_total_bulk(e, material)
            result = value if result is None else np.maximum(result, value)
        return result

    def restricted_range_bulk(self, energy: np.ndarray, material: int, delta_cut: float) -> np.ndarray:
        log_e, cumulative = self.restricted_grid(material, delta_cut)
        e = np.clip(np.asarray(energy, dtype=np.float64), delta_cut, RANGE_GRID_E_MAX)
        return np.interp(np.log(e), log_e, cumulative)

    def restricted_grid(self, material: int, delta_cut: float) -> tuple[np.ndarray, np.ndarray]:
        """(log E nodes, restricted range) from the cut to 64 MeV (water)."""
        self._check_material(material)
        grid = self._restricted_grids.get(delta_cut)
        if grid is None:
            grid = restricted_range_grid(WATER_DATA, delta_cut)
            self._restricted_grids[delta_cut] = grid
        return grid

    # -- scalar host API (CrossSectionSource surface) -------------------------------

    def mu_over_rho(self, energy: float, material: int, process: int) -> float:
        if energy <= 0.0:
            raise ValueError(f"non-positive photon energy {energy} MeV")
        channels = self.mu_channels(np.array([energy]), material)
        if not 0 <= process <= 3:
            raise ValueError(f"unknown photon process {process}")
        return float(channels[process][0])

    def mu_over_rho_total(self, energy: float, material: int) -> float:
        return float(self.mu_total_bulk(np.array([energy]), material)[0])

    def majorant(self, energy: float) -> float:
        return float(self.majorant_bulk(np.array([energy]))[0])

    def restricted_stopping_power(self, energy: float, material: int, delta_cut: float) -> float:
        self._check_material(material)
        return float(restricted_collision_stopping(np.array([energy]), WATER_DATA, delta_cut)[0])

    def radiative_stopping_power(self, energy: float, material: int) -> float:
        self._check_material(material)
        return float(radiative_stopping(np.array([energy]), self._radiative_coeffs)[0])

    def moller_cross_section(self, energy: float, material: int, delta_cut: float) -> float:
        self._check_material(material)
        return float(restricted_moller_cross_section(np.array([energy]), WATER_DATA, delta_cut)[0])

    def csda_range(self, energy: float, material: int) -> float:
        self._check_material(material)
        return float(np.interp(math.log(energy), self._range_log_e, self._range_values))

    def scattering_power(self, energy: float, material: int, delta_cut: float) -> float:
        self._check_material(material)
        return float(
            transport_moment_scattering_power(WATER_DATA.composition, np.array([energy]), delta_cut)[0]
        )

    def elastic_screening(self, energy: float, material: int) -> float:
        self._check_material(material)
        return float(moliere_screening(WATER_DATA.composition, np.array([energy]))[0])

    def restricted_range(self, energy: float, material: int, delta_cut: float) -> float:
        self._check_material(material)
        return float(self.restricted_range_bulk(np.array([energy]), material, delta_cut)[0])

    def energy_after_mass_path(
        self, energy: float, material: int, delta_cut: float, mass_path: float
    ) -> float:
        log_e, cumulative = self.restricted_grid(material, delta_cut)
        remaining = self.restricted_range(energy, material, delta_cut) - mass_path
        if remaining <= 0.0:
            return delta_cut
        e_end = float(np.exp(np.interp(remaining, cumulative, log_e)))
        return min(max(e_end, delta_cut), energy)

    # -- flattening -----------------------------------------------------------------

    def build_tables(self, ecut: float, pcut: float, e_max: float, n_points: int | None = None):
        from pyradmc_amd.data.tables import TABLE_POINTS, build_cross_section_tables

        return build_cross_section_tables(
            self, ecut=ecut, pcut=pcut, e_max=e_max,
            n_points=TABLE_POINTS if n_points is None else n_points,
        )

    # Bulk electron quantities for the table builder.
    def electron_bulk(self, energy: np.ndarray, material: int, ecut: float) -> dict[str, np.ndarray]:
        self._check_material(material)
        e = np.asarray(energy, dtype=np.float64)
        return {
            "stopping_restricted": restricted_collision_stopping(e, WATER_DATA, ecut),
            "stopping_radiative": radiative_stopping(e, self._radiative_coeffs),
            "moller": restricted_moller_cross_section(e, WATER_DATA, ecut),
            "csda_range": np.interp(np.log(e), self._range_log_e, self._range_values),
            "scattering_power": transport_moment_scattering_power(WATER_DATA.composition, e, ecut),
            "log_eta": np.log(moliere_screening(WATER_DATA.composition, e)),
        }
