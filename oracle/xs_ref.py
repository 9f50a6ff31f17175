"""ORACLE (test infrastructure only): scalar restatement of pyRadMC's analytic data layer.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import this package, and only as the checker; the product (``pyradmc_amd``) never
does. It restates, point by point in plain ``math``, the host query API and the table
flattening of the reference:

* Klein-Nishina / photoelectric / fitted pair - ``pyradmc/data/analytic.py:90-321``
* Berger-Seltzer stopping, density effect, Moller, radiative fit, CSDA range -
  ``pyradmc/data/berger_seltzer.py:47-240``
* Moliere screening, transport-moment scattering power -
  ``pyradmc/data/goudsmit_saunderson.py:140-399``
* restricted range and its inverse - ``pyradmc/data/interface.py:342-409``
* ``build_cross_section_tables`` - ``pyradmc/data/tables.py:222-362``

Pinned by: the reference's own table golden vectors (``tests/golden/
tables_water_analytic.npz``, written from the reference in the build container),
NIST XCOM water totals, the KN closed form at low energy (Thomson limit), and the
restricted-range identities r(ECUT)=0, dr/dE = 1/S (``tests/test_oracle_pins.py``).
"""

from __future__ import annotations

import math

import numpy as np

ELECTRON_MASS = 0.510_998_950_69      # pyradmc/__init__.py:97
GS_ELECTRON_MASS = 0.51099895          # goudsmit_saunderson.py:104
R_E = 2.817_940_3262e-13               # berger_seltzer.py:38
AVOGADRO = 6.022_140_76e23             # materials.py:46
ATOMIC_WEIGHT = {1: 1.008, 8: 15.999}  # materials.py:55-75
HBAR_C = 197.3269804
BOHR = 5.29177210903e4
ALPHA_FS = 7.2973525693e-3
TF = HBAR_C / (0.885 * BOHR)

WATER_NE = 10.0 * AVOGADRO / 18.01528
_M_H2O = 2.0 * ATOMIC_WEIGHT[1] + ATOMIC_WEIGHT[8]
WATER_COMPOSITION = ((1, 2.0 * ATOMIC_WEIGHT[1] / _M_H2O), (8, ATOMIC_WEIGHT[8] / _M_H2O))
WATER_I = 75.0e-6
WATER_STERNHEIMER = (0.09116, 3.4773, 0.2400, 2.8004, 3.5017, 0.097)
WATER_RADIATIVE = ((1.0, 0.0128), (10.0, 0.1813), (20.0, 0.4008))
NIST_TOTALS = ((2.0, 0.04942), (3.0, 0.03969), (4.0, 0.03403), (5.0, 0.03031), (6.0, 0.02770),
               (8.0, 0.02429), (10.0, 0.02219), (15.0, 0.01941), (20.0, 0.01813))
PAIR_THRESHOLD = 2.0 * ELECTRON_MASS
K_BS = 2.0 * math.pi * R_E**2 * ELECTRON_MASS


def kn_sigma(e: float) -> float:
    a = e / ELECTRON_MASS
    q = 1.0 + 2.0 * a
    lg = math.log(q)
    return 2.0 * math.pi * R_E**2 * (
        (1.0 + a) / a**2 * (2.0 * (1.0 + a) / q - lg / a) + lg / (2.0 * a) - (1.0 + 3.0 * a) / q**2
    )


def photo(e: float) -> float:
    return 0.030 * (0.050 / e) ** 3


def pair_coefficients() -> tuple[float, float, float]:
    e = np.array([a for a, _ in NIST_TOTALS])
    tot = np.array([b for _, b in NIST_TOTALS])
    kn = np.array([WATER_NE * kn_sigma(float(x)) for x in e])
    tau = np.array([photo(float(x)) for x in e])
    shape = (1.0 - PAIR_THRESHOLD / e) ** 3
    lg = np.log(e / PAIR_THRESHOLD)
    c, *_ = np.linalg.lstsq(np.stack([np.ones_like(lg), lg, lg**2], axis=1), (tot - kn - tau) / shape, rcond=None)
    return float(c[0]), float(c[1]), float(c[2])


def pair(e: float, c) -> float:
    if e <= PAIR_THRESHOLD:
        return 0.0
    lg = math.log(e / PAIR_THRESHOLD)
    return max(0.0, (1.0 - PAIR_THRESHOLD / e) ** 3 * (c[0] + c[1] * lg + c[2] * lg * lg))


def delta(tau: float) -> float:
    a, m, x0, x1, cbar, d0 = WATER_STERNHEIMER
    x = 0.5 * math.log10(tau * (tau + 2.0))
    if x >= x1:
        return 4.6052 * x - cbar
    if x >= x0:
        return 4.6052 * x - cbar + a * math.pow(x1 - x, m)
    return d0 * math.pow(10.0, 2.0 * (x - x0))


def s_col(e: float, cut: float) -> float:
    tau = e / ELECTRON_MASS
    b2 = tau * (tau + 2.0) / (tau + 1.0) ** 2
    w = min(cut / e, 0.5)
    g = (-1.0 - b2 + math.log(4.0 * (1.0 - w) * w) + 1.0 / (1.0 - w)
         + ((tau * w) ** 2 / 2.0 + (2.0 * tau + 1.0) * math.log(1.0 - w)) / (tau + 1.0) ** 2)
    lt = math.log(tau**2 * (tau + 2.0) / (2.0 * (WATER_I / ELECTRON_MASS) ** 2))
    return K_BS * WATER_NE / b2 * (lt + g - delta(tau))


def moller_sigma(e: float, cut: float) -> float:
    if e <= 2.0 * cut:
        return 0.0
    tau = e / ELECTRON_MASS
    b2 = tau * (tau + 2.0) / (tau + 1.0) ** 2
    w = cut / e
    br = ((tau / (tau + 1.0)) ** 2 * (0.5 - w) + 1.0 / w - 1.0 / (1.0 - w)
          + (2.0 * tau + 1.0) / (tau + 1.0) ** 2 * math.log(w / (1.0 - w)))
    return K_BS * WATER_NE / (b2 * e) * br


def radiative_coefficients() -> tuple[float, float, float]:
    le = np.log([a for a, _ in WATER_RADIATIVE])
    ls = np.log([b for _, b in WATER_RADIATIVE])
    c = np.linalg.solve(np.stack([np.ones_like(le), le, le**2], axis=1), ls)
    return float(c[0]), float(c[1]), float(c[2])


def s_rad(e: float, c) -> float:
    lg = math.log(e)
    return math.exp(c[0] + c[1] * lg + c[2] * lg * lg)


def csda_table(crad):
    e = np.geomspace(1.0e-3, 30.0, 600)
    inv = np.array([1.0 / (s_col(float(x), float(x) / 2.0) + s_rad(float(x), crad)) for x in e])
    return np.log(e), np.concatenate(([0.0], np.cumsum(np.diff(e) * 0.5 * (inv[1:] + inv[:-1]))))


def screening(e: float) -> float:
    m = GS_ELECTRON_MASS
    p = math.sqrt(e * (e + 2.0 * m))
    beta = p / (e + m)
    ws = 0.0
    ls = 0.0
    for z, f in WATER_COMPOSITION:
        w = f * z * (z + 1) / ATOMIC_WEIGHT[z]
        chi0 = TF * z ** (1.0 / 3.0) / p
        ws += w
        ls += w * math.log(0.25 * chi0 * chi0 * (1.13 + 3.76 * (ALPHA_FS * z / beta) ** 2))
    return math.exp(ls / ws)


def g1(eta: float) -> float:
    return 2.0 * eta * (1.0 + eta) * math.log((1.0 + eta) / eta) - 2.0 * eta


def _moller_antiderivative(u: float, tau: float) -> float:
    c = tau + 2.0
    return (2.0 * tau * c / (tau + 1.0) ** 2 * (u - 1.0) / (tau * u * u - c)
            + 2.0 * math.sqrt(tau * c) / (tau + 1.0) ** 2 * math.atanh(math.sqrt(tau / c) * u)
            + 2.0 / (c * (u + 1.0)) - math.log(1.0 - u) / c
            + (5.0 * tau * tau + 12.0 * tau + 5.0) / ((tau + 1.0) ** 2 * c) * math.log(u + 1.0)
            - 2.0 * (2.0 * tau + 1.0) / (tau + 1.0) ** 2 * math.log(u) + c / u - c / (2.0 * u * u))


def hard_moller_transport(e: float, cut: float) -> float:
    if e <= 2.0 * cut:
        return 0.0
    m = GS_ELECTRON_MASS
    tau = e / m
    b2 = tau * (tau + 2.0) / (tau + 1.0) ** 2

    def cos_p(t: float) -> float:
        r = e - t
        return math.sqrt(r * (e + 2.0 * m) / (e * (r + 2.0 * m)))

    dim = _moller_antiderivative(cos_p(cut), tau) - _moller_antiderivative(cos_p(0.5 * e), tau)
    return max(0.0, 2.0 * math.pi * R_E**2 * m / (b2 * e) * dim)


def scattering_power(e: float, cut: float) -> float:
    eta = screening(e)
    m = GS_ELECTRON_MASS
    tau = e / m
    pv = m * tau * (tau + 2.0) / (tau + 1.0)
    sig = math.pi * R_E**2 * (m / pv) ** 2 / (eta * (1.0 + eta)) * g1(eta)
    nuc = sum(f * AVOGADRO / ATOMIC_WEIGHT[z] * z * z for z, f in WATER_COMPOSITION)
    ne = sum(f * AVOGADRO / ATOMIC_WEIGHT[z] * z for z, f in WATER_COMPOSITION)
    return 2.0 * nuc * sig + 2.0 * ne * max(0.0, sig - hard_moller_transport(e, cut))


class WaterOracle:
    """The analytic water source, answered point by point."""

    def __init__(self, geometry_densities=((0, 1.0),)) -> None:
        self.geometry_densities = tuple(geometry_densities)
        self.pc = pair_coefficients()
        self.rc = radiative_coefficients()
        self.csda_log_e, self.csda_r = csda_table(self.rc)
        self._rr = {}

    def mu(self, e: float) -> tuple[float, float, float, float]:
        return WATER_NE * kn_sigma(e), photo(e), pair(e, self.pc), 0.0

    def majorant(self, e: float) -> float:
        return max(rho * sum(self.mu(e)) for _, rho in self.geometry_densities)

    def csda_range(self, e: float) -> float:
        return float(np.interp(math.log(e), self.csda_log_e, self.csda_r))

    def range_grid(self, cut: float):
        g = self._rr.get(cut)
        if g is None:
            e = np.geomspace(cut, 64.0, 4096)
            inv = np.array([1.0 / s_col(float(x), cut) for x in e])
            g = (np.log(e), np.concatenate(([0.0], np.cumsum(np.diff(e) * 0.5 * (inv[1:] + inv[:-1])))))
            self._rr[cut] = g
        return g

    def restricted_range(self, e: float, cut: float) -> float:
        le, r = self.range_grid(cut)
        return float(np.interp(math.log(min(max(e, cut), 64.0)), le, r))

    def energy_after(self, e: float, cut: float, mass_path: float) -> float:
        le, r = self.range_grid(cut)
        rem = self.restricted_range(e, cut) - mass_path
        if rem <= 0.0:
            return cut
        return min(max(float(np.exp(np.interp(rem, r, le))), cut), e)

    def tables(self, ecut: float, pcut: float, e_max: float, n: int = 2048) -> dict:
        """Flatten onto the photon, electron and range grids (float64 dict)."""
        pe = np.geomspace(0.5 * pcut, e_max, n)
        ee = np.geomspace(0.5 * ecut, e_max, n)
        ch = np.array([self.mu(float(x)) for x in pe]).T  # (4, n)
        out = {
            "mu_compton": ch[0][None], "mu_photo": ch[1][None], "mu_pair": ch[2][None],
            "mu_rayleigh": ch[3][None],
            "majorant": np.array([self.majorant(float(x)) for x in pe]) * (1.0 + 1.0e-6),
            "stopping_restricted": np.array([[s_col(float(x), ecut) for x in ee]]),
            "stopping_radiative": np.array([[s_rad(float(x), self.rc) for x in ee]]),
            "moller": np.array([[moller_sigma(float(x), ecut) for x in ee]]),
            "csda_range": np.array([[self.csda_range(float(x)) for x in ee]]),
            "scattering_power": np.array([[scattering_power(float(x), ecut) for x in ee]]),
            "log_eta": np.array([[math.log(screening(float(x))) for x in ee]]),
            "restricted_range": np.array([[self.restricted_range(float(x), ecut) for x in ee]]),
        }
        e_top = float(ee[-1])
        r_top = float(out["restricted_range"][0, -1])
        nodes = np.geomspace(r_top * 1.0e-6, r_top, n)
        out["energy_of_restricted_range"] = np.array(
            [[self.energy_after(e_top, ecut, r_top - float(r)) for r in nodes]]
        )
        x_max = 80.65543 * e_max
        out["coherent_x"] = np.geomspace(x_max * 1.0e-7, x_max, 512)
        out["coherent_cumulative"] = (0.5 * out["coherent_x"] ** 2)[None]
        lp0, lp1 = math.log(pe[0]), math.log(pe[-1])
        le0, le1 = math.log(ee[0]), math.log(ee[-1])
        out.update(
            n_points=n, ecut=ecut, pcut=pcut, e_max=e_max,
            photon_log_e_min=lp0, photon_inv_dlog=(n - 1) / (lp1 - lp0),
            electron_log_e_min=le0, electron_inv_dlog=(n - 1) / (le1 - le0),
            range_log_r_min=float(np.log(nodes[0])),
            range_inv_dlog=(n - 1) / float(np.log(nodes[-1]) - np.log(nodes[0])),
        )
        return out
