"""ORACLE (test infrastructure only): scalar restatement of pyRadMC's tabulated,
material-resolved cross-section source and its table flattening.

Only ``tests/`` load this module, as the checker. It restates point by point in plain
``math``:

* the compiled-table queries, linear in value over log E with flat clamping -
  ``TabulatedCrossSections`` (``pyradmc/data/tabulated/source.py:55-241``) over
  ``lookup_loglinear_2d`` (``pyradmc/data/tables.py:108-133``);
* the Woodcock majorant over the geometry's (material, maximum density) pairs
  (``source.py:174-179``);
* the elastic screening from each material's elemental composition
  (``pyradmc/data/interface.py:239-264``, ``goudsmit_saunderson.py`` Moliere form);
* the restricted range (trapezoid over 1/S, 4096 geometric nodes from the cut to
  64 MeV) and its inverse (``interface.py:342-409``);
* ``build_cross_section_tables`` for several materials, including the shared
  log-range grid of the inverse range (``pyradmc/data/tables.py:222-362``).

The input is the compiled data itself (a dict with the fields of the reference's
``TabulatedData``, ``pyradmc/data/tabulated/model.py:18-48``), so the oracle and the
product read the same numbers and differ only in how they evaluate them.

Pinned by (``tests/test_multimaterial.py``): grid-node exactness and the half-value
second row of the reference's own synthetic two-material table
(``tests/unit/test_tabulated_source.py:61-127`` of the reference), the majorant as the
maximum over materials, the range identities r(cut) = 0 and dr/dE = 1/S, the inverse
range as a round trip, screening equal to the analytic water oracle's for water, and
the two-material slab transported by the reference's ``ReferenceEngine``
(``tests/golden/ref_two_material_slab.npz``).
"""

from __future__ import annotations

import math

import numpy as np

from oracle import xs_ref

# Elemental mass fractions (pyradmc/data/materials.py:188-226) and the standard atomic
# weights they need (materials.py:55-75), for the materials the checks use.
ATOMIC_WEIGHT = {1: 1.008, 6: 12.011, 7: 14.007, 8: 15.999, 18: 39.95}
COMPOSITION = {
    "water": xs_ref.WATER_COMPOSITION,
    "air": ((6, 0.000124), (7, 0.755267), (8, 0.231781), (18, 0.012827)),
}
PROCESSES = (0, 1, 2, 3)  # Compton, photoelectric, pair, Rayleigh


def screening(e: float, composition) -> float:
    """Moliere screening parameter of a compound: the weighted geometric mean of the
    elements' eta with weights w_i Z_i (Z_i + 1) / A_i."""
    m = xs_ref.GS_ELECTRON_MASS
    p = math.sqrt(e * (e + 2.0 * m))
    beta = p / (e + m)
    ws = 0.0
    ls = 0.0
    for z, f in composition:
        w = f * z * (z + 1) / ATOMIC_WEIGHT[z]
        chi0 = xs_ref.TF * z ** (1.0 / 3.0) / p
        ws += w
        ls += w * math.log(0.25 * chi0 * chi0 * (1.13 + 3.76 * (xs_ref.ALPHA_FS * z / beta) ** 2))
    return math.exp(ls / ws)


class TabulatedOracle:
    """The tabulated source, answered point by point."""

    def __init__(self, data: dict, geometry_densities=((0, 1.0),)) -> None:
        self.d = data
        self.names = tuple(data["materials"])
        self.geometry_densities = tuple(geometry_densities)
        self.pe = np.asarray(data["photon_energies"], dtype=np.float64)
        self.ee = np.asarray(data["electron_energies"], dtype=np.float64)
        self._rr: dict = {}

    @property
    def n_materials(self) -> int:
        return len(self.names)

    @staticmethod
    def _look(row, grid, e: float) -> float:
        n = grid.size
        lo = math.log(grid[0])
        inv = (n - 1) / (math.log(grid[-1]) - lo)
        t = (math.log(e) - lo) * inv
        t = min(max(t, 0.0), float(n - 1))
        i = min(int(t), n - 2)
        w = t - i
        return float(row[i]) * (1.0 - w) + float(row[i + 1]) * w

    def mu(self, e: float, m: int) -> tuple[float, ...]:
        out = []
        for p in PROCESSES:
            table = self.d["mu_over_rho"].get(p)
            out.append(0.0 if table is None else self._look(table[m], self.pe, e))
        return tuple(out)

    def majorant(self, e: float) -> float:
        return max(rho * sum(self.mu(e, m)) for m, rho in self.geometry_densities)

    def electron(self, name: str, e: float, m: int) -> float:
        return self._look(self.d[name][m], self.ee, e)

    def range_grid(self, m: int, cut: float):
        key = (m, cut)
        g = self._rr.get(key)
        if g is None:
            e = np.geomspace(cut, 64.0, 4096)
            inv = np.array([1.0 / self.electron("restricted_stopping", float(x), m) for x in e])
            g = (np.log(e), np.concatenate(([0.0], np.cumsum(np.diff(e) * 0.5 * (inv[1:] + inv[:-1])))))
            self._rr[key] = g
        return g

    def restricted_range(self, e: float, m: int, cut: float) -> float:
        le, r = self.range_grid(m, cut)
        return float(np.interp(math.log(min(max(e, cut), 64.0)), le, r))

    def energy_after(self, e: float, m: int, cut: float, mass_path: float) -> float:
        le, r = self.range_grid(m, cut)
        rem = self.restricted_range(e, m, cut) - mass_path
        if rem <= 0.0:
            return cut
        return min(max(float(np.exp(np.interp(rem, r, le))), cut), e)

    def tables(self, ecut: float, pcut: float, e_max: float, n: int = 2048) -> dict:
        """Flatten every material onto the photon, electron and range grids (float64)."""
        if not math.isclose(ecut, float(self.d["delta_cut"]), rel_tol=1e-9, abs_tol=0.0):
            raise ValueError("tables compiled at another delta cut")
        nm = self.n_materials
        pe = np.geomspace(0.5 * pcut, e_max, n)
        ee = np.geomspace(0.5 * ecut, e_max, n)
        ch = np.array([[self.mu(float(x), m) for x in pe] for m in range(nm)])  # (nm, n, 4)
        out = {
            "mu_compton": ch[:, :, 0], "mu_photo": ch[:, :, 1], "mu_pair": ch[:, :, 2],
            "mu_rayleigh": ch[:, :, 3],
            "majorant": np.array([self.majorant(float(x)) for x in pe]) * (1.0 + 1.0e-6),
        }
        for key, name in (("stopping_restricted", "restricted_stopping"),
                          ("stopping_radiative", "radiative_stopping"), ("moller", "moller"),
                          ("csda_range", "csda_range"), ("scattering_power", "scattering_power")):
            out[key] = np.array([[self.electron(name, float(x), m) for x in ee] for m in range(nm)])
        out["log_eta"] = np.array([[math.log(screening(float(x), COMPOSITION[self.names[m]])) for x in ee]
                                   for m in range(nm)])
        out["restricted_range"] = np.array([[self.restricted_range(float(x), m, ecut) for x in ee]
                                            for m in range(nm)])
        e_top = float(ee[-1])
        r_top = float(out["restricted_range"][:, -1].max())
        nodes = np.geomspace(r_top * 1.0e-6, r_top, n)
        out["energy_of_restricted_range"] = np.array(
            [[self.energy_after(e_top, m, ecut, float(out["restricted_range"][m, -1]) - float(r)) for r in nodes]
             for m in range(nm)])
        x_max = 80.65543 * e_max
        out["coherent_x"] = np.geomspace(x_max * 1.0e-7, x_max, 512)
        out["coherent_cumulative"] = np.stack([0.5 * out["coherent_x"] ** 2 for _ in range(nm)])
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
