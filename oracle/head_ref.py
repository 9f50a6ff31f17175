"""ORACLE (test infrastructure only): numpy restatement of the treatment-head pre-solve.

Only tests/ use this module, as the checker of the device pre-solve
(``pyradmc_amd.head.presolve_head``). It restates pyradmc/geometry/head.py:
``_MuTables`` (:113-180, log of macroscopic total and Compton mu on a geometric energy
grid), ``_sample_compton_ratios`` (:183-217, Kahn's rejection for Klein-Nishina),
``_compton_cos_thetas`` / ``_compton_electron_cos_thetas`` (:220-241),
``_rotate_directions`` (:244-275), ``_roulette_below`` (:325-345), ``presolve_head``
(:377-560: forced first Compton in the stack at the chosen device's mid-plane, the
air column's forced Compton photon and recoil electron, the transmitted primary) and
``_propagate_and_add`` (:563-583: forward, above the cutoff, positive weight, moved to
the exit plane), with the population normalised per primary (:304-322).

Chord lengths come from ``pyradmc_amd.collimation.BeamLimitingStack.path_lengths``,
which tests/test_collimation.py pins against the scalar chords of
``oracle/collim_ref.py``. The random stream is numpy's PCG64 seeded from
``(seed, 0x48454144)`` (the reference's host stream); the device draws from its own
streams, so the two pre-solves agree in distribution, never bit for bit.

Pinned in tests/test_head.py: Kahn's sampler against the Klein-Nishina distribution,
the rotation against its defining angle, attenuation mode against the closed-form
Beer-Lambert weights, the roulette's expected weight, and the population's weight
totals against the forced-interaction identities.
"""

from __future__ import annotations

import numpy as np

ME = 0.51099895069


class MuTables:
    """Log of the macroscopic total and Compton attenuation (1/cm) per (material,
    density) row on ``n_table`` geometric energies in ``[e_min, e_max]``."""

    def __init__(self, cross_sections, rows, e_min: float, e_max: float, n_table: int) -> None:
        energies = np.geomspace(e_min, max(e_max, e_min * (1.0 + 1e-9)), n_table)
        self.log_e = np.log(energies)
        self.log_total = np.stack([np.log(rho * np.asarray(cross_sections.mu_total_bulk(energies, m)))
                                   for m, rho in rows])
        self.log_compton = np.stack([np.log(np.maximum(rho * np.asarray(cross_sections.mu_channels(energies, m)[0]),
                                                       1e-300)) for m, rho in rows])

    def total(self, row, e) -> np.ndarray:
        return self._read(self.log_total, row, e)

    def compton(self, row, e) -> np.ndarray:
        return self._read(self.log_compton, row, e)

    def _read(self, table, row, e) -> np.ndarray:
        le = np.log(np.asarray(e, dtype=np.float64))
        rows = np.broadcast_to(np.asarray(row), le.shape)
        out = np.empty(le.shape)
        for r in np.unique(rows):
            m = rows == r
            out[m] = np.exp(np.interp(le[m], self.log_e, table[int(r)]))
        return out


def kahn_ratios(e, gen) -> np.ndarray:
    """Klein-Nishina energy ratios E'/E by Kahn's rejection (PENELOPE-2018 sec. 2.3)."""
    e = np.asarray(e, dtype=np.float64)
    alpha = e / ME
    two_alpha = 2.0 * alpha
    p_branch = (1.0 + two_alpha) / (9.0 + two_alpha)
    out = np.empty_like(e)
    todo = np.ones(e.shape, dtype=bool)
    while todo.any():
        k = np.flatnonzero(todo)
        u1, u2, u3 = gen.random((3, k.size))
        ta = two_alpha[k]
        first = u1 <= p_branch[k]
        x = np.where(first, 1.0 + ta * u2, (1.0 + ta) / (1.0 + ta * u2))
        cos_t = 1.0 - (x - 1.0) / alpha[k]
        ok = np.where(first, u3 <= 4.0 * (x - 1.0) / (x * x), u3 <= 0.5 * (cos_t * cos_t + 1.0 / x))
        out[k[ok]] = 1.0 / x[ok]
        todo[k[ok]] = False
    return out


def photon_cos(e, ratio) -> np.ndarray:
    alpha = np.asarray(e, dtype=np.float64) / ME
    return np.clip(1.0 + 1.0 / alpha - 1.0 / (alpha * ratio), -1.0, 1.0)


def electron_cos(e, ratio) -> np.ndarray:
    e = np.asarray(e, dtype=np.float64)
    recoil = e * (1.0 - ratio)
    p_e = np.sqrt(recoil * (recoil + 2.0 * ME))
    return np.clip((e - ratio * e * photon_cos(e, ratio)) / p_e, -1.0, 1.0)


def rotate(d, cos_t, phi) -> np.ndarray:
    """Directions turned by polar cosine ``cos_t`` and azimuth ``phi`` (PENELOPE-2018
    eq. 1.131), with the pole branch when the transverse part vanishes."""
    d = np.asarray(d, dtype=np.float64)
    ux, uy, uz = d[:, 0], d[:, 1], d[:, 2]
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    cp, sp = np.cos(phi), np.sin(phi)
    tr2 = ux * ux + uy * uy
    pole = tr2 <= 1.0e-20
    tr = np.sqrt(np.where(pole, 1.0, tr2))
    k = sin_t / tr
    vx = ux * cos_t + k * (ux * uz * cp - uy * sp)
    vy = uy * cos_t + k * (uy * uz * cp + ux * sp)
    vz = uz * cos_t - tr * sin_t * cp
    sign = np.where(uz > 0.0, 1.0, -1.0)
    return np.stack([np.where(pole, sin_t * cp, vx), np.where(pole, sign * sin_t * sp, vy),
                     np.where(pole, sign * cos_t, vz)], axis=1)


def roulette(w, floor: float, gen) -> np.ndarray:
    """Below ``floor`` a weight survives at ``floor`` with probability ``w / floor``."""
    if floor <= 0.0:
        return w
    play = (w > 0.0) & (w < floor)
    if not play.any():
        return w
    live = gen.random(w.shape[0]) < w / floor
    out = np.where(play & live, floor, w)
    out[play & ~live] = 0.0
    return out


def _to_exit(parts, code, e, p, d, w, frame, exit_z, cutoff) -> float:
    pl, dl = frame.to_local(p, d)
    fwd = dl[:, 2] > 0.0
    t = np.where(fwd, (exit_z - pl[:, 2]) / np.where(fwd, dl[:, 2], 1.0), 0.0)
    keep = fwd & (t >= 0.0) & (e >= cutoff) & (w > 0.0)
    dropped = float(np.sum(w[~keep] * e[~keep]))
    if keep.any():
        q = p + t[:, None] * d
        parts.append(dict(particle_type=np.full(int(keep.sum()), code, np.int32), energy=e[keep], x=q[keep, 0],
                          y=q[keep, 1], z=q[keep, 2], ux=d[keep, 0], uy=d[keep, 1], uz=d[keep, 2], weight=w[keep]))
    return dropped


def presolve(columns: dict, *, cross_sections, max_energy: float, seed: int, exit_z: float, stack=None, frame=None,
             air=None, mode: str = "first_compton", pcut: float = 0.05, ecut: float = 0.2,
             weight_floor: float = 1.0e-3, e_min: float = 0.010, n_table: int = 512) -> dict:
    """The exit-plane population of the photon primaries ``columns`` (sample_batch
    columns), normalised per primary; ``air`` is a dict (z_start, z_end, material,
    density). Returns the population columns and the dropped weight-energy."""
    frame = stack.frame if stack is not None else frame
    e = np.asarray(columns["energy"], dtype=np.float64)
    n = e.size
    p = np.column_stack([columns["x"], columns["y"], columns["z"]]).astype(np.float64)
    d = np.column_stack([columns["ux"], columns["uy"], columns["uz"]]).astype(np.float64)
    w0 = np.asarray(columns["weight"], dtype=np.float64)
    gen = np.random.Generator(np.random.PCG64(np.random.SeedSequence((seed, 0x48454144))))
    rows = [] if stack is None else list(zip(stack.materials, stack.densities))
    air_row = len(rows)
    if air is not None:
        rows.append((air["material"], air["density"]))
    mu = MuTables(cross_sections, rows, e_min, max_energy, n_table) if rows else None
    parts: list = []
    dropped = 0.0
    transmitted = w0.copy()
    if stack is not None:
        t = stack.path_lengths(p, d)
        taus = np.column_stack([mu.total(i, e) * t[:, i] for i in range(len(stack.devices))])
        tau = taus.sum(axis=1)
        transmitted = w0 * np.exp(-tau)
        if mode == "first_compton" and np.any(tau > 0.0):
            k = np.flatnonzero(tau > 0.0)
            share = np.cumsum(taus[k], axis=1) / tau[k][:, None]
            dev = np.argmax(gen.random(k.size)[:, None] <= share, axis=1)
            z_mid = np.array([0.5 * (dv.z_top + dv.z_bottom) for dv in stack.devices])
            pl, dl = frame.to_local(p[k], d[k])
            site = p[k] + ((z_mid[dev] - pl[:, 2]) / dl[:, 2])[:, None] * d[k]
            ek = e[k]
            cf = mu.compton(dev, ek) / mu.total(dev, ek)
            w_int = w0[k] * (1.0 - np.exp(-tau[k]))
            ratio = kahn_ratios(ek, gen)
            e_out = ratio * ek
            phi = 2.0 * np.pi * gen.random(k.size)
            d_out = rotate(d[k], photon_cos(ek, ratio), phi)
            t_out = stack.path_lengths(site, d_out, from_origin=True)
            tau_out = np.column_stack([mu.total(i, e_out) * t_out[:, i] for i in range(len(stack.devices))]).sum(axis=1)
            w_s = roulette(w_int * cf * np.exp(-tau_out), weight_floor, gen)
            dropped += _to_exit(parts, 1, e_out, site, d_out, w_s, frame, exit_z, pcut)
    if air is not None:
        _, dl = frame.to_local(p, d)
        fwd = dl[:, 2] > 0.0
        length = np.where(fwd, (air["z_end"] - air["z_start"]) / np.where(fwd, dl[:, 2], 1.0), 0.0)
        mu_air = mu.total(air_row, e)
        tau_air = mu_air * length
        before = transmitted.copy()
        transmitted = transmitted * np.exp(-tau_air)
        if mode == "first_compton":
            k = np.flatnonzero(fwd & (tau_air > 0.0))
            if k.size:
                ek = e[k]
                pair_w = before[k] * (1.0 - np.exp(-tau_air[k])) * mu.compton(air_row, ek) / mu_air[k]
                u = gen.random(k.size)
                depth = -np.log(1.0 - u * (1.0 - np.exp(-tau_air[k]))) / mu_air[k]
                pl, dl_k = frame.to_local(p[k], d[k])
                site = p[k] + ((air["z_start"] - pl[:, 2]) / dl_k[:, 2] + depth)[:, None] * d[k]
                ratio = kahn_ratios(ek, gen)
                phi = 2.0 * np.pi * gen.random(k.size)
                d_ph = rotate(d[k], photon_cos(ek, ratio), phi)
                d_el = rotate(d[k], electron_cos(ek, ratio), phi + np.pi)
                dropped += _to_exit(parts, 1, ratio * ek, site, d_ph, roulette(pair_w, weight_floor, gen), frame,
                                    exit_z, pcut)
                dropped += _to_exit(parts, 2, ek - ratio * ek, site, d_el, roulette(pair_w, weight_floor, gen), frame,
                                    exit_z, ecut)
    dropped += _to_exit(parts, 1, e, p, d, transmitted, frame, exit_z, pcut)
    if not parts:
        raise ValueError("the pre-solve kept no particles")
    pop = {k: np.concatenate([x[k] for x in parts]) for k in parts[0]}
    pop["weight"] = pop["weight"] * (pop["weight"].size / n)
    return dict(population=pop, dropped=dropped, n_primaries=n)
