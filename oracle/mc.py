"""ORACLE (test infrastructure only): driver of the C transport restatement.

Builds the oracle's own tables (:mod:`oracle.xs_ref`) and GS grid
(:mod:`oracle.gs_ref`), runs the C history-stack transport (``mc_oracle.c``) batch by
batch, and scores exactly like the reference's ``BatchedDoseScorer`` /
``BatchedBeamletScorer`` / ``DijAssembler`` (``pyradmc/scoring/dose.py:169-207``,
``pyradmc/scoring/dij.py:155-189, 351-366, 397-434``), in float64 NumPy.

Also restates the device fold arithmetic on fixed-point quanta
(``pyradmc/backends/warp/kernels.py:771-909``) and the truncation rule, used to check
the HIP fold/finalize/truncation bit for bit on synthetic quanta maps, and the
planning product ``dose_for_weights`` (``pyradmc/scoring/dij.py:251-260``).
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess
from ctypes import POINTER, Structure, c_double, c_int, c_int32, c_int64, c_uint64
from pathlib import Path

import numpy as np

from oracle import gs_ref, xs_ref

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
CACHE = HERE / "_cache"

_dp = POINTER(c_double)


class OmConfig(Structure):
    _fields_ = [
        ("nx", c_int), ("ny", c_int), ("nz", c_int),
        ("ox", c_double), ("oy", c_double), ("oz", c_double),
        ("sx", c_double), ("sy", c_double), ("sz", c_double),
        ("density", _dp), ("material", POINTER(c_int32)),
        ("snx", c_int), ("sny", c_int), ("snz", c_int),
        ("sox", c_double), ("soy", c_double), ("soz", c_double),
        ("ssx", c_double), ("ssy", c_double), ("ssz", c_double),
        ("mass", _dp),
        ("n", c_int), ("n_mat", c_int),
        ("majorant", _dp), ("mu_c", _dp), ("mu_p", _dp), ("mu_pp", _dp), ("mu_r", _dp),
        ("csda", _dp), ("moller", _dp), ("rrange", _dp), ("srad", _dp), ("tpow", _dp),
        ("log_eta", _dp), ("inv_range", _dp),
        ("p_log_min", c_double), ("p_inv", c_double), ("e_log_min", c_double),
        ("e_inv", c_double), ("r_log_min", c_double), ("r_inv", c_double),
        ("gs", _dp), ("gs_ix0", c_int), ("gs_iy0", c_int), ("gs_neta", c_int),
        ("gs_nth", c_int), ("gs_nu", c_int),
        ("src_kind", c_int), ("particle", c_int), ("energy", c_double),
        ("pos", c_double * 3), ("dir", c_double * 3),
        ("z0", c_double), ("x0", c_double), ("xe", c_double), ("y0", c_double), ("ye", c_double),
        ("focal", c_double * 3), ("ua", c_double * 3), ("va", c_double * 3),
        ("wu", c_double), ("wv", c_double),
        ("n_bins", c_int), ("edges", _dp), ("cdf", _dp),
        ("n_beamlets", c_int), ("beamlet", _dp),
        ("pcut", c_double), ("ecut", c_double), ("step_fraction", c_double),
        ("resolution", c_double), ("transport_electrons", c_int), ("msc_gs", c_int),
        ("srest", _dp), ("dose_to_water", c_int),
        ("n_prim", c_int64), ("prim_kind", POINTER(c_int32)), ("prim", _dp),
        ("sigma_u", c_double), ("sigma_v", c_double),
    ]


class OmBooks(Structure):
    _fields_ = [
        ("emitted", c_double), ("deposited", c_double), ("escaped", c_double),
        ("unscored", c_double), ("vox", c_int64), ("deposits", c_int64), ("pushes", c_int64),
        ("photon_steps", c_int64), ("electron_steps", c_int64),
    ]


def build(verbose: bool = False) -> Path:
    """Compile ``mc_oracle.c`` with gcc into ``oracle/liboracle.so``."""
    tmp = HERE / f"liboracle.so.{os.getpid()}.tmp"  # concurrent builders each rename their own
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-pthread", "-o", str(tmp),
           str(HERE / "mc_oracle.c"), "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.is_file():
            build()
        h = ctypes.CDLL(str(LIB))
        h.om_batch.restype = c_int
        h.om_batch.argtypes = [POINTER(OmConfig), c_uint64, c_int, POINTER(c_int64),
                               POINTER(c_int64), c_int, _dp, POINTER(OmBooks)]
        _lib = h
    return _lib


def _p(a: np.ndarray, t=c_double):
    return a.ctypes.data_as(POINTER(t))


def gs_grid(tables: dict, materials=(0,), workers: int = 8) -> tuple[np.ndarray, int, int]:
    """The oracle's GS grid for the tables' window, cached under ``oracle/_cache``."""
    log_eta = np.asarray(tables["log_eta"])[list(materials)]
    ceiling = float(np.max(np.asarray(tables["scattering_power"])[list(materials)].max(axis=1)
                           * np.asarray(tables["restricted_range"])[list(materials)].max(axis=1)))
    ix0, ix1, iy0, iy1 = gs_ref.window(float(log_eta.min()), float(log_eta.max()), ceiling)
    path = CACHE / f"gs_{ix0}_{ix1}_{iy0}_{iy1}.npy"
    if path.is_file():
        values = np.load(path, allow_pickle=False)
    else:
        values = gs_ref.grid(ix0, ix1, iy0, iy1, workers=workers)
        CACHE.mkdir(exist_ok=True)
        tmp = path.with_suffix(f".partial-{os.getpid()}.npy")
        np.save(tmp, values)
        os.replace(tmp, path)
    return np.ascontiguousarray(values, dtype=np.float64), ix0, iy0


class Oracle:
    """CPU transport of one configuration; mirrors ReferenceEngine.run / run_dij."""

    def __init__(self, shape, spacing, origin, density, material, geometry_densities,
                 source: dict, *, pcut=0.05, ecut=0.2, transport_electrons=True, msc_model="gs",
                 step_energy_fraction=None, deposit_resolution_cm=None, scoring=None,
                 gs=None, tables=None, scoring_mode="dose_to_medium") -> None:
        """``tables``: the oracle's own flattened tables (a dict as
        :meth:`oracle.xs_ref.WaterOracle.tables` returns, e.g. from
        :mod:`oracle.tab_ref`); analytic water when omitted. ``scoring_mode``
        ``"dose_to_water"`` weights each deposit by the water/medium restricted
        stopping-power ratio."""
        self._keep = []
        c = OmConfig()
        c.nx, c.ny, c.nz = (int(v) for v in shape)
        c.ox, c.oy, c.oz = (float(v) for v in origin)
        c.sx, c.sy, c.sz = (float(v) for v in spacing)
        dens = self._hold(np.ascontiguousarray(density, dtype=np.float64))
        mat = self._hold(np.ascontiguousarray(material, dtype=np.int32))
        c.density, c.material = _p(dens), _p(mat, c_int32)
        if scoring is None:
            scoring = dict(shape=shape, spacing=spacing, origin=origin,
                           mass=np.asarray(density, dtype=np.float64) * float(np.prod(spacing)))
        self.scoring = scoring
        c.snx, c.sny, c.snz = (int(v) for v in scoring["shape"])
        c.sox, c.soy, c.soz = (float(v) for v in scoring["origin"])
        c.ssx, c.ssy, c.ssz = (float(v) for v in scoring["spacing"])
        self.mass = self._hold(np.ascontiguousarray(scoring["mass"], dtype=np.float64).reshape(-1))
        c.mass = _p(self.mass)
        e_max = float(source.get("max_energy", source.get("energy", 0.0))) * (1.0 + 1.0e-6)
        if tables is None:
            t = xs_ref.WaterOracle(geometry_densities).tables(ecut, pcut, e_max)
        else:
            t = tables
            if abs(t["e_max"] - e_max) > 1e-12 * e_max or t["ecut"] != ecut or t["pcut"] != pcut:
                raise ValueError("tables were flattened for other cuts or another e_max")
        self.tables = t
        n_mat = int(np.asarray(t["mu_compton"]).shape[0])
        if int(np.max(material)) >= n_mat or int(np.min(material)) < 0:
            raise ValueError("a voxel material has no table row")
        c.n, c.n_mat = int(t["n_points"]), n_mat
        for name, key in (("majorant", "majorant"), ("mu_c", "mu_compton"), ("mu_p", "mu_photo"),
                          ("mu_pp", "mu_pair"), ("mu_r", "mu_rayleigh"), ("csda", "csda_range"),
                          ("moller", "moller"), ("rrange", "restricted_range"),
                          ("srad", "stopping_radiative"), ("tpow", "scattering_power"),
                          ("log_eta", "log_eta"), ("inv_range", "energy_of_restricted_range"),
                          ("srest", "stopping_restricted")):
            setattr(c, name, _p(self._hold(np.ascontiguousarray(t[key], dtype=np.float64).reshape(-1))))
        c.p_log_min, c.p_inv = t["photon_log_e_min"], t["photon_inv_dlog"]
        c.e_log_min, c.e_inv = t["electron_log_e_min"], t["electron_inv_dlog"]
        c.r_log_min, c.r_inv = t["range_log_r_min"], t["range_inv_dlog"]
        c.msc_gs = 1 if msc_model == "gs" else 0
        if c.msc_gs:
            present = sorted(set(np.unique(np.asarray(material)).astype(int).tolist()))
            values, ix0, iy0 = gs if gs is not None else gs_grid(t, materials=present)
            values = self._hold(np.ascontiguousarray(values, dtype=np.float64))
            c.gs = _p(values)
            c.gs_ix0, c.gs_iy0 = int(ix0), int(iy0)
            c.gs_neta, c.gs_nth, c.gs_nu = (int(v) for v in values.shape)
        kinds = {"pencil": 0, "parallel": 1, "parallel_spectral": 2, "fan": 3, "lattice": 4, "given": 5,
                 "spot": 6}
        c.src_kind = kinds[source["kind"]]
        c.particle = {"photon": 0, "electron": 1}[source.get("particle", "photon")]
        c.energy = float(source.get("energy", 0.0))
        if "position" in source:
            c.pos[:] = source["position"]
            c.dir[:] = source["direction"]
        for key in ("z0", "x0", "xe", "y0", "ye", "wu", "wv", "sigma_u", "sigma_v"):
            if key in source:
                setattr(c, key, float(source[key]))
        for key in ("focal", "ua", "va"):
            if key in source:
                getattr(c, key)[:] = source[key]
        if "edges" in source:
            edges = self._hold(np.ascontiguousarray(source["edges"], dtype=np.float64))
            cdf = self._hold(np.ascontiguousarray(source["cdf"], dtype=np.float64))
            c.n_bins, c.edges, c.cdf = int(cdf.size), _p(edges), _p(cdf)
        if source["kind"] == "given":
            # source["primaries"]: sample_batch-style columns indexed by history key
            # (IAEA codes 1/2/3 -> photon, electron, positron; energy, x, y, z, ux, uy,
            # uz, weight).
            given = source["primaries"]
            code = np.asarray(given["particle_type"]).reshape(-1)
            pk = self._hold(np.ascontiguousarray(code - 1, dtype=np.int32))
            if np.any((pk < 0) | (pk > 2)):
                raise ValueError("given primaries must be photons, electrons or positrons (codes 1-3)")
            cols = np.stack([np.asarray(given[k], dtype=np.float64).reshape(-1)
                             for k in ("energy", "x", "y", "z", "ux", "uy", "uz", "weight")])
            prim = self._hold(np.ascontiguousarray(cols))
            c.n_prim, c.prim_kind, c.prim = int(code.size), _p(pk, c_int32), _p(prim)
        if "beamlets" in source:
            b = self._hold(np.ascontiguousarray(source["beamlets"], dtype=np.float64))
            c.n_beamlets, c.beamlet = int(b.shape[0]), _p(b)
        c.pcut, c.ecut = float(pcut), float(ecut)
        c.step_fraction = float(step_energy_fraction if step_energy_fraction is not None
                                else (0.20 if msc_model == "gs" else 0.05))
        c.resolution = 0.0 if deposit_resolution_cm is None else float(deposit_resolution_cm)
        c.transport_electrons = 1 if transport_electrons else 0
        if scoring_mode not in ("dose_to_medium", "dose_to_water"):
            raise ValueError(f"unknown scoring_mode {scoring_mode!r}")
        c.dose_to_water = 1 if scoring_mode == "dose_to_water" else 0
        self.cfg = c
        self.n_vox = int(np.prod(scoring["shape"]))

    def _hold(self, a):
        self._keep.append(a)
        return a

    def _batch(self, seed, first, count, threads):
        n_cols = len(first)
        energy = np.zeros(n_cols * self.n_vox)
        books = OmBooks()
        f = np.ascontiguousarray(first, dtype=np.int64)
        k = np.ascontiguousarray(count, dtype=np.int64)
        rc = lib().om_batch(ctypes.byref(self.cfg), c_uint64(seed), n_cols, _p(f, c_int64),
                            _p(k, c_int64), int(threads), _p(energy), ctypes.byref(books))
        if rc != 0:
            raise ValueError("history keys beyond the given primaries")
        return energy.reshape(n_cols, self.n_vox), books

    @staticmethod
    def _add(total: dict, books: OmBooks) -> None:
        for name, _ in OmBooks._fields_:
            total[name] = total.get(name, 0) + getattr(books, name)

    def run(self, n_histories: int, n_batches: int, seed: int, threads: int = 1) -> dict:
        """ReferenceEngine.run restated: history h (batch-major) keyed by h."""
        per = n_histories // n_batches
        covered = self.mass > 0.0
        s1 = np.zeros(self.n_vox)
        s2 = np.zeros(self.n_vox)
        books: dict = {}
        for b in range(n_batches):
            energy, bk = self._batch(seed, [b * per], [per], threads)
            self._add(books, bk)
            dose = np.divide(energy[0], self.mass * per, out=np.zeros(self.n_vox), where=covered)
            s1 += dose
            s2 += dose**2
        mean, sigma = finalize(s1, s2, n_batches)
        shape = tuple(self.scoring["shape"])
        return dict(dose=mean.reshape(shape), dose_sigma=sigma.reshape(shape), books=books,
                    n_batches=n_batches, n_histories=n_histories)

    def run_dij(self, counts, n_batches: int, seed: int, truncation: float, correlated=True,
                threads: int = 1) -> dict:
        """ReferenceEngine.run_dij restated (batch-outer, beamlet-inner), uneven counts
        allowed: beamlet j's r-th history has key r (correlated) or offset_j + r."""
        counts = np.asarray(counts, dtype=np.int64)
        per = counts // n_batches
        offsets = np.concatenate(([0], np.cumsum(counts)[:-1]))
        covered = self.mass > 0.0
        n_cols = counts.size
        s1 = np.zeros((n_cols, self.n_vox))
        s2 = np.zeros((n_cols, self.n_vox))
        books: dict = {}
        for b in range(n_batches):
            first = [int(b * per[j]) if correlated else int(offsets[j] + b * per[j]) for j in range(n_cols)]
            energy, bk = self._batch(seed, first, per, threads)
            self._add(books, bk)
            dose = np.divide(energy, self.mass[None, :] * per[:, None].astype(np.float64),
                             out=np.zeros_like(energy), where=covered[None, :])
            s1 += dose
            s2 += dose**2
        mean, sigma = finalize(s1, s2, n_batches)
        indptr, indices, d, s = truncate(mean, sigma, truncation)
        return dict(dense_dose=mean, dense_sigma=sigma, indptr=indptr, indices=indices, dose=d,
                    sigma=s, books=books)


# --------------------------------------------------------------- scoring restatement


def finalize(s1: np.ndarray, s2: np.ndarray, n: int):
    """Batch mean and standard error (pyradmc/scoring/dose.py:188-207)."""
    mean = s1 / n
    if n == 1:
        return mean, np.zeros_like(mean)
    variance = np.maximum(0.0, s2 / n - mean**2) * n / (n - 1)
    return mean, np.sqrt(variance / n)


def fold_quanta(quanta: np.ndarray, mass: np.ndarray, nhist, s1: np.ndarray, s2: np.ndarray):
    """Fold one batch of int64 quanta maps [columns, voxels] in place; the device fold's
    arithmetic ``d = (q * 1e-9) / (mass * nhist)``, zero where mass <= 0
    (pyradmc/backends/warp/kernels.py:771-878)."""
    q = np.asarray(quanta, dtype=np.int64)
    nh = np.asarray(nhist, dtype=np.float64).reshape(-1, 1)
    energy = q.astype(np.float64) * 1.0e-9
    denom = mass[None, :] * nh
    d = np.where(mass[None, :] > 0.0, energy / np.where(denom > 0.0, denom, 1.0), 0.0)
    d = np.where(q != 0, d, 0.0)
    s1 += d
    s2 += d * d
    return int(q.sum())


def truncate(mean: np.ndarray, sigma: np.ndarray, truncation: float):
    """DijAssembler.add_block + finalize: keep m > 0 and m >= t * column max, CSC."""
    idx, dose, sig, counts = [], [], [], []
    for row_d, row_s in zip(mean, sigma):
        col_max = float(row_d.max(initial=0.0))
        keep = (row_d >= truncation * col_max) & (row_d > 0.0)
        k = np.flatnonzero(keep)
        idx.append(k.astype(np.int64))
        dose.append(row_d[k])
        sig.append(row_s[k])
        counts.append(k.size)
    indptr = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    return indptr, np.concatenate(idx), np.concatenate(dose), np.concatenate(sig)


def source_dict(source) -> dict:
    """Describe a pyradmc_amd-style source object (attributes only) for the oracle."""
    name = type(source).__name__
    if name == "PencilBeamSource":
        return dict(kind="pencil", energy=source.energy, max_energy=source.energy,
                    position=list(source.position), direction=list(source.direction))
    if name == "ParallelBeamSource":
        return dict(kind="parallel", energy=source.energy, max_energy=source.energy, z0=source.z,
                    x0=source.x_range[0], xe=source.x_range[1] - source.x_range[0],
                    y0=source.y_range[0], ye=source.y_range[1] - source.y_range[0])
    if name == "SpectralParallelBeamSource":
        return dict(kind="parallel_spectral", max_energy=source.spectrum.max_energy, z0=source.z,
                    x0=source.x_range[0], xe=source.x_range[1] - source.x_range[0],
                    y0=source.y_range[0], ye=source.y_range[1] - source.y_range[0],
                    edges=np.asarray(source.spectrum.edges), cdf=np.asarray(source.spectrum.cdf))
    if name in ("GaussianSpotBeamSource", "GaussianSpotBeamletSource"):
        return dict(kind="spot", max_energy=source.spectrum.max_energy, focal=list(source.focal_point),
                    ua=list(source.u_axis), va=list(source.v_axis), wu=source.width_u, wv=source.width_v,
                    sigma_u=source.sigma_u, sigma_v=source.sigma_v, edges=np.asarray(source.spectrum.edges),
                    cdf=np.asarray(source.spectrum.cdf), beamlets=np.asarray(source.centers, dtype=np.float64))
    if name in ("SpectralBeamSource", "SpectralBeamletSource"):
        centers = source.centers
        return dict(kind="fan", max_energy=source.spectrum.max_energy, focal=list(source.focal_point),
                    ua=list(source.u_axis), va=list(source.v_axis), wu=source.width_u,
                    wv=source.width_v, edges=np.asarray(source.spectrum.edges),
                    cdf=np.asarray(source.spectrum.cdf), beamlets=np.asarray(centers, dtype=np.float64))
    if name == "BeamletGridSource":
        rows = []
        for j in range(source.n_beamlets):
            x0, x1, y0, y1 = source.beamlet_bounds(j)
            rows.append((x0, x1 - x0, y0, y1 - y0))
        return dict(kind="lattice", energy=source.energy, max_energy=source.energy, z0=source.z,
                    beamlets=np.asarray(rows, dtype=np.float64))
    raise ValueError(f"oracle has no generator for {name}")


def for_grid(grid, source, **kw) -> Oracle:
    """Convenience: an Oracle over a VoxelGrid-like object and a source object."""
    return Oracle(grid.shape, grid.spacing, grid.origin, grid.density, grid.material,
                  grid.max_density_by_material(), source_dict(source), **kw)


def dose_for_weights(indptr: np.ndarray, indices: np.ndarray, dose: np.ndarray, weights: np.ndarray,
                     n_voxels: int) -> np.ndarray:
    """DijResult.dose_for_weights (pyradmc/scoring/dij.py:251-260) for one weight
    vector [n_beamlets] or a matrix [n_beamlets][n_w]: float64 products of the float32
    entries and the weights, added column by column in ascending order (the entries of
    one column in CSC order). Returns [n_voxels] or [n_voxels][n_w] float64."""
    w = np.asarray(weights, dtype=np.float64)
    single = w.ndim == 1
    w2 = w[:, None] if single else w
    out = np.zeros((n_voxels, w2.shape[1]), dtype=np.float64)
    for j in range(w2.shape[0]):
        lo, hi = int(indptr[j]), int(indptr[j + 1])
        np.add.at(out, indices[lo:hi], np.asarray(dose[lo:hi])[:, None] * w2[j][None, :])
    return out[:, 0] if single else out
