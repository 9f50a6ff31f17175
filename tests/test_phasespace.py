"""IAEA phase-space reader, writer and sources (CPU), and the oracle's given-primaries
source that the phase-space GPU tests check against.

The reader is pinned to the IAEA(NDS)-0484 layout by records authored byte by byte with
``struct`` and hand-computed fields (the cases of pyradmc's tests/unit/test_iaea_phsp.py),
not by the writer, so a shared mistake in reader and writer cannot pass."""

from __future__ import annotations

import math
import struct
import warnings
from pathlib import Path

import numpy as np
import pytest

from pyradmc_amd.phasespace import (
    IAEAPhaseSpace,
    InMemoryPhaseSpaceSource,
    PhaseSpaceSource,
    read_iaea_header,
    write_iaea_phase_space,
)

PHOTON, ELECTRON, POSITRON = 1, 2, 3
FULL = 1 + 4 + 4 * 6  # typ, E, x y z u v weight
ME = 0.51099895069


def _header(*, stored=(1, 1, 1, 1, 1, 1, 1), constants="", n_f=0, n_i=0, byte_order=1234,
            length, particles) -> str:
    flags = "\n".join(f"    {s}     // flag" for s in stored)
    return (f"$IAEA_INDEX:\n1000\n\n$FILE_TYPE:\n0\n\n$RECORD_CONTENTS:\n{flags}\n"
            f"    {n_f}     // Extra floats stored ?\n    {n_i}     // Extra longs stored ?\n\n"
            f"$RECORD_CONSTANT:\n{constants}\n$RECORD_LENGTH:\n{length}\n\n"
            f"$BYTE_ORDER:\n{byte_order}\n\n$PARTICLES:\n{particles}\n")


def _files(tmp: Path, header: str, payload: bytes, stem: str = "case") -> Path:
    (tmp / f"{stem}.IAEAheader").write_text(header)
    (tmp / f"{stem}.IAEAphsp").write_bytes(payload)
    return tmp / f"{stem}.IAEAphsp"


def _one(typ, e, x=0.0, y=0.0, z=0.0, u=0.0, v=0.0, weight=1.0, order="<") -> bytes:
    return struct.pack(order + "bfffffff", typ, e, x, y, z, u, v, weight)


def _read_all(path) -> list:
    with IAEAPhaseSpace(path) as ps:
        return list(ps)


# ---------------------------------------------------------------- reader, by hand


def test_photon_record_with_an_extra_int(tmp_path):
    # typ -1: photon with w < 0; E -1: first particle of a history, 1 MeV; u = 0.6 and
    # v = 0 give |w| = 0.8 exactly; one extra int 13 (record of 33 bytes).
    payload = struct.pack("<bfffffffi", -PHOTON, -1.0, 3.0, 4.0, 5.0, 0.6, 0.0, 2.0, 13)
    assert len(payload) == 33
    path = _files(tmp_path, _header(n_i=1, length=33, particles=1), payload)
    (r,) = _read_all(path)
    assert (r.particle_type, r.new_history, r.energy) == (PHOTON, True, 1.0)
    assert (r.x, r.y, r.z, r.weight) == (3.0, 4.0, 5.0, 2.0)
    assert r.u == pytest.approx(0.6, abs=1e-7) and r.v == 0.0
    assert r.w == pytest.approx(-0.8, abs=1e-7)
    assert r.extra_ints == (13,) and r.extra_floats == ()


def test_positive_type_and_energy(tmp_path):
    path = _files(tmp_path, _header(length=FULL, particles=1), _one(ELECTRON, 2.5, v=0.8))
    (r,) = _read_all(path)
    assert (r.particle_type, r.new_history, r.energy) == (ELECTRON, False, 2.5)
    assert r.w == pytest.approx(0.6, abs=1e-7)


def test_extra_floats_precede_extra_ints(tmp_path):
    payload = struct.pack("<bfffffffffi", PHOTON, 1.0, 0, 0, 0, 0, 0, 1.0, 0.25, -3.5, 7)
    path = _files(tmp_path, _header(n_f=2, n_i=1, length=FULL + 12, particles=1), payload)
    (r,) = _read_all(path)
    assert r.extra_floats == (0.25, -3.5) and r.extra_ints == (7,)


def test_constant_coordinate_comes_from_the_header(tmp_path):
    # z not stored: the record is 4 bytes shorter and every z is the header constant.
    payload = struct.pack("<bffffff", PHOTON, 6.0, 1.0, 2.0, 0.0, 0.0, 1.0)
    hdr = _header(stored=(1, 1, 0, 1, 1, 1, 1), constants="-7.5     // z\n", length=FULL - 4,
                  particles=1)
    (r,) = _read_all(_files(tmp_path, hdr, payload))
    assert (r.x, r.y, r.z, r.w) == (1.0, 2.0, -7.5, 1.0)


def test_two_constants_in_coordinate_order(tmp_path):
    # x and weight constant: RECORD_CONSTANT lists them in x, y, z, u, v, weight order.
    payload = struct.pack("<bfffff", PHOTON, 6.0, 2.0, 3.0, 0.0, 0.0)
    hdr = _header(stored=(0, 1, 1, 1, 1, 1, 0), constants="1.25\n0.5\n", length=FULL - 8, particles=1)
    (r,) = _read_all(_files(tmp_path, hdr, payload))
    assert (r.x, r.y, r.z, r.weight) == (1.25, 2.0, 3.0, 0.5)


def test_history_flags_and_types(tmp_path):
    payload = _one(PHOTON, -1.0) + _one(ELECTRON, 0.3) + _one(-POSITRON, 0.2)
    recs = _read_all(_files(tmp_path, _header(length=FULL, particles=3), payload))
    assert [r.new_history for r in recs] == [True, False, False]
    assert [r.particle_type for r in recs] == [PHOTON, ELECTRON, POSITRON]


def test_header_counts(tmp_path):
    (tmp_path / "h.IAEAheader").write_text(_header(length=FULL, particles=7))
    h = read_iaea_header(tmp_path / "h.IAEAheader")
    assert (h.n_particles, h.record_length, h.byte_order) == (7, FULL, "<")


def test_big_endian(tmp_path):
    path = _files(tmp_path, _header(byte_order=4321, length=FULL, particles=1),
                  _one(PHOTON, 4.0, 1.0, 2.0, 3.0, order=">"))
    (r,) = _read_all(path)
    assert (r.energy, r.x, r.y, r.z) == (4.0, 1.0, 2.0, 3.0)


def test_degenerate_direction_is_renormalised(tmp_path):
    (r,) = _read_all(_files(tmp_path, _header(length=FULL, particles=1), _one(PHOTON, 1.0, u=0.8, v=0.8)))
    assert not math.isnan(r.w) and r.w == 0.0
    assert r.u == pytest.approx(math.sqrt(0.5), abs=1e-7) and r.u ** 2 + r.v ** 2 == pytest.approx(1.0)


@pytest.mark.parametrize("case,match", [
    (dict(length=999), "RECORD_LENGTH"),
    (dict(stored=(1, 1, 1, 1, 1, 0, 1)), "must store w"),
    (dict(stored=(1, 1, 0, 1, 1, 1, 1), length=FULL - 4), "RECORD_CONSTANT has 0 values"),
    (dict(byte_order=2143), "BYTE_ORDER"),
])
def test_corrupt_headers_are_rejected(tmp_path, case, match):
    kw = dict(length=FULL, particles=1)
    kw.update(case)
    (tmp_path / "bad.IAEAheader").write_text(_header(**kw))
    with pytest.raises(ValueError, match=match):
        read_iaea_header(tmp_path / "bad.IAEAheader")


def test_truncated_data_is_rejected(tmp_path):
    path = _files(tmp_path, _header(length=FULL, particles=1), _one(PHOTON, 1.0)[:-3])
    with pytest.raises(ValueError, match="not a multiple of record length"):
        read_iaea_header(path)


def test_stale_particle_count_warns_and_the_file_wins(tmp_path):
    path = _files(tmp_path, _header(length=FULL, particles=2), _one(PHOTON, 1.0))
    with pytest.warns(UserWarning, match="PARTICLES header says 2"):
        assert read_iaea_header(path).n_particles == 1


def test_no_count_anywhere_is_an_error(tmp_path):
    text = _header(length=FULL, particles=1).replace("$PARTICLES:\n1\n", "")
    (tmp_path / "x.IAEAheader").write_text(text)
    with pytest.raises(ValueError, match="cannot determine particle count"):
        read_iaea_header(tmp_path / "x")


def test_iteration_needs_the_with_block(tmp_path):
    ps = IAEAPhaseSpace(_files(tmp_path, _header(length=FULL, particles=1), _one(PHOTON, 1.0)))
    with pytest.raises(RuntimeError):
        list(ps)


# ------------------------------------------------------------------ writer


def _population(n, seed=0):
    rng = np.random.default_rng(seed)
    ct = rng.uniform(-1, 1, n)
    st = np.sqrt(1 - ct ** 2)
    ph = rng.uniform(0, 2 * np.pi, n)
    return {"particle_type": rng.choice([1, 1, 1, 2, 3], n).astype(np.int32),
            "energy": rng.uniform(0.1, 6.0, n), "x": rng.normal(0, 3, n), "y": rng.normal(0, 3, n),
            "z": np.full(n, -1.0), "ux": st * np.cos(ph), "uy": st * np.sin(ph), "uz": ct,
            "weight": rng.uniform(0.2, 2.0, n)}


@pytest.mark.parametrize("order", [1234, 4321])
def test_writer_round_trip(tmp_path, order):
    cols = _population(1000)
    fresh = np.arange(1000) % 3 == 0
    path = write_iaea_phase_space(tmp_path / "pop", cols, new_history=fresh, byte_order=order,
                                  constants={"z": -1.0})
    h = read_iaea_header(path)
    assert h.record_length == FULL - 4 and h.constants == {"z": -1.0}
    with IAEAPhaseSpace(path) as ps:
        got = ps.columns()
        recs = list(ps)
    np.testing.assert_array_equal(got["particle_type"], cols["particle_type"])
    np.testing.assert_array_equal(got["new_history"], fresh)
    for k in ("energy", "x", "y", "weight", "ux", "uy"):
        np.testing.assert_array_equal(got[k], cols[k].astype(np.float32))
    np.testing.assert_array_equal(got["z"], np.float32(-1.0))
    # w is rebuilt from float32 u and v: its square is exact to the rounding of u and v
    np.testing.assert_allclose(got["uz"].astype(np.float64) ** 2, cols["uz"] ** 2, atol=5e-7)
    assert np.all(np.sign(got["uz"]) == np.sign(cols["uz"]))
    assert [r.energy for r in recs[:5]] == [float(np.float32(e)) for e in cols["energy"][:5]]
    assert all(r.new_history == f for r, f in zip(recs, fresh))


# ------------------------------------------------------------------ sources


def _write_recs(tmp, recs, stem="s"):
    cols = {k: np.array([r[i] for r in recs], dtype=np.float64)
            for i, k in enumerate(("particle_type", "energy", "x", "y", "z", "ux", "uy", "uz", "weight"))}
    cols["particle_type"] = cols["particle_type"].astype(np.int32)
    return write_iaea_phase_space(tmp / stem, cols)


def _rec(code, e, x=0.0, y=0.0, z=0.0, weight=1.0):
    return (code, e, x, y, z, 0.6, 0.0, 0.8, weight)


def test_source_counts_and_energy(tmp_path):
    src = PhaseSpaceSource(_write_recs(tmp_path, [_rec(1, 6.0), _rec(2, 1.0), _rec(1, 3.0)]))
    assert len(src) == 3 and src.max_energy == 6.0
    src.close()
    src.close()


def test_unsupported_particles_are_dropped_or_refused(tmp_path):
    path = _write_recs(tmp_path, [_rec(1, 6.0), _rec(4, 2.0), _rec(1, 3.0)])
    with pytest.warns(UserWarning, match="unsupported particle types"):
        src = PhaseSpaceSource(path)
    assert len(src) == 2
    b = src.sample_batch(3, 0, 2)  # within the population: no reuse warning
    assert set(b["particle_type"].tolist()) == {1}
    with pytest.raises(ValueError, match="unsupported particle type 4"):
        PhaseSpaceSource(path, skip_unsupported=False)
    only_bad = _write_recs(tmp_path, [_rec(5, 6.0)], stem="bad")
    with pytest.raises(ValueError, match="no transportable"), pytest.warns(UserWarning):
        PhaseSpaceSource(only_bad)


def test_rank_to_position_map_skips_unsupported_records(tmp_path):
    rng = np.random.default_rng(4)
    codes = rng.choice([1, 2, 3, 4, 5], 500, p=[0.6, 0.2, 0.1, 0.05, 0.05])
    codes[:3] = 4  # unsupported at the start and in runs
    codes[-2:] = 5
    recs = [_rec(int(c), 1.0 + i) for i, c in enumerate(codes)]
    with pytest.warns(UserWarning):
        src = PhaseSpaceSource(_write_recs(tmp_path, recs))
    valid = np.flatnonzero(np.isin(codes, (1, 2, 3)))
    assert len(src) == valid.size
    np.testing.assert_array_equal(src.positions(np.arange(valid.size)), valid)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        b = src.sample_batch(11, 0, 4000)
    # energies identify the file positions: every one is a transportable record's
    assert set(np.round(b["energy"] - 1.0).astype(int).tolist()) <= set(valid.tolist())


def test_sampling_law_and_chunk_invariance(tmp_path):
    n = 37
    recs = [_rec(1 + i % 3, 1.0 + i, x=float(i), weight=0.5 + 0.01 * i) for i in range(n)]
    src = PhaseSpaceSource(_write_recs(tmp_path, recs))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        whole = src.sample_batch(2024, 100, 300)
        parts = [src.sample_batch(2024, 100 + a, b - a) for a, b in ((0, 7), (7, 128), (128, 300))]
    for k, v in whole.items():
        np.testing.assert_array_equal(v, np.concatenate([p[k] for p in parts]))
    # history h takes record floor(U_h * n), U_h the h-th double of Generator(PCG64(seed))
    u = np.random.Generator(np.random.PCG64(2024)).random(400)[100:]
    np.testing.assert_array_equal(whole["x"], np.floor(u * n))
    np.testing.assert_array_equal(whole["energy"], np.floor(u * n).astype(np.float32) + 1)
    assert whole["particle_type"].dtype == np.int32 and whole["energy"].dtype == np.float32
    np.testing.assert_allclose(whole["ux"] ** 2 + whole["uy"] ** 2 + whole["uz"] ** 2, 1.0, atol=1e-6)


def test_reuse_warning_fires_once(tmp_path):
    src = PhaseSpaceSource(_write_recs(tmp_path, [_rec(1, 2.0), _rec(1, 4.0)]))
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        src.sample_batch(1, 0, 2)
    with pytest.warns(UserWarning, match="latent variance"):
        src.sample_batch(1, 0, 3)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        src.sample_batch(1, 3, 4)


def test_closed_source_refuses_to_sample(tmp_path):
    with PhaseSpaceSource(_write_recs(tmp_path, [_rec(1, 2.0)])) as src:
        pass
    with pytest.raises(RuntimeError, match="closed"):
        src.sample_batch(1, 0, 1)


def _mem_cols(n=5):
    return dict(particle_type=np.array([1, 2, 3, 1, 1][:n]), energy=np.linspace(1, 5, n),
                x=np.arange(n, dtype=float), y=-np.arange(n, dtype=float), z=np.zeros(n),
                ux=np.full(n, 0.6), uy=np.zeros(n), uz=np.full(n, 0.8), weight=np.linspace(0.5, 1, n))


@pytest.mark.parametrize("change,match", [
    (dict(energy=np.ones(4)), "length"),
    (dict(particle_type=np.array([1, 2, 4, 1, 1])), "IAEA codes"),
    (dict(energy=np.array([1.0, 0.0, 1, 1, 1])), "positive energy"),
    (dict(uz=np.full(5, 0.79)), "unit vectors"),
    (dict(weight=np.array([1.0, -1, 1, 1, 1])), "non-negative"),
    (dict(weight=np.array([1.0, np.nan, 1, 1, 1])), "finite"),
])
def test_in_memory_validation(change, match):
    cols = _mem_cols()
    cols.update(change)
    with pytest.raises(ValueError, match=match):
        InMemoryPhaseSpaceSource(**cols)
    with pytest.raises(ValueError, match="at least one"):
        InMemoryPhaseSpaceSource(*[np.zeros(0)] * 9)


def test_in_memory_rows_and_file_twin_sample_alike(tmp_path):
    cols = _mem_cols()
    mem = InMemoryPhaseSpaceSource(**cols)
    assert len(mem) == 5 and mem.max_energy == 5.0
    got = mem.columns()
    got["x"][0] = 99.0  # a copy
    assert mem.columns()["x"][0] == 0.0
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        b = mem.sample_batch(7, 0, 64)
    rows = np.round(b["x"]).astype(int)
    np.testing.assert_array_equal(b["particle_type"], cols["particle_type"][rows])
    np.testing.assert_array_equal(b["weight"], cols["weight"][rows].astype(np.float32))
    # the same population written to a file samples the same records
    path = write_iaea_phase_space(tmp_path / "twin", cols)
    with PhaseSpaceSource(path) as f, warnings.catch_warnings():
        warnings.simplefilter("ignore")
        fb = f.sample_batch(7, 0, 64)
    for k in ("particle_type", "energy", "x", "y", "z", "weight"):
        np.testing.assert_array_equal(fb[k], b[k])
    np.testing.assert_allclose(fb["uz"], b["uz"], atol=1e-7)


def test_run_dij_refuses_a_phase_space(tmp_path):
    from pyradmc_amd.engine import HipEngine

    with PhaseSpaceSource(_write_recs(tmp_path, [_rec(1, 2.0)])) as src:
        with pytest.raises(TypeError, match="beamlet source"):
            HipEngine.run_dij(object(), src, 10, 1, seed=1)


# ----------------------------------------------- oracle given-primaries source


def _given(n, kind_code, energy, pos, direction, weight=1.0):
    ones = np.ones(n)
    return dict(kind="given", max_energy=energy, primaries=dict(
        particle_type=np.full(n, kind_code, np.int32), energy=energy * ones, x=pos[0] * ones,
        y=pos[1] * ones, z=pos[2] * ones, ux=direction[0] * ones, uy=direction[1] * ones,
        uz=direction[2] * ones, weight=weight * ones))


@pytest.mark.parametrize("particle,code", [("photon", 1), ("electron", 2)])
def test_oracle_given_primaries_equal_a_pencil_bit_for_bit(oracle_lib, particle, code):
    # A pencil draws nothing from a history's stream before transport, so primaries given
    # with the pencil's kind, energy, position and direction replay it exactly.
    from pyradmc_amd import VoxelGrid

    grid = VoxelGrid.uniform_water((8, 8, 12), (0.5, 0.5, 0.5))
    n, pos, d = 4000, (2.0, 2.0, -0.1), (0.0, 0.0, 1.0)
    pencil = dict(kind="pencil", particle=particle, energy=2.0, max_energy=2.0, position=list(pos),
                  direction=list(d))
    args = (grid.shape, grid.spacing, grid.origin, grid.density, grid.material,
            grid.max_density_by_material())
    a = oracle_lib.Oracle(*args, pencil).run(n, 4, seed=17, threads=2)
    b = oracle_lib.Oracle(*args, _given(n, code, 2.0, pos, d)).run(n, 4, seed=17, threads=2)
    np.testing.assert_array_equal(a["dose"], b["dose"])
    np.testing.assert_array_equal(a["dose_sigma"], b["dose_sigma"])
    assert a["books"] == b["books"]


def test_oracle_given_positrons_and_weights_close_the_ledger(oracle_lib):
    from pyradmc_amd import VoxelGrid

    grid = VoxelGrid.uniform_water((8, 8, 12), (0.5, 0.5, 0.5))
    n = 3000
    src = _given(n, 3, 1.5, (2.0, 2.0, 1.0), (0.0, 0.0, 1.0), weight=0.75)
    o = oracle_lib.Oracle(grid.shape, grid.spacing, grid.origin, grid.density, grid.material,
                          grid.max_density_by_material(), src)
    r = o.run(n, 3, seed=5, threads=2)
    bk = r["books"]
    assert bk["emitted"] == pytest.approx(n * 0.75 * (1.5 + 2 * ME), rel=1e-12)
    assert abs(bk["emitted"] - bk["deposited"] - bk["escaped"] - bk["unscored"]) < 1e-9 * bk["emitted"]
    # a positron started inside the phantom deposits most of its kinetic energy
    assert bk["deposited"] > 0.75 * n * 1.5 * 0.5
    with pytest.raises(ValueError, match="beyond the given primaries"):
        o.run(2 * n, 3, seed=5, threads=2)


# -------------------------------------------------- device-resident population


def test_population_blocks_are_the_transportable_records_in_order(tmp_path):
    from pyradmc_amd.phasespace import DevicePhaseSpace

    cols = _population(1000, seed=3)
    cols["particle_type"][[0, 5, 6, 999]] = 4  # unsupported, at both ends and in a run
    path = write_iaea_phase_space(tmp_path / "pop", cols)
    with pytest.warns(UserWarning):
        src = PhaseSpaceSource(path)
    with IAEAPhaseSpace(path) as ps:
        every = ps.columns()
    keep = np.isin(every["particle_type"], (1, 2, 3))
    for block in (7, 256, 4096):
        blocks = list(src.population_blocks(block))
        for k in ("particle_type", "energy", "x", "y", "z", "ux", "uy", "uz", "weight"):
            np.testing.assert_array_equal(np.concatenate([b[k] for b in blocks]), every[k][keep])
    dev = DevicePhaseSpace(src, block=300)
    assert len(dev) == keep.sum() and dev.max_energy == src.max_energy
    mem = dev.to_phase_space()
    np.testing.assert_array_equal(mem.columns()["x"], every["x"][keep])
    again = list(DevicePhaseSpace(mem, block=128).population_blocks())
    np.testing.assert_array_equal(np.concatenate([b["weight"] for b in again]), every["weight"][keep])
    assert DevicePhaseSpace(mem).token != dev.token
    with pytest.raises(TypeError, match="population_blocks"):
        DevicePhaseSpace(object())
