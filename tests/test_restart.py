"""Restart files (outputs_writer.jl:338-361 generate_filename, :1221-1273
write_restart!, :1275-1370 read_restart!, :1373-1402 find_restart_files),
host side: libgdhip's NetCDF classic writer and reader checked against
scipy.io.netcdf_file, an independent NetCDF classic implementation, in both
directions.  No GPU is used (gd_restart_*_arrays are host-only)."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

import gdhip

COMPONENTS = ["velocity_toroidal", "velocity_poloidal", "magnetic_toroidal", "magnetic_poloidal",
              "temperature_spectral"]


def _lm(lmax, mmax=None):
    mmax = lmax if mmax is None else mmax
    return [(l, m) for l in range(lmax + 1) for m in range(min(l, mmax) + 1)]


def _fields(seed, lmax, nr, mmax=None):
    rng = np.random.default_rng(seed)
    nlm = len(_lm(lmax, mmax))
    return rng.standard_normal((5, nr, nlm)) + 1j * rng.standard_normal((5, nr, nlm))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_scipy_reads_library_file(tmp_path, precision):
    L, nr = 11, 7
    st, h = _fields(1, L, nr), _fields(2, L, nr)
    r = np.linspace(0.54, 1.54, nr)
    tr = gdhip.TimeTracker(last_output_time=0.2, output_count=3, restart_count=1, grid_file_written=True)
    path = str(tmp_path / "a.nc")
    gdhip.write_restart_arrays(path, L, nr, r, st, h, step=7, time=0.25, tracker=tr, precision=precision)
    with open(path, "rb") as fh:
        assert fh.read(4) == b"CDF\x02"
    f = netcdf_file(path, "r", mmap=False)
    try:
        cast = np.float64 if precision == "fp64" else np.float32
        assert f.dimensions["spectral_mode"] == len(_lm(L)) and f.dimensions["r"] == nr
        assert f.dimensions["time"] == 1 and f.dimensions["scalar"] == 1 and f.dimensions["meta"] == 1
        lm = list(zip(f.variables["l_values"].data.tolist(), f.variables["m_values"].data.tolist()))
        assert lm == _lm(L)
        for k, c in enumerate(COMPONENTS):
            for part, arr in (("real", st[k].real), ("imag", st[k].imag)):
                v = f.variables["%s_%s" % (c, part)]
                assert v.dimensions == ("r", "spectral_mode")
                assert v.data.dtype == np.dtype(cast).newbyteorder(">")
                assert np.array_equal(v.data, arr.astype(cast))
                assert v.representation == b"spectral_space"
            for part, arr in (("real", h[k].real), ("imag", h[k].imag)):
                assert np.array_equal(f.variables["%s_prev_nonlinear_%s" % (c, part)].data, arr.astype(cast))
        assert f.variables["velocity_toroidal_real"].long_name == b"velocity_toroidal_real_coefficients"
        assert f.variables["step"].data.tolist() == [7]
        assert f.variables["time"].data.tolist() == [cast(0.25)]
        assert np.array_equal(f.variables["r"].data, r.astype(cast))
        assert f.variables["r"].long_name == b"radial_coordinate"
        assert b"".join(f.variables["geometry"].data.ravel().tolist()) == b"shell"
        assert f.variables["output_count"].data.tolist() == [3]
        assert f.variables["restart_count"].data.tolist() == [1]
        assert f.variables["grid_file_written"].data.tolist() == [1]
        assert f.variables["last_output_time"].data.tolist() == [cast(0.2)]
        assert f.title == b"GeoDynamo Simulation Output" and f.Conventions == b"CF-1.8"
        assert int(f.current_step) == 7 and float(f.current_time) == 0.25 and f.geometry == b"shell"
        assert int(f.mpi_rank) == 0 and int(f.mpi_total_ranks) == 1
    finally:
        f.close()


def _scipy_file(path, version, lmax, nr, st, order, dtype, step=3, time=0.5, history=None, record=False):
    lm = [_lm(lmax)[k] for k in order]
    f = netcdf_file(path, "w", version=version)
    f.createDimension("time", None if record else 1)
    f.createDimension("r", nr)
    f.createDimension("spectral_mode", len(lm))
    f.createDimension("scalar", 1)
    v = f.createVariable("time", "d", ("time",))
    v[:] = [time]
    v = f.createVariable("step", "i", ("time",))
    v[:] = [step]
    f.createVariable("l_values", "i", ("spectral_mode",))[:] = [a for a, _ in lm]
    f.createVariable("m_values", "i", ("spectral_mode",))[:] = [b for _, b in lm]
    f.createVariable("r", "d", ("r",))[:] = np.arange(nr, dtype=float)
    for k, c in enumerate(COMPONENTS):
        for part, arr in (("real", st[k].real), ("imag", st[k].imag)):
            f.createVariable("%s_%s" % (c, part), dtype, ("r", "spectral_mode"))[:] = arr[:, order]
        if history is not None:
            for part, arr in (("real", history[k].real), ("imag", history[k].imag)):
                f.createVariable("%s_prev_nonlinear_%s" % (c, part), dtype, ("r", "spectral_mode"))[:] = arr[:, order]
    f.createVariable("output_count", "i", ("scalar",))[:] = [4]
    f.createVariable("restart_count", "i", ("scalar",))[:] = [2]
    f.createVariable("last_output_time", "d", ("scalar",))[:] = [0.4]
    f.geometry = b"shell"
    f.mpi_rank = np.int32(0)
    f.close()


@pytest.mark.parametrize("version", [1, 2])
def test_library_reads_scipy_file_in_any_mode_order(tmp_path, version):
    L, nr = 9, 6
    st = _fields(3, L, nr)
    nlm = st.shape[2]
    # m-major order of the modes: the reader maps them by l_values / m_values
    order = sorted(range(nlm), key=lambda k: (_lm(L)[k][1], _lm(L)[k][0]))
    path = str(tmp_path / "b.nc")
    _scipy_file(path, version, L, nr, st, order, "d")
    s2, h2, step, r, md = gdhip.read_restart_arrays(path, L, nr)
    assert np.array_equal(s2, st)
    assert h2 is None  # the reference's restart stores no history
    assert step == 3 and md["current_time"] == 0.5 and md["output_count"] == 4 and md["restart_count"] == 2
    assert md["last_output_time"] == 0.4 and md["geometry"] == "shell"
    assert md["grid_file_written"]  # absent flag counts as written (outputs_writer.jl:1306-1312)
    assert np.array_equal(r, np.arange(nr, dtype=float))


def test_library_reads_fp32_history_from_scipy(tmp_path):
    L, nr = 5, 4
    st, h = _fields(4, L, nr), _fields(5, L, nr)
    path = str(tmp_path / "c.nc")
    _scipy_file(path, 2, L, nr, st, list(range(st.shape[2])), "f", history=h, step=12)
    s2, h2, step, _, _ = gdhip.read_restart_arrays(path, L, nr)
    f32 = lambda z: z.real.astype(np.float32).astype(float) + 1j * z.imag.astype(np.float32).astype(float)  # noqa: E731
    assert step == 12
    assert np.array_equal(s2, f32(st)) and np.array_equal(h2, f32(h))


def test_round_trip_and_history_rules(tmp_path):
    L, nr = 15, 9
    st, h = _fields(6, L, nr), _fields(7, L, nr)
    r = np.linspace(0.5, 1.5, nr)
    p1, p0 = str(tmp_path / "h.nc"), str(tmp_path / "n.nc")
    gdhip.write_restart_arrays(p1, L, nr, r, st, h, step=5)
    s2, h2, step, r2, _ = gdhip.read_restart_arrays(p1, L, nr)
    assert step == 5 and np.array_equal(s2, st) and np.array_equal(h2, h) and np.array_equal(r2, r)
    gdhip.write_restart_arrays(p0, L, nr, r, st, None, step=0)
    s3, h3, step3, _, _ = gdhip.read_restart_arrays(p0, L, nr)
    assert step3 == 0 and h3 is None and np.array_equal(s3, st)
    with pytest.raises(gdhip.GDError):  # a history needs a step > 0
        gdhip.write_restart_arrays(str(tmp_path / "x.nc"), L, nr, r, st, h, step=0)
    # mmax < lmax: nlm follows the truncation
    stm = _fields(8, L, nr, mmax=4)
    gdhip.write_restart_arrays(str(tmp_path / "m.nc"), L, nr, r, stm, None, mmax=4)
    s4, _, _, _, _ = gdhip.read_restart_arrays(str(tmp_path / "m.nc"), L, nr, mmax=4)
    assert np.array_equal(s4, stm)


def test_reader_rejects_bad_files(tmp_path):
    L, nr = 7, 5
    st = _fields(9, L, nr)
    r = np.linspace(0.5, 1.5, nr)
    good = str(tmp_path / "g.nc")
    gdhip.write_restart_arrays(good, L, nr, r, st)
    with pytest.raises(gdhip.GDError, match="spectral_mode"):
        gdhip.read_restart_arrays(good, L + 1, nr)
    with pytest.raises(gdhip.GDError, match="spectral_mode"):
        gdhip.read_restart_arrays(good, L, nr + 1)
    junk = tmp_path / "j.nc"
    junk.write_bytes(b"HDF\x89 not a classic file")
    with pytest.raises(gdhip.GDError, match="not a NetCDF classic file"):
        gdhip.read_restart_arrays(str(junk), L, nr)
    trunc = tmp_path / "t.nc"
    trunc.write_bytes(open(good, "rb").read()[:200])
    with pytest.raises(gdhip.GDError):
        gdhip.read_restart_arrays(str(trunc), L, nr)
    rec = str(tmp_path / "rec.nc")
    _scipy_file(rec, 2, L, nr, st, list(range(st.shape[2])), "d", record=True)
    with pytest.raises(gdhip.GDError, match="record variable"):
        gdhip.read_restart_arrays(rec, L, nr)
    miss = str(tmp_path / "miss.nc")
    f = netcdf_file(miss, "w", version=2)
    f.createDimension("r", nr)
    f.createDimension("spectral_mode", st.shape[2])
    f.createVariable("velocity_toroidal_real", "d", ("r", "spectral_mode"))[:] = st[0].real
    f.close()
    with pytest.raises(gdhip.GDError, match="velocity_toroidal_imag missing"):
        gdhip.read_restart_arrays(miss, L, nr)
    with pytest.raises(gdhip.GDError, match="cannot open"):
        gdhip.read_restart_arrays(str(tmp_path / "absent.nc"), L, nr)
    # a header whose dimension lengths multiply past 2^64: rejected, not read
    big = str(tmp_path / "big.nc")
    f = netcdf_file(big, "w", version=2)
    for d in ("xa", "xb", "xc"):
        f.createDimension(d, 2)
    f.createVariable("velocity_toroidal_real", "d", ("xa", "xb", "xc"))[:] = 1.0
    f.close()
    raw = bytearray(open(big, "rb").read())
    for d in (b"xa", b"xb", b"xc"):
        k = raw.index(d + b"\x00\x00") + 4
        raw[k:k + 4] = b"\xff\xff\xff\xff"
    open(big, "wb").write(bytes(raw))
    with pytest.raises(gdhip.GDError, match="larger than the file|outside the file"):
        gdhip.read_restart_arrays(big, L, nr)


def test_generate_filename_and_find_restart_files(tmp_path):
    cfg = gdhip.OutputConfig(output_dir=str(tmp_path), filename_prefix="run")
    name = gdhip.generate_filename(cfg, 1.5, 10, 3, "restart")
    assert name == os.path.join(str(tmp_path), "run_restart_geom_shell_rank_0003_time_1p500000.nc")
    cfg2 = gdhip.OutputConfig(output_dir=str(tmp_path), filename_prefix="run", naming_scheme=gdhip.TIME_RANK)
    assert os.path.basename(gdhip.generate_filename(cfg2, 0.0001234, 0, 12, "output", "ball")) == \
        "run_output_geom_ball_time_0p000123_rank_0012.nc"
    for t in (0.5, 1.0, 2.0):
        open(gdhip.generate_filename(cfg, t, 0, 3, "restart"), "wb").close()
    open(gdhip.generate_filename(cfg, 1.1, 0, 4, "restart"), "wb").close()  # another rank
    open(gdhip.generate_filename(cfg, 1.1, 0, 3, "output"), "wb").close()   # not a restart
    got = [os.path.basename(p) for p in gdhip.find_restart_files(str(tmp_path), 1.2, 3)]
    assert got == ["run_restart_geom_shell_rank_0003_time_1p000000.nc",
                   "run_restart_geom_shell_rank_0003_time_0p500000.nc",
                   "run_restart_geom_shell_rank_0003_time_2p000000.nc"]


def test_composition_round_trip_and_scipy(tmp_path):
    """composition_spectral (outputs_writer.jl:507-522, read back at :1327-1339)
    and its AB2 history."""
    L, nr = 9, 7
    st, h = _fields(10, L, nr), _fields(11, L, nr)
    c, ch = _fields(12, L, nr)[0], _fields(13, L, nr)[0]
    r = np.linspace(0.5, 1.5, nr)
    path = str(tmp_path / "c.nc")
    gdhip.write_restart_arrays(path, L, nr, r, st, h, step=3, composition=c, composition_history=ch)
    f = netcdf_file(path, "r", mmap=False)
    try:
        assert f.variables["composition_spectral_real"].dimensions == ("r", "spectral_mode")
        assert np.array_equal(f.variables["composition_spectral_imag"].data, c.imag)
        assert np.array_equal(f.variables["composition_spectral_prev_nonlinear_real"].data, ch.real)
    finally:
        f.close()
    s2, h2, step, _, md = gdhip.read_restart_arrays(path, L, nr)
    assert step == 3 and np.array_equal(s2, st) and np.array_equal(h2, h)
    assert np.array_equal(md["composition"], c) and np.array_equal(md["composition_history"], ch)
    # without composition: reported absent
    p0 = str(tmp_path / "n.nc")
    gdhip.write_restart_arrays(p0, L, nr, r, st, h, step=3)
    _, _, _, _, md0 = gdhip.read_restart_arrays(p0, L, nr)
    assert md0["composition"] is None and md0["composition_history"] is None
