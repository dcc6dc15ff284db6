"""Restart through the C ABI (write_restart! / read_restart!,
outputs_writer.jl:1221-1370, with the AB2 history beside the fields): a run
continued from a restart file takes exactly the uninterrupted run's steps,
bit for bit, on both arithmetic paths, eager and graph-replayed, into a fresh
step object and into one whose step graphs were already captured."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASE = dict(lmax=31, nlat=48, nlon=96, nr=17)


def _step(prec):
    import gdhip

    return gdhip.CNAB2Step(CASE["lmax"], CASE["nlat"], CASE["nlon"], CASE["nr"], Ra=200.0, dt=1e-4, precision=prec)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("graph", [True, False])
def test_continued_run_equals_uninterrupted(tmp_path, prec, graph):
    import gdhip

    print("case: restart %s graph=%s lmax=%d nr=%d" % (prec, graph, CASE["lmax"], CASE["nr"]), flush=True)
    st0 = gdhip.seeded_state(11, CASE["lmax"], CASE["nr"], amp=1e-3)
    a = _step(prec)
    a.set_state_complex(st0)
    a.run(7, graph=graph)
    ref = a.get_state_complex()

    b = _step(prec)
    b.set_state_complex(st0)
    b.run(4, graph=graph)
    cfg = gdhip.OutputConfig(output_dir=str(tmp_path), filename_prefix="gd")
    tr = gdhip.TimeTracker(last_output_time=0.0003, output_count=2)
    path = gdhip.write_restart(b, tr, {"current_time": 4e-4}, cfg)

    c = _step(prec)
    tr2 = gdhip.TimeTracker()
    md = gdhip.read_restart(c, tr2, str(tmp_path), 4e-4, cfg)
    assert md["file"] == path and md["current_step"] == 4 and c.step_count == 4
    assert tr2.output_count == 2 and tr2.last_output_time == 0.0003
    assert np.array_equal(c.get_state_complex(), b.get_state_complex())
    assert np.array_equal(c.get_history(), b.get_history())
    c.run(3, graph=graph)
    assert c.step_count == 7
    assert np.array_equal(c.get_state_complex(), ref)

    # b has captured its step graphs; it continues, then returns to step 4
    # from the file and must retrace the same steps
    b.run(2, graph=graph)
    gdhip.read_restart(b, gdhip.TimeTracker(), str(tmp_path), 4e-4, cfg)
    assert b.step_count == 4
    b.run(3, graph=graph)
    assert np.array_equal(b.get_state_complex(), ref)


def test_history_is_the_previous_steps_nonlinear_terms():
    import gdhip

    print("case: history = nonlinear of the previous state", flush=True)
    s = _step("fp64")
    s.set_state_complex(gdhip.seeded_state(12, CASE["lmax"], CASE["nr"], amp=1e-3))
    with pytest.raises(gdhip.GDError):
        s.get_history()  # no history before the first step
    s.run(2)
    n2 = s.nonlinear()  # N(state at step 2), which step 2 computes again
    s.run(1)
    assert np.array_equal(s.get_history(), n2)


def test_restart_without_history_bootstraps(tmp_path):
    """A file that stores no history (as the reference's restart) restores at
    step 0, so the continued run starts with an AB1 step like a fresh one."""
    import gdhip

    print("case: restart without history", flush=True)
    st0 = gdhip.seeded_state(13, CASE["lmax"], CASE["nr"], amp=1e-3)
    a = _step("fp64")
    a.set_state_complex(st0)
    a.run(3)
    mid = a.get_state_complex()
    a.run(2)
    path = str(tmp_path / "nohist.nc")
    gdhip.write_restart_arrays(path, CASE["lmax"], CASE["nr"], np.linspace(0.5, 1.5, CASE["nr"]), mid, None, step=0)
    c = _step("fp64")
    c.set_state_complex(gdhip.seeded_state(14, CASE["lmax"], CASE["nr"], amp=1e-3))
    import ctypes

    from gdhip._lib import RestartMeta, check, lib

    check(lib().gd_restart_read(c._h, path.encode(), ctypes.byref(RestartMeta())))
    assert c.step_count == 0
    b = _step("fp64")
    b.set_state_complex(mid)
    b.run(2)
    c.run(2)
    assert np.array_equal(c.get_state_complex(), b.get_state_complex())


@pytest.mark.parametrize("prec,out_prec", [("fp64", "fp64"), ("fp32", "fp32")])
def test_gpu_written_file_read_by_scipy(tmp_path, prec, out_prec):
    """A file written from a GPU step, opened by scipy's independent NetCDF
    classic reader: the fields (outputs_writer.jl:1221-1273) and the AB2
    history equal the step's own state and history in l-major order, and the
    l_values / m_values / r / step variables describe the step."""
    import gdhip
    from scipy.io import netcdf_file

    print("case: GPU restart file read by scipy %s -> %s" % (prec, out_prec), flush=True)
    st0 = gdhip.seeded_state(12, CASE["lmax"], CASE["nr"], amp=1e-3)
    a = _step(prec)
    a.set_state_complex(st0)
    a.run(3, graph=True)
    st, hist = a.get_state_complex(), a.get_history()
    cfg = gdhip.OutputConfig(output_dir=str(tmp_path), filename_prefix="gd", output_precision=out_prec)
    path = gdhip.write_restart(a, gdhip.TimeTracker(), {"current_time": 3e-4}, cfg)
    names = ("velocity_toroidal", "velocity_poloidal", "magnetic_toroidal", "magnetic_poloidal",
             "temperature_spectral")
    f = netcdf_file(path, "r", mmap=False)
    try:
        cast = np.float64 if out_prec == "fp64" else np.float32
        for k, nm in enumerate(names):
            assert f.variables[nm + "_real"].dimensions == ("r", "spectral_mode")
            for part, ref in (("_real", st[k].real), ("_imag", st[k].imag)):
                got = f.variables[nm + part].data
                assert np.array_equal(got, ref.astype(cast)), nm + part
            for part, ref in (("_real", hist[k].real), ("_imag", hist[k].imag)):
                got = f.variables[nm + "_prev_nonlinear" + part].data
                assert np.array_equal(got, ref.astype(cast)), nm + "_prev_nonlinear" + part
        lv = np.array([l for l in range(CASE["lmax"] + 1) for m in range(l + 1)])
        mv = np.array([m for l in range(CASE["lmax"] + 1) for m in range(l + 1)])
        assert np.array_equal(f.variables["l_values"].data, lv)
        assert np.array_equal(f.variables["m_values"].data, mv)
        # r is stored in the file's precision
        assert np.array_equal(f.variables["r"].data, a.radii().astype(cast))
        assert int(f.variables["step"].data.ravel()[0]) == 3
    finally:
        f.close()
