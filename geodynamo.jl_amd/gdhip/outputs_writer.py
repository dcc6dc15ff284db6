"""Restart files of the step (outputs_writer.jl: generate_filename :338-361,
write_restart! :1221-1273, read_restart! :1275-1370, find_restart_files
:1373-1402).

The files are NetCDF classic (CDF-2) written and parsed by libgdhip's C code
(`gd_restart_*`), with the reference's dimensions, variable names and
attributes, plus the AB2 history as `<component>_prev_nonlinear_{real,imag}`
so that a continued run is exactly the uninterrupted one.  This module holds
the host-side bookkeeping the reference keeps in Julia: the output
configuration, the time tracker and the choice of file.
"""
import ctypes
import dataclasses
import os
import re

import numpy as np

from ._lib import GD_PREC_FP32, GD_PREC_FP64, RestartComp, RestartMeta, check, lib

RANK_TIME, TIME_RANK = 0, 1  # FileNaming (outputs_writer.jl:25-28)


@dataclasses.dataclass
class OutputConfig:
    """The fields of OutputConfig (outputs_writer.jl:30-60) a restart uses."""
    output_dir: str = "."
    filename_prefix: str = "geodynamo"
    naming_scheme: int = RANK_TIME
    output_precision: str = "fp64"  # Float64 / Float32
    output_interval: float = 0.1
    restart_interval: float = 1.0


@dataclasses.dataclass
class TimeTracker:
    """The TimeTracker fields write_restart!/read_restart! touch."""
    last_output_time: float = 0.0
    last_restart_time: float = 0.0
    next_output_time: float = 0.0
    next_restart_time: float = 0.0
    output_count: int = 0
    restart_count: int = 0
    grid_file_written: bool = False


def _precision(p):
    if p in ("fp64", "float64", np.float64):
        return GD_PREC_FP64
    if p in ("fp32", "float32", np.float32):
        return GD_PREC_FP32
    raise ValueError("output_precision must be fp64 or fp32, not %r" % (p,))


def _meta(time, tracker, geometry, rank, nranks):
    m = RestartMeta()
    m.time = float(time)
    m.last_output_time = float(tracker.last_output_time)
    m.output_count = int(tracker.output_count)
    m.restart_count = int(tracker.restart_count)
    m.grid_file_written = int(bool(tracker.grid_file_written))
    m.mpi_rank = int(rank)
    m.mpi_total_ranks = int(nranks)
    m.geometry = geometry.encode()[:31]
    return m


def _meta_dict(m):
    return {"current_time": m.time, "last_output_time": m.last_output_time, "output_count": m.output_count,
            "restart_count": m.restart_count, "grid_file_written": bool(m.grid_file_written),
            "mpi_rank": m.mpi_rank, "mpi_total_ranks": m.mpi_total_ranks,
            "geometry": m.geometry.decode()}


def generate_filename(config, time, step, rank, file_type="output", geometry="shell"):
    """outputs_writer.jl:338-361 (the step number is not part of the name there either)."""
    buf = ctypes.create_string_buffer(4096)
    check(lib().gd_restart_filename(config.output_dir.encode(), config.filename_prefix.encode(),
                                    file_type.encode(), geometry.encode(), float(time), int(rank),
                                    int(config.naming_scheme == TIME_RANK), buf, len(buf)))
    return buf.value.decode()


def write_restart(step, tracker, metadata, config, rank=0, nranks=1):
    """write_restart! (outputs_writer.jl:1221-1273) for a CNAB2Step: metadata
    needs "current_time" (and optionally "geometry"); returns the file name."""
    time = float(metadata["current_time"])
    geometry = metadata.get("geometry", "shell")
    path = generate_filename(config, time, step.step_count, rank, "restart", geometry)
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    m = _meta(time, tracker, geometry, rank, nranks)
    check(lib().gd_restart_write(step._h, path.encode(), ctypes.byref(m), _precision(config.output_precision)))
    return path


def find_restart_files(restart_dir, target_time, rank):
    """outputs_writer.jl:1373-1402: this rank's restart files, nearest time first."""
    tag = "rank_%04d" % rank
    out = []
    for f in os.listdir(restart_dir):
        if not (f.endswith(".nc") and "restart" in f and tag in f):
            continue
        m = re.search(r"time_(\d+p\d+)", f)
        if m is None:
            continue
        try:
            t = float(m.group(1).replace("p", "."))
        except ValueError:
            continue
        out.append((os.path.join(restart_dir, f), t))
    out.sort(key=lambda x: abs(x[1] - target_time))
    return [p for p, _ in out]


def read_restart(step, tracker, restart_dir, restart_time, config, rank=0):
    """read_restart! (outputs_writer.jl:1275-1370) into a CNAB2Step: restores
    state, AB2 history and step counter, updates the tracker as the reference
    does, and returns the metadata dict (current_time, current_step, ...)."""
    files = find_restart_files(restart_dir, restart_time, rank)
    if not files:
        raise FileNotFoundError("Rank %d: No restart files found near time %s" % (rank, restart_time))
    m = RestartMeta()
    check(lib().gd_restart_read(step._h, files[0].encode(), ctypes.byref(m)))
    tracker.last_output_time = m.last_output_time
    tracker.output_count = m.output_count
    tracker.restart_count = m.restart_count
    tracker.grid_file_written = bool(m.grid_file_written)
    tracker.next_output_time = tracker.last_output_time + config.output_interval
    tracker.next_restart_time = tracker.last_restart_time + config.restart_interval
    md = _meta_dict(m)
    md["current_step"] = step.step_count
    md["file"] = files[0]
    return md


def write_restart_arrays(path, lmax, nr, r, state, history=None, step=0, time=0.0, tracker=None,
                         precision="fp64", mmax=None, geometry="shell", composition=None,
                         composition_history=None):
    """Host-only writer on l-major complex arrays (5, nr, nlm), the composition
    (nr, nlm) optional; no GPU used."""
    mmax = lmax if mmax is None else mmax
    tracker = tracker or TimeTracker()
    m = _meta(time, tracker, geometry, 0, 1)
    re_, im_ = np.ascontiguousarray(state.real, dtype=np.float64), np.ascontiguousarray(state.imag, dtype=np.float64)
    rr = np.ascontiguousarray(r, dtype=np.float64)
    if history is not None:
        hr = np.ascontiguousarray(history.real, dtype=np.float64)
        hi = np.ascontiguousarray(history.imag, dtype=np.float64)
        hp = (hr.ctypes.data, hi.ctypes.data)
    else:
        hp = (None, None)
    comp, keep = None, []
    if composition is not None:
        comp = RestartComp()
        parts = [composition.real, composition.imag]
        if composition_history is not None:
            parts += [composition_history.real, composition_history.imag]
        keep = [np.ascontiguousarray(x, dtype=np.float64) for x in parts]
        comp.re, comp.im = keep[0].ctypes.data, keep[1].ctypes.data
        if len(keep) == 4:
            comp.nl_re, comp.nl_im = keep[2].ctypes.data, keep[3].ctypes.data
    check(lib().gd_restart_write_arrays(path.encode(), ctypes.byref(m), lmax, mmax, nr, rr.ctypes.data, int(step),
                                        re_.ctypes.data, im_.ctypes.data, hp[0], hp[1], _precision(precision),
                                        ctypes.byref(comp) if comp is not None else None))


def read_restart_arrays(path, lmax, nr, mmax=None):
    """Host-only reader: (state, history or None, step, r, metadata dict); the
    dict holds "composition" and "composition_history" (None when absent)."""
    mmax = lmax if mmax is None else mmax
    nlm = sum(min(l, mmax) + 1 for l in range(lmax + 1))
    shape = (5, nr, nlm)
    re_, im_, hr, hi = (np.zeros(shape) for _ in range(4))
    r = np.zeros(nr)
    step, hist = ctypes.c_int(), ctypes.c_int()
    m = RestartMeta()
    cplanes = [np.zeros((nr, nlm)) for _ in range(4)]
    comp = RestartComp()
    comp.re, comp.im, comp.nl_re, comp.nl_im = (x.ctypes.data for x in cplanes)
    check(lib().gd_restart_read_arrays(path.encode(), lmax, mmax, nr, r.ctypes.data, ctypes.byref(step),
                                       re_.ctypes.data, im_.ctypes.data, hr.ctypes.data, hi.ctypes.data,
                                       ctypes.byref(hist), ctypes.byref(m), ctypes.byref(comp)))
    history = hr + 1j * hi if hist.value else None
    md = _meta_dict(m)
    md["composition"] = cplanes[0] + 1j * cplanes[1] if comp.present else None
    md["composition_history"] = cplanes[2] + 1j * cplanes[3] if comp.has_history else None
    return re_ + 1j * im_, history, step.value, r, md
