"""The device-resident CNAB2 step (the step of scripts/bench_timestepping.jl:6-18).

Field order: u_tor, u_pol, B_tor, B_pol, T.  State arrays are (5, nr, nlm)
float64 re/im in the reference's l-major spectral layout.
"""
import ctypes

import numpy as np

from ._arrays import context, prepare
from ._lib import (GD_FD_LITERAL, GD_FD_TRANSPOSE, StepDesc, check, lib)
from .shtnskit_transforms import PRECISIONS

FIELDS = ("u_tor", "u_pol", "B_tor", "B_pol", "T")
BATCH_MODES = {"streams": 0, "fused": 1}


class CNAB2Step:
    def __init__(self, lmax, nlat, nlon, nr, E=1e-3, Pm=5.0, Pr=1.0, Ra=100.0, rratio=0.35, dt=1e-5,
                 theta=0.5, bw=3, mmax=None, precision="fp64", literal_fd=False, radials=False,
                 coriolis_index=0, bc_rows=1, full_products=False, radial_force=False, device=0):
        """radial_force: the buoyancy (velocity.jl:534-537) enters the poloidal
        equation through u_r = l(l+1) P / r^2 (magnetic.jl:257-259); off by
        default, the reference's data flow, where it reaches no equation."""
        mmax = lmax if mmax is None else mmax
        self.desc = StepDesc(E, Pm, Pr, Ra, rratio, dt, theta, lmax, mmax, nlat, nlon, nr, bw,
                             PRECISIONS[precision], GD_FD_LITERAL if literal_fd else GD_FD_TRANSPOSE,
                             int(radials), int(coriolis_index), int(bc_rows), int(full_products),
                             int(radial_force))
        self.lmax, self.mmax, self.nlat, self.nlon, self.nr = lmax, mmax, nlat, nlon, nr
        self.precision = precision
        self.nlm = sum(min(l, mmax) + 1 for l in range(lmax + 1))
        self.device = device
        self._h = ctypes.c_void_p()
        check(lib().gd_step_create(context(device), ctypes.byref(self.desc), ctypes.byref(self._h)))

    @property
    def shape(self):
        return (5, self.nr, self.nlm)

    def set_state(self, re, im):
        ctx, flags, (a, b) = prepare(re, im, device=self.device, shapes=[self.shape] * 2)
        check(lib().gd_step_set_state(self._h, a, b, flags))

    def set_state_complex(self, st):
        self.set_state(np.ascontiguousarray(st.real), np.ascontiguousarray(st.imag))

    def get_state(self, re=None, im=None):
        if re is None:
            re, im = np.zeros(self.shape), np.zeros(self.shape)
        ctx, flags, (a, b) = prepare(re, im, device=self.device, shapes=[self.shape] * 2)
        check(lib().gd_step_get_state(self._h, a, b, flags))
        return re, im

    def get_state_complex(self):
        re, im = self.get_state()
        return re + 1j * im

    def get_history(self):
        """AB2 history: the previous step's nonlinear terms (the reference's
        prev_nonlinear, simulation.jl:1312-1329), complex (5, nr, nlm)."""
        re, im = np.zeros(self.shape), np.zeros(self.shape)
        ctx, flags, (a, b) = prepare(re, im, device=self.device)
        check(lib().gd_step_get_history(self._h, a, b, flags))
        return re + 1j * im

    def restore(self, state, history=None, step=0):
        """Sets state, AB2 history and step counter together (step 0: no history,
        the next step is the AB1 bootstrap)."""
        re, im = np.ascontiguousarray(state.real), np.ascontiguousarray(state.imag)
        if step > 0:
            if history is None:
                raise ValueError("step > 0 needs the nonlinear history")
            hre, him = np.ascontiguousarray(history.real), np.ascontiguousarray(history.imag)
            ctx, flags, (a, b, c, d) = prepare(re, im, hre, him, device=self.device, shapes=[self.shape] * 4)
        else:
            ctx, flags, (a, b) = prepare(re, im, device=self.device, shapes=[self.shape] * 2)
            c = d = None
        check(lib().gd_step_restore(self._h, a, b, c, d, int(step), flags))

    def set_temperature_bc(self, bc_type_inner=None, bc_type_outer=None, boundary_values=None):
        """Per-mode temperature conditions (thermal.jl:262-313): types 1
        (Dirichlet) / 2 (flux, tau method) per l-major mode and
        boundary_values (2, nlm); no arguments: back to the default (rows 0)."""
        if bc_type_inner is None:
            check(lib().gd_step_set_temperature_bc(self._h, None, None, None))
            return
        ti = np.ascontiguousarray(bc_type_inner, dtype=np.int32)
        to = np.ascontiguousarray(bc_type_outer, dtype=np.int32)
        v = np.ascontiguousarray(np.asarray(boundary_values, dtype=np.float64).T)
        if ti.shape != (self.nlm,) or to.shape != (self.nlm,) or v.shape != (self.nlm, 2):
            raise ValueError("bc types need nlm entries and boundary_values shape (2, nlm)")
        check(lib().gd_step_set_temperature_bc(self._h, ti.ctypes.data, to.ctypes.data, v.ctypes.data))

    def enable_composition(self, Sc=1.0):
        """Composition as a passive scalar of the step (compositional.jl), Pm/Sc
        diffusivity, default no-flux conditions; tangential step only."""
        check(lib().gd_step_enable_composition(self._h, float(Sc)))

    def set_composition(self, C):
        """C: complex (nr, nlm), l-major; call after set_state."""
        re, im = np.ascontiguousarray(C.real), np.ascontiguousarray(C.imag)
        ctx, flags, (a, b) = prepare(re, im, device=self.device, shapes=[(self.nr, self.nlm)] * 2)
        check(lib().gd_step_set_composition(self._h, a, b, flags))

    def get_composition(self):
        re, im = np.zeros((self.nr, self.nlm)), np.zeros((self.nr, self.nlm))
        ctx, flags, (a, b) = prepare(re, im, device=self.device)
        check(lib().gd_step_get_composition(self._h, a, b, flags))
        return re + 1j * im

    def get_composition_history(self):
        re, im = np.zeros((self.nr, self.nlm)), np.zeros((self.nr, self.nlm))
        ctx, flags, (a, b) = prepare(re, im, device=self.device)
        check(lib().gd_step_get_composition_history(self._h, a, b, flags))
        return re + 1j * im

    def restore_composition(self, C, history=None):
        """After restore(): the composition at the restored step and its history."""
        re, im = np.ascontiguousarray(C.real), np.ascontiguousarray(C.imag)
        if history is not None:
            hre, him = np.ascontiguousarray(history.real), np.ascontiguousarray(history.imag)
            ctx, flags, (a, b, c, d) = prepare(re, im, hre, him, device=self.device,
                                               shapes=[(self.nr, self.nlm)] * 4)
        else:
            ctx, flags, (a, b) = prepare(re, im, device=self.device, shapes=[(self.nr, self.nlm)] * 2)
            c = d = None
        check(lib().gd_step_restore_composition(self._h, a, b, c, d, flags))

    def set_composition_bc(self, bc_type_inner, bc_type_outer, boundary_values):
        ti = np.ascontiguousarray(bc_type_inner, dtype=np.int32)
        to = np.ascontiguousarray(bc_type_outer, dtype=np.int32)
        v = np.ascontiguousarray(np.asarray(boundary_values, dtype=np.float64).T)
        if ti.shape != (self.nlm,) or to.shape != (self.nlm,) or v.shape != (self.nlm, 2):
            raise ValueError("bc types need nlm entries and boundary_values shape (2, nlm)")
        check(lib().gd_step_set_composition_bc(self._h, ti.ctypes.data, to.ctypes.data, v.ctypes.data))

    def radii(self):
        r = np.zeros(self.nr)
        check(lib().gd_step_radii(self._h, r.ctypes.data))
        return r

    def nonlinear(self):
        re, im = np.zeros(self.shape), np.zeros(self.shape)
        ctx, flags, (a, b) = prepare(re, im, device=self.device)
        check(lib().gd_step_nonlinear(self._h, a, b, flags))
        return re + 1j * im

    def run(self, nsteps=1, graph=False):
        check((lib().gd_step_run_graph if graph else lib().gd_step_run)(self._h, nsteps))

    def synchronize(self):
        check(lib().gd_ctx_synchronize(context(self.device)))

    @property
    def step_count(self):
        return lib().gd_step_count(self._h)

    def energies(self):
        """(KE, ME, thermal) in FP64 (magnetic.jl:504-543, thermal.jl:413-438)."""
        out = np.zeros(3)
        check(lib().gd_step_energies(self._h, out.ctypes.data))
        return out

    KERNEL_CLASSES = ("spectral_pre", "synth_vector", "synth_scalar", "ring_fft_products",
                      "analysis_vector", "analysis_scalar", "cnab2_solve")

    def profile(self, nsteps=5):
        """Average device ms per step of each kernel class (HIP events, eager steps)."""
        out = np.zeros(len(self.KERNEL_CLASSES))
        check(lib().gd_step_profile(self._h, nsteps, out.ctypes.data))
        return dict(zip(self.KERNEL_CLASSES, out.tolist()))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gd_step_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StepBatch:
    """Independent shells advanced together (gd_batch_*): one graph per step
    parity holds every shell's step.  The shells must be at the same step
    count; results equal each shell's own run(graph=True) bit for bit.

    mode "streams": the shells' own steps spread over `nstreams` streams
    (FP64 shells only).
    mode "fused" (default): one launch per kernel form and stage over every shell of
    that form (synthesis, ring pass, both analyses, CNAB2 solves), on one
    stream, or split over `lanes` streams.  tiles (fused mode): "table"
    takes each shell's Legendre-synthesis tile from the committed tuning table
    (tile 0 where a shape is absent), "tune" runs the autotuner on this
    device, or a list gives them.  lanes: the fused shells' stream count
    (tune_lanes measures it; it advances the shells by real steps)."""

    def __init__(self, steps, nstreams=8, mode="fused", tiles="table", lanes=1):
        if not steps:
            raise ValueError("empty batch")
        if mode not in BATCH_MODES:
            raise ValueError("mode: one of %s" % sorted(BATCH_MODES))
        self.steps = list(steps)
        arr = (ctypes.c_void_p * len(self.steps))(*[s._h.value for s in self.steps])
        self._arr = arr
        self._h = ctypes.c_void_p()
        check(lib().gd_batch_create(context(self.steps[0].device), len(self.steps), arr, int(nstreams),
                                    ctypes.byref(self._h)))
        self.mode = mode
        check(lib().gd_batch_set_mode(self._h, BATCH_MODES[mode]))
        self.tune_ms = None
        if mode == "fused":
            if tiles == "tune":
                self.tune()
            elif tiles == "table":
                from .tuning import synth_tiles

                self.set_tiles(synth_tiles(self.steps))
            else:
                self.set_tiles(tiles)
        self.lanes_ms = None
        if mode == "fused":
            self.set_lanes(lanes)

    def run(self, nsteps=1):
        check(lib().gd_batch_run(self._h, nsteps))

    def set_lanes(self, lanes):
        check(lib().gd_batch_set_lanes(self._h, int(lanes)))

    def lanes(self):
        v = ctypes.c_int()
        check(lib().gd_batch_get_lanes(self._h, ctypes.byref(v)))
        return v.value

    def tune_lanes(self, nsteps=4):
        """Runs the lane autotuner (gd_batch_tune_lanes): advances every shell
        by 2 nsteps + 2 real steps per candidate lane count (1, 2, 4, 8 up to
        the stream and shell counts); returns the chosen count.
        self.lanes_ms: ms per step of lanes 1, 2, 4, 8 (-1: not run)."""
        ms = np.zeros(4)
        check(lib().gd_batch_tune_lanes(self._h, int(nsteps), ms.ctypes.data))
        self.lanes_ms = [float(v) for v in ms]
        return self.lanes()

    def set_tiles(self, tiles):
        t = np.ascontiguousarray(tiles, dtype=np.int32)
        if t.shape != (len(self.steps),):
            raise ValueError("tiles: one per shell (%d)" % len(self.steps))
        check(lib().gd_batch_set_tiles(self._h, t.ctypes.data))

    def tiles(self):
        t = np.zeros(len(self.steps), dtype=np.int32)
        check(lib().gd_batch_get_tiles(self._h, t.ctypes.data))
        return [int(v) for v in t]

    def tune(self, reps=5):
        """Runs the synthesis-tile autotuner (gd_batch_tune); returns the
        chosen tiles.  self.tune_ms: mean ms of the per-shell assignment and
        of the uniform tiles 0, 1, 2."""
        ms = np.zeros(4)
        check(lib().gd_batch_tune(self._h, int(reps), ms.ctypes.data))
        self.tune_ms = [float(v) for v in ms]
        return self.tiles()

    def synchronize(self):
        check(lib().gd_ctx_synchronize(context(self.steps[0].device)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gd_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

