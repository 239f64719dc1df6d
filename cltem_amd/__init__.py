"""cltem_amd: MI355X-native multislice engine behind clTEM's simulation interfaces.

The compute path is libcltem_amd.so (hand-written HIP kernels for gfx950 behind the
C ABI in include/cltem_amd.h).  This module is a thin ctypes binding of that ABI;
there is no Python or CPU fallback for any computation.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import MseDetector, MseGeometry, MseImaging, MseProbe, MseScenario
from .params import Parameterisation, load as load_params  # noqa: F401

__all__ = ["Engine", "MseError", "plan_geometry", "make_scenario", "make_probe", "make_detectors",
           "load_params", "Parameterisation"]


class MseError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _check(rc: int):
    if rc != _abi.MSE_OK:
        msg = _abi.load().mse_last_error()
        raise MseError(rc, msg.decode() if msg else "unknown error")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def make_scenario(**kw) -> MseScenario:
    """Fill an mse_scenario; keyword names follow the struct fields."""
    s = MseScenario()
    s.mode = kw.get("mode", 1)
    s.resolution = kw["resolution"]
    s.voltage_kv = kw.get("voltage_kv", 200.0)
    s.slice_thickness = kw.get("slice_thickness", 1.0)
    s.slice_offset = kw.get("slice_offset", 0.0)
    s.pad_xy[:] = kw.get("pad_xy", (-8.0, 8.0))
    s.pad_z[:] = kw.get("pad_z", (-3.0, 3.0))
    s.struct_min[:] = kw["struct_min"]
    s.struct_max[:] = kw["struct_max"]
    s.area[:] = kw["area"]
    s.area_padding = kw.get("area_padding", 0.0)
    s.stem_pixels[:] = kw.get("stem_pixels", (0, 0))
    s.beam_tilt_mrad = kw.get("beam_tilt_mrad", 0.0)
    s.beam_azimuth = kw.get("beam_azimuth", 0.0)
    s.limit_factor = kw.get("limit_factor", 2.0 / 3.0)
    s.full3d = int(kw.get("full3d", 0))
    s.full3d_integrals = kw.get("full3d_integrals", 20)
    s.moving_stem_frame = int(kw.get("moving_stem_frame", 0))
    return s


def plan_geometry(s: MseScenario) -> MseGeometry:
    g = MseGeometry()
    _check(_abi.load().mse_plan_geometry(C.byref(s), C.byref(g)))
    return g


def stem_pixel_frames(g: MseGeometry, pixels) -> np.ndarray:
    """Potential-grid origin (paddedSimLimitsX/Y(pixel)[0]) of each STEM pixel: its own field
    under the moving frame, sim_x[0], sim_y[0] otherwise.  Returns (n, 2) doubles."""
    lib = _abi.load()
    px = np.asarray(pixels, dtype=np.int64).reshape(-1)
    out = np.zeros((len(px), 2))
    for i, p in enumerate(px):
        _check(lib.mse_stem_pixel_frame(C.byref(g), int(p), _dp(out[i])))
    return out


def make_probe(cond_ap_mrad=20.0, smooth_mrad=0.0, taper=False, aberrations=None) -> MseProbe:
    """aberrations: dict like {"C10": 10.0, "C12": (mag, ang), "C30": 1e4, ...} in Angstrom/rad;
    complex terms are given as (magnitude, angle) like ComplexAberration."""
    p = MseProbe()
    ab = aberrations or {}
    for name in ("C10", "C30", "C50"):
        setattr(p, name, float(ab.get(name, 0.0)))
    for name in ("C12", "C21", "C23", "C32", "C34", "C41", "C43", "C45", "C52", "C54", "C56"):
        mag, ang = ab.get(name, (0.0, 0.0))
        z = mag * np.exp(1j * ang)
        getattr(p, name)[:] = (z.real, z.imag)
    p.cond_ap_mrad = cond_ap_mrad
    p.smooth_mrad = smooth_mrad
    p.edge_taper = int(taper)
    return p


def probe_aberration_vector(p: MseProbe) -> np.ndarray:
    """The 25 doubles of an mse_probe's aberrations in oracle order."""
    out = [p.C10, *p.C12, *p.C21, *p.C23, p.C30, *p.C32, *p.C34, *p.C41, *p.C43, *p.C45, p.C50,
           *p.C52, *p.C54, *p.C56]
    return np.array(out, dtype=np.float64)


def make_imaging(obj_ap_mrad=20.0, obj_smooth_mrad=0.0, beta_mrad=0.0, delta=0.0, aberrations=None) -> MseImaging:
    """Objective-lens parameters of the CTEM image (aberrations as in make_probe)."""
    im = MseImaging()
    ab = aberrations or {}
    for name in ("C10", "C30", "C50"):
        setattr(im, name, float(ab.get(name, 0.0)))
    for name in ("C12", "C21", "C23", "C32", "C34", "C41", "C43", "C45", "C52", "C54", "C56"):
        mag, ang = ab.get(name, (0.0, 0.0))
        z = mag * np.exp(1j * ang)
        getattr(im, name)[:] = (z.real, z.imag)
    im.obj_ap_mrad = obj_ap_mrad
    im.obj_smooth_mrad = obj_smooth_mrad
    im.beta_mrad = beta_mrad
    im.delta = delta
    return im


def imaging_vector(im: MseImaging) -> np.ndarray:
    """The 25 aberration doubles in the oracle's order (C10, C12 re/im, ..., C56 re/im)."""
    return probe_aberration_vector(im)


def make_detectors(rows):
    """rows: iterable of (inner, outer, xc, yc) in mrad -> ctypes array."""
    rows = [tuple(r) + (0.0,) * (4 - len(r)) for r in rows]
    arr = (MseDetector * len(rows))()
    for i, r in enumerate(rows):
        arr[i].inner_mrad, arr[i].outer_mrad, arr[i].xc_mrad, arr[i].yc_mrad = r
    return arr


class Engine:
    """One engine per GPU (C handle mse_engine*)."""

    def __init__(self, device: int = 0):
        self.lib = _abi.load()
        h = C.c_void_p()
        _check(self.lib.mse_create(device, C.byref(h)))
        self.h = h
        self.g: MseGeometry | None = None
        self._cfg_geom: dict = {}  # configurations with a geometry of their own
        self.n_waves = 0

    def close(self):
        if self.h:
            self.lib.mse_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup --------------------------------------------------------------------------
    def device_info(self):
        name = C.create_string_buffer(256)
        cu = C.c_int(0)
        bw = C.c_double(0)
        _check(self.lib.mse_device_info(self.h, name, 256, C.byref(cu), C.byref(bw)))
        return name.value.decode(), cu.value, bw.value

    def set_geometry(self, g: MseGeometry):
        _check(self.lib.mse_set_geometry(self.h, C.byref(g)))
        self.g = g
        self._cfg_geom = {}  # mse_set_geometry returns every configuration to the engine's

    def set_parameterisation(self, p: Parameterisation):
        v = np.ascontiguousarray(p.values, dtype=np.float64)
        _check(self.lib.mse_set_parameterisation(self.h, p.form, p.i_per_atom, _dp(v), len(v)))

    def config_geometry(self, config: int) -> MseGeometry:
        """The geometry a configuration is binned and sliced with (its own one after
        set_config_geometry, else the engine's)."""
        return self._cfg_geom.get(config, self.g)

    def sort_atoms(self, config: int, x, y, z, Z, want_ids: bool = True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        z = np.ascontiguousarray(z, dtype=np.float64)
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        n = len(x)
        g = self.config_geometry(config)
        bs = np.zeros(g.n_slices * g.blocks_x * g.blocks_y + 1, dtype=np.int32)
        bid = np.zeros(max(n, 1), dtype=np.int32)
        zid = np.zeros(max(n, 1), dtype=np.int32)
        nr = C.c_int(0)
        nullp = C.POINTER(C.c_int)()
        _check(self.lib.mse_sort_atoms(self.h, config, _dp(x), _dp(y), _dp(z), _ip(Z), n,
                                       _ip(bid) if want_ids else nullp,
                                       _ip(zid) if want_ids else nullp, _ip(bs), C.byref(nr)))
        return dict(block_ids=bid[:n], z_ids=zid[:n], block_start=bs, n_in_range=nr.value)

    def build_transmission(self, config: int, first: int = 0, count: int | None = None):
        count = self.config_geometry(config).n_slices - first if count is None else count
        _check(self.lib.mse_build_transmission(self.h, config, first, count))

    def read_transmission(self, config: int, slice_: int) -> np.ndarray:
        res = self.g.resolution
        out = np.zeros(2 * res * res, dtype=np.float32)
        _check(self.lib.mse_read_transmission(self.h, config, slice_,
                                              out.ctypes.data_as(C.POINTER(C.c_float))))
        return out.view(np.complex64).reshape(res, res)

    def set_config_geometry(self, config: int, g: MseGeometry):
        _check(self.lib.mse_set_config_geometry(self.h, config, C.byref(g)))
        self._cfg_geom[config] = g

    def set_config_beam(self, config: int, k_vec=None):
        """Beam wave vector of a configuration (plasmon tilt, modifyBeamTilt); None restores
        the geometry's.  Rebuild its transmission functions for the slices still to come."""
        if k_vec is None:
            _check(self.lib.mse_set_config_beam(self.h, config, None))
        else:
            k = np.ascontiguousarray(k_vec, dtype=np.float64).reshape(3)
            _check(self.lib.mse_set_config_beam(self.h, config, _dp(k)))

    def set_mixed_potentials(self, first_config: int = 0, n_sets: int = 0, seed: int = 0):
        """Mixed static potentials: every wave draws one of configurations first_config ..
        first_config + n_sets - 1 per slice (keyed by its probe index); n_sets < 2 turns it off."""
        _check(self.lib.mse_set_mixed_potentials(self.h, first_config, n_sets, seed & 0xFFFFFFFFFFFFFFFF))

    # -- waves --------------------------------------------------------------------------
    def set_wave_count(self, n: int):
        _check(self.lib.mse_set_wave_count(self.h, n))
        self.n_waves = n

    def set_wave_configs(self, configs):
        c = np.ascontiguousarray(configs, dtype=np.int32)
        _check(self.lib.mse_set_wave_configs(self.h, _ip(c), len(c)))

    def init_plane_wave(self, value: float = 1.0):
        _check(self.lib.mse_init_plane_wave(self.h, value))

    def init_probes(self, positions, probe: MseProbe):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 2)
        _check(self.lib.mse_init_probes(self.h, _dp(pos), len(pos), C.byref(probe)))

    def set_focus_spread(self, delta_c10=None):
        """Per-probe focus changes (Angstrom) added to C10 by later probe set-ups; None clears."""
        d = np.zeros(0) if delta_c10 is None else np.ascontiguousarray(delta_c10, dtype=np.float64)
        _check(self.lib.mse_set_focus_spread(self.h, _dp(d), len(d)))

    def set_reference_offset(self, config: int, dx: float, dy: float):
        """Shift the potential grid origin of `config` (source-size incoherence)."""
        _check(self.lib.mse_set_reference_offset(self.h, config, dx, dy))

    def init_probes_range(self, first_wave: int, positions, probe: MseProbe):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 2)
        _check(self.lib.mse_init_probes_range(self.h, first_wave, _dp(pos), len(pos), C.byref(probe)))

    def multislice(self, first: int = 0, count: int | None = None, graph: bool = False):
        """Run slices [first, first + count) for every wave; graph=True replays the whole
        range from one captured hipGraph."""
        count = self.g.n_slices - first if count is None else count
        fn = self.lib.mse_multislice_graph if graph else self.lib.mse_multislice
        _check(fn(self.h, first, count))

    def stem_detect(self, detectors, sqrt2: bool = True) -> np.ndarray:
        out = np.zeros((self.n_waves, len(detectors)))
        _check(self.lib.mse_stem_detect(self.h, detectors, len(detectors), int(sqrt2), _dp(out)))
        return out

    def diffraction(self, wave: int = 0) -> np.ndarray:
        res = self.g.resolution
        out = np.zeros((res, res))
        _check(self.lib.mse_diffraction(self.h, wave, _dp(out)))
        return out

    def exit_wave(self, wave: int = 0) -> np.ndarray:
        res = self.g.resolution
        out = np.zeros(2 * res * res)
        _check(self.lib.mse_exit_wave(self.h, wave, _dp(out)))
        return out.view(np.complex128).reshape(res, res)

    def ctem_image(self, imaging: MseImaging, accum=None) -> np.ndarray:
        """CTEM image (objective aperture, aberrations, coherence envelopes) of every wave,
        summed over the waves in FP64; the exit waves are kept."""
        res = self.g.resolution
        out = np.zeros((res, res)) if accum is None else accum
        assert out.shape == (res, res) and out.dtype == np.float64 and out.flags.c_contiguous
        _check(self.lib.mse_ctem_image(self.h, C.byref(imaging), int(accum is not None), _dp(out)))
        return out

    def ccd_image(self, image, dqe, ntf, binning: int = 1, dose_per_pixel: float = 100.0,
                  conversion: float = 1.0, seed: int = 0) -> np.ndarray:
        """CCD DQE filter, seeded Poisson dose and NTF filter applied to a CTEM image."""
        res = self.g.resolution
        img = np.ascontiguousarray(image, dtype=np.float64)
        d = np.ascontiguousarray(dqe, dtype=np.float64)
        n = np.ascontiguousarray(ntf, dtype=np.float64)
        assert img.shape == (res, res) and d.shape == n.shape
        out = np.zeros((res, res))
        _check(self.lib.mse_ccd_image(self.h, _dp(img), _dp(d), _dp(n), len(d), binning, dose_per_pixel,
                                      conversion, C.c_ulonglong(seed), _dp(out)))
        return out

    def exit_intensity_sum(self, accum=None) -> np.ndarray:
        res = self.g.resolution
        out = np.zeros((res, res)) if accum is None else accum
        assert out.shape == (res, res) and out.dtype == np.float64 and out.flags.c_contiguous
        _check(self.lib.mse_exit_intensity_sum(self.h, int(accum is not None), _dp(out)))
        return out

    def diffraction_sum(self, wave_group, n_groups: int, accum=None) -> np.ndarray:
        res = self.g.resolution
        grp = np.ascontiguousarray(wave_group, dtype=np.int32)
        out = np.zeros((n_groups, res, res)) if accum is None else accum
        assert out.shape == (n_groups, res, res) and out.dtype == np.float64 and out.flags.c_contiguous
        _check(self.lib.mse_diffraction_sum(self.h, _ip(grp), n_groups, int(accum is not None), _dp(out)))
        return out

    def forward_checkpointed(self, config: int = 0) -> np.ndarray:
        """Run every slice of `config` from the current entrance waves, keeping every state
        (within the checkpoint budget) or sqrt(n) checkpoints for adjoint_positions; returns
        |psi_exit|^2, shape (n_waves, res, res)."""
        res = self.g.resolution
        out = np.zeros((self.n_waves, res, res))
        _check(self.lib.mse_forward_checkpointed(self.h, config, _dp(out)))
        return out

    def set_checkpoint_budget(self, nbytes: int):
        """Device bytes forward_checkpointed may spend on keeping every forward state (default
        1 GiB); 0 always keeps sqrt(n) checkpoints and recomputes the states in between."""
        _check(self.lib.mse_set_checkpoint_budget(self.h, int(nbytes)))

    def adjoint_positions(self, dl_di, n_atoms: int) -> np.ndarray:
        """dL/d(x, y, z) of every atom of the last sort of the checkpointed configuration
        for a loss with dL/dI = dl_di (shape (n_waves, res, res)); returns (n_atoms, 3)."""
        res = self.g.resolution
        w = np.ascontiguousarray(dl_di, dtype=np.float64)
        assert w.size == self.n_waves * res * res
        out = np.zeros((max(n_atoms, 0), 3))
        _check(self.lib.mse_adjoint_positions(self.h, _dp(w), _dp(out)))
        return out

    def stem_batch(self, config: int, positions, batch: int, probe: MseProbe, detectors,
                   sqrt2: bool = True, use_graph: bool = False, accum=None) -> np.ndarray:
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 2)
        if accum is None:
            accum = np.zeros((len(pos), len(detectors)))
        assert accum.shape == (len(pos), len(detectors)) and accum.dtype == np.float64
        _check(self.lib.mse_stem_batch(self.h, config, _dp(pos), len(pos), batch, C.byref(probe),
                                       detectors, len(detectors), int(sqrt2), int(use_graph),
                                       _dp(accum)))
        self.n_waves = max(self.n_waves, batch)
        return accum

    def stem_batch_frames(self, config: int, positions, frames, batch: int, probe: MseProbe, detectors,
                          sqrt2: bool = True, accum=None) -> np.ndarray:
        """Moving STEM frame: probe p in its own field whose potential-grid origin is frames[p]."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 2)
        fr = np.ascontiguousarray(frames, dtype=np.float64).reshape(-1, 2)
        assert fr.shape == pos.shape
        if accum is None:
            accum = np.zeros((len(pos), len(detectors)))
        assert accum.shape == (len(pos), len(detectors)) and accum.dtype == np.float64
        _check(self.lib.mse_stem_batch_frames(self.h, config, _dp(pos), _dp(fr), len(pos), batch, C.byref(probe),
                                              detectors, len(detectors), int(sqrt2), _dp(accum)))
        self.n_waves = max(self.n_waves, batch)
        return accum

    # -- tuning -------------------------------------------------------------------------
    def set_potential_tables(self, enable: bool = True, lds: bool = True):
        """Tabulated potentials (default) or the formulas evaluated directly, from the next build;
        lds=False keeps the full-3D kernels' table reads in global memory (same values)."""
        _check(self.lib.mse_set_potential_tables(self.h, (1 if lds else 2) if enable else 0))

    def set_kernel_variant(self, kernel: int, variant: int):
        _check(self.lib.mse_set_kernel_variant(self.h, kernel, variant))

    def autotune(self, config: int = 0, n_waves: int = 64, reps: int = 5):
        row = np.zeros(16)
        col = np.zeros(16)
        br = C.c_int(0)
        bc = C.c_int(0)
        _check(self.lib.mse_autotune(self.h, config, n_waves, reps, _dp(row), _dp(col), C.byref(br),
                                     C.byref(bc)))
        self.n_waves = max(self.n_waves, n_waves)
        nr, nc = 7, 14  # row / column kernel variants
        return dict(row_us=row[:nr].tolist(), col_us=col[:nc].tolist(), best_row=br.value,
                    best_col=bc.value)

    def time_sweeps(self, slice_: int = 0, reps: int = 10):
        """Average microseconds per launch of the row and column passes (current variants,
        every wave with its own configuration); the waves' contents become undefined."""
        r = C.c_double(0)
        c = C.c_double(0)
        _check(self.lib.mse_time_sweeps(self.h, slice_, reps, C.byref(r), C.byref(c)))
        return r.value, c.value

    def bandwidth_probe(self, pattern: int, n_waves: int, reps: int = 5, variant: int = 0) -> float:
        us = C.c_double(0)
        _check(self.lib.mse_bandwidth_probe(self.h, pattern, variant, n_waves, reps, C.byref(us)))
        self.n_waves = max(self.n_waves, n_waves)
        return us.value

    # -- profiling ----------------------------------------------------------------------
    def profile(self, enable: bool = True):
        _check(self.lib.mse_profile_enable(self.h, int(enable)))

    def profile_read(self, kernel_id: int):
        n = C.c_int(0)
        ms = C.c_double(0)
        _check(self.lib.mse_profile_read(self.h, kernel_id, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def profile_reset(self):
        _check(self.lib.mse_profile_reset(self.h))

    def synchronize(self):
        _check(self.lib.mse_synchronize(self.h))


class EngineGroup:
    """Engines of one device (one per grid size) whose whole slice loops run as one graph
    launch (mse_group_multislice).  Close the group before its engines."""

    def __init__(self, engines):
        self.lib = _abi.load()
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h.value for e in self.engines])
        h = C.c_void_p()
        _check(self.lib.mse_group_create(arr, len(self.engines), C.byref(h)))
        self.h = h

    def multislice(self):
        _check(self.lib.mse_group_multislice(self.h))

    def close(self):
        if self.h:
            self.lib.mse_group_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# layout of the tabulated potentials (potential_table.hpp)
TAB_OCTAVE_LO, TAB_OCTAVE_HI, TAB_SUB_BITS = -20, 8, 5
TAB_BUCKETS = (TAB_OCTAVE_HI - TAB_OCTAVE_LO) << TAB_SUB_BITS
TAB_MANT_BITS = 23 - TAB_SUB_BITS


def potential_table(p: Parameterisation, Z: int, full3d: bool = False) -> np.ndarray:
    """The engine's tabulated potential of element Z: (TAB_BUCKETS, 4) cubic coefficients per
    bucket of r^2 (see mse_potential_table), evaluated by potential_table_eval."""
    out = np.zeros(4 * TAB_BUCKETS, dtype=np.float32)
    v = np.ascontiguousarray(p.values, dtype=np.float64)
    _check(_abi.load().mse_potential_table(p.form, int(full3d), _dp(v), len(v), p.i_per_atom, Z,
                                           out.ctypes.data_as(C.POINTER(C.c_float))))
    return out.reshape(TAB_BUCKETS, 4)


def potential_table_eval(tab: np.ndarray, r2) -> np.ndarray:
    """V at squared distances r2 from a table of potential_table, as the kernels evaluate it
    (float32 r^2 bits pick the bucket and its fraction t; cubic in t)."""
    bits = np.asarray(r2, dtype=np.float32).view(np.uint32).astype(np.int64)
    idx = np.clip((bits >> TAB_MANT_BITS) - ((127 + TAB_OCTAVE_LO) << TAB_SUB_BITS), 0, TAB_BUCKETS - 1)
    t = (bits & ((1 << TAB_MANT_BITS) - 1)).astype(np.float32) * np.float32(1.0 / (1 << TAB_MANT_BITS))
    c = tab[idx].astype(np.float32)
    return ((c[:, 3] * t + c[:, 2]) * t + c[:, 1]) * t + c[:, 0]
