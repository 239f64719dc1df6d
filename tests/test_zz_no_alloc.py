"""mse_stem_detect_noalloc (no allocation; fixed per-GPU workspace) against mse_stem_detect
on the same waves: two calls in a row, the second with fewer waves and detectors, and the
workspace limit refused with an error before any work is queued."""
import ctypes as C
import os

import numpy as np
import pytest

import cltem_amd
from cltem_amd import _abi
from cltem_amd import workloads as wl

pytestmark = pytest.mark.gpu

MSE_ERR_ARG = -1


def hip_runtime():
    """The HIP runtime libcltem_amd.so is bound to in this process (another library loaded
    earlier, torch's for one, may have brought its own copy: loading a second one fails)."""
    _abi.load()
    with open("/proc/self/maps") as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 6 and os.path.basename(parts[-1]).startswith("libamdhip64.so"):
                return C.CDLL(parts[-1])
    return C.CDLL("libamdhip64.so")


class DeviceBuffer:
    """Caller-owned device memory from the HIP runtime (the test allocates; the entry point
    under test does not)."""

    def __init__(self, host: np.ndarray):
        self.hip = hip_runtime()
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.host.nbytes)) == 0
        assert self.hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p),
                                  C.c_size_t(self.host.nbytes), 1) == 0   # host to device

    def read(self) -> np.ndarray:
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = None


def _noalloc(engine, dets, n_waves):
    buf = DeviceBuffer(np.full((n_waves, len(dets)), np.nan))
    try:
        rc = engine.lib.mse_stem_detect_noalloc(engine.h, dets, len(dets), 1, buf.ptr)
        assert rc == 0, engine.lib.mse_last_error()
        assert engine.lib.mse_synchronize(engine.h) == 0
        return buf.read()
    finally:
        buf.free()


def test_noalloc_detect_matches_stem_detect_twice():
    w = wl.stem_srtio3(nx=6, ny=6, nz=6, resolution=256, pixels=8, window_cells=(2, 3),
                       detectors=((60.0, 200.0), (0.0, 20.0), (10.0, 45.0, 7.5, -12.5)))
    spec = w["specimen"].displaced(0.08, 20261202, 0)
    with cltem_amd.Engine(0) as e:
        g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
        e.set_geometry(g)
        e.set_parameterisation(cltem_amd.load_params(w["params"]))
        e.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z, want_ids=False)
        e.build_transmission(0)
        pos = wl.stem_positions(g)
        probe = cltem_amd.make_probe(w["cond_ap_mrad"])

        # first call: 8 waves, three detectors (one of them off-centre)
        dets = cltem_amd.make_detectors(w["detectors"])
        e.set_wave_count(8)
        e.set_wave_configs([0] * 8)
        e.init_probes(pos[:8], probe)
        e.multislice()
        ref = e.stem_detect(dets)
        got = _noalloc(e, dets, 8)
        print(f"8 waves x 3 detectors: max |noalloc - stem_detect| = {np.max(np.abs(got - ref)):.3e}")
        assert np.all(np.isfinite(got)) and np.all(ref > 0)
        assert np.array_equal(got, ref)

        # second call: fewer waves and detectors, other probe positions
        dets2 = cltem_amd.make_detectors([w["detectors"][2]])
        e.set_wave_count(3)
        e.set_wave_configs([0] * 3)
        e.init_probes(pos[40:43], probe)
        e.multislice()
        ref2 = e.stem_detect(dets2)
        got2 = _noalloc(e, dets2, 3)
        print(f"3 waves x 1 detector: max |noalloc - stem_detect| = {np.max(np.abs(got2 - ref2)):.3e}")
        assert np.array_equal(got2, ref2)


def test_noalloc_detect_refuses_oversized_workspace():
    """8 detectors at 256^2 need 64 KiB of masks and 2 KiB per wave: 2100 waves exceed 4 MiB."""
    w = wl.stem_srtio3(nx=2, ny=2, nz=2, resolution=256, pixels=2, window_cells=(0, 1))
    with cltem_amd.Engine(0) as e:
        e.set_geometry(cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"])))
        e.set_wave_count(2100)
        dets = cltem_amd.make_detectors([(60.0, 200.0)] * 8)
        buf = DeviceBuffer(np.zeros(2100 * 8))
        try:
            rc = e.lib.mse_stem_detect_noalloc(e.h, dets, 8, 1, buf.ptr)
            assert rc == MSE_ERR_ARG
            assert b"exceeds the fixed 4194304 bytes" in e.lib.mse_last_error()
            # nothing was queued: the output is untouched and the engine still works
            assert e.lib.mse_synchronize(e.h) == 0
            assert not np.any(buf.read() != 0)
        finally:
            buf.free()


def test_noalloc_is_declared():
    assert "mse_stem_detect_noalloc" in _abi.EXPORTED
