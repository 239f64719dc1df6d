"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module.  It loads oracle/liboracle_mse.so (built by oracle/Makefile) and wraps the
functions declared in oracle/mse_oracle.h with numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_mse.so")


class OrcScenario(C.Structure):
    _fields_ = [
        ("mode", C.c_int),
        ("resolution", C.c_int),
        ("voltage_kv", C.c_double),
        ("slice_thickness", C.c_double),
        ("slice_offset", C.c_double),
        ("pad_xy", C.c_double * 2),
        ("pad_z", C.c_double * 2),
        ("struct_min", C.c_double * 3),
        ("struct_max", C.c_double * 3),
        ("area", C.c_double * 4),
        ("area_padding", C.c_double),
        ("stem_pixels", C.c_int * 2),
        ("beam_tilt_mrad", C.c_double),
        ("beam_azimuth", C.c_double),
        ("limit_factor", C.c_double),
        ("full3d", C.c_int),
        ("full3d_integrals", C.c_int),
        ("moving_stem_frame", C.c_int),
    ]


class OrcGeometry(C.Structure):
    _fields_ = [
        ("resolution", C.c_int),
        ("full_x", C.c_double * 2),
        ("full_y", C.c_double * 2),
        ("sim_x", C.c_double * 2),
        ("sim_y", C.c_double * 2),
        ("z_lims", C.c_double * 2),
        ("pad_z", C.c_double * 2),
        ("real_scale", C.c_double),
        ("inverse_scale", C.c_double),
        ("inverse_scale_angle", C.c_double),
        ("n_slices", C.c_int),
        ("pre_slices", C.c_int),
        ("blocks_x", C.c_int),
        ("blocks_y", C.c_int),
        ("block_scale_x", C.c_double),
        ("block_scale_y", C.c_double),
        ("load_x", C.c_int),
        ("load_y", C.c_int),
        ("load_z", C.c_int),
        ("slice_thickness", C.c_double),
        ("wavelength", C.c_double),
        ("wavenumber", C.c_double),
        ("sigma", C.c_double),
        ("sigma_eff", C.c_double),
        ("k_vec", C.c_double * 3),
        ("kmax_band", C.c_double),
        ("limit_factor", C.c_double),
        ("beam_tilt_mrad", C.c_double),
        ("beam_azimuth", C.c_double),
        ("stem_scale", C.c_double * 2),
        ("stem_start", C.c_double * 2),
        ("stem_pixels", C.c_int * 2),
        ("full3d", C.c_int),
        ("full3d_integrals", C.c_int),
        ("moving_frame", C.c_int),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.orc_wavelength.restype = C.c_double
        L.orc_wavelength.argtypes = [C.c_double]
        L.orc_sigma.restype = C.c_double
        L.orc_sigma.argtypes = [C.c_double]
        L.orc_geometry.argtypes = [C.POINTER(OrcScenario), C.POINTER(OrcGeometry)]
        L.orc_geometry_pixel.argtypes = [C.POINTER(OrcScenario), C.c_int, C.POINTER(OrcGeometry)]
        L.orc_sort_atoms.argtypes = [C.POINTER(OrcGeometry), dp, dp, dp, ip, C.c_int, dp, dp, dp,
                                     ip, ip, ip, ip, ip, ip]
        L.orc_bessk0.restype = C.c_double
        L.orc_bessk0.argtypes = [C.c_double]
        L.orc_bessk1.restype = C.c_double
        L.orc_bessk1.argtypes = [C.c_double]
        L.orc_vproj.restype = C.c_double
        L.orc_vproj.argtypes = [C.c_int, dp, C.c_int, C.c_int, C.c_double]
        L.orc_v3d.restype = C.c_double
        L.orc_v3d.argtypes = [C.c_int, dp, C.c_int, C.c_int, C.c_double]
        L.orc_kgrid.argtypes = [C.POINTER(OrcGeometry), dp, dp]
        L.orc_potential_slice.argtypes = [C.POINTER(OrcGeometry), dp, dp, dp, ip, ip, dp, C.c_int,
                                          C.c_int, C.c_int, dp, C.c_int]
        L.orc_transmission.argtypes = [C.POINTER(OrcGeometry), dp, dp, dp, ip, ip, dp, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int]
        L.orc_fft2.argtypes = [dp, C.c_int, C.c_int, C.c_int]
        L.orc_fft2_fast.argtypes = [dp, C.c_int, C.c_int, C.c_int]
        L.orc_band_limit.argtypes = [C.POINTER(OrcGeometry), dp]
        L.orc_propagator.argtypes = [C.POINTER(OrcGeometry), dp]
        L.orc_probe_kspace.argtypes = [C.POINTER(OrcGeometry), C.c_double, C.c_double, dp,
                                       C.c_double, C.c_double, C.c_int, dp]
        L.orc_slice_step.argtypes = [C.c_int, dp, dp, dp, dp]
        L.orc_stem_pixel.restype = C.c_double
        L.orc_stem_pixel.argtypes = [C.c_int, dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int]
        L.orc_diffraction.argtypes = [C.c_int, dp, dp]
        L.orc_translate_image.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, C.c_double]
        L.orc_rotate_vector_spherical.argtypes = [dp, dp, C.c_double, C.c_double]
        L.orc_rotate_vector_spherical.restype = None
        L.orc_run_plane.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp]
        L.orc_ctem_image.argtypes = [C.POINTER(OrcGeometry), dp, dp, dp]
        L.orc_ctem_intensity.argtypes = [C.POINTER(OrcGeometry), dp, dp, dp, ip, C.c_int, dp, C.c_int,
                                         C.c_int, dp, C.c_int]
        L.orc_ctem_loss_fd.argtypes = [C.POINTER(OrcGeometry), dp, dp, dp, ip, C.c_int, dp, C.c_int,
                                       C.c_int, dp, ip, ip, C.c_int, C.c_double, dp, dp, C.c_int]
        for f in (L.orc_run_stem, L.orc_run_stem_f32):
            f.argtypes = [C.POINTER(OrcGeometry), dp, dp, dp, C.c_int, dp, C.c_double, C.c_double,
                          C.c_int, dp, C.c_int, C.c_int, dp, C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _cplx(a):
    """complex128 array -> interleaved float64 view (contiguous copy)."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).view(np.float64)


def nthreads_default() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def wavelength(kv: float) -> float:
    return lib().orc_wavelength(kv)


def sigma(kv: float) -> float:
    return lib().orc_sigma(kv)


def scenario(**kw) -> OrcScenario:
    s = OrcScenario()
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


def geometry(s: OrcScenario) -> OrcGeometry:
    g = OrcGeometry()
    if lib().orc_geometry(C.byref(s), C.byref(g)) != 0:
        raise ValueError("orc_geometry rejected the scenario")
    return g


def geometry_pixel(s: OrcScenario, pixel: int) -> OrcGeometry:
    """The geometry of STEM pixel `pixel`'s job (its own field under the moving frame)."""
    g = OrcGeometry()
    if lib().orc_geometry_pixel(C.byref(s), int(pixel), C.byref(g)) != 0:
        raise ValueError("orc_geometry_pixel rejected the scenario or pixel")
    return g


def sort_atoms(g: OrcGeometry, x, y, z, Z):
    x, y, z, Z = _f64(x), _f64(y), _f64(z), _i32(Z)
    n = len(x)
    sx, sy, sz = (np.zeros(max(n, 1)) for _ in range(3))
    sZ, bid, zid = (np.zeros(max(n, 1), dtype=np.int32) for _ in range(3))
    bs = np.zeros(g.n_slices * g.blocks_x * g.blocks_y + 1, dtype=np.int32)
    nf, nr = C.c_int(0), C.c_int(0)
    lib().orc_sort_atoms(C.byref(g), _dp(x), _dp(y), _dp(z), _ip(Z), n, _dp(sx), _dp(sy), _dp(sz),
                         _ip(sZ), _ip(bid), _ip(zid), _ip(bs), C.byref(nf), C.byref(nr))
    k = nf.value
    return dict(x=sx[:k], y=sy[:k], z=sz[:k], Z=sZ[:k], block_ids=bid[:k], z_ids=zid[:k],
                block_start=bs, n_filtered=k, n_in_range=nr.value)


def kgrid(g: OrcGeometry):
    kx = np.zeros(g.resolution)
    ky = np.zeros(g.resolution)
    lib().orc_kgrid(C.byref(g), _dp(kx), _dp(ky))
    return kx, ky


def potential_slice(g, sorted_atoms, params, form, icount, slice_, nthreads=None):
    res = g.resolution
    out = np.zeros(res * res)
    a = sorted_atoms
    lib().orc_potential_slice(C.byref(g), _dp(_f64(a["x"])), _dp(_f64(a["y"])), _dp(_f64(a["z"])),
                              _ip(_i32(a["Z"])), _ip(_i32(a["block_start"])), _dp(_f64(params)),
                              form, icount, slice_, _dp(out), nthreads or nthreads_default())
    return out.reshape(res, res)


def transmission(g, sorted_atoms, params, form, icount, first=0, count=None, band_limit=True,
                 nthreads=None):
    res = g.resolution
    count = g.n_slices - first if count is None else count
    out = np.zeros(count * res * res * 2)
    a = sorted_atoms
    keep = [_f64(a["x"]), _f64(a["y"]), _f64(a["z"]), _i32(a["Z"]), _i32(a["block_start"]),
            _f64(params)]
    lib().orc_transmission(C.byref(g), _dp(keep[0]), _dp(keep[1]), _dp(keep[2]), _ip(keep[3]),
                           _ip(keep[4]), _dp(keep[5]), form, icount, first, count,
                           int(band_limit), _dp(out), nthreads or nthreads_default())
    return out.view(np.complex128).reshape(count, res, res)


def ctem_image(g, recip, aberr=None, obj_ap_mrad=20.0, obj_smooth_mrad=0.0, beta_mrad=0.0, delta=0.0):
    """CTEM image |IFFT(ctem_image(Psi))|^2 of the unshifted Psi(k) `recip`."""
    res = g.resolution
    lens = np.zeros(29)
    if aberr is not None:
        lens[:25] = aberr
    lens[25:] = (obj_ap_mrad, obj_smooth_mrad, beta_mrad, delta)
    r = _cplx(recip)
    out = np.zeros(res * res)
    assert lib().orc_ctem_image(C.byref(g), _dp(r), _dp(lens), _dp(out)) == 0
    return out.reshape(res, res)


def ctem_intensity(g, x, y, z, Z, params, form, icount, nthreads=None):
    """|psi_exit|^2 of a plane wave (value 1) through the whole path from unsorted atoms."""
    res = g.resolution
    x, y, z, Z, prm = _f64(x), _f64(y), _f64(z), _i32(Z), _f64(params)
    out = np.zeros(res * res)
    rc = lib().orc_ctem_intensity(C.byref(g), _dp(x), _dp(y), _dp(z), _ip(Z), len(x), _dp(prm), form,
                                  icount, _dp(out), nthreads or nthreads_default())
    assert rc == 0
    return out.reshape(res, res)


def ctem_loss_fd(g, x, y, z, Z, params, form, icount, target, atoms, axes, h=1e-3, nthreads=None):
    """Central differences of L = sum (|psi_exit|^2 - target)^2 in coordinate axes[i]
    (0 x, 1 y, 2 z) of atom atoms[i]; returns (grad, L at the given positions)."""
    x, y, z, Z, prm = _f64(x), _f64(y), _f64(z), _i32(Z), _f64(params)
    tgt = _f64(target).reshape(-1)
    at, ax = _i32(atoms), _i32(axes)
    grad = np.zeros(len(at))
    L0 = C.c_double(0.0)
    rc = lib().orc_ctem_loss_fd(C.byref(g), _dp(x), _dp(y), _dp(z), _ip(Z), len(x), _dp(prm), form, icount,
                                _dp(tgt), _ip(at), _ip(ax), len(at), h, _dp(grad), C.byref(L0),
                                nthreads or nthreads_default())
    assert rc == 0
    return grad, L0.value


def fft2(a, inverse=False):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    h, w = a.shape
    lib().orc_fft2(_dp(a.view(np.float64)), w, h, int(inverse))
    return a


def fft2_fast(a, inverse=False):
    """The CPU baseline's FFT (power-of-two sizes), in double."""
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    h, w = a.shape
    assert lib().orc_fft2_fast(_dp(a.view(np.float64)), w, h, int(inverse)) == 0
    return a


def band_limit(g, t):
    t = np.ascontiguousarray(t, dtype=np.complex128).copy()
    lib().orc_band_limit(C.byref(g), _dp(t.view(np.float64)))
    return t


def propagator(g):
    res = g.resolution
    P = np.zeros((res, res), dtype=np.complex128)
    lib().orc_propagator(C.byref(g), _dp(P.view(np.float64)))
    return P


def probe_kspace(g, px, py, aberr=None, cond_ap_mrad=20.0, smooth_mrad=0.0, taper=False):
    res = g.resolution
    ab = _f64(np.zeros(25) if aberr is None else aberr)
    out = np.zeros((res, res), dtype=np.complex128)
    lib().orc_probe_kspace(C.byref(g), px, py, _dp(ab), cond_ap_mrad, smooth_mrad, int(taper),
                           _dp(out.view(np.float64)))
    return out


def slice_step(psi, t, P):
    res = psi.shape[0]
    psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    recip = np.zeros_like(psi)
    t = np.ascontiguousarray(t, dtype=np.complex128)
    P = np.ascontiguousarray(P, dtype=np.complex128)
    lib().orc_slice_step(res, _dp(psi.view(np.float64)), _dp(recip.view(np.float64)),
                         _dp(t.view(np.float64)), _dp(P.view(np.float64)))
    return psi, recip


def stem_pixel(recip, inner, outer, xc, yc, angle_scale, sqrt2=True):
    recip = np.ascontiguousarray(recip, dtype=np.complex128)
    return lib().orc_stem_pixel(recip.shape[0], _dp(recip.view(np.float64)), inner, outer, xc, yc,
                                angle_scale, int(sqrt2))


def diffraction(recip):
    recip = np.ascontiguousarray(recip, dtype=np.complex128)
    res = recip.shape[0]
    out = np.zeros((res, res))
    lib().orc_diffraction(res, _dp(recip.view(np.float64)), _dp(out))
    return out


def translate_image(img, dkx, dky, inverse_scale):
    img = np.ascontiguousarray(img, dtype=np.float64)
    out = np.zeros_like(img)
    lib().orc_translate_image(img.shape[0], _dp(img), _dp(out), dkx, dky, inverse_scale)
    return out


def rotate_vector_spherical(z, y, theta, phi):
    z = _f64(z).copy()
    y = _f64(y).copy()
    lib().orc_rotate_vector_spherical(_dp(z), _dp(y), theta, phi)
    return z, y


def run_plane(trans, P):
    trans = np.ascontiguousarray(trans, dtype=np.complex128)
    ns, res, _ = trans.shape
    psi = np.zeros((res, res), dtype=np.complex128)
    recip = np.zeros((res, res), dtype=np.complex128)
    P = np.ascontiguousarray(P, dtype=np.complex128)
    lib().orc_run_plane(res, ns, _dp(trans.view(np.float64)), _dp(P.view(np.float64)),
                        _dp(psi.view(np.float64)), _dp(recip.view(np.float64)))
    return psi, recip


def run_stem(g, trans, P, positions, detectors, aberr=None, cond_ap_mrad=20.0, smooth_mrad=0.0,
             taper=False, sqrt2=True, nthreads=None, fp32=False):
    """positions: (n, 2) absolute Angstrom; detectors: (ndet, 4) inner, outer, xc, yc in mrad.
    Runs g.n_slices slices: trans must hold at least that many."""
    trans = np.ascontiguousarray(trans, dtype=np.complex128)
    P = np.ascontiguousarray(P, dtype=np.complex128)
    res = g.resolution
    if trans.ndim != 3 or trans.shape[0] < g.n_slices or trans.shape[1:] != (res, res) or P.shape != (res, res):
        raise ValueError("transmission functions or propagator do not match the geometry")
    pos = _f64(positions).reshape(-1, 2)
    det = _f64(detectors).reshape(-1, 4)
    ab = _f64(np.zeros(25) if aberr is None else aberr)
    out = np.zeros((len(pos), len(det)))
    f = lib().orc_run_stem_f32 if fp32 else lib().orc_run_stem
    f(C.byref(g), _dp(trans.view(np.float64)), _dp(P.view(np.float64)), _dp(pos), len(pos),
      _dp(ab), cond_ap_mrad, smooth_mrad, int(taper), _dp(det), len(det), int(sqrt2), _dp(out),
      nthreads or nthreads_default())
    return out
