"""ctypes mirror of include/cltem_amd.h (the C ABI of libcltem_amd.so)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcltem_amd.so")

MSE_OK = 0


class MseScenario(C.Structure):
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


class MseGeometry(C.Structure):
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
        ("frame_step", C.c_double),
        ("frame_half", C.c_double * 2),
        ("pad_xy", C.c_double * 2),
    ]


class MseProbe(C.Structure):
    _fields_ = [
        ("C10", C.c_double),
        ("C12", C.c_double * 2),
        ("C21", C.c_double * 2),
        ("C23", C.c_double * 2),
        ("C30", C.c_double),
        ("C32", C.c_double * 2),
        ("C34", C.c_double * 2),
        ("C41", C.c_double * 2),
        ("C43", C.c_double * 2),
        ("C45", C.c_double * 2),
        ("C50", C.c_double),
        ("C52", C.c_double * 2),
        ("C54", C.c_double * 2),
        ("C56", C.c_double * 2),
        ("cond_ap_mrad", C.c_double),
        ("smooth_mrad", C.c_double),
        ("edge_taper", C.c_int),
    ]


class MseImaging(C.Structure):
    _fields_ = [
        ("C10", C.c_double),
        ("C12", C.c_double * 2),
        ("C21", C.c_double * 2),
        ("C23", C.c_double * 2),
        ("C30", C.c_double),
        ("C32", C.c_double * 2),
        ("C34", C.c_double * 2),
        ("C41", C.c_double * 2),
        ("C43", C.c_double * 2),
        ("C45", C.c_double * 2),
        ("C50", C.c_double),
        ("C52", C.c_double * 2),
        ("C54", C.c_double * 2),
        ("C56", C.c_double * 2),
        ("obj_ap_mrad", C.c_double),
        ("obj_smooth_mrad", C.c_double),
        ("beta_mrad", C.c_double),
        ("delta", C.c_double),
    ]


class MseDetector(C.Structure):
    _fields_ = [
        ("inner_mrad", C.c_double),
        ("outer_mrad", C.c_double),
        ("xc_mrad", C.c_double),
        ("yc_mrad", C.c_double),
    ]


# (name, restype, argtypes); every function returns int status unless stated
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
_eng = C.c_void_p

SIGNATURES = [
    ("mse_create", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("mse_destroy", C.c_int, [_eng]),
    ("mse_last_error", C.c_char_p, []),
    ("mse_synchronize", C.c_int, [_eng]),
    ("mse_device_info", C.c_int, [_eng, C.c_char_p, C.c_int, _ip, _dp]),
    ("mse_plan_geometry", C.c_int, [C.POINTER(MseScenario), C.POINTER(MseGeometry)]),
    ("mse_stem_pixel_frame", C.c_int, [C.POINTER(MseGeometry), C.c_int, _dp]),
    ("mse_set_geometry", C.c_int, [_eng, C.POINTER(MseGeometry)]),
    ("mse_set_parameterisation", C.c_int, [_eng, C.c_int, C.c_int, _dp, C.c_int]),
    ("mse_sort_atoms", C.c_int, [_eng, C.c_int, _dp, _dp, _dp, _ip, C.c_int, _ip, _ip, _ip, _ip]),
    ("mse_build_transmission", C.c_int, [_eng, C.c_int, C.c_int, C.c_int]),
    ("mse_read_transmission", C.c_int, [_eng, C.c_int, C.c_int, _fp]),
    ("mse_set_config_geometry", C.c_int, [_eng, C.c_int, C.POINTER(MseGeometry)]),
    ("mse_set_config_beam", C.c_int, [_eng, C.c_int, C.POINTER(C.c_double)]),
    ("mse_set_mixed_potentials", C.c_int, [_eng, C.c_int, C.c_int, C.c_ulonglong]),
    ("mse_set_wave_count", C.c_int, [_eng, C.c_int]),
    ("mse_set_wave_configs", C.c_int, [_eng, _ip, C.c_int]),
    ("mse_init_plane_wave", C.c_int, [_eng, C.c_double]),
    ("mse_init_probes", C.c_int, [_eng, _dp, C.c_int, C.POINTER(MseProbe)]),
    ("mse_init_probes_range", C.c_int, [_eng, C.c_int, _dp, C.c_int, C.POINTER(MseProbe)]),
    ("mse_multislice", C.c_int, [_eng, C.c_int, C.c_int]),
    ("mse_stem_detect", C.c_int, [_eng, C.POINTER(MseDetector), C.c_int, C.c_int, _dp]),
    ("mse_stem_detect_noalloc", C.c_int, [_eng, C.POINTER(MseDetector), C.c_int, C.c_int, C.c_void_p]),
    ("mse_diffraction", C.c_int, [_eng, C.c_int, _dp]),
    ("mse_exit_wave", C.c_int, [_eng, C.c_int, _dp]),
    ("mse_ctem_image", C.c_int, [_eng, C.POINTER(MseImaging), C.c_int, _dp]),
    ("mse_ccd_image", C.c_int, [_eng, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_ulonglong,
                                _dp]),
    ("mse_exit_intensity_sum", C.c_int, [_eng, C.c_int, _dp]),
    ("mse_diffraction_sum", C.c_int, [_eng, _ip, C.c_int, C.c_int, _dp]),
    ("mse_multislice_graph", C.c_int, [_eng, C.c_int, C.c_int]),
    ("mse_group_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    ("mse_group_multislice", C.c_int, [C.c_void_p]),
    ("mse_group_destroy", C.c_int, [C.c_void_p]),
    ("mse_time_sweeps", C.c_int, [_eng, C.c_int, C.c_int, _dp, _dp]),
    ("mse_set_focus_spread", C.c_int, [_eng, _dp, C.c_int]),
    ("mse_set_reference_offset", C.c_int, [_eng, C.c_int, C.c_double, C.c_double]),
    ("mse_forward_checkpointed", C.c_int, [_eng, C.c_int, _dp]),
    ("mse_adjoint_positions", C.c_int, [_eng, _dp, _dp]),
    ("mse_set_checkpoint_budget", C.c_int, [_eng, C.c_ulonglong]),
    ("mse_stem_batch", C.c_int, [_eng, C.c_int, _dp, C.c_int, C.c_int, C.POINTER(MseProbe),
                                 C.POINTER(MseDetector), C.c_int, C.c_int, C.c_int, _dp]),
    ("mse_stem_batch_frames", C.c_int, [_eng, C.c_int, _dp, _dp, C.c_int, C.c_int, C.POINTER(MseProbe),
                                        C.POINTER(MseDetector), C.c_int, C.c_int, _dp]),
    ("mse_set_potential_tables", C.c_int, [_eng, C.c_int]),
    ("mse_potential_table", C.c_int, [C.c_int, C.c_int, _dp, C.c_int, C.c_int, C.c_int, _fp]),
    ("mse_set_kernel_variant", C.c_int, [_eng, C.c_int, C.c_int]),
    ("mse_autotune", C.c_int, [_eng, C.c_int, C.c_int, C.c_int, _dp, _dp, _ip, _ip]),
    ("mse_bandwidth_probe", C.c_int, [_eng, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    ("mse_profile_enable", C.c_int, [_eng, C.c_int]),
    ("mse_profile_read", C.c_int, [_eng, C.c_int, _ip, _dp]),
    ("mse_profile_reset", C.c_int, [_eng]),
]

EXPORTED = [s[0] for s in SIGNATURES]

_lib = None


def load(path: str = LIB_PATH):
    """Load libcltem_amd.so.  There is no fallback: a missing library is an error."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"cltem_amd: native library {path} is not built; run __graft_entry__.build() "
            "or make -C cltem_amd/csrc")
    lib = C.CDLL(path)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib
