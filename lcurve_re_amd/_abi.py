"""ctypes mirror of include/lcurve_amd.h and the loader of the built library.

The structures here must match the C header field for field; the library
exports ``lc_sizeof`` so tests can check the layouts agree.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "liblcurve_amd.so")

# lc_status
LC_OK = 0
LC_E_ROCHE_OVERFLOW = 1
LC_E_FACE = 2
LC_E_ECLIPSE_SEARCH = 3
LC_E_PARAM = 4
LC_E_DEVICE = 5
LC_E_UNSUPPORTED = 6
LC_E_ARG = 7
LC_E_LENSING = 8

LC_PREC_MIXED = 0
LC_PREC_FP64 = 1
LC_LDC_POLY = 0
LC_LDC_CLARET = 1


class lc_pparam(C.Structure):
    _fields_ = [("value", C.c_double), ("range", C.c_double), ("dstep", C.c_double),
                ("vary", C.c_int32), ("defined", C.c_int32)]


# Physical parameters in lc_model order (lcurve.h:340-370).
PPARAM_NAMES = [
    "q", "iangle", "r1", "r2", "cphi3", "cphi4", "spin1", "spin2", "t1", "t2",
    "ldc1_1", "ldc1_2", "ldc1_3", "ldc1_4", "ldc2_1", "ldc2_2", "ldc2_3", "ldc2_4",
    "velocity_scale", "beam_factor1", "beam_factor2", "t0", "period", "pdot", "deltat",
    "gravity_dark1", "gravity_dark2", "absorb", "slope", "quad", "cube", "third",
    "rdisc1", "rdisc2", "height_disc", "beta_disc", "temp_disc", "texp_disc",
    "lin_limb_disc", "quad_limb_disc", "temp_edge", "absorb_edge",
    "radius_spot", "length_spot", "height_spot", "expon_spot", "epow_spot", "angle_spot",
    "yaw_spot", "temp_spot", "tilt_spot", "cfrac_spot",
    "stsp11_long", "stsp11_lat", "stsp11_fwhm", "stsp11_tcen",
    "stsp12_long", "stsp12_lat", "stsp12_fwhm", "stsp12_tcen",
    "stsp13_long", "stsp13_lat", "stsp13_fwhm", "stsp13_tcen",
    "stsp21_long", "stsp21_lat", "stsp21_fwhm", "stsp21_tcen",
    "stsp22_long", "stsp22_lat", "stsp22_fwhm", "stsp22_tcen",
    "uesp_long1", "uesp_long2", "uesp_lathw", "uesp_taper", "uesp_temp",
]
COMP_DOUBLE = ["delta_phase", "lfudge", "llo", "lhi", "phase1", "phase2", "wavelength",
               "tperiod", "mucrit1", "mucrit2"]
COMP_INT = ["nlat1f", "nlat2f", "nlat1c", "nlat2c", "npole", "nlatfill", "nlngfill",
            "roche1", "roche2", "eclipse1", "eclipse2", "glens1", "use_radii",
            "gdark_bolom1", "gdark_bolom2", "limb1", "limb2", "mirror",
            "add_disc", "nrad", "opaque", "add_spot", "nspot", "iscale"]


class lc_model(C.Structure):
    _fields_ = ([(n, lc_pparam) for n in PPARAM_NAMES]
                + [(n, C.c_double) for n in COMP_DOUBLE]
                + [(n, C.c_int32) for n in COMP_INT]
                + [("reserved_", C.c_int32)])


class lc_datum(C.Structure):
    _fields_ = [("time", C.c_double), ("expose", C.c_double), ("flux", C.c_double),
                ("ferr", C.c_double), ("weight", C.c_double), ("ndiv", C.c_int32),
                ("reserved_", C.c_int32)]


class lc_scalars(C.Structure):
    _fields_ = [("sfac", C.c_double * 5), ("wdwarf", C.c_double), ("chisq", C.c_double),
                ("wnok", C.c_double), ("logg1", C.c_double), ("logg2", C.c_double),
                ("rv1", C.c_double), ("rv2", C.c_double)]


class lc_eval_opts(C.Structure):
    _fields_ = [("scale", C.c_int32), ("rdata", C.c_int32), ("info", C.c_int32),
                ("diagnostics", C.c_int32), ("max_model_points", C.c_int32),
                ("precision", C.c_int32)]


class lc_ldc(C.Structure):
    _fields_ = [("c1", C.c_double), ("c2", C.c_double), ("c3", C.c_double), ("c4", C.c_double),
                ("mucrit", C.c_double), ("type", C.c_int32), ("reserved_", C.c_int32)]


_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)


class lc_flat_grid(C.Structure):
    _fields_ = [("n", C.c_size_t), ("n0", C.c_size_t), ("n1", C.c_size_t),
                ("dx", _DP), ("dy", _DP), ("dz", _DP), ("px", _DP), ("py", _DP), ("pz", _DP),
                ("flux", _DP), ("in1", _DP), ("out1", _DP), ("moff", _IP),
                ("min_", _DP), ("mout_", _DP)]


class lc_phase_batch(C.Structure):
    _fields_ = [("n", C.c_size_t), ("ex", _DP), ("ey", _DP), ("ez", _DP), ("phin", _DP)]


class lc_star_faces(C.Structure):
    _fields_ = [("n", C.c_size_t), ("px", _DP), ("py", _DP), ("pz", _DP),
                ("dx", _DP), ("dy", _DP), ("dz", _DP),
                ("area", _FP), ("gravity", _FP), ("flux", _FP),
                ("necl", _IP), ("ingress", _DP), ("egress", _DP)]


class lc_profile(C.Structure):
    _fields_ = [("flux_ms_total", C.c_double), ("grid_ms_total", C.c_double),
                ("eclipse_ms_total", C.c_double), ("cont_ms_total", C.c_double),
                ("other_ms_total", C.c_double), ("flux_launches", C.c_int64),
                ("evaluations", C.c_int64), ("flux_elements", C.c_double),
                ("flux_flops", C.c_double), ("flux_bytes", C.c_double),
                ("host_plan_ms_total", C.c_double), ("call_ms_total", C.c_double),
                ("eclipse_candidates", C.c_double), ("eclipse_hits", C.c_double),
                ("flux_var_ms", C.c_double * 16), ("flux_var_flops", C.c_double * 16),
                ("flux_var_launches", C.c_int64 * 16), ("flux_var_evaluated", C.c_double * 16),
                ("grids_computed", C.c_int64), ("grids_reused", C.c_int64),
                ("graph_launches", C.c_int64), ("graph_captures", C.c_int64), ("graph_ms_total", C.c_double),
                ("last_rphase", C.c_int64), ("last_chunk", C.c_int64), ("continua_reused", C.c_int64),
                ("eclipse_bisected", C.c_double)]


# Every entry point include/lcurve_amd.h declares, with its signature.
_c = C
SIGNATURES = {
    "lc_ctx_create": (_c.c_int, [_c.c_int, _c.POINTER(_c.c_void_p)]),
    "lc_ctx_destroy": (None, [_c.c_void_p]),
    "lc_last_error": (_c.c_char_p, [_c.c_void_p]),
    "lc_ctx_get_profile": (_c.c_int, [_c.c_void_p, _c.POINTER(lc_profile)]),
    "lc_ctx_reset_profile": (_c.c_int, [_c.c_void_p]),
    "lc_ctx_set_stage_cache": (_c.c_int, [_c.c_void_p, _c.c_size_t]),
    "lc_ctx_set_graphs": (_c.c_int, [_c.c_void_p, _c.c_int]),
    "lc_ctx_autotune": (_c.c_int, [_c.c_void_p, _c.POINTER(lc_model), _c.c_size_t, _c.POINTER(lc_datum), _c.c_size_t,
                                   _c.POINTER(lc_eval_opts), _IP, _IP]),
    "lc_version": (_c.c_char_p, []),
    "lc_thread_error": (_c.c_char_p, []),
    "lc_sizeof": (_c.c_size_t, [_c.c_char_p]),
    "lc_model_from_json": (_c.c_int, [_c.c_char_p, _c.c_size_t, _c.POINTER(lc_model)]),
    "lc_model_nvary": (_c.c_int, [_c.POINTER(lc_model)]),
    "lc_model_set_param": (_c.c_int, [_c.POINTER(lc_model), _DP, _c.c_size_t]),
    "lc_model_get_param": (_c.c_int, [_c.POINTER(lc_model), _DP, _c.c_size_t]),
    "lc_model_get_range": (_c.c_int, [_c.POINTER(lc_model), _DP, _c.c_size_t]),
    "lc_model_get_dstep": (_c.c_int, [_c.POINTER(lc_model), _DP, _c.c_size_t]),
    "lc_model_get_limit": (_c.c_int, [_c.POINTER(lc_model), _DP, _DP, _c.c_size_t]),
    "lc_model_is_not_legal": (_c.c_int, [_c.POINTER(lc_model), _DP, _c.c_size_t]),
    "lc_model_is_not_legal_batch": (_c.c_int, [_c.POINTER(lc_model), _DP, _c.c_size_t, _c.c_size_t, _IP]),
    "lc_model_get_name": (_c.c_int, [_c.POINTER(lc_model), _c.c_int, _c.c_char_p, _c.c_size_t]),
    "lc_star_grid_size": (_c.c_int, [_c.POINTER(lc_model), _c.c_int, _c.c_int, _c.POINTER(_c.c_size_t)]),
    "lc_light_curve_comp": (_c.c_int, [_c.c_void_p, _c.POINTER(lc_model), _c.POINTER(lc_datum), _c.c_size_t,
                                       _c.POINTER(lc_eval_opts), _DP, _c.POINTER(lc_scalars)]),
    "lc_light_curve_comp_batch": (_c.c_int, [_c.c_void_p, _c.POINTER(lc_model), _c.c_size_t,
                                             _c.POINTER(lc_datum), _c.c_size_t, _c.POINTER(lc_eval_opts), _DP,
                                             _c.POINTER(lc_scalars), _IP]),
    "lc_light_curve_comp_batch_dev": (_c.c_int, [_c.c_void_p, _c.POINTER(lc_model), _c.c_size_t,
                                                 _c.POINTER(lc_datum), _c.c_size_t, _c.POINTER(lc_eval_opts),
                                                 _c.c_void_p, _c.POINTER(lc_scalars), _IP]),
    "lc_flat_sum_star1_multi": (_c.c_int, [_c.c_void_p, _c.POINTER(lc_flat_grid), _c.POINTER(lc_phase_batch),
                                           _c.POINTER(lc_ldc), _c.c_double, _c.c_double, _c.c_double,
                                           _c.c_double, _c.c_int, _DP]),
    "lc_flat_sum_star2_multi": (_c.c_int, [_c.c_void_p, _c.POINTER(lc_flat_grid), _c.POINTER(lc_phase_batch),
                                           _c.POINTER(lc_ldc), _c.c_double, _c.c_double, _c.c_double,
                                           _c.c_double, _c.c_int, _c.c_double, _c.c_int, _DP]),
    "lc_flat_sum_multi_noalloc": (_c.c_int, [_c.c_void_p, _c.c_int, _c.POINTER(lc_flat_grid),
                                             _c.POINTER(lc_phase_batch), _c.POINTER(lc_ldc), _c.c_double,
                                             _c.c_double, _c.c_double, _c.c_double, _c.c_int, _c.c_double,
                                             _c.c_int, _c.c_void_p]),
    "lc_set_star_grid": (_c.c_int, [_c.c_void_p, _c.POINTER(lc_model), _c.c_int, _c.c_int,
                                    _c.POINTER(lc_star_faces)]),
    "lc_set_star_continuum": (_c.c_int, [_c.c_void_p, _c.POINTER(lc_model), _c.POINTER(lc_star_faces),
                                         _c.POINTER(lc_star_faces)]),
    "lc_mlp_create": (_c.c_int, [_c.c_int, _IP, _c.c_int, _c.c_int, _c.POINTER(_c.c_void_p)]),
    "lc_mlp_destroy": (None, [_c.c_void_p]),
    "lc_mlp_last_error": (_c.c_char_p, [_c.c_void_p]),
    "lc_mlp_num_params": (_c.c_size_t, [_c.c_void_p]),
    "lc_mlp_set_params": (_c.c_int, [_c.c_void_p, _FP, _c.c_size_t]),
    "lc_mlp_get_params": (_c.c_int, [_c.c_void_p, _FP, _c.c_size_t]),
    "lc_mlp_forward": (_c.c_int, [_c.c_void_p, _FP, _c.c_size_t, _FP]),
    "lc_mlp_grad": (_c.c_int, [_c.c_void_p, _FP, _FP, _c.c_size_t, _FP, _DP]),
    "lc_mlp_set_dataset": (_c.c_int, [_c.c_void_p, _c.c_size_t]),
    "lc_mlp_load_dataset": (_c.c_int, [_c.c_void_p, _c.c_size_t, _FP, _FP, _c.c_size_t]),
    "lc_mlp_train": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_float, _c.c_uint64, _DP]),
    "lc_mlp_eval_dataset": (_c.c_int, [_c.c_void_p, _c.c_size_t, _c.c_size_t, _DP]),
    "lc_mlp_stats": (_c.c_int, [_c.c_void_p, _DP, _DP]),
    "lc_mlp_set_likelihood": (_c.c_int, [_c.c_void_p, _DP, _DP, _DP, _c.c_size_t, _FP, _c.c_float, _c.c_double,
                                         _c.c_int]),
    "lc_mlp_chisq": (_c.c_int, [_c.c_void_p, _FP, _DP, _c.c_size_t, _DP]),
    "lc_mlp_set_skip": (_c.c_int, [_c.c_void_p, _FP, _c.c_size_t]),
}

_lib = None


def load(path=None):
    """Loads the built library (raises if it has not been built); the
    environment variable LCURVE_AMD_LIB names another build to load."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("LCURVE_AMD_LIB") or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"lcurve_re_amd: native library not built ({p}); run __graft_entry__.build()")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def struct_sizes():
    return {"lc_pparam": lc_pparam, "lc_model": lc_model, "lc_datum": lc_datum,
            "lc_scalars": lc_scalars, "lc_eval_opts": lc_eval_opts, "lc_ldc": lc_ldc,
            "lc_flat_grid": lc_flat_grid, "lc_phase_batch": lc_phase_batch,
            "lc_star_faces": lc_star_faces, "lc_profile": lc_profile}
