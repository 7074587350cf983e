"""ctypes binding of libmnpbem_hip.so (declarations in include/mnpbem_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmnpbem_hip.so")

MNP_OK = 0
MNP_ERR_ARG, MNP_ERR_HIP, MNP_ERR_SINGULAR, MNP_ERR_NOCONV, MNP_ERR_INTERNAL, MNP_ERR_STATE = 1, 2, 3, 4, 5, 6
MNP_FP64, MNP_FP32, MNP_FP32_SPLIT_BF16 = 0, 1, 2
MNP_KEY = {"G": 0, "F": 1, "H1": 2, "H2": 3}
MNP_SOLVER_FP64, MNP_SOLVER_MIXED, MNP_SOLVER_MIXED_BF16 = 0, 1, 2


class MnpComplex(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]

    @classmethod
    def of(cls, z):
        z = complex(z)
        return cls(z.real, z.imag)


class GreenOptions(C.Structure):
    _fields_ = [("order", C.c_int), ("rel_cutoff", C.c_double), ("abs_cutoff", C.c_double),
                ("npol_rho", C.c_int), ("npol_phi", C.c_int), ("closed", C.c_int)]


class LayerOptions(C.Structure):
    _fields_ = [("z0", C.c_double), ("ztol", C.c_double), ("rmin", C.c_double), ("zmin", C.c_double),
                ("semi", C.c_double), ("ratio", C.c_double)]


class Timers(C.Structure):
    _fields_ = [("launches", C.c_int64 * 7), ("ms", C.c_double * 7), ("work", C.c_double * 7),
                ("bytes", C.c_double * 7)]


P, I64, D, INT = C.c_void_p, C.c_int64, C.c_double, C.c_int
PD = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes); the list is the exported C ABI
PROTOTYPES = {
    "mnp_last_error": (C.c_char_p, []),
    "mnp_version": (INT, []),
    "mnp_set_device": (INT, [INT]),
    "mnp_set_stream": (INT, [P]),
    "mnp_get_stream": (P, []),
    "mnp_synchronize": (INT, []),
    "mnp_timers_enable": (INT, [INT]),
    "mnp_timers_read": (INT, [C.POINTER(Timers)]),
    "mnp_green_options_default": (None, [C.POINTER(GreenOptions)]),
    "mnp_greenret_create": (INT, [P, I64, P, I64, C.POINTER(GreenOptions), C.POINTER(P)]),
    "mnp_greenret_create_like": (INT, [P, I64, P, I64, C.POINTER(GreenOptions), P, C.POINTER(P)]),
    "mnp_greenret_create_compound": (INT, [P, I64, P, I64, C.c_int32, P, P, P, C.POINTER(GreenOptions),
                                           C.POINTER(P)]),
    "mnp_greenret_destroy": (INT, [P]),
    "mnp_greenret_size": (INT, [P, PI64, PI64]),
    "mnp_greenret_geometry": (INT, [P, P, P, P, P, P]),
    "mnp_greenret_refined": (INT, [P, P, P, P, P]),
    "mnp_greenret_eval": (INT, [P, MnpComplex, INT, INT, P, I64]),
    "mnp_bemret_create": (INT, [P, INT, C.POINTER(P)]),
    "mnp_bemret_destroy": (INT, [P]),
    "mnp_bemret_initmat": (INT, [P, D, MnpComplex, MnpComplex]),
    "mnp_bemret_solve": (INT, [P, I64, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mnp_bemret_tangent_eps1": (INT, [P, I64, P, P, P, P, P, P, P, P]),
    "mnp_bemret_solve_transposed": (INT, [P, I64, P, P]),
    "mnp_bemret_extinction_gradient": (INT, [P, I64, P, P, P, P, P, P, P, P, P]),
    "mnp_bemret_batch_create": (INT, [I64, I64, INT, I64, I64, C.POINTER(P)]),
    "mnp_bemret_batch_destroy": (INT, [P]),
    "mnp_bemret_batch_set_options": (INT, [P, INT, INT, INT]),
    "mnp_bemret_batch_cross_sections": (INT, [P, I64, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mnp_bemret_batch_info": (INT, [P, C.POINTER(INT), P, PI64]),
    "mnp_bemret_initmat_media": (INT, [P, D, C.c_int32, P]),
    "mnp_bemret_initmat_many": (INT, [INT, C.POINTER(P), PD, C.POINTER(MnpComplex), C.POINTER(MnpComplex)]),
    "mnp_bemret_stats": (INT, [P, C.POINTER(INT), PD]),
    "mnp_bemret_set_refinement": (INT, [P, D, INT]),
    "mnp_gemm": (INT, [INT, I64, I64, I64, MnpComplex, P, I64, P, I64, MnpComplex, P, I64]),
    "mnp_getrf": (INT, [INT, I64, I64, P, I64, P, C.POINTER(INT)]),
    "mnp_getrf_nopiv": (INT, [INT, I64, I64, P, I64, C.POINTER(INT), C.POINTER(INT)]),
    "mnp_getrs": (INT, [INT, I64, I64, P, I64, P, P, I64]),
    "mnp_getrs_right": (INT, [INT, I64, I64, P, I64, P, P, I64]),
    "mnp_getri": (INT, [INT, I64, P, I64, P, P, I64]),
    "mnp_hmat_create": (INT, [P, INT, P, I64, P, I64, P, D, INT, C.POINTER(P)]),
    "mnp_hmat_destroy": (INT, [P]),
    "mnp_hmat_fill": (INT, [P, MnpComplex, INT]),
    "mnp_hmat_fill_layer": (INT, [P, MnpComplex, INT, INT, INT, P, D, MnpComplex, MnpComplex]),
    "mnp_hmat_matvec": (INT, [P, I64, P, P]),
    "mnp_hmat_stats": (INT, [P, PI64, PI64, C.POINTER(INT), PD]),
    "mnp_hmat_ranks": (INT, [P, P]),
    "mnp_hmat_coarsen": (INT, [P, PI64, P, P]),
    "mnp_bemretiter_create": (INT, [P, INT, P, I64, P, I64, P, D, INT, C.POINTER(P)]),
    "mnp_bemretiter_destroy": (INT, [P]),
    "mnp_bemretiter_set_coarse": (INT, [P, INT]),
    "mnp_bemretiter_set_options": (INT, [P, D, INT, INT]),
    "mnp_bemretiter_initmat": (INT, [P, D, MnpComplex, MnpComplex]),
    "mnp_bemretiter_create_layer": (INT, [P, P, INT, P, I64, P, I64, P, D, INT, C.POINTER(P)]),
    "mnp_bemretiter_initmat_layer": (INT, [P, D, MnpComplex, MnpComplex, MnpComplex]),
    "mnp_bemretiter_solve": (INT, [P, I64, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mnp_bemretiter_stats": (INT, [P, C.POINTER(INT), PD, C.POINTER(INT), PD, PD]),
    "mnp_bemstatlayer_create": (INT, [P, P, INT, C.POINTER(P)]),
    "mnp_bemstatlayer_initmat": (INT, [P, D, MnpComplex, MnpComplex, MnpComplex, MnpComplex]),
    "mnp_planewavestatlayer_potential": (INT, [P, I64, P, P]),
    "mnp_layer_options_default": (None, [C.POINTER(LayerOptions)]),
    "mnp_layertab_create": (INT, [C.POINTER(LayerOptions), I64, P, I64, P, D, C.POINTER(P)]),
    "mnp_layertab_destroy": (INT, [P]),
    "mnp_layertab_set": (INT, [P, I64, P, P, P]),
    "mnp_layertab_eval": (INT, [P, D, MnpComplex, MnpComplex, P]),
    "mnp_greenretlayer_create": (INT, [P, P, C.POINTER(P)]),
    "mnp_greenretlayer_destroy": (INT, [P]),
    "mnp_greenretlayer_refined": (INT, [P, PI64, PI64]),
    "mnp_greenretlayer_eval": (INT, [P, D, MnpComplex, MnpComplex, P, P]),
    "mnp_bemretlayer_create": (INT, [P, P, INT, C.POINTER(P)]),
    "mnp_bemretlayer_destroy": (INT, [P]),
    "mnp_bemretlayer_initmat": (INT, [P, D, MnpComplex, MnpComplex, MnpComplex]),
    "mnp_bemretlayer_solve": (INT, [P, I64, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mnp_bemretlayer_stats": (INT, [P, C.POINTER(INT), C.POINTER(INT), C.POINTER(D)]),
    "mnp_pointlayer_create": (INT, [P, P, I64, P, C.POINTER(P)]),
    "mnp_pointlayer_destroy": (INT, [P]),
    "mnp_pointlayer_fields": (INT, [P, D, MnpComplex, MnpComplex, I64, P, P, P, P, P, P]),
    "mnp_planewaveretlayer_potential": (INT, [P, C.POINTER(LayerOptions), D, MnpComplex, MnpComplex, I64, P, P,
                                              P, P, P, P]),
    "mnp_dipolelayer_potential": (INT, [P, P, D, MnpComplex, MnpComplex, I64, P, I64, P, P, P, P, P]),
    "mnp_dipolelayer_field": (INT, [P, D, MnpComplex, MnpComplex, I64, P, I64, P, I64, P, P, P]),
    "mnp_dipolelayer_farfield": (INT, [C.POINTER(LayerOptions), D, MnpComplex, MnpComplex, I64, P, I64, P, I64, P, P, P]),
    "mnp_farfield_layer": (INT, [P, C.POINTER(LayerOptions), D, MnpComplex, MnpComplex, I64, P, I64, P, P, P,
                                 P]),
    "mnp_stat_dipole_moment": (INT, [P, I64, P, P]),
    "mnp_bemretmirror_create": (INT, [P, INT, INT, C.POINTER(INT), C.POINTER(P)]),
    "mnp_bemretmirror_destroy": (INT, [P]),
    "mnp_bemretmirror_initmat": (INT, [P, D, MnpComplex, MnpComplex]),
    "mnp_bemretmirror_solve": (INT, [P, INT, INT, INT, I64] + [P] * 12),
    "mnp_bemretmirror_stats": (INT, [P, C.POINTER(INT), PD]),
    "mnp_bemretmirror_set_refinement": (INT, [P, D, INT]),
    "mnp_dipolestat_potential": (INT, [P, I64, P, I64, P, P, D, P, P]),
    "mnp_stat_field_own": (INT, [P, I64, P, I64, P, P]),
    "mnp_bemstatmirror_create": (INT, [P, INT, INT, C.POINTER(INT), C.POINTER(P)]),
    "mnp_bemstatmirror_destroy": (INT, [P]),
    "mnp_bemstatmirror_initmat": (INT, [P, D, MnpComplex, MnpComplex]),
    "mnp_bemstatmirror_solve": (INT, [P, INT, I64, P, P]),
    "mnp_bemstatmirror_stats": (INT, [P, INT, C.POINTER(INT), PD]),
    "mnp_bemstatmirror_set_refinement": (INT, [P, D, INT]),
    "mnp_mirror_full": (INT, [I64, INT, INT, I64, I64, C.POINTER(INT), C.POINTER(INT), PD, C.POINTER(INT), P, P, P]),
    "mnp_bemstatiter_create": (INT, [P, INT, P, I64, P, I64, P, D, INT, C.POINTER(P)]),
    "mnp_bemstatiter_destroy": (INT, [P]),
    "mnp_bemstatiter_set_options": (INT, [P, D, INT, INT]),
    "mnp_bemstatiter_initmat": (INT, [P, D, MnpComplex, MnpComplex]),
    "mnp_bemstatiter_solve": (INT, [P, I64, P, P]),
    "mnp_bemstatiter_stats": (INT, [P, C.POINTER(INT), PD, C.POINTER(INT), PD, PD]),
    "mnp_bemstat_create": (INT, [P, INT, C.POINTER(P)]),
    "mnp_bemstat_destroy": (INT, [P]),
    "mnp_bemstat_initmat": (INT, [P, D, MnpComplex, MnpComplex]),
    "mnp_bemstat_solve": (INT, [P, I64, P, P]),
    "mnp_bemstat_stats": (INT, [P, C.POINTER(INT), C.POINTER(D)]),
    "mnp_planewavestat_potential": (INT, [P, I64, P, P]),
    "mnp_planewavestat_cross_sections": (INT, [P, D, MnpComplex, I64, P, P, P, P, P]),
    "mnp_dipole_potential": (INT, [P, D, MnpComplex, I64, P, I64, P, P, P, P, P]),
    "mnp_field_points": (INT, [P, D, MnpComplex, I64, P, I64, P, P, P, P]),
    "mnp_potential_points": (INT, [P, D, MnpComplex, I64, P, I64, P, P, P, P]),
    "mnp_dipole_decayrate": (INT, [P, D, MnpComplex, I64, P, I64, P, I64, P, P, P, P, P, P, P]),
    "mnp_planewave_potential": (INT, [P, D, MnpComplex, I64, P, P, P, P]),
    "mnp_farfield": (INT, [P, D, MnpComplex, I64, P, I64, P, P, P, P]),
    "mnp_planewave_cross_sections": (INT, [P, D, MnpComplex, I64, P, P, I64, P, P, P, P,
                                           P, P, P]),
    "mnp_planewave_cross_sections_noalloc": (INT, [P, D, MnpComplex, I64, P, P, I64, P, P, P, P,
                                                   P, P, P]),
}


class MnpError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded library; raises if libmnpbem_hip.so is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MnpError(f"{LIB_PATH} not built: run `python -m mnpbem_amd.build` "
                           "(there is no CPU fallback)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(h, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def can_release():
    """True when a handle may be destroyed now: the library is loaded and the
    interpreter is not shutting down.  At shutdown the handles are left to the
    process exit (device memory goes with the process), so that no hipFree
    runs while the HIP runtime or an attached profiler is being torn down."""
    return _lib is not None and not sys.is_finalizing()


def check(status, what=""):
    if status != MNP_OK:
        msg = lib().mnp_last_error().decode(errors="replace")
        raise MnpError(f"{what} failed with status {status}: {msg}")


def call(name, *args):
    fn = getattr(lib(), name)
    check(fn(*args), name)
