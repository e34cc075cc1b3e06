"""ctypes binding of libgemini_fk.so.  There is no fallback: if the HIP library
is missing or cannot be loaded, every entry point raises."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgemini_fk.so")
_lib = None

D = C.POINTER(C.c_double)
F = C.POINTER(C.c_float)
I = C.POINTER(C.c_int)
LL = C.POINTER(C.c_longlong)
VP = C.c_void_p

# flags of the batch entry points (gemini_fk.h)
GM_FLAG_DEVICE_PTRS = 1
GM_FLAG_STRESS_SPECTRA = 2
GM_FLAG_DERIV_SPECTRA = 4
GM_FLAG_OCEAN_SPECTRA = 8

# every symbol declared in include/gemini_fk.h
EXPORTS = ["gm_ctx_create", "gm_ctx_destroy", "gm_set_option", "gm_last_error", "gm_model_set",
           "gm_nodes_set", "gm_minors_batch", "gm_gfk_batch", "gm_last_timing", "gm_last_system_stats",
           "gm_mlp_nparams", "gm_mlp_forward", "gm_mlp_grad", "gm_mlp_train", "gm_set_proposer",
           "gm_collect_steplog", "gm_minors_batch_wn", "gm_gfk_batch_wn", "gm_legendre_batch", "gm_moment_sums", "gm_synthetics", "gm_last_synth_timing",
           "gm_bessel_sums", "gm_synthetics_regional", "gm_force_sums", "gm_force_bessel_sums",
           "gm_event_filter", "gm_travel_times", "gm_gfk_batch_noalloc", "gm_node_properties", "gm_aski_batch", "gm_strain_spectra",
           "gm_strain_spectra_regional", "gm_ujgf_batch", "gm_ujgf_batch_wn"]


class GeminiFKError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gemini_fk error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GeminiFKError(-1, f"{LIB_PATH} not built; run __graft_entry__.build() or python -m gemini_unified_amd.build")
        L = C.CDLL(LIB_PATH)
        L.gm_ctx_create.argtypes = [C.c_int, C.POINTER(VP)]
        L.gm_ctx_destroy.argtypes = [VP]
        L.gm_set_option.argtypes = [VP, C.c_int, C.c_int]
        L.gm_last_error.argtypes = [VP, C.c_char_p, C.c_int]
        L.gm_model_set.argtypes = [VP, C.c_int] + [D] * 9 + [C.c_int, I, D, I, C.c_int, C.c_int, C.c_int, C.c_int,
                                                              C.c_double, C.c_double, C.c_double, C.c_double]
        L.gm_nodes_set.argtypes = [VP, C.c_int, D, I, I]
        L.gm_node_properties.argtypes = [VP, C.c_double, C.c_int, D, D, D]
        L.gm_minors_batch.argtypes = [VP, C.c_int, D, I, I, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                      C.c_int, D, D, D, D, D, D, D, D, I, LL]
        L.gm_gfk_batch.argtypes = [VP, C.c_int, D, I, I, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                   C.c_int, D, F, F, I, LL]
        L.gm_last_timing.argtypes = [VP, D]
        L.gm_last_system_stats.argtypes = [VP, C.c_int, LL]
        L.gm_mlp_nparams.restype = C.c_int
        L.gm_mlp_forward.argtypes = [C.c_int, F, C.c_int, F, F]
        L.gm_mlp_grad.argtypes = [C.c_int, F, C.c_int, F, F, C.c_float, C.c_float, F, F]
        L.gm_mlp_train.argtypes = [C.c_int, F, C.c_int, F, F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_float,
                                   C.c_uint, F]
        L.gm_set_proposer.argtypes = [VP, F, F, C.c_float, C.c_int, C.c_int]
        L.gm_collect_steplog.argtypes = [VP, C.c_int, D, I, I, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                         C.c_int, C.c_int, F, I, LL]
        L.gm_minors_batch_wn.argtypes = [VP, C.c_int, D, I, I, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                         C.c_int, C.c_int, D, D, D, D, D, D, D, D, I, LL]
        L.gm_aski_batch.argtypes = [VP, C.c_int, D, I, I, C.c_double, C.c_double, C.c_int, C.c_int, I, C.c_int,
                                    C.c_int, C.c_int, D, F, F, I, LL]
        L.gm_gfk_batch_wn.argtypes = [VP, C.c_int, D, I, I, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_int, C.c_int, D, F, F, I, LL]
        L.gm_ujgf_batch.argtypes = [VP, C.c_int, D, I, I, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, D, D,
                                    I, LL]
        L.gm_ujgf_batch_wn.argtypes = [VP, C.c_int, D, I, I, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                       C.c_int, D, D, I, LL]
        # array arguments of the synthetics entry points are void* so that device pointers (ints) pass too
        L.gm_legendre_batch.argtypes = [VP, C.c_int, D, C.c_int, D, D, D, D, D]
        L.gm_moment_sums.argtypes = [VP, C.c_int, C.c_int, I, C.c_int, C.c_longlong, VP, VP, C.c_int, D, C.c_int, VP]
        L.gm_synthetics.argtypes = [VP, C.c_int, C.c_int, C.c_double, C.c_double, I, C.c_int, C.c_longlong, VP, VP,
                                    C.c_int, D, D, D, D, D, D, C.c_int, D, C.c_int, C.c_int, C.c_double, C.c_int,
                                    VP, VP]
        L.gm_last_synth_timing.argtypes = [VP, D]
        L.gm_bessel_sums.argtypes = [VP, C.c_int, C.c_int, I, C.c_int, C.c_longlong, VP, VP, C.c_double, C.c_double,
                                     C.c_int, D, C.c_int, VP]
        L.gm_force_sums.argtypes = L.gm_moment_sums.argtypes
        L.gm_strain_spectra.argtypes = [VP, C.c_int, C.c_int, I, C.c_int, C.c_longlong, VP, VP, C.c_int, D, D,
                                        C.c_double, C.c_int, D, D, C.c_int, VP, VP, VP, VP]
        L.gm_strain_spectra_regional.argtypes = [VP, C.c_int, C.c_int, I, C.c_int, C.c_longlong, VP, VP, C.c_int, D,
                                                 D, C.c_double, C.c_double, C.c_double, C.c_int, D, D, C.c_int, VP,
                                                 VP, VP, VP]
        L.gm_force_bessel_sums.argtypes = L.gm_bessel_sums.argtypes
        L.gm_synthetics_regional.argtypes = [VP, C.c_int, C.c_int, C.c_double, C.c_double, I, C.c_int, C.c_longlong,
                                             VP, VP, C.c_double, C.c_double, C.c_int, D, D, D, D, D, D, C.c_int, D,
                                             C.c_int, C.c_int, C.c_double, C.c_int, VP, VP]
        L.gm_event_filter.argtypes = [C.c_int, D, C.c_int, C.c_int, C.c_double, C.c_double, D]
        L.gm_gfk_batch_noalloc.argtypes = [VP, C.c_int, D, I, I, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                           VP, VP, VP, VP, LL]
        L.gm_travel_times.argtypes = [VP, C.c_int, C.c_int, C.c_int, F, F, F, F, F, C.c_double, C.c_int, D, D, D, D, I]
        _lib = L
    return _lib


def last_error(h) -> str:
    buf = C.create_string_buffer(512)
    lib().gm_last_error(h, buf, 512)
    return buf.value.decode(errors="replace")


def check(h, rc):
    if rc != 0:
        raise GeminiFKError(rc, last_error(h))
    return rc
