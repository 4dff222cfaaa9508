"""ctypes declarations of include/rkx.h and the loader of rk_amd/librkx.so."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_FILE = os.path.join(HERE, "librkx.so")

RKX_MAX_STAGES = 16
RKX_BRUSS2D, RKX_LV, RKX_LV_CHAIN = 1, 2, 3
RKX_IMPL_D, RKX_IMPL_PIPED, RKX_IMPL_PIPEDLS = 0, 1, 2
RKX_IMPL_A, RKX_IMPL_E, RKX_IMPL_F = 3, 4, 5
RKX_CTRL_DEVICE, RKX_CTRL_HOST = 0, 1

ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "ELIMIT"}


class RkxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("rkx error %d (%s): %s" % (code, ERRORS.get(code, "?"), msg))
        self.code = code


class Ode(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int), ("grid", C.c_int), ("alpha", C.c_double),
                ("c0", C.c_double), ("d0", C.c_double), ("d1", C.c_double), ("c1", C.c_double),
                ("diff", C.c_double)]


class Method(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("s", C.c_int), ("ord", C.c_int), ("has_emb", C.c_int),
                ("A", (C.c_double * RKX_MAX_STAGES) * RKX_MAX_STAGES),
                ("b", C.c_double * RKX_MAX_STAGES), ("b_hat", C.c_double * RKX_MAX_STAGES),
                ("c", C.c_double * RKX_MAX_STAGES)]


class Opts(C.Structure):
    _fields_ = [("h0", C.c_double), ("step_limit", C.c_int), ("impl", C.c_int),
                ("ctrl", C.c_int), ("tile_rows", C.c_int), ("tile_waves", C.c_int),
                ("trace", C.POINTER(C.c_double)), ("trace_cap", C.c_int),
                ("stream", C.c_void_p), ("dump_path", C.c_char_p)]


class Stats(C.Structure):
    _fields_ = [("steps_acc", C.c_int), ("steps_rej", C.c_int), ("t", C.c_double),
                ("h", C.c_double), ("kernel_time", C.c_double), ("resident_bytes", C.c_size_t)]

    @property
    def steps(self):
        return self.steps_acc + self.steps_rej

    def as_dict(self):
        return dict(acc=self.steps_acc, rej=self.steps_rej, steps=self.steps, t=self.t, h=self.h,
                    kernel_time=self.kernel_time, resident_bytes=self.resident_bytes)


# (name, restype, argtypes) of every exported symbol of include/rkx.h
_dp = C.POINTER(C.c_double)
SIGNATURES = [
    ("rkx_ode_bruss2d", C.c_int, [C.POINTER(Ode), C.c_int, C.c_double]),
    ("rkx_ode_lv", C.c_int, [C.POINTER(Ode), C.c_int] + [C.c_double] * 4),
    ("rkx_ode_lv_chain", C.c_int, [C.POINTER(Ode), C.c_int] + [C.c_double] * 5),
    ("rkx_ode_acc_dist", C.c_int, [C.POINTER(Ode)]),
    ("rkx_ode_start", C.c_int, [C.POINTER(Ode), C.c_double, _dp]),
    ("rkx_ode_start_seeded", C.c_int, [C.POINTER(Ode), C.c_ulonglong, C.c_int, C.c_double, _dp]),
    ("rkx_eval_rng_device", C.c_int, [C.POINTER(Ode), C.c_int, C.c_int, C.c_double, C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
    ("rkx_eval_rng_host", C.c_int, [C.POINTER(Ode), C.c_int, C.c_int, C.c_double, _dp, _dp]),
    ("rkx_eval_comp_host", C.c_int, [C.POINTER(Ode), C.c_int, C.c_double, _dp, _dp]),
    ("rkx_method_get", C.c_int, [C.c_char_p, C.POINTER(Method)]),
    ("rkx_opts_default", None, [C.POINTER(Opts)]),
    ("rkx_solve", C.c_int, [C.POINTER(Ode), C.POINTER(Method), C.c_double, C.c_double,
                            C.c_void_p, C.c_void_p, C.c_double, C.POINTER(Opts),
                            C.POINTER(Stats)]),
    ("rkx_solve_host", C.c_int, [C.POINTER(Ode), C.POINTER(Method), C.c_double, C.c_double,
                                 _dp, _dp, C.c_double, C.POINTER(Opts), C.POINTER(Stats)]),
    ("rkx_batch_opts_default", None, [C.c_void_p]),
    ("rkx_batch_create", C.c_int, [C.POINTER(C.c_void_p), C.c_void_p]),
    ("rkx_batch_destroy", C.c_int, [C.c_void_p]),
    ("rkx_batch_footprint", C.c_size_t, [C.POINTER(Ode), C.POINTER(Method), C.c_int]),
    ("rkx_batch_add", C.c_int, [C.c_void_p, C.POINTER(Ode), C.POINTER(Method), C.c_double,
                                C.c_double, _dp, C.c_double, C.c_double, C.c_int, C.c_int]),
    ("rkx_batch_run", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    ("rkx_batch_status", C.c_int, [C.c_void_p, C.c_int, C.POINTER(Stats), C.POINTER(C.c_int)]),
    ("rkx_batch_result", C.c_int, [C.c_void_p, C.c_int, _dp]),
    ("rkx_batch_remove", C.c_int, [C.c_void_p, C.c_int]),
    ("rkx_batch_bytes", C.c_size_t, [C.c_void_p]),
    ("rkx_batch_tile", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("rkx_lv_ensemble", C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                  C.c_double, C.POINTER(Method), C.c_double, C.c_double, C.c_int,
                                  C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rkx_solve_noalloc", C.c_int, [C.POINTER(Ode), C.POINTER(Method), C.c_double, C.c_double,
                                    C.c_void_p, C.c_void_p, C.c_double, C.POINTER(Opts),
                                    C.POINTER(Stats)]),
    ("rkx_pow_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    ("rkx_pow_host", C.c_int, [_dp, _dp, _dp, C.c_size_t]),
    ("rkx_last_error", C.c_char_p, []),
    ("rkx_device_sync", C.c_int, []),
    ("rkx_solve_footprint", C.c_size_t, [C.POINTER(Ode), C.POINTER(Method), C.c_int]),
]

_lib = None


def lib_path():
    return _LIB_FILE


def lib():
    """Load librkx.so; raises if it has not been built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_FILE):
            raise RkxError(-3, "rk_amd/librkx.so is not built: run `python rk_amd/build.py` "
                               "(or __graft_entry__.build())")
        L = C.CDLL(_LIB_FILE)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RkxError(rc, lib().rkx_last_error().decode(errors="replace"))
    return rc
