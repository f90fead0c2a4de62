"""ctypes binding of the C-ABI in include/nlm.h (libnlm_hip.so).

This is the binding a Python caller without pybind11 would use; the test suite
uses it to check the exported symbols and to drive the engine directly.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnlm_hip.so")

# every entry point declared in include/nlm.h
SYMBOLS = [
    "nlm_model_ncoef", "nlm_plan_create", "nlm_plan_destroy", "nlm_plan_set_stream", "nlm_plan_stream",
    "nlm_plan_set_coef", "nlm_pump_linear", "nlm_pump_chi3", "nlm_pump_upload", "nlm_pump_download",
    "nlm_propagate_grid", "nlm_grid_download", "nlm_batch", "nlm_greens", "nlm_greens_download",
    "nlm_batch_dev", "nlm_plan_set_step_table", "nlm_plan_enable_timing", "nlm_plan_last_timing", "nlm_greens_compose", "nlm_fft",
    "nlm_plan_sync", "nlm_last_error", "nlm_version", "nlm_greens_forward", "nlm_greens_adjoint",
    "nlm_multi_create", "nlm_multi_launch", "nlm_multi_sync", "nlm_multi_output", "nlm_multi_classes",
    "nlm_multi_grid", "nlm_multi_destroy", "nlm_batch_forward", "nlm_batch_adjoint",
    "nlm_greens_compose_dev", "nlm_greens_cascade", "nlm_pump_import",
    "nlm_batch_grids", "nlm_gaussian_covariance", "nlm_gram", "nlm_gemm_nt", "nlm_greens_noalloc",
    "nlm_fft_schedule",
]

_lib = None


class PlanDesc(C.Structure):
    _fields_ = [("model", C.c_int32), ("n_freqs", C.c_int32), ("n_zsteps", C.c_int32),
                ("disp_step", C.c_void_p), ("half_in", C.c_void_p), ("half_out", C.c_void_p),
                ("step_coef", C.c_void_p), ("const_coef", C.c_void_p), ("raman", C.c_void_p),
                ("gnlse_filter", C.c_void_p)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libnlm_hip.so is not built; run python -m multimodenonlinearoptics_amd.build")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.nlm_last_error.restype = C.c_char_p
        L.nlm_version.restype = C.c_char_p
        L.nlm_fft.argtypes = [C.c_int, C.c_int, vp, vp, C.c_int]
        L.nlm_model_ncoef.argtypes = [C.c_int, ip, ip, ip, ip]
        L.nlm_plan_create.argtypes = [C.POINTER(PlanDesc), C.POINTER(vp)]
        L.nlm_plan_destroy.argtypes = [vp]
        L.nlm_plan_stream.argtypes = [vp]
        L.nlm_plan_stream.restype = vp
        L.nlm_plan_set_stream.argtypes = [vp, vp]
        L.nlm_plan_enable_timing.argtypes = [vp, C.c_int]
        L.nlm_plan_set_step_table.argtypes = [vp, C.c_int]
        L.nlm_plan_last_timing.argtypes = [vp, dp]
        L.nlm_plan_sync.argtypes = [vp]
        L.nlm_pump_linear.argtypes = [vp, C.c_int, vp, vp, C.c_double, vp]
        L.nlm_greens.argtypes = [vp, C.c_int, C.c_int, ip, C.c_int, ip, vp, vp, vp]
        L.nlm_greens_download.argtypes = [vp, vp, vp]
        L.nlm_greens_noalloc.argtypes = [vp, C.c_int, C.c_int, ip, C.c_int, ip, vp, vp]
        L.nlm_batch.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ip, vp]
        L.nlm_batch_dev.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ip, vp]
        L.nlm_greens_compose.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp]
        L.nlm_gaussian_covariance.argtypes = [C.c_int, vp, vp, vp, vp]
        L.nlm_gram.argtypes = [C.c_int, C.c_int, vp, vp]
        L.nlm_gemm_nt.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp]
        L.nlm_greens_compose_dev.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def check(rc):
    if rc == 0:
        return
    msg = lib().nlm_last_error().decode()
    if rc == 1:
        raise ValueError(msg)
    raise RuntimeError(msg)


def fft(x, inverse=False):
    """Device transform of the rows of x (complex128), reference conventions."""
    import numpy as np
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    n = x.shape[-1]
    out = np.empty_like(x)
    check(lib().nlm_fft(n, x.size // n, x.ctypes.data, out.ctypes.data, int(inverse)))
    return out


_hip = None


def hip():
    """The HIP runtime libnlm_hip.so is linked against (device buffers, streams
    and events for callers of the *_dev entry points that hold no framework
    tensors)."""
    global _hip
    if _hip is None:
        lib()
        H = C.CDLL("libamdhip64.so.7")
        vp = C.c_void_p
        H.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        H.hipStreamCreate.argtypes = [C.POINTER(vp)]
        H.hipStreamDestroy.argtypes = [vp]
        H.hipStreamSynchronize.argtypes = [vp]
        H.hipEventCreate.argtypes = [C.POINTER(vp)]
        H.hipEventDestroy.argtypes = [vp]
        H.hipEventRecord.argtypes = [vp, vp]
        H.hipEventSynchronize.argtypes = [vp]
        H.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        _hip = H
    return _hip


def _hip_check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with HIP error %d" % (what, rc))


class DeviceArray:
    """A device copy of a complex128 numpy array (freed with the object)."""

    def __init__(self, a=None, shape=None):
        import numpy as np
        H = hip()
        self.host = np.ascontiguousarray(a, dtype=np.complex128) if a is not None else np.zeros(shape, np.complex128)
        self.ptr = C.c_void_p()
        n = max(self.host.nbytes, 16)
        _hip_check(H.hipMalloc(C.byref(self.ptr), n), "hipMalloc")
        if a is not None and self.host.nbytes:
            _hip_check(H.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1), "hipMemcpy")

    def numpy(self):
        import numpy as np
        out = np.empty_like(self.host)
        if out.nbytes:
            _hip_check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2), "hipMemcpy")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value and _hip is not None:
            _hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()
