"""ctypes binding of libhs2.so, the HIP library (include/heatsim2_amd.h).

There is no fallback: if the library is missing or no GPU is visible, every
entry point raises.  Device buffers are torch tensors; their data_ptr() and
torch's current HIP stream are handed to the C ABI.
"""
import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhs2.so")

HS2_OK = 0
HS2_OPT_COMBINED_X = 1
# the version string hs2_version() of a matching libhs2.so starts with
LIBRARY_VERSION = "0.5"
_ERRNAMES = {1: "HS2_E_ARG", 2: "HS2_E_UNSUPPORTED", 3: "HS2_E_HIP", 4: "HS2_E_STATE",
             5: "HS2_E_ALLOC"}


class GridDesc(ctypes.Structure):
    _fields_ = [("nz", ctypes.c_int32), ("ny", ctypes.c_int32), ("nx", ctypes.c_int32),
                ("dz", ctypes.c_double), ("dy", ctypes.c_double), ("dx", ctypes.c_double),
                ("dt", ctypes.c_double)]


class HS2Error(RuntimeError):
    def __init__(self, func, code, msg):
        super().__init__("%s failed (%s): %s" % (func, _ERRNAMES.get(code, code), msg))
        self.code = code


# every symbol include/heatsim2_amd.h declares
EXPORTS = ("hs2_plan_create", "hs2_plan_destroy", "hs2_plan_set_compact", "hs2_plan_set_dense",
           "hs2_plan_set_dense_asym", "hs2_plan_expand", "hs2_plan_set_config", "hs2_plan_set_option",
           "hs2_plan_info",
           "hs2_step", "hs2_run",
           "hs2_run_timed", "hs2_grad_work_doubles", "hs2_grad_forward", "hs2_grad_backward",
           "hs2_batch_create", "hs2_batch_destroy", "hs2_batch_layout", "hs2_batch_set_dense",
           "hs2_batch_set_options", "hs2_batch_forward", "hs2_batch_backward",
           "hs2_step_noalloc", "hs2_surface_temperature", "hs2_last_error", "hs2_version")

_lib = None


def load():
    """Load libhs2.so (no GPU needed to load it)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("heatsim2_amd: HIP library %s is not built; run "
                          "`python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.hs2_plan_create.argtypes = [ctypes.POINTER(GridDesc), ctypes.POINTER(vp)]
    L.hs2_plan_create.restype = i32
    L.hs2_plan_destroy.argtypes = [vp]
    L.hs2_plan_destroy.restype = None
    L.hs2_plan_set_compact.argtypes = [vp, vp, vp, vp, vp, i32, i32, vp, vp, vp]
    L.hs2_plan_set_compact.restype = i32
    L.hs2_plan_set_dense.argtypes = [vp, vp, vp, vp, vp, vp]
    L.hs2_plan_set_dense.restype = i32
    L.hs2_plan_set_dense_asym.argtypes = [vp] * 9
    L.hs2_plan_set_dense_asym.restype = i32
    L.hs2_plan_expand.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.hs2_plan_expand.restype = i32
    L.hs2_plan_set_config.argtypes = [vp, i32, i32, i32, i32, i32, i32, i32, i32]
    L.hs2_plan_set_config.restype = i32
    L.hs2_plan_set_option.argtypes = [vp, i32, i32]
    L.hs2_plan_set_option.restype = i32
    L.hs2_plan_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.hs2_plan_info.restype = i32
    L.hs2_step.argtypes = [vp, vp, vp, vp, vp, i32, vp, vp, vp]
    L.hs2_step.restype = i32
    L.hs2_run.argtypes = [vp, vp, i32, vp, vp, i32, vp, vp]
    L.hs2_run.restype = i32
    L.hs2_run_timed.argtypes = [vp, vp, i32, i32, ctypes.POINTER(ctypes.c_double), vp]
    L.hs2_run_timed.restype = i32
    L.hs2_grad_work_doubles.argtypes = [vp, i32]
    L.hs2_grad_work_doubles.restype = ctypes.c_size_t
    L.hs2_grad_forward.argtypes = [vp, vp, i32, i32, vp, vp, vp, i32, vp, vp]
    L.hs2_grad_forward.restype = i32
    L.hs2_grad_backward.argtypes = [vp, i32, i32, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp,
                                    ctypes.c_size_t, vp]
    L.hs2_grad_backward.restype = i32
    L.hs2_batch_create.argtypes = [vp, i32, i32, ctypes.POINTER(vp)]
    L.hs2_batch_create.restype = i32
    L.hs2_batch_destroy.argtypes = [vp]
    L.hs2_batch_destroy.restype = None
    L.hs2_batch_layout.argtypes = [vp, vp, vp, vp]
    L.hs2_batch_layout.restype = i32
    L.hs2_batch_set_dense.argtypes = [vp, vp, vp, vp, vp, vp]
    L.hs2_batch_set_dense.restype = i32
    L.hs2_batch_set_options.argtypes = [vp, i32, i32]
    L.hs2_batch_set_options.restype = i32
    L.hs2_batch_forward.argtypes = [vp, vp, vp, vp, vp, ctypes.c_size_t, vp]
    L.hs2_batch_forward.restype = i32
    L.hs2_batch_backward.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_size_t, vp]
    L.hs2_batch_backward.restype = i32
    L.hs2_step_noalloc.argtypes = [ctypes.POINTER(GridDesc)] + [vp] * 10
    L.hs2_step_noalloc.restype = i32
    L.hs2_surface_temperature.argtypes = [vp, i32, i32, i32, dbl, vp, vp]
    L.hs2_surface_temperature.restype = i32
    L.hs2_last_error.argtypes = []
    L.hs2_last_error.restype = ctypes.c_char_p
    L.hs2_version.argtypes = []
    L.hs2_version.restype = ctypes.c_char_p
    ver = L.hs2_version().decode()
    if not ver.startswith("heatsim2_amd %s " % LIBRARY_VERSION):
        raise ImportError("heatsim2_amd %s: %s reports '%s'; rebuild it" %
                          (LIBRARY_VERSION, LIB_PATH, ver))
    _lib = L
    return L


def check(func, code):
    if code != HS2_OK:
        raise HS2Error(func, code, load().hs2_last_error().decode())


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("heatsim2_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """data_ptr of a contiguous device tensor (or None)."""
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    return ctypes.c_void_p(t.data_ptr())


def host_ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def surface_temperature_device(T, dz):
    if T.dim() != 3 or T.dtype != torch.float64:
        raise ValueError("T must be a (nz, ny, nx) float64 tensor")
    T = T.contiguous()
    nz, ny, nx = T.shape
    out = torch.empty((ny, nx), dtype=torch.float64, device=T.device)
    check("hs2_surface_temperature",
          load().hs2_surface_temperature(ptr(T), nz, ny, nx, float(dz), ptr(out), stream_handle()))
    return out


def plan_info(plan):
    info = (ctypes.c_int * 14)()
    check("hs2_plan_info", load().hs2_plan_info(plan, info))
    keys = ("mode", "classes_z", "classes_y", "classes_x", "x_cpl", "x_lanes", "y_seg",
            "y_lines", "z_seg", "z_lines", "x_lines", "x_lines_y", "y_from_x", "x_combined")
    return dict(zip(keys, list(info)))
