"""ctypes binding of libsgx.so (the C ABI declared in include/sgx.h).

The library is the only compute path: if it cannot be loaded, every call
raises.  There is no CPU fallback.
"""

from __future__ import annotations

import contextlib
import ctypes
import os
import threading

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SGX_LIB", os.path.join(os.path.dirname(_PKG), "libsgx.so"))

SGX_OK = 0
SGX_ERR_ARG = -1
SGX_ERR_SIZE = -2
SGX_ERR_VALUE = -3
SGX_ERR_DEVICE = -4
SGX_ERR_UNSUPPORTED = -5

SGX_HOST = 0x1
SGX_ASYNC = 0x2
SGX_Q_BF16 = 0x10
SGX_IN_F32 = 0x20

SGX_SINO_PAR = 0
SGX_SINO_FAN_ARC = 1
SGX_SINO_FAN_FLAT = 2
SGX_CT_FAN_ARC = 3
SGX_CT_FAN_FLAT = 4
SGX_SINO_MOJ = 5

SGX_PREC_FP32 = 0
SGX_PREC_BF16 = 1
SGX_PREC_F16S = 2

# Every symbol include/sgx.h declares (checked by the CPU test suite).
EXPORTS = (
    "sgx_version", "sgx_ctx_create", "sgx_ctx_destroy", "sgx_last_error",
    "sgx_ctx_set_stream", "sgx_ctx_synchronize", "sgx_back_par", "sgx_back_fan",
    "sgx_cbct_back", "sgx_cbct_back_opts", "sgx_plan_set_cbct_opts", "sgx_sino_filter",
    "sgx_parker_fan_short", "sgx_radon_ellipses", "sgx_radon_ellipsoids", "sgx_plan_create",
    "sgx_plan_destroy", "sgx_plan_set_view_weight", "sgx_plan_execute", "sgx_plan_filtered",
    "sgx_plan_view_weight", "sgx_plan_last_times", "sgx_plan_read_filtered",
    "sgx_plan_stored_bytes", "sgx_plan_read_stored", "sgx_plan_pitch",
    "sgx_ctx_set_option", "sgx_ctx_get_option", "sgx_plan_autotune", "sgx_plan_tuning",
    "sgx_batch_create", "sgx_batch_destroy", "sgx_batch_add", "sgx_batch_execute",
    "sgx_batch_last_ms", "sgx_batch_info", "sgx_batch_queues", "sgx_batch_choice",
    "sgx_plan_filter",
    "sgx_plan_back",
    "sgx_plan_back_t", "sgx_net_create", "sgx_net_destroy", "sgx_net_num_params",
    "sgx_net_params", "sgx_net_grads", "sgx_net_zero_grad", "sgx_net_forward",
    "sgx_net_backward", "sgx_net_adam", "sgx_mse",
    "sgx_project_bdd", "sgx_backproject_bdd", "sgx_cbct_back_noalloc",
)


class SgxGeom(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("parker", ctypes.c_int32),
        ("ns", ctypes.c_int64), ("nt", ctypes.c_int64), ("na", ctypes.c_int64),
        ("ds", ctypes.c_double), ("dt", ctypes.c_double),
        ("offset_s", ctypes.c_double), ("offset_t", ctypes.c_double),
        ("dsd", ctypes.c_double), ("dso", ctypes.c_double),
        ("source_offset", ctypes.c_double), ("gamma_max", ctypes.c_double),
        ("angle_weight", ctypes.c_double),
    ]


class SgxCbctOpts(ctypes.Structure):
    _fields_ = [("source_zs", ctypes.c_void_p), ("arc_t_exact", ctypes.c_int32),
                ("source_offset", ctypes.c_double)]


class SgxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None
_lock = threading.Lock()


def load():
    """Load libsgx.so, declaring argument types.  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libsgx.so not found at {LIB_PATH}; build it with "
                "`python sinograms.jl_amd/build.py` (there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    i64, f64, i32, u32 = ctypes.c_int64, ctypes.c_double, ctypes.c_int, ctypes.c_uint32
    vp, cp = ctypes.c_void_p, ctypes.c_char_p
    L.sgx_version.restype = i32
    L.sgx_ctx_create.argtypes = [i32, ctypes.POINTER(vp)]
    L.sgx_ctx_destroy.argtypes = [vp]
    L.sgx_ctx_destroy.restype = None
    L.sgx_last_error.argtypes = [vp]
    L.sgx_last_error.restype = cp
    L.sgx_ctx_set_stream.argtypes = [vp, vp]
    L.sgx_ctx_synchronize.argtypes = [vp]
    L.sgx_back_par.argtypes = [vp, vp, i64, i64, i64, vp, vp, f64, f64, vp, i64, vp, i64,
                               vp, vp, u32]
    L.sgx_back_fan.argtypes = [vp, vp, i64, i64, i64, vp, vp, f64, f64, f64, i32, f64, f64,
                               vp, i64, vp, i64, vp, vp, u32]
    L.sgx_cbct_back.argtypes = [vp, vp, i64, i64, i64, i64, vp, f64, f64, f64, f64, f64, f64,
                                i32, vp, i64, vp, i64, vp, i64, vp, vp, u32]
    L.sgx_cbct_back_opts.argtypes = [vp, vp, i64, i64, i64, i64, vp, f64, f64, f64, f64, f64,
                                     f64, i32, vp, i64, vp, i64, vp, i64, vp, vp, u32, vp]
    L.sgx_cbct_back_noalloc.argtypes = [vp, vp, i64, i64, i64, i64, vp, f64, f64, f64, f64, f64,
                                        f64, i32, vp, i64, vp, i64, vp, i64, vp, vp, u32]
    L.sgx_sino_filter.argtypes = [vp, vp, i64, i64, vp, i64, i64, i32, vp, u32]
    L.sgx_radon_ellipses.argtypes = [vp, i32, vp, i64, vp, i64, f64, f64, vp, i64, vp, u32]
    L.sgx_radon_ellipsoids.argtypes = [vp, i32, vp, i64, vp, i64, vp, vp, i64, f64, f64, vp,
                                       i64, vp, u32]
    L.sgx_parker_fan_short.argtypes = [vp, i64, i64, vp, vp, f64, vp, u32]
    L.sgx_plan_create.argtypes = [vp, ctypes.POINTER(SgxGeom), vp, vp, i64, vp, i64, vp, i64,
                                  vp, vp, i64, i32, ctypes.POINTER(vp)]
    L.sgx_plan_destroy.argtypes = [vp]
    L.sgx_plan_destroy.restype = None
    L.sgx_plan_set_view_weight.argtypes = [vp, vp]
    L.sgx_plan_set_cbct_opts.argtypes = [vp, vp]
    L.sgx_plan_execute.argtypes = [vp, vp, vp, u32]
    L.sgx_plan_filtered.argtypes = [vp]
    L.sgx_plan_filtered.restype = vp
    L.sgx_plan_view_weight.argtypes = [vp, vp]
    L.sgx_plan_read_filtered.argtypes = [vp, vp, vp]
    L.sgx_plan_stored_bytes.argtypes = [vp]
    L.sgx_plan_stored_bytes.restype = ctypes.c_int64
    L.sgx_plan_read_stored.argtypes = [vp, vp, ctypes.c_int64]
    L.sgx_ctx_set_option.argtypes = [vp, ctypes.c_int, ctypes.c_int64]
    L.sgx_ctx_get_option.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    L.sgx_plan_pitch.argtypes = [vp]
    L.sgx_plan_pitch.restype = ctypes.c_int64
    L.sgx_project_bdd.argtypes = [vp, vp, i64, i64, vp, i64, f64, f64, f64, f64, vp, u32]
    L.sgx_backproject_bdd.argtypes = [vp, vp, i64, i64, vp, i64, f64, f64, f64, f64, i32, vp,
                                      u32]
    L.sgx_plan_last_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float),
                                      ctypes.POINTER(ctypes.c_float)]
    L.sgx_plan_autotune.argtypes = [vp, vp, vp, i32, i32]
    L.sgx_plan_tuning.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.sgx_batch_create.argtypes = [vp, ctypes.POINTER(vp)]
    L.sgx_batch_destroy.argtypes = [vp]
    L.sgx_batch_destroy.restype = None
    L.sgx_batch_add.argtypes = [vp, vp, vp, vp]
    L.sgx_batch_execute.argtypes = [vp, u32]
    L.sgx_batch_last_ms.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    L.sgx_batch_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(ctypes.c_int64)]
    L.sgx_batch_queues.argtypes = [vp, ctypes.POINTER(ctypes.c_int64),
                                   ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.sgx_batch_choice.argtypes = [vp, ctypes.POINTER(ctypes.c_float),
                                   ctypes.POINTER(ctypes.c_float)]
    L.sgx_plan_filter.argtypes = [vp, vp, vp, u32]
    L.sgx_plan_back.argtypes = [vp, vp, vp, u32]
    L.sgx_plan_back_t.argtypes = [vp, vp, vp, u32]
    L.sgx_net_create.argtypes = [vp, i32, i32, ctypes.c_uint64, ctypes.POINTER(vp)]
    L.sgx_net_destroy.argtypes = [vp]
    L.sgx_net_destroy.restype = None
    L.sgx_net_num_params.argtypes = [vp]
    L.sgx_net_num_params.restype = ctypes.c_int64
    L.sgx_net_params.argtypes = [vp, vp, i32]
    L.sgx_net_grads.argtypes = [vp, vp]
    L.sgx_net_zero_grad.argtypes = [vp]
    L.sgx_net_forward.argtypes = [vp, vp, i64, i64, vp, i32]
    L.sgx_net_backward.argtypes = [vp, vp]
    L.sgx_net_adam.argtypes = [vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float]
    L.sgx_mse.argtypes = [vp, vp, vp, i64, ctypes.c_float, vp, ctypes.POINTER(ctypes.c_double)]


SGX_OPT_CBCT_KERNEL = 1
SGX_OPT_CBCT_MIRROR = 2
SGX_OPT_VIEW_CHUNKS = 3
CBCT_MIRROR = ("never", "fp16", "always")    # values of SGX_OPT_CBCT_MIRROR
# values of SGX_OPT_CBCT_KERNEL
CBCT_KERNELS = ("staged", "gather", "staged_packed16", "staged_z8", "staged_z8w")
# cone-beam kernels a plan reports (sgx_plan_tuning)
PLAN_CBCT_KERNELS = ("staged", "gather", "staged_packed16", "staged_mirror", "staged_z8",
                     "staged_z8w")
SGX_BATCH_NO_TUNE = 0x100
SGX_BATCH_NO_QUEUE = 0x200
SGX_BATCH_QUEUE = 0x400


class Context:
    """One libsgx context (device + stream)."""

    def __init__(self, device: int = 0, stream=None):
        L = load()
        h = ctypes.c_void_p()
        rc = L.sgx_ctx_create(device, ctypes.byref(h))
        if rc != SGX_OK:
            raise SgxError(rc, f"sgx_ctx_create({device}) failed ({rc})")
        self.h = h
        self.device = device
        self.stream = stream
        if stream is not None:
            self.check(L.sgx_ctx_set_stream(self.h, ctypes.c_void_p(stream)))

    def on_current_stream(self) -> bool:
        """True when this context's stream is torch's current stream of its
        device: device-array calls may then return without synchronising,
        ordered with torch's work like any torch operation."""
        if self.stream is None:
            return False
        import torch
        return torch.cuda.current_stream(self.device).cuda_stream == self.stream

    def _torch_stream(self):
        import torch
        if self.stream:
            return torch.cuda.ExternalStream(self.stream, device=self.device)
        return torch.cuda.default_stream(self.device)

    @contextlib.contextmanager
    def ordered(self, device_arrays: bool = True):
        """Order a call on this context's stream with torch's current stream.

        Before the call, the context's stream waits for the work queued so
        far on torch's current stream (an event recorded there), so that
        device arrays produced on that stream are read after they are
        written; after it, torch's current stream waits for the call's work.
        Yields True when the two streams are the same one (a call on device
        arrays may then return without a host wait).  With
        device_arrays=False (host arrays only) nothing is ordered."""
        if not device_arrays:
            yield False
            return
        if self.on_current_stream():
            yield True
            return
        import torch
        cur = torch.cuda.current_stream(self.device)
        mine = self._torch_stream()
        mine.wait_stream(cur)
        yield False
        cur.wait_stream(mine)

    def check(self, rc):
        if rc != SGX_OK:
            raise SgxError(rc, load().sgx_last_error(self.h).decode())

    @property
    def cbct_kernel(self) -> str:
        """Cone-beam backprojector of this context: "staged" (LDS-staged
        detector windows, the default), "gather", "staged_packed16" (staged,
        with scaled-FP16 plans interpolating in packed FP16), "staged_z8"
        (staged, blocks 8 voxels deep) or "staged_z8w" (8-deep blocks with
        windows of up to 64 bins x 32 rows, for coarse voxels)."""
        v = ctypes.c_int64()
        self.check(load().sgx_ctx_get_option(self.h, SGX_OPT_CBCT_KERNEL, ctypes.byref(v)))
        return CBCT_KERNELS[v.value]

    @cbct_kernel.setter
    def cbct_kernel(self, name: str):
        self.check(load().sgx_ctx_set_option(self.h, SGX_OPT_CBCT_KERNEL,
                                             CBCT_KERNELS.index(name)))

    @property
    def cbct_mirror(self) -> str:
        """When the staged cone-beam backprojector takes its mirror form (each
        voxel z > 0 paired with its image at -z, where the geometry allows):
        "fp16" (for scaled-FP16 plans, the default; the autotuner times it for
        the others), "always" or "never"."""
        v = ctypes.c_int64()
        self.check(load().sgx_ctx_get_option(self.h, SGX_OPT_CBCT_MIRROR, ctypes.byref(v)))
        return CBCT_MIRROR[v.value]

    @cbct_mirror.setter
    def cbct_mirror(self, when: str):
        self.check(load().sgx_ctx_set_option(self.h, SGX_OPT_CBCT_MIRROR,
                                             CBCT_MIRROR.index(when)))

    @property
    def view_chunks(self) -> int:
        """View chunks of a cone-beam plan execute (1 to 8, default 1): the
        filters of later chunks run beside the backprojections of earlier
        ones; 1 runs one filter launch, then one backprojection."""
        v = ctypes.c_int64()
        self.check(load().sgx_ctx_get_option(self.h, SGX_OPT_VIEW_CHUNKS, ctypes.byref(v)))
        return v.value

    @view_chunks.setter
    def view_chunks(self, k: int):
        self.check(load().sgx_ctx_set_option(self.h, SGX_OPT_VIEW_CHUNKS, int(k)))

    def close(self):
        if self.h:
            load().sgx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts = {}


def context(device: int = 0) -> Context:
    """The shared context of a device, running on torch's current stream."""
    c = _contexts.get(device)
    if c is None:
        import torch
        with torch.cuda.device(device):
            stream = torch.cuda.current_stream(device).cuda_stream
        c = Context(device, stream)
        _contexts[device] = c
    return c
