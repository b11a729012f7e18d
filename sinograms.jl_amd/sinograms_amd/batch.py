"""Batched reconstruction as one HIP graph, and plan autotuning.

The reference reconstructs a stack by a serial loop over slices
(``fbp(plan, aa)``, src/fbp2/fbp.jl:46-68).  Here many reconstructions of
uneven size -- each with its own plan -- are captured once into a HIP graph
whose independent plans run on parallel branches, and then replayed with one
launch.  Before capture every plan not yet tuned is autotuned: its candidate
kernel shapes are timed on the item's own data and the fastest is kept, per
size bucket, in the context's cache (so plans of one bucket are timed once).
"""

from __future__ import annotations

import ctypes

from . import _lib
from .recon import _Plan


def _dev_ptr(t, what):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32):
        raise TypeError(f"{what} must be a float32 CUDA tensor")
    if not t.permute(*range(t.dim() - 1, -1, -1)).is_contiguous():
        raise ValueError(f"{what} must be column-major (Julia order)")
    return t.data_ptr()


def autotune(plan: _Plan, sino, out, reps: int = 3, use_cache: bool = True) -> dict:
    """Time the plan's candidate kernel shapes on (sino, out) (column-major
    float32 CUDA tensors) and keep the fastest; returns the plan's tuning."""
    plan._ctx.check(_lib.load().sgx_plan_autotune(plan.h, _dev_ptr(sino, "sino"),
                                                  _dev_ptr(out, "out"), int(reps),
                                                  int(bool(use_cache))))
    return tuning(plan)


def tuning(plan: _Plan) -> dict:
    """The plan's kernel shape: cone-beam backprojector (one of
    ``_lib.PLAN_CBCT_KERNELS``),
    2-D view-grouping target and LDS view chunk, and whether it was tuned."""
    v = (ctypes.c_int64 * 4)()
    plan._ctx.check(_lib.load().sgx_plan_tuning(plan.h, v))
    return {"cbct_kernel": _lib.PLAN_CBCT_KERNELS[v[0]], "want_blocks": v[1], "view_chunk": v[2],
            "tuned": bool(v[3])}


class Batch:
    """Reconstructions (plan, sino, out) replayed as one HIP graph.

    Inputs and outputs are column-major float32 CUDA tensors that must stay
    alive (and keep their storage) while the batch is used: the graph is
    captured with their addresses."""

    def __init__(self, device: int = 0):
        self._ctx = _lib.context(device)
        h = ctypes.c_void_p()
        self._ctx.check(_lib.load().sgx_batch_create(self._ctx.h, ctypes.byref(h)))
        self.h = h
        self._keep = []

    def add(self, plan: _Plan, sino, out) -> None:
        if plan._ctx is not self._ctx:
            raise ValueError("the plan belongs to another device context")
        self._ctx.check(_lib.load().sgx_batch_add(self.h, plan.h, _dev_ptr(sino, "sino"),
                                                  _dev_ptr(out, "out")))
        self._keep.append((plan, sino, out))

    def execute(self, tune: bool = True, wait: bool = True, queues=None) -> None:
        """Run the batch's graph (autotuning and capturing it first when items
        were added or plans changed).  queues: None times both forms at
        capture and keeps the faster; True takes the queue launches (one
        launch per backprojector variant over a table of tasks); False the
        graph of whole plan executes."""
        flags = (0 if tune else _lib.SGX_BATCH_NO_TUNE) | (0 if wait else _lib.SGX_ASYNC)
        if queues is True:
            flags |= _lib.SGX_BATCH_QUEUE
        elif queues is False:
            flags |= _lib.SGX_BATCH_NO_QUEUE
        with self._ctx.ordered():
            self._ctx.check(_lib.load().sgx_batch_execute(self.h, flags))

    def last_ms(self) -> float:
        """Milliseconds of the last graph launch (HIP events)."""
        ms = ctypes.c_float()
        self._ctx.check(_lib.load().sgx_batch_last_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def info(self) -> dict:
        n, b = ctypes.c_int64(), ctypes.c_int64()
        self._ctx.check(_lib.load().sgx_batch_info(self.h, ctypes.byref(n), ctypes.byref(b)))
        q, qi, t = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._ctx.check(_lib.load().sgx_batch_queues(self.h, ctypes.byref(q), ctypes.byref(qi),
                                                     ctypes.byref(t)))
        mq, ml = ctypes.c_float(), ctypes.c_float()
        self._ctx.check(_lib.load().sgx_batch_choice(self.h, ctypes.byref(mq), ctypes.byref(ml)))
        return {"items": n.value, "branches": b.value, "queues": q.value,
                "queued_items": qi.value, "tasks": t.value,
                "timed_ms_queues": mq.value, "timed_ms_launches": ml.value}

    def __len__(self):
        return len(self._keep)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.load().sgx_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass
