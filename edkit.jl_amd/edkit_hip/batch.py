"""Multi-sector matvec batches replayed as one HIP graph.

MatvecBatch records y_s = (accumulate ? y_s : 0) + H_s x_s for every item s
(any operators and sectors on one context, blocks of nb <= 8 columns through
the block kernel) and replays them with one graph launch -- the batched form
of looping mul!(Y_s, H_s, X_s) over sectors (Operator.jl:322-328).
"""
import ctypes as C

import numpy as np

from ._lib import BatchItem, check, lib
from .device import DeviceArray, dev_ptr, dtype_code
from .matvec import device_operator


class MatvecBatch:
    """items: sequence of (H, x, y) or (H, x, y, accumulate) with device arrays
    x, y of shape (D,) or (D, nb), all of one dtype and one context.  streams:
    graph branches (0: one for sectors of >= 2^20 row-vectors, else 8).
    fuse: block products of sectors that take the same specialised block
    kernel run as one launch over all their chunks (else one launch each)."""

    def __init__(self, items, streams=0, fuse=True):
        if not items:
            raise ValueError("empty batch")
        self._keep = []
        arr = (BatchItem * len(items))()
        dtype = None
        ctx = None
        for i, it in enumerate(items):
            H, x, y = it[:3]
            acc = int(it[3]) if len(it) > 3 else 0
            if not isinstance(x, DeviceArray) or not isinstance(y, DeviceArray):
                raise TypeError("batch vectors must be DeviceArray")
            if x.dtype != y.dtype or x.shape != y.shape:
                raise ValueError("x and y must have the same shape and dtype")
            if dtype is None:
                dtype, ctx = x.dtype, H.B.ctx
            elif x.dtype != dtype:
                raise ValueError("all items of a batch share one dtype")
            if H.B.ctx is not ctx:
                raise ValueError("all items of a batch share one context")
            if x.shape[0] != H.B.D:
                raise ValueError("vector length does not match the basis dimension")
            nb = 1 if len(x.shape) == 1 else int(np.prod(x.shape[1:]))
            op = device_operator(H)
            arr[i] = BatchItem(op.h, dev_ptr(x), dev_ptr(y), nb, acc)
            self._keep.append((op, x, y))
        self.ctx = ctx
        self.dtype = dtype
        h = C.c_void_p()
        check(lib().edk_batch_create_ex(ctx.h, dtype_code(dtype), arr, len(items), int(streams),
                                        1 if fuse else 0, C.byref(h)))
        self.h = h

    @property
    def launches(self):
        """Kernel launches the batch's graph holds."""
        return int(lib().edk_batch_launches(self.h))

    def launch(self):
        """One graph launch (asynchronous on the context stream)."""
        check(lib().edk_batch_launch(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().edk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
