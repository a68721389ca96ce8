"""Device plumbing: torch-ROCm allocations, the stream handle, descriptor
uploads through a ring of pinned staging buffers."""
import numpy as np
import torch

from . import _lib

TORCH_DTYPE = {_lib.F64: torch.float64, _lib.F32: torch.float32, _lib.BF16: torch.bfloat16}
ITEMSIZE = {_lib.F64: 8, _lib.F32: 4, _lib.BF16: 2}


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError('the DMRG engine needs a ROCm GPU (torch.cuda.is_available() is False)')
    _lib.load()


def stream_handle():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(x):
    import ctypes
    return ctypes.c_void_p(int(x))


class Uploader:
    """Packs numpy record arrays into one byte buffer, copies it to the device
    on the current stream and returns the device address of each array."""

    def __init__(self, device, ring=4):
        self.device = device
        self.cap = 0
        self.ring = ring
        self.slot = 0
        self.host = [None] * ring
        self.events = [None] * ring
        self.dev = [None] * ring

    def _ensure(self, nbytes):
        if nbytes <= self.cap:
            return
        cap = max(nbytes, 2 * self.cap, 1 << 16)
        torch.cuda.current_stream().synchronize()
        for k in range(self.ring):
            self.host[k] = torch.empty(cap, dtype=torch.uint8, pin_memory=True)
            self.dev[k] = torch.empty(cap, dtype=torch.uint8, device=self.device)
            self.events[k] = None
        self.cap = cap

    def upload(self, arrays):
        offs = []
        total = 0
        for a in arrays:
            offs.append(total)
            total += (a.nbytes + 255) // 256 * 256
        self._ensure(max(total, 256))
        k = self.slot
        self.slot = (self.slot + 1) % self.ring
        if self.events[k] is not None:
            self.events[k].synchronize()
        hbuf = self.host[k].numpy()
        for a, o in zip(arrays, offs):
            if a.nbytes:
                hbuf[o:o + a.nbytes] = np.frombuffer(np.ascontiguousarray(a).tobytes(), np.uint8)
        self.dev[k][:total].copy_(self.host[k][:total], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self.events[k] = ev
        base = self.dev[k].data_ptr()
        # empty arrays still get a valid (unused) address: a GEMM problem may
        # legitimately have no K-segments (C = beta*C0)
        return [base + o for a, o in zip(arrays, offs)]
