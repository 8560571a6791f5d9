"""Device context: one HIP device + one stream, shared by all subsystems.

Replaces the luisa::compute::Device/Stream pair of the reference
(Solver/SimulationSolver/solver_interface.h:35-67).  By default kernels are
launched on torch's current stream so that torch-side work (allocation,
copies, reductions in tests) is ordered with them.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import check, lib


class Context:
    def __init__(self, device: int = 0, stream: "torch.cuda.Stream | None" = None):
        if not torch.cuda.is_available():
            raise RuntimeError("untangling26_amd needs a HIP device; none is visible")
        self.device = torch.device("cuda", device)
        self.torch_stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        h = C.c_void_p()
        rc = lib().ut_ctx_create(device, C.c_void_p(self.torch_stream.cuda_stream), 0, C.byref(h))
        if rc != 0:
            raise RuntimeError(f"ut_ctx_create failed with {rc}")
        self.handle = h

    def check(self, rc: int) -> int:
        return check(rc, self.handle)

    def synchronize(self) -> None:
        self.check(lib().ut_ctx_synchronize(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().ut_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def empty(self, shape, dtype):
        return torch.empty(shape, dtype=dtype, device=self.device)

    def zeros(self, shape, dtype):
        return torch.zeros(shape, dtype=dtype, device=self.device)


_default: Context | None = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default
