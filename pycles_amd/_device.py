"""Device plumbing: the compiled binding, torch allocation and streams.

torch (ROCm) is used only to allocate device memory and to name the current
HIP stream; every computation on the path runs in the HIP kernels of
``libpycles_amd.so``.  Importing this module fails loudly when the binding
has not been built, and ``require_gpu`` fails when no GPU is visible: there is
no CPU fallback.
"""
from __future__ import annotations

import numpy as np
import torch

try:
    from . import _capi
except ImportError as e:  # pragma: no cover - exercised only on a broken build
    raise ImportError("pycles_amd: the HIP library binding is not built; run "
                      "`python -m pycles_amd.build` (or __graft_entry__.build())") from e

F64 = _capi.F64
F32 = _capi.F32


def precision_code(p) -> int:
    if p in (F64, "f64", "fp64", "double", torch.float64):
        return F64
    if p in (F32, "f32", "fp32", "float", torch.float32):
        return F32
    raise ValueError(f"unknown precision {p!r}")


def torch_dtype(precision: int):
    return torch.float64 if precision == F64 else torch.float32


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("pycles_amd: no GPU visible; the dynamical core runs only on the GPU")
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def to_device(a: np.ndarray, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=require_gpu(), dtype=dtype)


def dims_of(Gr) -> "_capi.Dims":
    d = Gr.dims
    return _capi.Dims(tuple(d.n), d.gw, tuple(d.dx))


class DeviceRef:
    """Reference-state profiles in HBM (FP64) plus the ``pycles_ref`` view."""

    def __init__(self, RS):
        self.t = {name: to_device(getattr(RS, name)) for name in
                  ("rho0", "rho0_half", "alpha0", "alpha0_half", "p0", "p0_half")}
        self.c = _capi.Ref(*[self.t[n].data_ptr() for n in
                             ("rho0", "rho0_half", "alpha0", "alpha0_half", "p0", "p0_half")])


def device_ref(RS) -> DeviceRef:
    cached = getattr(RS, "_device_ref", None)
    if cached is None:
        cached = DeviceRef(RS)
        RS._device_ref = cached
    return cached
