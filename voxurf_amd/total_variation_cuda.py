"""Drop-in for the reference's `total_variation_cuda` extension
(lib/cuda/total_variation.cpp:29-32): in-place TV gradient accumulation, HIP kernels.

`param` and `grad` may be the reference's contiguous [1, C, X, Y, Z] tensors or the
channels-last layout voxurf_amd.grid.DenseGrid uses for multi-channel grids; both must
share one layout.
"""
import torch

from ._lib import call, dtype_code, ptr, stream_of
from .render_utils_cuda import _check_input  # noqa: F401  (same CHECK_CUDA message)


def grid_geom(t):
    """int64 CPU tensor {C, X, Y, Z, sc, sx, sy, sz} for a [1, C, X, Y, Z] grid tensor."""
    if t.dim() != 5 or t.size(0) != 1:
        raise RuntimeError("grid tensors must have shape [1, C, X, Y, Z]")
    C, X, Y, Z = t.shape[1:]
    s = t.stride()
    return torch.tensor([C, X, Y, Z, s[1], s[2], s[3], s[4]], dtype=torch.int64)


def _check_grid(t, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not (t.is_contiguous() or t.is_contiguous(memory_format=torch.channels_last_3d)):
        raise RuntimeError(f"{name} must be contiguous")


def _check_pair(param, grad):
    _check_grid(param, "param")
    _check_grid(grad, "grad")
    if param.shape != grad.shape or param.stride() != grad.stride() or param.dtype != grad.dtype:
        raise RuntimeError("param and grad must share shape, layout and dtype")


def total_variation_add_grad(param, grad, wx, wy, wz, dense_mode):
    """total_variation.cpp:16-20 (kernel total_variation_kernel.cu:13-35)."""
    _check_pair(param, grad)
    geom = grid_geom(param)
    call("vx_total_variation_add_grad", dtype_code(param), ptr(param), ptr(grad),
         geom.data_ptr(), float(wx), float(wy), float(wz), int(bool(dense_mode)),
         stream_of(param))


def total_variation_add_grad_new(param, grad, mask, wx, wy, wz, dense_mode):
    """total_variation.cpp:22-27 (kernel total_variation_kernel.cu:38-66)."""
    _check_pair(param, grad)
    _check_grid(mask, "mask")
    if mask.shape != param.shape or mask.stride() != param.stride() or mask.dtype != param.dtype:
        raise RuntimeError("mask must share the shape, layout and dtype of param")
    geom = grid_geom(param)
    call("vx_total_variation_add_grad_new", dtype_code(param), ptr(param), ptr(grad), ptr(mask),
         geom.data_ptr(), float(wx), float(wy), float(wz), int(bool(dense_mode)),
         stream_of(param))
