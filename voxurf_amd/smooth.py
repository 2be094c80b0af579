"""Dense per-step grid operators of the fine stage on the HIP kernels (csrc/smooth.hip).

* smooth_grad_tv / SmoothGradTV: the smooth_grad_tv term of
  Voxurf.density_total_variation (lib/voxurf_fine.py:412-421): weight * mean over the
  masked entries of (S(G) - G)^2, with G the 'interpolate' gradient grid of the SDF
  (:440-450) and S = tv_smooth_conv (:204-240, detached), as a loss differentiable in the
  SDF grid; smooth_grad_tv_add_grad adds its gradient straight into a gradient buffer.
* GaussianConv3d: the module _gaussian_3dconv builds (:246-258), used by smooth_conv,
  init_sdf_from_sdf(smooth=True) and smooth_scale, differentiable.
* resample_trilinear: F.interpolate(trilinear, align_corners=True), the resampling of
  scale_volume_grid (lib/grid.py:60-65, lib/voxurf_fine.py:385-396) and init_sdf_from_sdf.

Every kernel in these operators is separable: the 3x3x3 and ksize^3 kernels of the
reference are outer products of one 1-D filter with itself, so the 1-D taps are what
the kernels take.
"""
import numpy as np
import torch
import torch.nn as nn

from ._lib import call, dtype_code, lib, ptr, stream_of


def gaussian_taps(ksize, sigma):
    """1-D factor t of _gaussian_3dconv's kernel: kernel / kernel.sum() = t (x) t (x) t."""
    x = np.arange(-(ksize // 2), ksize // 2 + 1, 1, dtype=np.float64)
    g = np.exp(-(x ** 2) / (2 * sigma ** 2))
    return (g / g.sum()).tolist()


def tv_smooth_taps(sigma=0.0):
    """1-D factor of tv_smooth_conv's kernel ([1,2,1]^3 * exp(-sigma (r^2 - 1)), normalised)."""
    g = np.array([1.0, 2.0, 1.0]) * np.exp(-sigma * np.array([1.0, 0.0, 1.0]))
    return (g / g.sum()).tolist()


def _check_grid(t, name):
    if not t.is_cuda:
        raise RuntimeError(f"{name}: CUDA (HIP) tensor expected")
    if t.dim() != 5 or t.shape[0] != 1:
        raise RuntimeError(f"{name}: [1, C, X, Y, Z] grid expected")


def _mask_args(mask, shape, dev):
    """(contiguous bool mask or None, number of selected voxels)."""
    if mask is None:
        return None, shape[0] * shape[1] * shape[2]
    m = mask.reshape(shape).to(device=dev, dtype=torch.bool).contiguous()
    count = int(m.sum())
    if count == 0:
        raise RuntimeError("smooth_grad_tv: the mask selects no voxel")
    return m, count


def smooth_grad_tv_add_grad(sdf_grid, sdf_grad, voxel_size, mask, weight, sigma=0.0,
                            mask_count=None, loss=None, workspace=None):
    """Add d(loss)/d(sdf) of the smooth_grad_tv term into sdf_grad (same shape as the
    grid) and write the term's value into `loss` (a 1-element tensor) when given.
    mask_count: number of True entries of mask (computed when None)."""
    _check_grid(sdf_grid, "smooth_grad_tv")
    s = sdf_grid.detach()
    if s.shape[1] != 1 or not s.is_contiguous():
        raise RuntimeError("smooth_grad_tv: contiguous one-channel SDF grid expected")
    X, Y, Z = s.shape[-3:]
    if mask is not None and mask_count is not None:
        m = mask.reshape(X, Y, Z).to(torch.bool).contiguous()
        count = int(mask_count)
    else:
        m, count = _mask_args(mask, (X, Y, Z), s.device)
    if sdf_grad is not None and (sdf_grad.shape != s.shape or sdf_grad.dtype != s.dtype or
                                 not sdf_grad.is_contiguous()):
        raise RuntimeError("smooth_grad_tv: sdf_grad must match the grid")
    if workspace is None:
        workspace = smooth_grad_tv_workspace(s)
    dg, part = workspace
    taps = np.asarray(tv_smooth_taps(sigma), dtype=np.float64)
    call("vx_smooth_grad_tv", dtype_code(s), ptr(s), X, Y, Z, float(voxel_size),
         taps.ctypes.data, ptr(m), count, float(weight), ptr(dg), ptr(part), ptr(loss),
         ptr(sdf_grad), stream_of(s))
    return dg


def smooth_grad_tv_workspace(sdf_grid):
    """(dG [3, X, Y, Z], partial sums) buffers for smooth_grad_tv_add_grad."""
    X, Y, Z = sdf_grid.shape[-3:]
    dg = torch.empty((3, X, Y, Z), dtype=sdf_grid.dtype, device=sdf_grid.device)
    part = torch.empty(max(1, lib().vx_smooth_grad_tv_partials(X, Y, Z)), dtype=sdf_grid.dtype,
                       device=sdf_grid.device)
    return dg, part


class SmoothGradTV(torch.autograd.Function):
    """loss = smooth_grad_tv(sdf_grid, ...) with its gradient w.r.t. the SDF grid."""

    @staticmethod
    def forward(ctx, sdf_grid, voxel_size, mask, weight, sigma):
        loss = torch.empty(1, dtype=sdf_grid.dtype, device=sdf_grid.device)
        dg = smooth_grad_tv_add_grad(sdf_grid, None, voxel_size, mask, weight, sigma, loss=loss)
        ctx.save_for_backward(dg)
        ctx.vs = float(voxel_size)
        ctx.shape = sdf_grid.shape
        return loss[0]

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gout):
        (dg,) = ctx.saved_tensors
        X, Y, Z = ctx.shape[-3:]
        g = torch.zeros(ctx.shape, dtype=dg.dtype, device=dg.device)
        dgs = (dg * gout).contiguous()
        call("vx_dense_sdf_gradient_bwd", dtype_code(g), ptr(dgs), X, Y, Z, ctx.vs, ptr(g),
             stream_of(g))
        return g, None, None, None, None


def smooth_grad_tv(sdf_grid, voxel_size, mask=None, weight=1.0, sigma=0.0):
    """weight * mean((tv_smooth_conv(G).detach() - G)[mask]^2) (lib/voxurf_fine.py:417-420)."""
    return SmoothGradTV.apply(sdf_grid, voxel_size, mask, weight, sigma)


def _smooth3d(x, taps, transpose):
    C, X, Y, Z = x.shape[-4:]
    out = torch.empty_like(x)
    tmp = torch.empty_like(x)
    t = np.asarray(taps, dtype=np.float64)
    call("vx_smooth3d", dtype_code(x), ptr(x), ptr(out), ptr(tmp), C, X, Y, Z, t.ctypes.data,
         len(taps), int(transpose), stream_of(x))
    return out


class _Smooth3dFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, taps):
        ctx.taps = taps
        return _smooth3d(x.contiguous(), taps, False)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gout):
        return _smooth3d(gout.contiguous(), ctx.taps, True), None


class GaussianConv3d(nn.Module):
    """Drop-in for the Conv3d _gaussian_3dconv returns (lib/voxurf_fine.py:246-258):
    replicate-padded normalised Gaussian smoothing of a [1, C, X, Y, Z] grid (each
    channel on its own), differentiable w.r.t. its input."""

    def __init__(self, ksize=3, sigma=1.0):
        super().__init__()
        if ksize % 2 == 0 or ksize < 1 or ksize > 9:
            raise RuntimeError("GaussianConv3d: ksize must be odd and <= 9")
        self.ksize, self.sigma = ksize, sigma
        self.taps = gaussian_taps(ksize, sigma)

    def forward(self, x):
        _check_grid(x, "GaussianConv3d")
        return _Smooth3dFn.apply(x, self.taps)

    def extra_repr(self):
        return f"ksize={self.ksize}, sigma={self.sigma}"


def resample_trilinear(grid, size):
    """F.interpolate(grid, size, mode='trilinear', align_corners=True) for [1,C,X,Y,Z]."""
    _check_grid(grid, "resample_trilinear")
    g = grid.detach().contiguous()
    C, X, Y, Z = g.shape[1:]
    X2, Y2, Z2 = (int(v) for v in size)
    out = torch.empty((1, C, X2, Y2, Z2), dtype=g.dtype, device=g.device)
    call("vx_resample_trilinear", dtype_code(g), ptr(g), C, X, Y, Z, ptr(out), X2, Y2, Z2,
         stream_of(g))
    return out


def grid_resolution(xyz_min, xyz_max, num_voxels):
    """(voxel_size, world_size) of _set_grid_resolution (lib/voxurf_fine.py:315-320)."""
    xyz_min = torch.as_tensor(xyz_min, dtype=torch.float32)
    xyz_max = torch.as_tensor(xyz_max, dtype=torch.float32)
    voxel_size = ((xyz_max - xyz_min).prod() / num_voxels).pow(1 / 3)
    world_size = ((xyz_max - xyz_min) / voxel_size).long()
    return voxel_size, world_size


__all__ = ["gaussian_taps", "tv_smooth_taps", "smooth_grad_tv", "smooth_grad_tv_add_grad",
           "smooth_grad_tv_workspace", "SmoothGradTV", "GaussianConv3d", "resample_trilinear",
           "grid_resolution"]
