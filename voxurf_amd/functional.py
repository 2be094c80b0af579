"""torch.autograd.Functions over the HIP kernels of the Voxurf forward.

Each Function's forward and backward launch hand-written HIP kernels (libvoxhip.so);
PyTorch only allocates the tensors and records the graph.  The reference computes the
same quantities with F.grid_sample, elementwise torch ops and torch_scatter:
  GridSample        lib/grid.py:47-58 (DenseGrid.forward)
  SdfGradSample     lib/voxurf_fine.py:502-534 (grid_sampler, sample_ret+sample_grad)
  SampleSdfs        lib/voxurf_fine.py:537-577 (sample_sdfs)
  DenseSdfGradient  lib/voxurf_fine.py:440-450 (neus_sdf_gradient, 'interpolate')
  NeusAlpha         lib/voxurf_fine.py:463-500 (neus_alpha_from_sdf_scatter)
  Alphas2Weights    lib/voxurf_fine.py:981-997
  SegmentSum        torch_scatter.segment_coo(reduce='sum') at lib/voxurf_fine.py:753-777
"""
import torch

from . import render_utils_cuda as ru
from ._lib import call, dtype_code, ptr, stream_of
from .total_variation_cuda import grid_geom


def _zeros_like_grid(g):
    return torch.zeros_like(g, memory_format=torch.preserve_format)


def _dev_vec(v, like):
    if isinstance(v, torch.Tensor):
        return v.to(device=like.device, dtype=like.dtype).contiguous()
    return torch.tensor(v, device=like.device, dtype=like.dtype)


class GridSample(torch.autograd.Function):
    """out[P, C] = trilinear(grid [1,C,X,Y,Z], xyz [P,3]) with zero padding."""

    @staticmethod
    def forward(ctx, grid, xyz, xyz_min, xyz_max):
        xyz = xyz.contiguous()
        mn, mx = _dev_vec(xyz_min, grid), _dev_vec(xyz_max, grid)
        geom = grid_geom(grid)
        out = torch.empty((xyz.shape[0], grid.shape[1]), dtype=grid.dtype, device=grid.device)
        call("vx_grid_sample_fwd", dtype_code(grid), ptr(grid), geom.data_ptr(), ptr(xyz),
             ptr(mn), ptr(mx), xyz.shape[0], ptr(out), stream_of(grid))
        ctx.save_for_backward(xyz, mn, mx)
        ctx.grid_meta = (grid.shape, grid.stride(), grid.dtype, grid.device, geom)
        ctx.grid_like = grid
        return out

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gout):
        xyz, mn, mx = ctx.saved_tensors
        grid = ctx.grid_like
        ggrid = _zeros_like_grid(grid)
        gout = gout.contiguous()
        call("vx_grid_sample_bwd", dtype_code(grid), ptr(ggrid), ctx.grid_meta[4].data_ptr(),
             ptr(xyz), ptr(mn), ptr(mx), xyz.shape[0], ptr(gout), stream_of(grid))
        return ggrid, None, None, None


def grid_sample(grid, xyz, xyz_min, xyz_max):
    return GridSample.apply(grid, xyz, xyz_min, xyz_max)


class SdfGradSample(torch.autograd.Function):
    """Centre SDF, 6-neighbour gradient (xyz order) and the 6 neighbour values."""

    @staticmethod
    def forward(ctx, grid, xyz, xyz_min, xyz_max, voxel_size):
        xyz = xyz.contiguous()
        mn, mx = _dev_vec(xyz_min, grid), _dev_vec(xyz_max, grid)
        geom = grid_geom(grid)
        n = xyz.shape[0]
        sdf = torch.empty(n, dtype=grid.dtype, device=grid.device)
        grad = torch.empty((n, 3), dtype=grid.dtype, device=grid.device)
        feat = torch.empty((n, 6), dtype=grid.dtype, device=grid.device)
        call("vx_sdf_grad_fwd", dtype_code(grid), ptr(grid), geom.data_ptr(), ptr(xyz), ptr(mn),
             ptr(mx), float(voxel_size), n, ptr(sdf), ptr(grad), ptr(feat), stream_of(grid))
        ctx.save_for_backward(xyz, mn, mx)
        ctx.geom = geom
        ctx.vs = float(voxel_size)
        ctx.grid_like = grid
        return sdf, grad, feat

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gsdf, ggrad, gfeat):
        xyz, mn, mx = ctx.saved_tensors
        grid = ctx.grid_like
        ggrid = _zeros_like_grid(grid)
        c = lambda t: None if t is None else t.contiguous()  # noqa: E731
        call("vx_sdf_grad_bwd", dtype_code(grid), ptr(ggrid), ctx.geom.data_ptr(), ptr(xyz),
             ptr(mn), ptr(mx), ctx.vs, xyz.shape[0], ptr(c(gsdf)), ptr(c(ggrad)), ptr(c(gfeat)),
             stream_of(grid))
        return ggrid, None, None, None, None


class SampleSdfs(torch.autograd.Function):
    """feat [P, 6*nd], grad [P, 3*nd] in the reference's zyx component-major layout."""

    @staticmethod
    def forward(ctx, grid, xyz, xyz_min, xyz_max, displace, voxel_size, use_grad_norm):
        xyz = xyz.contiguous()
        mn, mx = _dev_vec(xyz_min, grid), _dev_vec(xyz_max, grid)
        disp = _dev_vec(list(displace), grid)
        nd = disp.numel()
        geom = grid_geom(grid)
        n = xyz.shape[0]
        feat = torch.empty((n, 6 * nd), dtype=grid.dtype, device=grid.device)
        grad = torch.empty((n, 3 * nd), dtype=grid.dtype, device=grid.device)
        call("vx_sample_sdfs_fwd", dtype_code(grid), ptr(grid), geom.data_ptr(), ptr(xyz),
             ptr(mn), ptr(mx), ptr(disp), nd, float(voxel_size), int(bool(use_grad_norm)), n,
             ptr(feat), ptr(grad), stream_of(grid))
        ctx.save_for_backward(grid, xyz, mn, mx, disp)
        ctx.geom = geom
        ctx.vs = float(voxel_size)
        ctx.norm = int(bool(use_grad_norm))
        return feat, grad

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gfeat, ggrad):
        grid, xyz, mn, mx, disp = ctx.saved_tensors
        ggrid = _zeros_like_grid(grid)
        c = lambda t: None if t is None else t.contiguous()  # noqa: E731
        call("vx_sample_sdfs_bwd", dtype_code(grid), ptr(ggrid), ptr(grid), ctx.geom.data_ptr(),
             ptr(xyz), ptr(mn), ptr(mx), ptr(disp), disp.numel(), ctx.vs, ctx.norm, xyz.shape[0],
             ptr(c(gfeat)), ptr(c(ggrad)), stream_of(grid))
        return ggrid, None, None, None, None, None, None


class DenseSdfGradient(torch.autograd.Function):
    """[1,1,X,Y,Z] -> [1,3,X,Y,Z] central differences, zero on boundary faces."""

    @staticmethod
    def forward(ctx, sdf, voxel_size):
        s = sdf.contiguous()
        X, Y, Z = s.shape[-3:]
        out = torch.empty((1, 3, X, Y, Z), dtype=s.dtype, device=s.device)
        call("vx_dense_sdf_gradient_fwd", dtype_code(s), ptr(s), X, Y, Z, float(voxel_size),
             ptr(out), stream_of(s))
        ctx.shape = (X, Y, Z)
        ctx.vs = float(voxel_size)
        ctx.like = s
        return out

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gout):
        X, Y, Z = ctx.shape
        gs = torch.zeros_like(ctx.like)
        gout = gout.contiguous()
        call("vx_dense_sdf_gradient_bwd", dtype_code(gs), ptr(gout), X, Y, Z, ctx.vs, ptr(gs),
             stream_of(gs))
        return gs, None


class NeusAlpha(torch.autograd.Function):
    @staticmethod
    def forward(ctx, viewdirs, ray_id, sdf, gradient, dist, inv_s):
        viewdirs, sdf, gradient = viewdirs.contiguous(), sdf.contiguous(), gradient.contiguous()
        ray_id = ray_id.contiguous()
        alpha = torch.empty_like(sdf)
        call("vx_neus_alpha_fwd", dtype_code(sdf), ptr(viewdirs), ptr(ray_id), ptr(sdf),
             ptr(gradient), sdf.numel(), float(dist), float(inv_s), ptr(alpha), stream_of(sdf))
        ctx.save_for_backward(viewdirs, ray_id, sdf, gradient)
        ctx.dist, ctx.inv_s = float(dist), float(inv_s)
        return alpha

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, galpha):
        viewdirs, ray_id, sdf, gradient = ctx.saved_tensors
        gsdf = torch.empty_like(sdf)
        ggrad = torch.empty_like(gradient)
        galpha = galpha.contiguous()
        call("vx_neus_alpha_bwd", dtype_code(sdf), ptr(viewdirs), ptr(ray_id), ptr(sdf),
             ptr(gradient), sdf.numel(), ctx.dist, ctx.inv_s, ptr(galpha), ptr(gsdf), ptr(ggrad),
             stream_of(sdf))
        return None, None, gsdf, ggrad, None, None


class Alphas2Weights(torch.autograd.Function):
    """lib/voxurf_fine.py:981-997, on the HIP alpha2weight kernels."""

    @staticmethod
    def forward(ctx, alpha, ray_id, N):
        weights, T, alphainv_last, i_start, i_end = ru.alpha2weight(alpha, ray_id, N)
        if alpha.requires_grad:
            ctx.save_for_backward(alpha, weights, T, alphainv_last, i_start, i_end)
            ctx.n_rays = N
        return weights, alphainv_last

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_weights, grad_last):
        alpha, weights, T, alphainv_last, i_start, i_end = ctx.saved_tensors
        grad = ru.alpha2weight_backward(alpha, weights, T, alphainv_last, i_start, i_end,
                                        ctx.n_rays, grad_weights.contiguous(),
                                        grad_last.contiguous())
        return grad, None, None


class SegmentSum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, src, index, n_out):
        src2 = src.reshape(src.shape[0], -1).contiguous()
        C = src2.shape[1]
        out = torch.zeros((n_out, C), dtype=src.dtype, device=src.device)
        call("vx_segment_sum", dtype_code(src), ptr(src2), ptr(index), src2.shape[0], C, n_out,
             ptr(out), stream_of(src))
        ctx.save_for_backward(index)
        ctx.shape = src.shape
        ctx.n_out = n_out
        return out.reshape((n_out,) + tuple(src.shape[1:]))

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gout):
        (index,) = ctx.saved_tensors
        g2 = gout.reshape(ctx.n_out, -1).contiguous()
        C = g2.shape[1]
        gsrc = torch.empty((ctx.shape[0], C), dtype=gout.dtype, device=gout.device)
        call("vx_segment_gather", dtype_code(gout), ptr(g2), ptr(index), ctx.shape[0], C,
             ctx.n_out, ptr(gsrc), stream_of(gout))
        return gsrc.reshape(ctx.shape), None, None


def segment_coo(src, index, out=None, reduce="sum", dim_size=None):
    """torch_scatter.segment_coo for reduce='sum' along dim 0 with a sorted index.

    With `out` given, the result is out + segment sums (torch_scatter accumulates into
    `out`); the reference always passes zeros.
    """
    if reduce != "sum":
        raise NotImplementedError("segment_coo: only reduce='sum' is used by Voxurf")
    if index.dtype != torch.int64:
        index = index.long()
    index = index.contiguous()
    n_out = out.shape[0] if out is not None else (dim_size if dim_size is not None else
                                                   (int(index[-1].item()) + 1 if index.numel() else 0))
    res = SegmentSum.apply(src, index, n_out)
    if out is not None:
        res = res + out.to(res.device)
    return res
