"""MaskCache: the coarse-geometry occupancy test of the fine stage
(lib/voxurf_fine.py:917-942, the same as lib/dvgo_ori.py:440-473).

The DVGO coarse density grid is max-pooled (kernel ks, stride 1, same padding) once at
construction; forward(xyz) returns alpha >= thres with
alpha = 1 - exp(-softplus(density(xyz) + act_shift) * voxel_size_ratio), the density
looked up trilinearly (or at the nearest lattice point with nearest=True) on the HIP
kernel vx_maskcache_alpha.  Constructed from a coarse checkpoint path like the
reference (loaded with weights_only=True), or directly from tensors.
"""
import torch
import torch.nn as nn
import torch.nn.functional as F

from ._lib import call, dtype_code, ptr, stream_of
from .total_variation_cuda import grid_geom


class MaskCache(nn.Module):
    def __init__(self, path=None, mask_cache_thres=1e-3, ks=3, density=None, xyz_min=None,
                 xyz_max=None, act_shift=0.0, voxel_size_ratio=1.0, nearest=False):
        super().__init__()
        if path is not None:
            st = torch.load(path, weights_only=True, map_location="cpu")
            kw = st["MaskCache_kwargs"]
            xyz_min, xyz_max = kw["xyz_min"], kw["xyz_max"]
            act_shift, voxel_size_ratio = kw["act_shift"], kw["voxel_size_ratio"]
            nearest = kw.get("nearest", False)
            density = st["model_state_dict"]["density"]
        self.mask_cache_thres = mask_cache_thres
        self.register_buffer("xyz_min", torch.as_tensor(xyz_min, dtype=torch.float32).clone())
        self.register_buffer("xyz_max", torch.as_tensor(xyz_max, dtype=torch.float32).clone())
        self.register_buffer("density", F.max_pool3d(density.float(), kernel_size=ks,
                                                     padding=ks // 2, stride=1).contiguous())
        self.act_shift = float(act_shift)
        self.voxel_size_ratio = float(voxel_size_ratio)
        self.nearest = bool(nearest)

    @classmethod
    def from_pooled(cls, density, xyz_min, xyz_max, act_shift, voxel_size_ratio,
                    mask_cache_thres=1e-3, nearest=False):
        """A MaskCache whose density is already max-pooled (its own `density` buffer, as
        a model's state dict holds it under mask_cache.density)."""
        mc = cls(density=torch.zeros(1, 1, 1, 1, 1), xyz_min=xyz_min, xyz_max=xyz_max,
                 act_shift=act_shift, voxel_size_ratio=voxel_size_ratio,
                 mask_cache_thres=mask_cache_thres, nearest=nearest)
        mc.density = density.float().contiguous().clone()
        return mc

    @torch.no_grad()
    def lattice_mask(self, xyz_min, xyz_max, world_size):
        """Voxurf._set_nonempty_mask (lib/voxurf_fine.py:354-367): the mask at the lattice
        points of a grid of world_size over [xyz_min, xyz_max], [X, Y, Z] bool."""
        dev = self.density.device
        axes = [torch.linspace(float(a), float(b), int(n), device=dev)
                for a, b, n in zip(xyz_min, xyz_max, world_size)]
        pts = torch.stack(torch.meshgrid(*axes, indexing="ij"), -1)
        return self(pts)

    @torch.no_grad()
    def forward(self, xyz):
        shape = xyz.shape[:-1]
        pts = xyz.reshape(-1, 3).to(self.density.dtype).contiguous()
        out = torch.empty(pts.shape[0], dtype=torch.bool, device=pts.device)
        geom = grid_geom(self.density)  # host array: must outlive the call
        call("vx_maskcache_alpha", dtype_code(self.density), ptr(self.density),
             geom.data_ptr(), ptr(pts), ptr(self.xyz_min), ptr(self.xyz_max),
             self.act_shift, self.voxel_size_ratio, float(self.mask_cache_thres),
             int(self.nearest), pts.shape[0], ptr(out), stream_of(pts))
        return out.reshape(shape)
