"""Drop-in DenseGrid (lib/grid.py:38-89) on the HIP kernels.

Same constructor, parameter name (`grid`, shape [1, C, X, Y, Z]), buffers
(`xyz_min`, `xyz_max`) and methods (forward, scale_volume_grid,
total_variation_add_grad, get_dense_grid, __isub__, extra_repr), so checkpoints and
call sites are interchangeable.  One layout choice differs: a grid with more than one
channel is held channels-last in memory (torch.channels_last_3d), so the C values of
one voxel are contiguous and a trilinear lookup reads 8 short runs instead of 8*C
scattered words.  The tensor's shape, indexing and state-dict entry are unchanged.
"""
import torch
import torch.nn as nn

from . import total_variation_cuda
from .functional import grid_sample


def create_grid(type, **kwargs):
    """lib/grid.py:27-33 (TensoRFGrid is not provided: nothing in Voxurf builds one)."""
    if type == "DenseGrid":
        return DenseGrid(**kwargs)
    raise NotImplementedError(type)


def _layout(t, channels_last):
    if channels_last and t.shape[1] > 1:
        return t.contiguous(memory_format=torch.channels_last_3d)
    return t.contiguous()


class DenseGrid(nn.Module):
    def __init__(self, channels, world_size, xyz_min, xyz_max, channels_last=True, **kwargs):
        super().__init__()
        self.channels = channels
        self.world_size = world_size
        self.channels_last = channels_last
        self.register_buffer("xyz_min", torch.as_tensor(xyz_min, dtype=torch.float32).clone())
        self.register_buffer("xyz_max", torch.as_tensor(xyz_max, dtype=torch.float32).clone())
        ws = [int(v) for v in (world_size.tolist() if torch.is_tensor(world_size) else world_size)]
        self.grid = nn.Parameter(_layout(torch.zeros([1, channels, *ws]), channels_last))

    def _apply(self, fn, *args, **kwargs):
        # keep the memory layout across .to()/.cuda()
        out = super()._apply(fn, *args, **kwargs)
        if self.grid.shape[1] > 1 and self.channels_last and not self.grid.is_contiguous(
                memory_format=torch.channels_last_3d):
            self.grid.data = _layout(self.grid.data, True)
        return out

    def forward(self, xyz):
        """Trilinear lookup at world coordinates xyz [..., 3] -> [..., C] (squeezed if C == 1)."""
        shape = xyz.shape[:-1]
        out = grid_sample(self.grid, xyz.reshape(-1, 3), self.xyz_min, self.xyz_max)
        out = out.reshape(*shape, self.channels)
        if self.channels == 1:
            out = out.squeeze(-1)
        return out

    def scale_volume_grid(self, new_world_size):
        """Resample the grid to new_world_size (lib/grid.py:60-65: trilinear,
        align_corners=True) on the HIP kernel vx_resample_trilinear."""
        from .smooth import resample_trilinear
        if self.channels == 0:
            self.grid = nn.Parameter(torch.zeros([1, self.channels, *new_world_size],
                                                 device=self.grid.device))
        else:
            self.grid = nn.Parameter(_layout(resample_trilinear(self.grid.data, new_world_size),
                                             self.channels_last))
        self.world_size = torch.as_tensor(new_world_size)

    def total_variation_add_grad(self, wx, wy, wz, dense_mode, mask=None):
        """Add the TV gradient to grid.grad in place (lib/grid.py:67-78)."""
        if mask is None:
            total_variation_cuda.total_variation_add_grad(self.grid, self.grid.grad, wx, wy, wz,
                                                          dense_mode)
        else:
            mask = mask.detach()
            if self.grid.size(1) > 1 and mask.size() != self.grid.size():
                mask = mask.repeat(1, self.grid.size(1), 1, 1, 1)
            assert mask.size() == self.grid.size()
            mask = mask.float().to(self.grid.dtype)
            mask = torch.empty_like(self.grid).copy_(mask)  # the grid's layout
            total_variation_cuda.total_variation_add_grad_new(self.grid, self.grid.grad, mask, wx,
                                                              wy, wz, dense_mode)

    def get_dense_grid(self):
        return self.grid

    @torch.no_grad()
    def __isub__(self, val):
        self.grid.data -= val
        return self

    def extra_repr(self):
        ws = self.world_size.tolist() if torch.is_tensor(self.world_size) else list(self.world_size)
        return f"channels={self.channels}, world_size={ws}"
