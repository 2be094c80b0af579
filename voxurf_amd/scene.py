"""Seeded synthetic scenes and ray batches for the coarse (96^3) and fine (256^3)
training steps and the render server.

The scene is an analytic sphere SDF on the bbox [-1, 1]^3,
    sdf(p) = |p| - 0.5 + 0.02 sin(8x) sin(8y) sin(8z),
sampled on the lattice x_i = -1 + 2 i / (R - 1) (np.mgrid[-1:1:Rj], the lattice the
reference uses for its ball initialisation, lib/voxurf_fine.py:85-86), stored as a
[1, 1, R, R, R] float32 grid.  Feature/colour grids are N(0, 0.1^2).  Rays start
uniformly on the radius-3 sphere and aim at the origin with N(0, 0.05^2) jitter.
Every tensor is drawn from its own torch.Generator so the scene does not depend on the
order tensors are created in.
"""
import dataclasses
import math

import torch


@dataclasses.dataclass
class SceneConfig:
    R: int                      # grid resolution per axis
    feat_dim: int               # channels of the feature/colour grid
    n_rays: int                 # rays per step
    inv_s: float                # NeuS sharpness 1/s_val
    stepsize: float = 0.5       # samples per voxel along a ray (in voxels)
    fast_color_thres: float = 1e-4
    seed: int = 0
    mlp_width: int = 192
    mlp_depth: int = 4          # rgbnet_depth: Linear layers (hidden = depth - 1)
    view_pe: int = 4            # view-direction PE frequencies 2^0..2^(view_pe-1)
    lr_sdf: float = 5e-3
    lr_k0: float = 0.1
    lr_mlp: float = 3e-3
    tv_weight: float = 0.0
    bg: float = 0.0

    @property
    def voxel_size(self) -> float:
        # float32 value of 2 / R (bbox side 2, R voxels, the reference's num_voxels route)
        return float(torch.tensor(2.0 / self.R, dtype=torch.float32))

    @property
    def stepdist(self) -> float:
        return float(torch.tensor(self.stepsize, dtype=torch.float32) *
                     torch.tensor(self.voxel_size, dtype=torch.float32))

    @property
    def mlp_in(self) -> int:
        return self.feat_dim + 3 + 3 + 6 * self.view_pe


# Plain coarse step: 96^3 SDF + RGB grid, s = 64, TV 1e-3, Adam lr 0.1.
COARSE = SceneConfig(R=96, feat_dim=3, n_rays=8192, inv_s=64.0, lr_sdf=0.1, lr_k0=0.1,
                     lr_mlp=0.0, tv_weight=1e-3, mlp_depth=0)
# Plain fine step: 256^3 SDF + 12-channel features, rgbnet 42->192x3->3, s = 256.
FINE = SceneConfig(R=256, feat_dim=12, n_rays=8192, inv_s=256.0)

# Grid regularisers of the fine stage (configs/dtu_e2e/fine.py:34-45, run.py:617-622 and
# 641-658): on steps with tv_from < step < tv_end and step % tv_every == 0, the loss gains
# weight_tv_density * smooth_grad_tv * mean((S(G) - G)^2) over the non-empty voxels, and
# the SDF gradient gains the TV term of weight weight_tv_density * sdf_tv / N_rays * R/128
# (dense while step < tv_dense_before).
FINE_TV = dict(weight_tv_density=0.01, sdf_tv=0.1, smooth_grad_tv=0.05, tv_every=3, tv_from=0,
               tv_end=30000, tv_dense_before=20000)


def tv_active(tv, step):
    """run.py:617, 641: whether the regularisers apply at (1-based) step `step`."""
    return tv is not None and tv["tv_from"] < step < tv["tv_end"] and step % tv["tv_every"] == 0


def _gen(seed, tag):
    g = torch.Generator()
    g.manual_seed((seed * 1000003 + tag) & 0x7FFFFFFF)
    return g


def sdf_grid(R, seed=0, dtype=torch.float32):
    """[1,1,R,R,R] analytic perturbed-sphere SDF (computed in float64, then cast)."""
    x = torch.linspace(-1.0, 1.0, R, dtype=torch.float64)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    s = torch.sqrt(X * X + Y * Y + Z * Z) - 0.5 + 0.02 * torch.sin(8 * X) * torch.sin(8 * Y) * torch.sin(8 * Z)
    return s.to(dtype)[None, None].contiguous()


def feature_grid(R, C, seed=0, dtype=torch.float32):
    return (torch.randn((1, C, R, R, R), generator=_gen(seed, 11), dtype=torch.float32) * 0.1).to(dtype)


def rays(n, seed=0, dtype=torch.float32, radius=3.0, jitter=0.05):
    """-> rays_o [n,3], rays_d [n,3] (unit), viewdirs [n,3] (= rays_d)."""
    g = _gen(seed, 21)
    u = torch.randn((n, 3), generator=g, dtype=torch.float64)
    u = u / u.norm(dim=-1, keepdim=True)
    xi = torch.randn((n, 3), generator=g, dtype=torch.float64)
    d = -u + jitter * xi
    d = d / d.norm(dim=-1, keepdim=True)
    o = radius * u
    return o.to(dtype).contiguous(), d.to(dtype).contiguous(), d.to(dtype).contiguous()


def targets(n, seed=0, dtype=torch.float32):
    return torch.rand((n, 3), generator=_gen(seed, 31), dtype=torch.float32).to(dtype)


def mlp_params(in_dim, width, depth, seed=0, dtype=torch.float32):
    """rgbnet weights as nn.Linear's default init (Kaiming-uniform weights, uniform
    biases), last bias zero (lib/voxurf_fine.py:132-149).  Returns [(W [out,in], b)]."""
    g = _gen(seed, 41)
    dims = [in_dim] + [width] * (depth - 1) + [3]
    out = []
    for i in range(len(dims) - 1):
        fan_in = dims[i]
        bound = 1.0 / math.sqrt(fan_in)
        # kaiming_uniform_(a=sqrt(5)) gives bound = sqrt(6 / ((1 + 5) * fan_in)) = 1/sqrt(fan_in)
        W = (torch.rand((dims[i + 1], fan_in), generator=g) * 2 - 1) * bound
        b = (torch.rand((dims[i + 1],), generator=g) * 2 - 1) * bound
        if i == len(dims) - 2:
            b.zero_()
        out.append((W.to(dtype).contiguous(), b.to(dtype).contiguous()))
    return out
