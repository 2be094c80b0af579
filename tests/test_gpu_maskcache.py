"""MaskCache occupancy test (GPU) against the CPU restatement of
lib/voxurf_fine.py:917-942: the boolean mask equals (alpha >= thres) computed with
F.grid_sample / softplus on the CPU, except where alpha lies within 1e-6 relative of the
threshold (exp/log1p rounding); trilinear and nearest lookups; loading from a
coarse-stage checkpoint dictionary (weights_only)."""
import io

import pytest
import torch
import torch.nn.functional as F

from oracle import torch_ref as tr
from voxurf_amd.maskcache import MaskCache

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nearest", [False, True])
def test_maskcache_matches_restatement(dev, nearest):
    g = torch.Generator().manual_seed(0)
    dens = torch.randn((1, 1, 20, 17, 23), generator=g) * 4 - 3
    mn, mx = torch.tensor([-1.0, -0.9, -1.1]), torch.tensor([1.0, 1.2, 0.8])
    act_shift, ratio, thres = -4.6, 0.5, 1e-3
    ckpt = {"MaskCache_kwargs": {"xyz_min": mn, "xyz_max": mx, "act_shift": act_shift,
                                 "voxel_size_ratio": ratio, "nearest": nearest},
            "model_state_dict": {"density": dens}}
    buf = io.BytesIO()
    torch.save(ckpt, buf)
    buf.seek(0)
    path = "/tmp/_maskcache_ckpt.tar"
    with open(path, "wb") as f:
        f.write(buf.getvalue())
    mc = MaskCache(path=path, mask_cache_thres=thres).to(dev)
    pts = (torch.rand((200000, 3), generator=g) * 2.4 - 1.2)
    out = mc(pts.to(dev)).cpu()
    pooled = F.max_pool3d(dens, kernel_size=3, padding=1, stride=1)
    alpha = tr.maskcache_alpha(pooled, pts, mn, mx, act_shift, ratio, nearest)
    ref = alpha >= thres
    diff = out != ref
    near_tie = (alpha - thres).abs() <= 1e-6 * thres
    print("mask true", int(ref.sum()), "differences", int(diff.sum()))
    assert not bool((diff & ~near_tie).any())


def test_fused_path_with_maskcache_equals_op_path(dev):
    """The fine step with a MaskCache: the fused kernels drop rejected samples exactly
    where the op-by-op model's compaction does (same survivors, weights, colours)."""
    import dataclasses
    from voxurf_amd import scene
    from voxurf_amd.fused import FusedFineTrainer
    from voxurf_amd.model import VoxurfFine
    cfg = dataclasses.replace(scene.FINE, R=64, n_rays=2048)
    # coarse density: occupied inside radius 0.55 only (a loose shell around the sphere)
    x = torch.linspace(-1, 1, 32)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    dens = torch.where((X * X + Y * Y + Z * Z).sqrt() < 0.55, 5.0, -20.0)[None, None]
    mc = MaskCache(density=dens, xyz_min=[-1, -1, -1], xyz_max=[1, 1, 1], act_shift=-4.6,
                   voxel_size_ratio=1.0, mask_cache_thres=1e-3).to(dev)
    sdf, k0 = scene.sdf_grid(cfg.R), scene.feature_grid(cfg.R, 12)
    mlp = scene.mlp_params(cfg.mlp_in, 192, 4)
    o, d, v = (t.to(dev) for t in scene.rays(cfg.n_rays, seed=4))
    f = FusedFineTrainer(cfg, sdf, k0, mlp, device=dev)
    f.mask_cache = mc
    m = f.forward(o, d, v)
    model = VoxurfFine(cfg, sdf, k0, mlp).to(dev)
    model.mask_cache = mc
    with torch.no_grad():
        out = model(o, d, v)
    assert m == out["n_samples"]
    assert torch.equal(f.rid[:m].long(), out["ray_id"])
    assert torch.equal(f.P4[:m, 3], out["weights"])
    # the mask removed samples: fewer in-box samples processed than without it
    f2 = FusedFineTrainer(cfg, sdf, k0, mlp, device=dev)
    f2.forward(o, d, v)
    assert int(f.stats[:, 2].sum()) < int(f2.stats[:, 2].sum())
