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
