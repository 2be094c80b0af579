"""Camera rays and the MaskCache ray pre-filter on the GPU (csrc/rays.hip) against the
oracle's restatement of lib/voxurf_fine.py:579-591, 1001-1068, 1127-1164.

Rays are float work: within 1e-5 of the largest component.  The hit test is a boolean
decision per ray: it equals the oracle's except for rays whose decision changes when the
mask threshold moves by 1e-6 relative (exp/log1p rounding in the MaskCache alpha).  The
per-view path (rays made inside the kernel) must equal hit_coarse_geo on the rays that
get_rays_of_a_view returns, bit for bit, and the training-ray compaction must equal
boolean indexing of those rays by those hits."""
import numpy as np
import pytest
import torch

from oracle import torch_ref as tr
from voxurf_amd import rays as vr
from voxurf_amd.maskcache import MaskCache

pytestmark = pytest.mark.gpu


def _pose(seed, radius=3.0):
    """A camera on a sphere looking at the origin (c2w [3, 4], OpenGL axes)."""
    g = torch.Generator().manual_seed(seed)
    u = torch.randn(3, generator=g)
    u = u / u.norm()
    eye = radius * u
    fwd = -u
    up = torch.tensor([0.0, 0.0, 1.0])
    right = torch.linalg.cross(fwd, up)
    right = right / right.norm()
    up2 = torch.linalg.cross(right, fwd)
    return torch.stack([right, up2, -fwd, eye], 1).float()


def _K(H, W, f):
    return np.array([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1]], dtype=np.float64)


def _close(a, b, tol=1e-5):
    scale = max(1.0, float(b.abs().max()))
    return float((a.cpu() - b).abs().max()) <= tol * scale


@pytest.mark.parametrize("ndc,inverse_y,flip_x,flip_y,mode", [
    (False, False, False, False, "center"),
    (False, True, True, False, "lefttop"),
    (True, False, False, True, "center"),
])
def test_rays_of_a_view(dev, ndc, inverse_y, flip_x, flip_y, mode):
    H, W = 37, 53
    K, c2w = _K(H, W, 45.0), _pose(3)
    o, d, v = vr.get_rays_of_a_view(H, W, K, c2w.to(dev), ndc, inverse_y, flip_x, flip_y, mode)
    ro, rd, rv = tr.get_rays_of_a_view(H, W, K, c2w, ndc, inverse_y, flip_x, flip_y, mode)
    assert o.shape == (H, W, 3)
    assert _close(o, ro) and _close(d, rd) and _close(v, rv)
    # ndc_rays on its own agrees with the warp inside get_rays_of_a_view
    if ndc:
        o0, d0 = vr.get_rays(H, W, K, c2w.to(dev), inverse_y, flip_x, flip_y, mode)
        on, dn = vr.ndc_rays(H, W, K[0][0], 1.0, o0, d0)
        assert torch.equal(on, o) and torch.equal(dn, d)


def _ball_cache(dev, R=32, radius=0.5):
    x = torch.linspace(-1, 1, R)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    dens = torch.where((X * X + Y * Y + Z * Z).sqrt() < radius, 5.0, -20.0)[None, None]
    mc = MaskCache(density=dens, xyz_min=[-1, -1, -1], xyz_max=[1, 1, 1], act_shift=-4.6,
                   voxel_size_ratio=1.0, mask_cache_thres=1e-3).to(dev)
    return mc, dens


def _oracle_hits(mc, o, d, stepdist, near=0.0):
    mn, mx = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
    return tr.hit_coarse_geo(o.cpu(), d.cpu(), mn, mx, near, stepdist, mc.density.cpu(),
                             torch.tensor(mn), torch.tensor(mx), mc.act_shift,
                             mc.voxel_size_ratio, mc.mask_cache_thres, tie=1e-6)


def test_hit_coarse_geo_matches_oracle(dev):
    from voxurf_amd import scene
    mc, _ = _ball_cache(dev)
    o, d, _ = scene.rays(3000, seed=7, jitter=0.3)
    stepdist = 0.5 * 2 / 64
    hit = vr.hit_coarse_geo(mc, mc.xyz_min, mc.xyz_max, o.to(dev), d.to(dev), 0.0, stepdist).cpu()
    strict, loose = _oracle_hits(mc, o, d, stepdist)
    print("hits", int(hit.sum()), "of", hit.numel(), "ties", int((strict != loose).sum()))
    assert bool((strict <= hit).all()) and bool((hit <= loose).all())
    assert 0 < int(hit.sum()) < hit.numel()
    # empty input
    e = vr.hit_coarse_geo(mc, mc.xyz_min, mc.xyz_max, o[:0].to(dev), d[:0].to(dev), 0.0, stepdist)
    assert e.shape == (0,)


def test_training_rays_in_maskcache_sampling(dev):
    mc, _ = _ball_cache(dev)

    class Model:
        pass
    model = Model()
    model.mask_cache, model.xyz_min, model.xyz_max = mc, mc.xyz_min, mc.xyz_max
    model.voxel_size = 2 / 64
    kw = {"near": 0.0, "stepsize": 0.5}
    HW = [(40, 60), (33, 47), (25, 30)]
    Ks = [_K(H, W, 50.0) for H, W in HW]
    poses = [_pose(s) for s in (1, 2, 5)]
    g = torch.Generator().manual_seed(0)
    imgs = [torch.rand((H, W, 3), generator=g).to(dev) for H, W in HW]
    # the third view looks away from the scene: nothing is kept from it
    poses[2] = torch.cat([-poses[2][:, :3], poses[2][:, 3:]], 1)
    out = vr.get_training_rays_in_maskcache_sampling(imgs, poses, HW, Ks, False, False, False,
                                                     False, model, kw)
    rgb_tr, o_tr, d_tr, v_tr, imsz = out
    assert imsz[2] == 0 and imsz[0] > 0 and imsz[1] > 0
    top = 0
    for img, (H, W), K, c2w, n in zip(imgs, HW, Ks, poses, imsz):
        o, d, v = vr.get_rays_of_a_view(H, W, K, c2w.to(dev), False, False, False, False)
        hit = vr.hit_coarse_geo(mc, mc.xyz_min, mc.xyz_max, o, d, kw["near"], 0.5 * 2 / 64)
        assert int(hit.sum()) == n
        assert torch.equal(rgb_tr[top:top + n], img[hit])
        assert torch.equal(o_tr[top:top + n], o[hit])
        assert torch.equal(d_tr[top:top + n], d[hit])
        assert torch.equal(v_tr[top:top + n], v[hit])
        strict, loose = _oracle_hits(mc, o.reshape(-1, 3), d.reshape(-1, 3), 0.5 * 2 / 64)
        h = hit.reshape(-1).cpu()
        assert bool((strict <= h).all()) and bool((h <= loose).all())
        top += n
    assert rgb_tr.shape[0] == top
