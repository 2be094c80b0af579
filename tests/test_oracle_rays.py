"""Pins of the oracle's camera rays and ray pre-filter (CPU): known answers for the
pinhole model (identity and rotated poses, flips, inverse_y, pixel offset), the NDC
warp checked against the projection of points on the original ray, and the pre-filter
on an occupied ball, where the answer follows from the ray's distance to the centre."""
import numpy as np
import torch

from oracle import torch_ref as tr


def _K(f, cx, cy):
    return np.array([[f, 0, cx], [0, f, cy], [0, 0, 1]], dtype=np.float64)


def test_rays_identity_pose_known_answer():
    H, W, f = 5, 7, 4.0
    c2w = torch.eye(4)[:3]
    o, d = tr.get_rays(H, W, _K(f, 3.0, 2.0), c2w, False, False, False)
    for r in range(H):
        for c in range(W):
            want = [(c + 0.5 - 3.0) / f, -(r + 0.5 - 2.0) / f, -1.0]
            assert d[r, c].tolist() == [np.float32(x) for x in want]
    assert torch.equal(o, torch.zeros_like(o))
    # inverse_y: +y and +z camera axes; lefttop: no half-pixel offset
    _, d2 = tr.get_rays(H, W, _K(f, 3.0, 2.0), c2w, True, False, False, mode="lefttop")
    assert d2[1, 2].tolist() == [np.float32((2 - 3.0) / f), np.float32((1 - 2.0) / f), 1.0]


def test_rays_flips_mirror_pixels():
    H, W = 4, 6
    K = _K(3.0, 2.5, 1.5)
    c2w = torch.eye(4)[:3]
    _, d = tr.get_rays(H, W, K, c2w, False, False, False)
    _, dx = tr.get_rays(H, W, K, c2w, False, True, False)
    _, dy = tr.get_rays(H, W, K, c2w, False, False, True)
    assert torch.equal(dx, d.flip(1))
    assert torch.equal(dy, d.flip(0))


def test_rays_rotation_and_translation():
    # 90 degrees about z: camera x -> world y, camera y -> world -x
    c2w = torch.tensor([[0.0, -1.0, 0.0, 1.0], [1.0, 0.0, 0.0, 2.0], [0.0, 0.0, 1.0, 3.0]])
    K = _K(2.0, 1.0, 1.0)
    _, d0 = tr.get_rays(3, 3, K, torch.eye(4)[:3], False, False, False)
    o, d = tr.get_rays(3, 3, K, c2w, False, False, False)
    assert torch.equal(d[..., 0], -d0[..., 1])
    assert torch.equal(d[..., 1], d0[..., 0])
    assert torch.equal(d[..., 2], d0[..., 2])
    assert torch.equal(o[1, 1], torch.tensor([1.0, 2.0, 3.0]))
    # the principal-point pixel looks along -z of the camera, i.e. -c2w[:, 2]
    _, dp = tr.get_rays(3, 3, _K(2.0, 1.5, 1.5), c2w, False, False, False)
    assert torch.equal(dp[1, 1], -c2w[:, 2])


def test_viewdirs_unit_and_ndc_projection():
    H, W, f = 6, 8, 5.0
    K = _K(f, 4.0, 3.0)
    c2w = torch.eye(4)[:3].clone()
    c2w[:, 3] = torch.tensor([0.1, -0.2, 0.5])
    o, d, v = tr.get_rays_of_a_view(H, W, K, c2w, False, False, False, False)
    assert torch.allclose(v.norm(dim=-1), torch.ones(H, W), atol=1e-6)
    on, dn, vn = tr.get_rays_of_a_view(H, W, K, c2w, True, False, False, False)
    assert torch.equal(vn, v)
    # NDC origins lie on the near plane z = -1
    assert torch.allclose(on[..., 2], torch.full((H, W), -1.0), atol=1e-5)
    # the NDC ray passes through the NDC projection of points on the world ray
    o64, d64 = o.double(), d.double()
    for s in (1.5, 3.0, 7.0):
        p = o64 + s * d64
        proj = torch.stack([-f / (W / 2) * p[..., 0] / p[..., 2], -f / (H / 2) * p[..., 1] / p[..., 2],
                            1 + 2 * 1.0 / p[..., 2]], -1)
        t = (proj[..., 2] - on[..., 2].double()) / dn[..., 2].double()
        q = on.double() + t[..., None] * dn.double()
        assert torch.allclose(q, proj, atol=1e-4)
    # a planted sign error in the projection fails this check
    p = o64 + 3.0 * d64
    bad = torch.stack([f / (W / 2) * p[..., 0] / p[..., 2], -f / (H / 2) * p[..., 1] / p[..., 2],
                       1 + 2 / p[..., 2]], -1)
    t = (bad[..., 2] - on[..., 2].double()) / dn[..., 2].double()
    assert not torch.allclose(on.double() + t[..., None] * dn.double(), bad, atol=1e-4)


def _ball(R=32, radius=0.5):
    x = torch.linspace(-1, 1, R)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    dens = torch.where((X * X + Y * Y + Z * Z).sqrt() < radius, 5.0, -20.0)[None, None]
    return torch.nn.functional.max_pool3d(dens, kernel_size=3, padding=1, stride=1)


def test_hit_coarse_geo_ball():
    g = torch.Generator().manual_seed(1)
    n = 400
    u = torch.randn((n, 3), generator=g)
    u = u / u.norm(dim=-1, keepdim=True)
    o = 3 * u
    # aim at points at a random distance (0 .. 1.2) from the centre, perpendicular to u
    w = torch.randn((n, 3), generator=g)
    w = w - (w * u).sum(-1, keepdim=True) * u
    w = w / w.norm(dim=-1, keepdim=True)
    dist = torch.rand(n, generator=g) * 1.2
    d = (dist[:, None] * w - o)
    d = d / d.norm(dim=-1, keepdim=True)
    mn, mx = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
    hit = tr.hit_coarse_geo(o, d, mn, mx, 0.0, 0.5 * 2 / 64, _ball(), torch.tensor(mn),
                            torch.tensor(mx), -4.6, 1.0, 1e-3)
    # perpendicular distance of the ray to the centre
    perp = (o - (o * d).sum(-1, keepdim=True) * d).norm(dim=-1)
    assert bool(hit[perp < 0.45].all())
    assert not bool(hit[perp > 0.8].any())
    assert 0 < int(hit.sum()) < n
    # rays pointing away from the box never hit
    away = tr.hit_coarse_geo(o, u, mn, mx, 0.0, 0.5 * 2 / 64, _ball(), torch.tensor(mn),
                             torch.tensor(mx), -4.6, 1.0, 1e-3)
    assert not bool(away.any())
