"""Pins of the oracle's dense per-step grid operators (CPU): the smooth_grad_tv term
against a brute-force loop (clamped neighbours, zero boundary faces, masked mean), its
gradient against central finite differences of the objective with the smoothed field
held fixed (the reference detaches it), the Gaussian and tv_smooth kernels against the
outer products of their 1-D factors (which the HIP kernels take), smoothing of constant
and linear fields, and trilinear resampling of a linear field."""
import itertools

import numpy as np
import torch

from oracle import torch_ref as tr
from voxurf_amd.smooth import gaussian_taps, tv_smooth_taps


def _brute_smooth_grad_tv(sdf, vs, mask, weight, sigma):
    s = sdf[0, 0].numpy().astype(np.float64)
    X, Y, Z = s.shape
    G = np.zeros((3, X, Y, Z))
    for x, y, z in itertools.product(range(X), range(Y), range(Z)):
        if 0 < x < X - 1:
            G[0, x, y, z] = (s[x + 1, y, z] - s[x - 1, y, z]) / 2 / vs
        if 0 < y < Y - 1:
            G[1, x, y, z] = (s[x, y + 1, z] - s[x, y - 1, z]) / 2 / vs
        if 0 < z < Z - 1:
            G[2, x, y, z] = (s[x, y, z + 1] - s[x, y, z - 1]) / 2 / vs
    w = tr.tv_smooth_weight(sigma, torch.float64)[0, 0].numpy()
    tot, cnt = 0.0, 0
    m = mask[0, 0].numpy()
    for c, x, y, z in itertools.product(range(3), range(X), range(Y), range(Z)):
        if not m[x, y, z]:
            continue
        sm = 0.0
        for a, b, d in itertools.product(range(3), range(3), range(3)):
            xx = min(max(x + a - 1, 0), X - 1)
            yy = min(max(y + b - 1, 0), Y - 1)
            zz = min(max(z + d - 1, 0), Z - 1)
            sm += w[a, b, d] * G[c, xx, yy, zz]
        tot += (sm - G[c, x, y, z]) ** 2
        cnt += 1
    return weight * tot / cnt


def _field(shape, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.randn((1, 1, *shape), generator=g, dtype=torch.float64)


def test_smooth_grad_tv_brute_force():
    sdf = _field((5, 4, 6), 0)
    mask = torch.rand((1, 1, 5, 4, 6), generator=torch.Generator().manual_seed(1)) > 0.3
    for sigma in (0.0, 0.5):
        got = float(tr.smooth_grad_tv(sdf, 0.25, mask, 0.7, sigma))
        want = _brute_smooth_grad_tv(sdf, 0.25, mask, 0.7, sigma)
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want))


def test_smooth_grad_tv_quadratic_interior_is_exact():
    # a quadratic SDF has a linear gradient field; the [1,2,1] smoothing reproduces it
    # exactly away from the faces, so only voxels near the boundary contribute
    x = torch.arange(8, dtype=torch.float64)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    sdf = (0.3 * X * X - 0.2 * Y * Z + 0.1 * Z)[None, None]
    interior = torch.zeros((1, 1, 8, 8, 8), dtype=torch.bool)
    interior[..., 2:-2, 2:-2, 2:-2] = True
    assert float(tr.smooth_grad_tv(sdf, 0.5, interior, 1.0)) < 1e-24
    full = torch.ones_like(interior)
    assert float(tr.smooth_grad_tv(sdf, 0.5, full, 1.0)) > 1e-3


def test_smooth_grad_tv_gradient_finite_differences():
    sdf = _field((6, 5, 7), 2).requires_grad_(True)
    mask = torch.rand((1, 1, 6, 5, 7), generator=torch.Generator().manual_seed(3)) > 0.4
    vs, weight = 0.3, 0.9
    loss = tr.smooth_grad_tv(sdf, vs, mask, weight)
    (grad,) = torch.autograd.grad(loss, sdf)
    # objective with the smoothed field frozen at the base point
    with torch.no_grad():
        g0 = tr.neus_sdf_gradient(sdf, vs).permute(1, 0, 2, 3, 4)
        sm0 = torch.nn.functional.conv3d(torch.nn.functional.pad(g0, (1,) * 6, mode="replicate"),
                                         tr.tv_smooth_weight(0.0, torch.float64))

    def frozen(s):
        g = tr.neus_sdf_gradient(s, vs).permute(1, 0, 2, 3, 4)
        e = (sm0 - g)[mask.repeat(3, 1, 1, 1, 1)] ** 2
        return float(e.mean() * weight)
    eps = 1e-6
    rng = np.random.default_rng(0)
    for _ in range(20):
        i = tuple(int(rng.integers(0, n)) for n in sdf.shape[2:])
        sp = sdf.detach().clone()
        sm = sdf.detach().clone()
        sp[0, 0][i] += eps
        sm[0, 0][i] -= eps
        fd = (frozen(sp) - frozen(sm)) / (2 * eps)
        assert abs(fd - float(grad[0, 0][i])) <= 1e-6 * max(1.0, abs(fd))


def test_kernels_are_outer_products_of_their_taps():
    for ksize, sigma in ((3, 1.0), (5, 0.8), (7, 1.3)):
        t = np.asarray(gaussian_taps(ksize, sigma))
        outer = np.einsum("i,j,k->ijk", t, t, t)
        w = tr.gaussian_3dconv_weight(ksize, sigma, torch.float64)[0, 0].numpy()
        assert np.allclose(outer, w, rtol=1e-13, atol=1e-16)
    for sigma in (0.0, 0.5):
        t = np.asarray(tv_smooth_taps(sigma))
        outer = np.einsum("i,j,k->ijk", t, t, t)
        assert np.allclose(outer, tr.tv_smooth_weight(sigma, torch.float64)[0, 0].numpy(),
                           rtol=1e-13, atol=1e-16)


def test_gaussian_smooth_constant_and_linear_fields():
    c = torch.full((1, 1, 6, 7, 5), 2.5, dtype=torch.float64)
    assert torch.allclose(tr.gaussian_smooth(c, 5, 0.8), c, rtol=1e-14)
    # a linear field is reproduced exactly away from the replicate-padded border
    x = torch.arange(9, dtype=torch.float64)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    lin = (0.5 * X - 0.25 * Y + 2 * Z)[None, None]
    out = tr.gaussian_smooth(lin, 5, 0.8)
    assert torch.allclose(out[..., 2:-2, 2:-2, 2:-2], lin[..., 2:-2, 2:-2, 2:-2], atol=1e-12)
    assert not torch.allclose(out[..., 0, :, :], lin[..., 0, :, :], atol=1e-3)


def test_scale_volume_grid_linear_field():
    x = torch.linspace(-1, 1, 5, dtype=torch.float64)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    f = lambda a, b, c: 0.3 * a - 0.7 * b + 0.2 * c + 0.1  # noqa: E731
    g = f(X, Y, Z)[None, None]
    out = tr.scale_volume_grid(g, (9, 7, 11))
    X2, Y2, Z2 = torch.meshgrid(torch.linspace(-1, 1, 9, dtype=torch.float64),
                                torch.linspace(-1, 1, 7, dtype=torch.float64),
                                torch.linspace(-1, 1, 11, dtype=torch.float64), indexing="ij")
    assert torch.allclose(out[0, 0], f(X2, Y2, Z2), atol=1e-12)
