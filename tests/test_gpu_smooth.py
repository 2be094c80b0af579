"""Dense per-step grid operators on the GPU (csrc/smooth.hip) against the oracle's
restatement of lib/voxurf_fine.py:204-258, 412-421, 440-450 and lib/grid.py:60-65.

Float32: values within 1e-5 and gradients within 1e-4 of the largest entry; float64
within 1e-10.  Grid shapes are not multiples of the kernels' 4x8x64 tiles and include a
dimension of one voxel."""
import pytest
import torch

from oracle import torch_ref as tr
from voxurf_amd import smooth as sm
from voxurf_amd.grid import DenseGrid

pytestmark = pytest.mark.gpu

TOL = {torch.float32: (1e-5, 1e-4), torch.float64: (1e-10, 1e-10)}


def _rel(a, b):
    return float((a.cpu() - b).abs().max()) / max(1e-30, float(b.abs().max()))


def _field(shape, seed, dtype):
    g = torch.Generator().manual_seed(seed)
    x = torch.linspace(-1, 1, shape[0])[:, None, None]
    base = (x * 0.7 + torch.randn(shape, generator=g) * 0.05).to(dtype)
    return base[None, None].contiguous()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("shape,masked,sigma", [((37, 29, 70), True, 0.0),
                                                 ((9, 17, 130), False, 0.0),
                                                 ((12, 10, 8), True, 0.5)])
def test_smooth_grad_tv(dev, dtype, shape, masked, sigma):
    tv, tg = TOL[dtype]
    sdf = _field(shape, 0, dtype)
    vs, weight = 2.0 / 64, 0.05 * 0.01
    mask = None
    if masked:
        mask = (torch.rand((1, 1, *shape), generator=torch.Generator().manual_seed(1)) > 0.3)
    ref_mask = mask if mask is not None else torch.ones((1, 1, *shape), dtype=torch.bool)
    s_ref = sdf.clone().requires_grad_(True)
    loss_ref = tr.smooth_grad_tv(s_ref, vs, ref_mask, weight, sigma)
    (g_ref,) = torch.autograd.grad(loss_ref, s_ref)
    s = sdf.to(dev)
    # accumulates into an existing gradient (of the same magnitude, so the subtraction
    # below loses nothing)
    base = torch.randn(s.shape, generator=torch.Generator().manual_seed(5)).to(dtype) * \
        float(g_ref.abs().max())
    grad = base.to(dev)
    loss = torch.empty(1, dtype=dtype, device=dev)
    sm.smooth_grad_tv_add_grad(s, grad, vs, None if mask is None else mask.to(dev), weight,
                               sigma, loss=loss)
    assert abs(float(loss) - float(loss_ref.detach())) <= tv * abs(float(loss_ref.detach()))
    assert _rel(grad.cpu() - base, g_ref) <= tg
    # the autograd form gives the same gradient
    s2 = s.clone().requires_grad_(True)
    l2 = sm.smooth_grad_tv(s2, vs, None if mask is None else mask.to(dev), weight, sigma)
    l2.backward()
    assert abs(float(l2.detach()) - float(loss_ref.detach())) <= tv * abs(float(loss_ref.detach()))
    assert _rel(s2.grad, g_ref) <= tg


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("ksize,sigma,shape", [(5, 0.8, (13, 9, 17)), (3, 1.0, (5, 1, 6)),
                                               (7, 1.3, (20, 11, 3))])
def test_gaussian_conv3d(dev, dtype, ksize, sigma, shape):
    tv, tg = TOL[dtype]
    g = torch.Generator().manual_seed(2)
    x = torch.randn((1, 2, *shape), generator=g).to(dtype)
    gout = torch.randn((1, 2, *shape), generator=g).to(dtype)
    xr = x.clone().requires_grad_(True)
    ref = torch.cat([tr.gaussian_smooth(xr[:, c:c + 1], ksize, sigma) for c in range(2)], 1)
    (gref,) = torch.autograd.grad(ref, xr, gout)
    m = sm.GaussianConv3d(ksize, sigma)
    xd = x.to(dev).requires_grad_(True)
    out = m(xd)
    out.backward(gout.to(dev))
    assert _rel(out.detach(), ref.detach()) <= tv
    assert _rel(xd.grad, gref) <= tg


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("src,dst", [((7, 9, 5), (13, 17, 10)), ((20, 12, 9), (8, 12, 4)),
                                     ((6, 1, 5), (11, 3, 5))])
def test_resample_trilinear(dev, dtype, src, dst):
    tv, _ = TOL[dtype]
    g = torch.Generator().manual_seed(3)
    x = torch.randn((1, 3, *src), generator=g).to(dtype)
    ref = tr.scale_volume_grid(x, dst)
    out = sm.resample_trilinear(x.to(dev), dst)
    assert out.shape == ref.shape
    assert _rel(out, ref) <= tv


def test_dense_grid_scale_volume_grid(dev):
    g = torch.Generator().manual_seed(4)
    grid = DenseGrid(12, [10, 8, 6], [-1, -1, -1], [1, 1, 1]).to(dev)
    with torch.no_grad():
        grid.grid.copy_(torch.randn(grid.grid.shape, generator=g).to(dev))
    ref = tr.scale_volume_grid(grid.grid.detach().cpu(), (15, 11, 9))
    grid.scale_volume_grid([15, 11, 9])
    assert grid.grid.shape == (1, 12, 15, 11, 9)
    assert grid.grid.is_contiguous(memory_format=torch.channels_last_3d)
    assert _rel(grid.grid.detach(), ref) <= 1e-5


def test_fused_step_with_regularisers(dev):
    """A fine step with the regularisers of configs/dtu_e2e/fine.py (weight raised so
    the terms are not lost in the render gradient): the change they make to the SDF
    gradient and to the loss equals the CPU restatement's (run.py:617-622, 641-658)."""
    import dataclasses
    from voxurf_amd import scene
    from voxurf_amd.fused import FusedFineTrainer
    cfg = dataclasses.replace(scene.FINE, R=48, n_rays=256, seed=5)
    tv = dict(scene.FINE_TV, weight_tv_density=1.0, tv_every=1)
    st = tr.make_state(cfg)
    o, d, v = scene.rays(cfg.n_rays, seed=5)
    tg = scene.targets(cfg.n_rays, seed=5)
    x = torch.linspace(-1, 1, cfg.R)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    r = (X * X + Y * Y + Z * Z).sqrt()
    mask = ((r > 0.2) & (r < 0.8))[None, None]

    def oracle(t):
        s = {k: (val.clone() if torch.is_tensor(val) else [(w.clone(), b.clone()) for w, b in val])
             for k, val in st.items()}
        return tr.train_step(s, {}, o, d, v, tg, cfg, 1, tv=t, nonempty_mask=mask)

    def fused(t):
        f = FusedFineTrainer(cfg, st["sdf"], st["k0"], st["mlp"], device=dev)
        f.tv, f.nonempty_mask = t, mask
        out = f.step(o.to(dev), d.to(dev), v.to(dev), tg.to(dev), profile=True)
        return float(out["loss"]), f.grads_snapshot["sdf"].cpu()
    r_on, r_off = oracle(tv), oracle(None)
    (l_on, g_on), (l_off, g_off) = fused(tv), fused(None)
    dref = r_on["grads"]["sdf"] - r_off["grads"]["sdf"]
    assert float(dref.abs().max()) > 0
    # compared where the render loss leaves the SDF gradient at zero (in both runs):
    # elsewhere the difference of two runs mixes in the atomics' rounding of the render
    # gradient, which the fine-step parity test covers on its own
    sel = (g_off == 0) & (r_off["grads"]["sdf"] == 0)
    assert int(sel.sum()) > sel.numel() // 2
    err = float((g_on[sel] - dref[sel]).abs().max())
    assert err <= 1e-4 * float(dref.abs().max())
    lref = r_on["loss"] - r_off["loss"]
    assert abs((l_on - l_off) - lref) <= 1e-4 * abs(lref)
