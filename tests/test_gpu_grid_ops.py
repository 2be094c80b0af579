"""HIP grid / NeuS / compositing / TV / Adam ops vs the CPU restatement (GPU).

Tolerances: float32 forward values within 1e-5 relative (plus an absolute floor of
1e-6 times the batch's magnitude, for values that pass through zero); gradients
within 1e-4 relative on the same footing; float64 within 1e-10.  The 6-neighbour
gradient is a difference quotient of nearby samples, so its float32 reference is
itself noisy at ~1e-5 of its magnitude: it is held to 1e-4 of the batch magnitude.
"""
import numpy as np
import pytest
import torch

from oracle import torch_ref as tr
from voxurf_amd import adam_upd_cuda, functional as fn, total_variation_cuda as tvc

pytestmark = pytest.mark.gpu

F32, F64 = torch.float32, torch.float64


def close(a, b, rtol, floor=1e-6):
    a = a.detach().cpu().to(b.dtype)
    scale = float(b.abs().max()) if b.numel() else 1.0
    torch.testing.assert_close(a, b, rtol=rtol, atol=max(floor * scale, 1e-30))


def tol(dtype, grad=False):
    if dtype == F64:
        return 1e-10
    return 1e-4 if grad else 1e-5


def rand_grid(C, X, Y, Z, dtype, channels_last, seed=0):
    g = torch.Generator().manual_seed(seed)
    t = torch.randn((1, C, X, Y, Z), generator=g, dtype=torch.float64).to(dtype)
    return t, (t.contiguous(memory_format=torch.channels_last_3d) if channels_last else t)


def rand_pts(n, dtype, seed=1, spread=1.1):
    g = torch.Generator().manual_seed(seed)
    return ((torch.rand((n, 3), generator=g, dtype=torch.float64) * 2 - 1) * spread).to(dtype)


@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("C,cl", [(1, False), (12, False), (12, True), (3, True)])
def test_grid_sample_fwd_bwd(dev, dtype, C, cl):
    cpu, layout = rand_grid(C, 17, 12, 9, dtype, cl)
    pts = rand_pts(20000, dtype)
    mn = torch.tensor([-1.0, -1.0, -1.0], dtype=dtype)
    mx = torch.tensor([1.0, 1.0, 1.0], dtype=dtype)
    ref = tr.dense_grid_forward(cpu, pts, mn, mx).reshape(len(pts), C)
    gpu_grid = layout.to(dev).requires_grad_(True)
    out = fn.grid_sample(gpu_grid, pts.to(dev), mn.to(dev), mx.to(dev))
    close(out, ref, tol(dtype))
    gout = torch.randn(len(pts), C, dtype=torch.float64).to(dtype)
    out.backward(gout.to(dev))
    c = cpu.clone().requires_grad_(True)
    (tr.dense_grid_forward(c, pts, mn, mx).reshape(len(pts), C) * gout).sum().backward()
    assert gpu_grid.grad.stride() == gpu_grid.stride()
    close(gpu_grid.grad, c.grad, tol(dtype, True))


@pytest.mark.parametrize("dtype", [F32, F64])
def test_sdf_grad_sample(dev, dtype):
    cpu, _ = rand_grid(1, 20, 16, 24, dtype, False, seed=3)
    cpu = cpu.cumsum(2).cumsum(3).cumsum(4) * 0.01  # smooth field
    pts = rand_pts(20000, dtype, seed=4, spread=1.0)
    pts[:3] = torch.tensor([[-1, -1, -1], [1, 1, 1], [0.999, -1, 0.5]], dtype=dtype)
    mn = torch.tensor([-1.0] * 3, dtype=dtype)
    mx = torch.tensor([1.0] * 3, dtype=dtype)
    vs = 2.0 / 19
    sdf, grad, feat = tr.grid_sampler(pts, cpu, mn, mx, vs)
    g = cpu.to(dev).requires_grad_(True)
    s2, g2, f2 = fn.SdfGradSample.apply(g, pts.to(dev), mn.to(dev), mx.to(dev), vs)
    close(s2, sdf, tol(dtype))
    close(f2, feat, tol(dtype))
    close(g2, grad, tol(dtype, True))
    ws = [torch.randn_like(sdf), torch.randn_like(grad), torch.randn_like(feat)]
    ((s2 * ws[0].to(dev)).sum() + (g2 * ws[1].to(dev)).sum() + (f2 * ws[2].to(dev)).sum()).backward()
    c = cpu.clone().requires_grad_(True)
    a, b, d = tr.grid_sampler(pts, c, mn, mx, vs)
    ((a * ws[0]).sum() + (b * ws[1]).sum() + (d * ws[2]).sum()).backward()
    close(g.grad, c.grad, tol(dtype, True))


@pytest.mark.parametrize("use_norm", [False, True])
@pytest.mark.parametrize("dtype", [F32, F64])
def test_sample_sdfs_hierarchical(dev, dtype, use_norm):
    cpu, _ = rand_grid(1, 14, 18, 11, dtype, False, seed=5)
    cpu = cpu.cumsum(2).cumsum(4) * 0.02
    pts = rand_pts(5000, dtype, seed=6, spread=1.0)
    mn = torch.tensor([-1.0] * 3, dtype=dtype)
    mx = torch.tensor([1.0] * 3, dtype=dtype)
    disp = [0.5, 1.0, 1.5, 2.0]
    feat, grad = tr.sample_sdfs(pts, cpu, disp, mn, mx, 0.1, use_norm)
    g = cpu.to(dev).requires_grad_(True)
    f2, g2 = fn.SampleSdfs.apply(g, pts.to(dev), mn.to(dev), mx.to(dev), disp, 0.1, use_norm)
    close(f2, feat, tol(dtype))
    close(g2, grad, tol(dtype, True))
    wf, wg = torch.randn_like(feat), torch.randn_like(grad)
    ((f2 * wf.to(dev)).sum() + (g2 * wg.to(dev)).sum()).backward()
    c = cpu.clone().requires_grad_(True)
    a, b = tr.sample_sdfs(pts, c, disp, mn, mx, 0.1, use_norm)
    ((a * wf).sum() + (b * wg).sum()).backward()
    close(g.grad, c.grad, tol(dtype, True))


@pytest.mark.parametrize("dtype", [F32, F64])
def test_dense_sdf_gradient(dev, dtype):
    cpu, _ = rand_grid(1, 9, 10, 11, dtype, False, seed=7)
    ref = tr.neus_sdf_gradient(cpu, 0.2)
    g = cpu.to(dev).requires_grad_(True)
    out = fn.DenseSdfGradient.apply(g, 0.2)
    close(out, ref, tol(dtype))
    w = torch.randn_like(ref)
    (out * w.to(dev)).sum().backward()
    c = cpu.clone().requires_grad_(True)
    (tr.neus_sdf_gradient(c, 0.2) * w).sum().backward()
    close(g.grad, c.grad, tol(dtype, True))


@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("C,cl", [(1, False), (12, False), (12, True)])
@pytest.mark.parametrize("dense", [True, False])
def test_total_variation(dev, dtype, C, cl, dense):
    cpu, layout = rand_grid(C, 13, 9, 10, dtype, cl, seed=8)
    grad = torch.randn_like(cpu) * (torch.rand_like(cpu) > 0.5)
    ref = grad.clone()
    tr.tv_add_grad(cpu, ref, 0.3, 0.5, 0.7, dense)
    p = layout.to(dev)
    g = torch.empty_like(p).copy_(grad.to(dev))
    tvc.total_variation_add_grad(p, g, 0.3, 0.5, 0.7, dense)
    close(g, ref, tol(dtype))
    mask = (torch.rand_like(cpu) > 0.3).to(dtype)
    ref = grad.clone()
    tr.tv_add_grad(cpu, ref, 0.3, 0.5, 0.7, dense, mask=mask)
    g = torch.empty_like(p).copy_(grad.to(dev))
    m = torch.empty_like(p).copy_(mask.to(dev))
    tvc.total_variation_add_grad_new(p, g, m, 0.3, 0.5, 0.7, dense)
    close(g, ref, tol(dtype))


@pytest.mark.parametrize("dtype", [F32, F64])
def test_adam_kernels(dev, dtype):
    gen = torch.Generator().manual_seed(9)
    n = 100003  # not a multiple of 4: exercises the vector body and the scalar tail
    p = torch.randn(n, generator=gen, dtype=F64).to(dtype)
    m = torch.randn(n, generator=gen, dtype=F64).to(dtype) * 0.1
    v = torch.rand(n, generator=gen, dtype=F64).to(dtype) * 0.01
    gr = torch.randn(n, generator=gen, dtype=F64).to(dtype)
    gr[::7] = 0
    perlr = torch.rand(n, generator=gen, dtype=F64).to(dtype)
    for step in (1, 5):
        for per in (None, perlr):
            P, M, V = p.clone(), m.clone(), v.clone()
            tr.adam_step(P, gr, M, V, step, 0.1, 0.9, 0.99, 1e-8, per_lr=per)
            Pd, Md, Vd, Gd = p.to(dev), m.to(dev), v.to(dev), gr.to(dev)
            from voxurf_amd._lib import call, dtype_code, ptr, stream_of
            call("vx_adam_step", dtype_code(Pd), ptr(Pd), ptr(Gd), ptr(Md), ptr(Vd),
                 ptr(None if per is None else per.to(dev)), n, step, 0.1, 0.9, 0.99, 1e-8, 1,
                 None, stream_of(Pd))
            close(Pd, P, tol(dtype))
            close(Md, M, tol(dtype))
            close(Vd, V, tol(dtype))
            assert int(torch.count_nonzero(Gd)) == 0  # gradient cleared in the same pass
    # the reference API has no masked per-lr variant (adam_upd.cpp:79-86)
    for masked, per in ((False, None), (True, None), (False, perlr)):
        if True:
            P, M, V = p.clone(), m.clone(), v.clone()
            tr.adam_upd(P, gr, M, V, 3, 0.9, 0.99, 0.05, 1e-8, masked=masked, perlr=per)
            Pd, Md, Vd = p.to(dev), m.to(dev), v.to(dev)
            if per is not None:
                adam_upd_cuda.adam_upd_with_perlr(Pd, gr.to(dev), Md, Vd, per.to(dev), 3, 0.9, 0.99, 0.05, 1e-8)
            elif masked:
                adam_upd_cuda.masked_adam_upd(Pd, gr.to(dev), Md, Vd, 3, 0.9, 0.99, 0.05, 1e-8)
            else:
                adam_upd_cuda.adam_upd(Pd, gr.to(dev), Md, Vd, 3, 0.9, 0.99, 0.05, 1e-8)
            close(Pd, P, tol(dtype))
            close(Md, M, tol(dtype))


@pytest.mark.parametrize("dtype", [F32, F64])
def test_neus_alpha(dev, dtype):
    gen = torch.Generator().manual_seed(10)
    N, n = 300, 50000
    v = torch.randn(N, 3, generator=gen, dtype=F64)
    v = (v / v.norm(dim=-1, keepdim=True)).to(dtype)
    rid = torch.randint(0, N, (n,), generator=gen).sort().values
    sdf = (torch.randn(n, generator=gen, dtype=F64) * 0.02).to(dtype)
    grad = torch.randn(n, 3, generator=gen, dtype=F64).to(dtype)
    ref = tr.neus_alpha(v, rid, 0.0039, sdf, grad, 256.0)
    s_d = sdf.to(dev).requires_grad_(True)
    g_d = grad.to(dev).requires_grad_(True)
    a = fn.NeusAlpha.apply(v.to(dev), rid.to(dev), s_d, g_d, 0.0039, 256.0)
    close(a, ref, tol(dtype))
    w = torch.randn(n, generator=gen, dtype=F64).to(dtype)
    (a * w.to(dev)).sum().backward()
    s_c = sdf.clone().requires_grad_(True)
    g_c = grad.clone().requires_grad_(True)
    (tr.neus_alpha(v, rid, 0.0039, s_c, g_c, 256.0) * w).sum().backward()
    close(s_d.grad, s_c.grad, tol(dtype, True))
    close(g_d.grad, g_c.grad, tol(dtype, True))


@pytest.mark.parametrize("dtype", [F32, F64])
def test_segment_sum_and_gather(dev, dtype):
    gen = torch.Generator().manual_seed(11)
    N = 1000
    lens = torch.randint(0, 200, (N,), generator=gen)
    idx = torch.repeat_interleave(torch.arange(N), lens)
    src = torch.randn(len(idx), 3, generator=gen, dtype=F64).to(dtype)
    ref = tr.segment_sum(src, idx, N)
    s = src.to(dev).requires_grad_(True)
    out = fn.segment_coo(s, idx.to(dev), out=torch.zeros(N, 3, device=dev, dtype=dtype))
    close(out, ref, tol(dtype))
    go = torch.randn(N, 3, dtype=dtype)
    out.backward(go.to(dev))
    close(s.grad, go[idx], 0)
    # 1-channel source
    out1 = fn.segment_coo(src[:, 0].to(dev), idx.to(dev), out=torch.zeros(N, device=dev, dtype=dtype))
    close(out1, tr.segment_sum(src[:, 0], idx, N), tol(dtype))


def test_alphas2weights_function(dev):
    from oracle import alpha2weight, alpha2weight_backward
    gen = torch.Generator().manual_seed(12)
    N = 500
    lens = torch.randint(0, 100, (N,), generator=gen)
    idx = torch.repeat_interleave(torch.arange(N), lens)
    alpha = torch.rand(len(idx), generator=gen) ** 4
    a = alpha.to(dev).requires_grad_(True)
    w, last = fn.Alphas2Weights.apply(a, idx.to(dev), N)
    rw, rT, rl, ris, rie = alpha2weight(alpha.numpy(), idx.numpy(), N)
    np.testing.assert_array_equal(w.detach().cpu().numpy(), rw)
    np.testing.assert_array_equal(last.detach().cpu().numpy(), rl)
    gw, gl = torch.randn(len(idx)), torch.randn(N)
    torch.autograd.backward([w, last], [gw.to(dev), gl.to(dev)])
    rg = alpha2weight_backward(alpha.numpy(), rw, rT, rl, ris, rie, N, gw.numpy(), gl.numpy())
    close(a.grad, torch.from_numpy(rg), 1e-4)
