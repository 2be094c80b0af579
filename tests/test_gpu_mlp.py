"""rgbnet on the matrix-core kernels vs nn.Linear chains (GPU).

FP32 path: logits within 1e-5 relative of a float64 CPU reference (floor 1e-6 of the
batch maximum), input/weight/bias gradients within 1e-4 relative (floor 1e-5).
BF16 path (BF16 MFMA operands, FP32 accumulation, BF16 stored activations): against a
float64 emulation that rounds to bfloat16 at exactly the points the kernels do (the
network input, the packed weights, each stored activation, the incoming gradient of each
data-gradient GEMM) every output and gradient is within 5e-3 of its largest entry (the
emulation cannot follow the FP32 summation order, so an activation near a bfloat16
rounding boundary can land one bfloat16 step away) -- a layout or indexing mistake gives
O(1) errors; the distance to plain float64 is printed (logits within 3e-2).
"""
import pytest
import torch
import torch.nn as nn

from voxurf_amd.mlp import RgbNet

pytestmark = pytest.mark.gpu


def ref_forward(x, net):
    h = x
    lins = net.linears()
    for i, l in enumerate(lins):
        h = h @ l.weight.T + l.bias
        if i < len(lins) - 1:
            h = torch.relu(h)
    return h


def bf(t):
    return t.to(torch.bfloat16).to(torch.float64)


def emulate_bf16(x, net, g):
    """float64 model of the BF16 kernels: logits and every gradient."""
    lins = net.linears()
    W = [l.weight.detach().double().cpu() for l in lins]
    b = [l.bias.detach().double().cpu() for l in lins]
    h = bf(x.detach().double().cpu())
    g = g.double().cpu()
    hs = [h]
    for i in range(len(lins) - 1):
        a = torch.relu(h @ bf(W[i]).T + b[i])
        h = bf(a)
        hs.append(h)
    logits = hs[-1] @ W[-1].T + b[-1]
    grads = [None] * (2 * len(lins))
    grads[-2], grads[-1] = g.T @ hs[-1], g.sum(0)
    dh = (g @ W[-1]) * (hs[-1] > 0)
    for li in range(len(lins) - 2, -1, -1):
        grads[2 * li], grads[2 * li + 1] = dh.T @ hs[li], dh.sum(0)
        d = bf(dh) @ bf(W[li])
        if li > 0:
            d = d * (hs[li] > 0)
        dh = d
    return logits, dh, grads


def rel_err(a, b):
    return float((a.double().cpu() - b).abs().max() / (b.abs().max() + 1e-30))


@pytest.mark.parametrize("din,M", [(42, 5000), (42, 31), (79, 4099), (42, 0)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_rgbnet_mfma(dev, din, M, dtype):
    torch.manual_seed(0)
    net = RgbNet(din, 192, 4, compute_dtype=dtype).to(dev)
    with torch.no_grad():
        for l in net.linears():
            l.bias.uniform_(-0.1, 0.1)
    x = torch.randn(M, din, device=dev).requires_grad_(True)
    out = net(x)
    g = torch.randn(M, 3, device=dev)
    out.backward(g)
    # float64 reference on the CPU
    ref = RgbNet(din, 192, 4).double()
    ref.load_state_dict({k: v.double().cpu() for k, v in net.state_dict().items()})
    xr = x.detach().double().cpu().requires_grad_(True)
    o_r = ref_forward(xr, ref)
    o_r.backward(g.double().cpu())
    if M == 0:
        assert out.shape == (0, 3)
        return
    pairs = [("x", x.grad, xr.grad)] + [
        (n, p.grad, q.grad) for (n, p), q in zip(net.named_parameters(), ref.parameters())]
    if dtype == torch.float32:
        torch.testing.assert_close(out.double().cpu(), o_r.detach(), rtol=1e-5,
                                   atol=1e-6 * float(o_r.abs().max()))
        for n, a, b in pairs:
            torch.testing.assert_close(a.double().cpu(), b, rtol=1e-4, atol=1e-5 * float(b.abs().max()),
                                       msg=lambda m: f"{n}: {m}")
    else:
        e = rel_err(out, o_r.detach())
        print(f"bf16 logits vs float64: rel err {e:.3g}")
        assert e < 3e-2
        for n, a, b in pairs:
            print(f"bf16 grad {n} vs float64: rel err {rel_err(a, b):.3g}")
        em_logits, em_dx, em_grads = emulate_bf16(x[:, :din], net, g)
        e = rel_err(out, em_logits)
        print(f"bf16 logits vs emulation: rel err {e:.3g}")
        assert e < 5e-3
        names = ["x"] + [n for n, _ in net.named_parameters()]
        for n, a, b in zip(names, [x.grad] + [p.grad for p in net.parameters()], [em_dx] + em_grads):
            e = rel_err(a, b)
            print(f"bf16 grad {n} vs emulation: rel err {e:.3g}")
            assert e < 5e-3, n
