"""rgbnet on the matrix-core kernels vs nn.Linear chains (GPU).

FP32 path: logits within 1e-5 relative of a float64 CPU reference (floor 1e-6 of the
batch maximum), input/weight/bias gradients within 1e-4 relative (floor 1e-5).
BF16 path (FP32 accumulation, BF16 activations): logits within 3e-2 of the largest
logit, gradients within 5e-2 of the largest gradient entry, against float64; the
measured errors are printed.
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
        print(f"bf16 logits rel err {e:.3g}")
        assert e < 3e-2
        for n, a, b in pairs:
            e = rel_err(a, b)
            print(f"bf16 grad {n} rel err {e:.3g}")
            assert e < 5e-2, n
