"""The page-skipping Adam passes against the dense one (GPU), bit for bit: pages never
touched with an all-zero gradient are left alone (their utils.Adam update is the
identity: m = v = g = 0), every other page is updated with the dense pass's arithmetic;
pages with a nonzero gradient become touched.  Sizes include a partial last page and a
tail of n % 4 elements."""
import pytest
import torch

from voxurf_amd._lib import call, lib, ptr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pages,extra", [(37, 0), (37, 125), (5, 6), (300, 2)])
def test_adam_step_touched_equals_dense(dev, pages, extra):
    P = int(lib().vx_adam_page_elems())
    n = P * pages + extra
    g = torch.Generator().manual_seed(n)
    npg = (n + P - 1) // P
    p = torch.randn(n, generator=g)
    m = torch.randn(n, generator=g) * 0.1
    v = torch.rand(n, generator=g) * 0.01
    gr = torch.randn(n, generator=g)
    touched = (torch.rand(npg, generator=g) < 0.4).int()
    zero_grad = torch.rand(npg, generator=g) < 0.6
    for k in range(npg):
        sl = slice(P * k, min(n, P * (k + 1)))
        if not touched[k]:
            m[sl] = 0
            v[sl] = 0
        if zero_grad[k]:
            gr[sl] = 0
    a = [t.clone().to(dev) for t in (p, gr, m, v)]
    b = [t.clone().to(dev) for t in (p, gr, m, v)]
    tch = touched.to(dev)
    for step in (1, 7):
        call("vx_adam_step", 0, ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(a[3]), None, n, step, 0.1,
             0.9, 0.99, 1e-8, 1, None, None)
        call("vx_adam_step_touched", ptr(b[0]), ptr(b[1]), ptr(b[2]), ptr(b[3]), None, n, step,
             0.1, 0.9, 0.99, 1e-8, ptr(tch), None)
        torch.cuda.synchronize()
        for x, y, name in zip(a, b, ("param", "grad", "exp_avg", "exp_avg_sq")):
            assert torch.equal(x, y), (name, step)
        # the pages that had a gradient are touched now
        want = (touched.bool() | ~zero_grad).int()
        assert torch.equal(tch.cpu(), want)
        # next step: fresh gradients on a different set of pages
        gr2 = torch.where((torch.rand(npg, generator=g) < 0.3).repeat_interleave(P)[:n],
                          torch.randn(n, generator=g), torch.zeros(n))
        a[1].copy_(gr2)
        b[1].copy_(gr2)
        touched = tch.cpu().clone()
        zero_grad = torch.tensor([not bool(gr2[P * k:P * (k + 1)].any()) for k in range(npg)])
