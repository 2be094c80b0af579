"""Host side of the rgbnet MLP on the matrix-core kernels (csrc/mlp.hip).

mlp_forward(x, linears, compute_dtype) -> logits [M, 3], differentiable w.r.t. x and
every Linear's weight and bias.  Shapes: width 192, at least one hidden layer of
192 x 192, input width <= 96 (the plain fine step's rgbnet takes 42 inputs, the
reference's fine rgbnet 79).  compute_dtype float32 runs FP32 MFMAs; bfloat16 runs the
forward and data-gradient GEMMs as BF16 MFMAs with FP32 accumulation, stores the
activations in BF16 and keeps weights, biases, gradients and the optimiser in FP32.
"""
import torch

from ._lib import call, ptr, stream_of, lib

WIDTH = 192


def _pad_width(d):
    for w in (64, 96):
        if d <= w:
            return w
    raise RuntimeError(f"rgbnet input width {d} > 96 is not supported by the MFMA kernels")


def _pack(W, kt, nto, transpose, bf16, st):
    n_out, k_in = W.shape
    n = nto * kt * 16 * 64
    out = torch.empty(n, dtype=torch.bfloat16 if bf16 else torch.float32, device=W.device)
    call("vx_mlp_pack", int(bf16), ptr(W), n_out, k_in, kt, nto, int(transpose), ptr(out), st)
    return out


class _MlpFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, bf16, *params):
        Ws = [p.detach().contiguous() for p in params[0::2]]
        bs = [p.detach().contiguous() for p in params[1::2]]
        if any(w.dtype != torch.float32 for w in Ws):
            raise RuntimeError("rgbnet parameters must be float32")
        M, din = x.shape
        dev = x.device
        st = stream_of(x)
        kin = _pad_width(Ws[0].shape[1])
        kt0 = kin // 32
        adt = torch.bfloat16 if bf16 else torch.float32
        if din == kin and x.is_contiguous():
            # already padded to the tile width (columns past in_features meet zero
            # weights in the packed operand)
            X = x.detach().to(adt)
        else:
            X = torch.nn.functional.pad(x.detach(), (0, kin - din)).to(adt).contiguous()
        hs = []
        h_in, kt = X, kt0
        nlin = len(Ws)
        logits = torch.empty((M, 3), dtype=torch.float32, device=dev)
        for li in range(nlin - 1):
            last = li == nlin - 2
            pk = _pack(Ws[li], kt, 6, False, bf16, st)
            H = torch.empty((M, WIDTH), dtype=adt, device=dev)
            call("vx_mlp_layer_fwd", int(bf16), kt, int(last), ptr(h_in), h_in.shape[1], M, ptr(pk),
                 ptr(bs[li]), ptr(H), WIDTH, ptr(Ws[-1]) if last else None,
                 ptr(bs[-1]) if last else None, ptr(logits) if last else None, st)
            hs.append(H)
            h_in, kt = H, 6
        ctx.save_for_backward(X, *hs)
        ctx.Ws = Ws
        ctx.bf16 = bf16
        ctx.din = din
        ctx.kt0 = kt0
        return logits

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, glogits):
        X, *hs = ctx.saved_tensors
        Ws, bf16, din, kt0 = ctx.Ws, ctx.bf16, ctx.din, ctx.kt0
        M = X.shape[0]
        dev = X.device
        st = stream_of(X)
        g = glogits.contiguous().float()
        nlin = len(Ws)
        grads = [None] * (2 * nlin)
        part = torch.empty(lib().vx_mlp_wgrad_part_elems(M), dtype=torch.float32, device=dev)
        partb = torch.empty(lib().vx_mlp_wgrad_partb_elems(M), dtype=torch.float32, device=dev)
        # output layer (K = 3) and the gradient of the last hidden activations
        dW = torch.empty_like(Ws[-1])
        db = torch.empty(3, dtype=torch.float32, device=dev)
        dH = torch.empty((M, WIDTH), dtype=torch.float32, device=dev)
        call("vx_mlp_last_bwd", int(bf16), ptr(g), ptr(hs[-1]), WIDTH, M, ptr(Ws[-1]), ptr(dH),
             WIDTH, ptr(part), ptr(partb), ptr(dW), ptr(db), st)
        grads[2 * (nlin - 1)], grads[2 * (nlin - 1) + 1] = dW, db
        dX = None
        for li in range(nlin - 2, -1, -1):
            h_prev = hs[li - 1] if li > 0 else X
            kt = 6 if li > 0 else kt0
            # weight gradient of layer li: dY = dH (its output grad), X = its input
            dW = torch.empty_like(Ws[li])
            db = torch.empty(WIDTH, dtype=torch.float32, device=dev)
            call("vx_mlp_wgrad", int(bf16), kt, Ws[li].shape[1], ptr(dH), WIDTH, ptr(h_prev),
                 h_prev.shape[1], M, ptr(part), ptr(partb), ptr(dW), ptr(db), st)
            grads[2 * li], grads[2 * li + 1] = dW, db
            # data gradient through layer li
            pkT = _pack(Ws[li], 6, kt, True, bf16, st)
            if li > 0:
                dprev = torch.empty((M, WIDTH), dtype=torch.float32, device=dev)
                call("vx_mlp_layer_bwd", int(bf16), kt, ptr(dH), WIDTH, M, ptr(pkT), ptr(h_prev),
                     WIDTH, ptr(dprev), WIDTH, st)
                dH = dprev
            else:
                dXp = torch.empty((M, 32 * kt), dtype=torch.float32, device=dev)
                call("vx_mlp_layer_bwd", int(bf16), kt, ptr(dH), WIDTH, M, ptr(pkT), None, 0,
                     ptr(dXp), 32 * kt, st)
                dX = dXp[:, :din]
        return (dX, None, *grads)


def mlp_forward(x, linears, compute_dtype=torch.float32, padded=False):
    """logits = rgbnet(x) on the MFMA kernels (float32 or bfloat16 compute).

    padded=True: x [M, W] float32 rows already padded to the kernels' tile width W
    (64 or 96, the smallest >= in_features) with finite values past in_features; they
    are read in place and the gradient has the same [M, W] shape."""
    if not x.is_cuda:
        raise RuntimeError("the rgbnet kernels need CUDA (HIP) tensors")
    if len(linears) < 3 or any(l.out_features != WIDTH for l in linears[:-1]) or \
            linears[-1].out_features != 3 or any(l.in_features != WIDTH for l in linears[1:]):
        raise RuntimeError("rgbnet kernels support width 192, depth >= 3, 3 outputs")
    bf16 = compute_dtype == torch.bfloat16
    params = []
    for l in linears:
        params += [l.weight, l.bias]
    if padded:
        kin = _pad_width(linears[0].in_features)
        if x.dim() != 2 or x.shape[1] != kin or x.dtype != torch.float32 or not x.is_contiguous():
            raise RuntimeError(f"padded rgbnet input must be contiguous float32 [M, {kin}]")
        x2 = x
    else:
        if x.shape[-1] != linears[0].in_features:
            raise RuntimeError(f"rgbnet input width {x.shape[-1]} != {linears[0].in_features}")
        x2 = x.reshape(-1, x.shape[-1]).float()
    out = _MlpFn.apply(x2, bf16, *params)
    return out.reshape(*x.shape[:-1], 3)
