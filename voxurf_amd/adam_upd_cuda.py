"""Drop-in for the reference's `adam_upd_cuda` extension (lib/cuda/adam_upd.cpp:79-86):
fused in-place Adam updates on HIP.  Signatures as the reference:
(param, grad, exp_avg, exp_avg_sq, step, beta1, beta2, lr, eps), and for the per-lr
variant (param, grad, exp_avg, exp_avg_sq, perlr, step, beta1, beta2, lr, eps).
"""
from ._lib import call, dtype_code, ptr, stream_of
from .render_utils_cuda import _check_input


def _prep(param, grad, exp_avg, exp_avg_sq, perlr=None):
    for t, n in ((param, "param"), (grad, "grad"), (exp_avg, "exp_avg"), (exp_avg_sq, "exp_avg_sq")):
        _check_input(t, n)
        if t.numel() != param.numel() or t.dtype != param.dtype:
            raise RuntimeError(f"{n} must match param in size and dtype")
    if perlr is not None:
        _check_input(perlr, "perlr")
        if perlr.numel() != param.numel() or perlr.dtype != param.dtype:
            raise RuntimeError("perlr must match param in size and dtype")
    return dtype_code(param)


def adam_upd(param, grad, exp_avg, exp_avg_sq, step, beta1, beta2, lr, eps):
    code = _prep(param, grad, exp_avg, exp_avg_sq)
    call("vx_adam_upd", code, ptr(param), ptr(grad), ptr(exp_avg), ptr(exp_avg_sq), None,
         param.numel(), int(step), float(beta1), float(beta2), float(lr), float(eps), 0,
         stream_of(param))


def masked_adam_upd(param, grad, exp_avg, exp_avg_sq, step, beta1, beta2, lr, eps):
    code = _prep(param, grad, exp_avg, exp_avg_sq)
    call("vx_adam_upd", code, ptr(param), ptr(grad), ptr(exp_avg), ptr(exp_avg_sq), None,
         param.numel(), int(step), float(beta1), float(beta2), float(lr), float(eps), 1,
         stream_of(param))


def adam_upd_with_perlr(param, grad, exp_avg, exp_avg_sq, perlr, step, beta1, beta2, lr, eps):
    code = _prep(param, grad, exp_avg, exp_avg_sq, perlr)
    call("vx_adam_upd", code, ptr(param), ptr(grad), ptr(exp_avg), ptr(exp_avg_sq), ptr(perlr),
         param.numel(), int(step), float(beta1), float(beta2), float(lr), float(eps), 0,
         stream_of(param))
