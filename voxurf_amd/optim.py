"""Adam with optional per-voxel learning rate (lib/utils.py:54-199) on one fused HIP
kernel per parameter, and the optimiser factory (lib/utils.py:202-229).

Semantics follow utils.Adam: bias-corrected first/second moments with
denom = sqrt(v)/sqrt(1-beta2^t) + eps and step = lr/(1-beta1^t); parameters whose
.grad is None are skipped; the per-voxel lr factor applies to the parameter whose shape
matches it.  `clear_grad=True` additionally zeroes each gradient in the same pass
(for callers that keep persistent gradient buffers instead of set_to_none).
"""
import torch
import torch.nn as nn

from ._lib import call, dtype_code, ptr, stream_of


class Adam(torch.optim.Optimizer):
    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=0,
                 amsgrad=False, clear_grad=False):
        if not 0.0 <= lr:
            raise ValueError("Invalid learning rate: {}".format(lr))
        if not 0.0 <= eps:
            raise ValueError("Invalid epsilon value: {}".format(eps))
        if not 0.0 <= betas[0] < 1.0:
            raise ValueError("Invalid beta parameter at index 0: {}".format(betas[0]))
        if not 0.0 <= betas[1] < 1.0:
            raise ValueError("Invalid beta parameter at index 1: {}".format(betas[1]))
        if not 0.0 <= weight_decay:
            raise ValueError("Invalid weight_decay value: {}".format(weight_decay))
        if amsgrad:
            raise NotImplementedError("amsgrad is not used by Voxurf and is not provided")
        defaults = dict(lr=lr, betas=betas, eps=eps, weight_decay=weight_decay, amsgrad=amsgrad)
        self.per_lr = None
        self.clear_grad = clear_grad
        super().__init__(params, defaults)

    def set_pervoxel_lr(self, count):
        assert self.param_groups[0]["params"][0].shape == count.shape
        self.per_lr = count.float() / count.max()

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        for group in self.param_groups:
            beta1, beta2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                if p.grad.is_sparse:
                    raise RuntimeError("Adam does not support sparse gradients")
                if not p.is_cuda:
                    raise RuntimeError("voxurf_amd.optim.Adam needs parameters on the GPU")
                state = self.state[p]
                if len(state) == 0:
                    state["step"] = 0
                    state["exp_avg"] = torch.zeros_like(p, memory_format=torch.preserve_format)
                    state["exp_avg_sq"] = torch.zeros_like(p, memory_format=torch.preserve_format)
                state["step"] += 1
                g = p.grad
                if g.stride() != p.stride():
                    g = torch.empty_like(p).copy_(g)
                    p.grad = g
                if group["weight_decay"] != 0:
                    g.add_(p, alpha=group["weight_decay"])
                per_lr = None
                if self.per_lr is not None and p.shape == self.per_lr.shape:
                    per_lr = torch.empty_like(p).copy_(self.per_lr)
                ev = getattr(self, "profile_events", None)
                if ev is not None:
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record()
                call("vx_adam_step", dtype_code(p), ptr(p), ptr(g), ptr(state["exp_avg"]),
                     ptr(state["exp_avg_sq"]), ptr(per_lr), p.numel(), state["step"],
                     float(group["lr"]), float(beta1), float(beta2), float(group["eps"]),
                     int(self.clear_grad), None, stream_of(p))
                if ev is not None:
                    e1.record()
                    ev[id(p)] = (e0, e1)
        return loss


def create_optimizer_or_freeze_model(model, cfg_train, global_step):
    """lib/utils.py:202-229: one param group per `lrate_<name>` entry of cfg_train."""
    decay_steps = cfg_train.lrate_decay * 1000
    decay_factor = 0.1 ** (global_step / decay_steps)
    param_group = []
    for k in cfg_train.keys():
        if not k.startswith("lrate_"):
            continue
        k = k[len("lrate_"):]
        if not hasattr(model, k):
            continue
        param = getattr(model, k)
        if param is None:
            continue
        lr = getattr(cfg_train, f"lrate_{k}") * decay_factor
        if lr > 0:
            if isinstance(param, nn.Module):
                param = param.parameters()
            param_group.append({"params": param, "lr": lr, "name": k})
        else:
            param.requires_grad = False
    return Adam(param_group, betas=(0.9, 0.99))
