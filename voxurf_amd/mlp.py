"""rgbnet: the colour MLP of lib/voxurf_fine.py:132-149.

RgbNet has the reference's module structure (nn.Sequential of Linear, ReLU,
(depth-2) x Sequential(Linear, ReLU), Linear), so its state-dict keys (`0.weight`,
`2.0.weight`, ..., `<depth>.bias`) are the reference's.  forward() runs the fused HIP
MLP kernels of libvoxhip.so (voxurf_amd.mlp_hip) for float32 inputs; compute_dtype and
f32_split select their precision (mlp_hip.prec_of).
"""
import torch
import torch.nn as nn


class RgbNet(nn.Sequential):
    def __init__(self, in_dim, width, depth, params=None, compute_dtype=torch.float32):
        layers = [nn.Linear(in_dim, width), nn.ReLU(inplace=True)]
        layers += [nn.Sequential(nn.Linear(width, width), nn.ReLU(inplace=True))
                   for _ in range(depth - 2)]
        layers += [nn.Linear(width, 3)]
        super().__init__(*layers)
        nn.init.constant_(self[-1].bias, 0)
        self.compute_dtype = compute_dtype
        # float32 compute: BF16-split MFMAs (FP32-accurate, the default) or FP32 MFMAs
        self.f32_split = True
        if params is not None:
            with torch.no_grad():
                for lin, (W, b) in zip(self.linears(), params):
                    lin.weight.copy_(W)
                    lin.bias.copy_(b)

    def linears(self):
        out = []
        for m in self:
            if isinstance(m, nn.Linear):
                out.append(m)
            elif isinstance(m, nn.Sequential):
                out.append(m[0])
        return out

    def forward(self, x):
        if self.compute_dtype == torch.float64:
            # float64 evaluation (the error reference of the precision study): the MFMA
            # kernels are float32/bfloat16 only, so this runs torch's own Linear layers.
            return super().forward(x)
        from .mlp_hip import mlp_forward
        return mlp_forward(x, self.linears(), self.compute_dtype, f32_split=self.f32_split)

    def forward_padded(self, x):
        """forward() of rows already padded to the kernels' tile width (64 for the
        42-wide fine input): x [M, 64] float32, zero past in_features."""
        from .mlp_hip import mlp_forward
        return mlp_forward(x, self.linears(), self.compute_dtype, padded=True,
                           f32_split=self.f32_split)
