"""Host side of the rgbnet MLP.

mlp_forward(x, linears, compute_dtype) -> logits [M, 3].
"""
import torch
import torch.nn.functional as F


def mlp_forward(x, linears, compute_dtype=torch.float32):
    h = x
    for i, lin in enumerate(linears):
        h = F.linear(h, lin.weight, lin.bias)
        if i < len(linears) - 1:
            h = F.relu(h)
    return h
