"""CPU reference of the learned phase-congruency network (test infrastructure).

The network has no counterpart in the reference; this module states it with
PyTorch's own CPU conv2d in float64 so that the HIP kernels (MFMA forward,
hand-written backward) are checked against a library routine:

  layer 1   conv3x3 1 -> 32, bias, ReLU
  layers 2-5 conv3x3 32 -> 32, bias, ReLU
  layer 6   conv3x3 32 -> 1, bias
  zero padding 1, cross-correlation (PyTorch's conv2d), OIHW weights.

`emulate=True` rounds what the HIP path stores in BF16 -- the weights of
every layer and the activations after each ReLU -- with torch's own
float32 -> bfloat16 conversion (round to nearest even); the backward treats
these roundings as identity (straight-through), as mixed-precision training
does.  The flat parameter order is the library's (see `unpack`).
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as Fn

CH = 32
MID = CH * CH * 9 + CH
OFF_W1, OFF_B1, OFF_MID = 0, CH * 9, CH * 9 + CH
OFF_W6 = OFF_MID + 4 * MID
OFF_B6 = OFF_W6 + CH * 9
TOTAL = OFF_B6 + 1


def unpack(p):
    """Flat parameter vector -> [(W, b)] for layers 1..6 (views of p)."""
    out = [(p[OFF_W1:OFF_B1].reshape(CH, 1, 3, 3), p[OFF_B1:OFF_MID])]
    for l in range(4):
        base = OFF_MID + l * MID
        out.append((p[base:base + CH * CH * 9].reshape(CH, CH, 3, 3), p[base + CH * CH * 9:base + MID]))
    out.append((p[OFF_W6:OFF_B6].reshape(1, CH, 3, 3), p[OFF_B6:TOTAL]))
    return out


def bf16(t: torch.Tensor) -> torch.Tensor:
    return t.to(torch.float32).to(torch.bfloat16).to(torch.float64)


class _Ste(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return bf16(x)

    @staticmethod
    def backward(ctx, g):
        return g


def forward(p: torch.Tensor, x: torch.Tensor, emulate: bool = True) -> torch.Tensor:
    """p: flat float64 parameters (may require grad); x: (n, h, w) float64."""
    q = _Ste.apply if emulate else (lambda t: t)
    layers = unpack(p)
    a = x[:, None]
    for i, (W, b) in enumerate(layers):
        z = Fn.conv2d(a, q(W), b, padding=1)
        if i == 5:
            return z[:, 0]
        a = q(torch.relu(z))
    raise AssertionError


def loss_and_grad(p_np: np.ndarray, x_np: np.ndarray, t_np: np.ndarray, emulate: bool = True,
                  weight_eps: float | None = None):
    """MSE over all pixels and its gradient with respect to the flat FP32
    master parameters.  weight_eps: each pixel's squared error weighted by
    max(t of image 2i, t of image 2i+1) + weight_eps (images in FSIM pairs),
    still divided by the pixel count."""
    p = torch.tensor(np.asarray(p_np, np.float64), requires_grad=True)
    x = torch.tensor(np.asarray(x_np, np.float64))
    t = torch.tensor(np.asarray(t_np, np.float64))
    out = forward(p, x, emulate)
    if weight_eps is None:
        loss = torch.mean((out - t) ** 2)
    else:
        pair = t.reshape(-1, 2, *t.shape[1:])
        wt = (torch.maximum(pair[:, 0], pair[:, 1]) + weight_eps)[:, None].expand_as(pair).reshape(t.shape)
        loss = torch.mean(wt * (out - t) ** 2)
    loss.backward()
    return float(loss.detach()), p.grad.numpy().copy(), out.detach().numpy()


def predict(p_np: np.ndarray, x_np: np.ndarray, emulate: bool = True) -> np.ndarray:
    with torch.no_grad():
        return forward(torch.tensor(np.asarray(p_np, np.float64)), torch.tensor(np.asarray(x_np, np.float64)),
                       emulate).numpy()


def he_uniform(seed: int) -> np.ndarray:
    """Seeded parameters of the same scale as the library's initialisation."""
    rng = np.random.default_rng(seed)
    p = np.zeros(TOTAL, np.float32)
    for i, (W, b) in enumerate(unpack(p)):
        fan_in = W.shape[1] * 9
        W[...] = rng.uniform(-1, 1, W.shape) * np.sqrt(6.0 / fan_in)
        b[...] = rng.uniform(-0.05, 0.05, b.shape)
    return p
