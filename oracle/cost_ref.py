"""CPU reference of the learned cost model (FP32 and BF16 numerics).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU cost
model kernels (map_folding_amd/csrc/cost_mlp.hip).  The product never imports
it.  The reference repository has no learned model; this is the plain PyTorch
FP32 statement of the same network, so its gradients come from autograd and
are independent of the hand-written backward kernels they check.

Parameter layout (flat FP32, row-major out x in): W1[128x80] b1 W2[128x128] b2
W3[128x128] b3 w4[128] b4.
"""
from __future__ import annotations

import numpy as np
import torch

FEATURES = 80
HIDDEN = 128
SHAPES = [("W1", (HIDDEN, FEATURES)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
          ("W3", (HIDDEN, HIDDEN)), ("b3", (HIDDEN,)), ("w4", (HIDDEN,)), ("b4", (1,))]
PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    """Flat parameter vector -> dict of torch tensors (views into one fp32 tensor)."""
    t = torch.as_tensor(np.asarray(flat, dtype=np.float32)).clone()
    out, off = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = t[off:off + k].view(shape)
        off += k
    assert off == PARAMS
    return t, out


def forward(p, x):
    h1 = torch.relu(x @ p["W1"].T + p["b1"])
    h2 = torch.relu(h1 @ p["W2"].T + p["b2"])
    h3 = torch.relu(h2 @ p["W3"].T + p["b3"])
    return h3 @ p["w4"] + p["b4"]


def loss_and_grad(flat, x, y):
    """Mean-squared-error loss and its gradient (flat layout), FP32 autograd."""
    t = torch.as_tensor(np.asarray(flat, dtype=np.float32)).clone().requires_grad_(True)
    _, p = _views(t)
    xx = torch.as_tensor(np.asarray(x, dtype=np.float32))
    yy = torch.as_tensor(np.asarray(y, dtype=np.float32))
    loss = ((forward(p, xx) - yy) ** 2).mean()
    loss.backward()
    return float(loss.detach()), t.grad.numpy().copy()


def _views(t):
    out, off = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = t[off:off + k].view(shape)
        off += k
    return t, out


def predict_fp32(flat, x):
    _, p = unpack(flat)
    with torch.no_grad():
        return forward(p, torch.as_tensor(np.asarray(x, dtype=np.float32))).numpy()


def _bf16(t):
    return t.to(torch.bfloat16).to(torch.float32)


def predict_bf16(flat, x):
    """BF16 inference semantics: inputs, weights and hidden activations rounded
    to BF16 (round to nearest even), products accumulated in FP32, biases FP32."""
    _, p = unpack(flat)
    with torch.no_grad():
        h = _bf16(torch.as_tensor(np.asarray(x, dtype=np.float32)))
        for w, b in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
            h = _bf16(torch.relu(h @ _bf16(p[w]).T + p[b]))
        return (h @ _bf16(p["w4"]) + p["b4"]).numpy()
