"""Numpy restatement of the learned predictor's network (checker only).

Forward/backward of the MLP used by hpsint_amd/predictor.py: tanh hidden
layers, linear output, MSE loss L = mean((Y - T)^2).  With round_bf16=True
the BF16 roundings of the GPU kernels are reproduced (inputs, weights,
hidden activations, dY and the back-propagated dZ are rounded to bfloat16
with round-to-nearest-even; products accumulate in float64 here, in FP32 on
the GPU), so the GPU and this restatement differ only by accumulation
order.  With round_bf16=False everything is float64, which is what the
finite-difference test pins.
"""
import numpy as np


def bf16(x):
    """round float64/float32 array to bfloat16 (RNE), returned as float64"""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(Ws, bs, X, round_bf16=True):
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    acts = [r(X)]
    h = acts[0]
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = h @ r(W).T + b
        if l < len(Ws) - 1:
            h = r(np.tanh(z))
        else:
            h = z
        acts.append(h)
    return acts[-1], acts


def loss_and_grad(Y, T, round_bf16=True):
    n = Y.size
    loss = np.mean((Y - T) ** 2)
    dY = 2.0 * (Y - T) / n
    return loss, (bf16(dY) if round_bf16 else dY)


def backward(Ws, acts, dY, round_bf16=True):
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    L = len(Ws)
    gW, gb = [None] * L, [None] * L
    dZ = dY
    for l in range(L - 1, -1, -1):
        gW[l] = dZ.T @ acts[l]
        gb[l] = dZ.sum(axis=0)
        if l > 0:
            dA = dZ @ r(Ws[l])
            dZ = r(dA * (1.0 - acts[l] ** 2))
    return gW, gb
