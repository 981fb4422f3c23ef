"""Learned Newton initial-guess predictor (build-only; fills the predictor
slot of the driver, S/driver.h:2002-2050 with S/ =
applications/sintering/include/pf-applications/sintering/).

Network: per node, 272 inputs (27-node stencil of c, mu, sum eta^2 and the
node's top-2 order parameters, at t_n and as the scaled increment
t_n - t_{n-1}, plus dt) -> tanh hidden layers of width 128 -> 4 outputs, the
correction of the linear extrapolation 2 u_n - u_{n-1} on (c, mu, eta_top1,
eta_top2).  All arithmetic runs in the library's BF16-MFMA kernels (FP32
accumulation); torch tensors only hold the memory.  The plain Newton solve
from the predicted guess keeps the same tolerance, so the plain method stays
both the fallback and the reference for the result.
"""
import ctypes
import math

import torch

from ._ffi import check, lib

N_IN = 272
N_OUT = 4


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class MLP:
    """weights W_l (out x in, FP32 master + bf16 copy), biases b_l (FP32)"""

    def __init__(self, op, dims=(N_IN, 128, 128, 128, N_OUT), seed=0, device="cuda"):
        self.op = op
        self.dims = list(dims)
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.W, self.Wb, self.b, self.mW, self.vW, self.mb, self.vb = [], [], [], [], [], [], []
        for l in range(len(dims) - 1):
            n_in, n_out = dims[l], dims[l + 1]
            bound = 1.0 / math.sqrt(n_in)
            w = (torch.rand(n_out, n_in, generator=g, dtype=torch.float64) * 2 - 1) * bound
            if l == len(dims) - 2:
                w *= 0.01   # start close to the plain linear extrapolation
            w = w.float().to(device)
            self.W.append(w)
            self.Wb.append(w.to(torch.bfloat16))
            self.b.append(torch.zeros(n_out, dtype=torch.float32, device=device))
            for lst in (self.mW, self.vW):
                lst.append(torch.zeros_like(w))
            for lst in (self.mb, self.vb):
                lst.append(torch.zeros(n_out, dtype=torch.float32, device=device))
        self.step = 0
        self._ws = {}

    @property
    def n_layers(self):
        return len(self.dims) - 1

    def _buf(self, key, shape, dtype):
        t = self._ws.get(key)
        if t is None or t.numel() < math.prod(shape) or t.dtype != dtype:
            t = torch.empty(math.prod(shape), dtype=dtype, device=self.W[0].device)
            self._ws[key] = t
        return t[:math.prod(shape)].view(*shape)

    def _gemm(self, ta, tb, epi, M, N, K, A, lda, B, ldb, C, ldc, bias=None, H=None, split=1):
        check(lib().hps_mlp_gemm(self.op.handle, ta, tb, epi, M, N, K, _p(A), lda, _p(B), ldb,
                                 _p(C), ldc, _p(bias), _p(H), split))

    def forward(self, X, M):
        """X: (M, 272) bf16 -> Y (M, 4) FP32; keeps the activations"""
        acts = [X]
        h = X
        for l in range(self.n_layers):
            n_in, n_out = self.dims[l], self.dims[l + 1]
            last = l == self.n_layers - 1
            out = self._buf(("h", l), (M, n_out), torch.float32 if last else torch.bfloat16)
            self._gemm(0, 1, 2 if last else 1, M, n_out, n_in, h, n_in, self.Wb[l], n_in, out,
                       n_out, bias=self.b[l])
            acts.append(out)
            h = out
        return acts[-1], acts

    def backward(self, acts, dY, M, split=None):
        """dY: (M, 4) bf16 -> gradients (gW, gb) per layer"""
        split = split or max(1, min(64, M // 2048))
        gW, gb = [None] * self.n_layers, [None] * self.n_layers
        dZ = dY
        for l in range(self.n_layers - 1, -1, -1):
            n_in, n_out = self.dims[l], self.dims[l + 1]
            A = acts[l]
            part = self._buf(("part", l), (split, n_out, n_in), torch.float32)
            self._gemm(1, 0, 0, n_out, n_in, M, dZ, n_out, A, n_in, part, n_in, split=split)
            g = self._buf(("gW", l), (n_out, n_in), torch.float32)
            check(lib().hps_mlp_sum_splits(self.op.handle, n_out * n_in, split, _p(part), _p(g)))
            gbl = self._buf(("gb", l), (n_out,), torch.float32)
            check(lib().hps_mlp_colsum(self.op.handle, M, n_out, _p(dZ), _p(gbl)))
            gW[l], gb[l] = g, gbl
            if l > 0:
                dA = self._buf(("dz", l), (M, n_in), torch.bfloat16)
                self._gemm(0, 0, 3, M, n_in, n_out, dZ, n_out, self.Wb[l], n_in, dA, n_in,
                           H=acts[l])
                dZ = dA
        return gW, gb

    def loss_grad(self, Y, T, M):
        dY = self._buf("dY", (M, N_OUT), torch.bfloat16)
        nb = 256
        part = self._buf("lpart", (nb,), torch.float32)
        loss = ctypes.c_float()
        check(lib().hps_mlp_loss(self.op.handle, M * N_OUT, _p(Y), _p(T), _p(dY), _p(part), nb,
                                 ctypes.byref(loss)))
        return loss.value, dY

    def adam(self, gW, gb, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        for l in range(self.n_layers):
            n_in, n_out = self.dims[l], self.dims[l + 1]
            check(lib().hps_mlp_adam(self.op.handle, n_out, n_in, _p(self.W[l]), _p(self.mW[l]),
                                     _p(self.vW[l]), _p(gW[l]), lr, b1, b2, eps, self.step,
                                     _p(self.Wb[l])))
            check(lib().hps_mlp_adam(self.op.handle, 1, n_out, _p(self.b[l]), _p(self.mb[l]),
                                     _p(self.vb[l]), _p(gb[l]), lr, b1, b2, eps, self.step,
                                     None))

    def train_step(self, X, T, M, lr):
        Y, acts = self.forward(X, M)
        loss, dY = self.loss_grad(Y, T, M)
        gW, gb = self.backward(acts, dY, M)
        self.adam(gW, gb, lr)
        return loss


class LearnedPredictor:
    """Newton initial guess u_pred = 2 u_n - u_{n-1} + ls * out_scale * MLP(features).

    ls is each node's local increment scale (RMS of the stencil increments
    of u_n - u_{n-1}, computed by the feature kernel); increment features are
    divided by it and the outputs multiplied by it, so the network sees the
    same scales in the quiet bulk and at moving interfaces.  out_scale holds
    one global constant per output, fixed at training."""

    def __init__(self, op, mlp: MLP = None, dt_scale=1.0, out_scale=None):
        self.op = op
        self.mlp = mlp or MLP(op)
        self.dt_scale = dt_scale
        dev = self.mlp.W[0].device
        self.out_scale = torch.tensor(list(out_scale) if out_scale is not None else [1.0] * N_OUT,
                                      dtype=torch.float32, device=dev)
        n = op.n_nodes
        self.X = torch.empty(n, N_IN, dtype=torch.bfloat16, device=dev)
        self.top2 = torch.empty(n, 2, dtype=torch.int32, device=dev)
        self.lscale = torch.empty(n, dtype=torch.float32, device=dev)
        self.diff = torch.empty(op.n_dofs, dtype=torch.float64, device=dev)
        self.use_network = True
        self.inc_scale = 1.0

    def step_scale(self, u_prev, u_cur):
        """RMS of u_cur - u_prev, computed by the library"""
        h = self.op.handle
        n = self.op.n_dofs
        check(lib().hps_copy(h, n, _p(self.diff), _p(u_cur)))
        check(lib().hps_axpy(h, n, -1.0, _p(u_prev), _p(self.diff)))
        return max(self.op.norm(self.diff) / math.sqrt(n), 1e-300)

    def features(self, u_prev, u_cur, dt):
        self.inc_scale = 1.0 / self.step_scale(u_prev, u_cur)
        check(lib().hps_mlp_features(self.op.handle, _p(u_prev), _p(u_cur), dt, self.dt_scale,
                                     self.inc_scale, _p(self.X), _p(self.top2),
                                     _p(self.lscale)))
        return self.X, self.top2

    def predict(self, u_prev, u_cur, dt, out):
        Y = None
        if self.use_network:
            X, _ = self.features(u_prev, u_cur, dt)
            Y, _ = self.mlp.forward(X, self.op.n_nodes)
        check(lib().hps_mlp_predict(self.op.handle, _p(u_prev), _p(u_cur), _p(Y),
                                    _p(self.top2), _p(self.out_scale), _p(self.lscale),
                                    _p(out)))
        return out

    def targets(self, u_prev, u_cur, u_next, T):
        """(u_next - (2 u_cur - u_prev)) / (ls * out_scale) on the 4 outputs;
        call after features() of the same (u_prev, u_cur)"""
        check(lib().hps_mlp_targets(self.op.handle, _p(u_prev), _p(u_cur), _p(u_next),
                                    _p(self.top2), _p(self.out_scale), _p(self.lscale), _p(T)))
        return T

    def __call__(self, x, states, dt):
        """TimeStepper hook: states = [u_n, u_{n-1}, ...]"""
        if len(states) < 2:
            return x
        return self.predict(states[1], states[0], dt, x)
