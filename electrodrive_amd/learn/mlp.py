"""Small MLP on native HIP kernels (forward, backward, Adam).

Architecture: Linear(F, W) -> SiLU -> Linear(W, W) -> SiLU -> Linear(W, W)
-> SiLU -> Linear(W, O)  (four linear layers, width W = 256 by default).
Parameters live in one flat FP32 device buffer; every forward / backward /
update step is a sequence of native kernels (``ed_gemm_f32`` on MFMA,
``ed_act_backward_f32``, ``ed_colsum_f32``, ``ed_adam_f32``, ``ed_wmse_f32``).
torch only provides device allocations.
"""
import ctypes

import numpy as np
import torch

from .. import _native as nat


class NativeMLP:
    def __init__(self, dims=(12, 256, 256, 256, 1), seed=0, device=None):
        nat.require_gpu()
        self.dims = tuple(int(d) for d in dims)
        self.dev = torch.device(device or "cuda")
        if self.dev.index is None:
            self.dev = torch.device("cuda", torch.cuda.current_device())
        rng = np.random.default_rng(seed)
        shapes, offs, tot = [], [], 0
        for i in range(len(self.dims) - 1):
            fi, fo = self.dims[i], self.dims[i + 1]
            shapes.append((fo, fi))
            offs.append((tot, tot + fo * fi))
            tot += fo * fi + fo
        self.shapes, self.offs, self.n_params = shapes, offs, tot
        host = np.zeros(tot, dtype=np.float32)
        for (fo, fi), (w0, b0) in zip(shapes, offs):
            bound = np.sqrt(6.0 / (fi + fo))  # Glorot uniform
            host[w0:w0 + fo * fi] = rng.uniform(-bound, bound, fo * fi)
        self.params = torch.tensor(host, device=self.dev)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self._cache = {}
        self._scratch = torch.empty(64 * max(self.dims) + 256, dtype=torch.float32, device=self.dev)

    # -- parameter views ---------------------------------------------------------
    def W(self, i, buf=None):
        buf = self.params if buf is None else buf
        fo, fi = self.shapes[i]
        w0, _ = self.offs[i]
        return buf[w0:w0 + fo * fi]

    def b(self, i, buf=None):
        buf = self.params if buf is None else buf
        fo, _ = self.shapes[i]
        _, b0 = self.offs[i]
        return buf[b0:b0 + fo]

    def state_dict(self):
        return {"dims": list(self.dims), "params": self.params.detach().cpu().numpy().copy()}

    def load_state_dict(self, sd):
        if tuple(sd["dims"]) != self.dims:
            raise ValueError("dims mismatch")
        self.params.copy_(torch.as_tensor(np.asarray(sd["params"], dtype=np.float32)))

    # -- kernels --------------------------------------------------------------------
    def _gemm(self, ta, tb, M, N, K, A, lda, B, ldb, C, ldc, bias=None, act=0, pre=None, beta=0.0):
        nat.check(nat.lib().ed_gemm_f32(int(ta), int(tb), int(M), int(N), int(K), nat.ptr(A), int(lda),
                                        nat.ptr(B), int(ldb), nat.ptr(C), int(ldc), nat.ptr(bias),
                                        int(act), nat.ptr(pre), float(beta),
                                        nat.stream_ptr(self.dev)), "ed_gemm_f32")

    def _buf(self, key, rows, cols):
        t = self._cache.get(key)
        if t is None or t.shape[0] < rows * cols:
            t = torch.empty(rows * cols, dtype=torch.float32, device=self.dev)
            self._cache[key] = t
        return t[:rows * cols]

    def forward(self, X, keep=False):
        """X: (B, F) float32 device tensor -> (B, O).  ``keep`` stores the
        activations for ``backward``."""
        X = X.contiguous()
        Bn = X.shape[0]
        L = len(self.shapes)
        h = X
        acts, pres = [X], []
        for i in range(L):
            fo, fi = self.shapes[i]
            last = i == L - 1
            out = self._buf(("h", i), Bn, fo)
            pre = None if last else self._buf(("z", i), Bn, fo)
            self._gemm(0, 1, Bn, fo, fi, h, fi, self.W(i), fi, out, fo, bias=self.b(i),
                       act=0 if last else 1, pre=pre)
            pres.append(pre)
            h = out
            acts.append(out)
        if keep:
            self._acts, self._pres, self._B = acts, pres, Bn
        return h.view(Bn, self.dims[-1])

    def backward(self, dout):
        """dout: (B, O) gradient of the loss w.r.t. the output; fills self.grads."""
        acts, pres, Bn = self._acts, self._pres, self._B
        L = len(self.shapes)
        g = dout.contiguous().view(-1)
        st = nat.stream_ptr(self.dev)
        for i in range(L - 1, -1, -1):
            fo, fi = self.shapes[i]
            if i < L - 1:
                dz = self._buf(("dz", i), Bn, fo)
                nat.check(nat.lib().ed_act_backward_f32(nat.ptr(pres[i]), nat.ptr(g), nat.ptr(dz),
                                                        Bn * fo, 1, st), "ed_act_backward_f32")
                g = dz
            # dW = g^T h_in  (fo x fi), db = colsum(g)
            self._gemm(1, 0, fo, fi, Bn, g, fo, acts[i], fi, self.W(i, self.grads), fi)
            nat.check(nat.lib().ed_colsum_f32(nat.ptr(g), Bn, fo, nat.ptr(self._scratch),
                                              nat.ptr(self.b(i, self.grads)), st), "ed_colsum_f32")
            if i > 0:
                dh = self._buf(("dh", i), Bn, fi)
                self._gemm(0, 0, Bn, fi, fo, g, fo, self.W(i), fi, dh, fi)
                g = dh

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, weight_decay=0.0):
        self.step += 1
        nat.check(nat.lib().ed_adam_f32(nat.ptr(self.params), nat.ptr(self.grads), nat.ptr(self.m),
                                        nat.ptr(self.v), self.n_params, float(lr), float(b1),
                                        float(b2), float(eps), self.step, float(weight_decay),
                                        nat.stream_ptr(self.dev)), "ed_adam_f32")

    def wmse(self, y, t, w=None):
        """Returns (loss, grad) for sum w (y - t)^2."""
        y = y.contiguous().view(-1)
        t = t.contiguous().view(-1)
        grad = torch.empty_like(y)
        part = self._buf(("part",), 256, 1)
        loss = ctypes.c_float(0.0)
        nat.check(nat.lib().ed_wmse_f32(nat.ptr(y), nat.ptr(t), nat.ptr(w), y.numel(),
                                        nat.ptr(grad), nat.ptr(part), ctypes.byref(loss),
                                        nat.stream_ptr(self.dev)), "ed_wmse_f32")
        return float(loss.value), grad
