"""WarmStartNet: the per-particle MLP of the learned warm start, trained and
run on the BF16 MFMA kernels of libpdx (include/pdx_nn.h).

Layers 64 -> 256 -> 256 -> 256 -> 32 (the first 3 outputs are u / s; the
other 29 columns pad the last GEMM to the 32-column tile and stay zero), ReLU
after the first three.  Parameters are FP32 masters with BF16 copies for the
GEMMs; activations are BF16, accumulation FP32.  torch provides device memory
and the stream only: every arithmetic step is a libpdx kernel.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _native as N

DIMS = (64, 256, 256, 256, 32)
NOUT = 3


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


def wgrad_rows_per_slice(M, slices=256):
    """Rows per weight-gradient slice: a multiple of 32 (the transpose-read
    kernel's stage) giving about `slices` partials, one block each."""
    return max(32, -(-M // (slices * 32)) * 32)


class WarmStartNet:
    def __init__(self, dims=DIMS, seed: int = 0, weights=None):
        import torch
        self.torch = torch
        self.lib = N.load()
        self.dims = tuple(int(d) for d in dims)
        rng = np.random.default_rng(seed)
        self.W, self.b, self.Wbf = [], [], []
        self.mW, self.vW, self.mb, self.vb = [], [], [], []
        for li, (k, n) in enumerate(zip(self.dims[:-1], self.dims[1:])):
            last = li == len(self.dims) - 2
            if weights is not None:
                w, bb = np.asarray(weights[f"W{li}"], np.float32), np.asarray(weights[f"b{li}"], np.float32)
            else:
                w = (rng.standard_normal((k, n)) * np.sqrt(2.0 / k) * (0.1 if last else 1.0)).astype(np.float32)
                bb = np.zeros(n, np.float32)
                if last:
                    w[:, NOUT:] = 0.0
            W = torch.from_numpy(np.ascontiguousarray(w)).cuda()
            B = torch.from_numpy(np.ascontiguousarray(bb)).cuda()
            self.W.append(W)
            self.b.append(B)
            self.Wbf.append(torch.empty((k, n), dtype=torch.int16, device="cuda"))
            for lst, ref in ((self.mW, W), (self.vW, W), (self.mb, B), (self.vb, B)):
                lst.append(torch.zeros_like(ref))
        self.step = 0
        self.refresh_bf16()

    # -- plumbing ----------------------------------------------------------------
    @property
    def stream(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def refresh_bf16(self):
        for W, Wb in zip(self.W, self.Wbf):
            _ok(self.lib.pdx_nn_to_bf16(_p(W), W.numel(), _p(Wb), self.stream), "pdx_nn_to_bf16")

    def to_bf16(self, x_f32):
        """(M, 64) float32 device tensor -> BF16 bits (int16 tensor)."""
        out = self.torch.empty(x_f32.shape, dtype=self.torch.int16, device="cuda")
        _ok(self.lib.pdx_nn_to_bf16(_p(x_f32), x_f32.numel(), _p(out), self.stream), "pdx_nn_to_bf16")
        return out

    def state(self) -> dict:
        out = {"dims": np.array(self.dims)}
        for li in range(len(self.W)):
            out[f"W{li}"] = self.W[li].cpu().numpy()
            out[f"b{li}"] = self.b[li].cpu().numpy()
        return out

    def save(self, path, **meta):
        np.savez(path, **self.state(), **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            w = {k: z[k] for k in z.files}
        return cls(dims=tuple(int(d) for d in w["dims"]), weights=w)

    # -- passes ------------------------------------------------------------------
    def forward(self, Xbf, keep=False):
        """Xbf: (M, 64) BF16 bits.  Returns the (M, 32) FP32 output (and the
        BF16 input of every layer when keep)."""
        torch = self.torch
        M = Xbf.shape[0]
        acts = [Xbf]
        h = Xbf
        L = len(self.W)
        yf = None
        for li in range(L):
            k, n = self.dims[li], self.dims[li + 1]
            last = li == L - 1
            if last:
                yf = torch.empty((M, n), dtype=torch.float32, device="cuda")
                _ok(self.lib.pdx_nn_linear_fwd(_p(h), _p(self.Wbf[li]), _p(self.b[li]), M, k, n, 0, None, _p(yf),
                                               n, self.stream), "pdx_nn_linear_fwd")
            else:
                y = torch.empty((M, n), dtype=torch.int16, device="cuda")
                _ok(self.lib.pdx_nn_linear_fwd(_p(h), _p(self.Wbf[li]), _p(self.b[li]), M, k, n, 1, _p(y), None,
                                               n, self.stream), "pdx_nn_linear_fwd")
                h = y
                acts.append(h)
        return (yf, acts) if keep else yf

    def gradients(self, Xbf, T, w, wsum, rows_per_slice=None):
        """Loss and parameter gradients of sum_i w_i |y_i - t_i|^2 / wsum.
        T: (M, 3) FP32 targets, w: (M,) FP32 weights (device tensors).
        rows_per_slice: rows of one weight-gradient partial (default: about one
        slice per CU, see wgrad_rows_per_slice)."""
        torch = self.torch
        M = Xbf.shape[0]
        if rows_per_slice is None:
            rows_per_slice = wgrad_rows_per_slice(M)
        y, acts = self.forward(Xbf, keep=True)
        L = len(self.W)
        nout = self.dims[-1]
        dY = torch.empty((M, nout), dtype=torch.int16, device="cuda")
        lp = torch.empty(((M + 255) // 256,), dtype=torch.float32, device="cuda")
        _ok(self.lib.pdx_nn_mse(_p(y), nout, _p(T), _p(w), M, float(wsum), _p(dY), _p(lp), self.stream), "pdx_nn_mse")
        ns = (M + rows_per_slice - 1) // rows_per_slice
        part = torch.empty((ns * max(a * b for a, b in zip(self.dims[:-1], self.dims[1:])),), dtype=torch.float32,
                           device="cuda")
        gW, gb = [None] * L, [None] * L
        for li in range(L - 1, -1, -1):
            k, n = self.dims[li], self.dims[li + 1]
            gW[li] = torch.empty((k, n), dtype=torch.float32, device="cuda")
            gb[li] = torch.empty((n,), dtype=torch.float32, device="cuda")
            _ok(self.lib.pdx_nn_linear_wgrad(_p(acts[li]), _p(dY), M, k, n, n, rows_per_slice, _p(part), _p(gW[li]),
                                             _p(gb[li]), self.stream), "pdx_nn_linear_wgrad")
            if li > 0:
                dX = torch.empty((M, k), dtype=torch.int16, device="cuda")
                _ok(self.lib.pdx_nn_linear_dgrad(_p(dY), _p(self.Wbf[li]), M, k, n, n, _p(acts[li]), _p(dX),
                                                 self.stream), "pdx_nn_linear_dgrad")
                dY = dX
        loss = float(lp.cpu().numpy().astype(np.float64).sum())
        return loss, gW, gb

    def adam(self, gW, gb, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        self.step += 1
        for li in range(len(self.W)):
            _ok(self.lib.pdx_nn_adam(_p(self.W[li]), _p(gW[li]), _p(self.mW[li]), _p(self.vW[li]), _p(self.Wbf[li]),
                                     self.W[li].numel(), lr, beta1, beta2, eps, self.step, weight_decay, self.stream),
                "pdx_nn_adam")
            _ok(self.lib.pdx_nn_adam(_p(self.b[li]), _p(gb[li]), _p(self.mb[li]), _p(self.vb[li]), None,
                                     self.b[li].numel(), lr, beta1, beta2, eps, self.step, 0.0, self.stream),
                "pdx_nn_adam")

    def train_step(self, Xbf, T, w, wsum, lr, **kw) -> float:
        loss, gW, gb = self.gradients(Xbf, T, w, wsum)
        self.adam(gW, gb, lr, **kw)
        return loss

    def predict(self, X_f32, chunk=1 << 20):
        """(M, 64) float32 device features -> (M, 3) float32 host array of u / s."""
        return self.predict_device(X_f32, chunk).cpu().numpy()

    def predict_device(self, X_f32, chunk=1 << 20):
        """(M, 64) float32 device features -> (M, 3) float32 device tensor of u / s."""
        import torch
        out = [self.forward(self.to_bf16(X_f32[a:a + chunk].contiguous()))[:, :NOUT]
               for a in range(0, X_f32.shape[0], chunk)]
        if not out:
            return torch.zeros((0, NOUT), dtype=torch.float32, device="cuda")
        return out[0].contiguous() if len(out) == 1 else torch.cat(out)
