"""Differentiable soft DP over (query, DB profile) pairs: torch binding of the
C ABI entry points cbp_sdp_prepare / cbp_sdp_scores / cbp_sdp_forward /
cbp_sdp_adjoint.

The model is the MAP forward model of the reference
(cumapdp/CuBatchMAPDP_fwd.cu:291-372) with exact log-sum-exp and five
global transition log-offsets theta = (insert open, insert extend, delete
open, delete extend, gap close); logZ = log sum_ij exp(MM(i,j)).  Score
matrices are packed in one flat float32 CUDA tensor, pair k's row-major
[Lq][Ldb] block at offsets[k].
"""
import ctypes as C
import numpy as np
from . import _ffi
from .host import CbpError


class SoftDP:
    """A prepared batch of pairs on one CuBatchProcessing handle.  One forward
    and its adjoint may be in flight at a time (they share the workspace)."""

    def __init__(self, cbp, pairs, ckpt=0):
        import torch
        self.cbp = cbp
        qlen = cbp._queries.lengths()
        dlen = cbp._chunk.lengths()
        self.pairs = [(int(q), int(p)) for q, p in pairs]
        self.shapes = [(int(qlen[q]), int(dlen[p])) for q, p in self.pairs]
        sizes = np.array([a * b for a, b in self.shapes], np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if len(sizes) else np.zeros(0, np.int64)
        self.total = int(sizes.sum())
        arr = (_ffi.SdpPair * max(1, len(self.pairs)))()
        for k, ((q, p), o) in enumerate(zip(self.pairs, self.offsets)):
            arr[k].qidx, arr[k].pronr, arr[k].soff = q, p, int(o)
        wsb = C.c_int64()
        cbp._check(cbp.L.cbp_sdp_prepare(cbp.h, arr, len(self.pairs), ckpt, C.byref(wsb)))
        self.ws_bytes = wsb.value
        self.ws = torch.empty(max(1, self.ws_bytes // 4), dtype=torch.float32, device="cuda")

    @staticmethod
    def _p(t):
        return C.c_void_p(t.data_ptr()) if t is not None else None

    @staticmethod
    def _stream():
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def scores(self):
        """Score matrices of the pairs (plain score + SS term), packed."""
        import torch
        S = torch.empty(self.total, dtype=torch.float32, device="cuda")
        self.cbp._check(self.cbp.L.cbp_sdp_scores(self.cbp.h, self._p(S), self._stream()))
        return S

    def forward(self, S, theta):
        import torch
        assert S.is_cuda and S.dtype == torch.float32 and S.numel() == self.total and S.is_contiguous()
        theta = theta.detach().float().contiguous()
        logZ = torch.empty(len(self.pairs), dtype=torch.float32, device="cuda")
        self.cbp._check(self.cbp.L.cbp_sdp_forward(self.cbp.h, self._p(S), self._p(theta), self._p(logZ),
                                                   self._p(self.ws), self._stream()))
        return logZ

    def adjoint(self, S, theta, logZ, gZ=None, want_post=False):
        """(dS, dtheta per pair [n][5], posterior or None)."""
        import torch
        theta = theta.detach().float().contiguous()
        dS = torch.empty_like(S)
        dth = torch.empty((len(self.pairs), 5), dtype=torch.float32, device="cuda")
        post = torch.empty_like(S) if want_post else None
        gz = gZ.detach().float().contiguous() if gZ is not None else None
        self.cbp._check(self.cbp.L.cbp_sdp_adjoint(self.cbp.h, self._p(S), self._p(theta), self._p(logZ),
                                                   self._p(gz), self._p(dS), self._p(dth), self._p(post),
                                                   self._p(self.ws), self._stream()))
        return dS, dth, post


def _make_function():
    import torch

    class SoftDPFunction(torch.autograd.Function):
        """logZ = SoftDPFunction.apply(S, theta, sdp): gradients for S (packed
        float32) and theta [5] by the hand-written adjoint kernel."""

        @staticmethod
        def forward(ctx, S, theta, sdp):
            logZ = sdp.forward(S.detach(), theta)
            ctx.save_for_backward(S, theta, logZ)
            ctx.sdp = sdp
            return logZ

        @staticmethod
        def backward(ctx, gZ):
            S, theta, logZ = ctx.saved_tensors
            dS, dth, _ = ctx.sdp.adjoint(S.detach(), theta, logZ, gZ.contiguous())
            return dS, dth.sum(0).to(theta.dtype), None

    return SoftDPFunction


_fn = None


def soft_dp(S, theta, sdp):
    """Differentiable logZ [npairs] for packed scores S and theta [5]."""
    global _fn
    if _fn is None:
        _fn = _make_function()
    return _fn.apply(S, theta, sdp)
