"""Learned activation-threshold surrogate with bisection fallback.

An ensemble of multilayer perceptrons maps a fiber's extracellular-potential
profile (Ve per compartment for 1 mA) and its diameter to its activation
threshold; the GPU bisection (``GPUThresholder``, the counterpart of the
reference's ``Thresholder.calculate_thresholds``,
dendra/models/instruments/thresholder.py:873-979) generates the training
thresholds and is the fallback and the judge of the surrogate's error.

Features: the profile divided by its largest magnitude (101 values), the
diameter mapped to [-1, 1], zero padding to 128.  Target: log(threshold *
max|profile|), standardised; since the cable is linear in the stimulus
current, the threshold of a profile scaled by c is the threshold divided by
c, so the magnitude enters exactly through max|profile|.

Members: 102(128) -> 512 -> 512 -> 512 -> 1, GELU, BF16 weights, float32
accumulation, Adam on float32 masters; every kernel is in dn_mlp.hip.  The
forward (inference, and the forward of a training step) is one fused launch
per 128-row tile and member (activations held in LDS, weights read in MFMA
fragment order from a packed BF16 copy); the backward runs the layer
products.
Prediction falls back to bisection where a feature lies outside the range
seen in training or where the ensemble's spread (max - min) / median exceeds
``spread`` (3 %).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib

F_IN, F_PAD, HID = 102, 128, 512
_BF = torch.bfloat16


def _p(t):
    return _lib.ptr(t)


class ThresholdSurrogate:
    def __init__(self, n_members: int = 4, seed: int = 0, device="cuda",
                 n_profile: int = 101, d_center: float = 9.0, d_scale: float = 7.0):
        self.lib = _lib.load()
        self.device = torch.device(device)
        _lib.require_gpu(torch.empty(0, device=self.device))
        if n_profile + 1 > F_PAD:
            raise ValueError("profile too long for the 128-feature input")
        self.nz, self.P = int(n_members), int(n_profile)
        self.dc, self.ds = float(d_center), float(d_scale)
        g = torch.Generator().manual_seed(seed)
        nz, dev = self.nz, self.device

        def init(shape, fan_in, gain=2.0):
            w = torch.randn(shape, generator=g, dtype=torch.float64) * math.sqrt(gain / fan_in)
            return w.to(torch.float32).to(dev)

        self.shapes = {"W1": (HID, F_PAD), "W2": (HID, HID), "W3": (HID, HID), "w4": (1, HID),
                       "b1": (1, HID), "b2": (1, HID), "b3": (1, HID), "b4": (1, 1)}
        W1 = init((nz, HID, F_PAD), F_IN)
        W1[:, :, F_IN:] = 0.0
        self.params = {
            "W1": W1, "W2": init((nz, HID, HID), HID), "W3": init((nz, HID, HID), HID),
            "w4": init((nz, 1, HID), HID, 1.0),
            "b1": torch.zeros((nz, 1, HID), device=dev), "b2": torch.zeros((nz, 1, HID), device=dev),
            "b3": torch.zeros((nz, 1, HID), device=dev), "b4": torch.zeros((nz, 1, 1), device=dev),
        }
        self.grads = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.m = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.bf = {k: torch.empty(v.shape, dtype=_BF, device=dev)
                   for k, v in self.params.items() if k[0] in "Ww"}
        self.bft = {k: torch.empty((nz, v.shape[2], v.shape[1]), dtype=_BF, device=dev)
                    for k, v in self.params.items() if k[0] == "W"}
        # MFMA fragment order for the fused forward (W) and backward (W^T)
        self.bfp = {k: torch.empty(v.shape, dtype=_BF, device=dev)
                    for k, v in self.params.items() if k[0] == "W"}
        self.bfpt = {k: torch.empty(v.shape, dtype=_BF, device=dev) for k in ("W2", "W3")
                     for v in (self.params[k],)}
        self.t_adam = 0
        self.tile = 64    # rows per workgroup of the fused inference forward
        self._refresh_bf16()
        self._plain_stale = False  # bf[W*] / bft behind the parameters
        self.ymean, self.ystd = 0.0, 1.0
        self.lo = self.hi = None
        self._bufs = {}

    # -- helpers ---------------------------------------------------------------
    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    @staticmethod
    def pack(w):
        """[nz][N][K] -> MFMA fragment order [nz][N/32][K/16][64 lanes][8]:
        lane l of block (nb, ks) holds W[32 nb + l % 32][16 ks + 8 (l // 32) + e]."""
        nz, N, K = w.shape
        return (w.reshape(nz, N // 32, 32, K // 16, 2, 8).permute(0, 1, 3, 4, 2, 5)
                .reshape(nz, N, K).contiguous())

    def _refresh_bf16(self):
        for k in self.bf:
            self.bf[k].copy_(self.params[k].to(_BF))
            if k in self.bft:
                self.bft[k].copy_(self.params[k].transpose(1, 2).to(_BF))
                self.bfp[k].copy_(self.pack(self.params[k].to(_BF)))
                if k in self.bfpt:
                    self.bfpt[k].copy_(self.pack(self.bft[k]))

    def _buf(self, name, shape, dtype):
        b = self._bufs.get(name)
        if b is None or b.shape != torch.Size(shape) or b.dtype != dtype:
            b = torch.empty(shape, dtype=dtype, device=self.device)
            self._bufs[name] = b
        return b

    def featurize(self, profile, diam, with_bounds=False):
        """BF16 features [n, 128], log max|profile| [n], out-of-range flags."""
        prof = torch.as_tensor(profile, dtype=torch.float32, device=self.device).contiguous()
        d = torch.as_tensor(diam, dtype=torch.float32, device=self.device).contiguous()
        n = prof.shape[0]
        if prof.shape[1] != self.P or d.shape != (n,):
            raise ValueError("profile must be (n, n_profile) and diam (n,)")
        X = torch.empty((n, F_PAD), dtype=_BF, device=self.device)
        lmag = torch.empty(n, dtype=torch.float32, device=self.device)
        flag = torch.zeros(n, dtype=torch.int32, device=self.device)
        use = with_bounds and self.lo is not None
        st = self.lib.dn_mlp_featurize(n, self.P, _p(prof), _p(d), self.dc, self.ds, _p(X),
                                       F_PAD, _p(lmag), _p(self.lo) if use else None,
                                       _p(self.hi) if use else None, 1e-3,
                                       _p(flag) if use else None, self._stream())
        _lib.check(st, "featurize")
        return X, lmag, flag

    def _gemm(self, epi, M, N, K, A, lda, sa, B, ldb, sb, *, splits=1, bias=None, Z=None,
              ldz=0, sz=0, H=None, ldh=0, sh=0, HT=None, ldht=0, sht=0, Zp=None, ldzp=0,
              szp=0, nz=None):
        st = self.lib.dn_mlp_gemm(epi, nz or self.nz, M, N, K, splits, _p(A), lda, sa, _p(B),
                                  ldb, sb, _p(bias), HID, _p(Z), ldz, sz, _p(H), ldh, sh,
                                  _p(HT), ldht, sht, _p(Zp), ldzp, szp, self._stream())
        _lib.check(st, "mlp gemm")

    # -- forward ----------------------------------------------------------------
    def _fused_forward(self, X, sx, n, y, train=None, tile=None):
        """One fused launch: X [n, 128] (sx = 0, shared) or [nz, n, 128]
        (sx = n * 128); y [nz, n].  train = (Z1, Z2, Z3, HT1, HT2, HT3, t, dy,
        part) with tile 64.  tile: rows per workgroup (128 or 64)."""
        b, bf, bp = self.params, self.bf, self.bfp
        tr = [_p(x) for x in train] if train is not None else [None] * 9
        tile = tile or self.tile
        st = self.lib.dn_mlp_forward(self.nz, n, tile, _p(X), sx, _p(bp["W1"]), _p(bp["W2"]),
                                     _p(bp["W3"]), HID * F_PAD, HID * HID, _p(b["b1"]),
                                     _p(b["b2"]), _p(b["b3"]), _p(bf["w4"]), _p(b["b4"]),
                                     _p(y), n, *tr, self._stream())
        _lib.check(st, "fused forward")

    def _forward_chunk(self, X, n):
        """Member predictions [nz, n] (standardised targets) for features X [n, 128]."""
        y = self._buf("i_y", (self.nz, n), torch.float32)
        self._fused_forward(X, 0, n, y)
        return y

    def _plain_copies(self):
        """The plain and transposed BF16 weights the layer products read; a
        fused training step's Adam leaves them behind (it writes only the
        packed copies the fused kernels read) and they are converted here,
        with the same rounding, when a layer product next needs them."""
        if self._plain_stale:
            for k in self.bft:
                self.bf[k].copy_(self.params[k].to(_BF))
                self.bft[k].copy_(self.params[k].transpose(1, 2).to(_BF))
            self._plain_stale = False

    def _forward_chunk_layers(self, X, n):
        """The same predictions through the layer products (dn_mlp_gemm +
        dn_mlp_out): the unfused form of _forward_chunk."""
        self._plain_copies()
        nz = self.nz
        h1 = self._buf("i_h1", (nz, n, HID), _BF)
        h2 = self._buf("i_h2", (nz, n, HID), _BF)
        y = self._buf("i_y2", (nz, n), torch.float32)
        b, bf = self.params, self.bf
        self._gemm(0, n, HID, F_PAD, X, F_PAD, 0, bf["W1"], F_PAD, HID * F_PAD, bias=b["b1"],
                   H=h1, ldh=HID, sh=n * HID)
        self._gemm(0, n, HID, HID, h1, HID, n * HID, bf["W2"], HID, HID * HID, bias=b["b2"],
                   H=h2, ldh=HID, sh=n * HID)
        self._gemm(0, n, HID, HID, h2, HID, n * HID, bf["W3"], HID, HID * HID, bias=b["b3"],
                   H=h1, ldh=HID, sh=n * HID)
        st = self.lib.dn_mlp_out(nz, n, HID, _p(h1), n * HID, _p(bf["w4"]), HID, _p(b["b4"]),
                                 None, 0, _p(y), None, n, self._stream())
        _lib.check(st, "mlp output")
        return y

    def predict(self, profile, diam, chunk: int = 65536, spread: float = 0.03):
        """Thresholds (mA, median of the members) and fallback flags."""
        X, lmag, flag = self.featurize(profile, diam, with_bounds=True)
        n = X.shape[0]
        thr = torch.empty(n, dtype=torch.float32, device=self.device)
        chunk = max(128, (chunk // 128) * 128)
        for s in range(0, n, chunk):
            e = min(n, s + chunk)
            m = -(-(e - s) // 128) * 128
            if m != e - s:
                Xc = torch.zeros((m, F_PAD), dtype=_BF, device=self.device)
                Xc[: e - s] = X[s:e]
            else:
                Xc = X[s:e]
            y = self._forward_chunk(Xc, m)
            st = self.lib.dn_mlp_ensemble(e - s, self.nz, _p(y), m, _p(lmag[s:]), self.ymean,
                                          self.ystd, spread, _p(thr[s:]), _p(flag[s:]),
                                          self._stream())
            _lib.check(st, "ensemble")
        return thr, flag

    def thresholds(self, profile, diam, judge, chunk: int = 65536, spread: float = 0.03):
        """Thresholds of every fiber: the ensemble median where every feature
        is inside the training range and the members agree within `spread`,
        bisection elsewhere.  judge(indices) returns the bisection thresholds
        of those fibers.  Returns (thresholds float64, fell_back bool)."""
        thr, flag = self.predict(profile, diam, chunk=chunk, spread=spread)
        out = thr.double().cpu().numpy()
        fb = flag.cpu().numpy().astype(bool)
        if fb.any():
            out[fb] = judge(np.flatnonzero(fb))
        return out, fb

    # -- training --------------------------------------------------------------
    def set_data(self, profile, diam, thr):
        """Featurise a training set; record feature bounds and target scaling."""
        X, lmag, _ = self.featurize(profile, diam)
        thr = torch.as_tensor(thr, dtype=torch.float64, device=self.device)
        ok = torch.isfinite(thr)
        if not bool(ok.all()):
            keep = ok.nonzero().squeeze(1)
            X, lmag, thr = X[keep], lmag[keep], thr[keep]
        y = torch.log(thr) + lmag.double()
        self.ymean, self.ystd = float(y.mean()), float(y.std())
        t = ((y - self.ymean) / self.ystd).float().contiguous()
        feats = X[:, : F_IN].float()
        lo = torch.cat([feats.min(0).values, lmag.min().view(1)])
        hi = torch.cat([feats.max(0).values, lmag.max().view(1)])
        self.lo, self.hi = lo.contiguous(), hi.contiguous()
        self.X, self.t = X.contiguous(), t
        return int(X.shape[0])

    def train_step(self, idx, B, lr, b1=0.9, b2=0.999, eps=1e-8, splits=8, fused=True):
        """One Adam step of every member on its minibatch idx[z] (int32 [nz, B]).

        fused: the forward and the backward data products run as one launch
        each over 64-row tiles (dn_mlp_forward, dn_mlp_backward), with the
        bias and output-weight gradients from per-tile partial sums; False
        runs the same arithmetic as layer products (dn_mlp_gemm epilogues 0
        and 1, dn_mlp_out, dn_mlp_out_bwd, dn_mlp_row_dot).  The weight
        gradients are split-K products either way."""
        nz = self.nz
        if fused and B % 64:
            raise ValueError("the fused step needs a batch that is a multiple of 64")
        # weight-gradient products split the batch in multiples of the 64-wide K tile
        while splits > 1 and B % (64 * splits):
            splits //= 2
        bf, bft, P, G = self.bf, self.bft, self.params, self.grads
        buf = self._buf
        X0 = buf("x0", (nz, B, F_PAD), _BF)
        X0T = buf("x0t", (nz, F_PAD, B), _BF)
        t0 = buf("t0", (nz, B), torch.float32)
        Z = [buf(f"z{l}", (nz, B, HID), torch.float32) for l in (1, 2, 3)]
        HT = [buf(f"ht{l}", (nz, HID, B), _BF) for l in (1, 2, 3)]
        dZT = [buf(f"dzt{l}", (nz, HID, B), _BF) for l in (1, 2, 3)]
        y = buf("y", (nz, B), torch.float32)
        dy = buf("dy", (nz, B), torch.float32)
        wpart = buf("wpart", (nz * splits, HID, HID), torch.float32)
        strm = self._stream()
        lib = self.lib
        _lib.check(lib.dn_mlp_gather(nz, B, F_PAD, _p(idx), _p(self.X), _p(self.t), _p(X0),
                                     _p(X0T), _p(t0), strm), "gather")
        if fused:
            tiles = B // 64
            part = buf("part", (nz, tiles, 5, HID), torch.float32)
            # forward: gelu'(pre-activations), transposed activations, dy and
            # the tiles' dw4 / db4 sums
            self._fused_forward(X0, B * F_PAD, B, y, (*Z, *HT, t0, dy, part), tile=64)
            # backward data products dZ3 -> dZ2 -> dZ1 and the tiles' bias sums
            _lib.check(lib.dn_mlp_backward(nz, B, 64, _p(dy), B, _p(bf["w4"]), _p(Z[0]),
                                           _p(Z[1]), _p(Z[2]), _p(self.bfpt["W2"]),
                                           _p(self.bfpt["W3"]), HID * HID, _p(dZT[0]),
                                           _p(dZT[1]), _p(dZT[2]), _p(part), strm), "backward")
            _lib.check(lib.dn_mlp_partials_reduce(nz, tiles, _p(part), _p(G["b1"]), _p(G["b2"]),
                                                  _p(G["b3"]), _p(G["w4"]), _p(G["b4"]), strm),
                       "partials")
        else:
            self._layers_step(X0, t0, Z, HT, dZT, y, dy, B)
        prevT = [X0T, HT[0], HT[1]]
        kin = [F_PAD, HID, HID]
        for l in (3, 2, 1):
            K_in = kin[l - 1]
            self._gemm(2, HID, K_in, B, dZT[l - 1], B, HID * B, prevT[l - 1], B, K_in * B,
                       splits=splits, Z=wpart, ldz=K_in, sz=HID * K_in)
            _lib.check(lib.dn_mlp_splitk_reduce(nz, HID * K_in, splits, _p(wpart),
                                                _p(G[f"W{l}"]), strm), "reduce")
        # Adam: every tensor in one launch
        self.t_adam += 1
        lr_t = lr * math.sqrt(1 - b2 ** self.t_adam) / (1 - b1 ** self.t_adam)
        tab = self._adam_table(fused)
        _lib.check(lib.dn_mlp_adam_multi(len(P), *[a.ctypes.data for a in tab], lr_t, b1, b2,
                                         eps, strm), "adam")
        self._plain_stale = self._plain_stale or fused
        return y, dy

    def _adam_table(self, fused=False):
        """Host arrays of dn_mlp_adam_multi (n, rows, cols, then the device
        addresses of p, g, m, v and the four BF16 copies; 0 = none).  fused:
        only the copies the fused kernels read (packed W, packed W^T, plain
        w4); the plain and transposed matrices are then refreshed lazily."""
        tabs = getattr(self, "_tabs", None)
        if tabs is None:
            tabs = self._tabs = {}
        if fused not in tabs:
            P = self.params
            addr = lambda d, k: (d[k].data_ptr() if k in d else 0)
            plain = {k: v for k, v in self.bf.items() if not fused or k not in self.bft}
            cols = [np.array([v.numel() for v in P.values()], np.int64),
                    np.array([v.shape[1] for v in P.values()], np.int32),
                    np.array([v.shape[2] for v in P.values()], np.int32)]
            for d in (P, self.grads, self.m, self.v, plain, {} if fused else self.bft,
                      self.bfp, self.bfpt):
                cols.append(np.array([addr(d, k) for k in P], np.uint64))
            tabs[fused] = cols
        return tabs[fused]

    def _layers_step(self, X0, t0, Z, HT, dZT, y, dy, B):
        """Forward and backward data products of a training step as layer
        products; fills Z, HT, y, dy, dZT and the bias / output gradients."""
        self._plain_copies()
        nz, lib, strm = self.nz, self.lib, self._stream()
        bf, bft, P, G = self.bf, self.bft, self.params, self.grads
        buf = self._buf
        H = [buf(f"h{l}", (nz, B, HID), _BF) for l in (1, 2, 3)]
        dZ = [buf(f"dz{l}", (nz, B, HID), _BF) for l in (1, 2, 3)]
        ins = [(X0, F_PAD), (H[0], HID), (H[1], HID)]
        for l in range(3):
            A, K = ins[l]
            self._gemm(0, B, HID, K, A, K, B * K, bf[f"W{l + 1}"], K, HID * K,
                       bias=P[f"b{l + 1}"], Z=Z[l], ldz=HID, sz=B * HID, H=H[l], ldh=HID,
                       sh=B * HID, HT=HT[l], ldht=B, sht=HID * B)
        _lib.check(lib.dn_mlp_out(nz, B, HID, _p(H[2]), B * HID, _p(bf["w4"]), HID,
                                  _p(P["b4"]), _p(t0), B, _p(y), _p(dy), B, strm), "out")
        _lib.check(lib.dn_mlp_out_bwd(nz, B, HID, _p(dy), B, _p(bf["w4"]), HID, _p(Z[2]),
                                      B * HID, _p(dZ[2]), _p(dZT[2]), B * HID, strm), "out bwd")
        _lib.check(lib.dn_mlp_row_dot(nz, HID, B, 1, _p(HT[2]), HID * B, _p(dy), B,
                                      _p(G["w4"]), HID, strm), "dw4")
        _lib.check(lib.dn_mlp_row_dot(nz, 1, B, 0, _p(dy), B, None, 0, _p(G["b4"]), 1, strm),
                   "db4")
        for l in (3, 2, 1):
            _lib.check(lib.dn_mlp_row_dot(nz, HID, B, 1, _p(dZT[l - 1]), HID * B, None, 0,
                                          _p(G[f"b{l}"]), HID, strm), "db")
            if l > 1:
                self._gemm(1, B, HID, HID, dZ[l - 1], HID, B * HID, bft[f"W{l}"], HID,
                           HID * HID, Zp=Z[l - 2], ldzp=HID, szp=B * HID, H=dZ[l - 2], ldh=HID,
                           sh=B * HID, HT=dZT[l - 2], ldht=B, sht=HID * B)

    def fit(self, profile, diam, thr, epochs: int = 20, batch: int = 8192,
            lr: float = 2e-3, lr_min: float = 2e-5, seed: int = 1, log=None):
        n = self.set_data(profile, diam, thr)
        if batch % 128 or n < batch:
            raise ValueError("batch must be a multiple of 128 and at most the data size")
        steps_per_epoch = n // batch
        total = epochs * steps_per_epoch
        g = torch.Generator(device=self.device).manual_seed(seed)
        step = 0
        for ep in range(epochs):
            perm = torch.stack([torch.randperm(n, generator=g, device=self.device)
                                for _ in range(self.nz)]).to(torch.int32)
            for s in range(steps_per_epoch):
                lr_s = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * step / total))
                idx = perm[:, s * batch:(s + 1) * batch].contiguous()
                y, dy = self.train_step(idx, batch, lr_s)
                step += 1
            if log is not None:
                # dy = 2 (y - t) / B  ->  mean squared residual of the last minibatch
                r = dy * (batch / 2.0)
                log(ep, float((r * r).mean()))
        return step
