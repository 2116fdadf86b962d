"""Learned proposal: an affine-coupling flow that replaces the harmonic rho as the sampling
distribution of the PIMC estimator.

Construction. Let R0 ~ rho (the reference's harmonic sampling distribution). Whiten it into
the ring's normal-mode coordinates,
Z = (R0 - m) U / sigma (U the circulant eigenvectors, sigma the harmonic standard deviations,
m the weighted mean of the surfaces' displacements), apply L affine coupling layers
(RealNVP, Dinh, Sohl-Dickstein & Bengio 2017) y = F(Z), and map back, R = m + (y sigma) U^T.
The linear maps cancel in the Jacobian, so the proposal density in the reference's
normalisation is

    log q(R) = log rho(R0) - log|det dF/dZ|,

and w = g(R) / q(R) has the same expectation Z/Z_rho as the harmonic estimator g/rho. The
last layer of every coupling network starts at zero, so an untrained flow is exactly rho.

Layers alternate the even-k and odd-k normal modes as the transformed half; each coupling
network is D/2 -> hidden -> hidden -> D (ReLU), giving raw scales and shifts for the other
half, with scales bounded by smax * tanh(raw / smax).

Draws start in the whitened coordinates: there rho is a mixture of unit Gaussians whose means
differ between surfaces (pib_flow_draw_base), so Z and log rho(R0) come from one kernel and R0
itself is never formed.

Training minimises KL(p || q) for p proportional to |g|: the loss of a batch with fixed
weights c_b (summing to one) is -sum_b c_b log q(R_b), differentiated through the inverse
flow and the analytic gradient of log rho. Samples and weights come from the method itself:
first from rho (weights g/rho), then from the flow (weights g/q).

Every operation runs in hand-written HIP kernels behind the C-ABI (MFMA GEMMs, coupling,
whitening, log rho, Adam); torch only holds the device buffers.
"""
import math

import numpy as np

from . import _native


class Flow:
    """gemm: "f32" (FP32 MFMA) or "bf16x3" (bf16 matrix cores with a hi/lo operand split,
    ~2^-16 relative error per product term) for the coupling networks' dense layers. The
    whitening transforms always run in FP32. Either way log q is exact for the map computed."""

    def __init__(self, data, layers=8, hidden=256, smax=2.0, seed=0, gemm="f32"):
        import torch
        if gemm not in ("f32", "bf16x3"):
            raise ValueError("gemm must be 'f32' or 'bf16x3'")
        self.gemm = gemm
        self.data = data
        self.ctx = data.context()
        self.N, self.P = data.modes, data.beads
        self.D = self.N * self.P
        self.H = self.D // 2
        self.layers, self.hidden, self.smax = layers, hidden, float(smax)
        self.dev = torch.device("cuda", data.device)
        rho = data.rho
        sigma = rho.standard_deviation[0]                 # (N, P), identical for all surfaces
        m = rho.state_weight @ rho.state_shift            # (N,)
        self.ctx.flow_set_basis(data.circulant_eigvects, sigma, m)
        self._set_base(rho, sigma, m)
        # parameter layout per layer: W1 (H x h), b1, W2 (h x h), b2, W3 (h x D), b3 (D)
        H, h, D = self.H, hidden, self.D
        self.shapes = [("W1", (H, h)), ("b1", (h,)), ("W2", (h, h)), ("b2", (h,)),
                       ("W3", (h, D)), ("b3", (D,))]
        self.per_layer = sum(int(np.prod(s)) for _, s in self.shapes)
        self.n_params = self.per_layer * layers
        rng = np.random.default_rng(seed)
        flat = np.zeros(self.n_params, dtype=np.float32)
        for l in range(layers):
            for name, shp in self.shapes:
                off, n = self._offset(l, name), int(np.prod(shp))
                if name in ("W1", "W2"):  # He initialisation for the ReLU layers
                    flat[off:off + n] = rng.normal(0.0, math.sqrt(2.0 / shp[0]), n)
        self.params = torch.from_numpy(flat).to(self.dev)
        self.grad = torch.zeros_like(self.params)
        self.adam_m = torch.zeros_like(self.params)
        self.adam_v = torch.zeros_like(self.params)
        self.step_count = 0
        self._bufs = {}
        self._version = 0          # bumped whenever the parameters change
        self.use_fused = True      # bf16x3 draws and densities: one launch per coupling layer where supported
        # draws only: "bf16x3" (the trained function to FP32 accuracy) or "bf16" (single bf16
        # products in the one-launch layers; log q stays exact for the map applied, see draw())
        self.draw_precision = "bf16x3"
        self._split_version = -1   # parameter version of the pre-split bf16 weights
        self._fsplit_version = -1  # parameter version of the fused layer's permuted weights
        self.use_frag = True       # single-bf16 layers at D = 512, 1024 on fragment-ordered weights
        self._frag_version = -1    # parameter version of those weights

    def _set_base(self, rho, sigma, m):
        """rho in the flow's coordinates Z = (R0 - m) U / sigma: the mixture over surfaces a
        (weights w_a) of N(mu_a, I), mu_a,jk = (d_aj - m_j) (sum_p U[p, k]) / sigma_jk, with
        log rho(R0) = logsumexp_a (c_a + Z . mu_a) - |Z|^2 / 2, c_a = P log F_a - |mu_a|^2 / 2
        (pimc.py:1087-1136: the O-matrix exponent summed over the ring is
        (x - d)^T (2 coth I - csch C) (x - d), whose eigenbasis is U with 1 / sigma^2 as
        eigenvalues)."""
        U = np.asarray(self.data.circulant_eigvects, dtype=np.float64)
        shift = np.asarray(rho.state_shift, dtype=np.float64)           # (A_rho, N)
        pref = np.asarray(rho.const.omatrix_prefactor, dtype=np.float64)  # (A_rho,)
        if np.any(pref <= 0.0):
            raise ValueError("flow base distribution needs positive O-matrix prefactors")
        colsum = U.sum(axis=0)                                            # (P,)
        # every eigenvector but the constant one is orthogonal to the constant vector: their
        # column sums are rounding noise (~1e-16); set them to exactly 0, so that mu is non-zero
        # only in the constant normal mode of each bead mode and the draw skips the rest
        colsum[np.abs(colsum) < 1e-12 * np.abs(colsum).max()] = 0.0
        mu = (shift - m[None, :])[:, :, None] * colsum[None, None, :] / np.asarray(sigma)[None]
        mu_p = np.concatenate([mu[:, :, 0::2].reshape(len(pref), -1), mu[:, :, 1::2].reshape(len(pref), -1)], axis=1)
        c = self.P * np.log(pref) - 0.5 * np.sum(mu_p * mu_p, axis=1)
        self.ctx.flow_set_base(mu_p, c)

    # ------------------------------------------------------------------ layout helpers
    def _offset(self, layer, name):
        off = layer * self.per_layer
        for n, shp in self.shapes:
            if n == name:
                return off
            off += int(np.prod(shp))
        raise KeyError(name)

    def p(self, layer, name, src=None):
        t = self.params if src is None else src
        shp = dict(self.shapes)[name]
        off = self._offset(layer, name)
        return t[off:off + int(np.prod(shp))].view(*shp)

    def halves(self, layer):
        """(a_off, b_off): even layers transform the odd-k half, odd layers the even-k half."""
        return (0, self.H) if layer % 2 == 0 else (self.H, 0)

    def set_params(self, flat):
        import torch
        self.params.copy_(torch.as_tensor(np.asarray(flat, dtype=np.float32)))
        self.invalidate()

    def invalidate(self):
        """Call after writing self.params directly (set_params and adam_step do it)."""
        self._version += 1

    def _buf(self, name, shape, dtype=None):
        import torch
        dtype = torch.float32 if dtype is None else dtype
        key = (name, tuple(shape), dtype)
        if key not in self._bufs:
            self._bufs[key] = torch.empty(shape, dtype=dtype, device=self.dev)
        return self._bufs[key]

    def _check_batch(self, B):
        if B % 64:
            raise ValueError("flow batches must be a multiple of 64 samples")

    def _gemm(self, *args, **kw):
        fn = self.ctx.gemm_bf16x3 if self.gemm == "bf16x3" else self.ctx.gemm_f32
        fn(*args, **kw)

    # ------------------------------------------------------------------ MLP of one layer
    def _mlp(self, layer, zp, B, h1, h2, o, stream):
        a_off, _ = self.halves(layer)
        H, h, D = self.H, self.hidden, self.D
        self._gemm(zp[:, a_off:], self.p(layer, "W1"), h1, B, h, H, D, h, h,
                   bias=self.p(layer, "b1"), relu=True, stream=stream)
        self._gemm(h1, self.p(layer, "W2"), h2, B, h, h, h, h, h,
                   bias=self.p(layer, "b2"), relu=True, stream=stream)
        self._gemm(h2, self.p(layer, "W3"), o, B, D, h, h, D, D,
                   bias=self.p(layer, "b3"), stream=stream)

    # ------------------------------------------------------------------ bf16x3 fast path
    def _split_weights(self, stream):
        """Weights split once per parameter version into bf16 hi/lo, stored transposed (N x K)
        as the pre-split product expects."""
        import torch
        if self._split_version == self._version:
            return
        H, h, D = self.H, self.hidden, self.D
        self._wsplit = []
        for l in range(self.layers):
            mats = []
            for name, (K, Nn) in (("W1", (H, h)), ("W2", (h, h)), ("W3", (h, D))):
                hi = self._buf(f"{name}h{l}", (Nn, K), torch.int16)
                lo = self._buf(f"{name}l{l}", (Nn, K), torch.int16)
                self.ctx.split_bf16(K, Nn, self.p(l, name), Nn, hi, lo, K, transpose=True, stream=stream)
                mats.append((hi, lo))
            self._wsplit.append(mats)
        self._split_version = self._version

    def _split_weights_fused(self, stream):
        """The fused layer's operands: W1 as for the pre-split products, W2 and W3 split with
        their input index permuted within blocks of 32 (PIB_SPLIT_PERM32)."""
        import torch
        self._split_weights(stream)
        if self._fsplit_version == self._version:
            return
        h, D = self.hidden, self.D
        self._wfused = []
        for l in range(self.layers):
            mats = [self._wsplit[l][0]]
            for name, (K, Nn) in (("W2", (h, h)), ("W3", (h, D))):
                hi = self._buf(f"{name}ph{l}", (Nn, K), torch.int16)
                lo = self._buf(f"{name}pl{l}", (Nn, K), torch.int16)
                self.ctx.split_bf16(K, Nn, self.p(l, name), Nn, hi, lo, K, transpose=True, perm32=True,
                                    stream=stream)
                mats.append((hi, lo))
            self._wfused.append(mats)
        self._fsplit_version = self._version

    def _mlp_fast(self, layer, zp, B, o, stream):
        """The coupling network on pre-split operands: the conditioning half of zp is split
        once, the hidden activations travel as bf16 hi/lo written by the previous product."""
        import torch
        a_off, _ = self.halves(layer)
        H, h, D = self.H, self.hidden, self.D
        ctx = self.ctx
        (w1h, w1l), (w2h, w2l), (w3h, w3l) = self._wsplit[layer]
        zh, zl = self._buf("zh", (B, H), torch.int16), self._buf("zl", (B, H), torch.int16)
        ah, al = self._buf("ah", (B, h), torch.int16), self._buf("al", (B, h), torch.int16)
        bh, bl = self._buf("bh", (B, h), torch.int16), self._buf("bl", (B, h), torch.int16)
        ctx.split_bf16(B, H, zp[:, a_off:], D, zh, zl, H, stream=stream)
        ctx.gemm_bf16x3_presplit(B, h, H, zh, zl, H, w1h, w1l, H, O_hi=ah, O_lo=al, ldo=h,
                                 bias=self.p(layer, "b1"), relu=True, stream=stream)
        ctx.gemm_bf16x3_presplit(B, h, h, ah, al, h, w2h, w2l, h, O_hi=bh, O_lo=bl, ldo=h,
                                 bias=self.p(layer, "b2"), relu=True, stream=stream)
        ctx.gemm_bf16x3_presplit(B, D, h, bh, bl, h, w3h, w3l, h, C=o, ldc=D,
                                 bias=self.p(layer, "b3"), stream=stream)

    def _fused_ok(self, B, single_bf16=False):
        """The one-launch layer runs hidden 256 at D = 512, and with single bf16 products
        also at D = 1024; B a multiple of 128."""
        return (self.use_fused and self.hidden == 256 and B % 128 == 0
                and (self.H == 256 or (single_bf16 and self.H == 512)))

    def _split_weights_frag(self, stream):
        """bf16 weights (hi parts) in the fragment order of pib_flow_layer_fused_frag, inputs in
        natural order."""
        import torch
        if self._frag_version == self._version:
            return
        H, h, D = self.H, self.hidden, self.D
        self._wfrag = []
        for l in range(self.layers):
            mats = []
            for name, (K, Nn) in (("W1", (H, h)), ("W2", (h, h)), ("W3", (h, D))):
                f = self._buf(f"{name}f{l}", (Nn * K,), torch.int16)
                self.ctx.split_bf16(K, Nn, self.p(l, name), Nn, f, None, K, transpose=True, frag32=True,
                                    stream=stream)
                mats.append(f)
            self._wfrag.append(mats)
        self._frag_version = self._version

    def _layer_fused(self, inverse, layer, zp, B, logdet, stream, single_bf16=False):
        """The whole coupling layer in one launch (pib_flow_layer_fused); single_bf16: one bf16
        product per weight block (the hi parts) instead of BF16x3, at D = 512 and 1024 (with
        use_frag) on fragment-ordered weights (pib_flow_layer_fused_frag)."""
        a_off, b_off = self.halves(layer)
        if single_bf16 and self.use_frag and self.D in (512, 1024):
            self._split_weights_frag(stream)
            self.ctx.flow_layer_fused_frag(inverse, B, self.D, self.hidden, a_off, b_off, self.smax, zp, logdet,
                                           self._wfrag[layer], self.p(layer, "b1"), self.p(layer, "b2"),
                                           self.p(layer, "b3"), stream)
            return
        self._split_weights_fused(stream)
        w = self._wfused[layer]
        if single_bf16:
            w = tuple((hi, None) for hi, _ in w)
        self.ctx.flow_layer_fused(inverse, B, self.D, self.hidden, a_off, b_off, self.smax, zp, logdet,
                                  w, self.p(layer, "b1"), self.p(layer, "b2"), self.p(layer, "b3"), stream)

    # ------------------------------------------------------------------ directions
    def forward(self, R0, stream=None):
        """Sampling direction: R0 (B, N, P) double -> R (B, N, P) double, log|det dF/dZ| (B,)."""
        import torch
        B = R0.shape[0]
        self._check_batch(B)
        D = self.D
        zp = self._buf("zp", (B, D))
        work = self._buf("work", (2 * B * D,))
        logdet = torch.zeros(B, dtype=torch.float64, device=self.dev)
        R = torch.empty_like(R0)
        self.ctx.flow_whiten(_native.FLOW_WHITEN, B, R0, work, zp, stream)
        self._layers_forward(zp, B, logdet, stream)
        self.ctx.flow_whiten(_native.FLOW_UNWHITEN, B, zp, work, R, stream)
        return R, logdet

    def _layers_forward(self, zp, B, logdet, stream, single_bf16=False):
        """The coupling layers in the sampling direction on zp (B, D) in place; adds
        log|det| to logdet. single_bf16 (one-launch layers only): single bf16 products."""
        D, h = self.D, self.hidden
        h1, h2, o = self._buf("h1", (B, h)), self._buf("h2", (B, h)), self._buf("o", (B, D))
        fast = self.gemm == "bf16x3"
        if single_bf16 and not (fast and self._fused_ok(B, True)):
            raise ValueError("single-bf16 layers need the one-launch layer (gemm='bf16x3', hidden 256, "
                             "D = 512 or 1024, B a multiple of 128)")
        if fast:
            self._split_weights(stream)
        for l in range(self.layers):
            a_off, b_off = self.halves(l)
            if fast and self._fused_ok(B, single_bf16):
                self._layer_fused(False, l, zp, B, logdet, stream, single_bf16)
                continue
            if fast:
                self._mlp_fast(l, zp, B, o, stream)
            else:
                self._mlp(l, zp, B, h1, h2, o, stream)
            self.ctx.flow_coupling(False, B, D, a_off, b_off, self.smax, zp, o, logdet, stream)

    def log_q(self, R, stream=None):
        """Density direction: log q(R) = log rho(R0) + log|det dF^-1/dy| with R0 = T^-1(R)."""
        import torch
        B = R.shape[0]
        self._check_batch(B)
        D, h = self.D, self.hidden
        zp = self._buf("zp", (B, D))
        work = self._buf("work", (2 * B * D,))
        h1, h2, o = self._buf("h1", (B, h)), self._buf("h2", (B, h)), self._buf("o", (B, D))
        logdet = torch.zeros(B, dtype=torch.float64, device=self.dev)
        R0 = torch.empty_like(R)
        self.ctx.flow_whiten(_native.FLOW_WHITEN, B, R, work, zp, stream)
        fast = self.gemm == "bf16x3"
        if fast:
            self._split_weights(stream)
        for l in reversed(range(self.layers)):
            a_off, b_off = self.halves(l)
            if fast and self._fused_ok(B):
                self._layer_fused(True, l, zp, B, logdet, stream)
                continue
            if fast:
                self._mlp_fast(l, zp, B, o, stream)
            else:
                self._mlp(l, zp, B, h1, h2, o, stream)
            self.ctx.flow_coupling(True, B, D, a_off, b_off, self.smax, zp, o, logdet, stream)
        self.ctx.flow_whiten(_native.FLOW_UNWHITEN, B, zp, work, R0, stream)
        logrho = torch.empty(B, dtype=torch.float64, device=self.dev)
        self.ctx.flow_logrho(B, R0, logrho=logrho, stream=stream)
        return logrho + logdet, R0

    def loss_and_grad(self, R, c, stream=None):
        """Loss -sum_b c_b log q(R_b) and its gradient (into self.grad) for fixed R, c."""
        import torch
        B = R.shape[0]
        self._check_batch(B)
        D, H, h, L = self.D, self.H, self.hidden, self.layers
        ctx = self.ctx
        c = c.to(dtype=torch.float64)
        ys = self._buf("ys", (L, B, D))
        h1s, h2s, os_ = self._buf("h1s", (L, B, h)), self._buf("h2s", (L, B, h)), self._buf("os", (L, B, D))
        zp = self._buf("zp", (B, D))
        work = self._buf("work", (2 * B * D,))
        logdet = torch.zeros(B, dtype=torch.float64, device=self.dev)
        R0 = self._buf("R0", (B, self.N, self.P), torch.float64)
        ctx.flow_whiten(_native.FLOW_WHITEN, B, R, work, zp, stream)
        for l in reversed(range(L)):
            a_off, b_off = self.halves(l)
            ys[l].copy_(zp)
            self._mlp(l, zp, B, h1s[l], h2s[l], os_[l], stream)
            ctx.flow_coupling(True, B, D, a_off, b_off, self.smax, zp, os_[l], logdet, stream)
        ctx.flow_whiten(_native.FLOW_UNWHITEN, B, zp, work, R0, stream)
        logrho = torch.empty(B, dtype=torch.float64, device=self.dev)
        dR0 = self._buf("dR0", (B, self.N, self.P))
        ctx.flow_logrho(B, R0, c=c, dR=dR0, logrho=logrho, stream=stream)
        loss = -(c * (logrho + logdet)).sum()
        dz = self._buf("dz", (B, D))
        ctx.flow_whiten(_native.FLOW_GRAD, B, dR0, work, dz, stream)
        dout = self._buf("dout", (B, D))
        dh2, dh1 = self._buf("dh2", (B, h)), self._buf("dh1", (B, h))
        g = self.grad
        for l in range(L):
            a_off, b_off = self.halves(l)
            ctx.flow_coupling_inverse_backward(B, D, a_off, b_off, self.smax, ys[l], os_[l], dz,
                                               dout, c, stream)
            # o = h2 W3 + b3
            self._gemm(h2s[l], dout, self.p(l, "W3", g), h, D, B, h, D, D, trans_a=True, stream=stream)
            ctx.colsum_f32(B, D, dout, D, self.p(l, "b3", g), stream)
            self._gemm(dout, self.p(l, "W3"), dh2, B, h, D, D, D, h, trans_b=True,
                         mask=h2s[l], ldmask=h, stream=stream)
            # h2 = relu(h1 W2 + b2)
            self._gemm(h1s[l], dh2, self.p(l, "W2", g), h, h, B, h, h, h, trans_a=True, stream=stream)
            ctx.colsum_f32(B, h, dh2, h, self.p(l, "b2", g), stream)
            self._gemm(dh2, self.p(l, "W2"), dh1, B, h, h, h, h, h, trans_b=True,
                         mask=h1s[l], ldmask=h, stream=stream)
            # h1 = relu(y_A W1 + b1)
            self._gemm(ys[l][:, a_off:], dh1, self.p(l, "W1", g), H, h, B, D, h, h, trans_a=True,
                         stream=stream)
            ctx.colsum_f32(B, h, dh1, h, self.p(l, "b1", g), stream)
            self._gemm(dh1, self.p(l, "W1"), dz[:, a_off:], B, H, h, h, h, D, trans_b=True,
                         accumulate=True, stream=stream)
        return loss

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        self.step_count += 1
        self._version += 1
        self.ctx.adam_f32(self.n_params, self.params, self.grad, self.adam_m, self.adam_v, lr,
                          beta1, beta2, eps, self.step_count, stream)

    # ------------------------------------------------------------------ proposal draws
    def offsets_ok(self):
        """Whether draws can be returned as float offsets from the basis centre and evaluated
        as such (pib_eval_block_flow_offsets: the FP64 lanes kernel, A_rho = A <= 8; the fused
        unwhitening: P = 64 or 128)."""
        d = self.data
        return self.P in (64, 128) and d.states <= 8 and d.rho.states == d.states

    def draw(self, B, seed, offset, stream=None, layer_events=None, offsets=False):
        """B proposal samples: R (B, N, P) double and log q(R) (B,). The base sample is drawn
        directly in the flow's coordinates with its log rho (pib_flow_draw_base).
        layer_events: optional pair of torch.cuda.Event recorded on `stream` around the
        coupling layers (to time them apart from the base draw and the unwhitening).
        offsets: return float offsets dR (B, N, P) from the basis centre m instead of R
        (R = m + dR exactly as the double path forms it; BoxData.evaluate_offsets_device takes
        them), half the bytes written and read.

        With draw_precision "bf16" the layers' networks run as single bf16 products. log q is
        still exact for the map applied: a coupling layer passes its conditioning half through
        unchanged, so its log-determinant is the sum of the scales it applied, whatever
        function of that half computed them; the weights g/q stay unbiased, and only how well
        q matches the trained function (the ESS) depends on the precision."""
        import torch
        self._check_batch(B)
        zp = self._buf("zp", (B, self.D))
        logrho = torch.empty(B, dtype=torch.float64, device=self.dev)
        self.ctx.flow_draw_base(seed, offset, B, zp, logrho, stream)
        logdet = torch.zeros(B, dtype=torch.float64, device=self.dev)
        if layer_events is not None:
            layer_events[0].record(stream)
        self._layers_forward(zp, B, logdet, stream, single_bf16=self.draw_precision == "bf16")
        if layer_events is not None:
            layer_events[1].record(stream)
        work = self._buf("work", (2 * B * self.D,))
        if offsets:
            dR = torch.empty((B, self.N, self.P), dtype=torch.float32, device=self.dev)
            self.ctx.flow_whiten(_native.FLOW_UNWHITEN_OFFSETS, B, zp, work, dR, stream)
            return dR, logrho - logdet
        R = torch.empty((B, self.N, self.P), dtype=torch.float64, device=self.dev)
        self.ctx.flow_whiten(_native.FLOW_UNWHITEN, B, zp, work, R, stream)
        return R, logrho - logdet


def log_weights(data, R, logq, stream=None):
    """log|g(R)| - log q(R) and sign(g) for proposal samples, g from the FP64 kernel."""
    import torch
    B = R.shape[0]
    out = torch.empty((_native.NUM_OUT, B), dtype=torch.float64, device=R.device)
    data.evaluate_device(R, B, False, out, B, noscale=False, stream=stream)
    g = out[_native.OUT_G]
    return torch.log(torch.abs(g)) + out[_native.OUT_LOGS] - logq, torch.sign(g)


def harmonic_log_weights(data, R, stream=None):
    """log|g/rho| and sign for rho samples (the harmonic estimator's weights)."""
    import torch
    B = R.shape[0]
    out = torch.empty((_native.NUM_OUT, B), dtype=torch.float64, device=R.device)
    data.evaluate_device(R, B, False, out, B, noscale=False, stream=stream)
    return torch.log(torch.abs(out[_native.OUT_G] / out[_native.OUT_RHO])), torch.sign(out[_native.OUT_G])
