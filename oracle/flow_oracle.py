"""Float64 reference of the learned proposal (TEST INFRASTRUCTURE ONLY).

Restates pibronic_amd.flow's construction with torch on the CPU in float64, with autograd
for the training gradient: whitening Z = (R - m) U / sigma into the permuted even-k/odd-k
layout, affine coupling layers y_B = z_B exp(s) + t with s = smax tanh(raw/smax) and an
MLP D/2 -> h -> h -> D (ReLU), log rho of the harmonic mixture (the O-matrix product of
pibronic/pimc/pimc.py:1087-1136 in log form), and log q(R) = log rho(R0) + log|det dF^-1|.
"""
import numpy as np
import torch

DT = torch.float64


class FlowRef:
    def __init__(self, U, sigma, m, coth, csch, pref, shift, layers, hidden, smax):
        self.U = torch.as_tensor(U, dtype=DT)
        self.sigma = torch.as_tensor(sigma, dtype=DT)
        self.m = torch.as_tensor(m, dtype=DT)
        self.coth = torch.as_tensor(coth, dtype=DT)
        self.csch = torch.as_tensor(csch, dtype=DT)
        self.pref = torch.as_tensor(pref, dtype=DT)
        self.shift = torch.as_tensor(shift, dtype=DT)
        self.N, self.P = self.sigma.shape
        self.D = self.N * self.P
        self.H = self.D // 2
        self.L, self.h, self.smax = layers, hidden, smax

    def unpack(self, flat):
        """Split the flat parameter vector (same layout as pibronic_amd.flow.Flow)."""
        flat = torch.as_tensor(flat, dtype=DT)
        H, h, D = self.H, self.h, self.D
        shapes = [(H, h), (h,), (h, h), (h,), (h, D), (D,)]
        out, off = [], 0
        for _ in range(self.L):
            layer = []
            for s in shapes:
                n = int(np.prod(s))
                layer.append(flat[off:off + n].view(*s))
                off += n
            out.append(layer)
        return out

    def halves(self, l):
        return (0, self.H) if l % 2 == 0 else (self.H, 0)

    # permuted layout: zp[:, h*D/2 + j*P/2 + kk] = Z[:, j, 2kk + h]
    def perm(self, Z):
        B = Z.shape[0]
        return torch.cat([Z[:, :, 0::2].reshape(B, -1), Z[:, :, 1::2].reshape(B, -1)], dim=1)

    def unperm(self, zp):
        B = zp.shape[0]
        Z = torch.empty((B, self.N, self.P), dtype=zp.dtype)
        Z[:, :, 0::2] = zp[:, :self.H].reshape(B, self.N, self.P // 2)
        Z[:, :, 1::2] = zp[:, self.H:].reshape(B, self.N, self.P // 2)
        return Z

    def whiten(self, R):
        return self.perm(((R - self.m[None, :, None]) @ self.U) / self.sigma)

    def unwhiten(self, zp):
        return self.m[None, :, None] + (self.unperm(zp) * self.sigma) @ self.U.T

    def mlp(self, x, W1, b1, W2, b2, W3, b3):
        h1 = torch.relu(x @ W1 + b1)
        h2 = torch.relu(h1 @ W2 + b2)
        return h2 @ W3 + b3

    def forward(self, R0, flat):
        prm = self.unpack(flat)
        zp = self.whiten(R0)
        logdet = torch.zeros(R0.shape[0], dtype=DT)
        for l in range(self.L):
            a, b = self.halves(l)
            o = self.mlp(zp[:, a:a + self.H], *prm[l])
            s = self.smax * torch.tanh(o[:, :self.H] / self.smax)
            t = o[:, self.H:]
            zb = zp[:, b:b + self.H] * torch.exp(s) + t
            zp = torch.cat([zp[:, :b], zb, zp[:, b + self.H:]], dim=1) if b else torch.cat([zb, zp[:, self.H:]], dim=1)
            logdet = logdet + s.sum(dim=1)
        return self.unwhiten(zp), logdet

    def log_rho(self, R0):
        x = R0[:, None, :, :] - self.shift[None, :, :, None]          # (B, A, N, P)
        xn = torch.roll(x, shifts=-1, dims=3)
        e = self.coth[None, None, :, None] * (x * x + xn * xn) - 2.0 * self.csch[None, None, :, None] * x * xn
        la = self.P * torch.log(self.pref)[None, :] - 0.5 * e.sum(dim=(2, 3))
        return torch.logsumexp(la, dim=1)

    def log_q(self, R, flat):
        prm = self.unpack(flat)
        zp = self.whiten(R)
        logdet = torch.zeros(R.shape[0], dtype=DT)
        for l in reversed(range(self.L)):
            a, b = self.halves(l)
            o = self.mlp(zp[:, a:a + self.H], *prm[l])
            s = self.smax * torch.tanh(o[:, :self.H] / self.smax)
            t = o[:, self.H:]
            zb = (zp[:, b:b + self.H] - t) * torch.exp(-s)
            zp = torch.cat([zp[:, :b], zb, zp[:, b + self.H:]], dim=1) if b else torch.cat([zb, zp[:, self.H:]], dim=1)
            logdet = logdet - s.sum(dim=1)
        R0 = self.unwhiten(zp)
        return self.log_rho(R0) + logdet, R0

    def loss_and_grad(self, R, c, flat):
        p = torch.as_tensor(flat, dtype=DT).clone().requires_grad_(True)
        lq, _ = self.log_q(torch.as_tensor(R, dtype=DT), p)
        loss = -(torch.as_tensor(c, dtype=DT) * lq).sum()
        loss.backward()
        return loss.item(), p.grad.numpy()
