"""ctypes binding of libpibronic_hip.so (the C-ABI declared in include/pibronic_hip.h).

The library is the only compute path of this package: there is no NumPy or PyTorch
fallback. If the shared object is missing or cannot be loaded, `load()` raises.
"""
import ctypes
import os

import numpy as np

LIB_NAME = "libpibronic_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

ABI_VERSION = 1

PIB_OK = 0
PIB_ERR_ARG = -1
PIB_ERR_HIP = -2
PIB_ERR_STATE = -3
PIB_ERR_UNSUPPORTED = -4

SET_RHO, SET_VIB, SET_VIB_PLUS, SET_VIB_MINUS = 0, 1, 2, 3
EVAL_PM, EVAL_NOSCALE, EVAL_WIDE, EVAL_FP32, EVAL_BF16X3 = 1, 2, 4, 8, 16
BATCH_LANES, BATCH_WIDE, BATCH_SEQUENTIAL = 0, 1, 2
SPLIT_TRANSPOSE, SPLIT_PERM32, SPLIT_FRAG32 = 1, 2, 4
OUT_RHO, OUT_G, OUT_GP, OUT_GM, OUT_LOGS = 0, 1, 2, 3, 4
NUM_OUT = 5
(SUM_N, SUM_R, SUM_R2, SUM_ABSR, SUM_NEG, SUM_D1, SUM_D1SQ, SUM_D2, SUM_D2SQ,
 SUM_RD1, SUM_RD2, SUM_D1D2, SUM_RP, SUM_RM) = range(14)
NUM_SUMS = 14

# exported symbol -> (restype, argtypes); must match include/pibronic_hip.h
_c_int, _c_i64, _c_u64, _c_dbl, _vp = ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double, ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "pib_abi_version": (_c_int, []),
    "pib_ctx_create": (_c_int, [_c_int, ctypes.POINTER(_vp)]),
    "pib_ctx_destroy": (_c_int, [_vp]),
    "pib_last_error": (ctypes.c_char_p, [_vp]),
    "pib_set_coupling": (_c_int, [_vp, _c_int, _c_int, _dp, _dp, _dp]),
    "pib_set_omatrix_set": (_c_int, [_vp, _c_int, _c_int, _c_int, _dp, _dp, _dp, _dp]),
    "pib_set_beads": (_c_int, [_vp, _c_int, _c_dbl]),
    "pib_set_lanes_per_sample": (_c_int, [_vp, _c_int]),
    "pib_get_lanes_per_sample": (_c_int, [_vp, _c_i64]),
    "pib_eval_block": (_c_int, [_vp, _vp, _c_i64, _c_int, _vp, _c_i64, _vp, _vp]),
    "pib_eval_block_flow_offsets": (_c_int, [_vp, _vp, _c_i64, _c_int, _vp, _c_i64, _vp]),
    "pib_set_sampler": (_c_int, [_vp, _c_int, _c_int, _c_int, _dp, _dp, _dp, _dp]),
    "pib_sample_block": (_c_int, [_vp, _c_u64, _c_i64, _c_i64, _vp, _vp, _vp]),
    "pib_block_sums": (_c_int, [_vp, _vp, _c_i64, _c_i64, _c_int, _c_dbl, _vp, _vp]),
    "pib_block_sums_q": (_c_int, [_vp, _vp, _c_i64, _c_i64, _c_int, _c_dbl, _vp, _vp, _vp]),
    "pib_block_sums_noalloc": (_c_int, [_vp, _vp, _c_i64, _c_i64, _c_int, _c_dbl, _vp, _vp, _vp]),
    "pib_gemm_f32": (_c_int, [_vp, _c_int, _c_int, _c_i64, _c_i64, _c_i64, _vp, _c_i64, _vp,
                              _c_i64, _vp, _c_i64, _vp, _c_int, _vp, _c_i64, _c_int, _vp]),
    "pib_gemm_bf16x3": (_c_int, [_vp, _c_int, _c_int, _c_i64, _c_i64, _c_i64, _vp, _c_i64, _vp,
                                 _c_i64, _vp, _c_i64, _vp, _c_int, _vp, _c_i64, _c_int, _vp]),
    "pib_gemm_bf16x3_presplit": (_c_int, [_vp, _c_i64, _c_i64, _c_i64, _vp, _vp, _c_i64, _vp, _vp,
                                          _c_i64, _vp, _c_i64, _vp, _vp, _c_i64, _vp, _c_int, _vp]),
    "pib_flow_set_base": (_c_int, [_vp, _c_int, _c_int, _vp, _vp]),
    "pib_flow_draw_base": (_c_int, [_vp, _c_u64, _c_i64, _c_i64, _vp, _vp, _vp]),
    "pib_split_bf16": (_c_int, [_vp, _c_i64, _c_i64, _vp, _c_i64, _c_int, _vp, _vp, _c_i64, _vp]),
    "pib_flow_layer_fused": (_c_int, [_vp, _c_int, _c_i64, _c_int, _c_int, _c_int, _c_int, ctypes.c_float,
                                      _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pib_flow_layer_fused_frag": (_c_int, [_vp, _c_int, _c_i64, _c_int, _c_int, _c_int, _c_int,
                                           ctypes.c_float, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pib_flow_set_basis": (_c_int, [_vp, _c_int, _c_int, _dp, _dp, _dp]),
    "pib_flow_whiten": (_c_int, [_vp, _c_int, _c_i64, _vp, _vp, _vp, _vp]),
    "pib_flow_coupling": (_c_int, [_vp, _c_int, _c_i64, _c_int, _c_int, _c_int, ctypes.c_float,
                                   _vp, _vp, _vp, _vp]),
    "pib_flow_coupling_inverse_backward": (_c_int, [_vp, _c_i64, _c_int, _c_int, _c_int,
                                                    ctypes.c_float, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pib_flow_logrho": (_c_int, [_vp, _c_i64, _vp, _vp, _vp, _vp, _vp]),
    "pib_colsum_f32": (_c_int, [_vp, _c_i64, _c_i64, _vp, _c_i64, _vp, _vp]),
    "pib_adam_f32": (_c_int, [_vp, _c_i64, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float,
                              ctypes.c_float, ctypes.c_float, _c_int, _vp]),
    "pib_batch_create": (_c_int, [_c_int, ctypes.POINTER(_vp), _c_i64, _c_int, _c_dbl, _c_int,
                                  ctypes.POINTER(_vp)]),
    "pib_batch_destroy": (_c_int, [_vp]),
    "pib_batch_last_error": (ctypes.c_char_p, [_vp]),
    "pib_batch_bind": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64]),
    "pib_batch_bind_sums": (_c_int, [_vp, _vp]),
    "pib_batch_set_schedule": (_c_int, [_vp, _c_int, ctypes.POINTER(_c_int), _c_int, _c_int]),
    "pib_batch_get_schedule": (_c_int, [_vp, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int),
                                        ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)]),
    "pib_batch_tune": (_c_int, [_vp, _c_u64, _dp, _dp, _vp]),
    "pib_batch_eval": (_c_int, [_vp, _vp]),
    "pib_batch_eval_sequential": (_c_int, [_vp, _vp]),
    "pib_batch_run": (_c_int, [_vp, _c_u64, _c_i64, _c_int, _vp]),
}

FLOW_WHITEN, FLOW_UNWHITEN, FLOW_GRAD, FLOW_UNWHITEN_OFFSETS = 0, 1, 2, 3

_lib = None


class NativeError(RuntimeError):
    """Raised for a non-zero status from the C-ABI (message from pib_last_error)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


def load(path=None):
    """Load the shared library once and attach the ctypes signatures. Raises when absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"{p} not found: build it with `python -m pibronic_amd.build`")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.pib_abi_version() != ABI_VERSION:
        raise ImportError(f"{p}: ABI version {lib.pib_abi_version()} != {ABI_VERSION}")
    if path is None:
        _lib = lib
    return lib


def _dptr(a):
    """Host float64 C-contiguous array -> double*; None -> NULL."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _host(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _devptr(t):
    """torch tensor (device) or integer address -> void* value."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Context:
    """One pib_ctx on one device; thin checked wrappers around every entry point."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        rc = self.lib.pib_ctx_create(int(device), ctypes.byref(h))
        if rc != PIB_OK:
            raise NativeError(rc, f"pib_ctx_create(device={device}) failed")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.lib.pib_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != PIB_OK:
            raise NativeError(rc, self.lib.pib_last_error(self.h).decode())

    # ------------------------------------------------------------------ constants
    def set_coupling(self, energy, linear, quadratic=None):
        E, L, Q = _host(energy), _host(linear), _host(quadratic)
        A, N = E.shape[0], L.shape[0]
        self._check(self.lib.pib_set_coupling(self.h, A, N, _dptr(E), _dptr(L), _dptr(Q)))

    def set_omatrix_set(self, slot, coth, csch, prefactor, state_shift):
        c, s, f, d = _host(coth), _host(csch), _host(prefactor), _host(state_shift)
        self._check(self.lib.pib_set_omatrix_set(self.h, slot, f.shape[0], c.shape[0],
                                                 _dptr(c), _dptr(s), _dptr(f), _dptr(d)))

    def set_beads(self, P, tau):
        self._check(self.lib.pib_set_beads(self.h, int(P), float(tau)))

    def set_lanes_per_sample(self, lanes):
        self._check(self.lib.pib_set_lanes_per_sample(self.h, int(lanes)))

    def lanes_per_sample(self, B):
        r = self.lib.pib_get_lanes_per_sample(self.h, int(B))
        if r < 0:
            self._check(r)
        return r

    def set_sampler(self, P, weights, coth, csch, state_shift):
        w, c, s, d = _host(weights), _host(coth), _host(csch), _host(state_shift)
        A_rho, N = d.shape
        self._check(self.lib.pib_set_sampler(self.h, A_rho, N, int(P), _dptr(w), _dptr(c),
                                             _dptr(s), _dptr(d)))

    # ------------------------------------------------------------------ per block
    def eval_block(self, q, B, flags, out, out_stride, debug=None, stream=None):
        self._check(self.lib.pib_eval_block(self.h, _devptr(q), int(B), int(flags), _devptr(out),
                                            int(out_stride), _devptr(debug), _stream(stream)))

    def eval_block_flow_offsets(self, dq, B, flags, out, out_stride, stream=None):
        """pib_eval_block on float offsets dq from the flow basis centre (R = m + dq)."""
        self._check(self.lib.pib_eval_block_flow_offsets(self.h, _devptr(dq), int(B), int(flags), _devptr(out),
                                                         int(out_stride), _stream(stream)))

    def sample_block(self, seed, offset, B, q, src=None, stream=None):
        self._check(self.lib.pib_sample_block(self.h, int(seed) & (2**64 - 1), int(offset), int(B),
                                              _devptr(q), _devptr(src), _stream(stream)))

    def block_sums(self, out, out_stride, B, pm, delta_beta, sums, stream=None):
        self._check(self.lib.pib_block_sums(self.h, _devptr(out), int(out_stride), int(B),
                                            int(bool(pm)), float(delta_beta), _devptr(sums),
                                            _stream(stream)))

    def block_sums_q(self, out, out_stride, B, pm, delta_beta, logq, sums, stream=None):
        self._check(self.lib.pib_block_sums_q(self.h, _devptr(out), int(out_stride), int(B),
                                              int(bool(pm)), float(delta_beta), _devptr(logq),
                                              _devptr(sums), _stream(stream)))

    def block_sums_noalloc(self, out, out_stride, B, pm, delta_beta, sums, logq=None, stream=None):
        """block_sums (logq None) or block_sums_q without device allocation: partial rows in the
        library's static per-GPU array; calls overlapping on one GPU race."""
        self._check(self.lib.pib_block_sums_noalloc(self.h, _devptr(out), int(out_stride), int(B),
                                                    int(bool(pm)), float(delta_beta), _devptr(logq),
                                                    _devptr(sums), _stream(stream)))

    # ------------------------------------------------------------------ learned proposal
    def gemm_f32(self, A, B, C, M, N, K, lda, ldb, ldc, trans_a=False, trans_b=False, bias=None,
                 relu=False, mask=None, ldmask=0, accumulate=False, stream=None):
        self._check(self.lib.pib_gemm_f32(self.h, int(trans_a), int(trans_b), int(M), int(N), int(K),
                                          _devptr(A), int(lda), _devptr(B), int(ldb), _devptr(C),
                                          int(ldc), _devptr(bias), int(bool(relu)), _devptr(mask),
                                          int(ldmask), int(bool(accumulate)), _stream(stream)))

    def gemm_bf16x3(self, A, B, C, M, N, K, lda, ldb, ldc, trans_a=False, trans_b=False, bias=None,
                    relu=False, mask=None, ldmask=0, accumulate=False, stream=None):
        """gemm_f32 on the bf16 matrix cores with a BF16x3 operand split (K % 32 == 0)."""
        self._check(self.lib.pib_gemm_bf16x3(self.h, int(trans_a), int(trans_b), int(M), int(N), int(K),
                                             _devptr(A), int(lda), _devptr(B), int(ldb), _devptr(C),
                                             int(ldc), _devptr(bias), int(bool(relu)), _devptr(mask),
                                             int(ldmask), int(bool(accumulate)), _stream(stream)))

    def gemm_bf16x3_presplit(self, M, N, K, A_hi, A_lo, lda, Bt_hi, Bt_lo, ldb, C=None, ldc=0,
                             O_hi=None, O_lo=None, ldo=0, bias=None, relu=False, stream=None):
        self._check(self.lib.pib_gemm_bf16x3_presplit(
            self.h, int(M), int(N), int(K), _devptr(A_hi), _devptr(A_lo), int(lda), _devptr(Bt_hi),
            _devptr(Bt_lo), int(ldb), _devptr(C), int(ldc), _devptr(O_hi), _devptr(O_lo), int(ldo),
            _devptr(bias), int(bool(relu)), _stream(stream)))

    def split_bf16(self, rows, cols, X, ldx, hi, lo, ldo, transpose=False, stream=None, perm32=False,
                   frag32=False):
        """lo may be None (only hi written); frag32: the transposed matrix in the fragment order of
        pib_flow_layer_fused_frag (with transpose=True)."""
        mode = (SPLIT_TRANSPOSE if transpose else 0) | (SPLIT_PERM32 if perm32 else 0) | \
            (SPLIT_FRAG32 if frag32 else 0)
        self._check(self.lib.pib_split_bf16(self.h, int(rows), int(cols), _devptr(X), int(ldx),
                                            mode, _devptr(hi), _devptr(lo), int(ldo), _stream(stream)))

    def flow_layer_fused(self, inverse, B, D, hidden, a_off, b_off, smax, z, logdet, w, b1, b2, b3,
                         stream=None):
        """w = ((W1h, W1l), (W2h, W2l), (W3h, W3l)) pre-split transposed weights, W2 and W3
        split with perm32=True. All three lo parts None: single bf16 products of the hi parts."""
        (w1h, w1l), (w2h, w2l), (w3h, w3l) = w
        self._check(self.lib.pib_flow_layer_fused(
            self.h, int(bool(inverse)), int(B), int(D), int(hidden), int(a_off), int(b_off), float(smax),
            _devptr(z), _devptr(logdet), _devptr(w1h), _devptr(w1l), _devptr(w2h), _devptr(w2l),
            _devptr(w3h), _devptr(w3l), _devptr(b1), _devptr(b2), _devptr(b3), _stream(stream)))

    def flow_layer_fused_frag(self, inverse, B, D, hidden, a_off, b_off, smax, z, logdet, w, b1, b2, b3,
                              stream=None):
        """Single-bf16 layer on w = (W1, W2, W3) split with transpose=True, frag32=True (hi
        parts)."""
        w1, w2, w3 = w
        self._check(self.lib.pib_flow_layer_fused_frag(
            self.h, int(bool(inverse)), int(B), int(D), int(hidden), int(a_off), int(b_off), float(smax),
            _devptr(z), _devptr(logdet), _devptr(w1), _devptr(w2), _devptr(w3), _devptr(b1), _devptr(b2),
            _devptr(b3), _stream(stream)))

    def flow_set_basis(self, U, sigma, m):
        Uh, sh, mh = _host(U), _host(sigma), _host(m)
        N, P = sh.shape
        self._check(self.lib.pib_flow_set_basis(self.h, N, P, _dptr(Uh), _dptr(sh), _dptr(mh)))

    def flow_whiten(self, mode, B, inp, work, out, stream=None):
        self._check(self.lib.pib_flow_whiten(self.h, int(mode), int(B), _devptr(inp), _devptr(work),
                                             _devptr(out), _stream(stream)))

    def flow_coupling(self, inverse, B, D, a_off, b_off, smax, z, o, logdet, stream=None):
        self._check(self.lib.pib_flow_coupling(self.h, int(bool(inverse)), int(B), int(D), int(a_off),
                                               int(b_off), float(smax), _devptr(z), _devptr(o),
                                               _devptr(logdet), _stream(stream)))

    def flow_coupling_inverse_backward(self, B, D, a_off, b_off, smax, y, o, dz, dout, c, stream=None):
        self._check(self.lib.pib_flow_coupling_inverse_backward(
            self.h, int(B), int(D), int(a_off), int(b_off), float(smax), _devptr(y), _devptr(o),
            _devptr(dz), _devptr(dout), _devptr(c), _stream(stream)))

    def flow_set_base(self, mu, c):
        """mu (A_rho, D) and c (A_rho,) of the flow's base distribution (float64 host arrays)."""
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        self._check(self.lib.pib_flow_set_base(self.h, int(mu.shape[0]), int(mu.shape[1]),
                                               mu.ctypes.data_as(ctypes.c_void_p),
                                               c.ctypes.data_as(ctypes.c_void_p)))

    def flow_draw_base(self, seed, offset, B, z, logrho, stream=None):
        self._check(self.lib.pib_flow_draw_base(self.h, int(seed) & (2**64 - 1), int(offset), int(B),
                                                _devptr(z), _devptr(logrho), _stream(stream)))

    def flow_logrho(self, B, R, c=None, dR=None, logrho=None, stream=None):
        self._check(self.lib.pib_flow_logrho(self.h, int(B), _devptr(R), _devptr(c), _devptr(dR),
                                             _devptr(logrho), _stream(stream)))

    def colsum_f32(self, M, N, X, ldx, out, stream=None):
        self._check(self.lib.pib_colsum_f32(self.h, int(M), int(N), _devptr(X), int(ldx), _devptr(out),
                                            _stream(stream)))

    def adam_f32(self, n, p, g, m, v, lr, beta1, beta2, eps, step, stream=None):
        self._check(self.lib.pib_adam_f32(self.h, int(n), _devptr(p), _devptr(g), _devptr(m), _devptr(v),
                                          float(lr), float(beta1), float(beta2), float(eps), int(step),
                                          _stream(stream)))


class Batch:
    """A pib_batch over Contexts (one block of B samples per problem, one persistent launch).

    The Contexts must stay alive, with unchanged constants, while the batch exists (the batch
    keeps references to them); after a setter call on one of them the batch refuses to launch
    (NativeError PIB_ERR_STATE) and must be created again."""

    def __init__(self, contexts, B, pm=False, noscale=False, delta_beta=1.0, max_blocks=1, fp32=False):
        self.lib = load()
        self.contexts = list(contexts)
        self.n, self.B, self.pm, self.max_blocks = len(self.contexts), int(B), bool(pm), int(max_blocks)
        arr = (_vp * self.n)(*[c.h.value for c in self.contexts])
        h = ctypes.c_void_p()
        flags = (EVAL_PM if pm else 0) | (EVAL_NOSCALE if noscale else 0) | (EVAL_FP32 if fp32 else 0)
        rc = self.lib.pib_batch_create(self.n, arr, self.B, flags, float(delta_beta), self.max_blocks,
                                       ctypes.byref(h))
        if rc != PIB_OK:
            msg = self.lib.pib_batch_last_error(h).decode() if h.value else "bad arguments"
            if h.value:
                self.lib.pib_batch_destroy(h)
            raise NativeError(rc, msg)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.pib_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != PIB_OK:
            raise NativeError(rc, self.lib.pib_batch_last_error(self.h).decode())

    def bind(self, i, q, out, out_stride):
        """Problem i reads q (B, N, P) and writes out (NUM_OUT rows of out_stride)."""
        self._check(self.lib.pib_batch_bind(self.h, int(i), _devptr(q), _devptr(out), int(out_stride)))

    def bind_sums(self, sums):
        """sums: (max_blocks, n, NUM_SUMS) float64 device tensor."""
        self._check(self.lib.pib_batch_bind_sums(self.h, _devptr(sums)))

    def set_schedule(self, family, lanes=None, tile_waves=4, grid_waves=2048):
        arr = None if lanes is None else (_c_int * self.n)(*[int(g) for g in lanes])
        self._check(self.lib.pib_batch_set_schedule(self.h, int(family), arr, int(tile_waves), int(grid_waves)))

    def schedule(self):
        fam, tile, grid = _c_int(), _c_int(), _c_int()
        lanes = (_c_int * self.n)()
        self._check(self.lib.pib_batch_get_schedule(self.h, ctypes.byref(fam), lanes, ctypes.byref(tile),
                                                    ctypes.byref(grid)))
        return {"family": {BATCH_LANES: "lanes", BATCH_WIDE: "wide", BATCH_SEQUENTIAL: "sequential"}[fam.value],
                "lanes": list(lanes),
                "tile": tile.value, "grid_waves": grid.value}

    def tune(self, seed=1, stream=None):
        best, seq = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.pib_batch_tune(self.h, ctypes.c_uint64(seed), ctypes.byref(best),
                                            ctypes.byref(seq), _stream(stream)))
        return best.value, seq.value

    def eval(self, stream=None):
        self._check(self.lib.pib_batch_eval(self.h, _stream(stream)))

    def eval_sequential(self, stream=None):
        self._check(self.lib.pib_batch_eval_sequential(self.h, _stream(stream)))

    def run(self, seed, first_block, nblocks, stream=None):
        self._check(self.lib.pib_batch_run(self.h, ctypes.c_uint64(seed), int(first_block), int(nblocks),
                                           _stream(stream)))
