"""Pibronic's block/run API for the PIMC estimator, computing on MI355X through
libpibronic_hip.so.

The classes keep the reference's names, attributes and call pattern
(pibronic/pimc/pimc.py:36-47): build a BoxData (or BoxDataPM), call preprocess(), make a
BoxResult (or BoxResultPM) and run block_compute (or block_compute_pm). Model loading,
folding and the temperature constants are host NumPy, the same expressions as the
reference, computed once per run. Everything per block -- bead sampling, O-matrices,
scaling, rho, V, its eigen-decomposition, M, the ring products and the block sums -- runs
in the HIP kernels; there is no CPU evaluation path. Without a GPU or without the built
library, block_compute raises.

Differences from the reference, all deliberate:
  * samples are drawn per block on the device (Philox, counter = global sample index) or,
    with data.sampler = "host", by a seeded numpy Generator on the host; the reference's
    legacy global MT19937 stream and its up-front draw of all surface sources (quirk Q5)
    are not reproduced (the distribution is the same). Like the reference, which reseeds
    numpy from OS entropy in every job process (pibronic/pimc/pimc.py:55), a run whose
    data.seed is None takes a fresh 64-bit seed from OS entropy; a job with result.id_job = J
    starts its counters at sample J * X, so jobs sharing one explicit seed draw disjoint
    slices of one stream. The seed and the first sample index are written to the result
    file as the extra keys "seed" and "sample_offset";
  * per-sample standard deviations are not materialised (the reference stores an (X,N,P)
    array); they are looked up per surface on the device;
  * every block also leaves FP64 block sums (result.block_sums) for the statistics, and
    the per-sample log of the removed scale (result.log_scale).
"""
import json
import os
import secrets
from functools import partial

import numpy as np
from numpy import float64 as F64
from numpy import newaxis as NEW
import scipy.linalg

from . import constants
from .constants import hbar
from . import file_structure
from . import vibronic_model_io as vIO
from .vibronic_model_keys import VibronicModelKeys as VMK
from . import _native

__all__ = ["ModelClass", "ModelVibronic", "ModelVibronicPM", "ModelSampling", "BoxData",
           "BoxDataPM", "BoxResult", "BoxResultPM", "block_compute", "block_compute_pm",
           "evaluate_samples", "TemperatureDependentClass"]


class ServerExecutionParameters:
    """Keys of the BoxData JSON parameter string (pibronic/server/server.py:11-23)."""
    X = "number_of_samples"
    nBlk = "number_of_blocks"
    A = "number_of_states"
    P = "number_of_beads"
    N = "number_of_modes"
    T = "temperature"
    BlkS = "block_size"
    dB = "delta_beta"
    D = "id_data"
    R = "id_rho"
    beta = "beta"
    tau = "tau"


SEP = ServerExecutionParameters


class TemperatureDependentClass:
    """coth, csch and the O-matrix prefactor of one model at one tau.

    pibronic/pimc/pimc.py:59-89. The reference also allocates per-block O-matrix caches
    here; the fused kernel never materialises them, so only the constants are kept:
    cothAN, cschAN (A, N) and omatrix_prefactor (A,) (the reference broadcasts the last one
    to (B, P, A)).
    """

    def __init__(self, model, tau):
        A, N = model.states, model.modes
        omega = np.broadcast_to(model.omega, (A, N))
        self.tau = tau
        self.cothAN = np.tanh(hbar * tau * omega) ** (-1.)
        self.cschAN = np.sinh(hbar * tau * omega) ** (-1.)
        energy = np.diag(model.energy) if len(model.energy.shape) > 1 else model.energy
        tilde_energy = energy + model.delta_weight
        pref = np.exp(-tau * tilde_energy.copy())
        pref *= np.prod(self.cschAN.copy(), axis=1) ** 0.5
        self.omatrix_prefactor = pref


class ModelClass:
    """A vibronic model: A surfaces, N modes (pibronic/pimc/pimc.py:92-149)."""

    def __init__(self, states=1, modes=1):
        self.states = states
        self.modes = modes
        self.omega = self.energy = self.linear = self.quadratic = None


class ModelVibronic(ModelClass):
    """The coupled model (pibronic/pimc/pimc.py:152-235)."""

    def __init__(self, data):
        super().__init__(data.states, data.modes)
        self.size = data.size
        self.beta = data.beta
        self.tau = data.tau
        A, N = self.states, self.modes
        self.omega = np.zeros(N, dtype=F64)
        self.energy = np.zeros((A, A), dtype=F64)
        self.linear = np.zeros((N, A, A), dtype=F64)
        self.quadratic = np.zeros((N, N, A, A), dtype=F64)
        self.delta_weight = np.zeros(A, dtype=F64)
        self.state_shift = np.zeros((A, N), dtype=F64)

    def load_model(self, path):
        kw = vIO.load_model_from_JSON(path)
        A, N = vIO._extract_dimensions_from_dictionary(kw)
        assert (A, N) == (self.states, self.modes), \
            f"model file has A={A}, N={N}; BoxData says A={self.states}, N={self.modes}"
        shape = vIO.model_shape_dict(A, N)
        self.energy = kw[VMK.E]
        self.omega = kw[VMK.w]
        self.linear = kw[VMK.G1] if VMK.G1 in kw else np.zeros(shape[VMK.G1], dtype=F64)
        self.quadratic = kw[VMK.G2] if VMK.G2 in kw else np.zeros(shape[VMK.G2], dtype=F64)

    def compute_linear_displacement(self, data):
        """delta_a = -1/2 sum_j (L^j_aa)^2 / w_j (pimc.py:172-177)."""
        for a in range(self.states):
            self.delta_weight[a] = -0.5 * (self.linear[:, a, a] ** 2. / self.omega).sum(axis=0)

    def initialize_TDP_object(self):
        self.const = TemperatureDependentClass(self, self.tau)

    def finish_folding_in_terms(self, data):
        """d_aj = -L^j_aa / w_j, then zero L^j_aa and E_aa (pimc.py:208-218)."""
        for a in range(self.states):
            self.state_shift[a, :] = -self.linear[:, a, a] / self.omega
            self.linear[:, a, a] = 0.0
            self.energy[a, a] = 0.0

    def precompute(self, data):
        """pimc.py:220-235 (the prefactor is built before the diagonal energy is zeroed)."""
        self.compute_linear_displacement(data)
        self.initialize_TDP_object()
        self.finish_folding_in_terms(data)


class ModelVibronicPM(ModelVibronic):
    """Coupled model with the beta +- delta_beta constants (pibronic/pimc/pimc.py:238-269)."""

    def __init__(self, data):
        super().__init__(data)
        self.delta_beta = data.delta_beta

    def initialize_TDP_object(self):
        super().initialize_TDP_object()
        self.const_plus = TemperatureDependentClass(self, self.tau_plus)
        self.const_minus = TemperatureDependentClass(self, self.tau_minus)

    def precompute(self, data):
        self.beta_plus = self.beta + self.delta_beta
        self.beta_minus = self.beta - self.delta_beta
        self.tau_plus = self.beta_plus / data.beads
        self.tau_minus = self.beta_minus / data.beads
        super().precompute(data)


class ModelSampling(ModelClass):
    """The harmonic sampling distribution rho (pibronic/pimc/pimc.py:272-424)."""

    def __init__(self, data):
        super().__init__()
        self.tau = data.tau
        self.beta = data.beta
        self.beads = data.beads

    def load_model(self, path):
        A, N = vIO.extract_dimensions_of_diagonal_model(path=path)
        self.states, self.modes = A, N
        kw = vIO.load_diagonal_model_from_JSON(path)
        shape = vIO.diagonal_model_shape_dict(A, N)
        self.energy = kw[VMK.E]
        self.omega = kw[VMK.w]
        self.linear = kw[VMK.G1] if VMK.G1 in kw else np.zeros(shape[VMK.G1], dtype=F64)
        self.quadratic = kw[VMK.G2] if VMK.G2 in kw else np.zeros(shape[VMK.G2], dtype=F64)
        self.delta_weight = np.zeros(A, dtype=F64)
        self.state_weight = np.zeros(A, dtype=F64)
        self.state_shift = np.zeros((A, N), dtype=F64)

    def compute_linear_displacement(self):
        """pimc.py:336-339"""
        self.delta_weight = -0.5 * (self.linear ** 2. / self.omega[:, NEW]).sum(axis=0)

    def compute_weight_for_each_state(self):
        """w_a ~ exp(-beta E~_a) / prod_j sinh(beta w_j / 2), normalised (pimc.py:341-351)."""
        self.state_weight = np.exp(-self.beta * (self.energy + self.delta_weight))
        self.state_weight /= np.prod(np.sinh((self.beta * self.omega) / 2.))
        self.state_weight /= self.state_weight.sum()

    def initialize_TDP_object(self):
        self.const = TemperatureDependentClass(self, self.tau)

    def finish_folding_in_terms(self):
        """pimc.py:372-381"""
        for a in range(self.states):
            self.state_shift[a, :] = -self.linear[:, a] / self.omega
            self.linear[:, a] = 0.0
        self.energy = np.full_like(self.energy, np.nan)

    def compute_sampling_constants(self, data):
        """Per-surface standard deviations of the normal-mode draws (pimc.py:398-406):
        1 / sqrt(2 coth_j - csch_j lambda_k), lambda_k the circulant eigenvalues."""
        self.inverse_covariance = (2. * self.const.cothAN[..., NEW]
                                   - self.const.cschAN[..., NEW] * data.circulant_eigvals)
        self.standard_deviation = np.sqrt(1. / self.inverse_covariance)

    def precompute(self, data):
        """pimc.py:409-424 (E~_a = E_a + delta_a is kept: finish_folding_in_terms blanks E)."""
        self.compute_linear_displacement()
        self.tilde_energy = self.energy + self.delta_weight
        self.compute_weight_for_each_state()
        self.initialize_TDP_object()
        self.finish_folding_in_terms()
        self.compute_sampling_constants(data)


class BoxData:
    """Run parameters, models, and the device state of one (model, rho, T, P) run.

    pibronic/pimc/pimc.py:427-692. Attributes as the reference: samples, blocks,
    block_size, states, modes, temperature, beads, beta, tau, delta_beta, id_data, id_rho,
    path_vib_model, path_rho_model, hash_vib, hash_rho. Extra: sampler ("device" or
    "host"), seed, device (CUDA ordinal).
    """

    pm = False
    _COMMA_REPLACEMENT = ";"
    _SEPARATORS = (_COMMA_REPLACEMENT, ":")

    def __init__(self):
        self.samples = self.blocks = self.block_size = 0
        self.states = self.modes = 0
        self.temperature = 0.0
        self.beads = 0
        self.beta = self.tau = 0.0
        if not hasattr(self, "delta_beta"):
            self.delta_beta = 0.0
        self.id_data = self.id_rho = 0
        self.path_vib_model = self.path_rho_model = ""
        self.hash_vib = self.hash_rho = None
        self.vib = self.rho = None
        self.sampler = "device"
        # True: run the one-wave-per-sample kernel (16x16 padded matrices) whatever A is;
        # it is chosen automatically for A > 8
        self.force_wide = False
        # True: the reduced-precision kernel (FP32 products, PIB_EVAL_FP32; A <= 8)
        self.fp32 = False
        # True: the wide kernel with BF16x3 ring products (PIB_EVAL_BF16X3; A <= 16)
        self.ring_bf16x3 = False
        self.seed = None
        # global index of the run's first sample in the (seed) Philox stream; block_compute
        # sets it to id_job * samples
        self.sample_offset = 0
        self.device = 0
        self._ctx = None
        self._dev = None

    # ------------------------------------------------------------------ constructors
    @classmethod
    def from_FileStructure(cls, FS):
        data = cls()
        data.id_data, data.id_rho = FS.id_data, FS.id_rho
        data.path_vib_model, data.path_rho_model = FS.path_vib_model, FS.path_rho_model
        data.states, data.modes = vIO.extract_dimensions_of_model(FS)
        return data

    @classmethod
    def build(cls, id_data, id_rho):
        data = cls()
        data.id_data, data.id_rho = id_data, id_rho
        return data

    @classmethod
    def from_json_file(cls, path):
        with open(path, mode="r", encoding="UTF8") as fh:
            return cls.from_json_string(fh.read())

    @classmethod
    def from_json_string(cls, json_str):
        data = cls()
        data.load_json_string(json_str)
        return data

    @classmethod
    def json_serialize(cls, params):
        return json.dumps(params, separators=cls._SEPARATORS)

    def encode_self(self, params=None):
        """JSON parameter string of this run (pimc.py:517-541)."""
        if params is None:
            params = {SEP.X: self.samples, SEP.nBlk: self.blocks, SEP.A: self.states,
                      SEP.P: self.beads, SEP.N: self.modes, SEP.T: self.temperature,
                      SEP.BlkS: self.block_size, SEP.dB: self.delta_beta, SEP.D: self.id_data,
                      SEP.R: self.id_rho, SEP.beta: self.beta, SEP.tau: self.tau}
        return self.json_serialize(params)

    def load_json_string(self, json_str):
        """Parse the job parameter string (';' stands for ',') (pimc.py:543-593).

        Needs "path_root" to locate the model files through FileStructure.
        """
        params = json.loads(json_str.replace(self._COMMA_REPLACEMENT, ","))
        self.samples = params[SEP.X]
        self.blocks = params[SEP.nBlk]
        self.states = params[SEP.A]
        self.beads = params[SEP.P]
        self.modes = params[SEP.N]
        self.temperature = params[SEP.T]
        self.block_size = params[SEP.BlkS]
        self.id_data = params[SEP.D]
        self.id_rho = params[SEP.R]
        self.beta = params[SEP.beta] if SEP.beta in params else \
            1.0 / (constants.boltzman * self.temperature)
        self.tau = params[SEP.tau] if SEP.tau in params else self.beta / self.beads
        if SEP.dB in params and type(params[SEP.dB]) is float:
            self.delta_beta = params[SEP.dB]
        FS = file_structure.FileStructure(params["path_root"], self.id_data, self.id_rho)
        self.path_vib_model = FS.path_vib_model
        self.path_rho_model = FS.path_rho_model
        FS.generate_model_hashes()
        self.hash_vib, self.hash_rho = FS.hash_vib, FS.hash_rho

    # ------------------------------------------------------------------ setup
    def _make_vib(self):
        return ModelVibronic(self)

    def initialize_models(self):
        """pimc.py:633-645"""
        self.vib = self._make_vib()
        self.vib.load_model(self.path_vib_model)
        self.vib.precompute(self)
        self.rho = ModelSampling(self)
        self.rho.load_model(self.path_rho_model)
        assert self.rho.modes == self.modes, "sampling model has a different number of modes"
        self.rho.precompute(self)

    def preprocess(self):
        """Sizes, beta, tau, the ring's circulant basis and the folded models (pimc.py:647-692).

        Host only; the device context is created on first use.
        """
        self.param_dict = {"X": self.samples, "A": self.states, "N": self.modes,
                           "P": self.beads, "B": self.block_size}
        self.size_list = ["X", "P", "N", "A", "B", "XP", "BP", "AN", "NA", "AA", "BNP", "BPA",
                          "BAA", "XNP", "NAA", "NNA", "ANP", "BANP", "BPAA", "NNAA", "BPAN"]
        self.size = {k: tuple(self.param_dict[c] for c in k) for k in self.size_list}
        self.beta = constants.beta(self.temperature)
        self.tau = self.beta / self.beads
        assert self.beads >= 3, "circulant matrix requires 3 or more beads"
        defining_vector = [0, 1] + [0] * (self.beads - 3) + [1]
        self.circulant_matrix = scipy.linalg.circulant(defining_vector)
        self.circulant_eigvals, self.circulant_eigvects = np.linalg.eigh(self.circulant_matrix,
                                                                         UPLO="L")
        self.initialize_models()
        self._ctx = None
        self._dev = None

    # ------------------------------------------------------------------ device
    def context(self):
        """Create (once) the device context and upload the run's constants."""
        if self._ctx is not None:
            return self._ctx
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("pibronic_amd needs a ROCm GPU: torch.cuda.is_available() is False")
        ctx = _native.Context(self.device)
        vib, rho = self.vib, self.rho
        Q = vib.quadratic if np.count_nonzero(vib.quadratic) else None
        ctx.set_coupling(vib.energy, vib.linear, Q)
        ctx.set_omatrix_set(_native.SET_RHO, rho.const.cothAN[0], rho.const.cschAN[0],
                            rho.const.omatrix_prefactor, rho.state_shift)
        ctx.set_omatrix_set(_native.SET_VIB, vib.const.cothAN[0], vib.const.cschAN[0],
                            vib.const.omatrix_prefactor, vib.state_shift)
        if self.pm:
            for slot, c in ((_native.SET_VIB_PLUS, vib.const_plus),
                            (_native.SET_VIB_MINUS, vib.const_minus)):
                ctx.set_omatrix_set(slot, c.cothAN[0], c.cschAN[0], c.omatrix_prefactor,
                                    vib.state_shift)
        ctx.set_beads(self.beads, self.tau)
        ctx.set_sampler(self.beads, rho.state_weight, rho.const.cothAN[0], rho.const.cschAN[0],
                        rho.state_shift)
        self._ctx = ctx
        return ctx

    def device_buffers(self, B):
        """q (B, N, P), out (NUM_OUT, B) and sums (NUM_SUMS,) as torch tensors on the GPU."""
        import torch
        if self._dev is None or self._dev["B"] < B:
            dev = torch.device("cuda", self.device)
            self._dev = dict(
                B=B,
                q=torch.empty((B, self.modes, self.beads), dtype=torch.float64, device=dev),
                src=torch.empty((B,), dtype=torch.int32, device=dev),
                out=torch.empty((_native.NUM_OUT, B), dtype=torch.float64, device=dev),
                sums=torch.zeros((_native.NUM_SUMS,), dtype=torch.float64, device=dev))
        return self._dev

    def draw_sample(self, sample_view):
        """Fill the device coordinate buffer with the block's bead coordinates R (B, N, P).

        pibronic/pimc/pimc.py:595-600 + 613-631. Device sampler by default; with
        sampler == "host", a numpy Generator seeded from (seed, block start) draws them.
        """
        import torch
        start, B = sample_view.start, sample_view.stop - sample_view.start
        ctx, dev = self.context(), self.device_buffers(B)
        seed = self.resolve_seed()
        first = int(self.sample_offset) + start
        if self.sampler == "device":
            ctx.sample_block(seed, first, B, dev["q"], dev["src"],
                             torch.cuda.current_stream(self.device))
        elif self.sampler == "host":
            q, src = self.draw_host_sample(np.random.default_rng([seed, first]), B)
            dev["q"][:B].copy_(torch.from_numpy(q))
            dev["src"][:B].copy_(torch.from_numpy(src.astype(np.int32)))
        else:
            raise ValueError(f"unknown sampler {self.sampler!r}")

    def resolve_seed(self):
        """The run's 64-bit sampler seed: data.seed, or, when it is None, a fresh one from OS
        entropy (kept in data.seed, so every block of the run and the result file use it)."""
        if self.seed is None:
            self.seed = secrets.randbits(64)
        return int(self.seed) & 0xFFFFFFFFFFFFFFFF

    def draw_host_sample(self, rng, B):
        """Host draw of R (B, N, P) and surface sources, the reference's algorithm
        (pimc.py:386-389, 329-333, 618-624) with a numpy Generator."""
        rho = self.rho
        src = rng.choice(rho.states, size=B, p=rho.state_weight)
        cc = rng.normal(0.0, rho.standard_deviation[src, ...], size=(B, self.modes, self.beads))
        R = np.einsum("ab,ijb->ija", self.circulant_eigvects, cc)
        R += rho.state_shift[src, :, NEW]
        return np.ascontiguousarray(R), src

    def evaluate_device(self, q, B, pm, out, out_stride, debug=None, noscale=False, stream=None):
        """One pib_eval_block call on device buffers (torch tensors)."""
        import torch
        flags = ((_native.EVAL_PM if pm else 0) | (_native.EVAL_NOSCALE if noscale else 0)
                 | (_native.EVAL_WIDE if self.force_wide else 0) | (_native.EVAL_FP32 if self.fp32 else 0)
                 | (_native.EVAL_BF16X3 if self.ring_bf16x3 else 0))
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        self.context().eval_block(q, B, flags, out, out_stride, debug, stream)

    def evaluate_offsets_device(self, dq, B, out, out_stride, noscale=False, stream=None):
        """pib_eval_block_flow_offsets: the FP64 evaluation (no +-dbeta) of coordinates given as
        float offsets from the flow basis centre (Flow.draw(offsets=True))."""
        import torch
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        self.context().eval_block_flow_offsets(dq, B, _native.EVAL_NOSCALE if noscale else 0, out, out_stride,
                                               stream)


class BoxDataPM(BoxData):
    """BoxData with the beta +- delta_beta path (pibronic/pimc/pimc.py:695-738)."""

    pm = True

    def __init__(self, delta_beta=None):
        self.delta_beta = constants.delta_beta if delta_beta is None else delta_beta
        super().__init__()

    def _make_vib(self):
        return ModelVibronicPM(self)

    def preprocess(self):
        self.beta_plus = self.beta + self.delta_beta
        self.beta_minus = self.beta - self.delta_beta
        self.tau_plus = self.beta_plus / self.beads
        self.tau_minus = self.beta_minus / self.beads
        super().preprocess()


class BoxResult:
    """Per-sample scaled g and rho of a run and their npz file (pimc.py:741-912)."""

    id_job = None
    hash_vib = None
    hash_rho = None
    key_list = ["number_of_samples", "s_rho", "s_g"]

    @classmethod
    def read_number_of_samples(cls, path_full):
        with np.load(path_full, mmap_mode="r") as d:
            return d["number_of_samples"]

    @classmethod
    def verify_result_keys_are_present(cls, path, fileObj):
        for k in cls.key_list:
            if k not in fileObj.keys():
                raise AssertionError(f"Expected key ({k}) not present in result file\n{path}\n")

    @classmethod
    def result_keys_are_present_in(cls, iterable):
        return all(k in iterable for k in cls.key_list)

    def initialize_arrays(self):
        self.scaled_g = np.full(self.samples, np.nan, dtype=F64)
        self.scaled_rho = np.full(self.samples, np.nan, dtype=F64)
        self.log_scale = np.full(self.samples, np.nan, dtype=F64)

    def __init__(self, data=None, X=None):
        self.path_root = None
        if data is not None:
            self.partial_name = partial(file_structure.pimc().format, P=data.beads,
                                        T=data.temperature)
            self.samples = data.samples
            self.hash_vib, self.hash_rho = data.hash_vib, data.hash_rho
        elif X is not None:
            self.samples = X
        else:
            self.samples = 0
        self.block_sums = None
        # sampler seed and first global sample index of the run (block_compute fills them)
        self.seed = None
        self.sample_offset = 0
        self.initialize_arrays()

    def compute_path_to_file(self):
        if self.samples == 0:
            raise AssertionError(f"{type(self).__name__} still has 0 samples - this should not happen")
        J = 0 if self.id_job is None else int(self.id_job)
        self.name = self.partial_name(J=J)
        return os.path.join(self.path_root, self.name)

    def _payload(self):
        return dict(s_rho=self.scaled_rho, s_g=self.scaled_g)

    def save_results(self, number_of_samples):
        """np.savez with the reference's keys (pimc.py:820-834)."""
        path = self.compute_path_to_file()
        assert self.hash_vib is not None and self.hash_rho is not None, \
            "we save hash values if they don't exist!"
        extra = {}
        if self.seed is not None:
            extra = dict(seed=np.uint64(self.seed), sample_offset=np.int64(self.sample_offset))
        np.savez(path, hash_vib=self.hash_vib, hash_rho=self.hash_rho,
                 number_of_samples=self.samples, **self._payload(), **extra)

    def _take(self, d, sl, dst):
        dst_map = {"s_g": "scaled_g", "s_rho": "scaled_rho", "s_gP": "scaled_gofr_plus",
                   "s_gM": "scaled_gofr_minus"}
        for k in self.key_list:
            if k in dst_map:
                getattr(self, dst_map[k])[dst] = d[k][sl]

    def load_results(self, path):
        with np.load(path, mmap_mode="r") as d:
            type(self).verify_result_keys_are_present(path, d)
            if self.samples == 0:
                self.samples = int(d["number_of_samples"])
            elif int(d["number_of_samples"]) != self.samples:
                raise AssertionError("BoxResult has a different number of samples that the input file - this should not happen")
            self.initialize_arrays()
            self._take(d, slice(None), slice(None))

    def load_multiple_results(self, list_of_paths, desired_number_of_samples=None):
        """Concatenate result files, skipping ones without the keys (pimc.py:853-912)."""
        assert len(list_of_paths) != 0, "list_of_paths cannot be empty"
        n, good = 0, []
        for path in list_of_paths:
            with np.load(path, mmap_mode="r") as d:
                if type(self).result_keys_are_present_in(d.keys()):
                    n += int(d["number_of_samples"])
                    good.append(path)
        assert len(good) > 0, f"none of the provided paths were good, i.e. none of them had all the required keys {self.key_list}"
        assert n != 0, "number of samples should have changed"
        want = n if desired_number_of_samples is None else min(n, desired_number_of_samples)
        self.samples = want
        self.initialize_arrays()
        start = 0
        for path in good:
            if start >= want:
                break
            with np.load(path) as d:
                length = min(int(d["number_of_samples"]), want - start)
                assert not np.any(d["s_rho"][0:length] == 0.0), "Zeros in the denominator"
                self._take(d, slice(0, length), slice(start, start + length))
                start += length


class BoxResultPM(BoxResult):
    """BoxResult with s_gP, s_gM (pimc.py:915-1040)."""

    key_list = ["s_gP", "s_gM"] + BoxResult.key_list

    def initialize_arrays(self):
        super().initialize_arrays()
        self.scaled_gofr_plus = np.full(self.samples, np.nan, dtype=F64)
        self.scaled_gofr_minus = np.full(self.samples, np.nan, dtype=F64)

    def _payload(self):
        d = super()._payload()
        d.update(s_gP=self.scaled_gofr_plus, s_gM=self.scaled_gofr_minus)
        return d


# --------------------------------------------------------------------------------------
# block loops
# --------------------------------------------------------------------------------------
def _run_blocks(data, result, pm):
    import torch
    assert data.blocks * data.block_size <= result.samples, "result is smaller than the run"
    B = data.block_size
    ctx = data.context()
    # independent job streams: an unset seed comes from OS entropy, and job J starts at
    # global sample J * X (pibronic/pimc/pimc.py:55; pibronic/server/job_boss.py:538-611)
    result.seed = data.resolve_seed()
    J = 0 if result.id_job is None else int(result.id_job)
    data.sample_offset = J * int(data.samples)
    result.sample_offset = data.sample_offset
    stream = torch.cuda.current_stream(data.device)
    dev = data.device_buffers(B)
    result.block_sums = np.zeros((data.blocks, _native.NUM_SUMS))
    sums_all = torch.zeros((data.blocks, _native.NUM_SUMS), dtype=torch.float64,
                           device=dev["out"].device)
    rows = [_native.OUT_RHO, _native.OUT_G, _native.OUT_LOGS]
    dst = [result.scaled_rho, result.scaled_g, result.log_scale]
    if pm:
        rows += [_native.OUT_GP, _native.OUT_GM]
        dst += [result.scaled_gofr_plus, result.scaled_gofr_minus]
    end = 0
    for block_index in range(data.blocks):
        start = block_index * B
        end = start + B
        data.draw_sample(slice(start, end))
        data.evaluate_device(dev["q"], B, pm, dev["out"], dev["out"].shape[1], stream=stream)
        ctx.block_sums(dev["out"], dev["out"].shape[1], B, pm, data.delta_beta if pm else 1.0,
                       sums_all[block_index], stream)
        host = dev["out"][:, :B].cpu().numpy()
        for r, d in zip(rows, dst):
            d[start:end] = host[r]
    result.block_sums[:] = sums_all.cpu().numpy()
    if result.path_root is not None:
        result.save_results(end)
    return result


def block_compute(data, result):
    """g and rho for every block (pibronic/pimc/pimc.py:1341-1385), on the GPU."""
    return _run_blocks(data, result, pm=False)


def block_compute_pm(data, result):
    """g, g(beta +- dbeta) and rho for every block (pibronic/pimc/pimc.py:1388-1462)."""
    assert isinstance(data, BoxDataPM), "incorrect object type"
    assert isinstance(result, BoxResultPM), "incorrect object type"
    return _run_blocks(data, result, pm=True)


def evaluate_samples(data, q, pm=None, noscale=False, debug=False):
    """Evaluate host-supplied bead coordinates q (B, N, P) on the GPU.

    The host-fed entry point: the same kernels as block_compute on coordinates the caller
    chose (for example the same host-seeded samples given to a CPU implementation). Returns
    numpy arrays s_rho, s_g, log_scale (and s_gP, s_gM); with debug=True also the
    per-(sample, bead) O_rho, O_vib, S and M arrays.
    """
    import torch
    pm = data.pm if pm is None else pm
    q = np.ascontiguousarray(q, dtype=F64)
    B, N, P = q.shape
    assert N == data.modes and P == data.beads, f"q has shape {q.shape}, expected (B, {data.modes}, {data.beads})"
    dev = torch.device("cuda", data.device)
    qd = torch.from_numpy(q).to(dev)
    out = torch.full((_native.NUM_OUT, max(B, 1)), float("nan"), dtype=torch.float64, device=dev)
    A, AR = data.states, data.rho.states
    rec = AR + A + 1 + A * A
    dbg = torch.full((B * P * rec,), float("nan"), dtype=torch.float64, device=dev) if debug else None
    data.evaluate_device(qd, B, pm, out, out.shape[1], debug=dbg, noscale=noscale)
    torch.cuda.synchronize(dev)
    o = out.cpu().numpy()
    res = dict(s_rho=o[_native.OUT_RHO, :B].copy(), s_g=o[_native.OUT_G, :B].copy(),
               log_scale=o[_native.OUT_LOGS, :B].copy())
    if pm:
        res.update(s_gP=o[_native.OUT_GP, :B].copy(), s_gM=o[_native.OUT_GM, :B].copy())
    if debug:
        d = dbg.cpu().numpy().reshape(B, P, rec)
        res.update(O_rho=d[..., :AR], O_vib=d[..., AR:AR + A], S=d[..., AR + A],
                   M=d[..., AR + A + 1:].reshape(B, P, A, A))
    return res


# --------------------------------------------------------------------------------------
# training-data hooks (pibronic/pimc/pimc.py:1216-1338)
# --------------------------------------------------------------------------------------
def _hook_context(data, kind):
    """A second context holding the constants the reference's training-data hooks use.

    Both hooks copy R into qTensor and subtract the coupled model's state shift d once before
    build_o_matrix subtracts a state shift again (pimc.py:1295 + 1092 and :610 + 1092: quirk
    Q6). For rho(R) (kind "rho") the sampling-model O-matrices therefore see R - d - d_rho;
    for g(R) (kind "g") the coupled-model O-matrices see R - 2d, and V is assembled from the
    shifted copies, row b at R - d_b and column c at R - d_c (diagonalize_coupling_matrix,
    :1147-1160). Expanding V in R gives E', L' with the shifts folded in (Q unchanged):
      L'[k,b,c] = L[k,b,c] - 1/2 sum_d Q[d,k,b,c] d_bd - 1/2 sum_e Q[k,e,b,c] d_ce,
      E'[b,c]   = E[b,c] - sum_d L[d,b,c] d_bd + 1/2 sum_de Q[d,e,b,c] d_ce d_bd;
    eigh(UPLO='L') reads the lower triangle, which is mirrored before upload.
    Returns (ctx, asym) with asym = (E'-E'^T, L'-L'^T), the part that makes V non-symmetric.
    """
    cache = data.__dict__.setdefault("_hook_ctx", {})
    if kind in cache:
        return cache[kind]
    vib, rho = data.vib, data.rho
    d = vib.state_shift  # (A, N)
    ctx = _native.Context(data.device)
    Qnz = vib.quadratic if np.count_nonzero(vib.quadratic) else None
    asym = None
    if kind == "g":
        Q, L, E = vib.quadratic, vib.linear, vib.energy
        Lp = (L - 0.5 * np.einsum("dkbc,bd->kbc", Q, d) - 0.5 * np.einsum("kebc,ce->kbc", Q, d))
        Ep = (E - np.einsum("dbc,bd->bc", L, d) + 0.5 * np.einsum("debc,ce,bd->bc", Q, d, d))
        asym = (Ep - Ep.T, Lp - Lp.transpose(0, 2, 1))
        low = np.tril(np.ones(E.shape, dtype=bool))
        Ep_m = np.where(low, Ep, Ep.T)
        Lp_m = np.where(low[NEW], Lp, Lp.transpose(0, 2, 1))
        ctx.set_coupling(Ep_m, Lp_m, Qnz)
        ctx.set_omatrix_set(_native.SET_RHO, rho.const.cothAN[0], rho.const.cschAN[0],
                            rho.const.omatrix_prefactor, rho.state_shift)
        ctx.set_omatrix_set(_native.SET_VIB, vib.const.cothAN[0], vib.const.cschAN[0],
                            vib.const.omatrix_prefactor, 2.0 * d)
    elif kind == "rho":
        if rho.states != data.states:
            raise NotImplementedError(
                "block_compute_rhoR_from_input_samples with a sampling model of another "
                "surface count reads the reference's qTempTensor, which that hook never fills "
                "from the input R (pibronic/pimc/pimc.py:1094-1096, 1293-1299)")
        ctx.set_coupling(vib.energy, vib.linear, Qnz)
        ctx.set_omatrix_set(_native.SET_RHO, rho.const.cothAN[0], rho.const.cschAN[0],
                            rho.const.omatrix_prefactor, rho.state_shift + d)
        ctx.set_omatrix_set(_native.SET_VIB, vib.const.cothAN[0], vib.const.cschAN[0],
                            vib.const.omatrix_prefactor, d)
    else:
        raise ValueError(kind)
    ctx.set_beads(data.beads, data.tau)
    cache[kind] = (ctx, asym)
    return cache[kind]


def _check_hook_symmetry(data, R, asym):
    """The reference asserts np.allclose(V^T, V) before eigh (pimc.py:1169). V^T - V is affine
    in R here; when it is not identically zero, V is formed on the host for the check."""
    dE, dL = asym
    if not (np.any(dE) or np.any(dL)):
        return
    vib = data.vib
    d = vib.state_shift
    for s in range(R.shape[0]):
        q = R[s][NEW] - d[:, :, NEW]  # (A, N, P): q_b = R - d_b
        V = (np.einsum("acef,debc,abdf->fbc", q[NEW], 0.5 * vib.quadratic, q[NEW])
             + np.einsum("dbc,abdf->fbc", vib.linear, q[NEW]) + vib.energy[NEW])
        assert np.allclose(V.transpose(0, 2, 1), V), \
            "V(R) is not symmetric for the shifted training coordinates (pibronic/pimc/pimc.py:1169)"


def _hook_eval(data, kind, R):
    import torch
    ctx, asym = _hook_context(data, kind)
    if kind == "g":
        _check_hook_symmetry(data, R, asym)
    B = R.shape[0]
    dev = torch.device("cuda", data.device)
    q = torch.from_numpy(np.ascontiguousarray(R, dtype=F64)).to(dev)
    out = torch.empty((_native.NUM_OUT, max(B, 1)), dtype=torch.float64, device=dev)
    ctx.eval_block(q, B, _native.EVAL_NOSCALE, out, out.shape[1],
                   None, torch.cuda.current_stream(data.device))
    torch.cuda.synchronize(dev)
    return out[_native.OUT_G if kind == "g" else _native.OUT_RHO, :B].cpu().numpy()


def block_compute_rhoR_from_input_samples(data, result, input_R_values):
    """rho(R) (unscaled) at the caller's bead coordinates R (X, N, P) into result.scaled_rho
    (pibronic/pimc/pimc.py:1273-1299), with the reference's shift of R by the coupled model's
    state shift (quirk Q6)."""
    R = np.asarray(input_R_values, dtype=F64)
    for block_index in range(data.blocks):
        sl = slice(block_index * data.block_size, (block_index + 1) * data.block_size)
        result.scaled_rho[sl] = _hook_eval(data, "rho", R[sl])
    return


def block_compute_gR_from_raw_samples(data, result, rng=None):
    """g(R) (unscaled) at uniform random R in [0, 1) (X, N, P), not drawn from rho, into
    result.scaled_g, then the training files of save_gR_with_samples
    (pibronic/pimc/pimc.py:1302-1338; quirk Q6 as in _hook_context). The reference draws R
    with the global legacy np.random; here `rng` (default: numpy Generator seeded with
    data.seed) makes a run reproducible."""
    rng = np.random.default_rng(data.seed) if rng is None else rng
    X = data.blocks * data.block_size
    input_R_values = np.empty((X, data.modes, data.beads), dtype=F64)
    end = 0
    for block_index in range(data.blocks):
        sl = slice(block_index * data.block_size, (block_index + 1) * data.block_size)
        input_R_values[sl] = rng.random((data.block_size, data.modes, data.beads))
        result.scaled_g[sl] = _hook_eval(data, "g", input_R_values[sl])
        end = sl.stop
    if result.path_root is not None:
        save_gR_with_samples(data, result, input_R_values)
    return input_R_values


def save_gR_with_samples(data, result, input_R_values):
    """The two npz files of a g(R) training run (pibronic/pimc/pimc.py:1248-1270): g with
    number_of_samples, and the input R values."""
    result.partial_name = partial(file_structure.training_data_g_output().format, P=data.beads,
                                  T=data.temperature)
    np.savez(result.compute_path_to_file(), number_of_samples=result.samples, g=result.scaled_g)
    result.partial_name = partial(file_structure.training_data_input().format, P=data.beads,
                                  T=data.temperature)
    np.savez(result.compute_path_to_file(), number_of_samples=result.samples,
             input_R_values=input_R_values)
