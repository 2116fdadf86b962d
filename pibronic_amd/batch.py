"""Many temperatures (and models) of a PIMC run evaluated together on one GPU.

The reference runs one (model, temperature) per process and farms the temperatures out as
separate jobs (pibronic/server/job_boss.py:538-611). Here the BoxData of every temperature
keeps its own context (constants, beads, sampler), and a TemperatureBatch evaluates one block
of every problem in a single persistent launch (libpibronic_hip's pib_batch_*), with the
block step (sample -> evaluate -> reduce) captured once as a hipGraph and replayed.

Per-problem results are the same quantities as block_compute / block_compute_pm: the
per-sample S12-scaled s_rho, s_g (s_gP, s_gM) and log scale of the last block, and the
per-block sums of the statistics layer.
"""
import numpy as np

from . import _native


class TemperatureBatch:
    """One block of `block_size` samples for each BoxData in `datas`, evaluated together.

    datas: preprocessed BoxData (or BoxDataPM, all the same kind) on one device; they may
    differ in temperature, beads and model. max_blocks: how many per-block sum rows pib run()
    may write. fp32: the reduced-precision kernel (PIB_EVAL_FP32)."""

    def __init__(self, datas, block_size=None, max_blocks=1, noscale=False, fp32=False):
        import torch
        self.datas = list(datas)
        if not self.datas:
            raise ValueError("TemperatureBatch needs at least one BoxData")
        self.pm = bool(self.datas[0].pm)
        if any(bool(d.pm) != self.pm for d in self.datas):
            raise ValueError("all problems must be BoxData or all BoxDataPM")
        self.B = int(block_size or self.datas[0].block_size)
        self.max_blocks = int(max_blocks)
        self.device = self.datas[0].device
        dev = torch.device("cuda", self.device)
        delta_beta = self.datas[0].delta_beta if self.pm else 1.0
        self.batch = _native.Batch([d.context() for d in self.datas], self.B, pm=self.pm,
                                   noscale=noscale, delta_beta=delta_beta, max_blocks=self.max_blocks,
                                   fp32=fp32)
        self.fp32 = bool(fp32)
        self.q = [torch.zeros((self.B, d.modes, d.beads), dtype=torch.float64, device=dev)
                  for d in self.datas]
        self.out = [torch.full((_native.NUM_OUT, self.B), float("nan"), dtype=torch.float64, device=dev)
                    for _ in self.datas]
        self.sums = torch.zeros((self.max_blocks, len(self.datas), _native.NUM_SUMS),
                                dtype=torch.float64, device=dev)
        for i in range(len(self.datas)):
            self.batch.bind(i, self.q[i], self.out[i], self.B)
        self.batch.bind_sums(self.sums)

    @property
    def n(self):
        return len(self.datas)

    def schedule(self):
        return self.batch.schedule()

    def set_schedule(self, family, lanes=None, tile_waves=4, grid_waves=2048):
        fam = {"lanes": _native.BATCH_LANES, "wide": _native.BATCH_WIDE,
               "sequential": _native.BATCH_SEQUENTIAL}.get(family, family)
        self.batch.set_schedule(fam, lanes, tile_waves, grid_waves)

    def tune(self, seed=1):
        """Autotune the schedule on the device; returns timings and the chosen schedule."""
        best, seq = self.batch.tune(seed)
        return {"batched_ms": best, "sequential_ms": seq, "schedule": self.schedule()}

    def evaluate(self, q_list=None, sequential=False):
        """Evaluate host-fed coordinates (one (B, N, P_i) array per problem) or, with
        q_list=None, whatever the q buffers hold. Returns per-problem dicts like
        pimc.evaluate_samples."""
        import torch
        if q_list is not None:
            for i, q in enumerate(q_list):
                q = np.ascontiguousarray(q, dtype=np.float64)
                assert q.shape == tuple(self.q[i].shape), (i, q.shape, tuple(self.q[i].shape))
                self.q[i].copy_(torch.from_numpy(q))
        if sequential:
            self.batch.eval_sequential()
        else:
            self.batch.eval()
        torch.cuda.synchronize(self.device)
        return [self._result(i) for i in range(self.n)]

    def _result(self, i):
        o = self.out[i].cpu().numpy()
        r = dict(s_rho=o[_native.OUT_RHO].copy(), s_g=o[_native.OUT_G].copy(),
                 log_scale=o[_native.OUT_LOGS].copy())
        if self.pm:
            r.update(s_gP=o[_native.OUT_GP].copy(), s_gM=o[_native.OUT_GM].copy())
        return r

    def run(self, seed, first_block=0, nblocks=1, stream=None):
        """Blocks first_block .. first_block+nblocks-1 of every problem: device sampling
        (the Philox counters of pib_sample_block at offset block*B), one persistent evaluation
        launch and the block sums, as hipGraph replays. Returns nothing; see block_sums()."""
        self.batch.run(seed, first_block, nblocks, stream)

    def block_sums(self):
        """(max_blocks, n, NUM_SUMS) numpy array of the per-block sums written so far."""
        import torch
        torch.cuda.synchronize(self.device)
        return self.sums.cpu().numpy()
