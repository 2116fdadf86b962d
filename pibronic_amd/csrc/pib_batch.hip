// Batched evaluation: many (model, temperature) problems in one persistent launch, a schedule
// autotuner, and the whole block step (sample, evaluate, reduce) captured once as a hipGraph.
//
// The reference runs one temperature per process (pibronic/server/job_boss.py:538-611 farms
// them out); here every problem is a context (model constants, O-matrix sets, beads, sampler)
// and the batch owns the device buffers of one block per problem. Problems may have different
// bead counts; the per-lane kernel family needs one (A, A_rho) for all of them, the wide family
// takes any A, A_rho <= 16.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/pibronic_hip.h"
#include "pib_ctx.hpp"
#include "pib_kernels.hpp"

using namespace pib;

namespace {

__global__ void k_set_i64(int64_t* p, int64_t v) {
  if (threadIdx.x == 0) *p = v;
}
__global__ void k_zero_i32(int* p) {
  if (threadIdx.x == 0) *p = 0;
}
__global__ void k_inc_i64(int64_t* p) {
  if (threadIdx.x == 0) *p += 1;
}

int pow2_floor(int x) {
  int g = 1;
  while (g * 2 <= x) g *= 2;
  return g;
}

}  // namespace

struct pib_batch {
  int device = 0, num_cu = 256;
  std::string err;
  int n = 0, flags = 0, pm = 0, max_blocks = 0;
  int64_t B = 0;
  double delta_beta = 1.0;
  std::vector<pib_ctx*> ctx;
  std::vector<uint64_t> gens;  // each context's setter generation when the batch copied its pointers
  std::vector<EvalArgs> args;  // per problem; q/out point into the batch buffers
  LaneProbs lprobs{};          // the lanes family's kernel-argument copy of args
  bool lane_ok = false, wide_ok = false;
  bool seq_lanes = true;  // every problem has a compiled lanes kernel (one launch per problem)
  int A = 0, AR = 0;
  // schedule
  int family = 0;  // PIB_BATCH_LANES or PIB_BATCH_WIDE
  std::vector<int> lanes;
  int tile_waves = 4, grid_waves = 0;
  int ntiles = 0;
  // caller-owned buffers (pib_batch_bind*), library-owned scratch
  std::vector<bool> bound;
  double* sums = nullptr;
  DevBuf d_args, d_tiles, d_counter, d_step, partials;
  int nparts = 1;
  // graph of one block step
  hipStream_t cap = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t graph_seed = 0;
  bool graph_valid = false;
};

namespace {

int bfail(pib_batch* b, int code, const std::string& m) {
  if (b) b->err = m;
  return code;
}

#define PIB_BHIP(b, call)                                                            \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) return bfail((b), PIB_ERR_HIP, std::string(#call) + ": " + \
                                                             hipGetErrorString(e_)); \
  } while (0)

int balloc(pib_batch* b, DevBuf& d, size_t bytes) {
  if (d.p) return PIB_OK;
  PIB_BHIP(b, hipMalloc(&d.p, bytes ? bytes : 8));
  d.bytes = bytes;
  return PIB_OK;
}

void drop_graph(pib_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  b->graph_valid = false;
}

// Tiles of `tile_waves` waves (lanes family) or samples (wide family), ordered by decreasing
// estimated cost so that the longest tiles start first.
int build_schedule(pib_batch* b) {
  struct T {
    BatchTile t;
    double cost;
  };
  std::vector<T> tiles;
  for (int i = 0; i < b->n; ++i) {
    EvalArgs& a = b->args[i];
    int64_t units;
    double unit_cost;
    if (b->family == PIB_BATCH_LANES || b->family == PIB_BATCH_SEQUENTIAL) {
      a.G = b->lanes[i];
      a.K = (a.P + a.G - 1) / a.G;
      const int spw = kWave / a.G;
      units = (b->B + spw - 1) / spw;
      unit_cost = (double)a.K + 2.0 * __builtin_ctz((unsigned)a.G);
    } else {
      units = b->B;
      unit_cost = (double)a.P;
    }
    for (int64_t w = 0; w < units; w += b->tile_waves) {
      T t;
      t.t.prob = i;
      t.t.wave0 = w;
      t.t.nwaves = (int)std::min<int64_t>(b->tile_waves, units - w);
      t.cost = unit_cost * t.t.nwaves;
      tiles.push_back(t);
    }
  }
  std::stable_sort(tiles.begin(), tiles.end(), [](const T& x, const T& y) { return x.cost > y.cost; });
  if (b->lane_ok) {
    b->lprobs.n = b->n;
    for (int i = 0; i < b->n; ++i) b->lprobs.p[i] = lane_prob(b->args[i]);
  }
  std::vector<BatchTile> ht(tiles.size());
  for (size_t k = 0; k < tiles.size(); ++k) ht[k] = tiles[k].t;
  if (ht.size() > (size_t)0x7fffffff) return bfail(b, PIB_ERR_ARG, "batch: too many tiles");
  b->ntiles = (int)ht.size();
  PIB_BHIP(b, hipDeviceSynchronize());  // no queued launch may still read the old schedule
  release(b->d_tiles);
  PIB_BHIP(b, hipMalloc(&b->d_tiles.p, ht.size() * sizeof(BatchTile)));
  b->d_tiles.bytes = ht.size() * sizeof(BatchTile);
  PIB_BHIP(b, hipMemcpy(b->d_tiles.p, ht.data(), ht.size() * sizeof(BatchTile), hipMemcpyHostToDevice));
  PIB_BHIP(b, hipMemcpy(b->d_args.p, b->args.data(), b->n * sizeof(EvalArgs), hipMemcpyHostToDevice));
  drop_graph(b);
  return PIB_OK;
}

hipError_t launch_batch_eval(pib_batch* b, hipStream_t s) {
  if (b->family == PIB_BATCH_SEQUENTIAL) {  // one launch per problem, tuned lanes
    for (int i = 0; i < b->n; ++i) {
      const EvalArgs& a = b->args[i];
      hipError_t e = (b->flags & PIB_EVAL_FP32) ? launch_eval_fp32(a.A, a.AR, b->pm, a, s)
                     : b->seq_lanes             ? launch_eval_fp64(a.A, a.AR, b->pm, a, s)
                                                : launch_eval_wide(b->pm, a, s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, (int*)b->d_counter.p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const EvalArgs* pa = (const EvalArgs*)b->d_args.p;
  const BatchTile* pt = (const BatchTile*)b->d_tiles.p;
  int* cnt = (int*)b->d_counter.p;
  if (b->family == PIB_BATCH_LANES)
    return launch_eval_batched(b->A, b->AR, b->pm, b->lprobs, pt, b->ntiles, cnt, b->grid_waves, s);
  return launch_eval_wide_batched(b->pm, pa, pt, b->ntiles, cnt, b->grid_waves, s);
}

int check_bound(pib_batch* b, const char* who) {
  for (int i = 0; i < b->n; ++i) {
    if (!b->bound[i]) return bfail(b, PIB_ERR_STATE, std::string(who) + ": buffers of problem " + std::to_string(i) + " not bound");
    // the argument blocks hold the context's device pointers as they were at batch_create;
    // a setter since then may have reallocated them
    if (b->ctx[i]->gen != b->gens[i])
      return bfail(b, PIB_ERR_STATE, std::string(who) + ": the constants of problem " + std::to_string(i) +
                                         " were set again after batch_create; create a new batch");
  }
  return PIB_OK;
}

int check_samplers(pib_batch* b, const char* who) {
  for (int i = 0; i < b->n; ++i) {
    const pib_ctx* c = b->ctx[i];
    if (!c->sampler_set || c->s_N != b->args[i].N || c->s_P != b->args[i].P)
      return bfail(b, PIB_ERR_STATE, std::string(who) + ": sampler not set, or its N/P differ from the beads, for problem " +
                                         std::to_string(i));
  }
  return PIB_OK;
}

SampleArgs sample_args(pib_batch* b, int i, uint64_t seed, const int64_t* step) {
  pib_ctx* c = b->ctx[i];
  SampleArgs a{};
  a.B = b->B;
  a.N = c->s_N;
  a.P = c->s_P;
  a.A_rho = c->s_A;
  a.offset = 0;
  a.key0 = (uint32_t)seed;
  a.key1 = (uint32_t)(seed >> 32);
  a.cdf = (const double*)c->s_cdf.p;
  a.chol = (const double*)c->s_chol.p;
  a.shift = (const double*)c->s_shift.p;
  a.q = const_cast<double*>(b->args[i].q);
  a.src = nullptr;
  a.step = step;
  return a;
}

SumArgs sum_args(pib_batch* b, int i, const int64_t* step) {
  SumArgs a{};
  a.out = b->args[i].out;
  a.out_stride = b->args[i].out_stride;
  a.B = b->B;
  a.pm = b->pm;
  a.delta_beta = b->pm ? b->delta_beta : 1.0;
  a.logq = nullptr;
  a.step = step;
  a.step_stride = (int64_t)b->n * NUM_SUMS;
  a.max_steps = b->max_blocks;
  return a;
}

double time_ms(pib_batch* b, hipStream_t s, int reps, int* rc) {
  hipEvent_t e0, e1;
  *rc = PIB_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    *rc = bfail(b, PIB_ERR_HIP, "tune: event create");
    return 0.0;
  }
  double best = 1e300;
  for (int r = 0; r < reps + 1; ++r) {
    (void)hipEventRecord(e0, s);
    if (launch_batch_eval(b, s) != hipSuccess) {
      *rc = bfail(b, PIB_ERR_HIP, "tune: launch");
      break;
    }
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) {
      *rc = bfail(b, PIB_ERR_HIP, "tune: synchronize");
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (r > 0) best = std::min(best, (double)ms);  // first run is a warm-up
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

}  // namespace

extern "C" {

int pib_batch_create(int n, pib_ctx* const* ctxs, int64_t B, int flags, double delta_beta,
                     int max_blocks, pib_batch** out) {
  if (!out) return PIB_ERR_ARG;
  *out = nullptr;
  if (n < 1 || !ctxs || B < 1 || max_blocks < 1) return PIB_ERR_ARG;
  if ((flags & PIB_EVAL_PM) && !(delta_beta > 0.0)) return PIB_ERR_ARG;
  if (flags & ~(PIB_EVAL_PM | PIB_EVAL_NOSCALE | PIB_EVAL_FP32)) return PIB_ERR_ARG;
  pib_batch* b = new pib_batch();
  *out = b;
  b->n = n;
  b->B = B;
  b->flags = flags;
  b->pm = (flags & PIB_EVAL_PM) ? 1 : 0;
  b->delta_beta = b->pm ? delta_beta : 1.0;
  b->max_blocks = max_blocks;
  b->ctx.assign(ctxs, ctxs + n);
  b->device = ctxs[0]->device;
  b->num_cu = ctxs[0]->num_cu;
  b->lane_ok = true;
  b->wide_ok = true;
  b->args.resize(n);
  b->lanes.resize(n);
  b->gens.resize(n);
  for (int i = 0; i < n; ++i) {
    pib_ctx* c = ctxs[i];
    if (!c) return bfail(b, PIB_ERR_ARG, "batch_create: null context");
    if (c->device != b->device) return bfail(b, PIB_ERR_ARG, "batch_create: contexts on different devices");
    bool wide = false;
    if (make_eval_args(c, flags, B, &b->args[i], &wide))
      return bfail(b, PIB_ERR_STATE, "batch_create: problem " + std::to_string(i) + ": " + c->err);
    b->gens[i] = c->gen;
    const int A = c->A, AR = b->args[i].AR;
    if (wide) b->seq_lanes = false;
    if (wide || (i > 0 && (A != b->A || AR != b->AR))) b->lane_ok = false;
    if (i == 0) { b->A = A; b->AR = AR; }
    if (A > kWideMaxStates || AR > kWideMaxStates) b->wide_ok = false;
    b->lanes[i] = auto_lanes(c, B, (flags & PIB_EVAL_FP32) != 0);
  }
  if (flags & PIB_EVAL_FP32) b->wide_ok = false;  // FP32 runs as one launch per problem
  if (n > kBatchLaunchProbs) b->lane_ok = false;   // the lanes kernel takes its problems by value
  if (!b->lane_ok && !b->wide_ok && !b->seq_lanes)
    return bfail(b, PIB_ERR_UNSUPPORTED, "batch_create: no kernel family fits every problem");
  if (hipSetDevice(b->device) != hipSuccess) return bfail(b, PIB_ERR_HIP, "batch_create: set device");
  b->nparts = (int)std::max<int64_t>(1, std::min<int64_t>((B + 255) / 256, 1024));
  int r;
  if ((r = balloc(b, b->d_args, (size_t)n * sizeof(EvalArgs)))) return r;
  if ((r = balloc(b, b->d_counter, sizeof(int)))) return r;
  if ((r = balloc(b, b->d_step, sizeof(int64_t)))) return r;
  if ((r = balloc(b, b->partials, (size_t)b->nparts * NUM_SUMS * sizeof(double)))) return r;
  b->bound.assign(n, false);
  for (int i = 0; i < n; ++i) {
    b->args[i].q = nullptr;
    b->args[i].out = nullptr;
    b->args[i].out_stride = B;
    b->args[i].dbg = nullptr;
  }
  b->family = (flags & PIB_EVAL_FP32) ? PIB_BATCH_SEQUENTIAL
              : b->lane_ok            ? PIB_BATCH_LANES
              : b->wide_ok            ? PIB_BATCH_WIDE
                                      : PIB_BATCH_SEQUENTIAL;
  b->tile_waves = 4;
  b->grid_waves = b->num_cu * 8;
  PIB_BHIP(b, hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
  return build_schedule(b);
}

int pib_batch_destroy(pib_batch* b) {
  if (!b) return PIB_OK;
  (void)hipSetDevice(b->device);
  (void)hipDeviceSynchronize();
  drop_graph(b);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  release(b->d_args);
  release(b->d_tiles);
  release(b->d_counter);
  release(b->d_step);
  release(b->partials);
  delete b;
  return PIB_OK;
}

const char* pib_batch_last_error(const pib_batch* b) { return b ? b->err.c_str() : "null batch"; }

int pib_batch_bind(pib_batch* b, int i, const double* q, double* out, int64_t out_stride) {
  if (!b) return PIB_ERR_ARG;
  if (i < 0 || i >= b->n) return bfail(b, PIB_ERR_ARG, "batch_bind: bad problem index");
  if (!q || !out || out_stride < b->B) return bfail(b, PIB_ERR_ARG, "batch_bind: null buffer or out_stride < B");
  b->args[i].q = q;
  b->args[i].out = out;
  b->args[i].out_stride = out_stride;
  b->bound[i] = true;
  return build_schedule(b);
}

int pib_batch_bind_sums(pib_batch* b, double* sums) {
  if (!b) return PIB_ERR_ARG;
  if (!sums) return bfail(b, PIB_ERR_ARG, "batch_bind_sums: null buffer");
  b->sums = sums;
  drop_graph(b);
  return PIB_OK;
}

int pib_batch_set_schedule(pib_batch* b, int family, const int* lanes, int tile_waves, int grid_waves) {
  if (!b) return PIB_ERR_ARG;
  if ((b->flags & PIB_EVAL_FP32) && family != PIB_BATCH_SEQUENTIAL)
    return bfail(b, PIB_ERR_UNSUPPORTED, "batch_set_schedule: FP32 batches run one launch per problem");
  if (family == PIB_BATCH_LANES && !b->lane_ok)
    return bfail(b, PIB_ERR_UNSUPPORTED, "batch_set_schedule: the lanes family needs one compiled (A, A_rho)");
  if (family == PIB_BATCH_WIDE && !b->wide_ok)
    return bfail(b, PIB_ERR_UNSUPPORTED, "batch_set_schedule: the wide family needs A, A_rho <= 16");
  if (family != PIB_BATCH_LANES && family != PIB_BATCH_WIDE && family != PIB_BATCH_SEQUENTIAL)
    return bfail(b, PIB_ERR_ARG, "batch_set_schedule: bad family");
  if (tile_waves < 1 || grid_waves < 1) return bfail(b, PIB_ERR_ARG, "batch_set_schedule: tile and grid must be >= 1");
  if (lanes) {
    for (int i = 0; i < b->n; ++i) {
      const int g = lanes[i];
      if (g < 1 || g > 64 || (g & (g - 1))) return bfail(b, PIB_ERR_ARG, "batch_set_schedule: lanes must be powers of two <= 64");
    }
    b->lanes.assign(lanes, lanes + b->n);
  }
  b->family = family;
  b->tile_waves = tile_waves;
  b->grid_waves = grid_waves;
  return build_schedule(b);
}

int pib_batch_get_schedule(pib_batch* b, int* family, int* lanes, int* tile_waves, int* grid_waves) {
  if (!b) return PIB_ERR_ARG;
  if (family) *family = b->family;
  if (lanes) std::copy(b->lanes.begin(), b->lanes.end(), lanes);
  if (tile_waves) *tile_waves = b->tile_waves;
  if (grid_waves) *grid_waves = b->grid_waves;
  return PIB_OK;
}

int pib_batch_eval(pib_batch* b, void* stream) {
  if (!b) return PIB_ERR_ARG;
  if (int r = check_bound(b, "batch_eval")) return r;
  PIB_BHIP(b, hipSetDevice(b->device));
  PIB_BHIP(b, launch_batch_eval(b, (hipStream_t)stream));
  return PIB_OK;
}

int pib_batch_eval_sequential(pib_batch* b, void* stream) {
  if (!b) return PIB_ERR_ARG;
  if (int r = check_bound(b, "batch_eval_sequential")) return r;
  PIB_BHIP(b, hipSetDevice(b->device));
  for (int i = 0; i < b->n; ++i) {
    EvalArgs a = b->args[i];
    bool wide = false;
    EvalArgs s{};
    if (make_eval_args(b->ctx[i], b->flags, b->B, &s, &wide))
      return bfail(b, PIB_ERR_STATE, "batch_eval_sequential: " + b->ctx[i]->err);
    s.q = a.q;
    s.out = a.out;
    s.out_stride = a.out_stride;
    s.dbg = nullptr;
    hipError_t e = (b->flags & PIB_EVAL_FP32) ? launch_eval_fp32(s.A, s.AR, b->pm, s, (hipStream_t)stream)
                   : wide ? launch_eval_wide(b->pm, s, (hipStream_t)stream)
                          : launch_eval_fp64(s.A, s.AR, b->pm, s, (hipStream_t)stream);
    if (e != hipSuccess) return bfail(b, PIB_ERR_HIP, std::string("batch_eval_sequential: ") + hipGetErrorString(e));
  }
  return PIB_OK;
}

int pib_batch_tune(pib_batch* b, uint64_t seed, double* best_ms, double* sequential_ms, void* stream) {
  if (!b) return PIB_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  PIB_BHIP(b, hipSetDevice(b->device));
  int rc0 = check_bound(b, "batch_tune");
  if (!rc0) rc0 = check_samplers(b, "batch_tune");
  if (rc0) return rc0;
  for (int i = 0; i < b->n; ++i) PIB_BHIP(b, launch_sample_fp64(sample_args(b, i, seed, nullptr), s));
  int rc = PIB_OK;
  // reference point: one launch per problem
  if (sequential_ms) {
    hipEvent_t e0, e1;
    PIB_BHIP(b, hipEventCreate(&e0));
    PIB_BHIP(b, hipEventCreate(&e1));
    double best = 1e300;
    for (int r = 0; r < 3 && rc == PIB_OK; ++r) {
      (void)hipEventRecord(e0, s);
      rc = pib_batch_eval_sequential(b, stream);
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (r > 0) best = std::min(best, (double)ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *sequential_ms = best;
  }
  struct Cand {
    int family, kt, tile, grid;
    std::vector<int> lanes;
    double ms;
  };
  auto lanes_for = [&](int kt) {
    std::vector<int> g(b->n);
    for (int i = 0; i < b->n; ++i) {
      const int P = b->args[i].P;
      g[i] = std::max(1, std::min(std::min(64, pow2_floor(P)), pow2_floor(std::max(1, P / kt))));
    }
    return g;
  };
  auto run = [&](Cand& c) -> int {
    b->family = c.family;
    if (c.family != PIB_BATCH_WIDE) b->lanes = c.lanes;
    b->tile_waves = c.tile;
    b->grid_waves = c.grid;
    int r = build_schedule(b);
    if (r) return r;
    c.ms = time_ms(b, s, 2, &r);
    return r;
  };
  Cand best{-1, 0, 0, 0, {}, 1e300};
  const int grids[3] = {b->num_cu * 4, b->num_cu * 8, b->num_cu * 16};
  for (int fam = 0; fam < 3; ++fam) {
    if (fam == PIB_BATCH_LANES && (!b->lane_ok || (b->flags & PIB_EVAL_FP32))) continue;
    if (fam == PIB_BATCH_WIDE && !b->wide_ok) continue;
    Cand cur{fam, 4, 4, b->num_cu * 8, lanes_for(4), 1e300};
    if (best.family >= 0) {  // prune a family whose default schedule is far behind
      Cand probe = cur;
      if ((rc = run(probe))) return rc;
      if (probe.ms > 1.5 * best.ms) continue;
    }
    if (fam == PIB_BATCH_SEQUENTIAL) {  // per-problem launches: only the lanes matter
      if (!b->seq_lanes) {
        if ((rc = run(cur))) return rc;
      } else {
        for (int kt : {2, 4, 8, 16}) {
          Cand c = cur;
          c.kt = kt;
          c.lanes = lanes_for(kt);
          if ((rc = run(c))) return rc;
          if (c.ms < cur.ms) cur = c;
        }
      }
      if (cur.ms < best.ms) best = cur;
      continue;
    }
    if (fam == PIB_BATCH_LANES) {  // beads per lane
      Cand bk = cur;
      for (int kt : {2, 4, 8, 16}) {
        Cand c = cur;
        c.kt = kt;
        c.lanes = lanes_for(kt);
        if ((rc = run(c))) return rc;
        if (c.ms < bk.ms) bk = c;
      }
      cur = bk;
    }
    {  // tile size
      Cand bt = cur;
      bt.ms = 1e300;
      for (int tile : {1, 4, 16}) {
        Cand c = cur;
        c.tile = tile;
        if ((rc = run(c))) return rc;
        if (c.ms < bt.ms) bt = c;
      }
      cur = bt;
    }
    {  // persistent grid size
      Cand bg = cur;
      bg.ms = 1e300;
      for (int grid : grids) {
        Cand c = cur;
        c.grid = grid;
        if ((rc = run(c))) return rc;
        if (c.ms < bg.ms) bg = c;
      }
      cur = bg;
    }
    if (cur.ms < best.ms) best = cur;
  }
  b->family = best.family;
  if (best.family != PIB_BATCH_WIDE) b->lanes = best.lanes;
  b->tile_waves = best.tile;
  b->grid_waves = best.grid;
  if ((rc = build_schedule(b))) return rc;
  if (best_ms) *best_ms = best.ms;
  return PIB_OK;
}

int pib_batch_run(pib_batch* b, uint64_t seed, int64_t first_block, int nblocks, void* stream) {
  if (!b) return PIB_ERR_ARG;
  if (first_block < 0 || nblocks < 0 || first_block + nblocks > b->max_blocks)
    return bfail(b, PIB_ERR_ARG, "batch_run: blocks outside [0, max_blocks)");
  int rc0 = check_bound(b, "batch_run");
  if (!rc0) rc0 = check_samplers(b, "batch_run");
  if (rc0) return rc0;
  if (!b->sums) return bfail(b, PIB_ERR_STATE, "batch_run: sums buffer not bound");
  PIB_BHIP(b, hipSetDevice(b->device));
  if (!b->graph_valid || b->graph_seed != seed) {
    drop_graph(b);
    // one block step: sample every problem at offset step*B, evaluate all in one persistent
    // launch, reduce every problem's block into sums[step][problem], advance step. Only
    // kernels are recorded (no memset or copy nodes).
    const int64_t* step = (const int64_t*)b->d_step.p;
    PIB_BHIP(b, hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int i = 0; i < b->n && e == hipSuccess; ++i) e = launch_sample_fp64(sample_args(b, i, seed, step), b->cap);
    if (e == hipSuccess) e = launch_batch_eval(b, b->cap);
    for (int i = 0; i < b->n && e == hipSuccess; ++i)
      e = launch_block_sums(sum_args(b, i, step), (double*)b->partials.p, b->nparts,
                            b->sums + (int64_t)i * NUM_SUMS, b->cap);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_inc_i64, dim3(1), dim3(64), 0, b->cap, (int64_t*)b->d_step.p);
      e = hipGetLastError();
    }
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->cap, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return bfail(b, PIB_ERR_HIP, std::string("batch_run: capture failed: ") +
                                       hipGetErrorString(e != hipSuccess ? e : e2));
    }
    hipError_t e3 = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e3 != hipSuccess) return bfail(b, PIB_ERR_HIP, std::string("batch_run: instantiate: ") + hipGetErrorString(e3));
    b->graph_seed = seed;
    b->graph_valid = true;
  }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(64), 0, s, (int64_t*)b->d_step.p, first_block);
  PIB_BHIP(b, hipGetLastError());
  for (int k = 0; k < nblocks; ++k) PIB_BHIP(b, hipGraphLaunch(b->exec, s));
  return PIB_OK;
}

}  // extern "C"
