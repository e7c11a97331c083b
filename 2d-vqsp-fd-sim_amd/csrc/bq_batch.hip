// bq_batch.hip - many grids advanced together on one MI355X.
//
// A batch groups the handles of independent runs (shots over models of
// different sizes, each a whole TIME_LOOP of its own, time_loop.F90:65-206).
// One time step of every shot is three launches in all, not three per shot:
//
//   kb_step      fused interior tiles of every shot, then the stress half-step
//                of every shot's other cells and the receiver samples of the
//                previous step
//   kb_velocity  the velocity half-step of every shot's other cells
//   kb_edges     row 1, row PMLZH and the bottom history of every shot
//   kb_lb_apply, kb_lb_extract   for the shots with a learned boundary
//                attached: the strip cells from the network, then the band
//                of the new level into the shot's ring
//
// A block finds its shot from the prefix sums of the shots' block counts and
// runs exactly the per-shot device code of the single-grid launches, so every
// shot's fields are bit-identical to a run of that shot alone. Two schedules:
// one block per work item (the hardware dispatcher balances), or a persistent
// grid of a few blocks per CU that take work items from per-XCD queues. The
// steps of a run are replayed from a HIP graph of two steps; a device counter
// gives each replay its step numbers. The tile height and the schedule are
// chosen by timing each candidate on the batch's own steps (bq_batch_autotune):
// all candidates compute the same cells with the same expressions, so tuning
// does not change the results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/biotfd.h"
#include "bq_handle.h"
#include "bq_learned.h"

namespace {

constexpr int kPersistStep = 2;   // persistent k_step blocks per CU (LDS-limited)
constexpr int kPersistVel = 8;    // persistent k_velocity blocks per CU
constexpr int kXcd = 8;

// One shot as the batched kernels see it.
struct ShotDev {
  KParams kp[2];               // parameters of a step of parity 0 and 1
  Fused fu;                    // fused tiles at the batch's tile height
  Split ss, sv;                // stress / velocity cells outside the tiles
  int nfused, nrest, nruns, nsample;   // kb_step blocks
  int vrest, vruns;                    // kb_velocity blocks
  int nedge;                           // kb_edges blocks
  int base_step, base_par;             // step number and parity at the run's start
  int lb_on;                           // learned boundary attached
  LbNet lb;
  float *ring;                         // its band levels, level n in slot n % 4
  int nlb, nlbx;                       // kb_lb_apply, kb_lb_extract blocks
};

// Prefix sums over the shots of each launch's blocks ([nshots + 1] each).
// C_FUSED: all fused blocks of a shot (persistent schedule); C_F1, C_F2, C_F4:
// its full-tile, half-tile and quarter-tile blocks (grid schedule).
enum { C_FUSED = 0, C_OTHER, C_VEL, C_EDGE, C_SAMPLE, C_LB, C_LBX, C_F1, C_F2, C_F4, C_N };

struct BatchArgs {
  const ShotDev *shots;
  const int *cum;      // [C_N][nshots + 1]
  int nshots;
  int total[C_N];
  int *dstep;          // steps of the current run done before this replay
  int *queue;          // persistent schedule: [0..7] fused bands, [8] others, [9] blocks done
};

// Largest s with cum[s] <= b (cum[0] = 0, non-decreasing): the shot of block b.
__device__ __forceinline__ int find_shot(const int *cum, int n, int b) {
  int lo = 0, hi = n;  // cum[lo] <= b < cum[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cum[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void other_item(const BatchArgs &B, int w, int kofs);

// Blocks of a grid-schedule kb_step launch: every XCD (block b runs on XCD
// b % 8) takes its share of the full-tile blocks of all shots, then of the
// half-tile blocks, then of the quarter-tile blocks, so the short blocks are
// dispatched last everywhere and fill in behind the long ones.
inline int batch_fused_launch(const int *total) {
  int l = 0;
  for (int c = C_F1; c <= C_F4; ++c) l += (total[c] + 7) / 8;
  return 8 * l;
}

// Shot and fused block (numbered as fused_block takes them) of grid-schedule
// block b; false past the XCD's share.
__device__ __forceinline__ bool batch_fused(const BatchArgs &B, int b, int *shot, int *blk) {
  const int x = b & 7;
  int q = b >> 3;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int n = B.total[C_F1 + c], per = n >> 3, rem = n & 7;
    const int cnt = per + (x < rem);
    if (q < cnt) {
      const int w = x * per + min(x, rem) + q;
      const int *cum = B.cum + (C_F1 + c) * (B.nshots + 1);
      const int s = find_shot(cum, B.nshots, w);
      const Fused &fu = B.shots[s].fu;
      const int t = w - cum[s];
      *shot = s;
      *blk = c == 0 ? t : fu.nfx * fu.nfz + (c == 2 ? fu.nb2 : 0) + t;
      return true;
    }
    q -= cnt;
  }
  return false;
}

// Item w of kb_step: items [0, total[C_FUSED]) are fused tiles (w is the tile
// number, tiles numbered shot after shot, row-major within a shot); the rest
// are the other cells' blocks and the sample blocks of every shot.
template <int FR>
__device__ __forceinline__ void step_item(const BatchArgs &B, int w, int kofs, FusedLds &L) {
  const int nf = B.total[C_FUSED];
  if (w < nf) {
    const int t = w;
    const int s = find_shot(B.cum + C_FUSED * (B.nshots + 1), B.nshots, t);
    const ShotDev &sd = B.shots[s];
    const KParams k = sd.kp[(sd.base_par + kofs) & 1];
    const Fused fu = sd.fu;
    fused_block<FR>(k, fu, t - B.cum[C_FUSED * (B.nshots + 1) + s], sd.base_step + kofs, L);
    return;
  }
  if (threadIdx.x >= kBlock) return;
  other_item(B, w - nf, kofs);
}

// Item w of the other cells' blocks and the sample blocks of every shot
// (kBlock threads).
__device__ __forceinline__ void other_item(const BatchArgs &B, int w, int kofs) {
  const int s = find_shot(B.cum + C_OTHER * (B.nshots + 1), B.nshots, w);
  const ShotDev &sd = B.shots[s];
  const KParams k = sd.kp[(sd.base_par + kofs) & 1];
  const int itile = sd.base_step + kofs;
  int b = w - B.cum[C_OTHER * (B.nshots + 1) + s];
  if (b < sd.nrest) {
    int i, l;
    const Split &sp = sd.ss;
    if (rest_cell(sp, b * kBlock + threadIdx.x, &i, &l)) stress_cell(k, i, l, itile);
    return;
  }
  b -= sd.nrest;
  if (b < sd.nruns) {
    int tile;
    const Split &sp = sd.ss;
    const int r = run_rect(sp, xcd_tile(b, sd.nruns), &tile);
    stress_run<kRun>(k, sp.qx0[r], sp.qxl[r], sp.qntx[r], sp.qz0[r], sp.qz1[r], tile, threadIdx.x >> 6,
                     threadIdx.x & 63, itile);
    return;
  }
  b -= sd.nruns;
  // samples of the previous step of this run (the run's first step has none:
  // the last sample of the previous run was taken at its end)
  if (kofs > 0) sample_receivers(k, b * kBlock + threadIdx.x, itile - 1, true);
}

__device__ __forceinline__ void velocity_item(const BatchArgs &B, int w, int kofs, float *lds) {
  const int s = find_shot(B.cum + C_VEL * (B.nshots + 1), B.nshots, w);
  const ShotDev &sd = B.shots[s];
  const KParams k = sd.kp[(sd.base_par + kofs) & 1];
  velocity_block(k, sd.sv, w - B.cum[C_VEL * (B.nshots + 1) + s], sd.vrest, sd.vruns, lds);
}

// Next item of a persistent block on XCD xcd (block b runs on XCD b % 8):
// first the fused tiles of its own band, then of the other bands, then the
// shared queue of the remaining items. -1: no work left.
__device__ __forceinline__ int next_item(int *q, int nf, int total, int xcd) {
  const int per = nf >> 3, rem = nf & 7;
  for (int d = 0; d < kXcd; ++d) {
    const int x = (xcd + d) & 7;
    const int size = per + (x < rem);
    if (__hip_atomic_load(&q[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= size) continue;
    const int c = atomicAdd(&q[x], 1);
    if (c < size) return x * per + min(x, rem) + c;
  }
  const int c = atomicAdd(&q[kXcd], 1);
  return nf + c < total ? nf + c : -1;
}

// The last block of a persistent launch to finish clears the queue for the
// next launch (stream order makes the clearing visible to it).
__device__ __forceinline__ void persistent_done(int *q) {
  __threadfence();
  if (atomicAdd(&q[kXcd + 1], 1) == (int)gridDim.x - 1) {
    for (int x = 0; x <= kXcd + 1; ++x) __hip_atomic_store(&q[x], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// (512 threads, at least 4 waves per SIMD: two blocks per CU, as the LDS allows)
template <int FR, bool PERSIST>
__global__ __launch_bounds__(kStepBlock, 4) void kb_step(BatchArgs B, int kl) {
  __shared__ FusedLds L;
  __shared__ int item;
  const int kofs = *B.dstep + kl;
  const int total = B.total[C_FUSED] + B.total[C_OTHER];
  if (!PERSIST) {
    int s, blk;
    if (batch_fused(B, blockIdx.x, &s, &blk)) {
      const ShotDev &sd = B.shots[s];
      const KParams k = sd.kp[(sd.base_par + kofs) & 1];   // a copy: held in registers
      const Fused fu = sd.fu;
      fused_block<FR>(k, fu, blk, sd.base_step + kofs, L);
    }
    return;
  }
  const int xcd = blockIdx.x & 7;
  for (;;) {
    if (threadIdx.x == 0) item = next_item(B.queue, B.total[C_FUSED], total, xcd);
    __syncthreads();
    const int w = item;
    __syncthreads();  // every thread has read item before thread 0 writes the next
    if (w < 0) break;
    step_item<FR>(B, w, kofs, L);
  }
  if (threadIdx.x == 0) persistent_done(B.queue);
}

// The other cells' stress blocks and the samples of every shot, without the
// fused tiles' LDS, so that many of these short blocks share a CU (grid
// schedule; kb_step then runs the fused tiles alone).
__global__ __launch_bounds__(kBlock) void kb_rest(BatchArgs B, int kl) {
  other_item(B, blockIdx.x, *B.dstep + kl);
}

template <bool PERSIST>
__global__ __launch_bounds__(kBlock) void kb_velocity(BatchArgs B, int kl) {
  __shared__ int item;
  __shared__ float lds[kEdgeLds];
  const int kofs = *B.dstep + kl;
  const int total = B.total[C_VEL];
  if (!PERSIST) {
    velocity_item(B, blockIdx.x, kofs, lds);
    return;
  }
  for (;;) {
    if (threadIdx.x == 0) {
      const int c = atomicAdd(&B.queue[kXcd], 1);
      item = c < total ? c : -1;
    }
    __syncthreads();
    const int w = item;
    __syncthreads();
    if (w < 0) break;
    velocity_item(B, w, kofs, lds);
  }
  if (threadIdx.x == 0) persistent_done(B.queue);
}

// Edge rows of every shot.
__global__ __launch_bounds__(kEdgeBlock) void kb_edges(BatchArgs B, int kl) {
  __shared__ float su[kEdgeBlock + 1], sq[kEdgeBlock + 1];
  const int kofs = *B.dstep + kl;
  const int s = find_shot(B.cum + C_EDGE * (B.nshots + 1), B.nshots, blockIdx.x);
  const ShotDev &sd = B.shots[s];
  const KParams k = sd.kp[(sd.base_par + kofs) & 1];
  edges_block(k, blockIdx.x - B.cum[C_EDGE * (B.nshots + 1) + s], su, sq);
}

// Learned boundary of the shots that have one, after the step (the step's
// parameters: its "next" set holds the new level). Block b: a tile of one
// side of one shot.
__global__ __launch_bounds__(kLbBlock) void kb_lb_apply(BatchArgs B, int kl) {
  __shared__ LbLds S;
  const int kofs = *B.dstep + kl;
  const int s = find_shot(B.cum + C_LB * (B.nshots + 1), B.nshots, blockIdx.x);
  const ShotDev &sd = B.shots[s];
  const KParams &k = sd.kp[(sd.base_par + kofs) & 1];
  const int it = sd.base_step + kofs;
  const int MX = k.MX, MZ = k.MZ, P = k.P, OFF = k.OFF;
  float *const U = k.U1, *const W = k.W1, *const QX = k.QX1, *const QZ = k.QZ1;
  float *const TXX = k.TXX1, *const TZZ = k.TZZ1, *const TXZ = k.TXZ1, *const PR = k.PR1;
  const int form = sd.lb.form;
  const size_t bs = lb_band_size(MX, MZ);
  const LbLevels lev{sd.ring + (size_t)((it + 0) & 3) * bs, sd.ring + (size_t)((it + 1) & 3) * bs,
                     sd.ring + (size_t)((it + 2) & 3) * bs, sd.ring + (size_t)((it + 3) & 3) * bs};
  const int b = blockIdx.x - B.cum[C_LB * (B.nshots + 1) + s];
  const int t0 = lb_tiles(lb_lz(MZ)), t1 = 2 * t0;
  const int side = b < t0 ? 0 : (b < t1 ? 1 : 2);
  const int tile = b - (side == 0 ? 0 : side == 1 ? t0 : t1);
  lb_tile(sd.lb, side, lb_side_len(side, MX, MZ), lb_side_off(side, MZ), tile * kLbP, lev, S,
          [=](int p, int f, int c, float prev, float d) {
            int i, l;
            lb_cell(side, p, c, MX, MZ, &i, &l);
            float *F = f == 0 ? U : f == 1 ? W : f == 2 ? QX : f == 3 ? QZ
                     : f == 4 ? TXX : f == 5 ? TZZ : f == 6 ? TXZ : PR;
            float &x = F[(size_t)l * P + (size_t)(i + OFF)];
            x = lb_update(form, x, prev, d);
          });
}

__global__ __launch_bounds__(kBlock) void kb_lb_extract(BatchArgs B, int kl) {
  const int kofs = *B.dstep + kl;
  const int s = find_shot(B.cum + C_LBX * (B.nshots + 1), B.nshots, blockIdx.x);
  const ShotDev &sd = B.shots[s];
  const KParams k = sd.kp[(sd.base_par + kofs) & 1];
  const int it = sd.base_step + kofs;
  const size_t bs = lb_band_size(k.MX, k.MZ);
  const size_t e = (size_t)(blockIdx.x - B.cum[C_LBX * (B.nshots + 1) + s]) * kBlock + threadIdx.x;
  if (e >= bs) return;
  const float *F[8] = {k.U1, k.W1, k.QX1, k.QZ1, k.TXX1, k.TZZ1, k.TXZ1, k.PR1};
  lb_extract(sd.ring + (size_t)(it & 3) * bs, F, k.P, k.OFF, k.MX, k.MZ, e, 0);
}

// Adds n steps to the run's step counter, after every launch that reads it.
__global__ void kb_advance(int *dstep, int n) {
  if (threadIdx.x == 0) *dstep += n;
}

// Samples of the last step of a run (step offset kl from the counter), from
// the set it wrote.
__global__ void kb_sample(BatchArgs B, int kl) {
  const int kofs = *B.dstep + kl;
  const int s = find_shot(B.cum + C_SAMPLE * (B.nshots + 1), B.nshots, blockIdx.x);
  const ShotDev &sd = B.shots[s];
  const KParams k = sd.kp[(sd.base_par + kofs) & 1];
  const int j = (blockIdx.x - B.cum[C_SAMPLE * (B.nshots + 1) + s]) * blockDim.x + threadIdx.x;
  sample_receivers(k, j, sd.base_step + kofs, false);
}

__global__ void kb_clear(int *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

}  // namespace

// ===================================================================== host

struct bq_batch {
  std::vector<bq_handle *> shots;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<ShotDev> host;
  ShotDev *dshots = nullptr;
  int *dcum = nullptr;     // [C_N][nshots + 1]
  int *dctl = nullptr;     // [0]: step counter, [8..19]: persistent queue
  int total[C_N] = {};
  int fr = 128;            // fused tile height
  int sched = BQ_SCHED_GRID;
  int use_graph = 1;
  int cus = 256;
  hipGraphExec_t graph = nullptr;
  int graph_fr = 0, graph_sched = -1;
  int graph_total[C_N] = {};   // block counts the graph was made for (none: graph null)
  hipEvent_t ev[2] = {nullptr, nullptr};
  int last_steps = 0;
  std::vector<double> tune_ms;   // per candidate, ms per step
  std::string err;
};

namespace {

int bfail(bq_batch *b, int code, const std::string &msg) {
  if (b) b->err = msg;
  return code;
}

#define BCHK(b, call)                                                               \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return bfail(b, -100, std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

thread_local std::string g_batch_create_err;

BatchArgs args_of(const bq_batch *b) {
  BatchArgs a{};
  a.shots = b->dshots;
  a.cum = b->dcum;
  a.nshots = (int)b->shots.size();
  for (int c = 0; c < C_N; ++c) a.total[c] = b->total[c];
  a.dstep = b->dctl;
  a.queue = b->dctl + 8;
  return a;
}

// Block counts of every shot for tile height fr, prefix sums, device copy.
int layout(bq_batch *b, int fr) {
  const int n = (int)b->shots.size();
  std::vector<int> cum((size_t)C_N * (n + 1), 0);
  for (int s = 0; s < n; ++s) {
    const bq_handle *h = b->shots[s];
    ShotDev &d = b->host[s];
    // the partition with fused tiles, whatever the shot's own bq_run uses:
    // the batch's launch holds enough tiles of all shots to fill the GPU
    const bq::Plan &p = h->tiles;
    d.fu = p.fu;
    bq::fused_rows(d.fu, fr);
    d.ss = p.ss;
    d.sv = p.sv;
    d.nfused = bq::fused_blocks(d.fu);
    d.nrest = (p.ss.cum[kMaxCells] + kBlock - 1) / kBlock;
    d.nruns = p.ss.qcum[kMaxRuns];
    d.nsample = (h->mr + kBlock - 1) / kBlock;
    d.vrest = (p.sv.cum[kMaxCells] + kBlock - 1) / kBlock;
    d.vruns = p.sv.qcum[kMaxRuns];
    d.nedge = p.sv.fold ? 0 : (h->g.NX + kEdgeBlock - 1) / kEdgeBlock;
    d.lb_on = h->lb != nullptr;
    d.nlb = d.lb_on ? 2 * lb_tiles(lb_lz(h->g.MZ)) + lb_tiles(lb_lx(h->g.MX)) : 0;
    d.nlbx = d.lb_on ? (int)((lb_band_size(h->g.MX, h->g.MZ) + kBlock - 1) / kBlock) : 0;
    const int cnt[C_N] = {d.nfused, d.nrest + d.nruns + d.nsample,
                          p.sv.etop + p.sv.ebot + d.vrest + d.vruns, d.nedge, d.nsample, d.nlb, d.nlbx,
                          d.fu.nfx * d.fu.nfz, d.fu.nb2, d.fu.nb4};
    for (int c = 0; c < C_N; ++c) cum[(size_t)c * (n + 1) + s + 1] = cum[(size_t)c * (n + 1) + s] + cnt[c];
  }
  for (int c = 0; c < C_N; ++c) b->total[c] = cum[(size_t)c * (n + 1) + n];
  BCHK(b, hipMemcpyAsync(b->dcum, cum.data(), cum.size() * sizeof(int), hipMemcpyHostToDevice,
                         b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  b->fr = fr;
  return 0;
}

int upload_shots(bq_batch *b) {
  const int n = (int)b->shots.size();
  for (int s = 0; s < n; ++s) {
    const bq_handle *h = b->shots[s];
    ShotDev &d = b->host[s];
    d.kp[0] = params_for(h, 0);
    d.kp[1] = params_for(h, 1);
    d.base_step = h->next_step;
    d.base_par = h->parity;
    d.ring = h->lb_ring;
    if (h->lb) d.lb = bq::lb_net_of(h->lb);
  }
  BCHK(b, hipMemcpyAsync(b->dshots, b->host.data(), b->host.size() * sizeof(ShotDev),
                         hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

// Queues the launches of one step (step offset kl from the counter) with the
// batch's tile height and schedule. advance: added to the counter at the end.
template <int FR>
void launch_step_fr(bq_batch *b, const BatchArgs &a, int kl, int advance) {
  const int nstep = a.total[C_FUSED] + a.total[C_OTHER];
  const int nvel = a.total[C_VEL];
  if (b->sched == BQ_SCHED_PERSISTENT) {
    if (nstep)
      hipLaunchKernelGGL((kb_step<FR, true>), dim3(std::min(nstep, kPersistStep * b->cus)),
                         dim3(kStepBlock), 0, b->stream, a, kl);
    if (nvel)
      hipLaunchKernelGGL((kb_velocity<true>), dim3(std::min(nvel, kPersistVel * b->cus)),
                         dim3(kBlock), 0, b->stream, a, kl);
  } else {
    if (a.total[C_FUSED])
      hipLaunchKernelGGL((kb_step<FR, false>), dim3(batch_fused_launch(a.total)), dim3(kStepBlock), 0,
                         b->stream, a, kl);
    if (a.total[C_OTHER])
      hipLaunchKernelGGL(kb_rest, dim3(a.total[C_OTHER]), dim3(kBlock), 0, b->stream, a, kl);
    if (nvel) hipLaunchKernelGGL((kb_velocity<false>), dim3(nvel), dim3(kBlock), 0, b->stream, a, kl);
  }
  if (a.total[C_EDGE])
    hipLaunchKernelGGL(kb_edges, dim3(a.total[C_EDGE]), dim3(kEdgeBlock), 0, b->stream, a, kl);
  if (a.total[C_LB]) {
    hipLaunchKernelGGL(kb_lb_apply, dim3(a.total[C_LB]), dim3(kLbBlock), 0, b->stream, a, kl);
    hipLaunchKernelGGL(kb_lb_extract, dim3(a.total[C_LBX]), dim3(kBlock), 0, b->stream, a, kl);
  }
  if (advance) hipLaunchKernelGGL(kb_advance, dim3(1), dim3(64), 0, b->stream, a.dstep, advance);
}

void launch_step(bq_batch *b, int kl, int advance) {
  const BatchArgs a = args_of(b);
  switch (b->fr) {
    case 64: launch_step_fr<64>(b, a, kl, advance); break;
    case 256: launch_step_fr<256>(b, a, kl, advance); break;
    default: launch_step_fr<128>(b, a, kl, advance); break;
  }
}

int build_graph(bq_batch *b) {
  if (b->graph && b->graph_fr == b->fr && b->graph_sched == b->sched &&
      std::equal(b->total, b->total + C_N, b->graph_total))
    return 0;
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  launch_step(b, 0, 0);
  launch_step(b, 1, 2);
  BCHK(b, hipStreamEndCapture(b->stream, &g));
  const hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->graph = nullptr;
    return bfail(b, -100, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  b->graph_fr = b->fr;
  b->graph_sched = b->sched;
  std::copy(b->total, b->total + C_N, b->graph_total);
  return 0;
}

// Steps of one run: counter cleared, nsteps steps, last samples; the shots'
// host state follows.
int run_steps(bq_batch *b, int nsteps, bool graph, bool timed) {
  // block counts follow the shots' receivers, which may be set after the
  // batch was made
  if (int rc = layout(b, b->fr)) return rc;
  if (int rc = upload_shots(b)) return rc;
  hipLaunchKernelGGL(kb_clear, dim3(1), dim3(64), 0, b->stream, b->dctl, 20);
  if (timed) BCHK(b, hipEventRecord(b->ev[0], b->stream));
  int done = 0;
  if (graph && nsteps >= 2) {
    if (int rc = build_graph(b)) return rc;
    for (; done + 2 <= nsteps; done += 2) BCHK(b, hipGraphLaunch(b->graph, b->stream));
  }
  for (; done < nsteps; ++done) launch_step(b, 0, 1);
  const BatchArgs a = args_of(b);
  if (a.total[C_SAMPLE])
    hipLaunchKernelGGL(kb_sample, dim3(a.total[C_SAMPLE]), dim3(kBlock), 0, b->stream, a, -1);
  if (timed) BCHK(b, hipEventRecord(b->ev[1], b->stream));
  BCHK(b, hipGetLastError());
  for (bq_handle *h : b->shots) {
    h->next_step += nsteps;
    h->parity ^= (nsteps & 1);
  }
  if (timed) b->last_steps = nsteps;
  return 0;
}

// Device time of the last timed run, ms (waits for it).
int run_ms(bq_batch *b, double *ms) {
  BCHK(b, hipEventSynchronize(b->ev[1]));
  float t = 0;
  BCHK(b, hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
  *ms = t;
  return 0;
}

int check_run(bq_batch *b, int nsteps) {
  if (nsteps < 0) return bfail(b, -2, "negative step count");
  for (size_t s = 0; s < b->shots.size(); ++s) {
    const bq_handle *h = b->shots[s];
    if (!h->model_ok || !h->pml_ok || !h->src_ok)
      return bfail(b, -6, "shot " + std::to_string(s) + ": model, PML and source must be set");
    if (h->next_step + nsteps - 1 > h->g.MT2)
      return bfail(b, -7, "shot " + std::to_string(s) + ": step beyond its MT2");
  }
  // work queued on the shots' own streams comes first
  for (bq_handle *h : b->shots) BCHK(b, hipStreamSynchronize(h->stream));
  return 0;
}

// Orders later work on the shots' own streams after the batch's.
int hand_back(bq_batch *b) {
  hipEvent_t e;
  BCHK(b, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  BCHK(b, hipEventRecord(e, b->stream));
  for (bq_handle *h : b->shots) BCHK(b, hipStreamWaitEvent(h->stream, e, 0));
  BCHK(b, hipEventDestroy(e));
  return 0;
}

}  // namespace

extern "C" {

int bq_batch_create(bq_handle *const *shots, int nshots, bq_batch **out) {
  if (!out) return -1;
  *out = nullptr;
  g_batch_create_err.clear();
  if (!shots || nshots < 1) {
    g_batch_create_err = "bq_batch_create: no shots";
    return -1;
  }
  bq_batch *b = new bq_batch();
  b->shots.assign(shots, shots + nshots);
  for (int s = 0; s < nshots; ++s) {
    if (!shots[s]) {
      g_batch_create_err = "bq_batch_create: null shot handle";
      delete b;
      return -1;
    }
    for (int r = 0; r < s; ++r)
      if (shots[r] == shots[s]) {
        g_batch_create_err = "bq_batch_create: a shot appears twice";
        delete b;
        return -1;
      }
  }
  b->device = shots[0]->device;
  for (int s = 1; s < nshots; ++s)
    if (shots[s]->device != b->device) {
      g_batch_create_err = "bq_batch_create: shots on different devices";
      delete b;
      return -1;
    }
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(b->device);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, b->device) == hipSuccess) b->cus = prop.multiProcessorCount;
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipMalloc(&b->dshots, sizeof(ShotDev) * nshots) == hipSuccess;
  ok = ok && hipMalloc(&b->dcum, sizeof(int) * C_N * (nshots + 1)) == hipSuccess;
  ok = ok && hipMalloc(&b->dctl, sizeof(int) * 20) == hipSuccess;
  ok = ok && hipMemset(b->dctl, 0, sizeof(int) * 20) == hipSuccess;
  ok = ok && hipEventCreate(&b->ev[0]) == hipSuccess && hipEventCreate(&b->ev[1]) == hipSuccess;
  b->host.assign(nshots, ShotDev{});
  if (ok) ok = layout(b, 128) == 0;
  (void)hipSetDevice(cur);
  if (!ok) {
    g_batch_create_err = "bq_batch_create: device allocation failed" + (b->err.empty() ? "" : ": " + b->err);
    bq_batch_destroy(b);
    return -101;
  }
  *out = b;
  return 0;
}

void bq_batch_destroy(bq_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  (void)hipFree(b->dshots);
  (void)hipFree(b->dcum);
  (void)hipFree(b->dctl);
  for (auto e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char *bq_batch_last_error(bq_batch *b) {
  if (b) return b->err.c_str();
  return g_batch_create_err.empty() ? "null batch" : g_batch_create_err.c_str();
}

int bq_batch_set_schedule(bq_batch *b, int tile_rows, int schedule, int use_graph) {
  if (!b) return -1;
  if (tile_rows != 64 && tile_rows != 128 && tile_rows != 256)
    return bfail(b, -2, "tile_rows must be 64, 128 or 256");
  if (schedule != BQ_SCHED_GRID && schedule != BQ_SCHED_PERSISTENT)
    return bfail(b, -2, "unknown schedule");
  b->sched = schedule;
  b->use_graph = use_graph != 0;
  if (tile_rows != b->fr) return layout(b, tile_rows);
  return 0;
}

int bq_batch_get_schedule(bq_batch *b, int *tile_rows, int *schedule, int *use_graph) {
  if (!b) return -1;
  if (tile_rows) *tile_rows = b->fr;
  if (schedule) *schedule = b->sched;
  if (use_graph) *use_graph = b->use_graph;
  return 0;
}

int bq_batch_run(bq_batch *b, int nsteps) {
  if (!b) return -1;
  if (int rc = check_run(b, nsteps)) return rc;
  if (nsteps == 0) return 0;
  if (int rc = run_steps(b, nsteps, b->use_graph != 0, true)) return rc;
  return hand_back(b);
}

int bq_batch_sync(bq_batch *b) {
  if (!b) return -1;
  BCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

int bq_batch_last_run_ms(bq_batch *b, double *ms) {
  if (!b || !ms) return -1;
  if (!b->last_steps) return bfail(b, -6, "no run yet");
  return run_ms(b, ms);
}

int bq_batch_autotune(bq_batch *b, int trial_steps, double *ms_per_step) {
  if (!b) return -1;
  if (trial_steps < 1) return bfail(b, -2, "trial_steps must be >= 1");
  static const int frs[3] = {64, 128, 256};
  static const int scheds[2] = {BQ_SCHED_GRID, BQ_SCHED_PERSISTENT};
  const int ncand = 6;
  if (int rc = check_run(b, ncand * trial_steps)) return rc;
  b->tune_ms.assign(ncand, 0.0);
  int best = -1;
  double best_ms = 0.0;
  for (int c = 0; c < ncand; ++c) {
    b->sched = scheds[c % 2];
    if (int rc = layout(b, frs[c / 2])) return rc;
    // the candidate's timed steps advance the shots like any other steps
    if (int rc = run_steps(b, trial_steps, false, true)) return rc;
    double run = 0.0;
    if (int rc = run_ms(b, &run)) return rc;
    const double ms = run / trial_steps;
    b->tune_ms[c] = ms;
    if (ms_per_step) ms_per_step[c] = ms;
    if (best < 0 || ms < best_ms) {
      best = c;
      best_ms = ms;
    }
  }
  b->sched = scheds[best % 2];
  if (int rc = layout(b, frs[best / 2])) return rc;
  return hand_back(b);
}

}  // extern "C"
