// bq_handle.h - the state one grid (one shot) holds on the device, shared by
// the single-grid entry points (bq_device.hip) and the batches (bq_batch.hip).
#pragma once
#include <string>
#include <vector>

#include "bq_fused2.h"
#include "bq_kernels.h"

struct bq_handle {
  GridDesc g{};
  int P = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  KParams kp{};
  std::vector<void *> allocs;
  float *fields[2] = {nullptr, nullptr};  // two full field sets, current and next
  int parity = 0;                         // which of the two is current
  // two steps per sweep: the region, the partition of the other cells, the
  // scratch set holding level n+1 of those cells and of the region's rim
  float *scratch_set = nullptr;
  bq::Fused2 f2{};
  Split ss2{}, sv2{};
  long long region_cells = 0, stress_cells2 = 0, velocity_cells2 = 0;
  int sweep = 1;            // steps per sweep allowed (bq_set_sweep)
  bool rec_sweep_ok = true; // every receiver node outside the region or in its rim
  int last_sweep = 1;       // steps per sweep of the last bq_run
  double ktime2[4] = {0, 0, 0, 0};  // two-step sweep, other stress, velocity, edges (ms)
  std::vector<int32_t> jm_host;
  std::vector<float> tab_host;
  bool model_ok = false, pml_ok = false, src_ok = false;
  int next_step = 1;
  int mr = 0;
  int timing = 0;  // 0: off; n: events around the launches of every n-th step
  std::vector<hipEvent_t> ev;
  double ktime[3] = {0, 0, 0};
  int timed_steps = 0;
  // the two partitions of a step: with fused tiles of fr_tiles rows (the
  // batches always take this one) and without them, every cell on the runs
  // and the per-cell path; tile_mode (bq_set_tiles) picks the one bq_run uses
  bq::Plan tiles{}, cells{};
  int fr_tiles = 128;
  int tile_mode = 0;
  bool tiles_auto = true;  // BQ_TILES_AUTO takes the fused tiles
  int z0 = 2;              // first row of both half-steps' cells (2 with stress imaging, else 0)
  // the partition bq_run uses
  Split ss{}, sv{};
  Fused fu{};
  int fr = 128;  // rows per fused block of the single-grid launches
  long long fused_cells = 0, stress_cells = 0, velocity_cells = 0;
  // learned absorbing boundary (bq_lb_attach): the network, and the band of
  // the last four time levels, level n in slot n % 4
  const void *lb = nullptr;
  float *lb_ring = nullptr;
  std::string err;
};

inline int fail_handle(bq_handle *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  return code;
}

namespace bq {
// Strip cells from the attached network after step `it`, whose parameters
// are k (bq_learned.hip).
int lb_after_step(bq_handle *h, const KParams &k, int it);
}  // namespace bq

// Kernel parameters of a step of the given parity (steps done, mod 2): after
// an odd number of steps the current and next field sets trade places.
inline KParams params_for(const bq_handle *h, int parity) {
  KParams k = h->kp;
  if (parity) {
    const size_t fsz = (size_t)h->P * h->g.NZ;
    const float **f0[8] = {&k.U0, &k.W0, &k.QX0, &k.QZ0, &k.TXX0, &k.TZZ0, &k.TXZ0, &k.PR0};
    float **f1[8] = {&k.U1, &k.W1, &k.QX1, &k.QZ1, &k.TXX1, &k.TZZ1, &k.TXZ1, &k.PR1};
    for (int f = 0; f < 8; ++f) {
      *f0[f] = h->fields[1] + f * fsz;
      *f1[f] = h->fields[0] + f * fsz;
    }
  }
  return k;
}

inline KParams step_params(const bq_handle *h) { return params_for(h, h->parity); }

// Kernel parameters reading set src and writing set dst (8 fields each).
inline KParams params_sets(const bq_handle *h, const float *src, float *dst) {
  KParams k = h->kp;
  const size_t fsz = (size_t)h->P * h->g.NZ;
  const float **f0[8] = {&k.U0, &k.W0, &k.QX0, &k.QZ0, &k.TXX0, &k.TZZ0, &k.TXZ0, &k.PR0};
  float **f1[8] = {&k.U1, &k.W1, &k.QX1, &k.QZ1, &k.TXX1, &k.TZZ1, &k.TXZ1, &k.PR1};
  for (int f = 0; f < 8; ++f) {
    *f0[f] = src + f * fsz;
    *f1[f] = dst + f * fsz;
  }
  return k;
}
