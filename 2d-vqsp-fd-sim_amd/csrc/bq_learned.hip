// bq_learned.hip - the learned absorbing boundary: band extraction, BF16 MFMA
// inference after each step, and the FP32 training kernels (forward, backward,
// Adam) that fit the networks to bands recorded from wide-PML runs.
//
// The reference absorbs with a CPML strip of TPML cells (ppml.F90:159-306). A
// grid built with TPML = 8 and a network attached (bq_lb_attach) has its 8
// strip cells per side overwritten after every step with the network's
// prediction from the last four levels of the 12-cell band; without the
// network the same grid is a plain 8-cell CPML. The network design and the
// band layout are described in bq_learned.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/biotfd.h"
#include "../../include/biotfd_learned.h"
#include "bq_handle.h"
#include "bq_learned.h"

namespace {

constexpr int kNW1 = kLbTap * kLbCin * kLbH;  // 122880
constexpr int kNW2 = kLbTap * kLbH * kLbH;    // 20480
constexpr int kNSide = kNW1 + 2 * kNW2;       // weights of one side's network
constexpr int kTB = 256;

struct F8 {
  const float *f[8];
};

F8 fields_of(const KParams &k, bool next) {
  F8 r;
  if (next) {
    const float *p[8] = {k.U1, k.W1, k.QX1, k.QZ1, k.TXX1, k.TZZ1, k.TXZ1, k.PR1};
    for (int i = 0; i < 8; ++i) r.f[i] = p[i];
  } else {
    const float *p[8] = {k.U0, k.W0, k.QX0, k.QZ0, k.TXX0, k.TZZ0, k.TXZ0, k.PR0};
    for (int i = 0; i < 8; ++i) r.f[i] = p[i];
  }
  return r;
}

__global__ __launch_bounds__(kTB) void k_lb_extract(F8 F, int P, int OFF, int MX, int MZ,
                                                    float *__restrict__ out, size_t n, int sub) {
  for (size_t e = (size_t)blockIdx.x * kTB + threadIdx.x; e < n; e += (size_t)gridDim.x * kTB)
    lb_extract(out, F.f, P, OFF, MX, MZ, e, sub);
}

// The strip cells (band cells 0..7 of the three sides, all fields) of grid
// a's fields from grid b's, at the same physical cells.
__global__ __launch_bounds__(kTB) void k_strip_copy(F8 A, int pa, int offa, F8 B, int pb, int offb,
                                                    int MX, int MZ) {
  const size_t lz = lb_lz(MZ), lx = lb_lx(MX);
  const size_t n = (2 * lz + lx) * kLbF * kLbS;
  for (size_t e = (size_t)blockIdx.x * kTB + threadIdx.x; e < n; e += (size_t)gridDim.x * kTB) {
    const size_t pos_all = e / (kLbF * kLbS);
    const int fc = (int)(e % (kLbF * kLbS)), f = fc / kLbS, c = fc % kLbS;
    const int side = pos_all < lz ? 0 : (pos_all < 2 * lz ? 1 : 2);
    const int pos = (int)(pos_all - (side == 0 ? 0 : side == 1 ? lz : 2 * lz));
    int i, l;
    lb_cell(side, pos, c, MX, MZ, &i, &l);
    ((float *)A.f[f])[(size_t)l * pa + (size_t)(i + offa)] = B.f[f][(size_t)l * pb + (size_t)(i + offb)];
  }
}


// After a step: strip cells of all three sides from the network; block b is a
// tile of 52 positions of one side.
__global__ __launch_bounds__(kLbBlock) void k_lb_apply(LbNet net, KParams k, LbLevels lev) {
  __shared__ LbLds S;
  const int MX = k.MX, MZ = k.MZ;
  const int t0 = lb_tiles(lb_lz(MZ)), t1 = 2 * t0;
  const int b = blockIdx.x;
  const int side = b < t0 ? 0 : (b < t1 ? 1 : 2);
  const int tile = b - (side == 0 ? 0 : side == 1 ? t0 : t1);
  const int L = lb_side_len(side, MX, MZ);
  lb_tile(net, side, L, lb_side_off(side, MZ), tile * kLbP, lev, S,
          [&](int p, int f, int c, float prev, float d) {
            int i, l;
            lb_cell(side, p, c, MX, MZ, &i, &l);
            float *F = f == 0 ? k.U1 : f == 1 ? k.W1 : f == 2 ? k.QX1 : f == 3 ? k.QZ1
                     : f == 4 ? k.TXX1 : f == 5 ? k.TZZ1 : f == 6 ? k.TXZ1 : k.PR1;
            float &x = F[(size_t)l * k.P + (size_t)(i + k.OFF)];
            // the CPML's update of this step is in x
            x = lb_update(net.form, x, prev, d);
          });
}

// The network on given bands of one side (length L, levels [4][L][96]):
// out[p][f][c] = the new strip values.
__global__ __launch_bounds__(kLbBlock) void k_lb_infer(LbNet net, int side, int L, LbLevels lev,
                                                       float *__restrict__ out) {
  __shared__ LbLds S;
  lb_tile(net, side, L, 0, blockIdx.x * kLbP, lev, S, [&](int p, int f, int c, float prev, float d) {
    out[((size_t)p * kLbF + f) * kLbS + c] = prev + d;
  });
}

// ------------------------------------------------------------- training
// Activations are [B][L][C] (channels fastest); weights [tap][in][out] (out
// fastest), FP32 throughout.

// Minibatch of one side from the recorded groups: x[b][p][lev*96 + fc] =
// band level lev / scale, y[b][p][f*8 + c] = (strip at level 4 - level 3) /
// scale, or with residual strip at level 4 / scale.
__global__ __launch_bounds__(kTB) void k_tr_gather(const float *__restrict__ store, size_t gsize,
                                                   size_t bsize, size_t soff, const int *__restrict__ groups,
                                                   int B, int L, const float *__restrict__ inv, int residual,
                                                   float *__restrict__ x, float *__restrict__ y) {
  const size_t nx = (size_t)B * L * kLbCin, ny = (size_t)B * L * kLbH;
  for (size_t e = (size_t)blockIdx.x * kTB + threadIdx.x; e < nx + ny; e += (size_t)gridDim.x * kTB) {
    if (e < nx) {
      const int ci = (int)(e % kLbCin);
      const size_t bp = e / kLbCin;
      const int b = (int)(bp / L), p = (int)(bp % L);
      const int lv = ci / kLbBand, fc = ci % kLbBand;
      const float *g = store + (size_t)groups[b] * gsize;
      x[e] = g[(size_t)lv * bsize + soff + (size_t)p * kLbBand + fc] * inv[fc / kLbCells];
    } else {
      const size_t q = e - nx;
      const int co = (int)(q % kLbH);
      const size_t bp = q / kLbH;
      const int b = (int)(bp / L), p = (int)(bp % L);
      const int f = co >> 3, c = co & 7;
      const float *g = store + (size_t)groups[b] * gsize + soff + (size_t)p * kLbBand + f * kLbCells + c;
      // residual: level 4 already holds the target (wide run - CPML update)
      y[q] = (residual ? g[4 * bsize] : g[4 * bsize] - g[3 * bsize]) * inv[f];
    }
  }
}

// ypre[b][p][co] = sum_t sum_ci w[t][ci][co] x[b][p+t-2][ci]; a = lrelu(ypre)
// (or a copy when act is 0).
__global__ __launch_bounds__(kTB) void k_tr_fwd(const float *__restrict__ x, const float *__restrict__ w,
                                                float *__restrict__ ypre, float *__restrict__ a, int B,
                                                int L, int cin, int act) {
  const size_t e = (size_t)blockIdx.x * kTB + threadIdx.x;
  if (e >= (size_t)B * L * kLbH) return;
  const int co = (int)(e % kLbH);
  const size_t bp = e / kLbH;
  const int b = (int)(bp / L), p = (int)(bp % L);
  float s = 0.f;
  for (int t = 0; t < kLbTap; ++t) {
    const int q = p + t - 2;
    if (q < 0 || q >= L) continue;
    const float *xr = x + ((size_t)b * L + q) * cin;
    const float *wr = w + (size_t)t * cin * kLbH + co;
    for (int ci = 0; ci < cin; ++ci) s += wr[(size_t)ci * kLbH] * xr[ci];
  }
  ypre[e] = s;
  a[e] = act ? (s > 0.f ? s : 0.1f * s) : s;
}

// dout = 2 (out - y) / n; part[block] = sum (out - y)^2 of the block. Only
// the output channels of the `inner` strip cells nearest the interior count.
__global__ __launch_bounds__(kTB) void k_tr_loss(const float *__restrict__ out, const float *__restrict__ y,
                                                 float *__restrict__ dout, size_t n, int inner,
                                                 double *__restrict__ part) {
  __shared__ double sh[kTB];
  const size_t e = (size_t)blockIdx.x * kTB + threadIdx.x;
  double d2 = 0.0;
  if (e < n) {
    const float d = (int)(e & 7) >= kLbS - inner ? out[e] - y[e] : 0.f;
    dout[e] = 2.0f * d / (float)n;
    d2 = (double)d * d;
  }
  sh[threadIdx.x] = d2;
  __syncthreads();
  for (int w = kTB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Weight gradient, split over nsplit chunks of the (b, p) sum:
// dwp[s][t][ci][co] = sum over chunk s of dy[b][p][co] x[b][p+t-2][ci].
__global__ __launch_bounds__(kTB) void k_tr_bwd_w(const float *__restrict__ x, const float *__restrict__ dy,
                                                  float *__restrict__ dwp, int B, int L, int cin, int nsplit) {
  const int nw = kLbTap * cin * kLbH;
  const int e = blockIdx.x * kTB + threadIdx.x;
  if (e >= nw) return;
  const int s = blockIdx.y;
  const int co = e % kLbH, ci = (e / kLbH) % cin, t = e / (kLbH * cin);
  const long long nbp = (long long)B * L;
  const long long b0 = nbp * s / nsplit, b1 = nbp * (s + 1) / nsplit;
  float acc = 0.f;
  for (long long bp = b0; bp < b1; ++bp) {
    const int b = (int)(bp / L), p = (int)(bp % L);
    const int q = p + t - 2;
    if (q < 0 || q >= L) continue;
    acc += dy[(size_t)bp * kLbH + co] * x[((size_t)b * L + q) * cin + ci];
  }
  dwp[(size_t)s * nw + e] = acc;
}

__global__ __launch_bounds__(kTB) void k_tr_sum_split(const float *__restrict__ dwp, float *__restrict__ dw,
                                                      int nw, int nsplit) {
  const int e = blockIdx.x * kTB + threadIdx.x;
  if (e >= nw) return;
  float s = 0.f;
  for (int k = 0; k < nsplit; ++k) s += dwp[(size_t)k * nw + e];
  dw[e] = s;
}

// Input gradient through a layer, times lrelu' of the previous layer's
// pre-activation: dx[b][q][ci] = lrelu'(xpre[b][q][ci]) *
// sum_t sum_co w[t][ci][co] dy[b][q-t+2][co].
__global__ __launch_bounds__(kTB) void k_tr_bwd_x(const float *__restrict__ dy, const float *__restrict__ w,
                                                  const float *__restrict__ xpre, float *__restrict__ dx,
                                                  int B, int L, int cin) {
  const size_t e = (size_t)blockIdx.x * kTB + threadIdx.x;
  if (e >= (size_t)B * L * cin) return;
  const int ci = (int)(e % cin);
  const size_t bq = e / cin;
  const int b = (int)(bq / L), q = (int)(bq % L);
  float s = 0.f;
  for (int t = 0; t < kLbTap; ++t) {
    const int p = q - t + 2;
    if (p < 0 || p >= L) continue;
    const float *dr = dy + ((size_t)b * L + p) * kLbH;
    const float *wr = w + ((size_t)t * cin + ci) * kLbH;
    for (int co = 0; co < kLbH; ++co) s += wr[co] * dr[co];
  }
  dx[e] = xpre[e] > 0.f ? s : 0.1f * s;
}

__global__ __launch_bounds__(kTB) void k_tr_adam(float *__restrict__ w, const float *__restrict__ g,
                                                 float *__restrict__ m, float *__restrict__ v, int n,
                                                 float lr, float c1, float c2) {
  const int e = blockIdx.x * kTB + threadIdx.x;
  if (e >= n) return;
  const float b1 = 0.9f, b2 = 0.999f;
  const float mm = b1 * m[e] + (1.f - b1) * g[e];
  const float vv = b2 * v[e] + (1.f - b2) * g[e] * g[e];
  m[e] = mm;
  v[e] = vv;
  w[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
}

// Sum of squares per field of one side over levels 0..3 of groups
// [g0, g1): part[block][8].
__global__ __launch_bounds__(kTB) void k_tr_rms(const float *__restrict__ store, size_t gsize,
                                                size_t bsize, size_t soff, int L, int g0, int g1,
                                                double *__restrict__ part) {
  __shared__ double sh[kLbF][kTB];
  double acc[kLbF] = {0, 0, 0, 0, 0, 0, 0, 0};
  const size_t per = (size_t)4 * L * kLbBand;
  const size_t n = per * (size_t)(g1 - g0);
  for (size_t e = (size_t)blockIdx.x * kTB + threadIdx.x; e < n; e += (size_t)gridDim.x * kTB) {
    const size_t g = g0 + e / per, r = e % per;
    const size_t lv = r / ((size_t)L * kLbBand), q = r % ((size_t)L * kLbBand);
    const float v = store[g * gsize + lv * bsize + soff + q];
    acc[(q % kLbBand) / kLbCells] += (double)v * v;
  }
  for (int f = 0; f < kLbF; ++f) sh[f][threadIdx.x] = acc[f];
  __syncthreads();
  for (int w = kTB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int f = 0; f < kLbF; ++f) sh[f][threadIdx.x] += sh[f][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int f = 0; f < kLbF; ++f) part[(size_t)blockIdx.x * kLbF + f] = sh[f][0];
}

// Misfit of one field between two grids over their common physical cells
// x = 1..MX, rows 0..MZ' (the grids may differ in PML width):
// part[block] = {sum (a-b)^2, sum b^2}.
__global__ __launch_bounds__(kTB) void k_field_misfit(const float *__restrict__ a, int pa, int offa,
                                                      const float *__restrict__ b, int pb, int offb,
                                                      int MX, int MZ, double *__restrict__ part) {
  __shared__ double sd[kTB], sb[kTB];
  const size_t n = (size_t)MX * (MZ + 1);
  double d2 = 0.0, b2 = 0.0;
  for (size_t e = (size_t)blockIdx.x * kTB + threadIdx.x; e < n; e += (size_t)gridDim.x * kTB) {
    const int i = 1 + (int)(e % MX), l = (int)(e / MX);
    const double x = a[(size_t)l * pa + i + offa], y = b[(size_t)l * pb + i + offb];
    d2 += (x - y) * (x - y);
    b2 += y * y;
  }
  sd[threadIdx.x] = d2;
  sb[threadIdx.x] = b2;
  __syncthreads();
  for (int w = kTB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sd[threadIdx.x] += sd[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sd[0];
    part[2 * blockIdx.x + 1] = sb[0];
  }
}

unsigned short host_bf16(float x) {
  unsigned int u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

}  // namespace

// ================================================================== host

struct bq_lb {
  int device = 0;
  unsigned short *w = nullptr;  // bf16: 3 sides x (W1, W2, W3), [tap][out][in]
  LbNet net{};
  std::string err;
};

struct bq_lb_trainer {
  int device = 0;
  hipStream_t stream = nullptr;
  int MX = 0, MZ = 0, max_groups = 0, batch = 0, maxL = 0;
  size_t bsize = 0, gsize = 0;
  float *store = nullptr;                   // [group][5][band]
  float *w[3] = {}, *m[3] = {}, *v[3] = {};  // per side: W1 W2 W3 packed (kNSide)
  float *gw = nullptr, *gpart = nullptr;     // gradient, split partials
  float *x = nullptr, *y = nullptr;          // minibatch input [B][L][384], target [B][L][64]
  float *y1 = nullptr, *a1 = nullptr, *y2 = nullptr, *a2 = nullptr, *y3 = nullptr, *o3 = nullptr;
  float *d3 = nullptr, *d2 = nullptr, *d1 = nullptr;
  float *inv = nullptr;                     // [3][8] on the device
  int *groups = nullptr;
  double *part = nullptr;
  int nparts = 0;
  int adam_t[3] = {0, 0, 0};
  int residual = 0;
  int inner = kLbS;
  float scale[3][kLbF];
  std::string err;
};

namespace bq {
const LbNet &lb_net_of(const void *lb) { return ((const bq_lb *)lb)->net; }

// After step `it` of a grid with a network attached: strip cells from the
// network (levels it-4 .. it-1 in the ring), then the band of level it into
// the ring. k: the step's parameters (the step wrote the "next" set).
int lb_after_step(bq_handle *h, const KParams &k, int it) {
  const bq_lb *lb = (const bq_lb *)h->lb;
  const size_t bs = lb_band_size(k.MX, k.MZ);
  const LbLevels lev{h->lb_ring + (size_t)((it + 0) & 3) * bs, h->lb_ring + (size_t)((it + 1) & 3) * bs,
                     h->lb_ring + (size_t)((it + 2) & 3) * bs, h->lb_ring + (size_t)((it + 3) & 3) * bs};
  const int tiles = 2 * lb_tiles(lb_lz(k.MZ)) + lb_tiles(lb_lx(k.MX));
  hipLaunchKernelGGL(k_lb_apply, dim3(tiles), dim3(kLbBlock), 0, h->stream, lb->net, k, lev);
  const int blocks = (int)std::min<size_t>((bs + kTB - 1) / kTB, 4096);
  hipLaunchKernelGGL(k_lb_extract, dim3(blocks), dim3(kTB), 0, h->stream, fields_of(k, true), k.P,
                     k.OFF, k.MX, k.MZ, h->lb_ring + (size_t)(it & 3) * bs, bs, 0);
  return hipGetLastError() == hipSuccess ? 0 : -100;
}
}  // namespace bq

namespace {

int tfail(bq_lb_trainer *t, int code, const std::string &msg) {
  if (t) t->err = msg;
  return code;
}

template <class Tp>
bool talloc(Tp **p, size_t n) {
  return hipMalloc((void **)p, (n ? n : 1) * sizeof(Tp)) == hipSuccess &&
         hipMemset(*p, 0, (n ? n : 1) * sizeof(Tp)) == hipSuccess;
}

thread_local std::string g_lb_err;

int grid_of(size_t n) { return (int)std::min<size_t>((n + kTB - 1) / kTB, 1 << 20); }

}  // namespace

extern "C" {

int bq_lb_create(const float *weights, const float *scales, int form, float gain, int inner,
                 bq_lb **out) {
  if (!out || !weights || !scales || inner < 1 || inner > kLbS || form < kLbDirect || form > kLbDamping)
    return -1;
  if (!(gain >= 0.f && gain <= 16.f)) {
    g_lb_err = "bq_lb_create: gain outside [0, 16]";
    return -2;
  }
  *out = nullptr;
  bq_lb *lb = new bq_lb();
  (void)hipGetDevice(&lb->device);
  std::vector<unsigned short> h((size_t)3 * kNSide);
  // [tap][in][out] FP32 -> BF16 in the fragment order of lb_layer: per tap,
  // output block v, input block kc, lane q the four values
  // W[t][16v + q%16][16kc + 4(q/16) .. +3]
  for (int s = 0; s < 3; ++s) {
    const float *src = weights + (size_t)s * kNSide;
    unsigned short *dst = h.data() + (size_t)s * kNSide;
    const int cins[3] = {kLbCin, kLbH, kLbH};
    size_t o = 0;
    for (int layer = 0; layer < 3; ++layer) {
      const int cin = cins[layer], nk = cin / 16;
      for (int t = 0; t < kLbTap; ++t)
        for (int v = 0; v < kLbH / 16; ++v)
          for (int kc = 0; kc < nk; ++kc)
            for (int q = 0; q < 64; ++q)
              for (int j = 0; j < 4; ++j) {
                const int co = 16 * v + (q & 15), ci = 16 * kc + 4 * (q >> 4) + j;
                const size_t at = ((((size_t)t * 4 + v) * nk + kc) * 64 + q) * 4 + j;
                dst[o + at] = host_bf16(src[o + ((size_t)t * cin + ci) * kLbH + co]);
              }
      o += (size_t)kLbTap * cin * kLbH;
    }
  }
  if (hipMalloc(&lb->w, h.size() * 2) != hipSuccess ||
      hipMemcpy(lb->w, h.data(), h.size() * 2, hipMemcpyHostToDevice) != hipSuccess) {
    g_lb_err = "bq_lb_create: device allocation failed";
    bq_lb_destroy(lb);
    return -101;
  }
  lb->net.form = form;
  lb->net.gain = gain;
  lb->net.inner = inner;
  for (int s = 0; s < 3; ++s) {
    const unsigned short *base = lb->w + (size_t)s * kNSide;
    lb->net.w1[s] = base;
    lb->net.w2[s] = base + kNW1;
    lb->net.w3[s] = base + kNW1 + kNW2;
    for (int f = 0; f < kLbF; ++f) {
      const float sc = scales[s * kLbF + f];
      if (!(sc > 0.f) || !std::isfinite(sc)) {
        g_lb_err = "bq_lb_create: scales must be positive and finite";
        bq_lb_destroy(lb);
        return -2;
      }
      lb->net.scale[s][f] = sc;
      lb->net.inv[s][f] = 1.0f / sc;
    }
  }
  *out = lb;
  return 0;
}

void bq_lb_destroy(bq_lb *lb) {
  if (!lb) return;
  (void)hipFree(lb->w);
  delete lb;
}

const char *bq_lb_last_error(void) { return g_lb_err.c_str(); }

int bq_lb_attach(bq_handle *h, bq_lb *lb) {
  if (!h || !lb) return -1;
  if (h->g.T != kLbS) return fail_handle(h, -2, "bq_lb_attach: the grid's TPML must be 8 (the strip)");
  if (h->next_step != 1) return fail_handle(h, -6, "bq_lb_attach: before the first step");
  if (h->device != lb->device) return fail_handle(h, -2, "bq_lb_attach: network on another device");
  const size_t bs = lb_band_size(h->g.MX, h->g.MZ);
  if (!h->lb_ring) {
    float *r = nullptr;
    if (hipMalloc(&r, 4 * bs * sizeof(float)) != hipSuccess)
      return fail_handle(h, -101, "bq_lb_attach: ring allocation failed");
    h->allocs.push_back(r);
    h->lb_ring = r;
  }
  if (hipMemsetAsync(h->lb_ring, 0, 4 * bs * sizeof(float), h->stream) != hipSuccess)
    return fail_handle(h, -100, "bq_lb_attach: clearing the ring failed");
  // level 0 (the fields before the first step) into slot 0
  const KParams k = step_params(h);
  hipLaunchKernelGGL(k_lb_extract, dim3(grid_of(bs)), dim3(kTB), 0, h->stream, fields_of(k, false),
                     k.P, k.OFF, k.MX, k.MZ, h->lb_ring, bs, 0);
  if (hipStreamSynchronize(h->stream) != hipSuccess)
    return fail_handle(h, -100, "bq_lb_attach: band extraction failed");
  h->lb = lb;
  return 0;
}

int bq_lb_detach(bq_handle *h) {
  if (!h) return -1;
  h->lb = nullptr;
  return 0;
}

int bq_band_size(int mx, int mzp) { return mx < 8 || mzp < 6 ? -1 : (int)lb_band_size(mx, mzp); }

static int band_extract(bq_handle *h, float *dst, int sub) {
  if (!h || !dst) return -1;
  if (h->g.T < kLbS) return fail_handle(h, -2, "bq_band_extract: the grid's TPML must be >= 8");
  const KParams k = step_params(h);  // the current set holds the latest fields
  const size_t bs = lb_band_size(k.MX, k.MZ);
  hipLaunchKernelGGL(k_lb_extract, dim3(grid_of(bs)), dim3(kTB), 0, h->stream, fields_of(k, false),
                     k.P, k.OFF, k.MX, k.MZ, dst, bs, sub);
  if (hipGetLastError() != hipSuccess) return fail_handle(h, -100, "bq_band_extract: launch failed");
  return 0;
}

int bq_band_extract(bq_handle *h, float *dst) { return band_extract(h, dst, 0); }

int bq_strip_copy(bq_handle *dst, bq_handle *src) {
  if (!dst || !src) return -1;
  if (dst->g.MX != src->g.MX || dst->g.MZ != src->g.MZ)
    return fail_handle(dst, -2, "bq_strip_copy: grids of different physical size");
  if (dst->g.T < kLbS || src->g.T < kLbS)
    return fail_handle(dst, -2, "bq_strip_copy: TPML must be >= 8");
  if (hipStreamSynchronize(src->stream) != hipSuccess) return fail_handle(dst, -100, "sync failed");
  const KParams kd = step_params(dst), ks = step_params(src);
  const size_t n = (2 * (size_t)lb_lz(kd.MZ) + lb_lx(kd.MX)) * kLbF * kLbS;
  hipLaunchKernelGGL(k_strip_copy, dim3(grid_of(n)), dim3(kTB), 0, dst->stream, fields_of(kd, false),
                     kd.P, kd.OFF, fields_of(ks, false), ks.P, ks.OFF, kd.MX, kd.MZ);
  if (hipGetLastError() != hipSuccess) return fail_handle(dst, -100, "bq_strip_copy: launch failed");
  return 0;
}

int bq_lb_infer(bq_lb *lb, int side, int len, const float *bands, float *out) {
  if (!lb || !bands || !out || side < 0 || side > 2 || len < 1) return -1;
  const size_t n = (size_t)len * kLbBand;
  float *d = nullptr, *o = nullptr;
  if (hipMalloc(&d, 4 * n * 4) != hipSuccess) return -101;
  if (hipMalloc(&o, (size_t)len * kLbF * kLbS * 4) != hipSuccess) {
    (void)hipFree(d);
    return -101;
  }
  int rc = hipMemcpy(d, bands, 4 * n * 4, hipMemcpyHostToDevice) == hipSuccess ? 0 : -100;
  if (!rc) {
    const LbLevels lev{d, d + n, d + 2 * n, d + 3 * n};
    hipLaunchKernelGGL(k_lb_infer, dim3(lb_tiles(len)), dim3(kLbBlock), 0, 0, lb->net, side, len, lev, o);
    if (hipDeviceSynchronize() != hipSuccess) rc = -100;
  }
  if (!rc && hipMemcpy(out, o, (size_t)len * kLbF * kLbS * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = -100;
  (void)hipFree(d);
  (void)hipFree(o);
  return rc;
}

int bq_field_misfit(bq_handle *a, bq_handle *b, int field_id, double out[2]) {
  if (!a || !b || !out || field_id < 0 || field_id > 7) return -1;
  if (a->g.MX != b->g.MX || a->g.MZ != b->g.MZ)
    return fail_handle(a, -2, "bq_field_misfit: grids of different physical size");
  if (hipStreamSynchronize(b->stream) != hipSuccess) return fail_handle(a, -100, "sync failed");
  const KParams ka = step_params(a), kb = step_params(b);
  const float *fa = fields_of(ka, false).f[field_id], *fb = fields_of(kb, false).f[field_id];
  const int blocks = 512;
  double *part = nullptr;
  if (hipMalloc(&part, 2 * blocks * sizeof(double)) != hipSuccess)
    return fail_handle(a, -101, "bq_field_misfit: allocation failed");
  hipLaunchKernelGGL(k_field_misfit, dim3(blocks), dim3(kTB), 0, a->stream, fa, ka.P, ka.OFF, fb,
                     kb.P, kb.OFF, ka.MX, ka.MZ, part);
  std::vector<double> h(2 * blocks);
  int rc = hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, a->stream) == hipSuccess &&
                   hipStreamSynchronize(a->stream) == hipSuccess
               ? 0
               : fail_handle(a, -100, "bq_field_misfit: copy failed");
  (void)hipFree(part);
  out[0] = out[1] = 0.0;
  for (int k = 0; k < blocks; ++k) {
    out[0] += h[2 * k];
    out[1] += h[2 * k + 1];
  }
  return rc;
}

// ---------------------------------------------------------------- trainer

int bq_lb_trainer_create(int mx, int mzp, int max_groups, int batch, int residual, int inner,
                         bq_lb_trainer **out) {
  if (!out || mx < 8 || mzp < 6 || max_groups < 1 || batch < 1 || inner < 1 || inner > kLbS)
    return -1;
  *out = nullptr;
  bq_lb_trainer *t = new bq_lb_trainer();
  t->residual = residual != 0;
  t->inner = inner;
  (void)hipGetDevice(&t->device);
  t->MX = mx;
  t->MZ = mzp;
  t->max_groups = max_groups;
  t->batch = batch;
  t->maxL = std::max(lb_lz(mzp), lb_lx(mx));
  t->bsize = lb_band_size(mx, mzp);
  t->gsize = 5 * t->bsize;
  const size_t nbl = (size_t)batch * t->maxL;
  bool ok = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && talloc(&t->store, t->gsize * max_groups);
  for (int s = 0; s < 3; ++s) ok = ok && talloc(&t->w[s], kNSide) && talloc(&t->m[s], kNSide) && talloc(&t->v[s], kNSide);
  ok = ok && talloc(&t->gw, kNSide) && talloc(&t->gpart, (size_t)8 * kNW1);
  ok = ok && talloc(&t->x, nbl * kLbCin) && talloc(&t->y, nbl * kLbH);
  float **acts[9] = {&t->y1, &t->a1, &t->y2, &t->a2, &t->y3, &t->o3, &t->d3, &t->d2, &t->d1};
  for (auto p : acts) ok = ok && talloc(p, nbl * kLbH);
  t->nparts = grid_of(nbl * kLbH);
  ok = ok && talloc(&t->inv, 3 * kLbF) && talloc(&t->groups, batch) &&
       talloc(&t->part, (size_t)std::max(t->nparts, 1024) * kLbF);
  if (!ok) {
    g_lb_err = "bq_lb_trainer_create: device allocation failed";
    bq_lb_trainer_destroy(t);
    return -101;
  }
  for (int s = 0; s < 3; ++s)
    for (int f = 0; f < kLbF; ++f) t->scale[s][f] = 1.0f;
  *out = t;
  return 0;
}

void bq_lb_trainer_destroy(bq_lb_trainer *t) {
  if (!t) return;
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  void *ps[] = {t->store, t->gw, t->gpart, t->x, t->y, t->y1, t->a1, t->y2, t->a2, t->y3, t->o3,
                t->d3, t->d2, t->d1, t->inv, t->groups, t->part};
  for (void *p : ps) (void)hipFree(p);
  for (int s = 0; s < 3; ++s) {
    (void)hipFree(t->w[s]);
    (void)hipFree(t->m[s]);
    (void)hipFree(t->v[s]);
  }
  if (t->stream) (void)hipStreamDestroy(t->stream);
  delete t;
}

const char *bq_lb_trainer_last_error(bq_lb_trainer *t) { return t ? t->err.c_str() : g_lb_err.c_str(); }

int bq_lb_trainer_capture(bq_lb_trainer *t, bq_handle *h, int group, int level) {
  if (!t || !h || group < 0 || group >= t->max_groups || level < 0 || level > 4) return -1;
  if (h->g.MX != t->MX || h->g.MZ != t->MZ)
    return tfail(t, -2, "bq_lb_trainer_capture: grid of another physical size");
  if (int rc = bq_band_extract(h, t->store + (size_t)group * t->gsize + (size_t)level * t->bsize))
    return tfail(t, rc, h->err);
  return 0;
}

int bq_lb_trainer_capture_residual(bq_lb_trainer *t, bq_handle *wide, bq_handle *trunc, int group) {
  if (!t || !wide || !trunc || group < 0 || group >= t->max_groups) return -1;
  if (wide->g.MX != t->MX || wide->g.MZ != t->MZ || trunc->g.MX != t->MX || trunc->g.MZ != t->MZ)
    return tfail(t, -2, "bq_lb_trainer_capture_residual: grid of another physical size");
  float *dst = t->store + (size_t)group * t->gsize + 4 * t->bsize;
  if (int rc = band_extract(wide, dst, 0)) return tfail(t, rc, wide->err);
  // the truncated grid's stream must see the wide band written
  if (hipStreamSynchronize(wide->stream) != hipSuccess) return tfail(t, -100, "sync failed");
  if (int rc = band_extract(trunc, dst, 1)) return tfail(t, rc, trunc->err);
  return 0;
}

int bq_lb_trainer_sync(bq_lb_trainer *t) {
  if (!t) return -1;
  return hipDeviceSynchronize() == hipSuccess ? 0 : tfail(t, -100, "device synchronize failed");
}

int bq_lb_trainer_scales(bq_lb_trainer *t, int ngroups, float *scales) {
  if (!t || !scales || ngroups < 1 || ngroups > t->max_groups) return -1;
  if (hipDeviceSynchronize() != hipSuccess) return tfail(t, -100, "synchronize failed");
  const int blocks = 1024;
  for (int s = 0; s < 3; ++s) {
    const int L = lb_side_len(s, t->MX, t->MZ);
    hipLaunchKernelGGL(k_tr_rms, dim3(blocks), dim3(kTB), 0, t->stream, t->store, t->gsize, t->bsize,
                       lb_side_off(s, t->MZ), L, 0, ngroups, t->part);
    std::vector<double> h((size_t)blocks * kLbF);
    if (hipMemcpyAsync(h.data(), t->part, h.size() * 8, hipMemcpyDeviceToHost, t->stream) != hipSuccess ||
        hipStreamSynchronize(t->stream) != hipSuccess)
      return tfail(t, -100, "bq_lb_trainer_scales: copy failed");
    const double cnt = 4.0 * L * kLbCells * ngroups;
    for (int f = 0; f < kLbF; ++f) {
      double acc = 0.0;
      for (int b = 0; b < blocks; ++b) acc += h[(size_t)b * kLbF + f];
      const double rms = std::sqrt(acc / cnt);
      t->scale[s][f] = rms > 0.0 ? (float)rms : 1.0f;
      scales[s * kLbF + f] = t->scale[s][f];
    }
  }
  float inv[3 * kLbF];
  for (int s = 0; s < 3; ++s)
    for (int f = 0; f < kLbF; ++f) inv[s * kLbF + f] = 1.0f / t->scale[s][f];
  if (hipMemcpy(t->inv, inv, sizeof inv, hipMemcpyHostToDevice) != hipSuccess)
    return tfail(t, -100, "bq_lb_trainer_scales: upload failed");
  return 0;
}

int bq_lb_trainer_set_weights(bq_lb_trainer *t, const float *weights) {
  if (!t || !weights) return -1;
  for (int s = 0; s < 3; ++s) {
    if (hipMemcpy(t->w[s], weights + (size_t)s * kNSide, kNSide * 4, hipMemcpyHostToDevice) != hipSuccess)
      return tfail(t, -100, "set_weights: copy failed");
    (void)hipMemset(t->m[s], 0, kNSide * 4);
    (void)hipMemset(t->v[s], 0, kNSide * 4);
    t->adam_t[s] = 0;
  }
  return 0;
}

int bq_lb_trainer_get_weights(bq_lb_trainer *t, float *weights) {
  if (!t || !weights) return -1;
  if (hipDeviceSynchronize() != hipSuccess) return tfail(t, -100, "synchronize failed");
  for (int s = 0; s < 3; ++s)
    if (hipMemcpy(weights + (size_t)s * kNSide, t->w[s], kNSide * 4, hipMemcpyDeviceToHost) != hipSuccess)
      return tfail(t, -100, "get_weights: copy failed");
  return 0;
}

// Forward (and with lr > 0 backward + Adam) on one minibatch of side `side`.
int bq_lb_trainer_step(bq_lb_trainer *t, int side, const int32_t *groups, float lr, double *loss) {
  if (!t || !groups || side < 0 || side > 2) return -1;
  for (int b = 0; b < t->batch; ++b)
    if (groups[b] < 0 || groups[b] >= t->max_groups) return tfail(t, -2, "group out of range");
  const int B = t->batch, L = lb_side_len(side, t->MX, t->MZ);
  const size_t nbl = (size_t)B * L, nout = nbl * kLbH;
  hipStream_t st = t->stream;
  if (hipMemcpyAsync(t->groups, groups, B * 4, hipMemcpyHostToDevice, st) != hipSuccess)
    return tfail(t, -100, "group upload failed");
  float *w1 = t->w[side], *w2 = w1 + kNW1, *w3 = w2 + kNW2;
  hipLaunchKernelGGL(k_tr_gather, dim3(grid_of(nbl * (kLbCin + kLbH))), dim3(kTB), 0, st, t->store,
                     t->gsize, t->bsize, lb_side_off(side, t->MZ), t->groups, B, L, t->inv + side * kLbF,
                     t->residual, t->x, t->y);
  const int go = grid_of(nout);
  hipLaunchKernelGGL(k_tr_fwd, dim3(go), dim3(kTB), 0, st, t->x, w1, t->y1, t->a1, B, L, kLbCin, 1);
  hipLaunchKernelGGL(k_tr_fwd, dim3(go), dim3(kTB), 0, st, t->a1, w2, t->y2, t->a2, B, L, kLbH, 1);
  hipLaunchKernelGGL(k_tr_fwd, dim3(go), dim3(kTB), 0, st, t->a2, w3, t->y3, t->o3, B, L, kLbH, 0);
  hipLaunchKernelGGL(k_tr_loss, dim3(go), dim3(kTB), 0, st, t->o3, t->y, t->d3, nout, t->inner,
                     t->part);
  std::vector<double> h(go);
  if (hipMemcpyAsync(h.data(), t->part, go * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
    return tfail(t, -100, "loss copy failed");
  if (lr > 0.f) {
    float *g1 = t->gw, *g2 = g1 + kNW1, *g3 = g2 + kNW2;
    const int sp12 = 16, sp1 = 4;
    // layer 3: dW3 from a2, d3; d2 = W3^T d3 * lrelu'(y2)
    hipLaunchKernelGGL(k_tr_bwd_w, dim3((kNW2 + kTB - 1) / kTB, sp12), dim3(kTB), 0, st, t->a2, t->d3,
                       t->gpart, B, L, kLbH, sp12);
    hipLaunchKernelGGL(k_tr_sum_split, dim3((kNW2 + kTB - 1) / kTB), dim3(kTB), 0, st, t->gpart, g3, kNW2, sp12);
    hipLaunchKernelGGL(k_tr_bwd_x, dim3(go), dim3(kTB), 0, st, t->d3, w3, t->y2, t->d2, B, L, kLbH);
    // layer 2
    hipLaunchKernelGGL(k_tr_bwd_w, dim3((kNW2 + kTB - 1) / kTB, sp12), dim3(kTB), 0, st, t->a1, t->d2,
                       t->gpart, B, L, kLbH, sp12);
    hipLaunchKernelGGL(k_tr_sum_split, dim3((kNW2 + kTB - 1) / kTB), dim3(kTB), 0, st, t->gpart, g2, kNW2, sp12);
    hipLaunchKernelGGL(k_tr_bwd_x, dim3(go), dim3(kTB), 0, st, t->d2, w2, t->y1, t->d1, B, L, kLbH);
    // layer 1
    hipLaunchKernelGGL(k_tr_bwd_w, dim3((kNW1 + kTB - 1) / kTB, sp1), dim3(kTB), 0, st, t->x, t->d1,
                       t->gpart, B, L, kLbCin, sp1);
    hipLaunchKernelGGL(k_tr_sum_split, dim3((kNW1 + kTB - 1) / kTB), dim3(kTB), 0, st, t->gpart, g1, kNW1, sp1);
    const int n = ++t->adam_t[side];
    const float c1 = 1.f - std::pow(0.9f, (float)n), c2 = 1.f - std::pow(0.999f, (float)n);
    hipLaunchKernelGGL(k_tr_adam, dim3((kNSide + kTB - 1) / kTB), dim3(kTB), 0, st, t->w[side], t->gw,
                       t->m[side], t->v[side], kNSide, lr, c1, c2);
  }
  if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess)
    return tfail(t, -100, "training step failed");
  if (loss) {
    double s = 0.0;
    for (double v : h) s += v;
    *loss = s / ((double)nout * t->inner / kLbS);
  }
  return 0;
}

}  // extern "C"
