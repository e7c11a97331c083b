// bq_learned.h - device code of the learned absorbing boundary, shared by the
// single-grid launches (bq_learned.hip) and the batches (bq_batch.hip).
//
// A grid whose PML is a thin strip of S = 8 cells on the left, right and bottom
// (the reference's CPML of ppml.F90:159-306 with TPML = 8) has its strip cells
// replaced after every step by a small network that stands in for a wide PML.
// For each side the network sees a band of 12 cells across the boundary - the
// 8 strip cells and the 4 interior cells next to them - for all 8 fields at
// the last 4 time levels, and predicts the change of the 8 strip cells' 8
// fields over the step:
//
//   input  x[p][c]  p along the boundary, c = level*96 + field*12 + cell
//                   (384 channels, each field divided by its scale)
//   layer1 h1 = lrelu(conv5(x,  W1: 384 -> 64))
//   layer2 h2 = lrelu(conv5(h1, W2:  64 -> 64))
//   layer3 d  =       conv5(h2, W3:  64 -> 64)     d[p][field*8 + cell]
//   strip(t+1) = strip(t) + gain * scale[field] * d            (direct)
//   strip(t+1) = cpml(t+1) + gain * scale[field] * d           (residual)
//   strip(t+1) = clamp(cpml(t+1) + gain * scale[field] * d,
//                      between 0 and cpml(t+1))                (damping)
//
// In the direct form the network is the strip's whole update; in the residual
// form it corrects the update the 8-cell CPML has just made, and gain 0 is
// the plain CPML. The damping form is the residual form limited so that the
// network can only pull each strip value toward zero, never past it: a
// sponge (a per-cell factor in [0, 1] after the CPML's update) whose factors
// the network chooses, so it cannot feed energy back into the grid. Only the
// `inner` strip cells nearest the interior receive the output; the others
// keep the value the 8-cell CPML computed.
//
// conv5 is a 5-tap convolution along the boundary with zero padding at the
// ends of each layer's input, no bias; lrelu(y) = max(y, 0.1 y). Inference
// runs the three layers in BF16 on the matrix cores (MFMA 16x16x16, FP32
// accumulation), all three in LDS for a tile of 52 positions per block.
//
// Band layout (one time level): side 0 (left, x = -7..4, rows l = 0..MZ'),
// side 1 (right, x = MX+8..MX-3, rows 0..MZ'), side 2 (bottom, rows
// l = MZ'+8..MZ'-3, x = -7..MX+8), each [pos][field][cell] with cell 0 the
// outermost strip cell, 7 the innermost, 8..11 the interior cells inward.
#pragma once
#include <hip/hip_runtime.h>

#include "bq_kernels.h"

namespace bqk {

constexpr int kLbS = 8;                 // strip cells per side
constexpr int kLbCells = 12;            // band cells across the boundary
constexpr int kLbF = 8;                 // fields
constexpr int kLbBand = kLbF * kLbCells;  // 96 values per band position and level
constexpr int kLbLev = 4;               // input time levels
constexpr int kLbCin = kLbLev * kLbBand;  // 384
constexpr int kLbH = 64;                // hidden channels = outputs (8 fields x 8 cells)
constexpr int kLbTap = 5;               // taps of each convolution
constexpr int kLbP = 52;                // valid output positions per inference block
constexpr int kLbRows = 68;             // LDS rows of each activation tile
constexpr int kLbXH = kLbCin / 2 + 8;   // LDS row stride of half the input tile (bf16)
constexpr int kLbHS = kLbH + 8;         // LDS row stride of the hidden tiles (bf16)
constexpr int kLbBlock = 256;
enum { kLbDirect = 0, kLbResidual = 1, kLbDamping = 2 };

// Weights of the three networks in BF16, [tap][out][in] (in contiguous), and
// the per-field scales of each side.
struct LbNet {
  const unsigned short *w1[3], *w2[3], *w3[3];
  float scale[3][kLbF], inv[3][kLbF];
  int form;       // kLbDirect, kLbResidual or kLbDamping (see above)
  float gain;
  int inner;      // strip cells the output is applied to, counted from the interior (1..8)
};

__host__ __device__ inline int lb_lz(int MZ) { return MZ + 1; }
__host__ __device__ inline int lb_lx(int MX) { return MX + 2 * kLbS; }
__host__ __device__ inline int lb_side_len(int side, int MX, int MZ) {
  return side < 2 ? lb_lz(MZ) : lb_lx(MX);
}
__host__ __device__ inline size_t lb_side_off(int side, int MZ) {
  return (size_t)side * lb_lz(MZ) * kLbBand;
}
__host__ __device__ inline size_t lb_band_size(int MX, int MZ) {
  return (size_t)(2 * lb_lz(MZ) + lb_lx(MX)) * kLbBand;
}

// Grid cell (i, l) of band position pos, cell c of a side.
__device__ __forceinline__ void lb_cell(int side, int pos, int c, int MX, int MZ, int *i, int *l) {
  if (side == 0) {
    *i = 1 - kLbS + c;
    *l = pos;
  } else if (side == 1) {
    *i = MX + kLbS - c;
    *l = pos;
  } else {
    *i = 1 - kLbS + pos;
    *l = MZ + kLbS - c;
  }
}

// Element e of the band of one level, from the fields F[8] of a grid with
// pitch P and column offset OFF (element (i,l) at l*P + i + OFF); with sub,
// subtracted from what out holds.
__device__ __forceinline__ void lb_extract(float *__restrict__ out, const float *const *F, int P,
                                           int OFF, int MX, int MZ, size_t e, int sub) {
  const size_t s1 = lb_side_off(1, MZ), s2 = lb_side_off(2, MZ);
  const int side = e < s1 ? 0 : (e < s2 ? 1 : 2);
  const size_t r = e - lb_side_off(side, MZ);
  const int pos = (int)(r / kLbBand), fc = (int)(r % kLbBand);
  const int f = fc / kLbCells, c = fc % kLbCells;
  int i, l;
  lb_cell(side, pos, c, MX, MZ, &i, &l);
  const float v = F[f][(size_t)l * P + (size_t)(i + OFF)];
  out[e] = sub ? out[e] - v : v;
}

__device__ __forceinline__ unsigned short lb_bf16(float x) {
  // round to nearest even (the values are finite)
  unsigned int u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

typedef short lb_bf16x4 __attribute__((ext_vector_type(4)));
typedef float lb_f32x4 __attribute__((ext_vector_type(4)));

// The band at the last four time levels, oldest first (level t = p3).
struct LbLevels {
  const float *p0, *p1, *p2, *p3;
};

__device__ __forceinline__ const float *lb_level(const LbLevels &v, int lv) {
  return lv == 0 ? v.p0 : lv == 1 ? v.p1 : lv == 2 ? v.p2 : v.p3;
}

// The first layer's input is staged in two halves of its channels (levels
// 0-1, then 2-3), so that three blocks fit in a CU's LDS.
struct alignas(16) LbLds {
  unsigned short x[kLbRows * kLbXH];
  unsigned short h1[kLbRows * kLbHS];
  unsigned short h2[kLbRows * kLbHS];
};

// One 5-tap convolution layer on the matrix cores for a tile of 64 output
// rows: out[r][co] = sum_t sum_ci in[r + t][ci] * w[t][co][ci], in and out in
// LDS (bf16, row stride is), w in global memory in fragment order (below).
// Wave v computes the 16 outputs 16v..16v+15 for all 64 rows (four 16-row
// tiles), so each weight fragment is loaded once per block and used four
// times. The MFMA 16x16x16 operand layout: lane q holds A[row q % 16][k
// 4(q/16) .. +3] and B[k 4(q/16) .. +3][col q % 16]; the result holds D[row
// 4(q/16) + j][col q % 16], j = 0..3.
//
// Fragment order: for each tap t, output block v (16 outputs) and input
// block kc (16 inputs), the 64 lanes' B operands one after another - lane q's
// four values W[t][16v + q%16][16kc + 4(q/16) .. +3] at (((t*4 + v) * CIN/16
// + kc) * 64 + q) * 4 - so that a wave loads each fragment as 512
// contiguous bytes.
// The input blocks KC0 .. KC0+NK-1 of a layer of CIN inputs (in: those
// NK*16 channels, row stride is), added to acc (zeroed first with ZERO).
template <int CIN, int KC0, int NK, bool ZERO>
__device__ __forceinline__ void lb_layer_part(const unsigned short *in, int is,
                                              const unsigned short *__restrict__ w,
                                              lb_f32x4 acc[4]) {
  const int lane = threadIdx.x & 63, v = threadIdx.x >> 6;
  const int qr = lane & 15, qk = 4 * (lane >> 4);
  if (ZERO)
    for (int m = 0; m < 4; ++m) acc[m] = lb_f32x4{0.f, 0.f, 0.f, 0.f};
  for (int t = 0; t < kLbTap; ++t) {
    const unsigned short *wfrag =
        w + (((size_t)(t * 4 + v) * (CIN / 16) + KC0) * 64 + lane) * 4;
    const unsigned short *arow = in + (size_t)(qr + t) * is + qk;
    // every weight fragment of the tap in flight at once (one memory
    // latency per tap instead of one per few fragments)
    lb_bf16x4 bw[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) bw[k] = *(const lb_bf16x4 *)(wfrag + (size_t)k * 256);
#pragma unroll
    for (int cb = 0; cb < NK * 16; cb += 16) {
      const lb_bf16x4 b = bw[cb / 16];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const lb_bf16x4 a = *(const lb_bf16x4 *)(arow + (size_t)16 * m * is + cb);
        acc[m] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[m], 0, 0, 0);
      }
    }
  }
}

template <int CIN>
__device__ __forceinline__ void lb_layer(const unsigned short *in, int is,
                                         const unsigned short *__restrict__ w, lb_f32x4 acc[4]) {
  lb_layer_part<CIN, 0, CIN / 16, true>(in, is, w, acc);
}

// Row and output channel of element j of acc[m] in lane `lane` of wave v.
__device__ __forceinline__ int lb_row(int m, int lane, int j) { return 16 * m + 4 * (lane >> 4) + j; }
__device__ __forceinline__ int lb_col(int v, int lane) { return 16 * v + (lane & 15); }

// Hidden layer result -> LDS (bf16), lrelu, zero outside the side's
// positions (row r is at position base + r).
__device__ __forceinline__ void lb_store_hidden(unsigned short *h, const lb_f32x4 acc[4], int base,
                                                int L) {
  const int lane = threadIdx.x & 63, v = threadIdx.x >> 6;
  const int co = lb_col(v, lane);
  for (int m = 0; m < 4; ++m)
    for (int j = 0; j < 4; ++j) {
      const int r = lb_row(m, lane, j);
      const int p = base + r;
      float y = acc[m][j];
      y = y > 0.f ? y : 0.1f * y;
      h[r * kLbHS + co] = (p >= 0 && p < L) ? lb_bf16(y) : (unsigned short)0;
    }
  // rows 64..67 only feed rows whose results are discarded
  for (int e = threadIdx.x; e < 4 * kLbHS; e += kLbBlock) h[64 * kLbHS + e] = 0;
}

// One tile of one side: positions p0 .. p0+51. lev: the band at the last
// four time levels, oldest first. For each strip
// cell put(p, f, c, prev, d) receives lev[3]'s value and the network's
// gain * scale * output.
template <class Put>
__device__ __forceinline__ void lb_tile(const LbNet &net, int side, int L, size_t soff, int p0,
                                        const LbLevels lev, LbLds &S, Put put) {
  const float *inv = net.inv[side];
  // input rows r at positions p0 - 6 + r, as float4 runs of 4 cells of one
  // field (12 cells per field): 24 per row and level; half h holds levels
  // 2h, 2h+1
  constexpr int kQ = kLbBand / 4, kN = kLbRows * 2 * kQ;
  constexpr int kIt = (kN + kLbBlock - 1) / kLbBlock;
  auto load_half = [&](int h, float4 *xs) {
#pragma unroll
    for (int k = 0; k < kIt; ++k) {
      const int q = threadIdx.x + k * kLbBlock;
      const int r = q / (2 * kQ), rem = q % (2 * kQ);
      const int lv = 2 * h + rem / kQ, c4 = rem % kQ;
      const int p = p0 - 6 + r;
      xs[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (q < kN && p >= 0 && p < L)
        xs[k] = *(const float4 *)(lb_level(lev, lv) + soff + (size_t)p * kLbBand + 4 * c4);
    }
  };
  auto store_half = [&](const float4 *xs) {
#pragma unroll
    for (int k = 0; k < kIt; ++k) {
      const int q = threadIdx.x + k * kLbBlock;
      if (q >= kN) break;
      const int r = q / (2 * kQ), rem = q % (2 * kQ);
      const int l2 = rem / kQ, c4 = rem % kQ;
      const float4 x = xs[k];
      const float sc = inv[(4 * c4) / kLbCells];
      const unsigned int lo = lb_bf16(x.x * sc) | ((unsigned int)lb_bf16(x.y * sc) << 16);
      const unsigned int hi = lb_bf16(x.z * sc) | ((unsigned int)lb_bf16(x.w * sc) << 16);
      *(uint2 *)(S.x + r * kLbXH + l2 * kLbBand + 4 * c4) = make_uint2(lo, hi);
    }
  };
  float4 xa[kIt];
  load_half(0, xa);
  store_half(xa);
  __syncthreads();
  lb_f32x4 acc[4];
  lb_layer_part<kLbCin, 0, kLbCin / 32, true>(S.x, kLbXH, net.w1[side], acc);
  load_half(1, xa);
  __syncthreads();
  store_half(xa);
  __syncthreads();
  lb_layer_part<kLbCin, kLbCin / 32, kLbCin / 32, false>(S.x, kLbXH, net.w1[side], acc);
  lb_store_hidden(S.h1, acc, p0 - 4, L);
  __syncthreads();
  lb_layer<kLbH>(S.h1, kLbHS, net.w2[side], acc);
  lb_store_hidden(S.h2, acc, p0 - 2, L);
  __syncthreads();
  lb_layer<kLbH>(S.h2, kLbHS, net.w3[side], acc);
  const int lane = threadIdx.x & 63, v = threadIdx.x >> 6;
  const int co = lb_col(v, lane), f = co >> 3, c = co & 7;
  if (c < kLbS - net.inner) return;
  const float g = net.gain * net.scale[side][f];
  for (int m = 0; m < 4; ++m)
    for (int j = 0; j < 4; ++j) {
      const int r = lb_row(m, lane, j);
      const int p = p0 + r;
      if (r < kLbP && p < L) {
        const float prev = lev.p3[soff + (size_t)p * kLbBand + f * kLbCells + c];
        put(p, f, c, prev, g * acc[m][j]);
      }
    }
}

__host__ __device__ inline int lb_tiles(int L) { return (L + kLbP - 1) / kLbP; }

// New value of a strip cell: cpml is what the 8-cell CPML has just computed
// there, prev its value before the step, d the network's gain * scale * output.
__device__ __forceinline__ float lb_update(int form, float cpml, float prev, float d) {
  if (form == kLbDirect) return prev + d;
  const float y = cpml + d;
  if (form == kLbResidual) return y;
  return cpml >= 0.f ? fminf(fmaxf(y, 0.f), cpml) : fmaxf(fminf(y, 0.f), cpml);
}

}  // namespace bqk

namespace bq {
// The networks of a bq_lb (bq_learned.hip).
const bqk::LbNet &lb_net_of(const void *lb);
}  // namespace bq
