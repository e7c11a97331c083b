// bq_probes.hip - bandwidth probes for gfx950 (diagnostics build,
// libbiotfd_probes.so). They have no counterpart in the reference and are not
// part of the time-loop library: they measure the device's achievable HBM
// bandwidth and calibrate the hardware traffic counters (FETCH_SIZE,
// WRITE_SIZE) against kernels whose bytes are known exactly.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/biotfd_probes.h"

namespace {

constexpr int kFW = 512;  // columns per block of the row-streaming probe
constexpr int kFR = 128;  // rows per block

// Block b -> tile, so that XCD b % 8 gets a contiguous band of tiles.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
  const int per = nblocks >> 3, rem = nblocks & 7, x = b & 7, q = b >> 3;
  return x * per + min(x, rem) + q;
}

// Plain streaming copy: each block moves 4 x 256 float4 (16 KiB), four
// independent loads in flight per lane.
__global__ __launch_bounds__(256) void k_copy_probe(const float4 *__restrict__ a,
                                                   float4 *__restrict__ b, size_t n4) {
  const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
  float4 v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const size_t t = base + 256 * u;
    v[u] = t < n4 ? a[t] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const size_t t = base + 256 * u;
    if (t < n4) b[t] = v[u];
  }
}

// Loads and stores with the streaming (non-temporal) hint, as the fused step
// tiles read and write their own columns' field rows.
__device__ __forceinline__ float stream_load(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void stream_store(float *p, float v) { __builtin_nontemporal_store(v, p); }

// Row-streaming probe: blocks of kFW columns x kFR rows march down nine input
// arrays and write eight output arrays two rows behind, with no stencil work -
// the access pattern of the fused step tiles without their arithmetic, one
// float per lane per array and row, with the tiles' streaming hint. Columns [0, nbx*kFW), rows
// [4, 4 + nbz*kFR + 4) of [nz][pitch] arrays. Bytes per launch are exact:
// nbx*kFW*(nbz*(kFR+4))*9*4 read (the 4 halo rows of each block read again),
// nbx*kFW*nbz*kFR*8*4 written.
struct MarchArrays {
  const float *in[9];
  float *out[8];
};

// Band-order probe: block b takes segment b % nbx of row band (b / nbx) %
// bands, row b / (nbx * bands) of that band (h rows per band), and reads nine
// arrays and writes eight at its 512 cells, with the streaming hint.
__global__ __launch_bounds__(kFW) void k_bands_probe(MarchArrays A, int pitch, int nbx, int bands, int h) {
  const int b = blockIdx.x;
  const int seg = b % nbx, band = (b / nbx) % bands, r = b / (nbx * bands);
  const size_t i = (size_t)(band * h + r) * pitch + (size_t)seg * kFW + threadIdx.x;
  float v = 0.0f;
#pragma unroll
  for (int f = 0; f < 9; ++f) v += stream_load(&A.in[f][i]);
#pragma unroll
  for (int f = 0; f < 8; ++f) stream_store(&A.out[f][i], v);
}

__global__ __launch_bounds__(kFW) void k_march_probe(MarchArrays A, int pitch, int nbx) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int bx = tile % nbx, bz = tile / nbx;
  const int c = bx * kFW + threadIdx.x;
  const int a = 4 + bz * kFR;
  for (int s = a; s < a + kFR + 4; ++s) {
    const size_t r = (size_t)s * pitch + c;
    float v = 0.0f;
#pragma unroll
    for (int f = 0; f < 9; ++f) v += stream_load(&A.in[f][r]);
    const int l = s - 2;
    if (l >= a && l < a + kFR) {
      const size_t rl = (size_t)l * pitch + c;
#pragma unroll
      for (int f = 0; f < 8; ++f) stream_store(&A.out[f][rl], v);
    }
  }
}

}  // namespace

extern "C" {

int bq_probe_copy(long long nfloats, int iters, double *gbps) {
  if (nfloats < 4 || iters < 1 || !gbps) return -1;
  const size_t n4 = (size_t)nfloats / 4;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n4 * 16) != hipSuccess) return -101;
  if (hipMalloc(&b, n4 * 16) != hipSuccess) { (void)hipFree(a); return -101; }
  (void)hipMemset(a, 0, n4 * 16);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const int blocks = (int)((n4 + 1023) / 1024);
  hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, 0, a, b, n4);
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < iters; ++r) hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, 0, a, b, n4);
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *gbps = 2.0 * n4 * 16 * iters / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

int bq_probe_march(int nx, int nz, int iters, double *gbps, double *bytes) {
  if (nx < kFW || nz < kFR + 12 || iters < 1 || !gbps) return -1;
  const int pitch = (nx + 63) / 64 * 64;
  const int nbx = nx / kFW, nbz = (nz - 8) / (kFR + 4);
  // rows touched: 4 .. 4 + nbz*kFR + 3 < nz; columns < nbx*kFW <= nx <= pitch
  const size_t n = (size_t)pitch * nz;
  std::vector<float *> buf(17, nullptr);
  int rc = 0;
  for (auto &p : buf) {
    if (hipMalloc(&p, n * sizeof(float)) != hipSuccess) { p = nullptr; rc = -101; break; }
    (void)hipMemset(p, 0, n * sizeof(float));
  }
  if (rc == 0) {
    MarchArrays A{};
    for (int f = 0; f < 9; ++f) A.in[f] = buf[f];
    for (int f = 0; f < 8; ++f) A.out[f] = buf[9 + f];
    const int blocks = nbx * nbz;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(k_march_probe, dim3(blocks), dim3(kFW), 0, 0, A, pitch, nbx);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < iters; ++r)
      hipLaunchKernelGGL(k_march_probe, dim3(blocks), dim3(kFW), 0, 0, A, pitch, nbx);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double cols = (double)nbx * kFW;
    const double moved = cols * nbz * (kFR + 4) * 9.0 * 4.0 + cols * nbz * kFR * 8.0 * 4.0;
    *gbps = moved * iters / (ms * 1e-3) / 1e9;
    if (bytes) *bytes = moved;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (hipGetLastError() != hipSuccess) rc = -100;
  }
  for (auto p : buf)
    if (p) (void)hipFree(p);
  return rc;
}

int bq_probe_bands(int nx, int nz, int bands, int iters, double *gbps) {
  if (nx < kFW || bands < 1 || bands > nz || iters < 1 || !gbps) return -1;
  const int pitch = (nx + 63) / 64 * 64;
  const int nbx = nx / kFW, h = nz / bands;
  // rows touched: band * h + r < bands * h <= nz; columns < nbx * kFW <= nx <= pitch
  const size_t n = (size_t)pitch * nz;
  std::vector<float *> buf(17, nullptr);
  int rc = 0;
  for (auto &p : buf) {
    if (hipMalloc(&p, n * sizeof(float)) != hipSuccess) { p = nullptr; rc = -101; break; }
    (void)hipMemset(p, 0, n * sizeof(float));
  }
  if (rc == 0) {
    MarchArrays A{};
    for (int f = 0; f < 9; ++f) A.in[f] = buf[f];
    for (int f = 0; f < 8; ++f) A.out[f] = buf[9 + f];
    const int blocks = nbx * bands * h;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(k_bands_probe, dim3(blocks), dim3(kFW), 0, 0, A, pitch, nbx, bands, h);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < iters; ++r)
      hipLaunchKernelGGL(k_bands_probe, dim3(blocks), dim3(kFW), 0, 0, A, pitch, nbx, bands, h);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *gbps = (double)blocks * kFW * 17.0 * 4.0 * iters / (ms * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (hipGetLastError() != hipSuccess) rc = -100;
  }
  for (auto p : buf)
    if (p) (void)hipFree(p);
  return rc;
}

}  // extern "C"
