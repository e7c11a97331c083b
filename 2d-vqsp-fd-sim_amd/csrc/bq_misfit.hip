// bq_misfit.hip - misfit of two runs' receiver traces, on the device.
//
// The reference writes its seismograms to files and has no comparison of its
// own; this is the measure the batched and boundary studies state their
// results in: for traces a and b of the same receivers and steps,
// sum (a - b)^2 and sum b^2, per receiver and in total, accumulated in
// binary64. The traces stay where the runs left them, in HBM.
//
// Two entry points share the kernels. bq_trace_misfit stages per-column
// partial sums in a buffer it allocates on each call. bq_trace_misfit_noalloc
// stages them in a file-scope device array instead and allocates nothing, for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops); it takes at most kMisfitMaxReceivers receivers.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/biotfd.h"
#include "bq_handle.h"

namespace {

constexpr int kMisfitBlock = 256;
constexpr int kMisfitMaxReceivers = 8192;
// staging of one call: 12*mr column sums, 2*mr receiver sums, 2 totals
constexpr int kMisfitStage = 14 * kMisfitMaxReceivers + 2;

// Staging of bq_trace_misfit_noalloc (917,520 bytes): one copy per device,
// held until the process ends and shared by every call on that device. Host
// code never passes this symbol to a kernel; it uses the address that
// hipGetSymbolAddress returned for the device.
static __device__ __align__(256) double scratch[kMisfitStage];

// Thread t = j*6 + c (receiver j, column c) sums its column over the steps:
// st[2t] = sum (a - b)^2, st[2t+1] = sum b^2. a and b are [nsteps][mr][6],
// so the threads of a wavefront read consecutive floats of each step.
__global__ __launch_bounds__(kMisfitBlock) void k_misfit_columns(const float *__restrict__ a,
                                                                 const float *__restrict__ b,
                                                                 int mr, int nsteps,
                                                                 double *__restrict__ st) {
  const int t = blockIdx.x * kMisfitBlock + threadIdx.x;
  const int ncol = mr * 6;
  if (t >= ncol) return;
  double d2 = 0.0, b2 = 0.0;
  for (int s = 0; s < nsteps; ++s) {
    const size_t o = (size_t)s * ncol + t;
    const double x = a[o], y = b[o];
    d2 += (x - y) * (x - y);
    b2 += y * y;
  }
  st[2 * t] = d2;
  st[2 * t + 1] = b2;
}

// Per receiver (over its 6 columns) into per[2j], per[2j+1] and the total
// into total[0..1]; one block, in a fixed order so that the sums do not
// depend on timing.
__global__ __launch_bounds__(kMisfitBlock) void k_misfit_receivers(int mr, double *__restrict__ st,
                                                                   double *__restrict__ per,
                                                                   double *__restrict__ total) {
  __shared__ double sd[kMisfitBlock], sb[kMisfitBlock];
  double d2 = 0.0, b2 = 0.0;
  for (int j = threadIdx.x; j < mr; j += kMisfitBlock) {
    double rd = 0.0, rb = 0.0;
    for (int c = 0; c < 6; ++c) {
      rd += st[2 * (j * 6 + c)];
      rb += st[2 * (j * 6 + c) + 1];
    }
    if (per) {
      per[2 * j] = rd;
      per[2 * j + 1] = rb;
    }
    d2 += rd;
    b2 += rb;
  }
  sd[threadIdx.x] = d2;
  sb[threadIdx.x] = b2;
  __syncthreads();
  for (int w = kMisfitBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sd[threadIdx.x] += sd[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    total[0] = sd[0];
    total[1] = sb[0];
  }
}

int mfail(bq_handle *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  return code;
}

// Checks the pair and the range. The staging buffer of a call holds
// 14*mr + 2 doubles: 12*mr column sums, 2*mr receiver sums, 2 totals.
int misfit_check(bq_handle *a, bq_handle *b, int first_step, int nsteps) {
  if (!a || !b) return -1;
  if (a->mr != b->mr) return mfail(a, -2, "bq_trace_misfit: the runs have different receiver counts");
  if (a->mr < 1 || !a->kp.traces || !b->kp.traces)
    return mfail(a, -2, "bq_trace_misfit: no receivers");
  if (a->device != b->device) return mfail(a, -2, "bq_trace_misfit: runs on different devices");
  if (first_step < 1 || nsteps < 1 || first_step + nsteps - 1 > a->next_step - 1 ||
      first_step + nsteps - 1 > b->next_step - 1)
    return mfail(a, -9, "bq_trace_misfit: steps not run by both runs");
  return 0;
}

int misfit_run(bq_handle *a, bq_handle *b, int first_step, int nsteps, double *stage,
               double *per_receiver, double total[2]) {
  const int mr = a->mr;
  // b's steps were queued on its own stream
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return mfail(a, -100, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  const size_t off = (size_t)(first_step - 1) * mr * 6;
  double *st = stage, *per = stage + 12 * (size_t)mr, *tot = per + 2 * (size_t)mr;
  const int blocks = (mr * 6 + kMisfitBlock - 1) / kMisfitBlock;
  hipLaunchKernelGGL(k_misfit_columns, dim3(blocks), dim3(kMisfitBlock), 0, a->stream,
                     a->kp.traces + off, b->kp.traces + off, mr, nsteps, st);
  hipLaunchKernelGGL(k_misfit_receivers, dim3(1), dim3(kMisfitBlock), 0, a->stream, mr, st, per, tot);
  e = hipGetLastError();
  if (e != hipSuccess) return mfail(a, -100, std::string("k_misfit: ") + hipGetErrorString(e));
  e = hipMemcpyAsync(total, tot, 2 * sizeof(double), hipMemcpyDeviceToHost, a->stream);
  if (e == hipSuccess && per_receiver)
    e = hipMemcpyAsync(per_receiver, per, 2 * sizeof(double) * mr, hipMemcpyDeviceToHost, a->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(a->stream);
  if (e != hipSuccess) return mfail(a, -100, std::string("bq_trace_misfit copy: ") + hipGetErrorString(e));
  return 0;
}

// Address of each device's copy of `scratch`, looked up once per device with
// that device current.
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_for(bq_handle *h, double **out) {
  if (h->device < 0 || h->device >= kMaxDevices) return mfail(h, -2, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[h->device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != h->device) e = hipSetDevice(h->device);
    void *p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != h->device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return mfail(h, -100, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[h->device] = (double *)p;
  }
  *out = g_scratch[h->device];
  return 0;
}

}  // namespace

extern "C" {

int bq_trace_misfit(bq_handle *a, bq_handle *b, int first_step, int nsteps, double *per_receiver,
                    double total[2]) {
  if (!total) return mfail(a, -1, "bq_trace_misfit: null output");
  if (int rc = misfit_check(a, b, first_step, nsteps)) return rc;
  double *stage = nullptr;
  const size_t n = 14 * (size_t)a->mr + 2;
  hipError_t e = hipMalloc(&stage, n * sizeof(double));
  if (e != hipSuccess) return mfail(a, -101, std::string("hipMalloc: ") + hipGetErrorString(e));
  const int rc = misfit_run(a, b, first_step, nsteps, stage, per_receiver, total);
  (void)hipFree(stage);
  return rc;
}

// Every value the kernels read from the staging array is written earlier in
// the same call (column sums, then receiver sums and totals), so nothing has
// to be cleared between calls. Calls that overlap on one device race on it.
int bq_trace_misfit_noalloc(bq_handle *a, bq_handle *b, int first_step, int nsteps,
                            double *per_receiver, double total[2]) {
  if (!total) return mfail(a, -1, "bq_trace_misfit_noalloc: null output");
  if (int rc = misfit_check(a, b, first_step, nsteps)) return rc;
  if (a->mr > kMisfitMaxReceivers)
    return mfail(a, -3, "bq_trace_misfit_noalloc: more than 8192 receivers; use bq_trace_misfit");
  double *stage = nullptr;
  if (int rc = scratch_for(a, &stage)) return rc;
  return misfit_run(a, b, first_step, nsteps, stage, per_receiver, total);
}

}  // extern "C"
