/*
 * biotfd_probes.h - bandwidth probes (diagnostics library
 * libbiotfd_probes.so, built beside libbiotfd.so).
 *
 * These entry points have no counterpart in the reference and are not part
 * of the time-loop ABI (biotfd.h). They measure the achievable HBM bandwidth
 * of the device and give kernels of exactly known bytes against which the
 * hardware traffic counters FETCH_SIZE / WRITE_SIZE are calibrated.
 */
#ifndef BIOTFD_PROBES_H
#define BIOTFD_PROBES_H

#ifdef __cplusplus
extern "C" {
#endif

/* Streams nfloats binary32 values from one device buffer to another iters
 * times (kernel k_copy_probe, float4 per lane) and reports read+write GB/s. */
int bq_probe_copy(long long nfloats, int iters, double *gbps);

/* The access pattern of the fused step tiles without their arithmetic
 * (kernel k_march_probe, one float per lane, loads and stores with the
 * streaming hint as the tiles' own columns use it): blocks of 512 columns x 128
 * rows march down nine [nz][nx] arrays (reading 4 halo rows more per block)
 * and write eight, one row at a time. Reports read+write GB/s and, if bytes
 * is not NULL, the exact bytes one launch reads and writes. Needs nx >= 512,
 * nz >= 140. */
int bq_probe_march(int nx, int nz, int iters, double *gbps, double *bytes);

/* The same nine-read, eight-write traffic per cell without the march (kernel
 * k_bands_probe, one cell per thread, streaming hint): each 512-thread block takes one
 * 512-column segment of one row of [nz][nx] arrays, and the blocks are
 * ordered so that those in flight cover `bands` row bands at once (bands = 1:
 * row-major order; the fused tiles of a grid with R rows of blocks march in R
 * bands). Reports read+write GB/s over columns [0, 512 * (nx / 512)) and the
 * first bands * (nz / bands) rows. Needs nx >= 512, 1 <= bands <= nz. */
int bq_probe_bands(int nx, int nz, int bands, int iters, double *gbps);

#ifdef __cplusplus
}
#endif
#endif /* BIOTFD_PROBES_H */
