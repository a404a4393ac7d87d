/*
 * libsnapgpu_probe — measurement aids for bench.py, kept out of libsnapgpu's
 * ABI (include/snapgpu.h): kernels of known traffic or access pattern that
 * calibrate profiler counters and give the gather ceiling of the operator's
 * sparse products. No reference counterpart. Return 0 or 1 (bad argument) /
 * 2 (HIP failure), with the message from snapgpu_probe_last_error().
 */
#ifndef SNAPGPU_PROBE_H
#define SNAPGPU_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* snapgpu_probe_last_error(void);

/* Streaming device-to-device copy of `bytes` (multiple of 16, 16-byte aligned
 * pointers): a kernel of known traffic, used to calibrate the hardware byte
 * counters that profiles report. */
int snapgpu_probe_copy_bytes(const void* d_src, void* d_dst, int64_t bytes, void* stream);

/* Gather-rate probe: reads one 16-float row of d_dense[rows][16] per entry of
 * d_idx (row ids < rows), with the access pattern of the operator's panel
 * products (16 lanes per row, 64 indices per group batch, eight rows in flight
 * per lane). With an L2-sized d_dense it gives the gather ceiling of those
 * products. d_out: 4096 * 256 floats (per-thread sums, so the loads stay live). */
int snapgpu_probe_gather(const uint32_t* d_idx, int64_t count, const float* d_dense, int64_t rows,
                         float* d_out, int64_t out_len, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SNAPGPU_PROBE_H */
