// Launchers for the step kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "layout.h"

namespace titan_mi355x {

hipError_t launch_step(const StepArgs &a, bool actuated, hipStream_t s);
const char *step_kernel_name(bool actuated);
hipError_t launch_step_tiles(const StepArgs &a, bool actuated, hipStream_t s);
const char *step_tiles_kernel_name(bool actuated);
size_t tiles_lds_bytes(int ldim);
// A body joins (bi.on) or leaves the launches.
hipError_t launch_set_body(const BodyTable &t, const BodyInit &bi, hipStream_t s);
// Actuation factors for the first step of a stepping call (launch slot a.slot).
hipError_t launch_actuate(const StepArgs &a, hipStream_t s);
// LDS of the resident kernel's two state buffers for `rows` masses.
size_t resident_lds_bytes(int rows);
hipError_t launch_rollout_resident(const StepArgs &a, bool actuated, int final_slot, hipStream_t s);
const char *resident_kernel_name(bool actuated);
// the resident launch whose bodies are all split-form bodies in LDS
const char *split_kernel_name(bool actuated);
// Raise the dynamic LDS limits of the tiled and resident kernels; *resident_max
// receives the resident kernel's limit (the device's per-workgroup maximum
// or what the kernel can use, the smaller), *split_max the split-form
// resident kernel's (less its static LDS).
hipError_t set_tiles_lds_limit(int *resident_max, int *split_max);
hipError_t launch_advance_T(double *Tm, const uint32_t *mflags, const int32_t *chunk_body,
                            const int64_t *bn, int64_t *bnsync, int64_t chunk_begin, int64_t chunk_end,
                            int64_t nb, double dt, hipStream_t s);
hipError_t launch_set_left(int64_t *bleft, const int32_t *b_chunk0, int64_t nb, int64_t n,
                           hipStream_t s);
hipError_t launch_set_i64(int64_t *p, int64_t v, hipStream_t s);
// HBM calibration: a streaming copy of n 16-byte words (read + non-temporal
// write) and a streaming read of n words (summed into out[0..gridsize), so
// that the loads are not removed).
hipError_t launch_stream_copy(const void *src, void *dst, int64_t n, hipStream_t s);
hipError_t launch_stream_read(const void *src, double *out, int64_t n, hipStream_t s);

}  // namespace titan_mi355x
