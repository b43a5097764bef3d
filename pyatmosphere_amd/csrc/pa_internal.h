// Internal declarations shared by the HIP translation units of libpa_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pa_hip.h"

namespace pa {

constexpr int kMaxPupils = PA_MAX_PUPILS;     // pupils evaluated per channel in one pass
constexpr int kMaxSums = 6 + kMaxPupils;      // partial sums per tile
constexpr int kChParams = 4;                  // per channel: delta, shift_x, shift_y, spare
constexpr int kMoments = PA_NUM_MOMENTS;      // finalized values per element (see pa_hip.h)

int set_error(int code, const char* msg);
int check_launch(const char* what);

// step passes (pa_step.hip)
int row_first(int n, float2* work, const float2* in, const float2* src, const int* chan,
              const float2* tw, int batch, hipStream_t s);
int col_step(int n, float2* work, const float2* htab, const int* chan, const int* nsteps, int step,
             int max_steps, const float2* tw, int batch, hipStream_t s);
int row_mid(int n, float2* work, const float2* fac, const int* nsteps, int step, const float2* tw, int batch,
            hipStream_t s);
int row_last(int n, const float2* work, float2* out, double* part, const double* chparam,
             const double* pup, int npup, const int* chan, const float2* tw, int batch,
             hipStream_t s);
int tile_rows(int n);

}  // namespace pa
