// Internal declarations shared by the HIP translation units of libcochlea_amd.
#ifndef CM_INTERNAL_H
#define CM_INTERNAL_H

#include <hip/hip_runtime.h>

#include "cm_types.h"
#include "../../include/cochlea_amd.h"

namespace cm {

struct BmLaunch {
  cm_bm_scalars sc;
  const float *__restrict__ input;
  int64_t input_len;
  const float *__restrict__ sec;     // 8 x 256
  const float *__restrict__ gamma;   // sets x 256
  const cm_bm_item *__restrict__ items;
  int32_t n_items;
  int32_t use_queue;
  int32_t *queue;                    // work counter (use_queue != 0)
  float *__restrict__ speeds;        // rows x 256
  int64_t rows;
  int32_t *failed_nodes;
  float *conv_time, *conv_loops, *conv_jacobi;
  int32_t *failed_blocks;
  float *conv_time_blocks, *conv_loops_blocks, *conv_jacobi_blocks;
  cm_item_stats *stats;
};

void set_error(const char *fmt, ...);
int check(hipError_t e, const char *what);
hipStream_t stream();
float *bm_speeds_ptr();
size_t bm_result_nodes();
float *bm_input_ptr();
size_t bm_input_cap();
int bm_ensure_input(size_t n);

}  // namespace cm

extern "C" int cm_launch_bm(const cm::BmLaunch *a, int waves_per_block, int grid_blocks, hipStream_t stream);

#endif
