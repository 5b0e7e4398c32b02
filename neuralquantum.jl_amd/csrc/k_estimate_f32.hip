// The per-sample estimator kernels in FP32 (template in k_estimate.h).
#include "k_estimate.h"

hipError_t nq_launch_estimate_f32(const est_args& a, int grid, hipStream_t st) {
    return nq_est::launch_prec<float>(a, grid, st);
}
