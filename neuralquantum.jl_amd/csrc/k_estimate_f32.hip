// The per-sample estimator kernels in FP32 (template in k_estimate.h).
#include "k_estimate.h"

hipError_t nq_launch_estimate_f32(const est_args& a, int grid, hipStream_t st) {
    return nq_est::launch_prec<float>(a, grid, st);
}

hipError_t nq_pk_estimate_f32(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st) {
    return nq_est::pk_launch_prec<float>(M, L, st);
}
