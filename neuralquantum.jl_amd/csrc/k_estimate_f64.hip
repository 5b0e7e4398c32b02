// The per-sample estimator kernels in FP64 (template in k_estimate.h), the
// LDS budget of a launch and the precision dispatch.
#include "k_estimate.h"

hipError_t nq_launch_estimate_f32(const est_args& a, int grid, hipStream_t st);

size_t est_lds_bytes(const est_args& a, size_t elt, int waves) {
    // params | exact force accumulator | waves x (W cache + flip masks) | table (16-aligned)
    const size_t b = (size_t)a.lds_param_bytes + ((a.Fpart || a.G) ? (size_t)fx_bytes(a.d.np) : 0) +
                     (size_t)waves * (size_t)a.max_conn * (2 * elt + 8);
    return ((b + 15) & ~(size_t)15) + (a.tab_lds ? (size_t)a.tab.bytes : 0);
}

hipError_t nq_launch_estimate(const est_args& a, bool fp64, int grid, hipStream_t st) {
    if (a.d.N > 16) return hipErrorInvalidValue;
    est_args b = a;
    if (b.rec_mode == 2) b.tab_lds = 0;
    if (est_lds_bytes(b, fp64 ? 8 : 4, b.G ? 1 : 4) > 160 * 1024) return hipErrorInvalidValue;
    if (fp64) return nq_est::launch_prec<double>(a, grid, st);
    return nq_launch_estimate_f32(a, grid, st);
}
