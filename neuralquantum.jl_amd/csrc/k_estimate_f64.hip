// The per-sample estimator kernels in FP64 (template in k_estimate.h), the
// LDS budget of a launch and the precision dispatch.
#include "k_estimate.h"

hipError_t nq_launch_estimate_f32(const est_args& a, int grid, hipStream_t st);

size_t est_lds_bytes(const est_args& a, size_t elt, int waves) {
    // params | exact force accumulator | waves x (W cache + flip masks) | table (16-aligned);
    // the FP32 forces launch stages its parameters in a padded layout of its own
    size_t pb = (size_t)a.lds_param_bytes;
    if (elt == 4 && a.rec_mode == 2 && !a.G)
        pb = std::max(pb, ((size_t)4 * nq_est::force_stage_floats(a.d) + 15) & ~(size_t)15);
    const size_t b = pb + ((a.Fpart || a.G) ? (size_t)fx_bytes(a.d.np) : 0) +
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

// the instantiation and block shape a walk (rec_mode 1) or force (rec_mode 2)
// launch of `a` uses, as nq_launch_estimate chooses them; -1 when the packed
// launches do not cover it (more than one unit or Pi row per lane)
int nq_est_variant(const est_args& a0, bool fp64, int* threads, size_t* lds) {
    est_args a = a0;
    if (a.d.N > 16 || a.G || (a.rec_mode != 1 && a.rec_mode != 2)) return -1;
    if ((2 * a.d.nh + 63) / 64 > 1 || (2 * a.d.na + 63) / 64 > 1) return -1;
    const bool force = a.rec_mode == 2;
    if (force) a.tab_lds = 0;
    if (!force && a.Fpart) return -1;
    const size_t elt = fp64 ? 8 : 4;
    const int wpb = nq_est::est_waves(a, elt);
    *threads = 64 * wpb;
    *lds = est_lds_bytes(a, elt, wpb);
    if (*lds > 160 * 1024) return -1;
    return (force ? 1 : 0) | (a.d.N > 8 ? 2 : 0) | (a.tab_lds ? 4 : 0) | (a.d.act == NQ_ACT_LOGCOSH ? 8 : 0) |
           (fp64 ? 16 : 0) | (force && !fp64 && nq_est::force_class(a.d) == 0 ? 32 : 0);
}

hipError_t nq_pk_estimate(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st) {
    if (L.variant & 16) return nq_est::pk_launch_prec<double>(M, L, st);
    return nq_pk_estimate_f32(M, L, st);
}
