// Device view of the flat connection table (include/nqhip.h, nqh_optable) and
// the wave-uniform walk over the connections of one doubled configuration.
#pragma once
#include "nq_device.h"

struct dev_table {
    int n_sites, n_terms;
    const int32_t* term_kind;
    const int32_t* term_op_l;
    const int32_t* term_op_r;
    const int32_t* op_site_off;
    const int32_t* op_sites;
    const int32_t* op_row_off;
    const int32_t* row_ent_off;
    const double* ent_mel;       // interleaved re, im
    const float* ent_mel_f;      // same values in FP32
    const uint32_t* ent_flip;    // physical flip mask (N <= 32)
    const void* base;            // the one allocation all arrays above live in
    int bytes;                   // its size (multiple of 16)
};

// Copy the table into LDS (all threads of the block take part; the caller
// synchronises) and return a view whose arrays point into the copy.
__device__ __forceinline__ dev_table table_to_lds(const dev_table& g, unsigned char* lds) {
    const uint4* src = reinterpret_cast<const uint4*>(g.base);
    uint4* dst = reinterpret_cast<uint4*>(lds);
    for (int i = threadIdx.x; i < (g.bytes >> 4); i += blockDim.x) dst[i] = src[i];
    const unsigned char* b = reinterpret_cast<const unsigned char*>(g.base);
    auto rb = [&](const void* q) { return lds + (reinterpret_cast<const unsigned char*>(q) - b); };
    dev_table t = g;
    t.term_kind = reinterpret_cast<const int32_t*>(rb(g.term_kind));
    t.term_op_l = reinterpret_cast<const int32_t*>(rb(g.term_op_l));
    t.term_op_r = reinterpret_cast<const int32_t*>(rb(g.term_op_r));
    t.op_site_off = reinterpret_cast<const int32_t*>(rb(g.op_site_off));
    t.op_sites = reinterpret_cast<const int32_t*>(rb(g.op_sites));
    t.op_row_off = reinterpret_cast<const int32_t*>(rb(g.op_row_off));
    t.row_ent_off = reinterpret_cast<const int32_t*>(rb(g.row_ent_off));
    t.ent_mel = reinterpret_cast<const double*>(rb(g.ent_mel));
    t.ent_mel_f = reinterpret_cast<const float*>(rb(g.ent_mel_f));
    t.ent_flip = reinterpret_cast<const uint32_t*>(rb(g.ent_flip));
    return t;
}

__device__ __forceinline__ int table_local_row(const dev_table& t, int o, uint32_t bits) {
    int r = 0;
    const int b = t.op_site_off[o], e = t.op_site_off[o + 1];
    for (int m = b; m < e; ++m) r |= (int)((bits >> t.op_sites[m]) & 1u) << (m - b);
    return r;
}

template <typename T>
__device__ __forceinline__ cplx<T> table_mel(const dev_table& t, int k);
template <>
__device__ __forceinline__ cplx<double> table_mel<double>(const dev_table& t, int k) {
    return cmk(t.ent_mel[2 * k], t.ent_mel[2 * k + 1]);
}
template <>
__device__ __forceinline__ cplx<float> table_mel<float>(const dev_table& t, int k) {
    return cmk(t.ent_mel_f[2 * k], t.ent_mel_f[2 * k + 1]);
}

// Visit the connections of (row bits ks, column bits ke) in the reference's
// order (KLocalLiouvillian.jl:46-52, KLocalOperatorTensor.jl:129-156), skipping
// exact-zero mels as the accumulator does (AccumulatorObsGrad.jl:67).
// All arguments must be wave-uniform; fn(mel, row_flip_mask, col_flip_mask).
template <typename T, class Fn>
__device__ __forceinline__ void table_walk(const dev_table& t, uint32_t ks, uint32_t ke, Fn&& fn) {
    for (int q = 0; q < t.n_terms; ++q) {
        const int kind = t.term_kind[q];
        if (kind == 0 || kind == 3) {
            const int o = t.term_op_l[q];
            const int row = t.op_row_off[o] + table_local_row(t, o, ks);
            for (int k = t.row_ent_off[row]; k < t.row_ent_off[row + 1]; ++k) {
                cplx<T> mel = table_mel<T>(t, k);
                if (mel.re == T(0) && mel.im == T(0)) continue;
                fn(mel, t.ent_flip[k], 0u);
            }
        } else if (kind == 1) {
            const int o = t.term_op_r[q];
            const int row = t.op_row_off[o] + table_local_row(t, o, ke);
            for (int k = t.row_ent_off[row]; k < t.row_ent_off[row + 1]; ++k) {
                cplx<T> mel = table_mel<T>(t, k);
                if (mel.re == T(0) && mel.im == T(0)) continue;
                fn(mel, 0u, t.ent_flip[k]);
            }
        } else {
            const int ol = t.term_op_l[q], orr = t.term_op_r[q];
            const int rl = t.op_row_off[ol] + table_local_row(t, ol, ks);
            const int rr = t.op_row_off[orr] + table_local_row(t, orr, ke);
            for (int k = t.row_ent_off[rl]; k < t.row_ent_off[rl + 1]; ++k) {
                const cplx<T> ml = table_mel<T>(t, k);
                for (int k2 = t.row_ent_off[rr]; k2 < t.row_ent_off[rr + 1]; ++k2) {
                    cplx<T> mel = cmul(ml, table_mel<T>(t, k2));
                    if (mel.re == T(0) && mel.im == T(0)) continue;
                    fn(mel, t.ent_flip[k], t.ent_flip[k2]);
                }
            }
        }
    }
}
