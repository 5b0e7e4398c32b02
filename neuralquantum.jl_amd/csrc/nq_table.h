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
    // per term 8 words (part of the same allocation): kind | nsites_l << 8 |
    // nsites_r << 16, row base of op_l, row base of op_r, 0, the sites of
    // op_l as bytes (2 words), those of op_r (2 words)
    const uint32_t* term_desc;
    int emax;                    // most connections one term yields (row entries, or their product)
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
    t.term_desc = reinterpret_cast<const uint32_t*>(rb(g.term_desc));
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

// The connections of (row bits ks, column bits ke) with the terms spread over
// the lanes (term q on lane q mod 64, ceil(n_terms / 64) rounds): each lane
// gathers its own term's local rows and entries, so a round costs a few
// dependent loads for 64 terms instead of a chain of loads per term.
// Unchanged configurations (no flips) go to diag(mel) in the lane that found
// them; the changed ones are numbered across the wave in (round, entry, lane)
// order and handed to put(slot, mel, ms, me) in their lane.  Returns the
// number of changed connections (wave-uniform).  The same connections and
// mels as table_walk (exact-zero mels skipped), in another order.
// ks, ke wave-uniform; every lane must be active.
template <typename T, class Diag, class Put>
__device__ __forceinline__ int table_walk_lanes(const dev_table& t, uint32_t ks, uint32_t ke, int lane, Diag&& diag,
                                                Put&& put) {
    int nc = 0;
    for (int q0 = 0; q0 < t.n_terms; q0 += 64) {
        const int q = q0 + lane;
        const bool act = q < t.n_terms;
        uint4 d0 = make_uint4(0u, 0u, 0u, 0u), d1 = d0;
        if (act) {
            d0 = reinterpret_cast<const uint4*>(t.term_desc)[2 * q];
            d1 = reinterpret_cast<const uint4*>(t.term_desc)[2 * q + 1];
        }
        const uint32_t kind = d0.x & 255u, nsl = (d0.x >> 8) & 255u, nsr = (d0.x >> 16) & 255u;
        const uint64_t sl = (uint64_t)d1.x | ((uint64_t)d1.y << 32);
        const uint64_t sr = (uint64_t)d1.z | ((uint64_t)d1.w << 32);
        uint32_t rl = 0, rr = 0;
        for (uint32_t j = 0; j < nsl; ++j) rl |= ((ks >> (uint32_t)((sl >> (8 * j)) & 255u)) & 1u) << j;
        for (uint32_t j = 0; j < nsr; ++j) rr |= ((ke >> (uint32_t)((sr >> (8 * j)) & 255u)) & 1u) << j;
        int kl0 = 0, cl = 0, kr0 = 0, cr = 0;
        if (act && kind != 1u) {
            kl0 = t.row_ent_off[d0.y + rl];
            cl = t.row_ent_off[d0.y + rl + 1] - kl0;
        }
        if (act && (kind == 1u || kind == 2u)) {
            kr0 = t.row_ent_off[d0.z + rr];
            cr = t.row_ent_off[d0.z + rr + 1] - kr0;
        }
        const int cnt = kind == 2u ? cl * cr : (kind == 1u ? cr : cl);
        int el = 0, er = 0;   // the pair (row entry, column entry) of a tensor term
        for (int e = 0; e < t.emax; ++e) {
            bool chg = false;
            cplx<T> mel = cmk(T(0), T(0));
            uint32_t ms = 0u, me = 0u;
            if (e < cnt) {
                if (kind == 2u) {
                    mel = cmul(table_mel<T>(t, kl0 + el), table_mel<T>(t, kr0 + er));
                    ms = t.ent_flip[kl0 + el];
                    me = t.ent_flip[kr0 + er];
                    if (++er == cr) { er = 0; ++el; }
                } else if (kind == 1u) {
                    mel = table_mel<T>(t, kr0 + e);
                    me = t.ent_flip[kr0 + e];
                } else {
                    mel = table_mel<T>(t, kl0 + e);
                    ms = t.ent_flip[kl0 + e];
                }
                if (!(mel.re == T(0) && mel.im == T(0))) {
                    if ((ms | me) == 0u) diag(mel);
                    else chg = true;
                }
            }
            const uint64_t m = __ballot(chg);
            if (chg)
                put(nc + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)),
                    mel, ms, me);
            nc += __popcll(m);
        }
    }
    return nc;
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
