// Loss statistics, force finish, the shifted SR solve and the parameter update.
//
//   stats ........ stat_analysis of |L_loc|^2 over chains x slices (stats.jl:52-77)
//   force ........ F = Re(conj(L_loc) dL_loc)/Ns - C Re(Obar)   (BatchedGradSampler.jl:108-118)
//   shift ........ S_ii += eps (sr_shift, SRDirect.jl:60-65) or eps * S_ii
//   Cholesky ..... right-looking, FP64, panels of 32 columns, one launch per
//                  panel: panel transform and trailing update on FP64 MFMA,
//                  look-ahead factorization of the next diagonal block in one
//                  wavefront's registers (cholesky!(Hermitian(S)) + ldiv!,
//                  SRDirect.jl:78-81)
//   update ....... p -= lr * Delta (Descent, src/Optimisers/rules.jl:15-17)
#include "nq_kernels.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int NB = 64;

__global__ void k_fill(double* x, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void k_fill_u32(uint32_t* x, int64_t n, uint32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void k_bump(uint32_t* x) {
    if (threadIdx.x == 0) x[0] += 1u;
}

// an iteration's outcome in one record of pinned host memory: the five loss
// statistics, the factorization's failed-pivot word and the step's
// finiteness flag, so that waiting for the iteration is one stream
// synchronisation and no copies
// the record's last word: out[7] = 1.0 once out[0..6] are visible to the
// host (each writer fences at system scope before the barrier, then one
// release store), so the host may read the record as soon as it sees it
__device__ __forceinline__ void status_done(double* out) {
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(reinterpret_cast<unsigned long long*>(out + 7), 0x3ff0000000000000ull, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void k_status_out(const double* stats, const int* info, const int* flag, double* out) {
    const int t = threadIdx.x;
    if (t < 5) out[t] = stats[t];
    if (t == 5) out[5] = (double)*info;
    if (t == 6) out[6] = (double)*flag;
    status_done(out);
}

template <typename T>
__device__ __forceinline__ double abs2_at(const T* L, int64_t s) {
    const double re = (double)L[2 * s], im = (double)L[2 * s + 1];
    return re * re + im * im;
}

__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < nw; ++k) t += red[k];
    return t;
}

template <typename T>
__device__ __forceinline__ void stats_body(const T* L, int64_t B, int64_t Lc, double* out) {
    __shared__ double red[16];
    const double invL = 1.0 / (double)Lc;
    // pass 1: mean of chain means (= overall mean, chains have equal length)
    double sm = 0.0;
    for (int64_t j = threadIdx.x; j < B; j += blockDim.x) {
        double m = 0.0;
        for (int64_t i = 0; i < Lc; ++i) m += abs2_at(L, j + i * B);
        sm += m * invL;
    }
    const double mu = block_sum(sm, red) / (double)B;
    // pass 2: chain variances, variance of chain means, variance of all values
    double vch = 0.0, vmu = 0.0, vall = 0.0;
    for (int64_t j = threadIdx.x; j < B; j += blockDim.x) {
        double m = 0.0;
        for (int64_t i = 0; i < Lc; ++i) m += abs2_at(L, j + i * B);
        m *= invL;
        double v = 0.0;
        for (int64_t i = 0; i < Lc; ++i) {
            const double x = abs2_at(L, j + i * B);
            v += (x - m) * (x - m);
            vall += (x - mu) * (x - mu);
        }
        vch += v / (double)(Lc - 1);
        vmu += (m - mu) * (m - mu);
    }
    vch = block_sum(vch, red);
    vmu = block_sum(vmu, red);
    vall = block_sum(vall, red);
    if (threadIdx.x == 0) {
        const double var_mu_ch = vmu / (double)(B - 1);
        const double var_mu = vall / (double)(B * Lc - 1);
        const double t = var_mu_ch / var_mu;
        out[0] = mu;
        out[1] = sqrt(var_mu_ch / (double)B);
        out[2] = vch / (double)B;
        out[3] = fmax(0.0, 0.5 * (t * (double)Lc - 1.0));
        out[4] = sqrt((double)(Lc - 1) / (double)Lc + t);
    }
}

// stage 1: part2[g][k] = sum of Fpart rows g*per .. (g+1)*per - 1 (grid: k chunks x groups)
__device__ __forceinline__ void force_part_body(const double* Fpart, int nblk, int per, int np, double* part2, int bx,
                                                int by) {
    const int k = bx * blockDim.x + threadIdx.x;
    const int g = by;
    if (k >= np) return;
    const int b0 = g * per, b1 = (b0 + per < nblk) ? b0 + per : nblk;
    double acc = 0.0;
    for (int b = b0; b < b1; ++b) acc += Fpart[(int64_t)b * np + k];
    part2[(int64_t)g * np + k] = acc;
}

// stage 2: F[k] = sum_g part2[g][k] - C Re(Obar_k)
__device__ __forceinline__ void force_finish_body(const double* part2, int ngroups, const double* mu,
                                                  const double* stats, const ndm_dims& d, double* F, int bx) {
    const int k = bx * blockDim.x + threadIdx.x;
    if (k >= d.np) return;
    double acc = 0.0;
    for (int g = 0; g < ngroups; ++g) acc += part2[(int64_t)g * d.np + k];
    const double reObar = (k < d.nlam + d.npi) ? mu[k] : 0.0;
    F[k] = acc - stats[0] * reObar;
}

__global__ void k_shift(double* S, int n, int lds, double eps, int kind) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double& s = S[(int64_t)k + (int64_t)k * lds];
    s += (kind == 0) ? eps : eps * s;
}

// ---- Cholesky in panels of CB = 32 columns.
//
// Storage during the factorization (S column-major, leading dimension lda,
// nr = n or n + 1 rows, row n carrying the right-hand side):
//   * diagonal blocks: only winv (below); their lower triangle keeps the
//     shifted matrix's values (read by nobody once the panel is factored);
//   * off-diagonal panels L_ik (i > k) written TRANSPOSED into the upper
//     triangle, A[k + c, i] = L[i, k + c]: the lower triangle of a panel is
//     still being read by the other tiles of the same step, the upper
//     triangle is read by nobody until the substitutions;
//   * winv[b] = L_bb^-T (32 x 32, column-major) of every diagonal block.
// Step k (one launch, k_chol_step): every trailing tile (I >= J, 64 x 64)
// stages its rows of the panel A_ik, transforms them to L_ik = A_ik L_kk^-T
// on the FP64 matrix cores, and applies A_IJ -= L_Ik L_Jk^T; the first tile
// of each row writes L_Ik out (upper triangle); tile (0, 0) transforms and
// updates the next diagonal block first, one 16 x 16 piece per wave, and
// factors it in one wavefront (look-ahead, chol_lookahead), so the serial
// column chain of block k + 1 runs while the other tiles of step k update.
// The right-hand side row takes part as an ordinary trailing row, which is
// the forward substitution L y = b; y ends in column n of the upper triangle.
constexpr int CB = 32;

// 1/sqrt(v) from the hardware estimate y refined by one third-order step:
// with e = 1 - v y^2, 1/sqrt(v) = y (1 - e)^(-1/2) = y (1 + e/2 + 3e^2/8 + ...),
// so y (1 + e (1/2 + 3e/8)) leaves a relative error O(e^3) plus rounding
// (~1 ulp): four dependent operations after the estimate instead of six for
// two Newton steps.  v <= 0 yields NaN or inf, which the caller reports
__device__ __forceinline__ double rsqrt_nr(double v) {
    const double y = __builtin_amdgcn_rsq(v);
    const double e = fma(-v * y, y, 1.0);
    return fma(y * e, fma(0.375, e, 0.5), y);
}

// factor the kb x kb block at (k, k) (identity padding to 32) in one
// wavefront: lane t < 32 holds row t of the block, lane 32 + t holds row t of
// the identity appended below it, so the same column sweep that produces
// L_kk turns the identity rows into I L_kk^-T = L_kk^-T (the bordering rule
// L_21 = A_21 L_11^-T with A_21 = I).
// The sweep runs in sub-panels of 8 columns: inside a sub-panel each column
// updates only the sub-panel's later columns, so the serial chain from one
// pivot to the next stays short: the next column at once (its multiplier
// broadcast from its lane by v_readlane), the others one pivot later (their
// multipliers broadcast from LDS, Lc[column][lane], written as each column
// is finished); the rest of the block then takes one rank-8 update from the
// same LDS copy.  Lane j's own entry of column j is v, so r[j] rsqrt(v) is
// L[j][j] = sqrt(v) there and needs no special case.  Writes L_kk^-T to `w`;
// L_kk itself is not stored: the panel transforms and the substitutions use
// L_kk^-T and the transposed off-diagonal panels only, and its 32 masked
// stores cost ~0.85 us of the pivot chain's 4.8 us.
// The block comes from A (global) or, when `src` is given, from LDS
// (src[row][column], rows and columns of the block).  The first diagonal
// block of a column range takes this one-wavefront form (k_chol_diag32);
// the look-ahead tile spreads the same sweep over three wavefronts
// (factor_diag32_split below).
constexpr int SUBP = 8;
template <bool FROM_LDS>
__device__ __forceinline__ void factor_diag32(double* A, int lda, int k, int kb, int nmax, double* w, int* info,
                                              bool report, double (*Lc)[64], const double (*src)[CB + 1]) {
    const int t = threadIdx.x & 63, tr = t & (CB - 1);
    const bool mrow = t < CB;
    double r[CB];
    const int64_t rrow = (int64_t)min(k + tr, nmax - 1);
#pragma unroll
    for (int c = 0; c < CB; ++c) {
        // unconditional load, then select
        const double a = FROM_LDS ? src[tr][c] : A[rrow + (int64_t)min(k + c, nmax - 1) * lda];
        r[c] = (mrow && tr < kb && c < kb && tr >= c) ? a : (tr == c ? 1.0 : 0.0);
    }
    int bad = 0;   // 1 + first column with a non-positive pivot (branch-free inside the sweep)
#pragma unroll
    for (int p0 = 0; p0 < CB; p0 += SUBP) {
        // column j updates the next column at once (on the pivot chain) and
        // the sub-panel's later columns one pivot later, after column j + 1's
        // square root has started; every column still receives its updates
        // in pivot order
        double lprev = 0.0;
#pragma unroll
        for (int j = p0; j < p0 + SUBP; ++j) {
            const double v = rl_d(r[j], j);                      // current A[j][j]
            bad = (bad == 0 && j < kb && !(v > 0.0)) ? j + 1 : bad;
            const double rp = rsqrt_nr(v);
            if (j > p0) {
#pragma unroll
                for (int m = j + 1; m < p0 + SUBP; ++m) r[m] -= lprev * Lc[j - 1][m];
            }
            const double lj = r[j] * rp;   // L[tr][j] (identity rows: (L^-T)[tr][j])
            r[j] = lj;
            Lc[j][t] = lj;
            if (j + 1 < p0 + SUBP) r[j + 1] -= lj * rl_d(lj, j + 1);
            lprev = lj;
        }
        if (p0 + SUBP < CB) {
            __builtin_amdgcn_wave_barrier();
            // column pairs, each pinned in place (otherwise the compiler sinks
            // every column's update to its first use, or hoists all the LDS
            // reads, holding hundreds of operands live); the next pair's
            // operands are loaded before the current pair's products, and each
            // column sums in two chains, so LDS latency and the FMA chains overlap
            double2 lv[2][SUBP];
#pragma unroll
            for (int q = 0; q < SUBP; ++q) lv[0][q] = *reinterpret_cast<const double2*>(&Lc[p0 + q][p0 + SUBP]);
#pragma unroll
            for (int c = p0 + SUBP; c < CB; c += 2) {
                const int b = ((c - p0 - SUBP) >> 1) & 1;
                if (c + 2 < CB) {
#pragma unroll
                    for (int q = 0; q < SUBP; ++q) lv[b ^ 1][q] = *reinterpret_cast<const double2*>(&Lc[p0 + q][c + 2]);
                }
                double a0 = r[c], a1 = r[c + 1], e0 = 0.0, e1 = 0.0;
#pragma unroll
                for (int q = 0; q < SUBP; q += 2) {
                    a0 -= r[p0 + q] * lv[b][q].x;
                    a1 -= r[p0 + q] * lv[b][q].y;
                    e0 -= r[p0 + q + 1] * lv[b][q + 1].x;
                    e1 -= r[p0 + q + 1] * lv[b][q + 1].y;
                }
                r[c] = a0 + e0;
                r[c + 1] = a1 + e1;
                asm volatile("" : "+v"(r[c]), "+v"(r[c + 1]));
            }
        }
    }
    if (report && bad != 0 && t == 0 && *info == 0) *info = k + bad;
    if (!mrow) {
        // (L^-T)[tr][c], column-major with ld 32
#pragma unroll
        for (int c = 0; c < CB; ++c) w[tr + CB * c] = r[c];
    }
}

// ---- the look-ahead tile's diagonal factor over three wavefronts: one
// sweeps sub-panels 0 and 1 (columns 0-15), the others sub-panels 2 and 3;
// while the first runs its pivots, the others apply each finished
// sub-panel's rank-8 update to their own columns, so only the next
// sub-panel's update stays on the pivot chain.  Same arithmetic, in the same
// order, as factor_diag32 (bit-identical L_kk^-T); the waves hand over
// sub-panels through a workgroup counter in LDS (release / acquire).

// bounded wait for the workgroup's sub-panel counter (the producer is a wave
// of the same workgroup, so it always arrives; the bound only keeps a fault
// from hanging the device)
__device__ __forceinline__ bool wait_prog(const int* prog, int v) {
    for (int it = 0; it < (1 << 20); ++it) {
        if (__hip_atomic_load(prog, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) >= v) return true;
        __builtin_amdgcn_s_sleep(1);
    }
    return false;
}

// pivots of sub-panel [p0, p0 + 8) on the register columns r[c - C0]
template <int C0>
__device__ __forceinline__ void sweep8(double* r, int p0, int kb, int t, int& bad, double (*Lc)[64]) {
    double lprev = 0.0;
#pragma unroll
    for (int j = p0; j < p0 + SUBP; ++j) {
        const double v = rl_d(r[j - C0], j);
        bad = (bad == 0 && j < kb && !(v > 0.0)) ? j + 1 : bad;
        const double rp = rsqrt_nr(v);
        if (j > p0) {
#pragma unroll
            for (int m = j + 1; m < p0 + SUBP; ++m) r[m - C0] -= lprev * Lc[j - 1][m];
        }
        const double lj = r[j - C0] * rp;
        r[j - C0] = lj;
        Lc[j][t] = lj;
        if (j + 1 < p0 + SUBP) r[j + 1 - C0] -= lj * rl_d(lj, j + 1);
        lprev = lj;
    }
}

// rank-8 update of the register columns [cbeg, cend) from sub-panel
// [q0, q0 + 8): the sub-panel's entries of this lane's row from registers
// (OWN) or from the LDS copy of the columns
template <int C0, bool OWN>
__device__ __forceinline__ void update8(double* r, int q0, int cbeg, int cend, int t, double (*Lc)[64]) {
    double lt[SUBP];
#pragma unroll
    for (int q = 0; q < SUBP; ++q) lt[q] = OWN ? r[q0 + q - C0] : Lc[q0 + q][t];
    double2 lv[2][SUBP];
#pragma unroll
    for (int q = 0; q < SUBP; ++q) lv[0][q] = *reinterpret_cast<const double2*>(&Lc[q0 + q][cbeg]);
#pragma unroll
    for (int c = cbeg; c < cend; c += 2) {
        const int b = ((c - cbeg) >> 1) & 1;
        if (c + 2 < cend) {
#pragma unroll
            for (int q = 0; q < SUBP; ++q) lv[b ^ 1][q] = *reinterpret_cast<const double2*>(&Lc[q0 + q][c + 2]);
        }
        double a0 = r[c - C0], a1 = r[c + 1 - C0], e0 = 0.0, e1 = 0.0;
#pragma unroll
        for (int q = 0; q < SUBP; q += 2) {
            a0 -= lt[q] * lv[b][q].x;
            a1 -= lt[q] * lv[b][q].y;
            e0 -= lt[q + 1] * lv[b][q + 1].x;
            e1 -= lt[q + 1] * lv[b][q + 1].y;
        }
        r[c - C0] = a0 + e0;
        r[c + 1 - C0] = a1 + e1;
        asm volatile("" : "+v"(r[c - C0]), "+v"(r[c + 1 - C0]));
    }
}

// the columns [C0, C0 + NC) of the kb x kb block src (rows in lanes < 32,
// identity rows in lanes 32-63), as factor_diag32 loads them
template <int C0, int NC>
__device__ __forceinline__ void load_cols(double* r, const double (*src)[CB + 1], int kb, int t) {
    const int tr = t & (CB - 1);
    const bool mrow = t < CB;
#pragma unroll
    for (int c = C0; c < C0 + NC; ++c) {
        const double a = src[tr][c];
        r[c - C0] = (mrow && tr < kb && c < kb && tr >= c) ? a : (tr == c ? 1.0 : 0.0);
    }
}

// one wavefront per part of the look-ahead workgroup: part 0 = sub-panels 0
// and 1, part 2 = sub-panel 2, part 3 = sub-panel 3 (prog = 0 on entry,
// fbad: 2 ints); L_kk^-T of the block at (k, k) to w, the first failing
// pivot to info
__device__ __forceinline__ void factor_diag32_split(int k, int kb, double* w, int* info, double (*Lc)[64],
                                                    const double (*src)[CB + 1], int* prog, int* fbad, int part) {
    const int t = threadIdx.x & 63, tr = t & (CB - 1);
    const bool mrow = t < CB;
    int bad = 0;
    if (part == 0) {
        double r[16];
        load_cols<0, 16>(r, src, kb, t);
        sweep8<0>(r, 0, kb, t, bad, Lc);
        __hip_atomic_store(prog, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        __builtin_amdgcn_wave_barrier();
        update8<0, true>(r, 0, 8, 16, t, Lc);
        sweep8<0>(r, 8, kb, t, bad, Lc);
        if (t == 0) fbad[0] = bad;
        __hip_atomic_store(prog, 2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (!mrow) {
#pragma unroll
            for (int c = 0; c < 16; ++c) w[tr + CB * c] = r[c];
        }
    } else if (part == 2) {
        double r[8];
        load_cols<16, 8>(r, src, kb, t);
        bool ok = wait_prog(prog, 1);
        update8<16, false>(r, 0, 16, 24, t, Lc);
        ok = wait_prog(prog, 2) && ok;
        update8<16, false>(r, 8, 16, 24, t, Lc);
        sweep8<16>(r, 16, kb, t, bad, Lc);
        if (!ok) bad = 17;   // (never: wave 0 always publishes)
        if (t == 0) fbad[1] = bad;
        __hip_atomic_store(prog, 3, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (!mrow) {
#pragma unroll
            for (int c = 16; c < 24; ++c) w[tr + CB * c] = r[c - 16];
        }
    } else if (part == 3) {
        double r[8];
        load_cols<24, 8>(r, src, kb, t);
        bool ok = wait_prog(prog, 1);
        update8<24, false>(r, 0, 24, 32, t, Lc);
        ok = wait_prog(prog, 2) && ok;
        update8<24, false>(r, 8, 24, 32, t, Lc);
        ok = wait_prog(prog, 3) && ok;
        update8<24, false>(r, 16, 24, 32, t, Lc);
        sweep8<24>(r, 24, kb, t, bad, Lc);
        if (!ok) bad = 25;
        if (!mrow) {
#pragma unroll
            for (int c = 24; c < 32; ++c) w[tr + CB * c] = r[c - 24];
        }
        if (t == 0) {
            // the first failing column of the block (waves 0 and 2 published
            // theirs before the counter this wave waited for)
            const int b0 = fbad[0], b2 = fbad[1];
            const int first = b0 ? b0 : (b2 ? b2 : bad);
            if (first != 0 && *info == 0) *info = k + first;
        }
    }
}

// the first diagonal block of a column range [k0, ...): winv block 0;
// blockIdx.y = instance (A and winv at the given strides)
__device__ __forceinline__ void chol_diag32_body(double* A, int n, int lda, int k, int kb, double* winv, int* info,
                                                 int64_t sA, int64_t sW, int by) {
    __shared__ __align__(16) double Ls[CB][64];
    A += by * sA;
    winv += by * sW;
    factor_diag32<false>(A, lda, k, kb, n, winv, info, true, Ls, nullptr);
}

// tiles of a step, column tile by column tile: column tile J (of TJ) holds
// row tiles J .. T - 1, so block b = 0 is tile (0, 0)
__host__ __device__ __forceinline__ int step_tiles(int T, int TJ) { return TJ * T - TJ * (TJ - 1) / 2; }

__device__ __forceinline__ void step_decode(int b, int T, int TJ, int* I, int* J) {
    int j = 0;
    while (j + 1 < TJ && step_tiles(T, j + 1) <= b) ++j;
    *J = j;
    *I = j + (b - step_tiles(T, j));
}

// The look-ahead tile (0, 0) of step k: rows and columns [o, o + 64),
// o = k + kb, whose first 32 rows and columns D are the next diagonal block.
// Its critical path is load -> L_D = A_Dk W (32 x 32) -> D -= L_D L_D^T ->
// factor D, so that part is spread over all four waves, one 16 x 16 tile
// each, with the MFMA operands loaded from memory in their lane layout (no
// LDS staging before the first product).  Wave 3, which has no tile of D to
// update (its quarter is above the diagonal), meanwhile issues the loads of
// the other 32 rows R (panel rows, the block below D, the lower half of the
// block right of it) and then transforms and stores them; waves 0, 1 and 2
// factor D (factor_diag32_split; waves 1 and 2 first store the panel rows
// of D, wave 0 afterwards those of R), all while the pivot chain runs.
// Operand layouts (f64 16x16x4 MFMA): A[i][k] from lane i + 16 k, B[k][j]
// from lane j + 16 k; output D[lane/16 + 4v][lane%16].  Products whose
// result goes to memory are formed transposed (operands swapped), so that a
// lane holds element (lane%16, lane/16 + 4v) of the quarter and a load or
// store instruction covers 16 consecutive rows of 4 columns.
__device__ __forceinline__ void chol_lookahead(double* A, int n, int nr, int ncu, int lda, int k, int kb,
                                               const double* w, double* wnext, int* info, double (*Lr)[CB + 1],
                                               double (*Lr2)[CB + 1], double (*Ld)[CB + 1], double (*Ls)[64],
                                               int* prog) {
    const int o = k + kb, kbn = min(CB, n - o), rb = o + CB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
    if (threadIdx.x == 0) {
        prog[0] = 0;   // the diagonal factor's sub-panel counter
        prog[3] = 0;   // L_R staged in Lr2 (wave 3 -> wave 0); published by the barriers below
    }
    // this wave's transform tile (rows rt, columns ct of L_D) and update tile
    // (ra, ca) of D: waves 0, 1, 2 own (0, 0), (1, 0), (1, 1); wave 3's
    // update tile (0, 1) is above the diagonal
    const int rt = wave >> 1, ct = wave & 1;
    const int ra = (wave == 1 || wave == 2) ? 1 : 0, ca = (wave >= 2) ? 1 : 0;
    double od[4];
    {
        const int row = o + rt * 16 + li;
        const int64_t rc = min(row, nr - 1);
        double xa[CB / 4], wb[CB / 4];
#pragma unroll
        for (int s = 0; s < CB / 4; ++s) {
            const int m = 4 * s + lk;
            xa[s] = A[rc + (int64_t)(k + min(m, kb - 1)) * lda];
            wb[s] = w[m + CB * (ct * 16 + li)];
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
            od[v] = A[(int64_t)min(o + ra * 16 + li, nr - 1) + (int64_t)min(o + ca * 16 + lk + 4 * v, nr - 1) * lda];
        d4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < CB / 4; ++s) {
            const double a = (4 * s + lk < kb && row < nr) ? xa[s] : 0.0;
            t = __builtin_amdgcn_mfma_f64_16x16x4f64(a, wb[s], t, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) Lr[rt * 16 + lk + 4 * v][ct * 16 + li] = t[v];
    }
    __syncthreads();
    double xb[2][CB / 4], wc2[2][CB / 4], q10[2][2][4], q11[2][2][4];
    if (wave == 3) {
        // the rows R: issued while waves 0-2 update D
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int s = 0; s < CB / 4; ++s) {
                const int m = 4 * s + lk;
                xb[h][s] = A[(int64_t)min(rb + h * 16 + li, nr - 1) + (int64_t)(k + min(m, kb - 1)) * lda];
                wc2[h][s] = w[m + CB * (h * 16 + li)];
            }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int64_t r = min(rb + a * 16 + li, nr - 1);
                    q10[a][b][v] = A[r + (int64_t)min(o + b * 16 + lk + 4 * v, nr - 1) * lda];
                    q11[a][b][v] = A[r + (int64_t)min(rb + b * 16 + lk + 4 * v, nr - 1) * lda];
                }
    } else {
        d4 u = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < CB / 4; ++s)
            u = __builtin_amdgcn_mfma_f64_16x16x4f64(Lr[ca * 16 + li][4 * s + lk], Lr[ra * 16 + li][4 * s + lk], u, 0,
                                                     0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int lr = ra * 16 + li, lc = ca * 16 + lk + 4 * v;
            const double val = od[v] - u[v];
            Ld[lr][lc] = val;
            // rows past the diagonal block (the right-hand side row, or rows
            // past n of a partial factorization) are ordinary trailing rows
            const int r = o + lr, c = o + lc;
            if (lr >= kbn && r < nr && c < ncu && r >= c) A[(int64_t)r + (int64_t)c * lda] = val;
        }
    }
    __syncthreads();
    if (wave == 0) {
        factor_diag32_split(o, kbn, wnext, info, Ls, Ld, prog, prog + 1, 0);
        if (rb < nr) {
            // then L_R's panel rows out (transposed), once wave 3 has staged
            // them (a wait that ran out, never expected, fails the factorization)
            if (wait_prog(prog + 3, 1)) {
                for (int e = lane; e < CB * CB; e += 64) {
                    const int c = e & (CB - 1), r = e >> 5;
                    if (c < kb && rb + r < nr) A[(int64_t)(k + c) + (int64_t)(rb + r) * lda] = Lr2[r][c];
                }
            } else if (lane == 0 && *info == 0) {
                *info = o + 1;
            }
        }
        return;
    }
    if (wave <= 2) {
        // L_D out, transposed into the upper triangle, then sub-panels 3
        // (wave 1) and 2 (wave 2) of the next diagonal block
        for (int e = (int)threadIdx.x - 64; e < CB * CB; e += 128) {
            const int c = e & (CB - 1), r = e >> 5;
            if (c < kb && o + r < nr) A[(int64_t)(k + c) + (int64_t)(o + r) * lda] = Lr[r][c];
        }
        factor_diag32_split(o, kbn, wnext, info, Ls, Ld, prog, prog + 1, wave == 1 ? 3 : 2);
        return;
    }
    // wave 3: rows R = [o + 32, o + 64)
    if (rb >= nr) return;
    // L_R = A_Rk W into Lr2 (its panel rows are stored by wave 0)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            d4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < CB / 4; ++s) {
                const double x = (4 * s + lk < kb && rb + a * 16 + li < nr) ? xb[a][s] : 0.0;
                t = __builtin_amdgcn_mfma_f64_16x16x4f64(x, wc2[b][s], t, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) Lr2[a * 16 + lk + 4 * v][b * 16 + li] = t[v];
        }
    __hip_atomic_store(prog + 3, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    __builtin_amdgcn_wave_barrier();
    // the block below D (rows R, columns [o, o + 32)) and the lower half of
    // the block R x R (transposed products: element (li, lk + 4 v))
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            d4 u = {0.0, 0.0, 0.0, 0.0}, x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < CB / 4; ++s) {
                const double ar = Lr2[a * 16 + li][4 * s + lk];
                u = __builtin_amdgcn_mfma_f64_16x16x4f64(Lr[b * 16 + li][4 * s + lk], ar, u, 0, 0, 0);
                if (b <= a) x = __builtin_amdgcn_mfma_f64_16x16x4f64(Lr2[b * 16 + li][4 * s + lk], ar, x, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = rb + a * 16 + li;
                const int c0 = o + b * 16 + lk + 4 * v, c1 = rb + b * 16 + lk + 4 * v;
                if (r < nr && c0 < ncu) A[(int64_t)r + (int64_t)c0 * lda] = q10[a][b][v] - u[v];
                if (b <= a && r < nr && c1 < ncu && r >= c1) A[(int64_t)r + (int64_t)c1 * lda] = q11[a][b][v] - x[v];
            }
        }
}

// step k: panel transform + trailing update of the lower tiles (I >= J) of
// the rows >= o = k + kb and the columns [o, ncu); n = columns of the
// matrix, nr = rows; columns [k0, n) are factored (n <= nr: a partial
// factorization leaves the Schur complement of the trailing rows / columns
// [n, nr), updated here when ncu = nr; with ncu = n those columns are left
// alone and only the panels' rows below are transformed, for one rank-n
// update afterwards); winv block (k - k0) / 32 holds panel k; blockIdx.y =
// instance (A, winv at strides)
__device__ __forceinline__ void chol_step_body(double* A, int n, int nr, int ncu, int lda, int k, int kb, double* winv,
                                               int* info, int k0, int64_t sA, int64_t sW, int bx, int by) {
    __shared__ double Xi[CB][65];               // staged A_Ik: [m][row], then L_Ik
    __shared__ double Xj[CB][65];               // the same for A_Jk (off-diagonal tiles only)
    __shared__ double Wt[CB][CB + 1];           // L_kk^-T
    __shared__ __align__(16) double Ls[CB][64];
    __shared__ double Ld[CB][CB + 1];           // the next diagonal block (look-ahead tile)
    __shared__ int prog[4];                     // its factor's sub-panel counter and failed pivots
    A += by * sA;
    winv += by * sW;
    const int o = k + kb;
    const double* w = winv + (size_t)((k - k0) / CB) * CB * CB;
    if (bx == 0 && o < n) {
        // tile (0, 0): the look-ahead tile (Xi and Xj serve as its LDS)
        chol_lookahead(A, n, nr, ncu, lda, k, kb, w, winv + (size_t)((o - k0) / CB) * CB * CB, info,
                       reinterpret_cast<double (*)[CB + 1]>(&Xi[0][0]), reinterpret_cast<double (*)[CB + 1]>(&Xj[0][0]),
                       Ld, Ls, prog);
        return;
    }
    int tI, tJ;
    {
        const int T = (nr - k - kb + 63) / 64, TJ = max(1, (ncu - k - kb + 63) / 64);
        step_decode(bx, T, TJ, &tI, &tJ);   // tJ <= tI
    }
    const int i0 = o + tI * 64, j0 = o + tJ * 64;
    const bool dg = tI == tJ;                   // diagonal tile: A_Jk = A_Ik
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wr = wave >> 1, wc = wave & 1;
    // every load of the kernel is issued up front, unconditionally (clamped
    // addresses, out-of-range values replaced afterwards), so the round trips
    // overlap instead of serializing one wait per guarded load
    // this lane's 16 entries of the tile A_IJ, in the layout of the
    // transposed product (row li, column lk + 4 v of each 16 x 16 quarter):
    // 16 consecutive rows of 4 columns per load, 4 cache lines, where the
    // product's own layout would touch 16
    double old[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = min(i0 + wr * 32 + a * 16 + li, nr - 1);
                const int c = min(j0 + wc * 32 + b * 16 + lk + 4 * v, nr - 1);
                old[a][b][v] = A[(int64_t)r + (int64_t)c * lda];
            }
    double si[64 * CB / 256], sj[64 * CB / 256], sw[CB * CB / 256];
#pragma unroll
    for (int q = 0; q < 64 * CB / 256; ++q) {
        const int e = threadIdx.x + 256 * q;
        const int r = e & 63, m = e >> 6;
        const int64_t col = (int64_t)(k + min(m, kb - 1)) * lda;
        si[q] = A[(int64_t)min(i0 + r, nr - 1) + col];
        if (!dg) sj[q] = A[(int64_t)min(j0 + r, nr - 1) + col];
    }
#pragma unroll
    for (int q = 0; q < CB * CB / 256; ++q) sw[q] = w[threadIdx.x + 256 * q];
#pragma unroll
    for (int q = 0; q < 64 * CB / 256; ++q) {
        const int e = threadIdx.x + 256 * q;
        const int r = e & 63, m = e >> 6;
        Xi[m][r] = (m < kb && i0 + r < nr) ? si[q] : 0.0;
        if (!dg) Xj[m][r] = (m < kb && j0 + r < nr) ? sj[q] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < CB * CB / 256; ++q) {
        const int e = threadIdx.x + 256 * q;
        Wt[e & (CB - 1)][e >> 5] = sw[q];
    }
    __syncthreads();
    // L_Ik = A_Ik W and L_Jk = A_Jk W (W = L_kk^-T): 2 x (4 x 2) output tiles
    // of 16 x 16, two per wave and side (one side on diagonal tiles); f64 MFMA
    // output D[lane/16 + 4v][lane%16]; the results replace the staged panels
    // once every wave has read them
    // (wave = row tile rt of 4, q = column tile ct of 2).  Every operand is
    // read from LDS first and the two cases run as separate branch-free
    // loops: a per-step test of dg between the products makes the compiler
    // move an accumulator between AGPRs and VGPRs and wait out every product
    d4 ai[2], aj[2];
    {
        double xi[CB / 4], xj[CB / 4], b0[CB / 4], b1[CB / 4];
#pragma unroll
        for (int s = 0; s < CB / 4; ++s) {
            const int m = 4 * s + lk;
            xi[s] = Xi[m][wave * 16 + li];
            xj[s] = Xj[m][wave * 16 + li];   // unused on diagonal tiles
            b0[s] = Wt[m][li];
            b1[s] = Wt[m][16 + li];
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) ai[q] = aj[q] = d4{0.0, 0.0, 0.0, 0.0};
        if (dg) {
#pragma unroll
            for (int s = 0; s < CB / 4; ++s) {
                ai[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[s], b0[s], ai[0], 0, 0, 0);
                ai[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[s], b1[s], ai[1], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int s = 0; s < CB / 4; ++s) {
                ai[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[s], b0[s], ai[0], 0, 0, 0);
                ai[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[s], b1[s], ai[1], 0, 0, 0);
                aj[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(xj[s], b0[s], aj[0], 0, 0, 0);
                aj[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(xj[s], b1[s], aj[1], 0, 0, 0);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int tile = wave * 2 + q;
        const int rt = tile >> 1, ct = tile & 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Xi[ct * 16 + li][rt * 16 + lk + 4 * v] = ai[q][v];
            if (!dg) Xj[ct * 16 + li][rt * 16 + lk + 4 * v] = aj[q][v];
        }
    }
    __syncthreads();
    double (*Pi)[65] = Xi;
    double (*Pj)[65] = dg ? Xi : Xj;
    // the quarter above the diagonal of a diagonal tile is never stored: its
    // wave has no update to make
    if (dg && wr == 0 && wc == 1) return;
    // A_IJ -= L_Ik L_Jk^T (lower part)
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[a][b][v] = 0.0;
#pragma unroll
    for (int m = 0; m < CB; m += 4) {
        double av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) av[a] = Pi[m + lk][wr * 32 + a * 16 + li];
#pragma unroll
        for (int b = 0; b < 2; ++b) bv[b] = Pj[m + lk][wc * 32 + b * 16 + li];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[b], av[a], acc[a][b], 0, 0, 0);
    }
    // (L_J L_I^T: the transposed quarter, element (li, lk + 4 v))
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = i0 + wr * 32 + a * 16 + li;
                const int c = j0 + wc * 32 + b * 16 + lk + 4 * v;
                if (r < nr && c < ncu && r >= c) A[(int64_t)r + (int64_t)c * lda] = old[a][b][v] - acc[a][b][v];
            }
    if (tJ == 0) {
        // L_Ik out, transposed into the upper triangle: A[k + c, i] = L[i, k + c],
        // by the first tile of the row of tiles: all four waves, or those still
        // running in a diagonal tile (0, 2 and 3)
        const int t0 = dg ? (wave == 0 ? lane : (int)threadIdx.x - 64) : (int)threadIdx.x;
        const int nt = dg ? 192 : 256;
        for (int e = t0; e < 64 * CB; e += nt) {
            const int c = e & (CB - 1), r = e >> 5;
            if (c < kb && i0 + r < nr) A[(int64_t)(k + c) + (int64_t)(i0 + r) * lda] = Pi[c][r];
        }
    }
}

// ---- substitutions with the stored factor
// Every load of a step is issued up front (clamped addresses, unused values
// zeroed), so a step costs about one memory round trip rather than one per
// guarded load.
//
// back substitution over a group of up to BG panels [k, kend) in one launch.
// Every workgroup first solves the group itself (redundantly, all in LDS and
// registers): for p = last .. 0, x_p = L_pp^-T y_p, then y_q -= L_pq^T x_p for
// the group's earlier panels q < p (the blocks A[q rows, p columns] of the
// upper triangle, staged once); then rows i < k take y_i -= sum_m
// L[k + m, i] x_m over all the group's columns (BR rows per workgroup, LPR
// lanes per row with BG CB / LPR columns each), their factor entries loaded
// before the group solve so that the two overlap.  One launch instead of BG.
constexpr int BG = 6, LPR = 8, BR = 256 / LPR, BCW = BG * CB / LPR;
__device__ __forceinline__ void back_group_body(const double* A, int lda, int k, int kend, int n, const double* winv,
                                                double* y, double* x, int k0, int64_t sA, int64_t sW, int64_t sY,
                                                int bx, int by) {
    __shared__ double Lq[BG * (BG - 1) / 2][CB][CB + 1];   // (p, q) block pairs, q < p: [row in q][column in p]
    __shared__ double ys[BG * CB], xs[BG * CB];
    A += by * sA;
    y += by * sY;
    x += by * sY;
    const int t = threadIdx.x;
    const int ng = kend - k;                       // columns of the group (<= BG * CB)
    const int npan = (ng + CB - 1) / CB;
    const int i = bx * BR + t / LPR, h = t % LPR;   // row below, its share of the columns
    // every load is issued before the first use: the group's upper-triangle
    // blocks (q < p), this thread's row of L_pp^-T (threads 32p .. 32p + 31),
    // y of the group, then the factor entries A[i, k + m] of the rows below
    // (clamped; unused ones multiply x = 0)
    constexpr int NQB = BG * (BG - 1) / 2, PER = CB * CB / 256;
    double lq[NQB][PER];
#pragma unroll
    for (int p = 1; p < BG; ++p)
#pragma unroll
        for (int q = 0; q < p; ++q)
#pragma unroll
            for (int v = 0; v < PER; ++v) {
                const int e = v * 256 + t, r = e & (CB - 1), m = e >> 5;
                lq[p * (p - 1) / 2 + q][v] =
                    A[(int64_t)min(k + q * CB + r, n - 1) + (int64_t)min(k + p * CB + m, n - 1) * lda];
            }
    const int pt = t / CB, c = t & (CB - 1);
    double wv[CB];
    if (pt < npan) {
        const double* w = winv + by * sW + (size_t)((k + pt * CB - k0) / CB) * CB * CB;
#pragma unroll
        for (int m = 0; m < CB; ++m) wv[m] = w[c + CB * m];
    }
    const double yv = y[k + min(t, ng - 1)];
    const int64_t ic = (int64_t)min(i, max(k - 1, 0));
    double a[BCW];
#pragma unroll
    for (int m = 0; m < BCW; ++m) a[m] = A[ic + (int64_t)min(k + h * BCW + m, n - 1) * lda];
#pragma unroll
    for (int p = 1; p < BG; ++p)
#pragma unroll
        for (int q = 0; q < p; ++q)
#pragma unroll
            for (int v = 0; v < PER; ++v) {
                const int e = v * 256 + t, r = e & (CB - 1), m = e >> 5;
                Lq[p * (p - 1) / 2 + q][r][m] = (p < npan) ? lq[p * (p - 1) / 2 + q][v] : 0.0;
            }
    if (t < BG * CB) ys[t] = (t < ng) ? yv : 0.0;
    __syncthreads();
    // the dot products of the group solve are on the serial chain: four
    // partial sums each (a dependent chain of CB / 4 FMAs, not CB)
    for (int p = npan - 1; p >= 0; --p) {
        if (pt == p) {
            double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int m = 0; m < CB; ++m) s[m & 3] = fma(wv[m], ys[p * CB + m], s[m & 3]);
            xs[t] = (t < ng) ? (s[0] + s[1]) + (s[2] + s[3]) : 0.0;
        }
        __syncthreads();
        if (t < p * CB) {
            const int q = t / CB, r = t & (CB - 1), b = p * (p - 1) / 2 + q;
            double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int m = 0; m < CB; ++m) s[m & 3] = fma(Lq[b][r][m], xs[p * CB + m], s[m & 3]);
            ys[t] -= (s[0] + s[1]) + (s[2] + s[3]);
        }
        __syncthreads();
    }
    if (bx == 0 && t < ng) x[k + t] = xs[t];
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < BCW; ++m) acc = fma(a[m], h * BCW + m < ng ? xs[h * BCW + m] : 0.0, acc);
#pragma unroll
    for (int d = 1; d < LPR; d <<= 1) acc += __shfl_xor(acc, d, 64);
    if (h == 0 && i < k) y[i] -= acc;
}

// forward step (refinement solves): z_k = L_kk^-1 y_k, then y_i -= sum_m L[i, k + m] z_m
// for i >= k + kb, the factor read from the upper triangle A[k + m, i]
__global__ __launch_bounds__(256) void k_fwd_step(const double* A, int n, int lda, int k, int kb, const double* winv,
                                                  double* y, double* z) {
    __shared__ double ys[CB], zs[CB];
    const double* w = winv + (size_t)(k / CB) * CB * CB;
    const int i = k + kb + blockIdx.x * blockDim.x + threadIdx.x;
    const double* col = A + (int64_t)min(i, n - 1) * lda + k;
    double a[CB];
#pragma unroll
    for (int m = 0; m < CB; ++m) a[m] = col[min(m, kb - 1)];
    const int c = threadIdx.x & (CB - 1);
    const double yv = y[k + min(c, kb - 1)];
    double wv[CB];
    if (threadIdx.x < CB) {
#pragma unroll
        for (int m = 0; m < CB; ++m) wv[m] = w[m + CB * c];   // (L^-1)[c][m] = (L^-T)[m][c]
        ys[c] = (c < kb) ? yv : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < CB) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < CB; ++m) s += wv[m] * ys[m];
        zs[c] = (c < kb) ? s : 0.0;
        if (blockIdx.x == 0 && c < kb) z[k + c] = s;
    }
    __syncthreads();
    if (i >= n) return;
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < CB; ++m) acc += a[m] * zs[m];
    y[i] -= acc;
}

// ---- split factorization of the NDM's S.  In P order S = [[Sll, Slp, 0],
// [Spl, Spp, Spm], [0, Smp, Smm]] (lambda, Pi, mu; the lambda-mu block is
// zero: O_lambda is real and O_mu imaginary).  Instance 0 = the [lambda | Pi]
// block and instance 1 = the [mu | Pi] block (mu first), each with its own
// right-hand side row.  Both factor their first nl columns in the same
// launches, with updates to their lambda / mu columns only: that gives the
// lambda and mu rows of L for every Pi column (and the right-hand side),
// U_s = L_s[Pi, 0:nl]^T, stored transposed in the upper triangles.  The Pi
// block's Schur complement in the whole matrix is then S_PiPi - U_0^T U_0 -
// U_1^T U_1, one rank-2nl update on the matrix cores (k_split_schur) in
// place of nl / 32 rank-32 updates of both instances' Pi blocks, and
// instance 0 factors it.  The serial pivot chain is nl + npi long instead of
// 2 nl + npi.
__device__ __forceinline__ int split_p(int inst, int i, int nl, int npi) {
    return (inst == 0 || i >= nl) ? i : nl + npi + i;
}

// lower triangle (with the shift on the diagonal) and the right-hand side row
// of both instances (m = nl + npi, lda = m + 1; instance 1 only its mu
// columns: its Pi columns are never read); grid (m columns, 2 instances)
// one workgroup per column j of an instance (rows j .. m).  Workgroup j = 0
// also factors the instance's first diagonal block (the same values, taken
// from S) into L_00^-T (winv block 0 of the instance), which the first
// factorization step needs, in place of a launch of its own
__device__ __forceinline__ void split_copy_body(const double* S, int lds, const double* F, int nl, int npi, double eps,
                                                int kind, double* A2, int lda, int j, int inst, double* winv,
                                                int64_t sW, int* info) {
    const int m = nl + npi;
    if (inst == 1 && j >= nl) return;
    double* A = A2 + (int64_t)inst * lda * lda;
    // the diagonal block first, so that its factor runs beside the other workgroups' copies
    if (j == 0) {
        __shared__ double Db[CB][CB + 1];
        __shared__ __align__(16) double Lc[CB][64];
        const int kb = nl < CB ? nl : CB;
        for (int e = threadIdx.x; e < CB * CB; e += blockDim.x) {
            const int r = e & (CB - 1), c = e >> 5;
            double v = 0.0;
            if (r < kb && c < kb && r >= c) {
                v = S[(int64_t)split_p(inst, r, nl, npi) + (int64_t)split_p(inst, c, nl, npi) * lds];
                if (r == c) v += (kind == 0) ? eps : eps * v;
            }
            Db[r][c] = v;
        }
        __syncthreads();
        if (threadIdx.x < 64) factor_diag32<true>(A, lda, 0, kb, nl, winv + inst * sW, info, true, Lc, Db);
    }
    // (in workgroup 0 the waves other than the factoring one copy the column)
    const int t0 = j == 0 ? (int)threadIdx.x - 64 : (int)threadIdx.x, nt = j == 0 ? (int)blockDim.x - 64 : blockDim.x;
    const int pj = split_p(inst, j, nl, npi);
    for (int i = j + t0; t0 >= 0 && i <= m; i += nt) {
        double v;
        if (i == m) {
            v = F[pj];
        } else {
            v = S[(int64_t)split_p(inst, i, nl, npi) + (int64_t)pj * lds];
            if (i == j) v += (kind == 0) ? eps : eps * v;
        }
        A[(int64_t)i + (int64_t)j * lda] = v;
    }
}

// instance 0's Pi rows and columns [nl, m] (the right-hand side row m
// included) -= U_0^T U_0 + U_1^T U_1, U_s = A_s[0:nl, nl:m+1] (row c of U_s:
// L_s[., c] of the Pi rows, from the upper triangle).  One 32 x 32 tile of
// the lower triangle per workgroup (tiles numbered as the factorization
// steps number theirs), each wave one 16 x 16 quarter over all 2 nl terms
// on the FP64 matrix cores.  The terms come in chunks of KC rows c: each
// chunk's two 32-column slices of U are read coalesced (64 consecutive rows
// of a column per wave) into registers PF chunks ahead, and staged in LDS
// for the products, so PF memory round trips are in flight while a chunk is
// multiplied.  Three chunks ahead keep the kernel within 168 VGPRs, three
// workgroups per CU, so that 768 tiles (the 4x4 XYZ network's Pi block has
// 595) run in one round.  f64 MFMA operands: A[i][k] from lane i + 16 k,
// B[k][j] from lane j + 16 k; output D[lane/16 + 4v][lane%16]
constexpr int KC = 64, PF = 3;
__device__ __forceinline__ void split_schur_body(double* A2, int lda, int nl, int npi, int bx, double* winvP,
                                                 int* info) {
    __shared__ __align__(16) double Xa[32][KC + 1];   // [tile column][chunk row]
    __shared__ __align__(16) double Xb[32][KC + 1];
    const int m = nl + npi, T = (npi + 1 + 31) / 32;
    int tI, tJ;
    step_decode(bx, T, T, &tI, &tJ);   // tJ <= tI
    const int i0 = nl + 32 * tI, j0 = nl + 32 * tJ;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, li = lane & 15, lk = lane >> 4;
    const int rt = wave >> 1, ct = wave & 1;
    const int cpi = (nl + KC - 1) / KC, nch = 2 * cpi;   // chunks per instance / in all
    const int r = t & (KC - 1), c4 = t >> 6;             // this thread's chunk row; columns c4 + 4 q
    // raw loads from clamped addresses; the rows past nl (and the chunks
    // past nch) are zeroed when staged, so no select waits on a load
    double va[PF][8], vb[PF][8];
    bool ok[PF];
    auto fetch = [&](int ch, double* a, double* b, bool& okc) {
        const int s = ch >= cpi ? 1 : 0;
        const int c = (ch - s * cpi) * KC + r;
        const double* U = A2 + (int64_t)s * lda * lda + min(c, nl - 1);
        okc = c < nl && ch < nch;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            a[q] = U[(int64_t)min(i0 + c4 + 4 * q, m) * lda];
            b[q] = U[(int64_t)min(j0 + c4 + 4 * q, m) * lda];
        }
    };
#pragma unroll
    for (int u = 0; u < PF; ++u) fetch(u, va[u], vb[u], ok[u]);
    d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    // whole rounds of PF chunks (the chunks past nch are zero), so that no
    // branch separates a chunk's loads from its use
    for (int ch0 = 0; ch0 < nch; ch0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            __syncthreads();   // every wave is done with the previous chunk
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                Xa[c4 + 4 * q][r] = ok[u] ? va[u][q] : 0.0;
                Xb[c4 + 4 * q][r] = ok[u] ? vb[u][q] : 0.0;
            }
            __syncthreads();
            fetch(ch0 + u + PF, va[u], vb[u], ok[u]);   // in flight during the next PF chunks
#pragma unroll
            for (int kk = 0; kk < KC; kk += 8) {
                acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(Xa[rt * 16 + li][kk + lk], Xb[ct * 16 + li][kk + lk], acc0,
                                                            0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(Xa[rt * 16 + li][kk + 4 + lk],
                                                            Xb[ct * 16 + li][kk + 4 + lk], acc1, 0, 0, 0);
            }
        }
    }
    // tile (0, 0) is the Pi block's first diagonal block: its workgroup
    // factors it into L^-T (winv block 0 of the Pi columns) from an LDS copy
    // of the updated values, in place of a launch of its own
    const bool diag = tI == 0 && tJ == 0;
    double (*Db)[CB + 1] = reinterpret_cast<double (*)[CB + 1]>(&Xb[0][0]);
    if (diag) __syncthreads();   // every wave is done with the last chunk
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = i0 + rt * 16 + lk + 4 * v, j = j0 + ct * 16 + li;
        if (i <= m && j <= m && i >= j) {
            const double nv = A2[(int64_t)i + (int64_t)j * lda] - (acc0[v] + acc1[v]);
            A2[(int64_t)i + (int64_t)j * lda] = nv;
            if (diag) Db[i - i0][j - j0] = nv;
        }
    }
    if (diag) {
        __syncthreads();
        const int kb = npi < CB ? npi : CB;
        if (wave == 0)
            factor_diag32<true>(A2, lda, nl, kb, m, winvP, info, true, reinterpret_cast<double (*)[64]>(&Xa[0][0]),
                                Db);
    }
}

// y1[i] -= sum_{c in [nl, m)} A1[i, c] x0[c] for i < nl (the mu rows' share of
// the Pi solution; A1's upper triangle holds L[c, i] transposed), in two
// fixed-order stages: k_split_upd gives every (64-row block, column chunk) pair
// a workgroup (64 rows x 4 column groups) and writes part[chunk][i];
// k_split_upd_fin sums the chunks in order
constexpr int kUpdChunks = 16;

__device__ __forceinline__ void split_upd_body(const double* A1, int lda, int nl, int m, const double* x0, double* part,
                                               int bx, int by) {
    __shared__ double red[4][64];
    const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int i = bx * 64 + r;
    const int ic = i < nl ? i : nl - 1;
    const int w = (m - nl + kUpdChunks - 1) / kUpdChunks;
    const int c0 = nl + by * w, c1 = min(m, c0 + w);
    double a0 = 0.0, a1 = 0.0;
    int c = c0 + g;
    for (; c + 4 < c1; c += 8) {
        a0 += A1[(int64_t)ic + (int64_t)c * lda] * x0[c];
        a1 += A1[(int64_t)ic + (int64_t)(c + 4) * lda] * x0[c + 4];
    }
    if (c < c1) a0 += A1[(int64_t)ic + (int64_t)c * lda] * x0[c];
    red[g][r] = a0 + a1;
    __syncthreads();
    if (g == 0 && i < nl) part[(int64_t)by * nl + i] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
}

__device__ __forceinline__ void split_upd_fin_body(const double* part, int nl, double* y1, int bx) {
    const int i = bx * blockDim.x + threadIdx.x;
    if (i >= nl) return;
    double t = 0.0;
    for (int q = 0; q < kUpdChunks; ++q) t += part[(int64_t)q * nl + i];
    y1[i] -= t;
}

// delta in P order from the two instances' solutions
__device__ __forceinline__ void split_gather_body(const double* x0, const double* x1, int nl, int npi, int np,
                                                  double* delta, int bx) {
    const int p = bx * blockDim.x + threadIdx.x;
    if (p >= np) return;
    delta[p] = (p < nl + npi) ? x0[p] : x1[p - nl - npi];
}

// A[n, j] = b[j] (the right-hand side as the extra row of the factorization)
__global__ void k_set_row(double* A, int n, int lda, const double* b) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) A[(int64_t)n + (int64_t)j * lda] = b[j];
}

// y[j] = A[j, n]: the forward-substituted right-hand side (upper column n)
__device__ __forceinline__ void copy_col_body(const double* A, int n, int lda, double* y, int bx) {
    const int j = bx * blockDim.x + threadIdx.x;
    if (j < n) y[j] = A[(int64_t)j + (int64_t)n * lda];
}

__global__ void k_residual(const double* S, int n, int lds, double eps, int kind, const double* x,
                           const double* b, double* r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc += S[(int64_t)i + (int64_t)j * lds] * x[j];
    const double sii = S[(int64_t)i + (int64_t)i * lds];
    acc += ((kind == 0) ? eps : eps * sii) * x[i];
    r[i] = b[i] - acc;
}

__global__ void k_axpy(double* x, const double* y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += y[i];
}

// P-order index -> reference flat index
__device__ __forceinline__ int p2ref(int k, const ndm_dims& d) {
    const int N = d.N, nh = d.nh, na = d.na;
    if (k < N) return d.bla + k;
    k -= N;
    if (k < nh) return d.hla + k;
    k -= nh;
    if (k < nh * N) return d.wla + k;
    k -= nh * N;
    if (k < na) return d.dla + k;
    k -= na;
    if (k < na * N) return d.ula + k;
    k -= na * N;
    if (k < na * N) return d.umu + k;
    k -= na * N;
    if (k < N) return d.bmu + k;
    k -= N;
    if (k < nh) return d.hmu + k;
    k -= nh;
    return d.wmu + k;
}

// the step is applied only when the solve succeeded: no failed pivot (info,
// PosDefException in SRDirect.jl:79 is thrown before any update), a finite
// step (flag) and a finite loss; otherwise the parameters stay as they were
__device__ __forceinline__ bool step_ok(const nq_step_guard& g) {
    return (!g.info || *g.info == 0) && (!g.flag || *g.flag == 0) && (!g.stats || isfinite(g.stats[0]));
}

__device__ __forceinline__ void update_body(double* p64, float* p32, const double* delta, const ndm_dims& d, double lr,
                                            const nq_step_guard& g, int bx) {
    const int k = bx * blockDim.x + threadIdx.x;
    if (k >= d.np || !step_ok(g)) return;
    const int r = p2ref(k, d);
    const double v = p64[r] - lr * delta[k];
    p64[r] = v;
    if (p32) p32[r] = (float)v;
}

__global__ void k_to_f32(const double* p64, float* p32, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p32[i] = (float)p64[i];
}

__device__ __forceinline__ void check_finite_body(const double* x, int64_t n, int* flag, int bx) {
    int64_t i = (int64_t)bx * blockDim.x + threadIdx.x;
    if (i < n && !isfinite(x[i])) atomicOr(flag, 1);
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---- the kernels (one problem per launch)
template <typename T>
__global__ __launch_bounds__(1024) void k_stats(const T* L, int64_t B, int64_t Lc, double* out) {
    stats_body<T>(L, B, Lc, out);
}

__global__ void k_force_part(const double* Fpart, int nblk, int per, int np, double* part2) {
    force_part_body(Fpart, nblk, per, np, part2, blockIdx.x, blockIdx.y);
}

__global__ void k_force_finish(const double* part2, int ngroups, const double* mu, const double* stats,
                               ndm_dims d, double* F) {
    force_finish_body(part2, ngroups, mu, stats, d, F, blockIdx.x);
}

__global__ __launch_bounds__(64) void k_chol_diag32(double* A, int n, int lda, int k, int kb, double* winv, int* info,
                                                    int64_t sA, int64_t sW) {
    chol_diag32_body(A, n, lda, k, kb, winv, info, sA, sW, blockIdx.y);
}

// the instances' tiles interleaved in one grid row (block b: tile b / ninst of
// instance b % ninst), so that every instance's look-ahead tile is among the
// first workgroups dispatched
__global__ __launch_bounds__(256) void k_chol_step(double* A, int n, int nr, int ncu, int lda, int k, int kb,
                                                   double* winv, int* info, int k0, int64_t sA, int64_t sW,
                                                   int ninst) {
    chol_step_body(A, n, nr, ncu, lda, k, kb, winv, info, k0, sA, sW, blockIdx.x / ninst, blockIdx.x % ninst);
}

__global__ __launch_bounds__(256) void k_back_group(const double* A, int lda, int k, int kend, int n,
                                                    const double* winv, double* y, double* x, int k0, int64_t sA,
                                                    int64_t sW, int64_t sY) {
    back_group_body(A, lda, k, kend, n, winv, y, x, k0, sA, sW, sY, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(256) void k_split_copy(const double* S, int lds, const double* F, int nl, int npi,
                                                    double eps, int kind, double* A2, int lda, double* winv,
                                                    int64_t sW, int* info) {
    split_copy_body(S, lds, F, nl, npi, eps, kind, A2, lda, blockIdx.x, blockIdx.y, winv, sW, info);
}

__global__ __launch_bounds__(256, 3) void k_split_schur(double* A2, int lda, int nl, int npi, double* winvP,
                                                        int* info) {
    split_schur_body(A2, lda, nl, npi, blockIdx.x, winvP, info);
}

__global__ __launch_bounds__(256) void k_split_upd(const double* A1, int lda, int nl, int m, const double* x0,
                                                   double* part) {
    split_upd_body(A1, lda, nl, m, x0, part, blockIdx.x, blockIdx.y);
}

__global__ void k_split_upd_fin(const double* part, int nl, double* y1) { split_upd_fin_body(part, nl, y1, blockIdx.x); }

__global__ void k_copy_col(const double* A, int n, int lda, double* y) { copy_col_body(A, n, lda, y, blockIdx.x); }

__global__ void k_update(double* p64, float* p32, const double* delta, ndm_dims d, double lr, nq_step_guard g) {
    update_body(p64, p32, delta, d, lr, g, blockIdx.x);
}

__global__ void k_check_finite(const double* x, int64_t n, int* flag) { check_finite_body(x, n, flag, blockIdx.x); }

// the right-hand side columns of both split instances (blockIdx.y)
__global__ void k_copy_col2(const double* A, int n, int lda, int64_t sA, double* y, int64_t sY) {
    copy_col_body(A + blockIdx.y * sA, n, lda, y + blockIdx.y * sY, blockIdx.x);
}

// the split solve's tail in one workgroup, with no launch between its parts:
// delta in P order from the two instances' solutions (split_gather_body),
// *flag = 1 if any entry is not finite, else 0 (as the flag clear and
// k_check_finite); then, when p64 is given, the step p -= lr delta under the
// same guard as k_update (no failed pivot, finite step, finite loss); then,
// when out is given, the status record [stats(5), info, flag] (as
// k_status_out)
__global__ __launch_bounds__(1024) void k_split_tail(const double* x0, const double* x1, int nl, int npi,
                                                     double* delta, int* flag, double* p64, float* p32, ndm_dims d,
                                                     double lr, const int* info, const double* stats, double* out) {
    // the first TPT entries of each thread stay in registers with their
    // parameters, and the guard words are read up front, so the step costs
    // no second round of dependent loads (the parameters are only read
    // before the guard is known)
    constexpr int TPT = 4;
    const int np = d.np, m = nl + npi, t = threadIdx.x, nt = blockDim.x;
    const bool need = p64 || out;
    const int inf = need ? *info : 0;
    const double loss = need ? stats[0] : 0.0;
    double v[TPT], q[TPT];
    int r[TPT];
    int bad = 0;
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
        const int p = t + u * nt;
        v[u] = 0.0;
        q[u] = 0.0;
        r[u] = 0;
        if (p < np) {
            v[u] = (p < m) ? x0[p] : x1[p - m];
            if (p64) {
                r[u] = p2ref(p, d);
                q[u] = p64[r[u]];
            }
        }
    }
    for (int p = t + TPT * nt; p < np; p += nt) {
        const double w = (p < m) ? x0[p] : x1[p - m];
        delta[p] = w;
        bad |= !isfinite(w);
    }
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
        if (t + u * nt < np) {
            delta[t + u * nt] = v[u];
            bad |= !isfinite(v[u]);
        }
    }
    bad = __syncthreads_or(bad);
    if (t == 0) *flag = bad;
    if (p64 && !bad && inf == 0 && isfinite(loss)) {
#pragma unroll
        for (int u = 0; u < TPT; ++u) {
            if (t + u * nt < np) {
                const double w = q[u] - lr * v[u];
                p64[r[u]] = w;
                if (p32) p32[r[u]] = (float)w;
            }
        }
        for (int p = t + TPT * nt; p < np; p += nt) {
            const double w0 = (p < m) ? x0[p] : x1[p - m];
            const int rr = p2ref(p, d);
            const double w = p64[rr] - lr * w0;
            p64[rr] = w;
            if (p32) p32[rr] = (float)w;
        }
    }
    if (out) {
        if (t < 5) out[t] = stats[t];
        if (t == 5) out[5] = (double)inf;
        if (t == 6) out[6] = (double)bad;
        status_done(out);
    }
}

// ---- packed forms: a member's part of the split solve, its arguments
// derived from its descriptor the way nq_split_factor / nq_split_substitute
// derive them (the host lays out the blocks with the same functions)
struct chol_desc {
    double* A;
    int n, nr, ncu, lda, k, kb;
    double* winv;
    int* info;
    int k0;
    int64_t sA, sW;
    int gx, gy;
};

__host__ __device__ inline int64_t winv_doubles(int n) { return (int64_t)((n + CB - 1) / CB) * CB * CB; }

__host__ __device__ inline int pk_panels(int nl, int npi, int phase) {
    return phase == 0 ? (nl + CB - 1) / CB : (npi + CB - 1) / CB;
}

__host__ __device__ inline int pk_groups(int nl, int npi, int phase) {
    return phase == 0 ? (npi + BG * CB - 1) / (BG * CB) : (nl + BG * CB - 1) / (BG * CB);
}

// phase 0: [lambda | Pi] and [mu | Pi] over their first nl columns (2 instances);
// phase 1: the Pi block; panel j of the phase (the diagonal block alone when j < 0)
__host__ __device__ inline chol_desc pk_chol(const nq_pk_member& m, int phase, int j) {
    chol_desc c;
    const int mm = m.nl + m.npi;
    c.A = m.A2;
    c.lda = mm + 1;
    c.nr = mm + 1;
    c.info = m.info;
    if (phase == 0) {
        c.n = m.nl;
        c.ncu = m.nl;
        c.k0 = 0;
        c.winv = m.winv;
        c.sA = (int64_t)c.lda * c.lda;
        c.sW = winv_doubles(m.nl);
        c.gy = 2;
    } else {
        c.n = mm;
        c.ncu = mm + 1;
        c.k0 = m.nl;
        c.winv = m.winv + 2 * winv_doubles(m.nl);
        c.sA = 0;
        c.sW = 0;
        c.gy = 1;
    }
    c.k = c.k0 + CB * (j < 0 ? 0 : j);
    c.kb = (c.n - c.k < CB) ? c.n - c.k : CB;
    if (j < 0) {
        c.gx = 1;
    } else {
        const int rest = c.nr - c.k - c.kb;
        const int T = (rest + 63) / 64, TJ = ((c.ncu - c.k - c.kb + 63) / 64 > 1) ? (c.ncu - c.k - c.kb + 63) / 64 : 1;
        c.gx = rest > 0 ? step_tiles(T, TJ) : 0;
    }
    return c;
}

struct back_desc {
    int k, kend, n, k0;
    const double* winv;
    int64_t sA, sW, sY;
    int gx, gy;
};

// phase 0: the Pi block's groups; phase 1: lambda and mu together; group j of
// the phase counts from the last group of panels
__host__ __device__ inline back_desc pk_back(const nq_pk_member& m, int phase, int j) {
    back_desc b;
    const int mm = m.nl + m.npi, lda = mm + 1, groups = pk_groups(m.nl, m.npi, phase);
    const int g = groups - 1 - j;
    if (phase == 0) {
        b.k0 = m.nl;
        b.n = mm;
        b.winv = m.winv + 2 * winv_doubles(m.nl);
        b.sA = b.sW = b.sY = 0;
        b.gy = 1;
    } else {
        b.k0 = 0;
        b.n = m.nl;
        b.winv = m.winv;
        b.sA = (int64_t)lda * lda;
        b.sW = winv_doubles(m.nl);
        b.sY = lda;
        b.gy = 2;
    }
    b.k = b.k0 + g * BG * CB;
    b.kend = (b.n - b.k < BG * CB) ? b.n : b.k + BG * CB;
    b.gx = g < 0 ? 0 : (b.k > 0 ? (b.k + BR - 1) / BR : 1);
    return b;
}

__device__ __forceinline__ const nq_pk_member& pk_m(const nq_pk_member* M, uint2 b) { return M[b.x & 0xffff]; }
__device__ __forceinline__ int pk_x(uint2 b) { return (int)(b.y & 0xffff); }
__device__ __forceinline__ int pk_y(uint2 b) { return (int)(b.y >> 16); }
__device__ __forceinline__ int pk_z(uint2 b) { return (int)(b.x >> 16); }

// the counters a member's iteration starts from: acceptance counts, info, flag
__global__ __launch_bounds__(256) void k_prep_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    const int64_t i = (int64_t)pk_x(b) * blockDim.x + threadIdx.x;
    if (i < m.chains) m.chain_acc[i] = 0u;
    if (i == 0) {
        m.acc[0] = 0ull;
        m.acc[1] = 0ull;
        *m.info = 0;
        *m.flag = 0;
    }
}

__global__ __launch_bounds__(256) void k_sfill_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    const int64_t i = (int64_t)pk_x(b) * blockDim.x + threadIdx.x;
    if (i < (int64_t)m.ldS * m.ldS) m.S[i] = 0.0;
}

template <typename T>
__global__ __launch_bounds__(1024) void k_stats_pk(const nq_pk_member* M, const uint2* blk) {
    const nq_pk_member& m = pk_m(M, blk[blockIdx.x]);
    stats_body<T>((const T*)m.Lloc, m.chains, m.Lc, m.stats);
}

__global__ __launch_bounds__(256) void k_force_part_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    const int per = (m.force_grid + NQ_FORCE_GROUPS - 1) / NQ_FORCE_GROUPS;
    force_part_body(m.Fpart, m.force_grid, per, m.np, m.fpart2, pk_x(b), pk_y(b));
}

__global__ __launch_bounds__(256) void k_force_finish_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    force_finish_body(m.fpart2, NQ_FORCE_GROUPS, m.mu, m.stats, m.sa.d, m.F, pk_x(b));
}

__global__ __launch_bounds__(256) void k_split_copy_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    split_copy_body(m.S, m.ldS, m.F, m.nl, m.npi, m.eps, m.kind, m.A2, m.nl + m.npi + 1, pk_x(b), pk_y(b), m.winv,
                    winv_doubles(m.nl), m.info);
}

__global__ __launch_bounds__(256) void k_chol_step_pk(const nq_pk_member* M, const uint2* blk, int phase, int j) {
    const uint2 b = blk[blockIdx.x];
    const chol_desc c = pk_chol(pk_m(M, b), phase, j);
    chol_step_body(c.A, c.n, c.nr, c.ncu, c.lda, c.k, c.kb, c.winv, c.info, c.k0, c.sA, c.sW, pk_x(b), pk_y(b));
}

__global__ __launch_bounds__(256, 3) void k_split_schur_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    split_schur_body(m.A2, m.nl + m.npi + 1, m.nl, m.npi, pk_x(b), m.winv + 2 * winv_doubles(m.nl), m.info);
}

// the forward-substituted right-hand sides of both instances (z = instance)
__global__ __launch_bounds__(256) void k_copy_cols_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    const int mm = m.nl + m.npi, lda = mm + 1, z = pk_z(b);
    copy_col_body(m.A2 + (int64_t)z * lda * lda, mm, lda, m.y2 + (int64_t)z * lda, pk_x(b));
}

__global__ __launch_bounds__(256) void k_back_group_pk(const nq_pk_member* M, const uint2* blk, int phase, int j) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    const back_desc d = pk_back(m, phase, j);
    back_group_body(m.A2, m.nl + m.npi + 1, d.k, d.kend, d.n, d.winv, m.y2, m.x2, d.k0, d.sA, d.sW, d.sY, pk_x(b),
                    pk_y(b));
}

__global__ __launch_bounds__(256) void k_split_upd_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    const int mm = m.nl + m.npi, lda = mm + 1;
    split_upd_body(m.A2 + (int64_t)lda * lda, lda, m.nl, mm, m.x2, m.x2 + 2 * (int64_t)lda, pk_x(b), pk_y(b));
}

__global__ __launch_bounds__(256) void k_split_upd_fin_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    const int lda = m.nl + m.npi + 1;
    split_upd_fin_body(m.x2 + 2 * (int64_t)lda, m.nl, m.y2 + lda, pk_x(b));
}

__global__ __launch_bounds__(256) void k_split_gather_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    const int lda = m.nl + m.npi + 1;
    split_gather_body(m.x2, m.x2 + lda, m.nl, m.npi, m.np, m.delta, pk_x(b));
}

__global__ __launch_bounds__(256) void k_check_finite_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    check_finite_body(m.delta, m.np, m.flag, pk_x(b));
}

__global__ __launch_bounds__(256) void k_update_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_m(M, b);
    const nq_step_guard g{m.info, m.flag, m.stats};
    update_body(m.p64, m.p32, m.delta, m.sa.d, m.lr, g, pk_x(b));
}

__global__ void k_bump_pk(const nq_pk_member* M, const uint2* blk) {
    const nq_pk_member& m = pk_m(M, blk[blockIdx.x]);
    if (threadIdx.x == 0) m.iter_dev[0] += 1u;
}

__global__ void k_collect_pk(const nq_pk_member* M, const uint2* blk) {
    const nq_pk_member& m = pk_m(M, blk[blockIdx.x]);
    const int t = threadIdx.x;
    if (t < 5) m.status[t] = m.stats[t];
    if (t == 5) m.status[5] = (double)*m.info;
    if (t == 6) m.status[6] = (double)*m.flag;
}

}  // namespace

hipError_t nq_launch_fill(double* x, int64_t n, double v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, st, x, n, v);
    return hipGetLastError();
}

hipError_t nq_launch_fill_u32(uint32_t* x, int64_t n, uint32_t v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(n, 256)), dim3(256), 0, st, x, n, v);
    return hipGetLastError();
}

hipError_t nq_launch_bump(uint32_t* x, hipStream_t st) {
    hipLaunchKernelGGL(k_bump, dim3(1), dim3(64), 0, st, x);
    return hipGetLastError();
}

hipError_t nq_launch_status_out(const double* stats, const int* info, const int* flag, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_status_out, dim3(1), dim3(64), 0, st, stats, info, flag, out);
    return hipGetLastError();
}

hipError_t nq_launch_stats(const void* Lloc, int fp64, int64_t B, int64_t Lc, double* out, hipStream_t st) {
    if (fp64)
        hipLaunchKernelGGL((k_stats<double>), dim3(1), dim3(1024), 0, st, (const double*)Lloc, B, Lc, out);
    else
        hipLaunchKernelGGL((k_stats<float>), dim3(1), dim3(1024), 0, st, (const float*)Lloc, B, Lc, out);
    return hipGetLastError();
}

hipError_t nq_launch_force_finish(const double* Fpart, int nb, double* part2, const double* mu,
                                  const double* stats, ndm_dims d, double* F, hipStream_t st) {
    const int groups = NQ_FORCE_GROUPS;
    const int per = (nb + groups - 1) / groups;
    hipLaunchKernelGGL(k_force_part, dim3(nblk(d.np, 256), groups), dim3(256), 0, st, Fpart, nb, per, d.np, part2);
    hipLaunchKernelGGL(k_force_finish, dim3(nblk(d.np, 256)), dim3(256), 0, st, part2, groups, mu, stats, d, F);
    return hipGetLastError();
}

hipError_t nq_launch_force_part(const double* Fpart, int nb, int n, double* part2, hipStream_t st) {
    const int groups = NQ_FORCE_GROUPS;
    const int per = (nb + groups - 1) / groups;
    hipLaunchKernelGGL(k_force_part, dim3(nblk(n, 256), groups), dim3(256), 0, st, Fpart, nb, per, n, part2);
    return hipGetLastError();
}

hipError_t nq_launch_shift(double* S, int n, int lds, double eps, int kind, hipStream_t st) {
    hipLaunchKernelGGL(k_shift, dim3(nblk(n, 256)), dim3(256), 0, st, S, n, lds, eps, kind);
    return hipGetLastError();
}

size_t nq_cholesky_winv_doubles(int n) { return (size_t)((n + CB - 1) / CB) * CB * CB; }

hipError_t nq_cholesky_range(double* A, int k0, int n, int nr, int ncu, int lda, int* info, double* winv, int ninst,
                              int64_t sA, int64_t sW, hipStream_t st, bool first_done) {
    if (k0 >= n) return hipSuccess;
    if (!first_done)
        hipLaunchKernelGGL(k_chol_diag32, dim3(1, ninst), dim3(64), 0, st, A, n, lda, k0, (n - k0 < CB ? n - k0 : CB),
                           winv, info, sA, sW);
    for (int k = k0; k < n; k += CB) {
        const int kb = (n - k < CB) ? n - k : CB;
        const int rest = nr - k - kb;
        if (rest <= 0) break;
        // row tiles of the trailing rows; column tiles of the updated columns
        // (at least one, whose tiles still transform the panel's rows)
        const int T = (rest + 63) / 64, TJ = std::max(1, (ncu - k - kb + 63) / 64);
        hipLaunchKernelGGL(k_chol_step, dim3(step_tiles(T, TJ) * ninst), dim3(256), 0, st, A, n, nr, ncu, lda, k, kb,
                           winv, info, k0, sA, sW, ninst);
    }
    return hipGetLastError();
}

hipError_t nq_cholesky(double* S, int n, int lds, int* info, double* winv, hipStream_t st) {
    // rows >= n (at most one extra row, the right-hand side) are carried along
    const int nr = (lds > n) ? n + 1 : n;
    return nq_cholesky_range(S, 0, n, nr, nr, lds, info, winv, 1, 0, 0, st, false);
}

hipError_t nq_cholesky_backsolve_range(const double* L, int k0, int n, int lds, const double* winv, double* y,
                                       double* x, int ninst, int64_t sA, int64_t sW, int64_t sY, hipStream_t st) {
    if (k0 >= n) return hipSuccess;
    // groups of BG panels from k0, the last group first
    const int groups = (n - k0 + BG * CB - 1) / (BG * CB);
    for (int g = groups - 1; g >= 0; --g) {
        const int k = k0 + g * BG * CB, kend = (n - k < BG * CB) ? n : k + BG * CB;
        hipLaunchKernelGGL(k_back_group, dim3(k > 0 ? nblk(k, BR) : 1, ninst), dim3(256), 0, st, L, lds, k, kend, n,
                           winv, y, x, k0, sA, sW, sY);
    }
    return hipGetLastError();
}

hipError_t nq_cholesky_backsolve(const double* L, int n, int lds, const double* winv, double* y, double* x,
                                 hipStream_t st) {
    return nq_cholesky_backsolve_range(L, 0, n, lds, winv, y, x, 1, 0, 0, 0, st);
}

hipError_t nq_copy_aug_col(const double* A, int n, int lds, double* y, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_col, dim3(nblk(n, 256)), dim3(256), 0, st, A, n, lds, y);
    return hipGetLastError();
}

hipError_t nq_cholesky_solve(const double* L, int n, int lds, const double* winv, double* y, double* tmp,
                             hipStream_t st) {
    // forward L z = y (z in tmp), then back L^T y = z
    for (int k = 0; k < n; k += CB) {
        const int kb = (n - k < CB) ? n - k : CB;
        const int rest = n - k - kb;
        hipLaunchKernelGGL(k_fwd_step, dim3(rest > 0 ? nblk(rest, 256) : 1), dim3(256), 0, st, L, n, lds, k, kb, winv,
                           y, tmp);
    }
    return nq_cholesky_backsolve(L, n, lds, winv, tmp, y, st);
}

size_t nq_split_x_doubles(int nl, int npi) { return 2 * (size_t)(nl + npi + 1) + (size_t)kUpdChunks * nl; }

size_t nq_split_winv_doubles(int nl, int npi) {
    return 2 * nq_cholesky_winv_doubles(nl) + nq_cholesky_winv_doubles(npi);
}

hipError_t nq_split_factor(const double* S, int lds, const double* F, int nl, int npi, double eps, int kind,
                           double* A2, double* winv, int* info, hipStream_t st) {
    const int m = nl + npi, lda = m + 1;
    const int64_t sA = (int64_t)lda * lda;
    const int64_t sW = (int64_t)nq_cholesky_winv_doubles(nl);
    // the copy and the Schur update also factor the first diagonal block of
    // the range that follows them
    hipLaunchKernelGGL(k_split_copy, dim3(m, 2), dim3(256), 0, st, S, lds, F, nl, npi, eps, kind, A2, lda, winv, sW,
                       info);
    // both [lambda | Pi] and [mu | Pi] factored over their first nl columns
    // together, their Pi columns left as they are
    hipError_t e = nq_cholesky_range(A2, 0, nl, m + 1, nl, lda, info, winv, 2, sA, sW, st, true);
    if (e != hipSuccess) return e;
    const int T = (npi + 1 + 31) / 32;
    hipLaunchKernelGGL(k_split_schur, dim3(step_tiles(T, T)), dim3(256), 0, st, A2, lda, nl, npi, winv + 2 * sW, info);
    // the Pi block of instance 0 now holds the whole matrix's Schur complement
    return nq_cholesky_range(A2, nl, m, m + 1, m + 1, lda, info, winv + 2 * sW, 1, 0, 0, st, true);
}

hipError_t nq_split_substitute(int nl, int npi, const double* A2, const double* winv, double* y2, double* x2,
                               hipStream_t st) {
    const int m = nl + npi, lda = m + 1;
    const int64_t sA = (int64_t)lda * lda, sY = lda;
    const int64_t sW = (int64_t)nq_cholesky_winv_doubles(nl);
    const double* winvP = winv + 2 * sW;
    hipError_t e;
    double *y0 = y2, *y1 = y2 + sY, *x0 = x2, *x1 = x2 + sY;
    hipLaunchKernelGGL(k_copy_col2, dim3(nblk(m, 256), 2), dim3(256), 0, st, A2, m, lda, sA, y0, sY);
    // x_Pi, with the lambda rows of y updated on the way; then the mu rows
    if ((e = nq_cholesky_backsolve_range(A2, nl, m, lda, winvP, y0, x0, 1, 0, 0, 0, st)) != hipSuccess) return e;
    double* part = x2 + 2 * sY;   // kUpdChunks x nl scratch after the two solutions
    hipLaunchKernelGGL(k_split_upd, dim3(nblk(nl, 64), kUpdChunks), dim3(256), 0, st, A2 + sA, lda, nl, m, x0, part);
    hipLaunchKernelGGL(k_split_upd_fin, dim3(nblk(nl, 256)), dim3(256), 0, st, part, nl, y1);
    // x_lambda and x_mu together
    return nq_cholesky_backsolve_range(A2, 0, nl, lda, winv, y0, x0, 2, sA, sW, sY, st);
}

hipError_t nq_split_tail(int nl, int npi, const double* x2, double* delta, int* flag, double* p64, float* p32,
                         const ndm_dims& d, double lr, const int* info, const double* stats, double* out,
                         hipStream_t st) {
    const int lda = nl + npi + 1;
    hipLaunchKernelGGL(k_split_tail, dim3(1), dim3(1024), 0, st, x2, x2 + lda, nl, npi, delta, flag, p64, p32, d, lr,
                       info, stats, out);
    return hipGetLastError();
}

hipError_t nq_set_aug_row(double* A, int n, int lds, const double* b, hipStream_t st) {
    hipLaunchKernelGGL(k_set_row, dim3(nblk(n, 256)), dim3(256), 0, st, A, n, lds, b);
    return hipGetLastError();
}

hipError_t nq_residual(const double* S, int n, int lds, double eps, int kind, const double* x,
                       const double* b, double* r, hipStream_t st) {
    hipLaunchKernelGGL(k_residual, dim3(nblk(n, 256)), dim3(256), 0, st, S, n, lds, eps, kind, x, b, r);
    return hipGetLastError();
}

hipError_t nq_launch_axpy(double* x, const double* y, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_axpy, dim3(nblk(n, 256)), dim3(256), 0, st, x, y, n);
    return hipGetLastError();
}

hipError_t nq_launch_update(double* p64, float* p32, const double* delta, ndm_dims d, double lr,
                            nq_step_guard g, hipStream_t st) {
    hipLaunchKernelGGL(k_update, dim3(nblk(d.np, 256)), dim3(256), 0, st, p64, p32, delta, d, lr, g);
    return hipGetLastError();
}

hipError_t nq_launch_to_f32(const double* p64, float* p32, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_to_f32, dim3(nblk(n, 256)), dim3(256), 0, st, p64, p32, n);
    return hipGetLastError();
}

hipError_t nq_launch_check_finite(const double* x, int64_t n, int* flag, hipStream_t st) {
    hipLaunchKernelGGL(k_check_finite, dim3(nblk(n, 256)), dim3(256), 0, st, x, n, flag);
    return hipGetLastError();
}

int nq_pk_split_panels(int nl, int npi, int phase) { return pk_panels(nl, npi, phase); }
int nq_pk_split_groups(int nl, int npi, int phase) { return pk_groups(nl, npi, phase); }

hipError_t nq_pk_plan(const nq_pk_member& m, int kernel, int param, nq_pk_shape* o) {
    *o = nq_pk_shape{0, 1, 1, 0, 256, 0};
    const int mm = m.nl + m.npi;
    const int phase = param >> 16, j = param & 0xffff;
    switch (kernel) {
        case NQ_PK_PREP: o->gx = (int)nblk(m.chains > 0 ? m.chains : 1, 256); break;
        case NQ_PK_METROPOLIS: {
            int grid = 0;
            o->variant = nq_sampler_variant(m.sa, m.fp64 != 0, &o->threads, &o->lds, &grid);
            if (o->variant < 0) return hipErrorInvalidValue;
            o->gx = grid;
            break;
        }
        case NQ_PK_EST_WALK:
        case NQ_PK_EST_FORCE:
            o->variant = nq_est_variant(kernel == NQ_PK_EST_WALK ? m.ew : m.ef, m.fp64 != 0, &o->threads, &o->lds);
            if (o->variant < 0) return hipErrorInvalidValue;
            o->gx = kernel == NQ_PK_EST_WALK ? m.est_grid : m.force_grid;
            break;
        case NQ_PK_COLSUM:
            o->variant = m.fp64;
            o->gx = (m.R + 255) / 256;
            o->gy = m.col_chunks;
            break;
        case NQ_PK_COLSUM_FIN: o->gx = (m.R + 63) / 64; break;
        case NQ_PK_STATS:
            o->variant = m.fp64;
            o->threads = 1024;
            o->gx = 1;
            break;
        case NQ_PK_FORCE_PART:
            o->gx = (int)nblk(m.np, 256);
            o->gy = NQ_FORCE_GROUPS;
            break;
        case NQ_PK_FORCE_FIN: o->gx = (int)nblk(m.np, 256); break;
        case NQ_PK_SFILL:
            if (m.s_mode >= 0) o->gx = (int)nblk((int64_t)m.ldS * m.ldS, 256);
            break;
        case NQ_PK_GRAM:
            if (m.s_mode >= 0) {
                o->variant = m.s_mode | (m.fp64 << 1);
                o->gx = nq_pk_gram_blocks(m, param);
            }
            break;
        case NQ_PK_CENTER:
            if (m.s_mode >= 0) break;
            o->variant = m.fp64;
            o->gx = (m.R + 63) / 64;
            o->gy = (int)((m.nsp + 63) / 64);
            break;
        case NQ_PK_GRAM_PART:
            if (m.s_mode >= 0) break;
            o->gx = m.tiles;
            o->gy = m.splits;
            break;
        case NQ_PK_GRAM_FIN: {
            if (m.s_mode >= 0) break;
            const int TS = (m.np + 1 + 63) / 64;
            o->gx = TS * (TS + 1) / 2;
            break;
        }
        case NQ_PK_SPLIT_COPY:
            o->gx = mm;
            o->gy = 2;
            break;
        case NQ_PK_CHOL_STEP: {
            if (j >= pk_panels(m.nl, m.npi, phase)) break;
            const chol_desc c = pk_chol(m, phase, j);
            o->gx = c.gx;
            o->gy = c.gy;
            break;
        }
        case NQ_PK_SCHUR: {
            const int T = (m.npi + 1 + 31) / 32;
            o->gx = step_tiles(T, T);
            break;
        }
        case NQ_PK_COPY_COLS:
            o->gx = (int)nblk(mm, 256);
            o->gy = 1;
            o->gz = 2;
            break;
        case NQ_PK_BACK: {
            if (j >= pk_groups(m.nl, m.npi, phase)) break;
            const back_desc d = pk_back(m, phase, j);
            o->gx = d.gx;
            o->gy = d.gy;
            break;
        }
        case NQ_PK_SPLIT_UPD:
            o->gx = (int)nblk(m.nl, 64);
            o->gy = kUpdChunks;
            break;
        case NQ_PK_UPD_FIN: o->gx = (int)nblk(m.nl, 256); break;
        case NQ_PK_GATHER: o->gx = (int)nblk(m.np, 256); break;
        case NQ_PK_CHECK: o->gx = (int)nblk(m.np, 256); break;
        case NQ_PK_UPDATE: o->gx = (int)nblk(m.np, 256); break;
        case NQ_PK_BUMP:
        case NQ_PK_COLLECT:
            o->threads = 64;
            o->gx = 1;
            break;
        default: return hipErrorInvalidValue;
    }
    if (o->gx > 0xffff || o->gy > 0xffff || o->gz > 0xffff) return hipErrorInvalidValue;
    return hipSuccess;
}

hipError_t nq_pk_solve(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st) {
    if (L.nblk <= 0) return hipSuccess;
    const dim3 g((unsigned)L.nblk), b((unsigned)L.threads);
    const int phase = L.param >> 16, j = L.param & 0xffff;
    switch (L.kernel) {
        case NQ_PK_PREP: hipLaunchKernelGGL(k_prep_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_STATS:
            if (L.variant) hipLaunchKernelGGL((k_stats_pk<double>), g, b, 0, st, M, L.blk);
            else hipLaunchKernelGGL((k_stats_pk<float>), g, b, 0, st, M, L.blk);
            break;
        case NQ_PK_FORCE_PART: hipLaunchKernelGGL(k_force_part_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_FORCE_FIN: hipLaunchKernelGGL(k_force_finish_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_SFILL: hipLaunchKernelGGL(k_sfill_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_SPLIT_COPY: hipLaunchKernelGGL(k_split_copy_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_CHOL_STEP: hipLaunchKernelGGL(k_chol_step_pk, g, b, 0, st, M, L.blk, phase, j); break;
        case NQ_PK_SCHUR: hipLaunchKernelGGL(k_split_schur_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_COPY_COLS: hipLaunchKernelGGL(k_copy_cols_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_BACK: hipLaunchKernelGGL(k_back_group_pk, g, b, 0, st, M, L.blk, phase, j); break;
        case NQ_PK_SPLIT_UPD: hipLaunchKernelGGL(k_split_upd_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_UPD_FIN: hipLaunchKernelGGL(k_split_upd_fin_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_GATHER: hipLaunchKernelGGL(k_split_gather_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_CHECK: hipLaunchKernelGGL(k_check_finite_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_UPDATE: hipLaunchKernelGGL(k_update_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_BUMP: hipLaunchKernelGGL(k_bump_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_COLLECT: hipLaunchKernelGGL(k_collect_pk, g, b, 0, st, M, L.blk); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t nq_pk_run(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st) {
    switch (L.kernel) {
        case NQ_PK_METROPOLIS: return nq_pk_sampler(M, L, st);
        case NQ_PK_EST_WALK:
        case NQ_PK_EST_FORCE: return nq_pk_estimate(M, L, st);
        case NQ_PK_COLSUM:
        case NQ_PK_COLSUM_FIN:
        case NQ_PK_CENTER:
        case NQ_PK_GRAM_PART:
        case NQ_PK_GRAM_FIN:
        case NQ_PK_GRAM: return nq_pk_sbuild(M, L, st);
        default: return nq_pk_solve(M, L, st);
    }
}
