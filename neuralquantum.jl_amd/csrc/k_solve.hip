// Loss statistics, force finish, the shifted SR solve and the parameter update.
//
//   stats ........ stat_analysis of |L_loc|^2 over chains x slices (stats.jl:52-77)
//   force ........ F = Re(conj(L_loc) dL_loc)/Ns - C Re(Obar)   (BatchedGradSampler.jl:108-118)
//   shift ........ S_ii += eps (sr_shift, SRDirect.jl:60-65) or eps * S_ii
//   Cholesky ..... blocked right-looking, FP64, panels of 32 columns: diagonal
//                  block factored in registers of one wavefront (v_readlane
//                  broadcasts), panel triangular solve, trailing update on
//                  FP64 MFMA (cholesky!(Hermitian(S)) + ldiv!, SRDirect.jl:78-81)
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
__global__ __launch_bounds__(256) void k_stats(const T* L, int64_t B, int64_t Lc, double* out) {
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
__global__ void k_force_part(const double* Fpart, int nblk, int per, int np, double* part2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;
    if (k >= np) return;
    const int b0 = g * per, b1 = (b0 + per < nblk) ? b0 + per : nblk;
    double acc = 0.0;
    for (int b = b0; b < b1; ++b) acc += Fpart[(int64_t)b * np + k];
    part2[(int64_t)g * np + k] = acc;
}

// stage 2: F[k] = sum_g part2[g][k] - C Re(Obar_k)
__global__ void k_force_finish(const double* part2, int ngroups, const double* mu, const double* stats,
                               ndm_dims d, double* F) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
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

// ---- Cholesky step k with a panel of CB = 32 columns, one wavefront per block.
// Every block factors the kb x kb diagonal block (padded with the identity)
// with lane t < 32 holding row t in registers; column j is broadcast with
// v_readlane (no LDS and no barrier on the dependent chain).  Block 0 writes
// the factor back; block b >= 1 solves rows k + kb + 64 (b - 1) + lane of the
// panel, X = A[i, k:k+kb] L_kk^-T, by column-oriented substitution against
// the factor in LDS.
constexpr int CB = 32;

__global__ __launch_bounds__(64, 1) void k_chol_panel32(double* A, int n, int lda, int k, int kb, int* info) {
    __shared__ double L[CB][CB + 1];
    __shared__ double rdiag[CB];
    const int t = threadIdx.x, tr = t & (CB - 1);
    // loads are unconditional (clamped addresses) and selected afterwards: a
    // load under a lane-dependent condition would be waited for one by one
    double r[CB];
    const int64_t rrow = (int64_t)min(k + tr, n - 1);
#pragma unroll
    for (int c = 0; c < CB; ++c) {
        const double a = A[rrow + (int64_t)min(k + c, n - 1) * lda];
        r[c] = (tr < kb && c < kb && tr >= c) ? a : (tr == c ? 1.0 : 0.0);
    }
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        const double v = rl_d(r[j], j);                      // current A[j][j]
        if (j < kb && t == 0 && blockIdx.x == 0 && !(v > 0.0) && *info == 0) *info = k + j + 1;
        const double piv = sqrt(v);
        const double rp = 1.0 / piv;
        const double lj = (tr == j) ? piv : r[j] * rp;      // L[tr][j] (zero above the diagonal)
        r[j] = lj;
        // r[m] -= L[tr][j] L[m][j] for every m > j; in rows tr < m this writes
        // the upper triangle, which no lane reads (lane m only ever hands out
        // its own entries r[j] with j <= m), so no mask is needed
#pragma unroll
        for (int m = j + 1; m < CB; ++m) r[m] -= lj * rl_d(lj, m);
    }
    if (t < CB) {
#pragma unroll
        for (int c = 0; c < CB; ++c) L[t][c] = r[c];
    }
    __syncthreads();
    if (t < CB) rdiag[t] = 1.0 / L[t][t];
    __syncthreads();
    if (blockIdx.x == 0) {
        if (t < kb) {
#pragma unroll
            for (int c = 0; c < CB; ++c)
                if (c < kb && t >= c) A[(int64_t)(k + t) + (int64_t)(k + c) * lda] = r[c];
        }
        return;
    }
    const int i = k + kb + (blockIdx.x - 1) * 64 + t;
    if (i >= n) return;
    double x[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        const double a = A[(int64_t)i + (int64_t)min(k + j, n - 1) * lda];
        x[j] = (j < kb) ? a : 0.0;
    }
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        if (j < kb) {
            x[j] *= rdiag[j];
#pragma unroll
            for (int m = j + 1; m < CB; ++m)
                if (m < kb) x[m] -= x[j] * L[m][j];
        }
    }
#pragma unroll
    for (int j = 0; j < CB; ++j)
        if (j < kb) A[(int64_t)i + (int64_t)(k + j) * lda] = x[j];
}

__device__ __forceinline__ void tri_decode(int b, int* I, int* J) {
    int j = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
    while ((j + 1) * (j + 2) / 2 <= b) ++j;
    while (j * (j + 1) / 2 > b) --j;
    *J = j;
    *I = b - j * (j + 1) / 2;
}

// ---- trailing update: A[i, j] -= sum_m X[i, m] X[j, m] for tiles I >= J of the
// trailing matrix (origin o = k + kb), X = A[:, k:k+kb]; FP64 MFMA 16x16x4
template <int KW>
__global__ __launch_bounds__(256) void k_syrk_trailing(double* A, int n, int lda, int k, int kb) {
    __shared__ double Xi[KW][65];   // [m][row]
    __shared__ double Xj[KW][65];
    int tI, tJ;
    tri_decode(blockIdx.x, &tJ, &tI);   // tJ <= tI
    const int o = k + kb;
    const int i0 = o + tI * 64, j0 = o + tJ * 64;
    for (int e = threadIdx.x; e < 64 * KW; e += blockDim.x) {
        const int r = e & 63, m = e >> 6;
        const int64_t col = (int64_t)(k + min(m, kb - 1)) * lda;
        const double vi = A[(int64_t)min(i0 + r, n - 1) + col];
        const double vj = A[(int64_t)min(j0 + r, n - 1) + col];
        Xi[m][r] = (m < kb && i0 + r < n) ? vi : 0.0;
        Xj[m][r] = (m < kb && j0 + r < n) ? vj : 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int li = lane & 15, lk = lane >> 4;
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[a][b][v] = 0.0;
    for (int m = 0; m < KW; m += 4) {
        double av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) av[a] = Xi[m + lk][wr * 32 + a * 16 + li];
#pragma unroll
        for (int b = 0; b < 2; ++b) bv[b] = Xj[m + lk][wc * 32 + b * 16 + li];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    // f64 MFMA output: D[lane / 16 + 4 v][lane % 16]
    const int lj = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = i0 + wr * 32 + a * 16 + lg + 4 * v;
                const int c = j0 + wc * 32 + b * 16 + lj;
                if (r < n && c < n && r >= c) A[(int64_t)r + (int64_t)c * lda] -= acc[a][b][v];
            }
}

// ---- forward substitution L y = b, one block of 64 per step
__global__ __launch_bounds__(64) void k_fwd_diag(const double* L, int lda, int k, int kb, double* y) {
    const int lane = threadIdx.x;
    double row[NB];   // row[j] = L[k + lane][k + j]
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const double a = L[(int64_t)(k + min(lane, kb - 1)) + (int64_t)(k + min(j, kb - 1)) * lda];
        row[j] = (lane < kb && j < kb) ? a : 1.0;
    }
    const double y0 = y[k + min(lane, kb - 1)];
    double v = (lane < kb) ? y0 : 0.0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (j < kb) {
            if (lane == j) v = v / row[j];
            const double yj = __shfl(v, j, 64);
            if (lane > j && lane < kb) v -= row[j] * yj;
        }
    }
    if (lane < kb) y[k + lane] = v;
}

// y[i] -= sum_{m < kb} L[i][k + m] y[k + m] for k + kb <= i < n (coalesced over i)
__global__ __launch_bounds__(256) void k_fwd_update(const double* L, int n, int lda, int k, int kb, double* y) {
    __shared__ double ys[NB];
    if (threadIdx.x < kb) ys[threadIdx.x] = y[k + threadIdx.x];
    __syncthreads();
    const int i = k + kb + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int m = 0; m < kb; ++m) acc += L[(int64_t)i + (int64_t)(k + m) * lda] * ys[m];
    y[i] -= acc;
}

// A[n, j] = b[j] (the right-hand side as the extra row of the factorization)
__global__ void k_set_row(double* A, int n, int lda, const double* b) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) A[(int64_t)n + (int64_t)j * lda] = b[j];
}

// ---- back substitution L^T x = y, one block of 64 per step (last block first)
// diag step: x[k:k+kb] = L_kk^-T y[k:k+kb] (one wave, the block's columns in registers)
__global__ __launch_bounds__(64) void k_back_diag(const double* L, int lda, int k, int kb, double* y) {
    const int lane = threadIdx.x;
    double col[NB];   // col[j] = L[k + j][k + lane]
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const double a = L[(int64_t)(k + min(j, kb - 1)) + (int64_t)(k + min(lane, kb - 1)) * lda];
        col[j] = (lane < kb && j < kb) ? a : 1.0;
    }
    const double y0 = y[k + min(lane, kb - 1)];
    double v = (lane < kb) ? y0 : 0.0;
#pragma unroll
    for (int jj = 0; jj < NB; ++jj) {
        const int j = NB - 1 - jj;
        if (j < kb) {
            if (lane == j) v = v / col[j];
            const double xj = __shfl(v, j, 64);
            if (lane < j) v -= col[j] * xj;
        }
    }
    if (lane < kb) y[k + lane] = v;
}

// update step: y[i] -= sum_{m < kb} L[k + m][i] x[k + m] for i < k; one row i
// per thread, its 64 factor entries are contiguous (column i, rows k..k+kb-1)
__global__ __launch_bounds__(256) void k_back_update(const double* L, int lda, int k, int kb, double* y) {
    __shared__ double xs[NB];
    if (threadIdx.x < kb) xs[threadIdx.x] = y[k + threadIdx.x];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    const double* c = L + (int64_t)i * lda + k;
    double acc = 0.0;
#pragma unroll 8
    for (int m = 0; m < kb; ++m) acc += c[m] * xs[m];
    y[i] -= acc;
}

// the forward half of the solve is the factorization's extra row n: copy it out
__global__ void k_copy_row(const double* A, int n, int lda, double* y) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) y[j] = A[(int64_t)n + (int64_t)j * lda];
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

__global__ void k_update(double* p64, float* p32, const double* delta, ndm_dims d, double lr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.np) return;
    const int r = p2ref(k, d);
    const double v = p64[r] - lr * delta[k];
    p64[r] = v;
    if (p32) p32[r] = (float)v;
}

__global__ void k_to_f32(const double* p64, float* p32, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p32[i] = (float)p64[i];
}

__global__ void k_check_finite(const double* x, int64_t n, int* flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !isfinite(x[i])) atomicOr(flag, 1);
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

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

hipError_t nq_launch_stats(const void* Lloc, int fp64, int64_t B, int64_t Lc, double* out, hipStream_t st) {
    if (fp64)
        hipLaunchKernelGGL((k_stats<double>), dim3(1), dim3(256), 0, st, (const double*)Lloc, B, Lc, out);
    else
        hipLaunchKernelGGL((k_stats<float>), dim3(1), dim3(256), 0, st, (const float*)Lloc, B, Lc, out);
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

hipError_t nq_cholesky(double* S, int n, int lds, int* info, hipStream_t st) {
    // rows >= n (at most one extra row, the right-hand side) are carried along:
    // the panel solves and trailing updates turn row n into (L^-1 b)^T
    const int nr = (lds > n) ? n + 1 : n;
    for (int k = 0; k < n; k += CB) {
        const int kb = (n - k < CB) ? n - k : CB;
        const int rest = nr - k - kb;
        hipLaunchKernelGGL(k_chol_panel32, dim3(1 + nblk(rest, 64)), dim3(64), 0, st, S, nr, lds, k, kb, info);
        if (rest > 0) {
            const int T = (rest + 63) / 64;
            hipLaunchKernelGGL(k_syrk_trailing<CB>, dim3(T * (T + 1) / 2), dim3(256), 0, st, S, nr, lds, k, kb);
        }
    }
    return hipGetLastError();
}

hipError_t nq_cholesky_backsolve(const double* L, int n, int lds, double* y, hipStream_t st) {
    const int last = ((n - 1) / NB) * NB;
    for (int k = last; k >= 0; k -= NB) {
        const int kb = (n - k < NB) ? n - k : NB;
        hipLaunchKernelGGL(k_back_diag, dim3(1), dim3(64), 0, st, L, lds, k, kb, y);
        if (k > 0) hipLaunchKernelGGL(k_back_update, dim3(nblk(k, 256)), dim3(256), 0, st, L, lds, k, kb, y);
    }
    return hipGetLastError();
}

hipError_t nq_copy_aug_row(const double* A, int n, int lds, double* y, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_row, dim3(nblk(n, 256)), dim3(256), 0, st, A, n, lds, y);
    return hipGetLastError();
}

hipError_t nq_cholesky_solve(const double* L, int n, int lds, double* y, hipStream_t st) {
    for (int k = 0; k < n; k += NB) {
        const int kb = (n - k < NB) ? n - k : NB;
        hipLaunchKernelGGL(k_fwd_diag, dim3(1), dim3(64), 0, st, L, lds, k, kb, y);
        const int rest = n - k - kb;
        if (rest > 0) hipLaunchKernelGGL(k_fwd_update, dim3(nblk(rest, 256)), dim3(256), 0, st, L, n, lds, k, kb, y);
    }
    return nq_cholesky_backsolve(L, n, lds, y, st);
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
                            hipStream_t st) {
    hipLaunchKernelGGL(k_update, dim3(nblk(d.np, 256)), dim3(256), 0, st, p64, p32, delta, d, lr);
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
