// S build on the matrix cores: S = Re(Oc Oc^H) / Ns (SRDirect.jl:26-49) from
// the compact O rows.  With the lambda block real, the mu block imaginary and
// the Pi block complex, Re(Oc Oc^H) splits into two symmetric Gram matrices of
// column blocks of Y (rows = samples): A = [Re lambda | Re Pi] and
// B = [Im Pi | Im mu]; S in P order is A^T A embedded at [0, nA) plus B^T B
// embedded at [nlam, np) (the lambda-mu block is identically zero).
//
// Each workgroup computes one 64x64 upper-triangular tile of a Gram block;
// the 4 waves own 32x32 quadrants of 2x2 MFMA 16x16 tiles.  Columns are
// centred (Y - mean) while staging through LDS, which is what the reference's
// O - Obar does (BaseIterativeSampler.jl:19-26), before any down-conversion.
//   FP64: v_mfma_f64_16x16x4f64     FP32: v_mfma_f32_16x16x4f32
//   BF16: v_mfma_f32_16x16x32_bf16 (FP32 accumulation)
#include "nq_kernels.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

template <int MODE> struct gram_traits;
template <> struct gram_traits<NQ_GRAM_F64> { typedef double st; typedef d4 acc; static constexpr int KC = 16; };
template <> struct gram_traits<NQ_GRAM_F32> { typedef float st; typedef f4 acc; static constexpr int KC = 16; };
template <> struct gram_traits<NQ_GRAM_BF16> { typedef __bf16 st; typedef f4 acc; static constexpr int KC = 32; };

__device__ __forceinline__ void tri_decode(int b, int* I, int* J) {
    int j = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
    while ((j + 1) * (j + 2) / 2 <= b) ++j;
    while (j * (j + 1) / 2 > b) --j;
    *J = j;
    *I = b - j * (j + 1) / 2;
}

template <int MODE, typename YT>
__global__ __launch_bounds__(256) void k_gram(gram_args a) {
    typedef typename gram_traits<MODE>::st ST;
    typedef typename gram_traits<MODE>::acc AT;
    constexpr int KC = gram_traits<MODE>::KC;
    // f64/f32: [k][64 + 1] (k-major, rows contiguous); bf16: [row][KC + 8] (k contiguous)
    constexpr int PITCH = (MODE == NQ_GRAM_BF16) ? (KC + 8) : 65;
    constexpr int TILE = (MODE == NQ_GRAM_BF16) ? 64 * PITCH : KC * PITCH;
    __shared__ ST As[TILE];
    __shared__ ST Bs[TILE];

    int I, J;
    tri_decode(blockIdx.x, &I, &J);
    const int r0 = I * 64, q0 = J * 64;
    const YT* Y = reinterpret_cast<const YT*>(a.Y);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;

    AT acc[2][2];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int n = 0; n < 2; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[m][n][v] = 0;

    // this thread's staging column and its mean
    const int lc = tid & 63;
    const int ca = r0 + lc, cb = q0 + lc;
    const bool va = ca < a.n, vb = cb < a.n;
    const double mua = (va && a.mu) ? a.mu[a.c0 + ca] : 0.0;
    const double mub = (vb && a.mu) ? a.mu[a.c0 + cb] : 0.0;

    for (int64_t k0 = 0; k0 < a.Ns; k0 += KC) {
#pragma unroll
        for (int kk = tid >> 6; kk < KC; kk += 4) {
            const int64_t s = k0 + kk;
            double xa = 0.0, xb = 0.0;
            if (s < a.Ns) {
                const YT* row = Y + s * (int64_t)a.ldy + a.c0;
                if (va) xa = (double)row[ca] - mua;
                if (vb) xb = (double)row[cb] - mub;
            }
            if constexpr (MODE == NQ_GRAM_BF16) {
                As[lc * PITCH + kk] = (__bf16)(float)xa;
                Bs[lc * PITCH + kk] = (__bf16)(float)xb;
            } else {
                As[kk * PITCH + lc] = (ST)xa;
                Bs[kk * PITCH + lc] = (ST)xb;
            }
        }
        __syncthreads();
        const int li = lane & 15, lk = lane >> 4;
        if constexpr (MODE == NQ_GRAM_BF16) {
            bf8 av[2], bv[2];
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const int row = wr * 32 + m * 16 + li;
#pragma unroll
                for (int e = 0; e < 8; ++e) av[m][e] = As[row * PITCH + lk * 8 + e];
            }
#pragma unroll
            for (int n = 0; n < 2; ++n) {
                const int col = wc * 32 + n * 16 + li;
#pragma unroll
                for (int e = 0; e < 8; ++e) bv[n][e] = Bs[col * PITCH + lk * 8 + e];
            }
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int n = 0; n < 2; ++n)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[n], acc[m][n], 0, 0, 0);
        } else {
#pragma unroll
            for (int k = 0; k < KC; k += 4) {
                ST av[2], bv[2];
#pragma unroll
                for (int m = 0; m < 2; ++m) av[m] = As[(k + lk) * PITCH + wr * 32 + m * 16 + li];
#pragma unroll
                for (int n = 0; n < 2; ++n) bv[n] = Bs[(k + lk) * PITCH + wc * 32 + n * 16 + li];
#pragma unroll
                for (int m = 0; m < 2; ++m)
#pragma unroll
                    for (int n = 0; n < 2; ++n) {
                        if constexpr (MODE == NQ_GRAM_F64)
                            acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bv[n], acc[m][n], 0, 0, 0);
                        else
                            acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[m], bv[n], acc[m][n], 0, 0, 0);
                    }
            }
        }
        __syncthreads();
    }

    // D[i][j] of a 16x16 tile, j = lane % 16:
    //   f32 / bf16: i = 4 (lane / 16) + v;   f64: i = lane / 16 + 4 v
    const int lj = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int n = 0; n < 2; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ti = (MODE == NQ_GRAM_F64) ? lg + 4 * v : 4 * lg + v;
                const int r = r0 + wr * 32 + m * 16 + ti;
                const int c = q0 + wc * 32 + n * 16 + lj;
                if (r < a.n && c < a.n) {
                    const double val = a.scale * (double)acc[m][n][v];
                    a.S[(int64_t)(a.o + r) + (int64_t)(a.o + c) * a.lds] += val;
                    if (I != J) a.S[(int64_t)(a.o + c) + (int64_t)(a.o + r) * a.lds] += val;
                }
            }
}

template <typename YT>
__global__ __launch_bounds__(256) void k_colsum(const YT* Y, int64_t Ns, int R, double* part, int chunks) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    if (c >= R) return;
    const int64_t per = (Ns + chunks - 1) / chunks;
    const int64_t s0 = chunk * per, s1 = (s0 + per < Ns) ? s0 + per : Ns;
    double acc = 0.0;
    for (int64_t s = s0; s < s1; ++s) acc += (double)Y[s * R + c];
    part[(int64_t)chunk * R + c] = acc;
}

__global__ __launch_bounds__(256) void k_colsum_finish(const double* part, int chunks, int R, double scale,
                                                       double* mu) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= R) return;
    double acc = 0.0;
    for (int k = 0; k < chunks; ++k) acc += part[(int64_t)k * R + c];
    mu[c] = acc * scale;
}

template <int MODE>
hipError_t launch_mode(const gram_args& a, hipStream_t st) {
    const int T = (a.n + 63) / 64;
    const int blocks = T * (T + 1) / 2;
    if (blocks == 0) return hipSuccess;
    if (a.y_fp64)
        hipLaunchKernelGGL((k_gram<MODE, double>), dim3(blocks), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_gram<MODE, float>), dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace

hipError_t nq_launch_gram(const gram_args& a, int mode, hipStream_t st) {
    switch (mode) {
        case NQ_GRAM_F64: return launch_mode<NQ_GRAM_F64>(a, st);
        case NQ_GRAM_F32: return launch_mode<NQ_GRAM_F32>(a, st);
        case NQ_GRAM_BF16: return launch_mode<NQ_GRAM_BF16>(a, st);
    }
    return hipErrorInvalidValue;
}

hipError_t nq_launch_colsum(const void* Y, int y_fp64, int64_t Ns, int R, double* part, int chunks,
                            hipStream_t st) {
    dim3 grid((R + 255) / 256, chunks);
    if (y_fp64)
        hipLaunchKernelGGL((k_colsum<double>), grid, dim3(256), 0, st, (const double*)Y, Ns, R, part, chunks);
    else
        hipLaunchKernelGGL((k_colsum<float>), grid, dim3(256), 0, st, (const float*)Y, Ns, R, part, chunks);
    return hipGetLastError();
}

hipError_t nq_launch_colsum_finish(const double* part, int chunks, int R, double scale, double* mu,
                                   hipStream_t st) {
    hipLaunchKernelGGL(k_colsum_finish, dim3((R + 255) / 256), dim3(256), 0, st, part, chunks, R, scale, mu);
    return hipGetLastError();
}
