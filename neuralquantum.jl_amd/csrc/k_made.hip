// Learned Metropolis proposal: a MADE over the D = 2N doubled bits (two hidden
// ReLU layers of width H = 128, sigmoid Bernoulli conditionals), trained by
// maximum likelihood with Adam on the sampler's own output, and used as an
// independence proposal mixed into Metropolis-Hastings with the exact ratio
// p(x') q(x) / (p(x) q(x')) (the reference's `log_prob_bias` hook,
// Metropolis.jl:54-67,149).  Master weights FP32 (masked entries held at zero);
// every matrix product runs on v_mfma_f32_16x16x32_bf16 with BF16 operands and
// FP32 accumulation, operands staged in LDS.
//
// One workgroup (4 waves) handles 64 configurations.  Parameter layout
// (row-major): W1[H][D] b1[H] W2[H][H] b2[H] W3[D][H] b3[D].
#include "nq_kernels.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int MH = 128;   // hidden width
constexpr int MB = 64;    // configurations per workgroup

// C (M x N) = A (M x K) B (K x N); element (m, k) of A at A[m * sam + k * sak],
// (k, n) of B at B[k * sbk + n * sbn]; M, N multiples of 16, K of 32.  The 16x16
// output tiles are spread over the 4 waves; epi(m, n, value) consumes them.
// An operand whose K index is contiguous (stride 1, rows 16-byte aligned: the
// forward products) is read 8 values at a time (ds_read_b128); the
// transposed operands of the backward products are gathered element by element
template <class Epi>
__device__ __forceinline__ void mma_lds(const __bf16* A, int sam, int sak, const __bf16* B, int sbk, int sbn,
                                        int M, int N, int K, Epi&& epi) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int tn = N / 16, tiles = (M / 16) * tn;
    const bool av = sak == 1 && (sam & 7) == 0, bv = sbk == 1 && (sbn & 7) == 0;
    for (int t = wave; t < tiles; t += 4) {
        const int m0 = (t / tn) * 16, n0 = (t % tn) * 16;
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < K; k0 += 32) {
            bf8 a, b;
            if (av) {
                a = *reinterpret_cast<const bf8*>(A + (m0 + li) * sam + k0 + lk * 8);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) a[e] = A[(m0 + li) * sam + (k0 + lk * 8 + e) * sak];
            }
            if (bv) {
                b = *reinterpret_cast<const bf8*>(B + (n0 + li) * sbn + k0 + lk * 8);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) b[e] = B[(k0 + lk * 8 + e) * sbk + (n0 + li) * sbn];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) epi(m0 + 4 * lk + v, n0 + li, acc[v]);
    }
}

__device__ __forceinline__ int mdeg(int k, int D) { return (k % (D - 1)) + 1; }

struct made_lds {
    __bf16 W1[MH * 32];
    __bf16 W2[MH * MH];
    __bf16 W3[32 * MH];
    __bf16 X[MB * 32];
    __bf16 h1[MB * MH];
    __bf16 h2[MB * MH];
    __bf16 dl[MB * 32];
    __bf16 da2[MB * MH];
    __bf16 da1[MB * MH];
    float Lg[MB * 32];
    float Ls[MB * 32];    // sampling: logit d of each row, saved at step d
    float b1[MH], b2[MH], b3[32];
    float red[MH];
};

__device__ void load_weights(made_lds& L, const float* th, int D) {
    const float* W1 = th;
    const float* b1 = W1 + MH * D;
    const float* W2 = b1 + MH;
    const float* b2 = W2 + MH * MH;
    const float* W3 = b2 + MH;
    const float* b3 = W3 + D * MH;
    for (int e = threadIdx.x; e < MH * 32; e += blockDim.x) {
        const int h = e / 32, d = e % 32;
        L.W1[e] = (__bf16)(d < D ? W1[h * D + d] : 0.f);
    }
    for (int e = threadIdx.x; e < MH * MH; e += blockDim.x) L.W2[e] = (__bf16)W2[e];
    for (int e = threadIdx.x; e < 32 * MH; e += blockDim.x) {
        const int d = e / MH;
        L.W3[e] = (__bf16)(d < D ? W3[e] : 0.f);
    }
    for (int e = threadIdx.x; e < MH; e += blockDim.x) { L.b1[e] = b1[e]; L.b2[e] = b2[e]; }
    for (int e = threadIdx.x; e < 32; e += blockDim.x) L.b3[e] = e < D ? b3[e] : 0.f;
}

// h1 = relu(X W1^T + b1), h2 = relu(h1 W2^T + b2) for the 64 rows of L.X
__device__ void made_hidden(made_lds& L) {
    mma_lds(L.X, 32, 1, L.W1, 1, 32, MB, MH, 32, [&](int m, int n, float v) {
        L.h1[m * MH + n] = (__bf16)fmaxf(v + L.b1[n], 0.f);
    });
    __syncthreads();
    mma_lds(L.h1, MH, 1, L.W2, 1, MH, MB, MH, MH, [&](int m, int n, float v) {
        L.h2[m * MH + n] = (__bf16)fmaxf(v + L.b2[n], 0.f);
    });
    __syncthreads();
}

__device__ __forceinline__ float softplusf(float x) { return fmaxf(x, 0.f) + log1pf(expf(-fabsf(x))); }

// all D logits of the 64 rows into L.Lg (the one code path every log q uses,
// so sampling-time and evaluation-time log q agree bit for bit)
__device__ void made_logits(made_lds& L) {
    mma_lds(L.h2, MH, 1, L.W3, 1, MH, MB, 32, MH, [&](int m, int n, float v) { L.Lg[m * 32 + n] = v + L.b3[n]; });
    __syncthreads();
}

// log q of row m from L.Lg and L.X, summed over d in order
__device__ __forceinline__ float made_row_logq(const made_lds& L, int m, int D) {
    float acc = 0.f;
    for (int d = 0; d < D; ++d) {
        const float l = L.Lg[m * 32 + d];
        acc += (float)L.X[m * 32 + d] * l - softplusf(l);
    }
    return acc;
}

// ---- one training step: per-workgroup partial gradients of the mean NLL
__global__ __launch_bounds__(256) void k_made_grad(const float* th, int D, const uint64_t* keys, int B,
                                                   float scale, float* gpart, float* lpart) {
    extern __shared__ __align__(16) unsigned char smem[];
    made_lds& L = *reinterpret_cast<made_lds*>(smem);
    load_weights(L, th, D);
    const int s0 = blockIdx.x * MB;
    for (int e = threadIdx.x; e < MB * 32; e += blockDim.x) {
        const int s = e / 32, d = e % 32;
        const uint64_t k = (s0 + s < B) ? keys[s0 + s] : 0ull;
        L.X[e] = (__bf16)((d < D && s0 + s < B) ? (float)((k >> d) & 1ull) : 0.f);
    }
    __syncthreads();
    made_hidden(L);
    const int P = MH * D + MH + MH * MH + MH + D * MH + D;
    float* g = gpart + (size_t)blockIdx.x * P;
    float* gW1 = g;
    float* gb1 = gW1 + MH * D;
    float* gW2 = gb1 + MH;
    float* gb2 = gW2 + MH * MH;
    float* gW3 = gb2 + MH;
    float* gb3 = gW3 + D * MH;
    // logits, dl = (sigma(l) - x) scale, log-likelihood
    made_logits(L);
    for (int e = threadIdx.x; e < MB * 32; e += blockDim.x) {
        const int m = e / 32, n = e % 32;
        const bool live = n < D && s0 + m < B;
        const float l = L.Lg[e];
        const float x = (float)L.X[e];
        L.dl[e] = (__bf16)(live ? (1.f / (1.f + expf(-l)) - x) * scale : 0.f);
    }
    if (threadIdx.x < MB) L.red[threadIdx.x] = (s0 + threadIdx.x < B) ? made_row_logq(L, threadIdx.x, D) : 0.f;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int m = 0; m < MB; ++m) t += L.red[m];
        lpart[blockIdx.x] = t;
    }
    // dW3 = dl^T h2 (masked: d + 1 > m2(k)), db3 = column sums of dl
    mma_lds(L.dl, 1, 32, L.h2, MH, 1, 32, MH, MB, [&](int d, int k, float v) {
        if (d < D) gW3[d * MH + k] = (d + 1 > mdeg(k, D)) ? v : 0.f;
    });
    if (threadIdx.x < D) {
        float acc = 0.f;
        for (int s = 0; s < MB; ++s) acc += (float)L.dl[s * 32 + threadIdx.x];
        gb3[threadIdx.x] = acc;
    }
    // dh2 = dl W3, da2 = dh2 * relu'(a2)
    mma_lds(L.dl, 32, 1, L.W3, MH, 1, MB, MH, 32, [&](int m, int n, float v) {
        L.da2[m * MH + n] = (__bf16)(((float)L.h2[m * MH + n] > 0.f) ? v : 0.f);
    });
    __syncthreads();
    // dW2 = da2^T h1 (masked: m2(k') >= m1(k)), db2
    mma_lds(L.da2, 1, MH, L.h1, MH, 1, MH, MH, MB, [&](int kp, int k, float v) {
        gW2[kp * MH + k] = (mdeg(kp, D) >= mdeg(k, D)) ? v : 0.f;
    });
    if (threadIdx.x < MH) {
        float acc = 0.f;
        for (int s = 0; s < MB; ++s) acc += (float)L.da2[s * MH + threadIdx.x];
        gb2[threadIdx.x] = acc;
    }
    // dh1 = da2 W2, da1 = dh1 * relu'(a1)
    mma_lds(L.da2, MH, 1, L.W2, MH, 1, MB, MH, MH, [&](int m, int n, float v) {
        L.da1[m * MH + n] = (__bf16)(((float)L.h1[m * MH + n] > 0.f) ? v : 0.f);
    });
    __syncthreads();
    // dW1 = da1^T X (masked: m1(k) >= d + 1), db1
    mma_lds(L.da1, 1, MH, L.X, 32, 1, MH, 32, MB, [&](int k, int d, float v) {
        if (d < D) gW1[k * D + d] = (mdeg(k, D) >= d + 1) ? v : 0.f;
    });
    if (threadIdx.x < MH) {
        float acc = 0.f;
        for (int s = 0; s < MB; ++s) acc += (float)L.da1[s * MH + threadIdx.x];
        gb1[threadIdx.x] = acc;
    }
}

// ---- Adam on the summed partial gradients (Kingma & Ba, 2015)
__global__ void k_made_adam(float* th, float* m, float* v, const float* gpart, int nparts, int P, float lr,
                            float b1, float b2, float eps, int t) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P) return;
    float g = 0.f;
    for (int w = 0; w < nparts; ++w) g += gpart[(size_t)w * P + k];
    const float mk = b1 * m[k] + (1.f - b1) * g;
    const float vk = b2 * v[k] + (1.f - b2) * g * g;
    m[k] = mk;
    v[k] = vk;
    const float mh = mk / (1.f - powf(b1, (float)t));
    const float vh = vk / (1.f - powf(b2, (float)t));
    th[k] -= lr * mh / (sqrtf(vh) + eps);
}

// ---- log q of given configurations
__global__ __launch_bounds__(256) void k_made_logq(const float* th, int D, const uint64_t* keys, int64_t n,
                                                   float* lq) {
    extern __shared__ __align__(16) unsigned char smem[];
    made_lds& L = *reinterpret_cast<made_lds*>(smem);
    load_weights(L, th, D);
    const int64_t s0 = (int64_t)blockIdx.x * MB;
    for (int e = threadIdx.x; e < MB * 32; e += blockDim.x) {
        const int s = e / 32, d = e % 32;
        const uint64_t k = (s0 + s < n) ? keys[s0 + s] : 0ull;
        L.X[e] = (__bf16)((d < D) ? (float)((k >> d) & 1ull) : 0.f);
    }
    __syncthreads();
    made_hidden(L);
    made_logits(L);
    if (threadIdx.x < MB && s0 + threadIdx.x < n) lq[s0 + threadIdx.x] = made_row_logq(L, threadIdx.x, D);
}

// ---- ancestral sampling of n configurations with their log q;
// uniforms from Philox, counter (sample, d, iteration, 3 + 16 tag)
__global__ __launch_bounds__(256) void k_made_sample(const float* th, int D, int64_t n, uint64_t seed, uint32_t iter,
                                                     const uint32_t* iter_dev, uint32_t tag, uint64_t* keys,
                                                     float* lq) {
    extern __shared__ __align__(16) unsigned char smem[];
    made_lds& L = *reinterpret_cast<made_lds*>(smem);
    load_weights(L, th, D);
    const int64_t s0 = (int64_t)blockIdx.x * MB;
    for (int e = threadIdx.x; e < MB * 32; e += blockDim.x) L.X[e] = (__bf16)0.f;
    __syncthreads();
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t it = iter_dev ? *iter_dev : iter;
    // logit d depends on x_<d only (exactly: masked weights are zero), so the
    // logits computed at step d equal those of the finished configuration
    for (int d = 0; d < D; ++d) {
        made_hidden(L);
        made_logits(L);
        if (threadIdx.x < MB) {
            const int row = threadIdx.x;
            const float l = L.Lg[row * 32 + d];
            const int64_t sid = s0 + row;
            philox_out o = philox4x32_10((uint32_t)sid, (uint32_t)d, it, 3u + 16u * tag, k0, k1);
            const float u = u01_24(o.x[0]);
            const float p = 1.f / (1.f + expf(-l));
            L.X[row * 32 + d] = (__bf16)((u < p) ? 1.f : 0.f);
            L.Ls[row * 32 + d] = l;
        }
        __syncthreads();
    }
    if (threadIdx.x < MB && s0 + threadIdx.x < n) {
        const int row = threadIdx.x;
        uint64_t k = 0;
        float acc = 0.f;
        for (int d = 0; d < D; ++d) {
            const float x = (float)L.X[row * 32 + d];
            const float l = L.Ls[row * 32 + d];
            acc += x * l - softplusf(l);
            if (x > 0.5f) k |= 1ull << d;
        }
        keys[s0 + row] = k;
        if (lq) lq[s0 + row] = acc;
    }
}

// ---- sum of the per-workgroup partial gradients (and log-likelihoods)
__global__ void k_made_sum_parts(const float* gpart, int nparts, int P, float* g, const float* lpart, float* nll,
                                 float inv_b) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < P) {
        float acc = 0.f;
        for (int w = 0; w < nparts; ++w) acc += gpart[(size_t)w * P + k];
        g[k] = acc;
    }
    if (k == 0 && nll) {
        float acc = 0.f;
        for (int w = 0; w < nparts; ++w) acc += lpart[w];
        *nll = -acc * inv_b;
    }
}

// ---- independence move: chain c proposes y_c (drawn from q) and accepts
// iff u - exp(log p(y) - log p(x) + log q(x) - log q(y)) >= 0 is false,
// log p = 2 Re log rho; keys[0, n) are the chains' states, keys[n, 2n) the
// proposals, lp and lq hold their log rho (complex, T) and log q.
template <typename T>
__global__ void k_made_accept(const uint64_t* keys, const T* lp, const float* lq, int64_t n, uint64_t seed,
                              uint32_t iter, const uint32_t* iter_dev, uint32_t slice, uint64_t* chain_state,
                              unsigned long long* accepted, uint32_t* chain_acc) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool acc = false;
    if (c < n) {
        const uint32_t it = iter_dev ? *iter_dev : iter;
        const double dlp = 2.0 * ((double)lp[2 * (n + c)] - (double)lp[2 * c]);
        const double dlq = (double)lq[c] - (double)lq[n + c];
        philox_out o = philox4x32_10((uint32_t)c, slice, it, 4u, (uint32_t)seed, (uint32_t)(seed >> 32));
        const double u = u01_53(o.x[1], o.x[2]);
        acc = !(u - exp(dlp + dlq) >= 0.0);
        if (acc) {
            chain_state[c] = keys[n + c];
            if (chain_acc) chain_acc[c] += 1u;
        }
    }
    const unsigned long long m = __ballot(acc);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(accepted, (unsigned long long)__popcll(m));
}

// ---- dst[i] = src[i], i < n
__global__ void k_copy_keys(const uint64_t* src, int64_t n, uint64_t* dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ---- append n keys to a ring of `cap` entries; pos[0] = write position,
// pos[1] = number of valid entries (both kept on the device so the push can
// sit inside a captured graph)
__global__ void k_ring_push(const uint64_t* src, int64_t n, uint64_t* ring, int64_t cap, int64_t* pos) {
    const int64_t p0 = pos[0];
    const int64_t m = n < cap ? n : cap;
    const int64_t off = n - m;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        ring[(p0 + i) % cap] = src[off + i];
}

__global__ void k_ring_advance(int64_t n, int64_t cap, int64_t* pos) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int64_t m = n < cap ? n : cap;
        pos[0] = (pos[0] + m) % cap;
        pos[1] = (pos[1] + m) < cap ? (pos[1] + m) : cap;
    }
}

inline unsigned nwg(int64_t n) { return (unsigned)((n + MB - 1) / MB); }

}  // namespace

namespace {
template <class K>
hipError_t allow_big_lds(K kern) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)sizeof(made_lds));
}
hipError_t made_lds_setup() {
    static hipError_t once = [] {
        hipError_t e = allow_big_lds(k_made_grad);
        if (e == hipSuccess) e = allow_big_lds(k_made_logq);
        if (e == hipSuccess) e = allow_big_lds(k_made_sample);
        return e;
    }();
    return once;
}
}  // namespace

size_t nq_made_lds_bytes() { return sizeof(made_lds); }
hipError_t nq_made_prepare() { return made_lds_setup(); }

int nq_made_param_count(int D) { return MH * D + MH + MH * MH + MH + D * MH + D; }
size_t nq_made_grad_parts(int64_t B) { return nwg(B); }

hipError_t nq_launch_made_grad(const float* th, int D, const uint64_t* keys, int B, float* gpart, float* lpart,
                               hipStream_t st) {
    if (D < 2 || D > 32 || B < 1) return hipErrorInvalidValue;
    if (hipError_t e = made_lds_setup()) return e;
    hipLaunchKernelGGL(k_made_grad, dim3(nwg(B)), dim3(256), sizeof(made_lds), st, th, D, keys, B,
                       1.0f / (float)B, gpart, lpart);
    return hipGetLastError();
}

hipError_t nq_launch_made_adam(float* th, float* m, float* v, const float* gpart, int nparts, int D, float lr, int t,
                               hipStream_t st) {
    const int P = nq_made_param_count(D);
    hipLaunchKernelGGL(k_made_adam, dim3((P + 255) / 256), dim3(256), 0, st, th, m, v, gpart, nparts, P, lr, 0.9f,
                       0.999f, 1e-8f, t);
    return hipGetLastError();
}

hipError_t nq_launch_made_logq(const float* th, int D, const uint64_t* keys, int64_t n, float* lq, hipStream_t st) {
    if (D < 2 || D > 32 || n < 1) return hipErrorInvalidValue;
    if (hipError_t e = made_lds_setup()) return e;
    hipLaunchKernelGGL(k_made_logq, dim3(nwg(n)), dim3(256), sizeof(made_lds), st, th, D, keys, n, lq);
    return hipGetLastError();
}

hipError_t nq_launch_made_sample(const float* th, int D, int64_t n, uint64_t seed, uint32_t iter,
                                 const uint32_t* iter_dev, uint32_t tag, uint64_t* keys, float* lq, hipStream_t st) {
    if (D < 2 || D > 32 || n < 1) return hipErrorInvalidValue;
    if (hipError_t e = made_lds_setup()) return e;
    hipLaunchKernelGGL(k_made_sample, dim3(nwg(n)), dim3(256), sizeof(made_lds), st, th, D, n, seed, iter, iter_dev,
                       tag, keys, lq);
    return hipGetLastError();
}

hipError_t nq_launch_made_sum_parts(const float* gpart, int nparts, int D, float* g, const float* lpart, float* nll,
                                    int64_t B, hipStream_t st) {
    const int P = nq_made_param_count(D);
    hipLaunchKernelGGL(k_made_sum_parts, dim3((P + 255) / 256), dim3(256), 0, st, gpart, nparts, P, g, lpart, nll,
                       1.0f / (float)B);
    return hipGetLastError();
}

hipError_t nq_launch_made_accept(const uint64_t* keys, const void* lp, bool lp_fp64, const float* lq, int64_t n,
                                 uint64_t seed, uint32_t iter, const uint32_t* iter_dev, uint32_t slice,
                                 uint64_t* chain_state, unsigned long long* accepted, uint32_t* chain_acc,
                                 hipStream_t st) {
    const unsigned g = (unsigned)((n + 255) / 256);
    if (lp_fp64)
        hipLaunchKernelGGL(k_made_accept<double>, dim3(g), dim3(256), 0, st, keys, (const double*)lp, lq, n, seed,
                           iter, iter_dev, slice, chain_state, accepted, chain_acc);
    else
        hipLaunchKernelGGL(k_made_accept<float>, dim3(g), dim3(256), 0, st, keys, (const float*)lp, lq, n, seed, iter,
                           iter_dev, slice, chain_state, accepted, chain_acc);
    return hipGetLastError();
}

hipError_t nq_launch_copy_keys(const uint64_t* src, int64_t n, uint64_t* dst, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, n, dst);
    return hipGetLastError();
}

hipError_t nq_launch_ring_push(const uint64_t* src, int64_t n, uint64_t* ring, int64_t cap, int64_t* pos,
                               hipStream_t st) {
    const int64_t m = n < cap ? n : cap;
    int64_t g = (m + 255) / 256;
    if (g > 1024) g = 1024;
    hipLaunchKernelGGL(k_ring_push, dim3((unsigned)g), dim3(256), 0, st, src, n, ring, cap, pos);
    hipLaunchKernelGGL(k_ring_advance, dim3(1), dim3(64), 0, st, n, cap, pos);
    return hipGetLastError();
}
