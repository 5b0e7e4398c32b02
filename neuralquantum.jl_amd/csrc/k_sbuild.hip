// S build on the matrix cores: S = Re(Oc Oc^H) / Ns (SRDirect.jl:26-49) from
// the compact O rows.  With the lambda block real, the mu block imaginary and
// the Pi block complex, Re(Oc Oc^H) splits into two symmetric Gram matrices of
// column blocks of Y (rows = samples): A = [Re lambda | Re Pi] and
// B = [Im Pi | Im mu]; S in P order is A^T A embedded at [0, nA) plus B^T B
// embedded at [nlam, np) (the lambda-mu block is identically zero).
//
// FP64 / FP32 (k_gram): each workgroup computes one 64x64 upper-triangular
// tile of a Gram block; the 4 waves own 32x32 quadrants of 2x2 MFMA 16x16
// tiles.  Columns are centred (Y - mean) while staging through LDS, which is
// what the reference's O - Obar does (BaseIterativeSampler.jl:19-26).
//   FP64: v_mfma_f64_16x16x4f64     FP32: v_mfma_f32_16x16x4f32
// BF16 (k_center_t + k_gram_bf16_part + k_gram_finish): a centred, transposed
// BF16 copy, then a split-K Gram of 128x128 tiles with
// v_mfma_f32_16x16x32_bf16 (FP32 accumulation) and one pass that writes S.
#include "nq_kernels.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

template <int MODE> struct gram_traits;
template <> struct gram_traits<NQ_GRAM_F64> { typedef double st; typedef d4 acc; static constexpr int KC = 16; };
template <> struct gram_traits<NQ_GRAM_F32> { typedef float st; typedef f4 acc; static constexpr int KC = 16; };

__device__ __forceinline__ void tri_decode(int b, int* I, int* J) {
    int j = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
    while ((j + 1) * (j + 2) / 2 <= b) ++j;
    while (j * (j + 1) / 2 > b) --j;
    *J = j;
    *I = b - j * (j + 1) / 2;
}

template <int MODE, typename YT>
__device__ __forceinline__ void gram_body(const gram_args& a, int bx) {
    typedef typename gram_traits<MODE>::st ST;
    typedef typename gram_traits<MODE>::acc AT;
    constexpr int KC = gram_traits<MODE>::KC;
    constexpr int PITCH = 65;                 // [k][64 + 1]: k-major, rows contiguous
    constexpr int TILE = KC * PITCH;
    constexpr int NJ = KC / 4;                // sample rows staged per thread and k-step
    __shared__ ST As[2 * TILE];               // double-buffered: one barrier per k-step
    __shared__ ST Bs[2 * TILE];

    int I, J;
    tri_decode(bx, &I, &J);
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
    double ya[NJ], yb[NJ];
    // unconditional loads from clamped addresses (one k-step ahead), centred
    // and selected when stored
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int64_t s = k0 + (tid >> 6) + 4 * j;
            const YT* row = Y + (s < a.Ns ? s : a.Ns - 1) * (int64_t)a.ldy + a.c0;
            ya[j] = (double)row[va ? ca : 0];
            yb[j] = (double)row[vb ? cb : 0];
        }
    };
    auto store = [&](int buf, int64_t k0) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int kk = (tid >> 6) + 4 * j;
            const int64_t s = k0 + kk;
            As[buf * TILE + kk * PITCH + lc] = (ST)((s < a.Ns && va) ? ya[j] - mua : 0.0);
            Bs[buf * TILE + kk * PITCH + lc] = (ST)((s < a.Ns && vb) ? yb[j] - mub : 0.0);
        }
    };
    const int64_t nsteps = (a.Ns + KC - 1) / KC;
    load(0);
    store(0, 0);
    if (nsteps > 1) load(KC);
    __syncthreads();
    const int li = lane & 15, lk = lane >> 4;
    for (int64_t i = 0; i < nsteps; ++i) {
        if (i + 1 < nsteps) store((int)((i + 1) & 1), (i + 1) * KC);
        if (i + 2 < nsteps) load((i + 2) * KC);
        const ST* A = As + (i & 1) * TILE;
        const ST* B = Bs + (i & 1) * TILE;
#pragma unroll
        for (int k = 0; k < KC; k += 4) {
            ST av[2], bv[2];
#pragma unroll
            for (int m = 0; m < 2; ++m) av[m] = A[(k + lk) * PITCH + wr * 32 + m * 16 + li];
#pragma unroll
            for (int n = 0; n < 2; ++n) bv[n] = B[(k + lk) * PITCH + wc * 32 + n * 16 + li];
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
        __syncthreads();
    }

    // D[i][j] of a 16x16 tile, j = lane % 16:  f32: i = 4 (lane / 16) + v;  f64: i = lane / 16 + 4 v
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

// ---- BF16 path: centred, transposed copy Yt[c][s] = bf16(Y[s][c0 + c] - mu),
// samples padded with zeros to a multiple of 64, then a 128x128-tile Gram.
template <typename YT>
__device__ __forceinline__ void center_t_body(const YT* Y, int64_t Ns, int ldy, int c0, int n, const double* mu,
                                              __bf16* Yt, int64_t nsp, int bx, int by) {
    __shared__ float tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int cb = bx * 64;
    const int64_t sb = (int64_t)by * 64;
    const int c = cb + tx;
    const double m = (c < n && mu) ? mu[c0 + c] : 0.0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int64_t si = sb + ty + 4 * r;
        float v = 0.f;
        const double y = (double)Y[min(si, Ns - 1) * ldy + c0 + min(c, n - 1)];   // unconditional load
        if (c < n && si < Ns) v = (float)(y - m);
        tile[ty + 4 * r][tx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int cc = cb + ty + 4 * r;
        const int64_t si = sb + tx;
        if (cc < n && si < nsp) Yt[(int64_t)cc * nsp + si] = (__bf16)tile[tx][ty + 4 * r];
    }
}

// ---- split-K BF16 Gram.  Pass 1: each workgroup takes one 128x128 upper
// tile of a Gram block (block A = Yt rows [0, nA), block B = rows
// [nA, nA + nB)) over one chunk of the samples and writes its FP32 partial
// sums (row-major 128 x 128) to `part`; splitting the sample axis keeps
// every CU busy with a few tiles in flight.  LDS is double-buffered with a
// one-step register prefetch (one barrier per k-step of 32).
// Pass 2 (k_gram_finish): sums the chunks in a fixed order (deterministic,
// no atomics), scales, embeds block A at S[0, nA) and block B at
// S[offB, offB + nB) (adding where they overlap), writes every element of
// the (np + 1)^2 matrix (zeros outside the blocks) and mirrors the lower
// triangle through an LDS transpose so both writes are coalesced.
__device__ __forceinline__ void gram_bf16_part_body(const __bf16* Yt, int64_t nsp, int nA, int TA, int nB,
                                                    int64_t kchunk, int tiles, float* part, int bx, int by) {
    constexpr int KS = 32, PITCH = 40, BUF = 128 * PITCH;
    __shared__ __align__(16) __bf16 As[2 * BUF];
    __shared__ __align__(16) __bf16 Bs[2 * BUF];
    const int tilesA = TA * (TA + 1) / 2;
    const int tile = bx;
    const __bf16* Y = Yt;
    int n = nA, tl = tile;
    if (tile >= tilesA) {
        Y = Yt + (int64_t)nA * nsp;
        n = nB;
        tl = tile - tilesA;
    }
    int I, J;
    tri_decode(tl, &I, &J);
    const int r0 = I * 128, q0 = J * 128;
    const int64_t kbeg = (int64_t)by * kchunk;
    const int64_t kend = (kbeg + kchunk < nsp) ? kbeg + kchunk : nsp;
    const int nsteps = kend > kbeg ? (int)((kend - kbeg) / KS) : 0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int li = lane & 15, lk = lane >> 4;

    f4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[a][b][v] = 0.f;

    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    const int row0 = tid >> 2, ch = tid & 3;            // rows row0 and row0 + 64, 16 B chunk ch
    const bool va0 = r0 + row0 < n, va1 = r0 + row0 + 64 < n, vb0 = q0 + row0 < n, vb1 = q0 + row0 + 64 < n;
    const __bf16* ya0 = Y + (int64_t)min(r0 + row0, n - 1) * nsp + ch * 8;
    const __bf16* ya1 = Y + (int64_t)min(r0 + row0 + 64, n - 1) * nsp + ch * 8;
    const __bf16* yb0 = Y + (int64_t)min(q0 + row0, n - 1) * nsp + ch * 8;
    const __bf16* yb1 = Y + (int64_t)min(q0 + row0 + 64, n - 1) * nsp + ch * 8;
    const u4 z4 = u4{0u, 0u, 0u, 0u};
    // k-steps are loaded PFD steps ahead into a register ring: raw loads from
    // clamped rows and steps (no select on a loaded value, which would wait
    // for it at once), zeroed for rows past the block when staged in LDS
    constexpr int PFD = 4;
    u4 ra0[PFD], ra1[PFD], rb0[PFD], rb1[PFD];
    auto load = [&](int step, int slot) {
        const int64_t k0 = kbeg + (int64_t)min(step, max(nsteps - 1, 0)) * KS;
        ra0[slot] = *reinterpret_cast<const u4*>(ya0 + k0);
        ra1[slot] = *reinterpret_cast<const u4*>(ya1 + k0);
        rb0[slot] = *reinterpret_cast<const u4*>(yb0 + k0);
        rb1[slot] = *reinterpret_cast<const u4*>(yb1 + k0);
    };
    auto store = [&](int buf, int slot) {
        __bf16* A = As + buf * BUF;
        __bf16* B = Bs + buf * BUF;
        *reinterpret_cast<u4*>(&A[row0 * PITCH + ch * 8]) = va0 ? ra0[slot] : z4;
        *reinterpret_cast<u4*>(&A[(row0 + 64) * PITCH + ch * 8]) = va1 ? ra1[slot] : z4;
        *reinterpret_cast<u4*>(&B[row0 * PITCH + ch * 8]) = vb0 ? rb0[slot] : z4;
        *reinterpret_cast<u4*>(&B[(row0 + 64) * PITCH + ch * 8]) = vb1 ? rb1[slot] : z4;
    };
    if (nsteps > 0) {
#pragma unroll
        for (int u = 0; u < PFD; ++u) load(u, u);
        store(0, 0);
        __syncthreads();
        // whole rounds of PFD steps (LDS buffer i & 1 = u & 1, ring slot i %
        // PFD = u); the products of steps past nsteps are skipped
        for (int i0 = 0; i0 < nsteps; i0 += PFD) {
#pragma unroll
            for (int u = 0; u < PFD; ++u) {
                const int i = i0 + u;
                const __bf16* A = As + (u & 1) * BUF;
                const __bf16* B = Bs + (u & 1) * BUF;
                bf8 av[4], bv[4];
#pragma unroll
                for (int a = 0; a < 4; ++a)
                    av[a] = *reinterpret_cast<const bf8*>(&A[(wr * 64 + a * 16 + li) * PITCH + lk * 8]);
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    bv[b] = *reinterpret_cast<const bf8*>(&B[(wc * 64 + b * 16 + li) * PITCH + lk * 8]);
                store((u + 1) & 1, (u + 1) % PFD);   // step i + 1, loaded PFD - 1 steps ago
                load(i + PFD, u);                    // slot u (step i) was staged one step ago
                if (i < nsteps) {
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int b = 0; b < 4; ++b)
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
                }
                __syncthreads();
            }
        }
    }
    float* P = part + ((int64_t)by * tiles + tile) * (128 * 128);
    const int lj = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                P[(wr * 64 + a * 16 + 4 * lg + v) * 128 + wc * 64 + b * 16 + lj] = acc[a][b][v];
}

// sum over the chunks of element (rr, cc) of the Gram block whose tiles start at toff
// offset of element (rr, cc) of a Gram block in split 0 of the partials
// (upper tiles only: the lower half is read transposed)
__device__ __forceinline__ int gram_part_off(int toff, int rr, int cc) {
    int I = rr >> 7, J = cc >> 7, x = rr & 127, y = cc & 127;
    if (I > J) {
        const int t0 = I; I = J; J = t0;
        const int t1 = x; x = y; y = t1;
    }
    return (toff + J * (J + 1) / 2 + I) * (128 * 128) + x * 128 + y;
}

__device__ __forceinline__ void gram_finish_body(const float* part, int splits, int tiles, int nA, int TA, int nB,
                                                 int offB, double scale, double* S, int np, int lds, int bx) {
    __shared__ double T[64][65];
    int bi, bj;
    tri_decode(bx, &bi, &bj);   // bi <= bj: 64 x 64 tiles of the (np + 1)^2 matrix
    const int y = threadIdx.x & 63, x0 = threadIdx.x >> 6;
    const int c = bj * 64 + y;
    const int tilesA = TA * (TA + 1) / 2;
    // the 16 elements' offsets in each block (-1: outside it), then the
    // chunks in order with all 16 (or 32) loads of a chunk in flight
    int oa[16], ob[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int r = bi * 64 + x0 + 4 * m;
        const bool in = r < np && c < np;
        oa[m] = (in && r < nA && c < nA) ? gram_part_off(0, r, c) : -1;
        ob[m] = (in && nB > 0 && r >= offB && c >= offB && r - offB < nB && c - offB < nB)
                    ? gram_part_off(tilesA, r - offB, c - offB)
                    : -1;
    }
    const int64_t stride = (int64_t)tiles * (128 * 128);
    double sa[16], sb[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) sa[m] = sb[m] = 0.0;
    for (int sp = 0; sp < splits; ++sp) {
        const float* p = part + sp * stride;
        float la[16], lb[16];   // unconditional loads (offset 0 stands in), selected after
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            la[m] = p[oa[m] >= 0 ? oa[m] : 0];
            lb[m] = p[ob[m] >= 0 ? ob[m] : 0];
        }
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            sa[m] += oa[m] >= 0 ? (double)la[m] : 0.0;
            sb[m] += ob[m] >= 0 ? (double)lb[m] : 0.0;
        }
    }
    double v[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = scale * (sa[m] + sb[m]);
    // S[c, r] (coalesced in c); on a diagonal tile this is every element
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int r = bi * 64 + x0 + 4 * m;
        if (r <= np && c <= np) S[(int64_t)c + (int64_t)r * lds] = v[m];
    }
    if (bi != bj) {
        // S[r, c] through an LDS transpose (coalesced in r)
#pragma unroll
        for (int m = 0; m < 16; ++m) T[x0 + 4 * m][y] = v[m];
        __syncthreads();
        const int r = bi * 64 + y;
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int cc = bj * 64 + x0 + 4 * m;
            if (r <= np && cc <= np) S[(int64_t)r + (int64_t)cc * lds] = T[y][x0 + 4 * m];
        }
    }
}

template <typename YT>
__device__ __forceinline__ void colsum_body(const YT* Y, int64_t Ns, int R, double* part, int chunks, int bx, int by) {
    const int c = bx * blockDim.x + threadIdx.x;
    const int chunk = by;
    if (c >= R) return;
    const int64_t per = (Ns + chunks - 1) / chunks;
    const int64_t s0 = chunk * per, s1 = (s0 + per < Ns) ? s0 + per : Ns;
    // four independent sums, unrolled: many loads in flight per thread
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t s = s0;
#pragma unroll 4
    for (; s + 4 <= s1; s += 4) {
        a0 += (double)Y[s * R + c];
        a1 += (double)Y[(s + 1) * R + c];
        a2 += (double)Y[(s + 2) * R + c];
        a3 += (double)Y[(s + 3) * R + c];
    }
    for (; s < s1; ++s) a0 += (double)Y[s * R + c];
    part[(int64_t)chunk * R + c] = (a0 + a1) + (a2 + a3);
}

// mu[c] = scale * sum_k part[k][c]: 64 columns per workgroup, the chunks split
// over 4 thread groups (coalesced along the columns), combined in a fixed order
__device__ __forceinline__ void colsum_finish_body(const double* part, int chunks, int R, double scale, double* mu,
                                                   int bx) {
    __shared__ double red[4][64];
    const int x = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int c = bx * 64 + x;
    const int cc = c < R ? c : R - 1;
    double a0 = 0.0, a1 = 0.0;
    int k = g;
    for (; k + 4 < chunks; k += 8) {
        a0 += part[(int64_t)k * R + cc];
        a1 += part[(int64_t)(k + 4) * R + cc];
    }
    if (k < chunks) a0 += part[(int64_t)k * R + cc];
    red[g][x] = a0 + a1;
    __syncthreads();
    if (g == 0 && c < R) mu[c] = ((red[0][x] + red[1][x]) + (red[2][x] + red[3][x])) * scale;
}

template <int MODE, typename YT>
__global__ __launch_bounds__(256) void k_gram(gram_args a) {
    gram_body<MODE, YT>(a, blockIdx.x);
}

template <typename YT>
__global__ __launch_bounds__(256) void k_center_t(const YT* Y, int64_t Ns, int ldy, int c0, int n,
                                                  const double* mu, __bf16* Yt, int64_t nsp) {
    center_t_body<YT>(Y, Ns, ldy, c0, n, mu, Yt, nsp, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(256) void k_gram_bf16_part(const __bf16* Yt, int64_t nsp, int nA, int TA, int nB,
                                                        int64_t kchunk, int tiles, float* part) {
    gram_bf16_part_body(Yt, nsp, nA, TA, nB, kchunk, tiles, part, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(256) void k_gram_finish(const float* part, int splits, int tiles, int nA, int TA, int nB,
                                                     int offB, double scale, double* S, int np, int lds) {
    gram_finish_body(part, splits, tiles, nA, TA, nB, offB, scale, S, np, lds, blockIdx.x);
}

template <typename YT>
__global__ __launch_bounds__(256) void k_colsum(const YT* Y, int64_t Ns, int R, double* part, int chunks) {
    colsum_body<YT>(Y, Ns, R, part, chunks, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(256) void k_colsum_finish(const double* part, int chunks, int R, double scale,
                                                       double* mu) {
    colsum_finish_body(part, chunks, R, scale, mu, blockIdx.x);
}

// ---- packed forms (block records and descriptors as in nq_pk_member)
__device__ __forceinline__ const nq_pk_member& pk_member(const nq_pk_member* M, uint2 b) { return M[b.x & 0xffff]; }

// the FP64 / FP32 Gram block `which` of a member (build_s: block A then B, adding into S)
__host__ __device__ inline gram_args pk_gram_args(const nq_pk_member& m, int which) {
    gram_args g{};
    g.Y = m.Y;
    g.y_fp64 = m.fp64;
    g.Ns = m.Ns;
    g.ldy = m.R;
    g.mu = m.mu;
    g.S = m.S;
    g.lds = m.ldS;
    g.scale = 1.0 / (double)m.Ns;
    g.c0 = which == 0 ? 0 : m.nA;
    g.n = which == 0 ? m.nA : m.nB;
    g.o = which == 0 ? 0 : m.offB;
    return g;
}

template <int MODE, typename YT>
__global__ __launch_bounds__(256) void k_gram_pk(const nq_pk_member* M, const uint2* blk, int which) {
    const uint2 b = blk[blockIdx.x];
    const gram_args a = pk_gram_args(pk_member(M, b), which);
    gram_body<MODE, YT>(a, b.y & 0xffff);
}

template <typename YT>
__global__ __launch_bounds__(256) void k_colsum_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_member(M, b);
    colsum_body<YT>((const YT*)m.Y, m.Ns, m.R, m.colpart, m.col_chunks, b.y & 0xffff, b.y >> 16);
}

__global__ __launch_bounds__(256) void k_colsum_finish_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_member(M, b);
    colsum_finish_body(m.colpart, m.col_chunks, m.R, 1.0 / (double)m.Ns, m.mu, b.y & 0xffff);
}

template <typename YT>
__global__ __launch_bounds__(256) void k_center_t_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_member(M, b);
    center_t_body<YT>((const YT*)m.Y, m.Ns, m.R, 0, m.R, m.mu, m.Yt, m.nsp, b.y & 0xffff, b.y >> 16);
}

__global__ __launch_bounds__(256) void k_gram_bf16_part_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_member(M, b);
    gram_bf16_part_body(m.Yt, m.nsp, m.nA, m.TA, m.nB, m.kchunk, m.tiles, m.gpart, b.y & 0xffff, b.y >> 16);
}

__global__ __launch_bounds__(256) void k_gram_finish_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = pk_member(M, b);
    gram_finish_body(m.gpart, m.splits, m.tiles, m.nA, m.TA, m.nB, m.offB, 1.0 / (double)m.Ns, m.S, m.np, m.ldS,
                     b.y & 0xffff);
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
    }
    return hipErrorInvalidValue;
}

hipError_t nq_launch_center_t(const void* Y, int y_fp64, int64_t Ns, int ldy, int c0, int n, const double* mu,
                              __bf16* Yt, int64_t nsp, hipStream_t st) {
    dim3 grid((n + 63) / 64, (unsigned)((nsp + 63) / 64));
    if (y_fp64)
        hipLaunchKernelGGL((k_center_t<double>), grid, dim3(256), 0, st, (const double*)Y, Ns, ldy, c0, n, mu, Yt, nsp);
    else
        hipLaunchKernelGGL((k_center_t<float>), grid, dim3(256), 0, st, (const float*)Y, Ns, ldy, c0, n, mu, Yt, nsp);
    return hipGetLastError();
}

// split-K chunks of the BF16 Gram: max_splits when given (a tuned shape),
// otherwise about three tiles in flight per CU, with the partial sums (64 KB
// per tile and chunk) kept within twice the BF16 input so that small
// problems (many of them sharing the device) do not turn into partial-sum
// traffic; at most one chunk per 64 samples and 64 chunks
int nq_gram_split_plan(int nA, int nB, int64_t nsp, int max_splits, int64_t* kchunk) {
    const int TA = (nA + 127) / 128, TB = (nB + 127) / 128;
    const int tiles = TA * (TA + 1) / 2 + TB * (TB + 1) / 2;
    int s;
    if (max_splits > 0) {
        s = max_splits;
    } else {
        s = tiles > 0 ? (3 * 256 + tiles - 1) / tiles : 1;
        const int64_t in_bytes = 2 * nsp * (int64_t)(nA + nB);
        const int64_t cap = tiles > 0 ? 2 * in_bytes / ((int64_t)tiles * 128 * 128 * 4) : 1;
        if (cap < s) s = cap > 1 ? (int)cap : 1;
    }
    const int maxs = nsp >= 64 ? (int)(nsp / 64) : 1;
    s = s < maxs ? s : maxs;
    s = s < 64 ? s : 64;
    s = s > 1 ? s : 1;
    int64_t kc = (nsp + s - 1) / s;
    kc = (kc + 63) / 64 * 64;
    if (kc < 64) kc = 64;
    *kchunk = kc;
    return (int)((nsp + kc - 1) / kc);
}

size_t nq_gram_split_part_floats(int nA, int nB, int64_t nsp, int max_splits) {
    int64_t kc;
    const int splits = nq_gram_split_plan(nA, nB, nsp, max_splits, &kc);
    const int TA = (nA + 127) / 128, TB = (nB + 127) / 128;
    const int tiles = TA * (TA + 1) / 2 + TB * (TB + 1) / 2;
    return (size_t)splits * tiles * 128 * 128;
}

hipError_t nq_launch_gram2_bf16(const __bf16* Yt, int64_t nsp, int nA, int nB, int offB, double scale, float* part,
                                double* S, int np, int lds, int max_splits, hipStream_t st) {
    int64_t kc;
    const int splits = nq_gram_split_plan(nA, nB, nsp, max_splits, &kc);
    const int TA = (nA + 127) / 128, TB = (nB + 127) / 128;
    const int tiles = TA * (TA + 1) / 2 + TB * (TB + 1) / 2;
    if (tiles > 0)
        hipLaunchKernelGGL(k_gram_bf16_part, dim3(tiles, splits), dim3(256), 0, st, Yt, nsp, nA, TA, nB, kc, tiles,
                           part);
    const int TS = (np + 1 + 63) / 64;
    hipLaunchKernelGGL(k_gram_finish, dim3(TS * (TS + 1) / 2), dim3(256), 0, st, part, splits, tiles, nA, TA, nB, offB,
                       scale, S, np, lds);
    return hipGetLastError();
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
    hipLaunchKernelGGL(k_colsum_finish, dim3((R + 63) / 64), dim3(256), 0, st, part, chunks, R, scale, mu);
    return hipGetLastError();
}

void nq_gram_shape(int nA, int nB, int64_t nsp, int max_splits, int* TA, int* tiles, int* splits, int64_t* kchunk) {
    *splits = nq_gram_split_plan(nA, nB, nsp, max_splits, kchunk);
    *TA = (nA + 127) / 128;
    const int TB = (nB + 127) / 128;
    *tiles = *TA * (*TA + 1) / 2 + TB * (TB + 1) / 2;
}

hipError_t nq_pk_sbuild(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st) {
    if (L.nblk <= 0) return hipSuccess;
    const dim3 g((unsigned)L.nblk), b((unsigned)L.threads);
    const bool f64 = L.variant != 0;
    switch (L.kernel) {
        case NQ_PK_COLSUM:
            if (f64) hipLaunchKernelGGL((k_colsum_pk<double>), g, b, 0, st, M, L.blk);
            else hipLaunchKernelGGL((k_colsum_pk<float>), g, b, 0, st, M, L.blk);
            break;
        case NQ_PK_COLSUM_FIN: hipLaunchKernelGGL(k_colsum_finish_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_CENTER:
            if (f64) hipLaunchKernelGGL((k_center_t_pk<double>), g, b, 0, st, M, L.blk);
            else hipLaunchKernelGGL((k_center_t_pk<float>), g, b, 0, st, M, L.blk);
            break;
        case NQ_PK_GRAM_PART: hipLaunchKernelGGL(k_gram_bf16_part_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_GRAM_FIN: hipLaunchKernelGGL(k_gram_finish_pk, g, b, 0, st, M, L.blk); break;
        case NQ_PK_GRAM:
            // variant: S-build mode | FP64 rows << 1
            switch (L.variant) {
                case 0: hipLaunchKernelGGL((k_gram_pk<NQ_GRAM_F64, float>), g, b, 0, st, M, L.blk, L.param); break;
                case 1: hipLaunchKernelGGL((k_gram_pk<NQ_GRAM_F32, float>), g, b, 0, st, M, L.blk, L.param); break;
                case 2: hipLaunchKernelGGL((k_gram_pk<NQ_GRAM_F64, double>), g, b, 0, st, M, L.blk, L.param); break;
                case 3: hipLaunchKernelGGL((k_gram_pk<NQ_GRAM_F32, double>), g, b, 0, st, M, L.blk, L.param); break;
                default: return hipErrorInvalidValue;
            }
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int nq_pk_gram_blocks(const nq_pk_member& m, int which) {
    const int T = (pk_gram_args(m, which).n + 63) / 64;
    return T * (T + 1) / 2;
}
