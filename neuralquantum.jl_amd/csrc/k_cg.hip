// Matrix-free SR: S v = (1/Ns) sum_s Oc(s) (Oc(s) . v) + shift, without S.
//
// Reference semantics: SrMatrixR (SR_notfull.jl:84-148): mul!(res, O, O' v)
// over the real and imaginary parts, res / N + shift v; solved by the
// iterative solvers of SRIterative.jl:71-152 (here conjugate gradients,
// sr_cg, with IterativeSolvers' relative-residual tolerance).
//
// The compact O rows Y (Ns x R, row per sample, P-order column blocks) hold
// block A = columns [0, nA) (the real parts) and block B = columns [nA, R)
// (the imaginary parts); column c belongs to parameter c (block A) or
// offB + c - nA (block B), so Re(Oc^H Oc) v = A^T (A v_A) + B^T (B v_B).
// One product is two passes over Y:
//   k_yv      w_A[s] = sum_{c < nA} Yc[s, c] v[c],  w_B[s] likewise over B
//             (one wavefront per sample row, coalesced along the row);
//   k_ytw     per chunk of samples, part[c] = sum_s Yc[s, c] w_blk(c)[s]
//             (threads along the columns, coalesced);
//   k_ytw_finish  q[k] = (sum of the chunks of the columns of k) / Ns
//             + shift(k) v[k].
// Yc = Y - mu (centred on the fly, FP64 arithmetic on FP32 or FP64 rows).
#include "nq_device.h"
#include "nq_kernels.h"

namespace {

template <typename YT>
__global__ __launch_bounds__(256) void k_yv(const YT* Y, int64_t Ns, int R, int nA, int offB, const double* mu,
                                            const double* v, double* w) {
    const int lane = threadIdx.x & 63;
    const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= Ns) return;
    const YT* row = Y + s * R;
    double a = 0.0, b = 0.0;
    for (int c = lane; c < R; c += 64) {
        const double y = (double)row[c] - mu[c];
        if (c < nA) a += y * v[c];
        else b += y * v[offB + c - nA];
    }
    a = wave_total(a);
    b = wave_total(b);
    if (lane == 0) {
        w[2 * s] = a;
        w[2 * s + 1] = b;
    }
}

template <typename YT>
__global__ __launch_bounds__(256) void k_ytw(const YT* Y, int64_t Ns, int R, int nA, const double* mu, const double* w,
                                             int chunks, double* part) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= R) return;
    const int64_t per = (Ns + chunks - 1) / chunks;
    const int64_t s0 = (int64_t)blockIdx.y * per, s1 = (s0 + per < Ns) ? s0 + per : Ns;
    const int blk = c < nA ? 0 : 1;
    const double m = mu[c];
    double acc = 0.0;
    for (int64_t s = s0; s < s1; ++s) acc += ((double)Y[s * R + c] - m) * w[2 * s + blk];
    part[(int64_t)blockIdx.y * R + c] = acc;
}

// q[k] = scale * (column sums of parameter k) + shift(k) v[k]; a parameter
// owns one column of A (k < nA), one of B (k >= offB, k - offB < nB) or both
__global__ __launch_bounds__(256) void k_ytw_finish(const double* part, int chunks, int R, int nA, int offB, int np,
                                                    double scale, double eps, int kind, const double* diag,
                                                    const double* v, double* q) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const int nB = R - nA;
    double acc = 0.0;
    if (k < nA)
        for (int i = 0; i < chunks; ++i) acc += part[(int64_t)i * R + k];
    if (k >= offB && k - offB < nB)
        for (int i = 0; i < chunks; ++i) acc += part[(int64_t)i * R + nA + (k - offB)];
    double sh = 0.0;
    if (v) sh = ((kind == 0) ? eps : eps * diag[k]) * v[k];
    q[k] = scale * acc + sh;
}

// diag(S)[k] = (1/Ns) sum_s (sum over the columns of k of Yc[s, c]^2)
template <typename YT>
__global__ __launch_bounds__(256) void k_colsq(const YT* Y, int64_t Ns, int R, const double* mu, int chunks,
                                               double* part) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= R) return;
    const int64_t per = (Ns + chunks - 1) / chunks;
    const int64_t s0 = (int64_t)blockIdx.y * per, s1 = (s0 + per < Ns) ? s0 + per : Ns;
    const double m = mu[c];
    double acc = 0.0;
    for (int64_t s = s0; s < s1; ++s) {
        const double y = (double)Y[s * R + c] - m;
        acc += y * y;
    }
    part[(int64_t)blockIdx.y * R + c] = acc;
}

// out[slot] = sum_i x[i] y[i] (two stages: per-block partials, then one block)
__global__ __launch_bounds__(256) void k_dot_part(const double* x, const double* y, int n, double* part) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += x[i] * y[i];
    acc = wave_total(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ __launch_bounds__(64) void k_dot_finish(const double* part, int nparts, double* out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 64) acc += part[i];
    acc = wave_total(acc);
    if (threadIdx.x == 0) *out = acc;
}

// CG scalars live in device memory: sc[0] = r.r, sc[1] = p.q, sc[2] = new r.r
// x += alpha p, r -= alpha q with alpha = sc[0] / sc[1]
__global__ __launch_bounds__(256) void k_cg_xr(double* x, double* r, const double* p, const double* q, int n,
                                               const double* sc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double alpha = sc[0] / sc[1];
    x[i] += alpha * p[i];
    r[i] -= alpha * q[i];
}

// p = r + beta p with beta = sc[2] / sc[0]
__global__ __launch_bounds__(256) void k_cg_p(double* p, const double* r, int n, const double* sc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = r[i] + (sc[2] / sc[0]) * p[i];
}

__global__ void k_copy_scalar(const double* src, double* dst) { *dst = *src; }

// MINRES vectors (the Lanczos basis v and the directions w = V R^-1)
// y = a y
__global__ __launch_bounds__(256) void k_mr_scale(double* y, int n, double a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] *= a;
}

// v_next -= h v_prev (h on the host), then v_next -= sc[0] v_curr (the
// projection, on the device): the two orthogonalisations of one Lanczos step,
// the second applied by k_mr_orth after the dot product
__global__ __launch_bounds__(256) void k_mr_axpy(double* y, const double* x, int n, double a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

__global__ __launch_bounds__(256) void k_mr_orth(double* vn, const double* vc, int n, const double* proj) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) vn[i] -= proj[0] * vc[i];
}

// v_next /= h4; w_next = (v_curr - h2 w_curr - h1 w_prev) / h3; x += rhs w_next
__global__ __launch_bounds__(256) void k_mr_step(double* vn, const double* vc, double* wn, const double* wc,
                                                 const double* wp, double* x, int n, double inv_h4, double h2,
                                                 double h1, double inv_h3, double rhs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vn[i] *= inv_h4;
    const double w = (vc[i] - h2 * wc[i] - h1 * wp[i]) * inv_h3;
    wn[i] = w;
    x[i] += rhs * w;
}

// LSQR: v *= inv_alfa (the new right vector); x += t1 w; w = v + t2 w
__global__ __launch_bounds__(256) void k_ls_step(double* v, double* w, double* x, int n, double inv_alfa, double t1,
                                                 double t2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double vi = v[i] * inv_alfa, wi = w[i];
    v[i] = vi;
    x[i] += t1 * wi;
    w[i] = vi + t2 * wi;
}

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

hipError_t nq_launch_sv(const void* Y, int y_fp64, int64_t Ns, int R, int nA, int offB, int np, const double* mu,
                        const double* v, double eps, int kind, const double* diag, double* w, double* part,
                        int chunks, double* q, hipStream_t st) {
    const dim3 g1(nblk(Ns, 4)), g2(nblk(R, 256), chunks);
    if (y_fp64) {
        hipLaunchKernelGGL((k_yv<double>), g1, dim3(256), 0, st, (const double*)Y, Ns, R, nA, offB, mu, v, w);
        hipLaunchKernelGGL((k_ytw<double>), g2, dim3(256), 0, st, (const double*)Y, Ns, R, nA, mu, w, chunks, part);
    } else {
        hipLaunchKernelGGL((k_yv<float>), g1, dim3(256), 0, st, (const float*)Y, Ns, R, nA, offB, mu, v, w);
        hipLaunchKernelGGL((k_ytw<float>), g2, dim3(256), 0, st, (const float*)Y, Ns, R, nA, mu, w, chunks, part);
    }
    hipLaunchKernelGGL(k_ytw_finish, dim3(nblk(np, 256)), dim3(256), 0, st, part, chunks, R, nA, offB, np,
                       1.0 / (double)Ns, eps, kind, diag, v, q);
    return hipGetLastError();
}

hipError_t nq_launch_sdiag(const void* Y, int y_fp64, int64_t Ns, int R, int nA, int offB, int np, const double* mu,
                           double* part, int chunks, double* diag, hipStream_t st) {
    const dim3 g(nblk(R, 256), chunks);
    if (y_fp64)
        hipLaunchKernelGGL((k_colsq<double>), g, dim3(256), 0, st, (const double*)Y, Ns, R, mu, chunks, part);
    else
        hipLaunchKernelGGL((k_colsq<float>), g, dim3(256), 0, st, (const float*)Y, Ns, R, mu, chunks, part);
    // diag[k] = scale * column sums of k (the finish kernel without a shift term)
    hipLaunchKernelGGL(k_ytw_finish, dim3(nblk(np, 256)), dim3(256), 0, st, part, chunks, R, nA, offB, np,
                       1.0 / (double)Ns, 0.0, 0, (const double*)nullptr, (const double*)nullptr, diag);
    return hipGetLastError();
}

hipError_t nq_launch_dot(const double* x, const double* y, int n, double* part, int nparts, double* out,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_dot_part, dim3(nparts), dim3(256), 0, st, x, y, n, part);
    hipLaunchKernelGGL(k_dot_finish, dim3(1), dim3(64), 0, st, part, nparts, out);
    return hipGetLastError();
}

hipError_t nq_launch_cg_xr(double* x, double* r, const double* p, const double* q, int n, const double* sc,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_cg_xr, dim3(nblk(n, 256)), dim3(256), 0, st, x, r, p, q, n, sc);
    return hipGetLastError();
}

hipError_t nq_launch_cg_p(double* p, const double* r, int n, const double* sc, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_p, dim3(nblk(n, 256)), dim3(256), 0, st, p, r, n, sc);
    return hipGetLastError();
}

hipError_t nq_launch_copy_scalar(const double* src, double* dst, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_scalar, dim3(1), dim3(1), 0, st, src, dst);
    return hipGetLastError();
}

hipError_t nq_launch_mr_scale(double* y, int n, double a, hipStream_t st) {
    hipLaunchKernelGGL(k_mr_scale, dim3(nblk(n, 256)), dim3(256), 0, st, y, n, a);
    return hipGetLastError();
}

hipError_t nq_launch_mr_axpy(double* y, const double* x, int n, double a, hipStream_t st) {
    hipLaunchKernelGGL(k_mr_axpy, dim3(nblk(n, 256)), dim3(256), 0, st, y, x, n, a);
    return hipGetLastError();
}

hipError_t nq_launch_mr_orth(double* vn, const double* vc, int n, const double* proj, hipStream_t st) {
    hipLaunchKernelGGL(k_mr_orth, dim3(nblk(n, 256)), dim3(256), 0, st, vn, vc, n, proj);
    return hipGetLastError();
}

hipError_t nq_launch_mr_step(double* vn, const double* vc, double* wn, const double* wc, const double* wp, double* x,
                             int n, double inv_h4, double h2, double h1, double inv_h3, double rhs, hipStream_t st) {
    hipLaunchKernelGGL(k_mr_step, dim3(nblk(n, 256)), dim3(256), 0, st, vn, vc, wn, wc, wp, x, n, inv_h4, h2, h1,
                       inv_h3, rhs);
    return hipGetLastError();
}

hipError_t nq_launch_ls_step(double* v, double* w, double* x, int n, double inv_alfa, double t1, double t2,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_ls_step, dim3(nblk(n, 256)), dim3(256), 0, st, v, w, x, n, inv_alfa, t1, t2);
    return hipGetLastError();
}
