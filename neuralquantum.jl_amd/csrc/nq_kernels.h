// Argument blocks and launchers of the libnqhip kernels (host <-> device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nq_device.h"
#include "nq_table.h"

// ---- fused estimator (k_estimate.hip)
struct est_args {
    ndm_dims d;
    const void* p;            // parameters, reference order, precision T
    dev_table tab;
    const uint64_t* keys;     // Ns packed configurations
    int64_t Ns;
    double invNs;
    int do_conn;              // walk connections (L_loc, force)
    int lds_param_bytes;      // LDS bytes reserved for the staged parameters
    int max_conn;             // upper bound of connections per configuration (LDS W cache)
    int tab_lds;              // stage the connection table in LDS (tab.bytes small enough)
    void* lp;                 // 2 Ns T: log rho (re, im), nullable
    void* Y;                  // Ns x R T: compact O rows, nullable
    void* Lloc;               // 2 Ns T, nullable
    double* Fpart;            // gridDim x np: per-workgroup force partials (P order), nullable
    void* G;                  // Ns x np complex FP64: per-sample dL_loc (reference order), nullable, pre-zeroed
    // two launches instead of one: rec_mode 1 walks the connections and
    // writes L_loc and, per sample, the changed connections (W, flip masks),
    // their count and the unchanged connections' mel sum; rec_mode 2 reads
    // those (and L_loc) and adds the forces only.  0: everything in one launch
    int rec_mode;
    void* crw;                // Ns x max_conn complex T: W of the changed connections
    uint2* crm;               // Ns x max_conn: their (row, column) flip masks
    int* cnum;                // Ns: changed connections per sample
    void* cwd;                // 2 Ns T: sum of the unchanged connections' mels
    int waves;                // waves per workgroup, 4 or 8 (0: 8 while the LDS allows two workgroups per CU)
};
size_t est_lds_bytes(const est_args& a, size_t elt, int waves);
hipError_t nq_launch_estimate(const est_args& a, bool fp64, int grid, hipStream_t st);

// ---- Metropolis sampler (k_sampler.hip)
struct samp_args {
    ndm_dims d;
    const void* p;
    int64_t chains;
    int passes, burn, stored;  // calls = burn + stored
    uint32_t t0;               // index of this launch's first proposal within the iteration
    int slice0;                // slice index of this launch's first stored sample
    int diag;                  // chain on the diagonal (sigma, sigma): p ~ rho(sigma, sigma), both sides flipped
    uint32_t tag;              // random stream: counter word 3 = {0 start, 1 proposal} + 16 tag
    int lpc;                   // lanes per chain (power of two)
    int restart;               // draw a fresh uniform configuration per chain
    int compat_stale_last_slice;
    uint64_t seed;
    uint32_t iter;
    const uint32_t* iter_dev;  // when set, the iteration index is read here (graph replays)
    uint64_t* chain_state;     // chains, in/out (nullable when restart)
    uint64_t* samples;         // (stored [+1]) x chains, chain fastest
    unsigned long long* accepted;
    uint32_t* chain_acc;       // chains: accepted proposals per chain, added to (nullable)
    int waves;                 // waves per workgroup, 1, 2 or 4 (0: 4)
};
int nq_sampler_lanes_per_chain(const ndm_dims& d);
hipError_t nq_launch_metropolis(const samp_args& a, bool fp64, hipStream_t st);

// ---- S build (k_sbuild.hip)
enum { NQ_GRAM_F64 = 0, NQ_GRAM_F32 = 1 };
// S[o + r, o + c] += scale * sum_s (Y[s, c0 + r] - mu[c0 + r]) (Y[s, c0 + c] - mu[c0 + c])
// for r, c < n; S is column-major with leading dimension lds.  Y is Ns x ldy of
// type double (y_fp64) or float.
struct gram_args {
    const void* Y;
    int y_fp64;
    int64_t Ns;
    int ldy, c0, n;
    const double* mu;          // column means (length >= c0 + n), nullable
    double* S;
    int lds, o;
    double scale;
};
hipError_t nq_launch_gram(const gram_args& a, int mode, hipStream_t st);

// BF16 S build: Yt[c][s] = bf16(Y[s, c0 + c] - mu[c0 + c]) for c < n, s < nsp
// (zero beyond Ns; nsp a multiple of 64), then the split-K Gram below
hipError_t nq_launch_center_t(const void* Y, int y_fp64, int64_t Ns, int ldy, int c0, int n, const double* mu,
                              __bf16* Yt, int64_t nsp, hipStream_t st);
// split-K BF16 Gram of up to two column blocks of Yt (block A: rows [0, nA),
// block B: rows [nA, nA + nB)) written into all of the (np + 1)^2 matrix S:
// scale * A^T A at [0, nA) plus scale * B^T B at [offB, offB + nB), zeros
// elsewhere (row and column np included); `part` holds
// nq_gram_split_part_floats(nA, nB, nsp, max_splits) floats of partial sums;
// max_splits: split-K chunks (0: chosen from the shape)
size_t nq_gram_split_part_floats(int nA, int nB, int64_t nsp, int max_splits);
hipError_t nq_launch_gram2_bf16(const __bf16* Yt, int64_t nsp, int nA, int nB, int offB, double scale, float* part,
                                double* S, int np, int lds, int max_splits, hipStream_t st);

// matrix-free SR (k_cg.hip): q = S v + shift(v) with S = Re(Oc^H Oc) / Ns from
// the compact rows Y (block A = columns [0, nA) -> parameters [0, nA), block
// B = columns [nA, R) -> parameters [offB, offB + R - nA)); w: 2 Ns doubles,
// part: chunks x R doubles; kind 0: eps v, 1: eps diag(S) v (diag from
// nq_launch_sdiag)
hipError_t nq_launch_sv(const void* Y, int y_fp64, int64_t Ns, int R, int nA, int offB, int np, const double* mu,
                        const double* v, double eps, int kind, const double* diag, double* w, double* part,
                        int chunks, double* q, hipStream_t st);
hipError_t nq_launch_sdiag(const void* Y, int y_fp64, int64_t Ns, int R, int nA, int offB, int np, const double* mu,
                           double* part, int chunks, double* diag, hipStream_t st);
// *out = x . y (part: nparts doubles)
hipError_t nq_launch_dot(const double* x, const double* y, int n, double* part, int nparts, double* out,
                         hipStream_t st);
// CG steps with the scalars in device memory sc = [r.r, p.q, new r.r]
hipError_t nq_launch_cg_xr(double* x, double* r, const double* p, const double* q, int n, const double* sc,
                           hipStream_t st);
hipError_t nq_launch_cg_p(double* p, const double* r, int n, const double* sc, hipStream_t st);
// MINRES vector updates (k_cg.hip): y *= a; y += a x; vn -= proj[0] vc;
// vn /= h4, wn = (vc - h2 wc - h1 wp) / h3, x += rhs wn
hipError_t nq_launch_mr_scale(double* y, int n, double a, hipStream_t st);
hipError_t nq_launch_mr_axpy(double* y, const double* x, int n, double a, hipStream_t st);
hipError_t nq_launch_mr_orth(double* vn, const double* vc, int n, const double* proj, hipStream_t st);
hipError_t nq_launch_mr_step(double* vn, const double* vc, double* wn, const double* wc, const double* wp, double* x,
                             int n, double inv_h4, double h2, double h1, double inv_h3, double rhs, hipStream_t st);
// LSQR step (k_cg.hip): v *= inv_alfa; x += t1 w; w = v + t2 w
hipError_t nq_launch_ls_step(double* v, double* w, double* x, int n, double inv_alfa, double t1, double t2,
                             hipStream_t st);
hipError_t nq_launch_copy_scalar(const double* src, double* dst, hipStream_t st);

// column partial sums: part[chunk][c] = sum_{s in chunk} Y[s, c]
hipError_t nq_launch_colsum(const void* Y, int y_fp64, int64_t Ns, int R, double* part, int chunks,
                            hipStream_t st);
// mu[c] = scale * sum_chunk part[chunk][c]
hipError_t nq_launch_colsum_finish(const double* part, int chunks, int R, double scale, double* mu,
                                   hipStream_t st);

// ---- small kernels and the solve (k_solve.hip)
hipError_t nq_launch_fill(double* x, int64_t n, double v, hipStream_t st);
// 32-bit fill (counters and flags; used instead of memset so the sequence can be graph-captured)
hipError_t nq_launch_fill_u32(uint32_t* x, int64_t n, uint32_t v, hipStream_t st);
// *x += 1 (the iteration counter of a captured iteration)
hipError_t nq_launch_bump(uint32_t* x, hipStream_t st);
// statistics (5 doubles), failed-pivot word and finiteness flag into out[0..6] (pinned host memory)
hipError_t nq_launch_status_out(const double* stats, const int* info, const int* flag, double* out, hipStream_t st);
// stats of |L_loc|^2 over B chains x L slices (stats.jl:52-77): out[5]
hipError_t nq_launch_stats(const void* Lloc, int fp64, int64_t B, int64_t Lc, double* out, hipStream_t st);
// F[k] = sum_b Fpart[b][k] - C Re(Obar_k), P order; C = stats[0]; part2: NQ_FORCE_GROUPS x np scratch
#define NQ_FORCE_GROUPS 32
hipError_t nq_launch_force_finish(const double* Fpart, int nblk, double* part2, const double* mu,
                                  const double* stats, ndm_dims d, double* F, hipStream_t st);
// part2[g][k] = sum of the Fpart rows of group g (NQ_FORCE_GROUPS groups), k < n
hipError_t nq_launch_force_part(const double* Fpart, int nb, int n, double* part2, hipStream_t st);
// S[k,k] += eps (kind 0) or eps * S[k,k] (kind 1)
hipError_t nq_launch_shift(double* S, int n, int lds, double eps, int kind, hipStream_t st);
// in-place Cholesky of S (column-major, n x n, leading dimension lds): the
// panels below the diagonal 32 x 32 blocks transposed into the upper triangle
// (A[j, i] = L[i, j] for i beyond j's block), the diagonal blocks only as
// winv[b] = L_bb^-T (32 x 32 column-major, nq_cholesky_winv_doubles(n) doubles);
// *info = first failing column + 1 or 0.  With lds > n, row n holds a
// right-hand side b^T on entry and L^-1 b ends in column n (read it with
// nq_copy_aug_col).
size_t nq_cholesky_winv_doubles(int n);
// factor columns [k0, n) of the nr x nr lower matrix A (rows past n are carried
// along: a partial factorization leaves their Schur complement in columns
// [n, nr) when ncu = nr, and leaves those columns untouched when ncu = n);
// winv block (k - k0) / 32 holds panel k; ninst instances at strides sA (A),
// sW (winv); first_done: winv block 0 is already there (the launch that
// produced the range's first columns factored its diagonal block)
hipError_t nq_cholesky_range(double* A, int k0, int n, int nr, int ncu, int lda, int* info, double* winv, int ninst,
                              int64_t sA, int64_t sW, hipStream_t st, bool first_done);
hipError_t nq_cholesky_backsolve_range(const double* L, int k0, int n, int lds, const double* winv, double* y,
                                       double* x, int ninst, int64_t sA, int64_t sW, int64_t sY, hipStream_t st);
// the NDM's shifted SR solve through the lambda / mu split (k_solve.hip):
// A2 2 (m+1)^2 doubles, winv nq_split_winv_doubles, y2 2 (m+1), x2
// nq_split_x_doubles (the two solutions, then scratch), m = nl + npi; delta in
// P order (np = 2 nl + npi)
size_t nq_split_winv_doubles(int nl, int npi);
size_t nq_split_x_doubles(int nl, int npi);
hipError_t nq_split_factor(const double* S, int lds, const double* F, int nl, int npi, double eps, int kind,
                           double* A2, double* winv, int* info, hipStream_t st);
hipError_t nq_split_substitute(int nl, int npi, const double* A2, const double* winv, double* y2, double* x2,
                               hipStream_t st);
hipError_t nq_cholesky(double* S, int n, int lds, int* info, double* winv, hipStream_t st);
// x = L^-T y with the factor from nq_cholesky (y is overwritten)
hipError_t nq_cholesky_backsolve(const double* L, int n, int lds, const double* winv, double* y, double* x,
                                 hipStream_t st);
// y[j] = A[j, n] (the forward-substituted right-hand side)
hipError_t nq_copy_aug_col(const double* A, int n, int lds, double* y, hipStream_t st);
// y <- (L L^T)^-1 y with the factor from nq_cholesky; tmp: n doubles
hipError_t nq_cholesky_solve(const double* L, int n, int lds, const double* winv, double* y, double* tmp,
                             hipStream_t st);
// A[n, j] = b[j]: the right-hand side as the extra row nq_cholesky carries along
hipError_t nq_set_aug_row(double* A, int n, int lds, const double* b, hipStream_t st);
// r = b - (S + shift) x   (S the unfactored matrix, shift as in nq_launch_shift)
hipError_t nq_residual(const double* S, int n, int lds, double eps, int kind, const double* x,
                       const double* b, double* r, hipStream_t st);
// x += y
hipError_t nq_launch_axpy(double* x, const double* y, int n, hipStream_t st);
// params[perm_ref(k)] -= lr * delta[k] (delta in P order); refresh the FP32 copy
// device status words an update checks before it applies the step (each nullable):
// info = failed pivot column of the factorization, flag = non-finite step,
// stats[0] = the loss of the gradient's samples
struct nq_step_guard {
    const int* info;
    const int* flag;
    const double* stats;
};
hipError_t nq_launch_update(double* p64, float* p32, const double* delta, ndm_dims d, double lr,
                            nq_step_guard g, hipStream_t st);
// the split solve's last launch (one workgroup): delta in P order from x2
// (nq_split_substitute's solutions), *flag = any non-finite entry (0 or 1);
// with p64, the step p -= lr delta guarded as nq_launch_update's (info, the
// flag, stats[0]); with out, the status record [stats(5), info, flag]
hipError_t nq_split_tail(int nl, int npi, const double* x2, double* delta, int* flag, double* p64, float* p32,
                         const ndm_dims& d, double lr, const int* info, const double* stats, double* out,
                         hipStream_t st);
// FP32 copy of the parameters
hipError_t nq_launch_to_f32(const double* p64, float* p32, int64_t n, hipStream_t st);
// NaN / Inf check: *flag |= 1 if any non-finite in x
hipError_t nq_launch_check_finite(const double* x, int64_t n, int* flag, hipStream_t st);

// ---- connection listing (k_network.hip)
hipError_t nq_launch_connections(dev_table tab, int N, const uint64_t* keys, int64_t B, int cap,
                                 int* counts, uint64_t* keys_out, double* mel_out, hipStream_t st);

// ---- learned proposal: MADE over the D = 2N doubled bits (k_made.hip)
// parameters (FP32, row-major): W1[128][D] b1[128] W2[128][128] b2[128] W3[D][128] b3[D]
int nq_made_param_count(int D);
size_t nq_made_lds_bytes();
// raise the dynamic-LDS limit of the MADE kernels (once per process, before any capture)
hipError_t nq_made_prepare();
// number of per-workgroup partials a gradient over B configurations produces
size_t nq_made_grad_parts(int64_t B);
// gpart[w][P]: partial gradients of the mean NLL; lpart[w]: partial sums of log q
hipError_t nq_launch_made_grad(const float* th, int D, const uint64_t* keys, int B, float* gpart, float* lpart,
                               hipStream_t st);
// g = sum_w gpart[w]; *nll = -sum_w lpart[w] / B (nll nullable)
hipError_t nq_launch_made_sum_parts(const float* gpart, int nparts, int D, float* g, const float* lpart, float* nll,
                                    int64_t B, hipStream_t st);
// one Adam step on the summed partials, t counted from 1
hipError_t nq_launch_made_adam(float* th, float* m, float* v, const float* gpart, int nparts, int D, float lr, int t,
                               hipStream_t st);
hipError_t nq_launch_made_logq(const float* th, int D, const uint64_t* keys, int64_t n, float* lq, hipStream_t st);
// n ancestral samples and their log q (lq nullable); iteration from iter_dev when set
hipError_t nq_launch_made_sample(const float* th, int D, int64_t n, uint64_t seed, uint32_t iter,
                                 const uint32_t* iter_dev, uint32_t tag, uint64_t* keys, float* lq, hipStream_t st);
// independence Metropolis-Hastings step of n chains (keys[0,n) states, keys[n,2n) proposals)
hipError_t nq_launch_made_accept(const uint64_t* keys, const void* lp, bool lp_fp64, const float* lq, int64_t n,
                                 uint64_t seed, uint32_t iter, const uint32_t* iter_dev, uint32_t slice,
                                 uint64_t* chain_state, unsigned long long* accepted, uint32_t* chain_acc,
                                 hipStream_t st);
hipError_t nq_launch_copy_keys(const uint64_t* src, int64_t n, uint64_t* dst, hipStream_t st);
// append the last min(n, cap) keys of src to the ring; pos = {write position, fill}
hipError_t nq_launch_ring_push(const uint64_t* src, int64_t n, uint64_t* ring, int64_t cap, int64_t* pos,
                               hipStream_t st);

// ---- RBMSplit network (k_rbm.hip): complex parameters ar[N] ac[N] b[nh] Wr[nh x N]
// Wc[nh x N], interleaved (re, im) in precision T, reference order
struct rbm_args {
    int N, nh, npc;           // npc = 2N + nh + 2 nh N complex parameters (ket: N + nh + nh N)
    int ket;                  // RBM on kets (a[N] b[nh] W[nh x N]), keys of N bits
    const void* p;
    dev_table tab;
    const uint64_t* keys;
    int64_t Ns;
    double invNs;
    int do_conn;
    int lds_param_bytes;      // >= 2 npc sizeof(T), multiple of 16
    int max_conn;
    int tab_lds;              // stage the connection table in LDS
    void* lp;                 // 2 Ns T, nullable
    void* Y;                  // Ns x 2 npc T: [Re O | Im O], nullable
    void* Lloc;               // 2 Ns T, nullable
    double* Fpart;            // gridDim x 2 npc: [Re | Im] force partials, nullable
    double* G;                // Ns x npc complex FP64 dL_loc (parity mode), nullable, pre-zeroed
};
size_t rbm_est_lds_bytes(const rbm_args& a, size_t elt);
hipError_t nq_launch_rbm_estimate(const rbm_args& a, bool fp64, int grid, hipStream_t st);
struct rbm_samp_args {
    int N, nh, npc;
    int ket;
    int diag;                 // diagonal chain (density matrices): p ~ |rho(sigma, sigma)|
    uint32_t tag;
    const void* p;
    int64_t chains;
    int passes, burn, stored;
    uint32_t t0;
    int slice0;
    int lpc;
    int restart;
    int compat_stale_last_slice;
    uint64_t seed;
    uint32_t iter;
    const uint32_t* iter_dev;
    uint64_t* chain_state;
    uint64_t* samples;
    unsigned long long* accepted;
    uint32_t* chain_acc;       // chains: accepted proposals per chain, added to (nullable)
};
hipError_t nq_launch_rbm_metropolis(const rbm_samp_args& a, bool fp64, hipStream_t st);
// in place: [[G_rr, G_ri], [G_ir, G_ii]] -> [[A, -B], [B, A]] (A = G_rr + G_ii, B = G_ri - G_ir)
hipError_t nq_launch_embed_complex(double* S, int npc, int ld, hipStream_t st);
// F[k] = sum_b Fpart[b][k] - C mu[k], k < n
hipError_t nq_launch_force_all(const double* Fpart, int nblk, double* part2, const double* mu, const double* stats,
                               int n, double* F, hipStream_t st);
// p -= lr (delta[0:npc] + i delta[npc:2 npc]); refresh the FP32 copy
hipError_t nq_launch_update_complex(double* p64, float* p32, const double* delta, int npc, double lr,
                                    nq_step_guard g, hipStream_t st);
// energy path: statistics of complex values, out = [Re mean, err, var, tau, R, Im mean]
hipError_t nq_launch_stats_cplx(const void* E, int fp64, int64_t B, int64_t Lc, double* out, hipStream_t st);
// F = [Re | Im] of sum_s E_s conj(Y_s - mu) / Ns (Y rows [Re O | Im O]); part: chunks x 2 npc scratch
hipError_t nq_launch_energy_force(const void* Y, const void* E, int fp64, const double* mu, int64_t Ns, int npc,
                                  double* part, int chunks, double* F, hipStream_t st);
// ---- exact full sums (k_rbm.hip)
hipError_t nq_launch_iota_keys(uint64_t* keys, int64_t n, uint64_t start, hipStream_t st);
// out[0..4] = {max log p, sum w, sum w |L|^2, sum w Re L, sum w Im L}, w = exp(2 Re lp - max)
hipError_t nq_launch_exact_partial(const void* lp, const void* L, int fp64, int64_t n, double* out, hipStream_t st);
// log p = 2 Re log rho (complex, network precision) of n configurations, as doubles
hipError_t nq_launch_logp_of(const void* lp, int fp64, int64_t n, double* out, hipStream_t st);
// combine nch chunk partials of k_exact_partial: out = [sum p |L|^2, Re sum p L, Im sum p L, log Z]
hipError_t nq_launch_exact_combine(const double* part, int64_t nch, double* out, hipStream_t st);

// ---- packed launches: the workgroups of many independent problems (batch
// members) in one grid.  Block b runs the block body of member blk[b].x &
// 0xffff, instance / z index blk[b].x >> 16, at (x, y) = (blk[b].y & 0xffff,
// blk[b].y >> 16) of that member's own grid, with the member's arguments from
// a device array of descriptors: each block does exactly what it does in the
// member's own launch, so a member's results do not depend on the batch.
// One SR iteration (the split Cholesky solve of an NDM, BF16 S build) in the
// order enqueue_iteration runs it; the host lays the blocks out per launch.
struct nq_pk_member {
    samp_args sa;              // the sampler launch (iteration number from sa.iter_dev)
    est_args ew, ef;           // estimator: walk (rec_mode 1), forces (rec_mode 2)
    int est_grid;              // workgroups of the walk launch
    int force_grid;            // workgroups of the forces launch
    int fp64;                  // network precision of Y, L_loc
    const void* Y;             // Ns x R compact O rows
    const void* Lloc;
    int64_t Ns, chains, Lc;
    int R, np;
    double* colpart;           // kColChunks x R column-sum partials
    int col_chunks;
    double* mu;
    double* stats;
    double* Fpart;             // force_grid x np force partials
    double* fpart2;            // NQ_FORCE_GROUPS x np
    double* F;
    int s_mode;                // S build: NQ_GRAM_F64 / NQ_GRAM_F32 (fill + two k_gram), -1: BF16 split-K
    // BF16 S build
    __bf16* Yt;
    int64_t nsp, kchunk;
    int nA, nB, offB, TA, tiles, splits, ldS;
    float* gpart;
    double* S;
    // split Cholesky solve (nq_split_factor, nq_split_substitute)
    double* A2;
    double* winv;
    int* info;
    int* flag;
    int nl, npi;
    double eps;
    int kind;
    double* y2;
    double* x2;
    double* delta;
    // update
    double* p64;
    float* p32;
    double lr;
    uint32_t* iter_dev;
    unsigned long long* acc;
    uint32_t* chain_acc;
    double* status;            // 8 doubles: the statistics, info, flag (read back once for the whole batch)
};

enum {
    NQ_PK_PREP = 0,      // zero the acceptance counters, info and flag
    NQ_PK_METROPOLIS,
    NQ_PK_EST_WALK,
    NQ_PK_EST_FORCE,
    NQ_PK_COLSUM,
    NQ_PK_COLSUM_FIN,
    NQ_PK_STATS,
    NQ_PK_FORCE_PART,
    NQ_PK_FORCE_FIN,
    NQ_PK_CENTER,
    NQ_PK_GRAM_PART,
    NQ_PK_GRAM_FIN,
    NQ_PK_SFILL,         // S = 0 before the FP64 / FP32 Gram blocks
    NQ_PK_GRAM,          // param: Gram block (0: A, 1: B)
    NQ_PK_SPLIT_COPY,
    NQ_PK_CHOL_STEP,     // param: phase << 16 | panel number within the phase
    NQ_PK_SCHUR,
    NQ_PK_COPY_COLS,
    NQ_PK_BACK,          // param: phase << 16 | group number within the phase
    NQ_PK_SPLIT_UPD,
    NQ_PK_UPD_FIN,
    NQ_PK_GATHER,
    NQ_PK_CHECK,
    NQ_PK_UPDATE,
    NQ_PK_BUMP,
    NQ_PK_COLLECT,       // statistics and status words into the member's status slot
    NQ_PK_COUNT
};

// one packed launch
struct nq_pk_launch {
    int kernel;                // NQ_PK_*
    int param;                 // kernel-specific (see the enum)
    int variant;               // instantiation (nq_pk_plan)
    int threads;               // threads per workgroup
    size_t lds;                // dynamic LDS bytes
    const uint2* blk;          // nblk block records (device)
    int nblk;
};

// a member's grid for one launch kind (x, y, z; 0 blocks: the member sits it
// out), its instantiation and block shape: the same choices its own launch
// makes.  Host side, from the host copy of the descriptor.
struct nq_pk_shape {
    int gx, gy, gz;
    int variant, threads;
    size_t lds;
};
hipError_t nq_pk_plan(const nq_pk_member& m, int kernel, int param, nq_pk_shape* out);
// panels / back-substitution groups of each phase of the split solve
int nq_pk_split_panels(int nl, int npi, int phase);
int nq_pk_split_groups(int nl, int npi, int phase);
hipError_t nq_pk_run(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st);
// per file (nq_pk_run dispatches)
hipError_t nq_pk_sampler(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st);
hipError_t nq_pk_estimate(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st);
hipError_t nq_pk_estimate_f32(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st);
hipError_t nq_pk_sbuild(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st);
hipError_t nq_pk_solve(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st);
int nq_est_variant(const est_args& a, bool fp64, int* threads, size_t* lds);
int nq_sampler_variant(const samp_args& a, bool fp64, int* threads, size_t* lds, int* grid);
void nq_gram_shape(int nA, int nB, int64_t nsp, int max_splits, int* TA, int* tiles, int* splits, int64_t* kchunk);
int nq_pk_gram_blocks(const nq_pk_member& m, int which);
