// libnqhip host runtime: the C ABI of include/nqhip.h on top of the kernels.
// One context owns one HIP stream and all device buffers of one SR loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/nqhip.h"
#include "nq_kernels.h"

namespace {

struct dbuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct nqh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // a second stream (default priority) for the forces launch beside the S
    // build, forked from and joined back into `stream` by the two events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    double* hstat = nullptr;   // pinned: the last iteration's statistics, pivot word, finiteness flag, done word
    int inflight = 0;          // iterations enqueued by sr_enqueue and not yet finished
    uint64_t seed = 0;
    std::string err;

    // problem
    bool has_problem = false;
    int N_table = 0;
    dbuf table_mem;
    dev_table tab{};
    int max_conn = 0;

    bool problem_ket = false;          // every term acts on kets (a Hamiltonian), not (row, column) pairs
    // network: 0 = NDM (d), 1 = RBMSplit, 2 = RBM on kets (rbm_*; complex parameters interleaved)
    bool has_net = false;
    int net_kind = 0;
    ndm_dims d{};
    int rbm_N = 0, rbm_nh = 0, rbm_npc = 0;
    dbuf p64, p32;
    int net_prec = NQH_PREC_FP64;
    int s_prec = NQH_PREC_FP64;

    // sampler
    bool has_sampler = false;
    int64_t chains = 0;
    int chain_len = 0, passes = 1, burn = 0, flags = 0;
    uint32_t iter = 0;
    bool chains_initialised = false;
    dbuf chain_state, samples, acc;
    dbuf chain_acc;            // accepted proposals per chain in the last sampling call

    // per-sample buffers (capacity Ns_cap)
    int64_t ns_cap = 0;
    dbuf keys_in, lp, Y, Lloc, G, Yt;
    dbuf crw, crm, cnum, cwd;          // changed connections per sample, between the two estimator launches
    int est_grid = 0;
    // launch shapes (0: chosen from the problem; nqh_set_launch_shape)
    int tune_spb = 0;          // estimator samples per workgroup
    int tune_est_waves = 0;    // estimator waves per workgroup (4 or 8)
    int tune_splits = 0;       // BF16 Gram split-K chunks
    int tune_samp_waves = 0;   // sampler waves per workgroup (1, 2, 4)
    dbuf Fpart;
    // iteration buffers
    dbuf colpart, mu, stats, F, S, S0, delta, resid, info, flag, fpart2, winv, tmp, gpart;
    dbuf A2, winv2, y2, x2;   // the split lambda / mu factorization of the NDM's S

    // SR solver: NQH_SOLVER_CHOLESKY (S built and factored), NQH_SOLVER_CG or
    // NQH_SOLVER_MINRES (matrix-free on the O rows, S never formed)
    int solver = NQH_SOLVER_CHOLESKY;
    double cg_tol = 1e-4;
    int cg_maxiter = 0;
    int cg_iters = 0;
    double cg_resid = 0.0;
    dbuf cg_w, cg_part, cg_p, cg_q, cg_r, cg_sc, cg_diag, cg_dpart;
    dbuf mr_vw;   // MINRES: three Lanczos vectors and three directions, Np doubles each
    int64_t last_ns = 0;
    bool grad_ready = false;
    bool solve_ok = false;    // the last precondition / iteration solved (its step may be applied)
    // configuration generation: bumped by every reallocation of a device
    // buffer, every dropped graph (which every reconfiguration call does) and
    // nqh_set_iteration; a batch's captured graph is valid only for the
    // generations it was captured at
    uint64_t gen = 0;

    // graph capture of the SR iteration
    bool use_graph = false;
    bool graph_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double g_eps = 0.0, g_lr = 0.0;
    int g_kind = -1, g_refine = -1;
    dbuf iter_dev;

    // learned proposal (MADE over the doubled bits)
    bool made_ready = false;           // parameters allocated and initialised
    double made_mix = 0.0;             // fraction of stored slices preceded by a global move
    int made_every = 0, made_steps = 0, made_batch = 1024;
    float made_lr = 1e-3f;
    int made_D = 0, made_P = 0, made_t = 0;
    int64_t ring_cap = 0, ring_fill = 0, ring_pos = 0;   // host mirror of the device ring state
    int64_t made_calls = 0;            // sampler invocations through the public entry points
    int64_t made_proposed = 0;         // global moves proposed by the last sampler run
    dbuf made_th, made_m, made_v, made_gpart, made_lpart, made_g, made_nll, made_keys, made_lp, made_lq, ring,
        ring_posd, made_tmp;

    // observables (diagonal chain + ket operator on the row index)
    dbuf obs_table_mem, obs_keys, obs_state, obs_lp, obs_L, obs_stats;
    dev_table obs_tab{};

    // per-phase HIP-event timing
    bool profiling = false;
    hipEvent_t ev[NQH_NUM_PHASES][2] = {};
    bool pending[NQH_NUM_PHASES] = {};
    double phase_ms[NQH_NUM_PHASES] = {};
    int64_t phase_n[NQH_NUM_PHASES] = {};
};

namespace {

constexpr int kColChunks = 256;
constexpr int64_t kSplitSamples = 4096;   // batches from this size run the forces launch on a side stream
constexpr int kTableLdsMax = 32 * 1024;   // connection tables up to this size are staged in LDS

nqh_status fail(nqh_ctx* c, nqh_status s, const std::string& msg) {
    if (c) c->err = msg;
    return s;
}

nqh_status hip_fail(nqh_ctx* c, hipError_t e, const char* where) {
    return fail(c, NQH_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(c, expr)                                    \
    do {                                                   \
        hipError_t _e = (expr);                            \
        if (_e != hipSuccess) return hip_fail(c, _e, #expr); \
    } while (0)

// Host/device copies and clears are queued on the context's stream and waited for there. That
// stream is non-blocking, so a plain hipMemcpy or hipMemset (null stream) would not be ordered
// against its kernels: a clear could land after a later upload or conversion.
static hipError_t copy_sync(hipStream_t st, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
    return e == hipSuccess ? hipStreamSynchronize(st) : e;
}

static hipError_t zero_sync(hipStream_t st, void* dst, size_t bytes) {
    hipError_t e = hipMemsetAsync(dst, 0, bytes, st);
    return e == hipSuccess ? hipStreamSynchronize(st) : e;
}

enum { PH_SAMPLER = 0, PH_ESTIMATE, PH_POST, PH_SBUILD, PH_SOLVE, PH_UPDATE, PH_FACTOR };

void phase_begin(nqh_ctx* c, int ph) {
    if (c->profiling) (void)hipEventRecord(c->ev[ph][0], c->stream);
}
void phase_end(nqh_ctx* c, int ph) {
    if (c->profiling) {
        (void)hipEventRecord(c->ev[ph][1], c->stream);
        c->pending[ph] = true;
    }
}
// after a stream synchronisation: fold the recorded phases into the totals
void phase_collect(nqh_ctx* c) {
    if (!c->profiling) return;
    for (int ph = 0; ph < NQH_NUM_PHASES; ++ph) {
        if (!c->pending[ph]) continue;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev[ph][0], c->ev[ph][1]) == hipSuccess) {
            c->phase_ms[ph] += ms;
            c->phase_n[ph] += 1;
        }
        c->pending[ph] = false;
    }
}

nqh_status ensure(nqh_ctx* c, dbuf& b, size_t bytes) {
    if (b.bytes >= bytes && b.p) return NQH_OK;
    if (b.p) hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (bytes == 0) return NQH_OK;
    c->gen++;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return fail(c, NQH_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    b.bytes = bytes;
    return NQH_OK;
}

void release(dbuf& b) {
    if (b.p) hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

template <typename T>
T* P(dbuf& b) { return reinterpret_cast<T*>(b.p); }

bool fp64_net(const nqh_ctx* c) { return c->net_prec == NQH_PREC_FP64; }
size_t net_elt(const nqh_ctx* c) { return fp64_net(c) ? sizeof(double) : sizeof(float); }
bool is_rbm(const nqh_ctx* c) { return c->net_kind == 1 || c->net_kind == 2; }
bool is_ket(const nqh_ctx* c) { return c->net_kind == 2; }
int net_N(const nqh_ctx* c) { return is_rbm(c) ? c->rbm_N : c->d.N; }
// columns of a stored O row: NDM compact P order, RBMSplit [Re O | Im O]
int ycols(const nqh_ctx* c) { return is_rbm(c) ? 2 * c->rbm_npc : c->d.R; }
// size of the real SR system = doubles in the flat parameter vector
int nsys(const nqh_ctx* c) { return is_rbm(c) ? 2 * c->rbm_npc : c->d.np; }
// complex entries of a reference-order O row
int nderiv(const nqh_ctx* c) { return is_rbm(c) ? c->rbm_npc : c->d.np; }

// P-order index -> reference index (same map as k_update's p2ref)
int p2ref_host(int k, const ndm_dims& d) {
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

nqh_status need_net(nqh_ctx* c) {
    if (!c->has_net) return fail(c, NQH_E_STATE, "network not set (nqh_set_ndm)");
    return NQH_OK;
}
nqh_status need_problem(nqh_ctx* c) {
    if (!c->has_problem) return fail(c, NQH_E_STATE, "problem not set (nqh_set_problem)");
    if (c->has_net && c->N_table != net_N(c))
        return fail(c, NQH_E_ARG, "problem and network have different numbers of sites");
    if (c->has_net && c->problem_ket != is_ket(c))
        return fail(c, NQH_E_ARG, c->problem_ket ? "a Hamiltonian (ket operator) needs the ket RBM (nqh_set_rbm)"
                                                 : "a Liouvillian needs a density-matrix network (NDM or RBMSplit)");
    return NQH_OK;
}

// estimator workgroups for Ns samples: two samples per wave at 8 waves per
// workgroup, at most 1024 workgroups (every workgroup stages the parameters
// and, for the forces, writes np partials: few enough that this stays small
// next to the samples' work, also when many problems share the device)
int est_grid_for(int64_t Ns, int spb = 0) {
    if (spb <= 0) spb = 16;
    int64_t g = (Ns + spb - 1) / spb;
    if (g > 1024) g = 1024;
    if (g < 1) g = 1;
    return (int)g;
}

// workgroups of the forces launch: from kSplitSamples samples on, where it
// runs beside the S build, the FP32 network's matrix-core launch takes four
// times the walk's samples per workgroup (staging, accumulator clear and
// partial write amortised over more samples, fewer partials to sum, and more
// of the GPU left to the S build); otherwise, and in FP64, the walk's grid.
// A function of the batch alone, so that every path (graph, eager, profiled,
// packed) sums the same partials
int force_grid_of(const nqh_ctx* c, int64_t Ns) {
    return (fp64_net(c) || Ns < kSplitSamples) ? c->est_grid : std::max(1, (c->est_grid + 3) / 4);
}

// column-sum chunks of the NDM's O rows: 64 samples or more per chunk, at most kColChunks
int col_chunks_for(int64_t Ns) {
    int64_t k = Ns / 64;
    if (k > kColChunks) k = kColChunks;
    if (k < 1) k = 1;
    return (int)k;
}

// the Gram blocks of the S build: block A = columns [0, nA) of the compact
// O rows, block B = the next nB columns, embedded in S at offB
void gram_blocks(const nqh_ctx* c, int* nA, int* nB, int* offB) {
    if (is_rbm(c)) {
        *nA = ycols(c);
        *nB = 0;
        *offB = 0;
    } else {
        *nA = c->d.nlam + c->d.npi;
        *nB = c->d.npi + c->d.nmu;
        *offB = c->d.nlam;
    }
}

// the NDM's S has a zero lambda-mu block: factor it through the split
bool use_split(const nqh_ctx* c) {
    return !is_rbm(c) && c->d.nlam > 0 && c->d.nlam == c->d.nmu && c->d.npi > 0;
}

// per-sample buffers for Ns samples
nqh_status ensure_samples(nqh_ctx* c, int64_t Ns) {
    nqh_status s;
    const size_t e = net_elt(c);
    if ((s = ensure(c, c->keys_in, sizeof(uint64_t) * Ns))) return s;
    if ((s = ensure(c, c->lp, 2 * e * Ns))) return s;
    if ((s = ensure(c, c->Y, e * (size_t)ycols(c) * Ns))) return s;
    if ((s = ensure(c, c->Lloc, 2 * e * Ns))) return s;
    if (!is_rbm(c)) {
        // the changed connections between the two estimator launches
        const size_t nrec = (size_t)Ns * (size_t)std::max(c->max_conn, 1);
        if ((s = ensure(c, c->crw, nrec * 2 * e)) || (s = ensure(c, c->crm, nrec * sizeof(uint2))) ||
            (s = ensure(c, c->cnum, (size_t)Ns * sizeof(int))) || (s = ensure(c, c->cwd, (size_t)Ns * 2 * e)))
            return s;
    }
    if (!is_rbm(c) && Ns >= kSplitSamples) {
        // the side stream (created on first use, so that contexts that never
        // fork do not take a hardware queue from the others)
        if (!c->side) {
            HIPCHK(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
            HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
            HIPCHK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
        }
    }
    c->est_grid = est_grid_for(Ns, c->tune_spb);
    if ((s = ensure(c, c->Fpart, sizeof(double) * (size_t)c->est_grid * nsys(c)))) return s;
    if ((s = ensure(c, c->colpart, sizeof(double) * (size_t)kColChunks * ycols(c)))) return s;
    if ((s = ensure(c, c->mu, sizeof(double) * ycols(c)))) return s;
    if ((s = ensure(c, c->stats, sizeof(double) * 8))) return s;
    if ((s = ensure(c, c->F, sizeof(double) * nsys(c)))) return s;
    // S has one extra row and column: row np carries the right-hand side
    // through the factorization (forward substitution for free); the
    // matrix-free solver never forms it
    const size_t ldS = (size_t)nsys(c) + 1;
    if (c->solver == NQH_SOLVER_CHOLESKY) {
        if ((s = ensure(c, c->S, sizeof(double) * ldS * ldS))) return s;
        if ((s = ensure(c, c->S0, sizeof(double) * ldS * ldS))) return s;
        if (use_split(c)) {
            const size_t lda = (size_t)c->d.nlam + c->d.npi + 1;
            if ((s = ensure(c, c->A2, sizeof(double) * 2 * lda * lda))) return s;
            if ((s = ensure(c, c->winv2, sizeof(double) * nq_split_winv_doubles(c->d.nlam, c->d.npi)))) return s;
            if ((s = ensure(c, c->y2, sizeof(double) * 2 * lda))) return s;
            if ((s = ensure(c, c->x2, sizeof(double) * nq_split_x_doubles(c->d.nlam, c->d.npi)))) return s;
        }
    }
    if ((s = ensure(c, c->fpart2, sizeof(double) * (size_t)NQ_FORCE_GROUPS * nsys(c)))) return s;
    if ((s = ensure(c, c->delta, sizeof(double) * nsys(c)))) return s;
    if ((s = ensure(c, c->resid, sizeof(double) * nsys(c)))) return s;
    if ((s = ensure(c, c->tmp, sizeof(double) * nsys(c)))) return s;
    if ((s = ensure(c, c->winv, sizeof(double) * nq_cholesky_winv_doubles(nsys(c))))) return s;
    if ((s = ensure(c, c->info, sizeof(int) * 4))) return s;
    if (c->s_prec == NQH_PREC_BF16 && c->solver == NQH_SOLVER_CHOLESKY) {
        const int64_t nsp = (Ns + 63) / 64 * 64;
        if ((s = ensure(c, c->Yt, sizeof(uint16_t) * (size_t)ycols(c) * nsp))) return s;
        int nA, nB, offB;
        gram_blocks(c, &nA, &nB, &offB);
        if ((s = ensure(c, c->gpart, sizeof(float) * nq_gram_split_part_floats(nA, nB, nsp, c->tune_splits))))
            return s;
    }
    if ((s = ensure(c, c->flag, sizeof(int) * 4))) return s;
    return NQH_OK;
}

est_args make_est(nqh_ctx* c, const uint64_t* keys_dev, int64_t Ns) {
    est_args a{};
    a.d = c->d;
    a.p = fp64_net(c) ? c->p64.p : c->p32.p;
    a.tab = c->tab;
    a.keys = keys_dev;
    a.Ns = Ns;
    a.invNs = 1.0 / (double)Ns;
    a.do_conn = 0;
    a.lds_param_bytes = (int)(((size_t)c->d.np * net_elt(c) + 15) / 16 * 16);
    a.max_conn = std::max(c->max_conn, 1);
    a.tab_lds = (c->has_problem && c->tab.bytes <= kTableLdsMax) ? 1 : 0;
    a.waves = c->tune_est_waves;
    return a;
}

rbm_args make_rbm(nqh_ctx* c, const uint64_t* keys_dev, int64_t Ns) {
    rbm_args a{};
    a.N = c->rbm_N;
    a.nh = c->rbm_nh;
    a.npc = c->rbm_npc;
    a.ket = is_ket(c) ? 1 : 0;
    a.p = fp64_net(c) ? c->p64.p : c->p32.p;
    a.tab = c->tab;
    a.keys = keys_dev;
    a.Ns = Ns;
    a.invNs = 1.0 / (double)Ns;
    a.lds_param_bytes = (int)((2 * (size_t)c->rbm_npc * net_elt(c) + 15) / 16 * 16);
    a.max_conn = std::max(c->max_conn, 1);
    a.tab_lds = (c->has_problem && c->tab.bytes <= kTableLdsMax) ? 1 : 0;
    return a;
}

// log rho (and optionally O rows) of Ns device keys
hipError_t launch_logpsi(nqh_ctx* c, const uint64_t* keys_dev, int64_t Ns, void* lp, void* Y) {
    if (is_rbm(c)) {
        rbm_args a = make_rbm(c, keys_dev, Ns);
        a.lp = lp;
        a.Y = Y;
        return nq_launch_rbm_estimate(a, fp64_net(c), est_grid_for(Ns), c->stream);
    }
    est_args a = make_est(c, keys_dev, Ns);
    a.lp = lp;
    a.Y = Y;
    return nq_launch_estimate(a, fp64_net(c), est_grid_for(Ns), c->stream);
}

// the update reads the solve's status words on the device and applies no
// step after a failed factorization, a non-finite step or a non-finite loss
hipError_t launch_update(nqh_ctx* c, double lr) {
    const nq_step_guard g{P<int>(c->info), P<int>(c->flag), P<double>(c->stats)};
    if (is_rbm(c))
        return nq_launch_update_complex(P<double>(c->p64), P<float>(c->p32), P<double>(c->delta), c->rbm_npc, lr,
                                        g, c->stream);
    return nq_launch_update(P<double>(c->p64), P<float>(c->p32), P<double>(c->delta), c->d, lr, g, c->stream);
}

hipError_t launch_metropolis(nqh_ctx* c, const samp_args& a) {
    if (!is_rbm(c)) return nq_launch_metropolis(a, fp64_net(c), c->stream);
    rbm_samp_args r{};
    r.N = c->rbm_N;
    r.nh = c->rbm_nh;
    r.npc = c->rbm_npc;
    r.ket = is_ket(c) ? 1 : 0;
    r.diag = a.diag;
    r.tag = a.tag;
    r.p = a.p;
    r.chains = a.chains;
    r.passes = a.passes;
    r.burn = a.burn;
    r.stored = a.stored;
    r.t0 = a.t0;
    r.slice0 = a.slice0;
    int l = 1;
    while (l < c->rbm_nh && l < 64) l <<= 1;
    r.lpc = l;
    r.restart = a.restart;
    r.compat_stale_last_slice = a.compat_stale_last_slice;
    r.seed = a.seed;
    r.iter = a.iter;
    r.iter_dev = a.iter_dev;
    r.chain_state = a.chain_state;
    r.samples = a.samples;
    r.accepted = a.accepted;
    r.chain_acc = a.chain_acc;
    return nq_launch_rbm_metropolis(r, fp64_net(c), c->stream);
}

// RBMSplit: estimator, means, statistics, complex force, and the real
// embedding [[A, -B], [B, A]] of the Hermitian S = A + i B
nqh_status run_grad_pipeline_rbm(nqh_ctx* c, const uint64_t* keys_dev, int64_t chains, int64_t Lc) {
    const int64_t Ns = chains * Lc;
    const bool f64 = fp64_net(c);
    const int npc = c->rbm_npc, n = 2 * npc, R = n;
    const bool ket = is_ket(c);
    rbm_args a = make_rbm(c, keys_dev, Ns);
    a.do_conn = 1;
    a.lp = c->lp.p;
    a.Y = c->Y.p;
    a.Lloc = c->Lloc.p;                       // L_loc, or the local energy on kets
    a.Fpart = ket ? nullptr : P<double>(c->Fpart);
    phase_begin(c, PH_ESTIMATE);
    HIPCHK(c, nq_launch_rbm_estimate(a, f64, c->est_grid, c->stream));
    phase_end(c, PH_ESTIMATE);
    phase_begin(c, PH_POST);
    HIPCHK(c, nq_launch_colsum(c->Y.p, f64, Ns, R, P<double>(c->colpart), kColChunks, c->stream));
    HIPCHK(c, nq_launch_colsum_finish(P<double>(c->colpart), kColChunks, R, 1.0 / (double)Ns, P<double>(c->mu),
                                      c->stream));
    if (ket) {
        // energy: statistics of E_loc, grad C = sum E conj(Oc) / Ns (BatchedValSampler.jl:100-116)
        HIPCHK(c, nq_launch_stats_cplx(c->Lloc.p, f64, chains, Lc, P<double>(c->stats), c->stream));
        HIPCHK(c, nq_launch_energy_force(c->Y.p, c->Lloc.p, f64, P<double>(c->mu), Ns, npc, P<double>(c->colpart),
                                         kColChunks, P<double>(c->F), c->stream));
    } else {
        HIPCHK(c, nq_launch_stats(c->Lloc.p, f64, chains, Lc, P<double>(c->stats), c->stream));
        HIPCHK(c, nq_launch_force_all(P<double>(c->Fpart), c->est_grid, P<double>(c->fpart2), P<double>(c->mu),
                                      P<double>(c->stats), n, P<double>(c->F), c->stream));
    }
    phase_end(c, PH_POST);
    phase_begin(c, PH_SBUILD);
    const int ldS = n + 1;
    if (c->s_prec == NQH_PREC_BF16) {
        const int64_t nsp = (Ns + 63) / 64 * 64;
        __bf16* Yt = reinterpret_cast<__bf16*>(c->Yt.p);
        HIPCHK(c, nq_launch_center_t(c->Y.p, f64, Ns, R, 0, R, P<double>(c->mu), Yt, nsp, c->stream));
        HIPCHK(c, nq_launch_gram2_bf16(Yt, nsp, R, 0, 0, 1.0 / (double)Ns, P<float>(c->gpart), P<double>(c->S), n,
                                       ldS, c->tune_splits, c->stream));
    } else {
        HIPCHK(c, nq_launch_fill(P<double>(c->S), (int64_t)ldS * ldS, 0.0, c->stream));
        gram_args g{};
        g.Y = c->Y.p;
        g.y_fp64 = f64;
        g.Ns = Ns;
        g.ldy = R;
        g.mu = P<double>(c->mu);
        g.S = P<double>(c->S);
        g.lds = ldS;
        g.scale = 1.0 / (double)Ns;
        g.c0 = 0;
        g.n = R;
        g.o = 0;
        HIPCHK(c, nq_launch_gram(g, c->s_prec == NQH_PREC_FP64 ? NQ_GRAM_F64 : NQ_GRAM_F32, c->stream));
    }
    HIPCHK(c, nq_launch_embed_complex(P<double>(c->S), npc, ldS, c->stream));
    phase_end(c, PH_SBUILD);
    c->last_ns = Ns;
    c->grad_ready = true;
    c->solve_ok = false;
    return NQH_OK;
}

static nqh_status build_s(nqh_ctx* c, int64_t Ns, bool f64);

// log rho, O rows, L_loc and the changed connections (first estimator
// launch), column means; then the forces (second launch), stats and F on the
// side stream while the main stream builds S; joined before the solve
nqh_status run_grad_pipeline(nqh_ctx* c, const uint64_t* keys_dev, int64_t chains, int64_t Lc) {
    if (is_rbm(c)) return run_grad_pipeline_rbm(c, keys_dev, chains, Lc);
    const int64_t Ns = chains * Lc;
    const bool f64 = fp64_net(c);
    nqh_status s;

    est_args a = make_est(c, keys_dev, Ns);
    a.do_conn = 1;
    a.Lloc = c->Lloc.p;
    a.crw = c->crw.p;
    a.crm = P<uint2>(c->crm);
    a.cnum = P<int>(c->cnum);
    a.cwd = c->cwd.p;
    est_args a1 = a, a2 = a;
    a1.rec_mode = 1;
    a1.lp = c->lp.p;
    a1.Y = c->Y.p;
    a2.rec_mode = 2;
    a2.Fpart = P<double>(c->Fpart);
    // Two launches: the walk (log rho, O rows, L_loc, changed connections),
    // then the forces.  Large batches run the forces beside the S build (in
    // order when per-phase times are recorded, or when S is not built (CG));
    // below kSplitSamples both run in order on the context's stream (a second
    // stream costs more than the overlap gains, e.g. for many small problems
    // each on their own stream)
    const bool beside = Ns >= kSplitSamples && !c->profiling && c->solver == NQH_SOLVER_CHOLESKY;
    phase_begin(c, PH_ESTIMATE);
    HIPCHK(c, nq_launch_estimate(a1, f64, c->est_grid, c->stream));
    const int fgrid = force_grid_of(c, Ns);
    if (!beside) HIPCHK(c, nq_launch_estimate(a2, f64, fgrid, c->stream));
    phase_end(c, PH_ESTIMATE);
    phase_begin(c, PH_POST);
    HIPCHK(c, nq_launch_colsum(c->Y.p, f64, Ns, c->d.R, P<double>(c->colpart), col_chunks_for(Ns), c->stream));
    HIPCHK(c, nq_launch_colsum_finish(P<double>(c->colpart), col_chunks_for(Ns), c->d.R, 1.0 / (double)Ns,
                                      P<double>(c->mu), c->stream));
    hipStream_t fs = c->stream;
    if (beside) {
        HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
        fs = c->side;
        HIPCHK(c, nq_launch_estimate(a2, f64, fgrid, fs));
    }
    HIPCHK(c, nq_launch_stats(c->Lloc.p, f64, chains, Lc, P<double>(c->stats), fs));
    HIPCHK(c, nq_launch_force_finish(P<double>(c->Fpart), fgrid, P<double>(c->fpart2), P<double>(c->mu),
                                     P<double>(c->stats), c->d, P<double>(c->F), fs));
    if (beside) HIPCHK(c, hipEventRecord(c->ev_join, c->side));
    phase_end(c, PH_POST);
    s = build_s(c, Ns, f64);
    if (beside) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    if (s) return s;
    c->last_ns = Ns;
    c->grad_ready = true;
    c->solve_ok = false;
    return NQH_OK;
}

// S (P order: A-gram at [0, nA), B-gram at [nlam, np)) on the main stream
static nqh_status build_s(nqh_ctx* c, int64_t Ns, bool f64) {
    if (c->solver != NQH_SOLVER_CHOLESKY) {   // matrix-free: the solve works on Y and mu directly
        c->last_ns = Ns;
        c->grad_ready = true;
        c->solve_ok = false;
        return NQH_OK;
    }
    // S in P order: A-gram at [0, nA), B-gram at [nlam, np)
    const ndm_dims& d = c->d;
    const int np = d.np;
    phase_begin(c, PH_SBUILD);
    const int ldS = np + 1;
    if (c->s_prec == NQH_PREC_BF16) {
        // centred BF16 copy with samples contiguous, then both Gram blocks in
        // one split-K launch and one pass that writes all of S
        const int64_t nsp = (Ns + 63) / 64 * 64;
        __bf16* Yt = reinterpret_cast<__bf16*>(c->Yt.p);
        HIPCHK(c, nq_launch_center_t(c->Y.p, f64, Ns, d.R, 0, d.R, P<double>(c->mu), Yt, nsp, c->stream));
        int nA, nB, offB;
        gram_blocks(c, &nA, &nB, &offB);
        HIPCHK(c, nq_launch_gram2_bf16(Yt, nsp, nA, nB, offB, 1.0 / (double)Ns, P<float>(c->gpart), P<double>(c->S),
                                       np, ldS, c->tune_splits, c->stream));
        phase_end(c, PH_SBUILD);
        c->last_ns = Ns;
        c->grad_ready = true;
        c->solve_ok = false;
        return NQH_OK;
    }
    HIPCHK(c, nq_launch_fill(P<double>(c->S), (int64_t)ldS * ldS, 0.0, c->stream));
    const int mode = c->s_prec == NQH_PREC_FP64 ? NQ_GRAM_F64 : NQ_GRAM_F32;
    gram_args g{};
    g.Y = c->Y.p;
    g.y_fp64 = f64;
    g.Ns = Ns;
    g.ldy = d.R;
    g.mu = P<double>(c->mu);
    g.S = P<double>(c->S);
    g.lds = ldS;
    g.scale = 1.0 / (double)Ns;
    g.c0 = 0;
    g.n = d.nlam + d.npi;
    g.o = 0;
    HIPCHK(c, nq_launch_gram(g, mode, c->stream));
    g.c0 = d.nlam + d.npi;
    g.n = d.npi + d.nmu;
    g.o = d.nlam;
    HIPCHK(c, nq_launch_gram(g, mode, c->stream));
    phase_end(c, PH_SBUILD);
    c->last_ns = Ns;
    c->grad_ready = true;
    c->solve_ok = false;
    return NQH_OK;
}

nqh_status upload_keys(nqh_ctx* c, const uint64_t* keys, int64_t B) {
    nqh_status s = ensure(c, c->keys_in, sizeof(uint64_t) * (size_t)B);
    if (s) return s;
    HIPCHK(c, hipMemcpyAsync(c->keys_in.p, keys, sizeof(uint64_t) * B, hipMemcpyHostToDevice, c->stream));
    return NQH_OK;
}

nqh_status check_keys(nqh_ctx* c, const uint64_t* keys, int64_t B) {
    const int nb = is_ket(c) ? net_N(c) : 2 * net_N(c);
    const uint64_t lim = (nb >= 64) ? ~0ull : ((1ull << nb) - 1);
    for (int64_t b = 0; b < B; ++b)
        if (keys[b] & ~lim) return fail(c, NQH_E_ARG, "configuration key has bits beyond 2N");
    return NQH_OK;
}

// compact P-order row (real components) -> reference-order complex vector
template <typename T>
void expand_row(const T* y, const ndm_dims& d, double* out /* 2 np */) {
    const int nA = d.nlam + d.npi;
    for (int k = 0; k < d.np; ++k) {
        const int r = p2ref_host(k, d);
        double re = 0.0, im = 0.0;
        if (k < d.nlam) re = (double)y[k];
        else if (k < nA) { re = (double)y[k]; im = (double)y[k + d.npi]; }
        else im = (double)y[k + d.npi];
        out[2 * r] = re;
        out[2 * r + 1] = im;
    }
}

template <typename T>
void copy_cplx(const T* src, double* dst, int64_t n2) {
    for (int64_t k = 0; k < n2; ++k) dst[k] = (double)src[k];
}

}  // namespace

extern "C" {

static void drop_graph(nqh_ctx* c);
static nqh_status made_chain_bufs(nqh_ctx* c);
static nqh_status made_train_async(nqh_ctx* c, const uint64_t* src, int64_t n, int steps, int batch, float lr);

// retrain the learned proposal on the sample buffer every `made_every`
// sampler invocations (eagerly, before the invocation is enqueued)
static nqh_status made_maybe_retrain(nqh_ctx* c) {
    if (!c->made_ready) return NQH_OK;
    const bool due = c->made_every > 0 && c->made_steps > 0 && c->made_calls % c->made_every == 0 &&
                     c->ring_cap > 0 && c->ring_fill >= c->made_batch;
    c->made_calls += 1;
    if (!due) return NQH_OK;
    return made_train_async(c, reinterpret_cast<const uint64_t*>(c->ring.p), c->ring_fill, c->made_steps,
                            c->made_batch, c->made_lr);
}

int32_t nqh_abi_version(void) { return NQH_ABI_VERSION; }

static nqh_status scratch_setup(nqh_ctx* c);

nqh_status nqh_create(nqh_ctx** out, int32_t device, uint64_t seed) {
    if (!out) return NQH_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return NQH_E_HIP;
    if (device < 0 || device >= n) return NQH_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return NQH_E_HIP;
    nqh_ctx* c = new nqh_ctx();
    c->device = device;
    c->seed = seed;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return NQH_E_HIP;
    }
    // coherent: the kernel's stores reach the host as they are made (the
    // done word is polled while the last kernel may still be running)
    if (hipHostMalloc((void**)&c->hstat, 8 * sizeof(double), hipHostMallocCoherent) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return NQH_E_HIP;
    }
    // the allocation-free entry point's scratch address on this device (a
    // failure here is reported by that entry point, which tries again)
    (void)scratch_setup(c);
    *out = c;
    return NQH_OK;
}

void nqh_destroy(nqh_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    dbuf* all[] = {&c->table_mem, &c->p64, &c->p32, &c->chain_state, &c->samples, &c->acc, &c->chain_acc, &c->keys_in, &c->lp,
                   &c->Y, &c->Lloc, &c->G, &c->Yt, &c->Fpart, &c->colpart, &c->mu, &c->stats, &c->F, &c->S, &c->S0,
                   &c->delta, &c->resid, &c->info, &c->flag, &c->fpart2, &c->winv, &c->tmp, &c->gpart, &c->A2,
                   &c->winv2, &c->y2, &c->x2, &c->cg_w,
                   &c->cg_part, &c->cg_p, &c->cg_q, &c->cg_r, &c->cg_sc, &c->cg_diag, &c->cg_dpart, &c->mr_vw, &c->made_th, &c->made_m, &c->made_v,
                   &c->made_gpart, &c->made_lpart, &c->made_g, &c->made_nll, &c->made_keys, &c->made_lp,
                   &c->made_lq, &c->ring, &c->ring_posd, &c->made_tmp, &c->obs_table_mem, &c->obs_keys,
                   &c->obs_state, &c->obs_lp, &c->obs_L, &c->obs_stats, &c->crw, &c->crm, &c->cnum, &c->cwd};
    drop_graph(c);
    release(c->iter_dev);
    for (dbuf* b : all) release(*b);
    for (int ph = 0; ph < NQH_NUM_PHASES; ++ph)
        for (int k = 0; k < 2; ++k)
            if (c->ev[ph][k]) (void)hipEventDestroy(c->ev[ph][k]);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) {
        (void)hipStreamSynchronize(c->side);
        (void)hipStreamDestroy(c->side);
    }
    (void)hipStreamDestroy(c->stream);
    if (c->hstat) (void)hipHostFree(c->hstat);
    delete c;
}

const char* nqh_last_error(const nqh_ctx* c) { return c ? c->err.c_str() : "null context"; }

// validate a flat table and upload it as one device allocation; with
// `ket_as_row` the ket terms (kind 3) are stored as row-side terms (kind 0)
static nqh_status upload_table(nqh_ctx* c, const nqh_optable* t, bool ket_as_row, dbuf& mem, dev_table& tab,
                               int* maxc_out, int* nket_out) {
    const int N = t->n_sites;
    if (N < 1 || N > 32) return fail(c, NQH_E_ARG, "n_sites must be in 1..32");
    if (t->n_terms < 0 || t->n_ops < 0 || t->n_rows < 0 || t->n_entries < 0) return fail(c, NQH_E_ARG, "negative sizes");
    // validation of the flat table
    if (t->op_site_off[0] != 0 || t->op_row_off[0] != 0 || t->row_ent_off[0] != 0)
        return fail(c, NQH_E_ARG, "offsets must start at 0");
    if (t->op_row_off[t->n_ops] != t->n_rows || t->row_ent_off[t->n_rows] != t->n_entries)
        return fail(c, NQH_E_ARG, "offset tables inconsistent with n_rows / n_entries");
    const int nsites_total = t->op_site_off[t->n_ops];
    for (int o = 0; o < t->n_ops; ++o) {
        const int k = t->op_site_off[o + 1] - t->op_site_off[o];
        if (k < 1 || k > 8) return fail(c, NQH_E_ARG, "local operators must act on 1..8 sites");
        if (t->op_row_off[o + 1] - t->op_row_off[o] != (1 << k)) return fail(c, NQH_E_ARG, "operator must have 2^k rows");
    }
    for (int m = 0; m < nsites_total; ++m)
        if (t->op_sites[m] < 0 || t->op_sites[m] >= N) return fail(c, NQH_E_ARG, "site index out of range");
    for (int q = 0; q < t->n_rows; ++q)
        if (t->row_ent_off[q + 1] < t->row_ent_off[q]) return fail(c, NQH_E_ARG, "row offsets not monotone");
    const uint64_t lim = (N >= 64) ? ~0ull : ((1ull << N) - 1);
    for (int e = 0; e < t->n_entries; ++e)
        if (t->ent_flip[e] & ~lim) return fail(c, NQH_E_ARG, "flip mask beyond n_sites");
    int maxc = 0, n_ket = 0, emax = 0;
    for (int q = 0; q < t->n_terms; ++q) n_ket += t->term_kind[q] == 3;
    if (n_ket != 0 && n_ket != t->n_terms)
        return fail(c, NQH_E_ARG, "ket terms (kind 3) cannot be mixed with row/column terms");
    for (int q = 0; q < t->n_terms; ++q) {
        const int k = t->term_kind[q], ol = t->term_op_l[q], orr = t->term_op_r[q];
        auto ok = [&](int o) { return o >= 0 && o < t->n_ops; };
        auto maxrow = [&](int o) {
            int m = 0;
            for (int r = t->op_row_off[o]; r < t->op_row_off[o + 1]; ++r)
                m = std::max(m, t->row_ent_off[r + 1] - t->row_ent_off[r]);
            return m;
        };
        int e = 0;
        if (k == 0 || k == 3) { if (!ok(ol)) return fail(c, NQH_E_ARG, "bad row operator"); e = maxrow(ol); }
        else if (k == 1) { if (!ok(orr)) return fail(c, NQH_E_ARG, "bad column operator"); e = maxrow(orr); }
        else if (k == 2) {
            if (!ok(ol) || !ok(orr)) return fail(c, NQH_E_ARG, "bad tensor operators");
            e = maxrow(ol) * maxrow(orr);
        } else return fail(c, NQH_E_ARG, "bad term kind");
        maxc += e;
        emax = std::max(emax, e);
    }
    // one device allocation: int32 arrays, then doubles, floats, flips
    const size_t ni = (size_t)3 * t->n_terms + (t->n_ops + 1) + nsites_total + (t->n_ops + 1) + (t->n_rows + 1);
    const size_t bytes_i = (ni * 4 + 15) / 16 * 16;
    const size_t bytes_d = (size_t)2 * t->n_entries * sizeof(double);
    const size_t bytes_f = ((size_t)2 * t->n_entries * sizeof(float) + 15) / 16 * 16;
    const size_t bytes_m = (size_t)t->n_entries * sizeof(uint32_t);
    const size_t bytes_tab = (bytes_i + bytes_d + bytes_f + bytes_m + 16 + 15) / 16 * 16;
    // term descriptors (8 words per term) after the table proper
    const size_t bytes_desc = (size_t)8 * sizeof(uint32_t) * std::max(t->n_terms, 1);
    std::vector<unsigned char> host(bytes_tab + bytes_desc, 0);
    int32_t* hi = reinterpret_cast<int32_t*>(host.data());
    size_t off = 0;
    auto put = [&](const int32_t* src, int n) { std::memcpy(hi + off, src, sizeof(int32_t) * n); size_t o = off; off += n; return o; };
    const size_t o_kind = put(t->term_kind, t->n_terms);
    if (ket_as_row)
        for (int q = 0; q < t->n_terms; ++q)
            if (hi[o_kind + q] == 3) hi[o_kind + q] = 0;
    const size_t o_l = put(t->term_op_l, t->n_terms);
    const size_t o_r = put(t->term_op_r, t->n_terms);
    const size_t o_so = put(t->op_site_off, t->n_ops + 1);
    const size_t o_s = put(t->op_sites, nsites_total);
    const size_t o_ro = put(t->op_row_off, t->n_ops + 1);
    const size_t o_eo = put(t->row_ent_off, t->n_rows + 1);
    double* hd = reinterpret_cast<double*>(host.data() + bytes_i);
    std::memcpy(hd, t->ent_mel, bytes_d);
    float* hf = reinterpret_cast<float*>(host.data() + bytes_i + bytes_d);
    for (int e = 0; e < 2 * t->n_entries; ++e) hf[e] = (float)t->ent_mel[e];
    uint32_t* hm = reinterpret_cast<uint32_t*>(host.data() + bytes_i + bytes_d + bytes_f);
    for (int e = 0; e < t->n_entries; ++e) hm[e] = (uint32_t)t->ent_flip[e];
    uint32_t* hdesc = reinterpret_cast<uint32_t*>(host.data() + bytes_tab);
    for (int q = 0; q < t->n_terms; ++q) {
        const int k = hi[o_kind + q];
        const bool use_l = k != 1, use_r = k == 1 || k == 2;
        auto pack = [&](int o, uint32_t* w) {
            for (int m = t->op_site_off[o]; m < t->op_site_off[o + 1]; ++m) {
                const int j = m - t->op_site_off[o];
                w[j >> 2] |= (uint32_t)t->op_sites[m] << (8 * (j & 3));
            }
            return (uint32_t)(t->op_site_off[o + 1] - t->op_site_off[o]);
        };
        uint32_t* dq = hdesc + 8 * q;
        const uint32_t nsl = use_l ? pack(t->term_op_l[q], dq + 4) : 0u;
        const uint32_t nsr = use_r ? pack(t->term_op_r[q], dq + 6) : 0u;
        dq[0] = (uint32_t)k | (nsl << 8) | (nsr << 16);
        dq[1] = use_l ? (uint32_t)t->op_row_off[t->term_op_l[q]] : 0u;
        dq[2] = use_r ? (uint32_t)t->op_row_off[t->term_op_r[q]] : 0u;
    }
    nqh_status s = ensure(c, mem, host.size());
    if (s) return s;
    HIPCHK(c, copy_sync(c->stream, mem.p, host.data(), host.size(), hipMemcpyHostToDevice));
    unsigned char* base = reinterpret_cast<unsigned char*>(mem.p);
    const int32_t* di = reinterpret_cast<const int32_t*>(base);
    tab.n_sites = N;
    tab.n_terms = t->n_terms;
    tab.term_kind = di + o_kind;
    tab.term_op_l = di + o_l;
    tab.term_op_r = di + o_r;
    tab.op_site_off = di + o_so;
    tab.op_sites = di + o_s;
    tab.op_row_off = di + o_ro;
    tab.row_ent_off = di + o_eo;
    tab.ent_mel = reinterpret_cast<const double*>(base + bytes_i);
    tab.ent_mel_f = reinterpret_cast<const float*>(base + bytes_i + bytes_d);
    tab.ent_flip = reinterpret_cast<const uint32_t*>(base + bytes_i + bytes_d + bytes_f);
    tab.base = base;
    tab.bytes = (int)host.size();
    tab.term_desc = reinterpret_cast<const uint32_t*>(base + bytes_tab);
    tab.emax = emax;
    *maxc_out = maxc;
    *nket_out = n_ket;
    return NQH_OK;
}

nqh_status nqh_set_problem(nqh_ctx* c, const nqh_optable* t) {
    if (!c || !t) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    int maxc = 0, n_ket = 0;
    nqh_status s = upload_table(c, t, false, c->table_mem, c->tab, &maxc, &n_ket);
    if (s) return s;
    c->N_table = t->n_sites;
    c->max_conn = maxc;
    c->problem_ket = n_ket > 0;
    c->has_problem = true;
    drop_graph(c);
    return NQH_OK;
}

nqh_status nqh_set_ndm(nqh_ctx* c, int32_t N, int32_t nh, int32_t na, int32_t act) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    // the estimator's limits: 2N-bit keys in 32 bits of lane state, two angle registers per lane
    if (N < 1 || N > 16) return fail(c, NQH_E_ARG, "N must be in 1..16");
    if (nh < 1 || na < 1 || nh > 64 || na > 64) return fail(c, NQH_E_ARG, "nh, na must be in 1..64");
    if (act != NQH_ACT_SOFTPLUS && act != NQH_ACT_LOGCOSH)
        return fail(c, NQH_E_ARG, "act must be NQH_ACT_SOFTPLUS (af_softplus) or NQH_ACT_LOGCOSH (af_logcosh)");
    c->d = make_ndm_dims(N, nh, na);
    c->d.act = act;
    c->net_kind = 0;
    drop_graph(c);
    nqh_status s;
    if ((s = ensure(c, c->p64, sizeof(double) * c->d.np))) return s;
    if ((s = ensure(c, c->p32, sizeof(float) * c->d.np))) return s;
    HIPCHK(c, zero_sync(c->stream, c->p64.p, sizeof(double) * c->d.np));
    HIPCHK(c, zero_sync(c->stream, c->p32.p, sizeof(float) * c->d.np));
    c->has_net = true;
    c->ns_cap = 0;
    c->grad_ready = false;
    return NQH_OK;
}

nqh_status nqh_set_rbmsplit(nqh_ctx* c, int32_t N, int32_t nh) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    if (N < 1 || N > 16) return fail(c, NQH_E_ARG, "N must be in 1..16");
    if (nh < 1 || nh > 128) return fail(c, NQH_E_ARG, "nh must be in 1..128");
    drop_graph(c);
    c->net_kind = 1;
    c->rbm_N = N;
    c->rbm_nh = nh;
    c->rbm_npc = 2 * N + nh + 2 * nh * N;
    c->d = ndm_dims{};
    c->d.N = N;
    const int n = 2 * c->rbm_npc;
    nqh_status s;
    if ((s = ensure(c, c->p64, sizeof(double) * n))) return s;
    if ((s = ensure(c, c->p32, sizeof(float) * n))) return s;
    HIPCHK(c, zero_sync(c->stream, c->p64.p, sizeof(double) * n));
    HIPCHK(c, zero_sync(c->stream, c->p32.p, sizeof(float) * n));
    c->has_net = true;
    c->ns_cap = 0;
    c->grad_ready = false;
    c->made_ready = false;
    return NQH_OK;
}

nqh_status nqh_set_rbm(nqh_ctx* c, int32_t N, int32_t nh) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    if (N < 1 || N > 16) return fail(c, NQH_E_ARG, "N must be in 1..16");
    if (nh < 1 || nh > 128) return fail(c, NQH_E_ARG, "nh must be in 1..128");
    drop_graph(c);
    c->net_kind = 2;
    c->rbm_N = N;
    c->rbm_nh = nh;
    c->rbm_npc = N + nh + nh * N;
    c->d = ndm_dims{};
    c->d.N = N;
    const int n = 2 * c->rbm_npc;
    nqh_status s;
    if ((s = ensure(c, c->p64, sizeof(double) * n))) return s;
    if ((s = ensure(c, c->p32, sizeof(float) * n))) return s;
    HIPCHK(c, zero_sync(c->stream, c->p64.p, sizeof(double) * n));
    HIPCHK(c, zero_sync(c->stream, c->p32.p, sizeof(float) * n));
    c->has_net = true;
    c->ns_cap = 0;
    c->grad_ready = false;
    c->made_ready = false;
    return NQH_OK;
}

int64_t nqh_num_params(const nqh_ctx* c) { return (c && c->has_net) ? nsys(c) : -1; }

int64_t nqh_num_derivs(const nqh_ctx* c) { return (c && c->has_net) ? nderiv(c) : -1; }


nqh_status nqh_set_params(nqh_ctx* c, const double* flat, int64_t np) {
    if (!c || !flat) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s = need_net(c);
    if (s) return s;
    if (np != nsys(c)) return fail(c, NQH_E_ARG, "parameter count mismatch");
    for (int64_t k = 0; k < np; ++k)
        if (!std::isfinite(flat[k])) return fail(c, NQH_E_NAN, "non-finite parameter");
    HIPCHK(c, hipMemcpyAsync(c->p64.p, flat, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, nq_launch_to_f32(P<double>(c->p64), P<float>(c->p32), np, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return NQH_OK;
}

nqh_status nqh_get_params(const nqh_ctx* cc, double* flat, int64_t np) {
    nqh_ctx* c = const_cast<nqh_ctx*>(cc);
    if (!c || !flat) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s = need_net(c);
    if (s) return s;
    if (np != nsys(c)) return fail(c, NQH_E_ARG, "parameter count mismatch");
    HIPCHK(c, hipMemcpyAsync(flat, c->p64.p, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return NQH_OK;
}

nqh_status nqh_set_precision(nqh_ctx* c, int32_t net_prec, int32_t sbuild_prec) {
    if (!c) return NQH_E_ARG;
    drop_graph(c);
    if (net_prec != NQH_PREC_FP64 && net_prec != NQH_PREC_FP32) return fail(c, NQH_E_ARG, "net precision FP64 or FP32");
    if (sbuild_prec < NQH_PREC_FP64 || sbuild_prec > NQH_PREC_BF16) return fail(c, NQH_E_ARG, "bad S-build precision");
    if (net_prec != c->net_prec) {
        c->ns_cap = 0;   // element size of the per-sample buffers changes
        c->grad_ready = false;
    }
    c->net_prec = net_prec;
    c->s_prec = sbuild_prec;
    return NQH_OK;
}

static nqh_status eval_host_keys(nqh_ctx* c, const uint64_t* keys, int64_t B, bool want_y, bool conn,
                                 bool want_g) {
    nqh_status s;
    if ((s = check_keys(c, keys, B))) return s;
    if ((s = ensure_samples(c, B))) return s;
    if ((s = upload_keys(c, keys, B))) return s;
    double* G = nullptr;
    if (want_g) {
        const size_t gb = (size_t)2 * sizeof(double) * nderiv(c) * B;
        if ((s = ensure(c, c->G, gb))) return s;
        HIPCHK(c, hipMemsetAsync(c->G.p, 0, gb, c->stream));
        G = P<double>(c->G);
    }
    if (is_rbm(c)) {
        rbm_args a = make_rbm(c, P<uint64_t>(c->keys_in), B);
        a.lp = c->lp.p;
        a.Y = want_y ? c->Y.p : nullptr;
        a.do_conn = conn ? 1 : 0;
        a.Lloc = conn ? c->Lloc.p : nullptr;
        a.G = G;
        HIPCHK(c, nq_launch_rbm_estimate(a, fp64_net(c), est_grid_for(B), c->stream));
        return NQH_OK;
    }
    est_args a = make_est(c, P<uint64_t>(c->keys_in), B);
    a.lp = c->lp.p;
    a.Y = want_y ? c->Y.p : nullptr;
    a.do_conn = conn ? 1 : 0;
    a.Lloc = conn ? c->Lloc.p : nullptr;
    a.G = G;
    HIPCHK(c, nq_launch_estimate(a, fp64_net(c), est_grid_for(B), c->stream));
    return NQH_OK;
}

static nqh_status download_cplx(nqh_ctx* c, dbuf& b, double* dst, int64_t n2) {
    if (fp64_net(c)) {
        HIPCHK(c, hipMemcpyAsync(dst, b.p, sizeof(double) * n2, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> tmp(n2);
        HIPCHK(c, hipMemcpyAsync(tmp.data(), b.p, sizeof(float) * n2, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        copy_cplx(tmp.data(), dst, n2);
    }
    return NQH_OK;
}

nqh_status nqh_logpsi(nqh_ctx* c, const uint64_t* keys, int64_t B, double* logpsi) {
    if (!c || !keys || !logpsi || B < 0) return NQH_E_ARG;
    if (B == 0) return NQH_OK;
    (void)hipSetDevice(c->device);
    nqh_status s = need_net(c);
    if (s) return s;
    if ((s = eval_host_keys(c, keys, B, false, false, false))) return s;
    return download_cplx(c, c->lp, logpsi, 2 * B);
}

nqh_status nqh_logpsi_grad(nqh_ctx* c, const uint64_t* keys, int64_t B, double* logpsi, double* O) {
    if (!c || !keys || B < 0) return NQH_E_ARG;
    if (B == 0) return NQH_OK;
    (void)hipSetDevice(c->device);
    nqh_status s = need_net(c);
    if (s) return s;
    if ((s = eval_host_keys(c, keys, B, true, false, false))) return s;
    if (logpsi && (s = download_cplx(c, c->lp, logpsi, 2 * B))) return s;
    if (O && is_rbm(c)) {
        const int npc = c->rbm_npc;
        std::vector<double> y((size_t)2 * npc * B);
        if (fp64_net(c)) {
            HIPCHK(c, copy_sync(c->stream, y.data(), c->Y.p, sizeof(double) * y.size(), hipMemcpyDeviceToHost));
        } else {
            std::vector<float> yf(y.size());
            HIPCHK(c, copy_sync(c->stream, yf.data(), c->Y.p, sizeof(float) * yf.size(), hipMemcpyDeviceToHost));
            for (size_t k = 0; k < y.size(); ++k) y[k] = yf[k];
        }
        for (int64_t b = 0; b < B; ++b)
            for (int k = 0; k < npc; ++k) {
                O[(size_t)2 * (b * npc + k)] = y[(size_t)b * 2 * npc + k];
                O[(size_t)2 * (b * npc + k) + 1] = y[(size_t)b * 2 * npc + npc + k];
            }
    } else if (O) {
        const ndm_dims& d = c->d;
        if (fp64_net(c)) {
            std::vector<double> y((size_t)d.R * B);
            HIPCHK(c, copy_sync(c->stream, y.data(), c->Y.p, sizeof(double) * y.size(), hipMemcpyDeviceToHost));
            for (int64_t b = 0; b < B; ++b) expand_row(y.data() + (size_t)b * d.R, d, O + (size_t)2 * b * d.np);
        } else {
            std::vector<float> y((size_t)d.R * B);
            HIPCHK(c, copy_sync(c->stream, y.data(), c->Y.p, sizeof(float) * y.size(), hipMemcpyDeviceToHost));
            for (int64_t b = 0; b < B; ++b) expand_row(y.data() + (size_t)b * d.R, d, O + (size_t)2 * b * d.np);
        }
    }
    return NQH_OK;
}

nqh_status nqh_connections(nqh_ctx* c, const uint64_t* keys, int64_t B, int32_t cap, int32_t* counts,
                           uint64_t* keys_out, double* mel_out) {
    if (!c || !keys || !counts || B < 0 || cap < 0) return NQH_E_ARG;
    if (B == 0) return NQH_OK;
    (void)hipSetDevice(c->device);
    nqh_status s = need_problem(c);
    if (s) return s;
    const int N = c->N_table;
    const uint64_t lim = (2 * N >= 64) ? ~0ull : ((1ull << (2 * N)) - 1);
    for (int64_t b = 0; b < B; ++b)
        if (keys[b] & ~lim) return fail(c, NQH_E_ARG, "configuration key has bits beyond 2N");
    const int dcap = std::max(cap, 1);
    dbuf kin, cnt, ko, mo;
    auto cleanup = [&]() { release(kin); release(cnt); release(ko); release(mo); };
    if ((s = ensure(c, kin, sizeof(uint64_t) * B)) || (s = ensure(c, cnt, sizeof(int) * B)) ||
        (s = ensure(c, ko, sizeof(uint64_t) * B * dcap)) || (s = ensure(c, mo, 2 * sizeof(double) * B * dcap))) {
        cleanup();
        return s;
    }
    hipError_t e = copy_sync(c->stream, kin.p, keys, sizeof(uint64_t) * B, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = nq_launch_connections(c->tab, N, P<uint64_t>(kin), B, cap, P<int>(cnt), P<uint64_t>(ko), P<double>(mo),
                                  c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = copy_sync(c->stream, counts, cnt.p, sizeof(int) * B, hipMemcpyDeviceToHost);
    if (e == hipSuccess && keys_out && cap > 0) e = copy_sync(c->stream, keys_out, ko.p, sizeof(uint64_t) * B * cap, hipMemcpyDeviceToHost);
    if (e == hipSuccess && mel_out && cap > 0) e = copy_sync(c->stream, mel_out, mo.p, 2 * sizeof(double) * B * cap, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return hip_fail(c, e, "nqh_connections");
    for (int64_t b = 0; b < B; ++b)
        if (counts[b] > cap) return fail(c, NQH_E_CAPACITY, "connection capacity too small");
    return NQH_OK;
}

nqh_status nqh_local_estimator(nqh_ctx* c, const uint64_t* keys, int64_t B, double* Lloc, double* dLloc) {
    if (!c || !keys || !Lloc || B < 0) return NQH_E_ARG;
    if (B == 0) return NQH_OK;
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_net(c)) || (s = need_problem(c))) return s;
    if (dLloc && is_ket(c)) return fail(c, NQH_E_ARG, "the energy estimator has no per-sample gradient");
    if ((s = eval_host_keys(c, keys, B, false, true, dLloc != nullptr))) return s;
    if ((s = download_cplx(c, c->Lloc, Lloc, 2 * B))) return s;
    if (dLloc) {
        HIPCHK(c, copy_sync(c->stream, dLloc, c->G.p, sizeof(double) * 2 * B * nderiv(c), hipMemcpyDeviceToHost));
    }
    return NQH_OK;
}

nqh_status nqh_grad_from_samples(nqh_ctx* c, const uint64_t* keys, int64_t chains, int64_t Lc, double* stats,
                                 double* S, double* F, double* Obar) {
    if (!c || !keys || chains < 2 || Lc < 2) return fail(c, NQH_E_ARG, "need >= 2 chains and chain length >= 2");
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_net(c)) || (s = need_problem(c))) return s;
    if (S && c->solver != NQH_SOLVER_CHOLESKY)
        return fail(c, NQH_E_STATE, "S is not formed with the matrix-free solver (pass S = NULL)");
    const int64_t Ns = chains * Lc;
    if ((s = check_keys(c, keys, Ns))) return s;
    if ((s = ensure_samples(c, Ns))) return s;
    if ((s = upload_keys(c, keys, Ns))) return s;
    if ((s = run_grad_pipeline(c, P<uint64_t>(c->keys_in), chains, Lc))) return s;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    phase_collect(c);
    const ndm_dims& d = c->d;
    if (stats) HIPCHK(c, copy_sync(c->stream, stats, c->stats.p, sizeof(double) * 5, hipMemcpyDeviceToHost));
    if (is_rbm(c)) {
        // F and Obar interleaved complex (reference order); S as the real
        // embedding [[Re S, -Im S], [Im S, Re S]] (2 npc x 2 npc, column-major)
        const int npc = c->rbm_npc, n = 2 * npc;
        if (F) {
            std::vector<double> f(n);
            HIPCHK(c, copy_sync(c->stream, f.data(), c->F.p, sizeof(double) * n, hipMemcpyDeviceToHost));
            for (int k = 0; k < npc; ++k) { F[2 * k] = f[k]; F[2 * k + 1] = f[npc + k]; }
        }
        if (Obar) {
            std::vector<double> mu(n);
            HIPCHK(c, copy_sync(c->stream, mu.data(), c->mu.p, sizeof(double) * n, hipMemcpyDeviceToHost));
            for (int k = 0; k < npc; ++k) { Obar[2 * k] = mu[k]; Obar[2 * k + 1] = mu[npc + k]; }
        }
        if (S) {
            const size_t ldS = (size_t)n + 1;
            std::vector<double> sp(ldS * ldS);
            HIPCHK(c, copy_sync(c->stream, sp.data(), c->S.p, sizeof(double) * sp.size(), hipMemcpyDeviceToHost));
            for (int l = 0; l < n; ++l)
                for (int k = 0; k < n; ++k) S[(size_t)k + (size_t)l * n] = sp[(size_t)k + (size_t)l * ldS];
        }
        return NQH_OK;
    }
    if (F) {

        std::vector<double> f(d.np);
        HIPCHK(c, copy_sync(c->stream, f.data(), c->F.p, sizeof(double) * d.np, hipMemcpyDeviceToHost));
        for (int k = 0; k < d.np; ++k) F[p2ref_host(k, d)] = f[k];
    }
    if (Obar) {
        std::vector<double> mu(d.R);
        HIPCHK(c, copy_sync(c->stream, mu.data(), c->mu.p, sizeof(double) * d.R, hipMemcpyDeviceToHost));
        expand_row(mu.data(), d, Obar);
    }
    if (S) {
        const size_t ldS = (size_t)d.np + 1;
        std::vector<double> sp(ldS * ldS);
        HIPCHK(c, copy_sync(c->stream, sp.data(), c->S.p, sizeof(double) * sp.size(), hipMemcpyDeviceToHost));
        std::vector<int> r(d.np);
        for (int k = 0; k < d.np; ++k) r[k] = p2ref_host(k, d);
        for (int l = 0; l < d.np; ++l)
            for (int k = 0; k < d.np; ++k) S[(size_t)r[k] + (size_t)r[l] * d.np] = sp[(size_t)k + (size_t)l * ldS];
    }
    return NQH_OK;
}

nqh_status nqh_set_sampler(nqh_ctx* c, int64_t chains, int32_t chain_length, int32_t passes, int32_t burn,
                           int32_t flags) {
    if (!c) return NQH_E_ARG;
    if (chains < 2 || chain_length < 2 || passes < 1 || burn < 0)
        return fail(c, NQH_E_ARG, "need chains >= 2, chain_length >= 2, passes >= 1, burn >= 0");
    if ((flags & NQH_SAMPLER_COMPAT_STALE_LAST_SLICE) && (flags & NQH_SAMPLER_PERSISTENT))
        return fail(c, NQH_E_ARG, "compat and persistent sampler flags are exclusive");
    (void)hipSetDevice(c->device);
    c->chains = chains;
    c->chain_len = chain_length;
    c->passes = passes;
    c->burn = burn;
    c->flags = flags;
    nqh_status s;
    if ((s = ensure(c, c->chain_state, sizeof(uint64_t) * chains))) return s;
    if ((s = ensure(c, c->samples, sizeof(uint64_t) * chains * chain_length))) return s;
    if ((s = ensure(c, c->acc, sizeof(unsigned long long) * 2))) return s;
    if ((s = ensure(c, c->chain_acc, sizeof(uint32_t) * chains))) return s;
    c->chains_initialised = false;
    c->has_sampler = true;
    drop_graph(c);
    return made_chain_bufs(c);
}

nqh_status nqh_set_iteration(nqh_ctx* c, uint32_t iter) {
    if (!c) return NQH_E_ARG;
    drop_graph(c);
    c->iter = iter;
    return NQH_OK;
}

// one global (independence) move of every chain, proposals drawn from the MADE
static nqh_status made_global_move(nqh_ctx* c, uint32_t slice, const uint32_t* iter_dev) {
    const int64_t n = c->chains;
    uint64_t* kk = P<uint64_t>(c->made_keys);
    HIPCHK(c, nq_launch_copy_keys(P<uint64_t>(c->chain_state), n, kk, c->stream));
    HIPCHK(c, nq_launch_made_sample(P<float>(c->made_th), c->made_D, n, c->seed, c->iter, iter_dev, slice, kk + n,
                                    nullptr, c->stream));
    HIPCHK(c, nq_launch_made_logq(P<float>(c->made_th), c->made_D, kk, 2 * n, P<float>(c->made_lq), c->stream));
    HIPCHK(c, launch_logpsi(c, kk, 2 * n, c->made_lp.p, nullptr));
    HIPCHK(c, nq_launch_made_accept(kk, c->made_lp.p, fp64_net(c), P<float>(c->made_lq), n, c->seed, c->iter,
                                    iter_dev, slice, P<uint64_t>(c->chain_state),
                                    P<unsigned long long>(c->acc) + 1, P<uint32_t>(c->chain_acc), c->stream));
    return NQH_OK;
}

// host mirror of the device ring state after one push of a sample set
static void ring_mirror_advance(nqh_ctx* c) {
    const int64_t m = std::min(c->chains * c->chain_len, c->ring_cap);
    c->ring_pos = (c->ring_pos + m) % c->ring_cap;
    c->ring_fill = std::min(c->ring_cap, c->ring_fill + m);
}

// slice i is preceded by a global move iff floor((i + 1) mix) > floor(i mix)
static bool made_scheduled(double mix, int i) {
    return std::floor((i + 1) * mix + 1e-12) > std::floor(i * mix + 1e-12);
}

static nqh_status run_sampler(nqh_ctx* c) {
    if (!c->has_sampler) return fail(c, NQH_E_STATE, "sampler not set (nqh_set_sampler)");
    samp_args a{};
    a.d = c->d;
    a.p = fp64_net(c) ? c->p64.p : c->p32.p;
    a.chains = c->chains;
    a.passes = c->passes;
    const bool compat = (c->flags & NQH_SAMPLER_COMPAT_STALE_LAST_SLICE) != 0;
    const bool persistent = (c->flags & NQH_SAMPLER_PERSISTENT) != 0;
    const bool mixed = c->made_ready && c->made_mix > 0.0;
    if (mixed && compat) return fail(c, NQH_E_STATE, "the learned proposal and the compat sampler flag are exclusive");
    a.restart = (!persistent || !c->chains_initialised) ? 1 : 0;
    a.burn = a.restart ? c->burn : 0;
    a.stored = compat ? c->chain_len - 1 : c->chain_len;
    a.compat_stale_last_slice = compat ? 1 : 0;
    a.lpc = is_rbm(c) ? 1 : nq_sampler_lanes_per_chain(c->d);
    a.seed = c->seed;
    a.iter = c->iter;
    a.chain_state = P<uint64_t>(c->chain_state);
    a.samples = P<uint64_t>(c->samples);
    a.accepted = P<unsigned long long>(c->acc);
    a.chain_acc = P<uint32_t>(c->chain_acc);
    a.iter_dev = c->use_graph ? P<uint32_t>(c->iter_dev) : nullptr;
    a.waves = c->tune_samp_waves;
    phase_begin(c, PH_SAMPLER);
    HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->acc), 4, 0u, c->stream));
    HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->chain_acc), c->chains, 0u, c->stream));
    if (!mixed) {
        HIPCHK(c, launch_metropolis(c, a));
    } else {
        // burn-in (and restart) first, then per stored slice: [global move],
        // `passes` local proposals, store.  Proposal and slice indices continue
        // across the launches, so the local moves use the same random streams
        // as one uninterrupted launch.
        const int stored = a.stored, burn = a.burn;
        a.stored = 0;
        HIPCHK(c, launch_metropolis(c, a));
        a.restart = 0;
        a.burn = 0;
        a.stored = 1;
        c->made_proposed = 0;
        for (int i = 0; i < stored; ++i) {
            if (made_scheduled(c->made_mix, i)) {
                nqh_status s = made_global_move(c, (uint32_t)i, a.iter_dev);
                if (s) return s;
                c->made_proposed += c->chains;
            }
            a.t0 = (uint32_t)((burn + i) * c->passes);
            a.slice0 = i;
            HIPCHK(c, launch_metropolis(c, a));
        }
    }
    if (c->made_ready && c->ring_cap > 0) {
        const int64_t Ns = c->chains * c->chain_len;
        HIPCHK(c, nq_launch_ring_push(P<uint64_t>(c->samples), Ns, P<uint64_t>(c->ring), c->ring_cap,
                                      P<int64_t>(c->ring_posd), c->stream));
        ring_mirror_advance(c);
    }
    phase_end(c, PH_SAMPLER);
    c->chains_initialised = true;
    c->iter += 1;
    return NQH_OK;
}

nqh_status nqh_sample(nqh_ctx* c, uint64_t* keys_out, int64_t* accepted) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_net(c))) return s;
    if ((s = made_maybe_retrain(c))) return s;
    if ((s = run_sampler(c))) return s;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    phase_collect(c);
    if (keys_out)
        HIPCHK(c, copy_sync(c->stream, keys_out, c->samples.p, sizeof(uint64_t) * c->chains * c->chain_len, hipMemcpyDeviceToHost));
    if (accepted) {
        unsigned long long v = 0;
        HIPCHK(c, copy_sync(c->stream, &v, c->acc.p, sizeof(v), hipMemcpyDeviceToHost));
        *accepted = (int64_t)v;
    }
    return NQH_OK;
}

nqh_status nqh_get_chain_acceptance(nqh_ctx* c, uint32_t* accepted_out, int64_t chains) {
    if (!c || !accepted_out) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    if (!c->has_sampler || !c->chain_acc.p) return fail(c, NQH_E_STATE, "sampler not set (nqh_set_sampler)");
    if (chains != c->chains) return fail(c, NQH_E_ARG, "chains must equal the sampler's chain count");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, copy_sync(c->stream, accepted_out, c->chain_acc.p, sizeof(uint32_t) * (size_t)chains, hipMemcpyDeviceToHost));
    return NQH_OK;
}

static nqh_status sample_grad_async(nqh_ctx* c) {
    nqh_status s;
    if ((s = need_net(c)) || (s = need_problem(c))) return s;
    if (!c->has_sampler) return fail(c, NQH_E_STATE, "sampler not set (nqh_set_sampler)");
    const int64_t Ns = c->chains * c->chain_len;
    if ((s = ensure_samples(c, Ns))) return s;
    if ((s = run_sampler(c))) return s;
    return run_grad_pipeline(c, P<uint64_t>(c->samples), c->chains, c->chain_len);
}

nqh_status nqh_sample_grad(nqh_ctx* c, double* stats) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s = made_maybe_retrain(c);
    if (s) return s;
    s = sample_grad_async(c);
    if (s) return s;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    phase_collect(c);
    double st[5];
    HIPCHK(c, copy_sync(c->stream, st, c->stats.p, sizeof(st), hipMemcpyDeviceToHost));
    if (stats) std::memcpy(stats, st, sizeof(st));
    if (!std::isfinite(st[0])) return fail(c, NQH_E_NAN, "non-finite loss <L^dag L>");
    return NQH_OK;
}

// matrix-free conjugate gradients on (S + shift) x = F (SRIterative.jl:71-152
// with sr_cg; S v from SrMatrixR, SR_notfull.jl:123-148).  Converged when
// ||r|| <= tol ||F|| (IterativeSolvers' relative-residual rule).  Without
// convergence in maxiter products the step is zero, as the reference sets
// Delta w = 0 after its retries (SRIterative.jl:134-150).  Synchronises with
// the host once per product (the convergence test).
static nqh_status cg_solve(nqh_ctx* c, double eps, int kind) {
    if (is_rbm(c)) return fail(c, NQH_E_STATE, "the matrix-free CG solver supports NDM networks only");
    const ndm_dims& d = c->d;
    const int np = d.np, R = d.R;
    const int64_t Ns = c->last_ns;
    const bool f64 = c->net_prec == NQH_PREC_FP64;
    int nA, nB, offB;
    gram_blocks(c, &nA, &nB, &offB);
    constexpr int chunks = 64, nparts = 64;
    nqh_status s;
    if ((s = ensure(c, c->cg_w, sizeof(double) * 2 * (size_t)Ns)) ||
        (s = ensure(c, c->cg_part, sizeof(double) * (size_t)chunks * R)) ||
        (s = ensure(c, c->cg_p, sizeof(double) * np)) || (s = ensure(c, c->cg_q, sizeof(double) * np)) ||
        (s = ensure(c, c->cg_r, sizeof(double) * np)) || (s = ensure(c, c->cg_diag, sizeof(double) * np)) ||
        (s = ensure(c, c->cg_sc, sizeof(double) * 8)) || (s = ensure(c, c->cg_dpart, sizeof(double) * nparts)))
        return s;
    double *x = P<double>(c->delta), *r = P<double>(c->cg_r), *p = P<double>(c->cg_p), *q = P<double>(c->cg_q);
    double* sc = P<double>(c->cg_sc);
    double* diag = P<double>(c->cg_diag);
    const double* mu = P<double>(c->mu);
    const double* F = P<double>(c->F);
    hipStream_t st = c->stream;
    HIPCHK(c, nq_launch_fill(x, np, 0.0, st));
    HIPCHK(c, hipMemcpyAsync(r, F, sizeof(double) * np, hipMemcpyDeviceToDevice, st));
    HIPCHK(c, hipMemcpyAsync(p, F, sizeof(double) * np, hipMemcpyDeviceToDevice, st));
    if (kind == NQH_SHIFT_DIAG)
        HIPCHK(c, nq_launch_sdiag(c->Y.p, f64, Ns, R, nA, offB, np, mu, P<double>(c->cg_part), chunks, diag, st));
    HIPCHK(c, nq_launch_dot(F, F, np, P<double>(c->cg_dpart), nparts, sc, st));
    double f2 = 0.0, rr = 0.0;
    HIPCHK(c, hipMemcpyAsync(&f2, sc, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    const int maxit = c->cg_maxiter > 0 ? c->cg_maxiter : 10 * np;
    const double target = c->cg_tol * c->cg_tol * f2;
    int it = 0;
    rr = f2;
    bool ok = f2 == 0.0;
    while (!ok && it < maxit) {
        HIPCHK(c, nq_launch_sv(c->Y.p, f64, Ns, R, nA, offB, np, mu, p, eps, kind, diag, P<double>(c->cg_w),
                               P<double>(c->cg_part), chunks, q, st));
        HIPCHK(c, nq_launch_dot(p, q, np, P<double>(c->cg_dpart), nparts, sc + 1, st));
        HIPCHK(c, nq_launch_cg_xr(x, r, p, q, np, sc, st));
        HIPCHK(c, nq_launch_dot(r, r, np, P<double>(c->cg_dpart), nparts, sc + 2, st));
        HIPCHK(c, hipMemcpyAsync(&rr, sc + 2, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        ++it;
        if (!std::isfinite(rr)) break;
        if (rr <= target) {
            ok = true;
            break;
        }
        HIPCHK(c, nq_launch_cg_p(p, r, np, sc, st));
        HIPCHK(c, nq_launch_copy_scalar(sc + 2, sc, st));
    }
    c->cg_iters = it;
    c->cg_resid = f2 > 0.0 ? std::sqrt(rr / f2) : 0.0;
    if (!ok && std::isfinite(rr)) HIPCHK(c, nq_launch_fill(x, np, 0.0, st));
    HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->info), 1, 0u, st));
    HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->flag), 1, 0u, st));
    HIPCHK(c, nq_launch_check_finite(x, np, P<int>(c->flag), st));
    return NQH_OK;
}

// Givens rotation [c s; -s c] [f; g] = [r; 0] with LAPACK dlartg's sign
// convention (c > 0 when |f| > |g|), as LinearAlgebra.givensAlgorithm
static void givens(double f, double g, double* cs, double* sn, double* r) {
    if (g == 0.0) {
        *cs = 1.0, *sn = 0.0, *r = f;
    } else if (f == 0.0) {
        *cs = 0.0, *sn = 1.0, *r = g;
    } else {
        const double h = std::hypot(f, g);
        *cs = f / h, *sn = g / h, *r = h;
        if (std::fabs(f) > std::fabs(g) && *cs < 0.0) *cs = -*cs, *sn = -*sn, *r = -*r;
    }
}

// The scalar part of the reference's MINRES-QLP (minresqlp.jl, sr_qlp),
// real, unpreconditioned, unshifted: one Lanczos step's (alfa, betan) in, the
// solution-update coefficients and the stopping flags out.  As that file is
// written: its MINRES/QLP branch test compares `flag != flag0`, so it always
// updates x through the QLP branch, whose iterates equal those of the MINRES
// branch used here in exact arithmetic; gmin keeps its first-step value
// (minresqlp.jl:344 assigns a local); flags 2, 4, 6, 7 take the step count
// back; any flag counts as converged, maxiter (flag 8) included.
struct qlp_scalars {
    static constexpr double kTranCond = 10e6, kMaxXnorm = 10e6, kAcondLim = 10e14;
    double tol;
    int maxiter;
    int flag0 = -2, flag = -2, iters = 0;
    double beta = 0, tau = 0, taul = 0, taul2 = 0, phi, beta1, betan, gmin = 0;
    double cs = -1, sn = 0, cr1 = -1, sr1 = 0, cr2 = -1, sr2 = 0, dltan = 0, eplnn = 0;
    double gama = 0, gamal = 0, gamal2 = 0, eta = 0, etal = 0, etal2 = 0, vepln = 0, veplnl = 0, veplnl2 = 0;
    double ul3 = 0, ul2 = 0, ul = 0, u = 0, rnorm, xnorm = 0, relares = 0, xl2norm = 0, anorm = 0, acond = 1;
    double relres, resnorm = 123, pnorm = 0, gbar = 0;

    qlp_scalars(double b1, double t, int mi)
        : tol(t), maxiter(mi), phi(b1), beta1(b1), betan(b1), rnorm(b1), relres(b1 / (b1 + 1e-50)) {}

    // SymGivens (minresqlp.jl:484-509): r >= 0
    static void sym_givens(double a, double b, double* c, double* s, double* r) {
        if (b == 0.0) {
            *c = a == 0.0 ? 1.0 : std::copysign(1.0, a), *s = 0.0, *r = std::fabs(a);
        } else if (a == 0.0) {
            *c = 0.0, *s = std::copysign(1.0, b), *r = std::fabs(b);
        } else if (std::fabs(b) > std::fabs(a)) {
            const double t = a / b;
            *s = std::copysign(1.0, b) / std::sqrt(1 + t * t), *c = *s * t, *r = b / *s;
        } else {
            const double t = b / a;
            *c = std::copysign(1.0, a) / std::sqrt(1 + t * t), *s = *c * t, *r = a / *c;
        }
    }
    bool converged() const { return resnorm <= tol || flag != flag0; }
    bool done(int iteration) const { return iteration > maxiter || converged(); }
    void begin(double* betal, double* b) {
        ++iters;
        *betal = beta;
        beta = betan;
        *b = beta;
    }
    // minresqlp.jl:200-261; returns w = (v - epln w_{k-2} - dlta w_{k-1}) / gama, x += tau w
    void rotate(double betal, double alfa, double bn, double* epln, double* dlta_qlp, double* gama_w, double* tau_w) {
        betan = bn;
        pnorm = std::sqrt(betal * betal + alfa * alfa + betan * betan);
        const double dbar = dltan;
        double dlta = cs * dbar + sn * alfa;
        *epln = eplnn;
        gbar = sn * dbar - cs * alfa;
        eplnn = sn * betan;
        dltan = -cs * betan;
        *dlta_qlp = dlta;
        gamal2 = gamal;
        gamal = gama;
        sym_givens(gbar, betan, &cs, &sn, &gama);
        *gama_w = gama;
        taul2 = taul;
        taul = tau;
        tau = cs * phi;
        phi = sn * phi;
        *tau_w = tau;
        if (iters > 2) {
            veplnl2 = veplnl;
            etal2 = etal;
            etal = eta;
            const double d = sr2 * vepln - cr2 * dlta;
            veplnl = cr2 * vepln + sr2 * dlta;
            dlta = d;
            eta = sr2 * gama;
            gama = -cr2 * gama;
        }
        if (iters > 1) {
            sym_givens(gamal, dlta, &cr1, &sr1, &gamal);
            vepln = sr1 * gama;
            gama = -cr1 * gama;
        }
        const double ul4 = ul3;
        ul3 = ul2;
        if (iters > 2) ul2 = (taul2 - etal2 * ul4 - veplnl2 * ul3) / gamal2;
        if (iters > 1) ul = (taul - etal * ul3 - veplnl * ul2) / gamal;
        const double xn = std::sqrt(xl2norm * xl2norm + ul2 * ul2 + ul * ul);
        if (std::fabs(gama) > DBL_MIN && xn < kMaxXnorm) {
            u = (tau - eta * ul2 - vepln * ul) / gama;
            if (std::sqrt(xn * xn + u * u) > kMaxXnorm) u = 0.0, flag = 6;
        } else {
            u = 0.0, flag = 9;
        }
        xl2norm = std::sqrt(xl2norm * xl2norm + ul2 * ul2);
        xnorm = std::sqrt(xl2norm * xl2norm + ul * ul + u * u);
    }
    // minresqlp.jl:328-396
    void finish() {
        sym_givens(gamal, eplnn, &cr2, &sr2, &gamal);
        anorm = std::max(std::max(anorm, pnorm), std::max(gamal, std::fabs(gama)));
        if (iters == 1) gmin = gama;
        const double acondl = acond, rnorml = rnorm, relresl = relres;
        acond = anorm / gmin;
        if (flag != 9) rnorm = phi;
        relres = rnorm / (anorm * xnorm + beta1);
        const double rootl = std::sqrt(gbar * gbar + dltan * dltan);
        relares = rootl / anorm;
        const double epsx = anorm * xnorm * DBL_EPSILON;
        if (flag == flag0 || flag == 9) {
            const double t1 = 1 + relres, t2 = 1 + relares;
            if (iters >= maxiter) flag = 8;
            if (acond >= kAcondLim) flag = 7;
            if (xnorm >= kMaxXnorm) flag = 6;
            if (epsx >= beta1) flag = 5;
            if (t2 <= 1) flag = 4;
            if (t1 <= 1) flag = 3;
            if (relares <= tol) flag = 2;
            if (relres <= tol) flag = 1;
        }
        if (flag == 2 || flag == 4 || flag == 6 || flag == 7) {
            --iters;
            acond = acondl, rnorm = rnorml, relres = relresl;
        }
        resnorm = std::min(relres, relares);
    }
};

// sr_qlp: the Lanczos vectors and the x update of minres_solve, the
// coefficients and the stopping rule of qlp_scalars; x = F / alfa when F is an
// eigenvector (betan = 0 at the first step, the file's flag -1 intent)
static nqh_status qlp_loop(nqh_ctx* c, double eps, int kind, double res0, int maxit, double* vw, double* x,
                           const double* diag, const double* mu, bool f64, int64_t Ns, int R, int nA, int offB,
                           int chunks, int nparts) {
    const int np = c->d.np;
    double *vp = vw, *vc = vw + np, *vn = vw + 2 * (size_t)np;
    double *wp = vw + 3 * (size_t)np, *wc = vw + 4 * (size_t)np, *wn = vw + 5 * (size_t)np;
    double* sc = P<double>(c->cg_sc);
    hipStream_t st = c->stream;
    qlp_scalars q(res0, c->cg_tol, maxit);
    bool finite = std::isfinite(res0);
    int iteration = 1;
    if (res0 == 0.0) q.flag = 0;   // F = 0: x = 0
    while (finite && !q.done(iteration)) {
        double betal, beta;
        q.begin(&betal, &beta);
        HIPCHK(c, nq_launch_sv(c->Y.p, f64, Ns, R, nA, offB, np, mu, vc, eps, kind, diag, P<double>(c->cg_w),
                               P<double>(c->cg_part), chunks, vn, st));
        if (q.iters > 1) HIPCHK(c, nq_launch_mr_axpy(vn, vp, np, -beta, st));
        HIPCHK(c, nq_launch_dot(vc, vn, np, P<double>(c->cg_dpart), nparts, sc, st));
        HIPCHK(c, nq_launch_mr_orth(vn, vc, np, sc, st));
        HIPCHK(c, nq_launch_dot(vn, vn, np, P<double>(c->cg_dpart), nparts, sc + 1, st));
        double hv[2];
        HIPCHK(c, hipMemcpyAsync(hv, sc, sizeof(hv), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if (!std::isfinite(hv[0]) || !std::isfinite(hv[1])) {
            finite = false;
            break;
        }
        const double alfa = hv[0], betan = std::sqrt(hv[1]);
        if (q.iters == 1 && betan == 0.0) {
            HIPCHK(c, nq_launch_mr_axpy(x, vc, np, res0 / alfa, st));   // vc = F / ||F||
            q.flag = -1;
            break;
        }
        double epln, dlta, gama, tau;
        q.rotate(betal, alfa, betan, &epln, &dlta, &gama, &tau);
        double rhs = tau;
        if (!(q.xnorm < qlp_scalars::kMaxXnorm)) rhs = 0.0, q.flag = 6;
        HIPCHK(c, nq_launch_mr_step(vn, vc, wn, wc, wp, x, np, betan > 0.0 ? 1.0 / betan : 0.0, dlta, epln,
                                    1.0 / gama, rhs, st));
        double* t = vp;
        vp = vc, vc = vn, vn = t;
        t = wp;
        wp = wc, wc = wn, wn = t;
        q.finish();
        iteration = q.iters;
    }
    if (!finite) return fail(c, NQH_E_NAN, "non-finite value in the MINRES-QLP recurrence");
    c->cg_iters = q.iters;
    c->cg_resid = q.relres;
    HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->info), 1, 0u, st));
    HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->flag), 1, 0u, st));
    HIPCHK(c, nq_launch_check_finite(x, np, P<int>(c->flag), st));
    return NQH_OK;
}

// matrix-free MINRES on (S + shift) x = F (SRIterative.jl:101-108, sr_minres:
// IterativeSolvers.jl 0.8.1 `minres`, the Lanczos recurrence with Givens
// rotations of Paige and Saunders): per product one Lanczos step (two dot
// products, read by the host together), the rotations on the host, and one
// fused update of v, w and x.  Converged when the recurrence's residual norm
// |rhs_2| <= tol ||F||; without convergence in maxiter products the step is
// zero, as with CG.  Synchronises with the host once per product.
static nqh_status minres_solve(nqh_ctx* c, double eps, int kind, bool qlp) {
    if (is_rbm(c)) return fail(c, NQH_E_STATE, "the matrix-free MINRES solvers support NDM networks only");
    const ndm_dims& d = c->d;
    const int np = d.np, R = d.R;
    const int64_t Ns = c->last_ns;
    const bool f64 = c->net_prec == NQH_PREC_FP64;
    int nA, nB, offB;
    gram_blocks(c, &nA, &nB, &offB);
    constexpr int chunks = 64, nparts = 64;
    nqh_status s;
    if ((s = ensure(c, c->cg_w, sizeof(double) * 2 * (size_t)Ns)) ||
        (s = ensure(c, c->cg_part, sizeof(double) * (size_t)chunks * R)) ||
        (s = ensure(c, c->mr_vw, sizeof(double) * 6 * (size_t)np)) ||
        (s = ensure(c, c->cg_diag, sizeof(double) * np)) || (s = ensure(c, c->cg_sc, sizeof(double) * 8)) ||
        (s = ensure(c, c->cg_dpart, sizeof(double) * nparts)))
        return s;
    double* vw = P<double>(c->mr_vw);
    double *vp = vw, *vc = vw + np, *vn = vw + 2 * (size_t)np;
    double *wp = vw + 3 * (size_t)np, *wc = vw + 4 * (size_t)np, *wn = vw + 5 * (size_t)np;
    double* x = P<double>(c->delta);
    double* sc = P<double>(c->cg_sc);
    double* diag = P<double>(c->cg_diag);
    const double* mu = P<double>(c->mu);
    const double* F = P<double>(c->F);
    hipStream_t st = c->stream;
    HIPCHK(c, nq_launch_fill(x, np, 0.0, st));
    HIPCHK(c, nq_launch_fill(vw, 6 * np, 0.0, st));   // w_prev, w_curr enter with zero weight: keep them finite
    HIPCHK(c, hipMemcpyAsync(vc, F, sizeof(double) * np, hipMemcpyDeviceToDevice, st));
    if (kind == NQH_SHIFT_DIAG)
        HIPCHK(c, nq_launch_sdiag(c->Y.p, f64, Ns, R, nA, offB, np, mu, P<double>(c->cg_part), chunks, diag, st));
    HIPCHK(c, nq_launch_dot(F, F, np, P<double>(c->cg_dpart), nparts, sc, st));
    double f2 = 0.0;
    HIPCHK(c, hipMemcpyAsync(&f2, sc, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    const double res0 = std::sqrt(f2);
    const int maxit = c->cg_maxiter > 0 ? c->cg_maxiter : 10 * np;
    const double tolerance = c->cg_tol * res0;
    double resnorm = res0;
    if (res0 > 0.0) HIPCHK(c, nq_launch_mr_scale(vc, np, 1.0 / res0, st));
    if (qlp) return qlp_loop(c, eps, kind, res0, maxit, vw, x, diag, mu, f64, Ns, R, nA, offB, chunks, nparts);
    // H = the active column of the tridiagonal (h1 .. h4), rhs = the rotated right-hand side
    double h1 = 0.0, h2 = 0.0, h3 = 0.0, h4 = 0.0, rhs1 = res0, rhs2 = 0.0;
    double c_prev = 0.0, s_prev = 0.0, c_curr = 0.0, s_curr = 0.0;
    int it = 1;
    bool finite = std::isfinite(res0);
    while (finite && !(it > maxit || resnorm <= tolerance)) {
        HIPCHK(c, nq_launch_sv(c->Y.p, f64, Ns, R, nA, offB, np, mu, vc, eps, kind, diag, P<double>(c->cg_w),
                               P<double>(c->cg_part), chunks, vn, st));
        if (it > 1) HIPCHK(c, nq_launch_mr_axpy(vn, vp, np, -h2, st));
        HIPCHK(c, nq_launch_dot(vc, vn, np, P<double>(c->cg_dpart), nparts, sc, st));
        HIPCHK(c, nq_launch_mr_orth(vn, vc, np, sc, st));
        HIPCHK(c, nq_launch_dot(vn, vn, np, P<double>(c->cg_dpart), nparts, sc + 1, st));
        double hv[2];
        HIPCHK(c, hipMemcpyAsync(hv, sc, sizeof(hv), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if (!std::isfinite(hv[0]) || !std::isfinite(hv[1])) {
            finite = false;
            break;
        }
        h3 = hv[0];
        h4 = std::sqrt(hv[1]);
        if (it > 2) {
            h1 = s_prev * h2;
            h2 = c_prev * h2;
        }
        if (it > 1) {
            const double t = -s_curr * h2 + c_curr * h3;
            h2 = c_curr * h2 + s_curr * h3;
            h3 = t;
        }
        double cn, sn;
        givens(h3, h4, &cn, &sn, &h3);
        rhs2 = -sn * rhs1;
        rhs1 = cn * rhs1;
        HIPCHK(c, nq_launch_mr_step(vn, vc, wn, wc, wp, x, np, h4 > 0.0 ? 1.0 / h4 : 0.0, it > 1 ? h2 : 0.0,
                                    it > 2 ? h1 : 0.0, 1.0 / h3, rhs1, st));
        double* t = vp;
        vp = vc, vc = vn, vn = t;
        t = wp;
        wp = wc, wc = wn, wn = t;
        c_prev = c_curr, s_prev = s_curr, c_curr = cn, s_curr = sn;
        rhs1 = rhs2;
        h2 = h4;
        resnorm = std::fabs(rhs2);
        ++it;
    }
    if (!finite) return fail(c, NQH_E_NAN, "non-finite value in the MINRES recurrence");
    const bool ok = resnorm <= tolerance;
    c->cg_iters = it - 1;
    c->cg_resid = res0 > 0.0 ? resnorm / res0 : 0.0;
    if (!ok) HIPCHK(c, nq_launch_fill(x, np, 0.0, st));
    HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->info), 1, 0u, st));
    HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->flag), 1, 0u, st));
    HIPCHK(c, nq_launch_check_finite(x, np, P<int>(c->flag), st));
    return NQH_OK;
}

// matrix-free LSQR on (S + shift) x = F (SRIterative.jl:109-116, sr_lsq:
// IterativeSolvers.jl 0.8.1 `lsqr` with its defaults, which the reference
// does not override: atol = btol = sqrt(eps), conlim = 1/sqrt(eps); the
// published method of Paige and Saunders without damping).  Per step two
// products (A v and A^T u = A u: S is symmetric), two host reads (beta, then
// alfa with ||w||^2), the rotation and the stopping tests on the host, one
// fused update of v, x and w.  Converged unless the run stops on the
// condition estimate or maxiter (istop 3, 6, 7): then the step is zero, as
// with the other iterative solvers.
static nqh_status lsqr_solve(nqh_ctx* c, double eps, int kind) {
    if (is_rbm(c)) return fail(c, NQH_E_STATE, "the matrix-free LSQR solver supports NDM networks only");
    const ndm_dims& d = c->d;
    const int np = d.np, R = d.R;
    const int64_t Ns = c->last_ns;
    const bool f64 = c->net_prec == NQH_PREC_FP64;
    int nA, nB, offB;
    gram_blocks(c, &nA, &nB, &offB);
    constexpr int chunks = 64, nparts = 64;
    nqh_status s;
    if ((s = ensure(c, c->cg_w, sizeof(double) * 2 * (size_t)Ns)) ||
        (s = ensure(c, c->cg_part, sizeof(double) * (size_t)chunks * R)) ||
        (s = ensure(c, c->mr_vw, sizeof(double) * 6 * (size_t)np)) ||
        (s = ensure(c, c->cg_diag, sizeof(double) * np)) || (s = ensure(c, c->cg_sc, sizeof(double) * 8)) ||
        (s = ensure(c, c->cg_dpart, sizeof(double) * nparts)))
        return s;
    double* vw = P<double>(c->mr_vw);
    double *u = vw, *v = vw + np, *w = vw + 2 * (size_t)np, *t = vw + 3 * (size_t)np;
    double* x = P<double>(c->delta);
    double* sc = P<double>(c->cg_sc);
    double* diag = P<double>(c->cg_diag);
    const double* mu = P<double>(c->mu);
    const double* F = P<double>(c->F);
    hipStream_t st = c->stream;
    auto sv = [&](const double* in, double* out) {
        return nq_launch_sv(c->Y.p, f64, Ns, R, nA, offB, np, mu, in, eps, kind, diag, P<double>(c->cg_w),
                            P<double>(c->cg_part), chunks, out, st);
    };
    auto read = [&](double* h, int n) -> hipError_t {
        hipError_t e = hipMemcpyAsync(h, sc, sizeof(double) * n, hipMemcpyDeviceToHost, st);
        return e != hipSuccess ? e : hipStreamSynchronize(st);
    };
    const double atol = std::sqrt(DBL_EPSILON), btol = atol, ctol = atol;   // ctol = 1 / conlim
    HIPCHK(c, nq_launch_fill(x, np, 0.0, st));
    if (kind == NQH_SHIFT_DIAG)
        HIPCHK(c, nq_launch_sdiag(c->Y.p, f64, Ns, R, nA, offB, np, mu, P<double>(c->cg_part), chunks, diag, st));
    // u = F / beta, v = A u / alfa, w = v
    HIPCHK(c, hipMemcpyAsync(u, F, sizeof(double) * np, hipMemcpyDeviceToDevice, st));
    HIPCHK(c, nq_launch_dot(F, F, np, P<double>(c->cg_dpart), nparts, sc, st));
    double h[3];
    HIPCHK(c, read(h, 1));
    double beta = std::sqrt(h[0]), alfa = 0.0;
    const double bnorm = beta;
    if (beta > 0.0) {
        HIPCHK(c, nq_launch_mr_scale(u, np, 1.0 / beta, st));
        HIPCHK(c, sv(u, v));
        HIPCHK(c, nq_launch_dot(v, v, np, P<double>(c->cg_dpart), nparts, sc, st));
        HIPCHK(c, read(h, 1));
        alfa = std::sqrt(h[0]);
    }
    if (alfa > 0.0) HIPCHK(c, nq_launch_mr_scale(v, np, 1.0 / alfa, st));
    HIPCHK(c, hipMemcpyAsync(w, v, sizeof(double) * np, hipMemcpyDeviceToDevice, st));
    double anorm = 0.0, ddnorm = 0.0, xxnorm = 0.0, z = 0.0, cs2 = -1.0, sn2 = 0.0, rnorm = beta;
    double rhobar = alfa, phibar = beta;
    const int maxit = c->cg_maxiter > 0 ? c->cg_maxiter : 10 * np;
    int itn = 0, istop = 0;
    bool finite = std::isfinite(alfa) && std::isfinite(beta);
    if (finite && alfa * beta != 0.0) {
        while (itn < maxit) {
            ++itn;
            // beta u = A v - alfa u
            HIPCHK(c, sv(v, t));
            HIPCHK(c, nq_launch_mr_axpy(t, u, np, -alfa, st));
            std::swap(u, t);
            HIPCHK(c, nq_launch_dot(u, u, np, P<double>(c->cg_dpart), nparts, sc, st));
            HIPCHK(c, read(h, 1));
            beta = std::sqrt(h[0]);
            double inv_alfa = 1.0;
            if (beta > 0.0) {
                // alfa v = A u - beta v (v normalised by the step kernel)
                HIPCHK(c, nq_launch_mr_scale(u, np, 1.0 / beta, st));
                anorm = std::sqrt(anorm * anorm + alfa * alfa + beta * beta);
                HIPCHK(c, sv(u, t));
                HIPCHK(c, nq_launch_mr_axpy(t, v, np, -beta, st));
                std::swap(v, t);
                HIPCHK(c, nq_launch_dot(v, v, np, P<double>(c->cg_dpart), nparts, sc, st));
                HIPCHK(c, nq_launch_dot(w, w, np, P<double>(c->cg_dpart), nparts, sc + 1, st));
                HIPCHK(c, read(h, 2));
                alfa = std::sqrt(h[0]);
                if (alfa > 0.0) inv_alfa = 1.0 / alfa;
            } else {
                HIPCHK(c, nq_launch_dot(w, w, np, P<double>(c->cg_dpart), nparts, sc + 1, st));
                HIPCHK(c, read(h, 2));
            }
            if (!std::isfinite(beta) || !std::isfinite(alfa) || !std::isfinite(h[1])) {
                finite = false;
                break;
            }
            const double rho = std::hypot(rhobar, beta), cs = rhobar / rho, sn = beta / rho;
            const double theta = sn * alfa;
            rhobar = -cs * alfa;
            const double phi = cs * phibar;
            phibar = sn * phibar;
            const double tau = sn * phi;
            const double t1 = phi / rho, t2 = -theta / rho;
            ddnorm += h[1] / (rho * rho);
            HIPCHK(c, nq_launch_ls_step(v, w, x, np, inv_alfa, t1, t2, st));
            const double delta = sn2 * rho, gambar = -cs2 * rho, rhs = phi - delta * z, zbar = rhs / gambar;
            const double xnorm = std::sqrt(xxnorm + zbar * zbar), gamma = std::hypot(gambar, theta);
            cs2 = gambar / gamma, sn2 = theta / gamma;
            z = rhs / gamma;
            xxnorm += z * z;
            const double acond = anorm * std::sqrt(ddnorm);
            rnorm = std::fabs(phibar);
            const double arnorm = alfa * std::fabs(tau);
            const double test1 = rnorm / bnorm;
            const double test2 = anorm * rnorm > 0.0 ? arnorm / (anorm * rnorm) : 0.0;
            const double test3 = acond > 0.0 ? 1.0 / acond : HUGE_VAL;
            const double t1n = test1 / (1 + anorm * xnorm / bnorm), rtol = btol + atol * anorm * xnorm / bnorm;
            if (itn >= maxit) istop = 7;
            if (1 + test3 <= 1) istop = 6;
            if (1 + test2 <= 1) istop = 5;
            if (1 + t1n <= 1) istop = 4;
            if (test3 <= ctol) istop = 3;
            if (test2 <= atol) istop = 2;
            if (test1 <= rtol) istop = 1;
            if (istop) break;
        }
    }
    if (!finite) return fail(c, NQH_E_NAN, "non-finite value in the LSQR recurrence");
    const bool ok = istop != 3 && istop != 6 && istop != 7;
    c->cg_iters = itn;
    c->cg_resid = bnorm > 0.0 ? rnorm / bnorm : 0.0;
    if (!ok) HIPCHK(c, nq_launch_fill(x, np, 0.0, st));
    HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->info), 1, 0u, st));
    HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->flag), 1, 0u, st));
    HIPCHK(c, nq_launch_check_finite(x, np, P<int>(c->flag), st));
    return NQH_OK;
}

// lr given (a whole iteration): the split solve's last launch also applies the
// guarded step and writes the status record, and *stepped says so
static nqh_status precondition_async(nqh_ctx* c, double eps, int kind, int refine, const double* lr = nullptr,
                                     bool* stepped = nullptr) {
    if (!c->grad_ready) return fail(c, NQH_E_STATE, "no gradient: call nqh_sample_grad first");
    if (kind != NQH_SHIFT_ADD && kind != NQH_SHIFT_DIAG) return fail(c, NQH_E_ARG, "bad shift kind");
    if (!(eps >= 0.0) || refine < 0 || refine > 8) return fail(c, NQH_E_ARG, "eps >= 0, refine in 0..8");
    if (c->solver != NQH_SOLVER_CHOLESKY) {
        phase_begin(c, PH_SOLVE);
        const nqh_status s = c->solver == NQH_SOLVER_CG     ? cg_solve(c, eps, kind)
                             : c->solver == NQH_SOLVER_LSQR ? lsqr_solve(c, eps, kind)
                                                            : minres_solve(c, eps, kind, c->solver == NQH_SOLVER_MINRES_QLP);
        phase_end(c, PH_SOLVE);
        return s;
    }
    const int np = nsys(c), ldS = np + 1;
    double* S = P<double>(c->S);
    phase_begin(c, PH_SOLVE);
    if (refine == 0 && use_split(c)) {
        // lambda and mu factored side by side, then the Pi block (no refinement)
        HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->info), 1, 0u, c->stream));
        phase_begin(c, PH_FACTOR);
        HIPCHK(c, nq_split_factor(S, ldS, P<double>(c->F), c->d.nlam, c->d.npi, eps, kind, P<double>(c->A2),
                                  P<double>(c->winv2), P<int>(c->info), c->stream));
        phase_end(c, PH_FACTOR);
        HIPCHK(c, nq_split_substitute(c->d.nlam, c->d.npi, P<double>(c->A2), P<double>(c->winv2), P<double>(c->y2),
                                      P<double>(c->x2), c->stream));
        const bool step = lr != nullptr;
        HIPCHK(c, nq_split_tail(c->d.nlam, c->d.npi, P<double>(c->x2), P<double>(c->delta), P<int>(c->flag),
                                step ? P<double>(c->p64) : nullptr, P<float>(c->p32), c->d, step ? *lr : 0.0,
                                P<int>(c->info), P<double>(c->stats), step ? c->hstat : nullptr, c->stream));
        if (stepped) *stepped = step;
        phase_end(c, PH_SOLVE);
        return NQH_OK;
    }
    if (refine > 0)
        HIPCHK(c, hipMemcpyAsync(c->S0.p, S, sizeof(double) * (size_t)ldS * ldS, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, nq_launch_shift(S, np, ldS, eps, kind, c->stream));
    HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->info), 1, 0u, c->stream));
    HIPCHK(c, nq_set_aug_row(S, np, ldS, P<double>(c->F), c->stream));
    // the factorization also forward-substitutes F (carried as row np)
    phase_begin(c, PH_FACTOR);
    HIPCHK(c, nq_cholesky(S, np, ldS, P<int>(c->info), P<double>(c->winv), c->stream));
    phase_end(c, PH_FACTOR);
    HIPCHK(c, nq_copy_aug_col(S, np, ldS, P<double>(c->resid), c->stream));
    HIPCHK(c, nq_cholesky_backsolve(S, np, ldS, P<double>(c->winv), P<double>(c->resid), P<double>(c->delta),
                                    c->stream));
    for (int it = 0; it < refine; ++it) {
        HIPCHK(c, nq_residual(P<double>(c->S0), np, ldS, eps, kind, P<double>(c->delta), P<double>(c->F),
                              P<double>(c->resid), c->stream));
        HIPCHK(c, nq_cholesky_solve(S, np, ldS, P<double>(c->winv), P<double>(c->resid), P<double>(c->tmp),
                                    c->stream));
        HIPCHK(c, nq_launch_axpy(P<double>(c->delta), P<double>(c->resid), np, c->stream));
    }
    HIPCHK(c, nq_launch_fill_u32(P<uint32_t>(c->flag), 1, 0u, c->stream));
    HIPCHK(c, nq_launch_check_finite(P<double>(c->delta), np, P<int>(c->flag), c->stream));
    phase_end(c, PH_SOLVE);
    return NQH_OK;
}

static nqh_status check_solve(nqh_ctx* c) {
    int info = 0, flag = 0;
    HIPCHK(c, copy_sync(c->stream, &info, c->info.p, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(c, copy_sync(c->stream, &flag, c->flag.p, sizeof(int), hipMemcpyDeviceToHost));
    if (info != 0) {
        char buf[128];
        std::snprintf(buf, sizeof(buf), "S + shift is not positive definite (column %d): no step applied", info);
        return fail(c, NQH_E_NOTPD, buf);
    }
    if (flag) return fail(c, NQH_E_NAN, "non-finite SR update: no step applied");
    return NQH_OK;
}

nqh_status nqh_precondition(nqh_ctx* c, double eps, int32_t shift_kind, int32_t refine, double* delta_out) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    c->solve_ok = false;
    if (!c->grad_ready) return fail(c, NQH_E_STATE, "no gradient: call nqh_sample_grad first");
    {
        // a non-finite loss (NaN in L_loc or in the network) is reported before any solve
        double loss = 0.0;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, copy_sync(c->stream, &loss, c->stats.p, sizeof(double), hipMemcpyDeviceToHost));
        if (!std::isfinite(loss)) return fail(c, NQH_E_NAN, "non-finite loss <L^dag L>: no solve");
    }
    nqh_status s = precondition_async(c, eps, shift_kind, refine);
    if (s) return s;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    phase_collect(c);
    if ((s = check_solve(c))) return s;
    c->solve_ok = true;
    if (delta_out && is_rbm(c)) {
        const int npc = c->rbm_npc;
        std::vector<double> dv(2 * npc);
        HIPCHK(c, copy_sync(c->stream, dv.data(), c->delta.p, sizeof(double) * 2 * npc, hipMemcpyDeviceToHost));
        for (int k = 0; k < npc; ++k) { delta_out[2 * k] = dv[k]; delta_out[2 * k + 1] = dv[npc + k]; }
    } else if (delta_out) {
        const ndm_dims& d = c->d;
        std::vector<double> dv(d.np);
        HIPCHK(c, copy_sync(c->stream, dv.data(), c->delta.p, sizeof(double) * d.np, hipMemcpyDeviceToHost));
        for (int k = 0; k < d.np; ++k) delta_out[p2ref_host(k, d)] = dv[k];
    }
    return NQH_OK;
}

nqh_status nqh_update(nqh_ctx* c, double lr) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    if (!c->grad_ready || !c->solve_ok)
        return fail(c, NQH_E_STATE, "no successful nqh_precondition for the current gradient: no step applied");
    phase_begin(c, PH_UPDATE);
    HIPCHK(c, launch_update(c, lr));
    phase_end(c, PH_UPDATE);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    phase_collect(c);
    return NQH_OK;
}

static void drop_graph(nqh_ctx* c) {
    c->gen++;
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->graph_valid = false;
}

// enqueue one whole SR iteration (no host synchronisation inside)
static nqh_status enqueue_iteration(nqh_ctx* c, double eps, int kind, int refine, double lr) {
    nqh_status s = sample_grad_async(c);
    if (s) return s;
    bool stepped = false;
    if ((s = precondition_async(c, eps, kind, refine, &lr, &stepped))) return s;
    if (stepped) return NQH_OK;
    phase_begin(c, PH_UPDATE);
    HIPCHK(c, launch_update(c, lr));
    phase_end(c, PH_UPDATE);
    HIPCHK(c, nq_launch_status_out(P<double>(c->stats), P<int>(c->info), P<int>(c->flag), c->hstat, c->stream));
    return NQH_OK;
}

static nqh_status graph_iteration(nqh_ctx* c, double eps, int kind, int refine, double lr) {
    const bool persistent = (c->flags & NQH_SAMPLER_PERSISTENT) != 0;
    if (persistent && !c->chains_initialised) {
        // the restart + burn-in iteration runs eagerly (iteration number from
        // the host, not from a possibly stale device copy); later ones replay a graph
        drop_graph(c);
        c->use_graph = false;
        const nqh_status s = enqueue_iteration(c, eps, kind, refine, lr);
        c->use_graph = true;
        return s;
    }
    if (!c->graph_valid || c->g_eps != eps || c->g_kind != kind || c->g_refine != refine || c->g_lr != lr) {
        drop_graph(c);
        nqh_status s;
        const int64_t Ns = c->chains * c->chain_len;
        if ((s = ensure_samples(c, Ns))) return s;
        if ((s = ensure(c, c->iter_dev, sizeof(uint32_t) * 4))) return s;
        HIPCHK(c, hipMemcpyAsync(c->iter_dev.p, &c->iter, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const bool prof = c->profiling;
        c->profiling = false;
        const uint32_t iter0 = c->iter;
        const int64_t ring_pos0 = c->ring_pos, ring_fill0 = c->ring_fill;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        s = enqueue_iteration(c, eps, kind, refine, lr);
        hipError_t e1 = nq_launch_bump(P<uint32_t>(c->iter_dev), c->stream);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(c->stream, &g);
        c->profiling = prof;
        c->iter = iter0;          // the capture did not run anything
        c->ring_pos = ring_pos0;
        c->ring_fill = ring_fill0;
        if (s) { if (g) (void)hipGraphDestroy(g); return s; }
        if (e1 != hipSuccess) { if (g) (void)hipGraphDestroy(g); return hip_fail(c, e1, "bump"); }
        if (e2 != hipSuccess) return hip_fail(c, e2, "hipStreamEndCapture");
        c->graph = g;
        HIPCHK(c, hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
        c->graph_valid = true;
        c->g_eps = eps; c->g_kind = kind; c->g_refine = refine; c->g_lr = lr;
    }
    HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    if (c->made_ready && c->ring_cap > 0) ring_mirror_advance(c);
    c->iter += 1;
    c->chains_initialised = true;
    c->grad_ready = true;
    c->solve_ok = false;
    c->last_ns = c->chains * c->chain_len;
    return NQH_OK;
}

nqh_status nqh_set_sr_solver(nqh_ctx* c, int32_t solver, double tol, int32_t maxiter) {
    if (!c) return NQH_E_ARG;
    if (solver != NQH_SOLVER_CHOLESKY && solver != NQH_SOLVER_CG && solver != NQH_SOLVER_MINRES &&
        solver != NQH_SOLVER_MINRES_QLP && solver != NQH_SOLVER_LSQR)
        return fail(c, NQH_E_ARG, "unknown solver");
    if (solver != NQH_SOLVER_CHOLESKY && !(tol > 0.0)) return fail(c, NQH_E_ARG, "tol must be > 0");
    if (maxiter < 0) return fail(c, NQH_E_ARG, "maxiter must be >= 0");
    (void)hipSetDevice(c->device);
    drop_graph(c);
    c->solver = solver;
    c->cg_tol = tol;
    c->cg_maxiter = maxiter;
    c->grad_ready = false;   // the next gradient is prepared for this solver
    return NQH_OK;
}

nqh_status nqh_get_sr_solver_info(nqh_ctx* c, int32_t* iters, double* rel_resid) {
    if (!c) return NQH_E_ARG;
    if (iters) *iters = c->cg_iters;
    if (rel_resid) *rel_resid = c->cg_resid;
    return NQH_OK;
}

nqh_status nqh_set_graph(nqh_ctx* c, int32_t on) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    drop_graph(c);
    c->use_graph = on != 0;
    return NQH_OK;
}

nqh_status nqh_set_launch_shape(nqh_ctx* c, int32_t est_samples_per_block, int32_t est_waves, int32_t gram_splits,
                                int32_t sampler_waves) {
    if (!c) return NQH_E_ARG;
    if (est_samples_per_block < 0 || est_samples_per_block > 4096)
        return fail(c, NQH_E_ARG, "estimator samples per workgroup in 0..4096");
    if (est_waves != 0 && est_waves != 4 && est_waves != 8) return fail(c, NQH_E_ARG, "estimator waves 0, 4 or 8");
    if (gram_splits < 0 || gram_splits > 64) return fail(c, NQH_E_ARG, "Gram split-K chunks in 0..64");
    if (sampler_waves != 0 && sampler_waves != 1 && sampler_waves != 2 && sampler_waves != 4)
        return fail(c, NQH_E_ARG, "sampler waves 0, 1, 2 or 4");
    (void)hipSetDevice(c->device);
    drop_graph(c);
    c->tune_spb = est_samples_per_block;
    c->tune_est_waves = est_waves;
    c->tune_splits = gram_splits;
    c->tune_samp_waves = sampler_waves;
    c->ns_cap = 0;   // per-sample buffers (force partials, Gram partials) resized on the next call
    c->gen++;        // batches capture their graphs again
    return NQH_OK;
}

nqh_status nqh_get_launch_shape(const nqh_ctx* c, int32_t* out4) {
    if (!c || !out4) return NQH_E_ARG;
    out4[0] = c->tune_spb;
    out4[1] = c->tune_est_waves;
    out4[2] = c->tune_splits;
    out4[3] = c->tune_samp_waves;
    return NQH_OK;
}

// enqueue one SR iteration on the context's stream without waiting for it
static nqh_status sr_enqueue_one(nqh_ctx* c, double eps, int32_t shift_kind, int32_t refine, double lr);
// the record's done word is cleared before the only iteration in flight is
// enqueued (every iteration ends with the kernel that writes the record)
static nqh_status sr_enqueue(nqh_ctx* c, double eps, int32_t shift_kind, int32_t refine, double lr) {
    if (c->inflight == 0) reinterpret_cast<volatile double*>(c->hstat)[7] = 0.0;
    const nqh_status s = sr_enqueue_one(c, eps, shift_kind, refine, lr);
    if (!s) ++c->inflight;
    return s;
}

static nqh_status sr_enqueue_one(nqh_ctx* c, double eps, int32_t shift_kind, int32_t refine, double lr) {
    if (nqh_status sm = made_maybe_retrain(c)) return sm;
    if (c->use_graph && !c->profiling && c->solver == NQH_SOLVER_CHOLESKY) {
        nqh_status s0 = need_net(c);
        if (!s0) s0 = need_problem(c);
        if (!s0 && !c->has_sampler) s0 = fail(c, NQH_E_STATE, "sampler not set (nqh_set_sampler)");
        if (!s0 && (shift_kind != NQH_SHIFT_ADD && shift_kind != NQH_SHIFT_DIAG)) s0 = fail(c, NQH_E_ARG, "bad shift kind");
        if (!s0 && (!(eps >= 0.0) || refine < 0 || refine > 8)) s0 = fail(c, NQH_E_ARG, "eps >= 0, refine in 0..8");
        if (s0) return s0;
        return graph_iteration(c, eps, shift_kind, refine, lr);
    }
    return enqueue_iteration(c, eps, shift_kind, refine, lr);
}

// wait for the enqueued iteration; statistics and solve checks
// (the update kernel has already skipped the step when the loss, the
// factorization or the step failed: the parameters are then unchanged)
// an iteration's outcome from its statistics and the solve's status words
static nqh_status finish_status(nqh_ctx* c, const double* st, int info, int flag, double* stats) {
    c->solve_ok = false;
    if (stats) std::memcpy(stats, st, 5 * sizeof(double));
    if (!std::isfinite(st[0])) return fail(c, NQH_E_NAN, "non-finite loss <L^dag L>: no step applied");
    if (info != 0) {
        char buf[128];
        std::snprintf(buf, sizeof(buf), "S + shift is not positive definite (column %d): no step applied", info);
        return fail(c, NQH_E_NOTPD, buf);
    }
    if (flag) return fail(c, NQH_E_NAN, "non-finite SR update: no step applied");
    c->solve_ok = true;
    return NQH_OK;
}

static nqh_status sr_finish(nqh_ctx* c, double* stats) {
    // the iteration's last kernel writes its outcome into the pinned record
    // and then the done word: with one iteration in flight and no phase
    // events to read, a short spin on that word replaces the runtime's wait;
    // a slow or failed iteration (no done word within 200 ms) falls back to
    // the stream synchronisation, which also reports errors
    const volatile double* h = c->hstat;
    bool seen = false;
    if (c->inflight == 1 && !c->profiling) {
        const auto t0 = std::chrono::steady_clock::now();
        while (!(seen = h[7] != 0.0)) {
            if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) break;
            __builtin_ia32_pause();
        }
    }
    c->inflight = 0;
    if (!seen) HIPCHK(c, hipStreamSynchronize(c->stream));
    phase_collect(c);
    double st[5];
    for (int k = 0; k < 5; ++k) st[k] = h[k];
    const int info = (int)h[5], flag = (int)h[6];
    return finish_status(c, st, info, flag, stats);
}

nqh_status nqh_sr_iteration(nqh_ctx* c, double eps, int32_t shift_kind, int32_t refine, double lr, double* stats) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s = sr_enqueue(c, eps, shift_kind, refine, lr);
    if (s) return s;
    return sr_finish(c, stats);
}

nqh_status nqh_sr_iteration_async(nqh_ctx* c, double eps, int32_t shift_kind, int32_t refine, double lr) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    return sr_enqueue(c, eps, shift_kind, refine, lr);
}

nqh_status nqh_wait(nqh_ctx* c, double* stats) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    return sr_finish(c, stats);
}

// ---- batches of independent problems: one graph launch per SR iteration
struct nqh_batch {
    int device = 0;
    std::vector<nqh_ctx*> m;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join, pre;
    hipEvent_t post = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool valid = false;
    double eps = 0.0, lr = 0.0;
    int kind = -1, refine = -1;
    std::vector<uint64_t> gens;   // members' configuration generations at capture
    std::string err;
    int mode = NQH_BATCH_PACKED;
    bool packed = false;          // the current graph is the packed form
    void* pk_dev = nullptr;       // its member descriptors and block records
    double* pk_status = nullptr;  // pinned host memory: 8 doubles per member (statistics, info, flag)
    std::vector<double> pk_host;
    int pk_launches = 0;
};

namespace {

nqh_status bfail(nqh_batch* b, nqh_status s, const std::string& msg) {
    b->err = msg;
    return s;
}

nqh_status bmember(nqh_batch* b, size_t i, nqh_status s) {
    return bfail(b, s, "member " + std::to_string(i) + ": " + b->m[i]->err);
}

void batch_drop_graph(nqh_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->valid = false;
    if (b->pk_dev) {
        // no launch of the packed graph is in flight: every batch iteration
        // ends with each member's statistics read back
        (void)hipFree(b->pk_dev);
        b->pk_dev = nullptr;
    }
    if (b->pk_status) {
        (void)hipHostFree(b->pk_status);
        b->pk_status = nullptr;
    }
    b->packed = false;
    b->pk_launches = 0;
}

#define BHIPCHK(b, expr)                                                                           \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess) return bfail(b, NQH_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// capture every member's iteration, each on its own stream forked from and
// joined back into the batch stream, as one graph
nqh_status batch_build(nqh_batch* b, double eps, int kind, int refine, double lr) {
    batch_drop_graph(b);
    const size_t n = b->m.size();
    std::vector<uint32_t> iter0(n);
    std::vector<char> prof(n);
    for (size_t i = 0; i < n; ++i) {
        nqh_ctx* c = b->m[i];
        nqh_status s;
        if ((s = need_net(c)) || (s = need_problem(c))) return bmember(b, i, s);
        if (!c->has_sampler) return bmember(b, i, fail(c, NQH_E_STATE, "sampler not set (nqh_set_sampler)"));
        if ((s = ensure_samples(c, c->chains * c->chain_len))) return bmember(b, i, s);
        if ((s = ensure(c, c->iter_dev, sizeof(uint32_t) * 4))) return bmember(b, i, s);
        BHIPCHK(b, hipMemcpyAsync(c->iter_dev.p, &c->iter, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
        BHIPCHK(b, hipStreamSynchronize(c->stream));
        iter0[i] = c->iter;
        prof[i] = c->profiling;
        c->profiling = false;
    }
    nqh_status s = NQH_OK;
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bfail(b, NQH_E_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    e = hipEventRecord(b->fork, b->stream);
    for (size_t i = 0; i < n && e == hipSuccess && !s; ++i) {
        nqh_ctx* c = b->m[i];
        e = hipStreamWaitEvent(c->stream, b->fork, 0);
        if (e == hipSuccess) {
            s = enqueue_iteration(c, eps, kind, refine, lr);
            if (s) s = bmember(b, i, s);
        }
        if (e == hipSuccess && !s) e = nq_launch_bump(P<uint32_t>(c->iter_dev), c->stream);
        if (e == hipSuccess && !s) e = hipEventRecord(b->join[i], c->stream);
        if (e == hipSuccess && !s) e = hipStreamWaitEvent(b->stream, b->join[i], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    for (size_t i = 0; i < n; ++i) {
        b->m[i]->iter = iter0[i];      // the capture ran nothing
        b->m[i]->profiling = prof[i];
    }
    if (s || e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (s) return s;
        return bfail(b, NQH_E_HIP, std::string("batch capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    b->graph = g;
    BHIPCHK(b, hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->valid = true;
    b->eps = eps; b->kind = kind; b->refine = refine; b->lr = lr;
    b->gens.resize(n);
    for (size_t i = 0; i < n; ++i) b->gens[i] = b->m[i]->gen;
    return NQH_OK;
}

// ---- the packed form: every kernel of the iteration launched once for all
// members.  A member qualifies when its iteration is the NDM's split Cholesky
// solve without refinement, with the BF16 S build, one hidden and one ancilla
// row per lane (nh, na <= 32) and no learned proposal; the descriptors hold
// exactly the arguments enqueue_iteration passes in graph form.
bool pk_qualifies(nqh_ctx* c, int refine) {
    if (is_rbm(c) || c->solver != NQH_SOLVER_CHOLESKY || refine != 0 || !use_split(c)) return false;
    if (c->made_ready && c->made_mix > 0.0) return false;
    if (c->d.nh > 32 || c->d.na > 32) return false;
    return true;
}

nq_pk_member pk_describe(nqh_ctx* c, double eps, int kind, double lr) {
    nq_pk_member m{};
    const ndm_dims& d = c->d;
    const int64_t Ns = c->chains * c->chain_len;
    const bool f64 = fp64_net(c);
    // the sampler (run_sampler in graph form: iteration number from iter_dev)
    samp_args& a = m.sa;
    a.d = d;
    a.p = f64 ? c->p64.p : c->p32.p;
    a.chains = c->chains;
    a.passes = c->passes;
    const bool compat = (c->flags & NQH_SAMPLER_COMPAT_STALE_LAST_SLICE) != 0;
    const bool persistent = (c->flags & NQH_SAMPLER_PERSISTENT) != 0;
    a.restart = (!persistent || !c->chains_initialised) ? 1 : 0;
    a.burn = a.restart ? c->burn : 0;
    a.stored = compat ? c->chain_len - 1 : c->chain_len;
    a.compat_stale_last_slice = compat ? 1 : 0;
    a.lpc = nq_sampler_lanes_per_chain(d);
    a.seed = c->seed;
    a.iter = c->iter;
    a.iter_dev = P<uint32_t>(c->iter_dev);
    a.chain_state = P<uint64_t>(c->chain_state);
    a.samples = P<uint64_t>(c->samples);
    a.accepted = P<unsigned long long>(c->acc);
    a.chain_acc = P<uint32_t>(c->chain_acc);
    a.waves = c->tune_samp_waves;
    // the estimator's two launches (run_grad_pipeline)
    est_args e = make_est(c, P<uint64_t>(c->samples), Ns);
    e.do_conn = 1;
    e.Lloc = c->Lloc.p;
    e.crw = c->crw.p;
    e.crm = P<uint2>(c->crm);
    e.cnum = P<int>(c->cnum);
    e.cwd = c->cwd.p;
    m.ew = e;
    m.ew.rec_mode = 1;
    m.ew.lp = c->lp.p;
    m.ew.Y = c->Y.p;
    m.ef = e;
    m.ef.rec_mode = 2;
    m.ef.tab_lds = 0;   // the records replace the walk
    m.ef.Fpart = P<double>(c->Fpart);
    m.est_grid = c->est_grid;
    m.force_grid = force_grid_of(c, Ns);
    m.fp64 = f64 ? 1 : 0;
    m.Y = c->Y.p;
    m.Lloc = c->Lloc.p;
    m.Ns = Ns;
    m.chains = c->chains;
    m.Lc = c->chain_len;
    m.R = d.R;
    m.np = d.np;
    m.colpart = P<double>(c->colpart);
    m.col_chunks = col_chunks_for(Ns);
    m.mu = P<double>(c->mu);
    m.stats = P<double>(c->stats);
    m.Fpart = P<double>(c->Fpart);
    m.fpart2 = P<double>(c->fpart2);
    m.F = P<double>(c->F);
    // the S build (build_s): BF16 split-K, or FP64 / FP32 Gram blocks
    m.s_mode = c->s_prec == NQH_PREC_BF16 ? -1 : (c->s_prec == NQH_PREC_FP64 ? NQ_GRAM_F64 : NQ_GRAM_F32);
    m.nsp = (Ns + 63) / 64 * 64;
    m.Yt = reinterpret_cast<__bf16*>(c->Yt.p);
    gram_blocks(c, &m.nA, &m.nB, &m.offB);
    nq_gram_shape(m.nA, m.nB, m.nsp, c->tune_splits, &m.TA, &m.tiles, &m.splits, &m.kchunk);
    m.gpart = P<float>(c->gpart);
    m.S = P<double>(c->S);
    m.ldS = d.np + 1;
    // the split solve (precondition_async) and the update
    m.A2 = P<double>(c->A2);
    m.winv = P<double>(c->winv2);
    m.info = P<int>(c->info);
    m.flag = P<int>(c->flag);
    m.nl = d.nlam;
    m.npi = d.npi;
    m.eps = eps;
    m.kind = kind;
    m.y2 = P<double>(c->y2);
    m.x2 = P<double>(c->x2);
    m.delta = P<double>(c->delta);
    m.p64 = P<double>(c->p64);
    m.p32 = P<float>(c->p32);
    m.lr = lr;
    m.iter_dev = P<uint32_t>(c->iter_dev);
    m.acc = P<unsigned long long>(c->acc);
    m.chain_acc = P<uint32_t>(c->chain_acc);
    return m;
}

// the launches of one packed iteration, in the order enqueue_iteration runs a member's
std::vector<std::pair<int, int>> pk_sequence(const std::vector<nq_pk_member>& M) {
    int pa = 0, pb = 0, ga = 0, gb = 0;
    for (const nq_pk_member& m : M) {
        pa = std::max(pa, nq_pk_split_panels(m.nl, m.npi, 0));
        pb = std::max(pb, nq_pk_split_panels(m.nl, m.npi, 1));
        ga = std::max(ga, nq_pk_split_groups(m.nl, m.npi, 0));
        gb = std::max(gb, nq_pk_split_groups(m.nl, m.npi, 1));
    }
    std::vector<std::pair<int, int>> q;
    for (int k : {NQ_PK_PREP, NQ_PK_METROPOLIS, NQ_PK_EST_WALK, NQ_PK_EST_FORCE, NQ_PK_COLSUM, NQ_PK_COLSUM_FIN,
                  NQ_PK_STATS, NQ_PK_FORCE_PART, NQ_PK_FORCE_FIN, NQ_PK_CENTER, NQ_PK_GRAM_PART, NQ_PK_GRAM_FIN,
                  NQ_PK_SFILL})
        q.push_back({k, 0});
    q.push_back({NQ_PK_GRAM, 0});
    q.push_back({NQ_PK_GRAM, 1});
    q.push_back({NQ_PK_SPLIT_COPY, 0});
    for (int j = 0; j < pa; ++j) q.push_back({NQ_PK_CHOL_STEP, j});
    q.push_back({NQ_PK_SCHUR, 0});
    for (int j = 0; j < pb; ++j) q.push_back({NQ_PK_CHOL_STEP, (1 << 16) | j});
    q.push_back({NQ_PK_COPY_COLS, 0});
    for (int j = 0; j < ga; ++j) q.push_back({NQ_PK_BACK, j});
    q.push_back({NQ_PK_SPLIT_UPD, 0});
    q.push_back({NQ_PK_UPD_FIN, 0});
    for (int j = 0; j < gb; ++j) q.push_back({NQ_PK_BACK, (1 << 16) | j});
    for (int k : {NQ_PK_GATHER, NQ_PK_CHECK, NQ_PK_UPDATE, NQ_PK_BUMP, NQ_PK_COLLECT}) q.push_back({k, 0});
    return q;
}

nqh_status batch_build_packed(nqh_batch* b, double eps, int kind, double lr) {
    batch_drop_graph(b);
    const size_t n = b->m.size();
    if (n > 0xffff) return bfail(b, NQH_E_ARG, "at most 65535 members in a packed batch");
    std::vector<nq_pk_member> M(n);
    for (size_t i = 0; i < n; ++i) {
        nqh_ctx* c = b->m[i];
        nqh_status s;
        if ((s = need_net(c)) || (s = need_problem(c))) return bmember(b, i, s);
        if (!c->has_sampler) return bmember(b, i, fail(c, NQH_E_STATE, "sampler not set (nqh_set_sampler)"));
        if ((s = ensure_samples(c, c->chains * c->chain_len))) return bmember(b, i, s);
        if ((s = ensure(c, c->iter_dev, sizeof(uint32_t) * 4))) return bmember(b, i, s);
        BHIPCHK(b, hipMemcpyAsync(c->iter_dev.p, &c->iter, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
        BHIPCHK(b, hipStreamSynchronize(c->stream));
        M[i] = pk_describe(c, eps, kind, lr);
    }
    // block records per launch: members with more parameters first (their
    // blocks run longest), instances / z, then y, then x, as in their own grids
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return M[x].np > M[y].np; });
    struct plan { nq_pk_launch L; size_t off; };
    std::vector<plan> launches;
    std::vector<uint2> rec;
    for (const auto& kp : pk_sequence(M)) {
        // members grouped by instantiation and block shape: one launch each
        std::vector<std::pair<nq_pk_shape, size_t>> sh;
        for (size_t i : order) {
            nq_pk_shape o;
            const hipError_t e = nq_pk_plan(M[i], kp.first, kp.second, &o);
            if (e != hipSuccess)
                return bfail(b, NQH_E_STATE, "member " + std::to_string(i) + ": no packed form of launch " +
                                                 std::to_string(kp.first));
            if ((int64_t)o.gx * o.gy * o.gz > 0) sh.push_back({o, i});
        }
        std::vector<char> done(sh.size(), 0);
        for (size_t a = 0; a < sh.size(); ++a) {
            if (done[a]) continue;
            plan p{};
            p.L.kernel = kp.first;
            p.L.param = kp.second;
            p.L.variant = sh[a].first.variant;
            p.L.threads = sh[a].first.threads;
            p.L.lds = 0;
            p.off = rec.size();
            for (size_t q = a; q < sh.size(); ++q) {
                const nq_pk_shape& o = sh[q].first;
                if (done[q] || o.variant != p.L.variant || o.threads != p.L.threads) continue;
                done[q] = 1;
                p.L.lds = std::max(p.L.lds, o.lds);
                // a Cholesky step's instances interleaved (both look-ahead
                // tiles among the first workgroups dispatched)
                const bool inter = kp.first == NQ_PK_CHOL_STEP;
                for (int z = 0; z < o.gz; ++z)
                    for (int u = 0; u < o.gx * o.gy; ++u) {
                        const int x = inter ? u / o.gy : u % o.gx, y = inter ? u % o.gy : u / o.gx;
                        rec.push_back(make_uint2((uint32_t)sh[q].second | ((uint32_t)z << 16),
                                                 (uint32_t)x | ((uint32_t)y << 16)));
                    }
            }
            p.L.nblk = (int)(rec.size() - p.off);
            launches.push_back(p);
        }
    }
    // device layout: descriptors | block records; the status slots (8
    // doubles per member, written by the last launch) in pinned host memory,
    // so that the batch iteration ends with one synchronisation and no copy
    const size_t mbytes = (sizeof(nq_pk_member) * n + 255) / 256 * 256;
    const size_t bytes = mbytes + sizeof(uint2) * rec.size();
    BHIPCHK(b, hipHostMalloc((void**)&b->pk_status, sizeof(double) * 8 * n, hipHostMallocDefault));
    BHIPCHK(b, hipMalloc(&b->pk_dev, bytes));
    for (size_t i = 0; i < n; ++i) M[i].status = b->pk_status + 8 * i;
    unsigned char* base = static_cast<unsigned char*>(b->pk_dev);
    const nq_pk_member* Md = reinterpret_cast<const nq_pk_member*>(base);
    const uint2* Rd = reinterpret_cast<const uint2*>(base + mbytes);
    BHIPCHK(b, hipMemcpyAsync(base, M.data(), sizeof(nq_pk_member) * n, hipMemcpyHostToDevice, b->stream));
    BHIPCHK(b, hipMemcpyAsync(base + mbytes, rec.data(), sizeof(uint2) * rec.size(), hipMemcpyHostToDevice, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bfail(b, NQH_E_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    for (size_t k = 0; k < launches.size() && e == hipSuccess; ++k) {
        nq_pk_launch L = launches[k].L;
        L.blk = Rd + launches[k].off;
        e = nq_pk_run(Md, L, b->stream);
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return bfail(b, NQH_E_HIP, std::string("packed batch capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    b->graph = g;
    BHIPCHK(b, hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->valid = true;
    b->packed = true;
    b->pk_launches = (int)launches.size();
    b->eps = eps; b->kind = kind; b->refine = 0; b->lr = lr;
    b->gens.resize(n);
    for (size_t i = 0; i < n; ++i) b->gens[i] = b->m[i]->gen;
    return NQH_OK;
}

// a member was reconfigured, reallocated or given a new iteration number
// since the capture: the graph holds stale buffers, shapes or counters
bool batch_stale(const nqh_batch* b) {
    if (b->gens.size() != b->m.size()) return true;
    for (size_t i = 0; i < b->m.size(); ++i)
        if (b->m[i]->gen != b->gens[i]) return true;
    return false;
}

}  // namespace

nqh_status nqh_batch_create(nqh_batch** out, nqh_ctx* const* members, int32_t n) {
    if (!out || !members || n < 1) return NQH_E_ARG;
    *out = nullptr;
    std::vector<nqh_ctx*> m(members, members + n);
    for (int i = 0; i < n; ++i) {
        if (!m[i] || m[i]->device != m[0]->device) return NQH_E_ARG;
        for (int j = 0; j < i; ++j)
            if (m[j] == m[i]) return NQH_E_ARG;
    }
    if (hipSetDevice(m[0]->device) != hipSuccess) return NQH_E_HIP;
    nqh_batch* b = new nqh_batch();
    b->device = m[0]->device;
    b->m = m;
    b->join.assign(n, nullptr);
    b->pre.assign(n, nullptr);
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&b->post, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < n && ok; ++i) {
        ok = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&b->pre[i], hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
        nqh_batch_destroy(b);
        return NQH_E_HIP;
    }
    for (nqh_ctx* c : m) {
        // members read the iteration number from device memory, as in their own graphs
        drop_graph(c);
        c->use_graph = true;
    }
    *out = b;
    return NQH_OK;
}

void nqh_batch_destroy(nqh_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->pre)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->post) (void)hipEventDestroy(b->post);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* nqh_batch_last_error(const nqh_batch* b) { return b ? b->err.c_str() : "null batch"; }

nqh_status nqh_batch_set_mode(nqh_batch* b, int32_t mode) {
    if (!b) return NQH_E_ARG;
    if (mode != NQH_BATCH_PACKED && mode != NQH_BATCH_GRAPH) return bfail(b, NQH_E_ARG, "unknown batch mode");
    if (mode != b->mode) {
        (void)hipSetDevice(b->device);
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        batch_drop_graph(b);
        b->mode = mode;
    }
    return NQH_OK;
}

int32_t nqh_batch_launches(const nqh_batch* b) {
    if (!b || !b->valid) return -1;
    return b->packed ? b->pk_launches : 0;
}

nqh_status nqh_batch_sr_iteration(nqh_batch* b, double eps, int32_t kind, int32_t refine, double lr, double* stats) {
    if (!b) return NQH_E_ARG;
    (void)hipSetDevice(b->device);
    if (kind != NQH_SHIFT_ADD && kind != NQH_SHIFT_DIAG) return bfail(b, NQH_E_ARG, "bad shift kind");
    if (!(eps >= 0.0) || refine < 0 || refine > 8) return bfail(b, NQH_E_ARG, "eps >= 0, refine in 0..8");
    const size_t n = b->m.size();
    bool eager = false;
    for (size_t i = 0; i < n; ++i) {
        nqh_ctx* c = b->m[i];
        if (c->made_ready && c->made_mix > 0.0)
            return bfail(b, NQH_E_STATE, "member " + std::to_string(i) + ": learned proposals are not batched");
        if (c->solver != NQH_SOLVER_CHOLESKY)
            return bfail(b, NQH_E_STATE, "member " + std::to_string(i) + ": the iterative solver is not batched");
        if ((c->flags & NQH_SAMPLER_PERSISTENT) && !c->chains_initialised) eager = true;
    }
    if (eager) {
        // the restart + burn-in iteration runs eagerly, members concurrently on their streams
        batch_drop_graph(b);
        for (size_t i = 0; i < n; ++i) {
            nqh_ctx* c = b->m[i];
            c->use_graph = false;    // iteration number from the host
            const nqh_status s = enqueue_iteration(c, eps, kind, refine, lr);
            c->use_graph = true;
            if (s) return bmember(b, i, s);
        }
    } else {
        bool pk = b->mode == NQH_BATCH_PACKED;
        for (size_t i = 0; i < n && pk; ++i) pk = pk_qualifies(b->m[i], refine);
        if (!b->valid || batch_stale(b) || b->packed != pk || b->eps != eps || b->kind != kind ||
            b->refine != refine || b->lr != lr) {
            nqh_status s = pk ? batch_build_packed(b, eps, kind, lr) : batch_build(b, eps, kind, refine, lr);
            if (s) return s;
        }
        // work already queued on a member's stream precedes the batch's graph,
        // and the member's later work follows it
        for (size_t i = 0; i < n; ++i) {
            BHIPCHK(b, hipEventRecord(b->pre[i], b->m[i]->stream));
            BHIPCHK(b, hipStreamWaitEvent(b->stream, b->pre[i], 0));
        }
        BHIPCHK(b, hipGraphLaunch(b->gexec, b->stream));
        BHIPCHK(b, hipEventRecord(b->post, b->stream));
        for (size_t i = 0; i < n; ++i) {
            nqh_ctx* c = b->m[i];
            BHIPCHK(b, hipStreamWaitEvent(c->stream, b->post, 0));
            c->iter += 1;
            c->chains_initialised = true;
            c->grad_ready = true;
            c->solve_ok = false;
            c->last_ns = c->chains * c->chain_len;
        }
    }
    nqh_status first = NQH_OK;
    if (!eager && b->packed) {
        // every member's statistics and status words, from the pinned slots
        b->pk_host.resize(8 * n);
        BHIPCHK(b, hipStreamSynchronize(b->stream));
        const volatile double* h = b->pk_status;
        for (size_t k = 0; k < 8 * n; ++k) b->pk_host[k] = h[k];
        for (size_t i = 0; i < n; ++i) {
            const double* q = b->pk_host.data() + 8 * i;
            nqh_status s = finish_status(b->m[i], q, (int)q[5], (int)q[6], stats ? stats + 5 * i : nullptr);
            if (s && !first) first = bmember(b, i, s);
        }
        return first;
    }
    for (size_t i = 0; i < n; ++i) {
        nqh_status s = sr_finish(b->m[i], stats ? stats + 5 * i : nullptr);
        if (s && !first) first = bmember(b, i, s);
    }
    return first;
}

nqh_status nqh_get_samples(nqh_ctx* c, uint64_t* keys_out, int64_t n) {
    if (!c || !keys_out) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    if (!c->has_sampler) return fail(c, NQH_E_STATE, "sampler not set");
    if (n != c->chains * c->chain_len) return fail(c, NQH_E_ARG, "n must be chains * chain_length");
    HIPCHK(c, copy_sync(c->stream, keys_out, c->samples.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    return NQH_OK;
}

nqh_status nqh_set_profiling(nqh_ctx* c, int32_t on) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    if (on && !c->ev[0][0]) {
        for (int ph = 0; ph < NQH_NUM_PHASES; ++ph)
            for (int k = 0; k < 2; ++k) HIPCHK(c, hipEventCreate(&c->ev[ph][k]));
    }
    c->profiling = on != 0;
    for (int ph = 0; ph < NQH_NUM_PHASES; ++ph) {
        c->phase_ms[ph] = 0.0;
        c->phase_n[ph] = 0;
        c->pending[ph] = false;
    }
    return NQH_OK;
}

nqh_status nqh_get_phase_times(nqh_ctx* c, double* ms_out, int64_t* counts_out) {
    if (!c || !ms_out) return NQH_E_ARG;
    for (int ph = 0; ph < NQH_NUM_PHASES; ++ph) {
        ms_out[ph] = c->phase_ms[ph];
        if (counts_out) counts_out[ph] = c->phase_n[ph];
    }
    return NQH_OK;
}

// exact sums over every configuration, in chunks of at most 2^20 keys
nqh_status nqh_exact(nqh_ctx* c, double* out, double* logp_out) {
    if (!c || !out) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_net(c)) || (s = need_problem(c))) return s;
    const int bits = is_ket(c) ? net_N(c) : 2 * net_N(c);
    if (bits > 30) return fail(c, NQH_E_ARG, "the exact sum is limited to 2^30 configurations");
    const int64_t total = (int64_t)1 << bits;
    const int64_t chunk = std::min<int64_t>(total, (int64_t)1 << 20);
    const int64_t nch = (total + chunk - 1) / chunk;
    const size_t e = net_elt(c);
    dbuf kb, lpb, Lb, pb;
    auto cleanup = [&]() { release(kb); release(lpb); release(Lb); release(pb); };
    if ((s = ensure(c, kb, sizeof(uint64_t) * chunk)) || (s = ensure(c, lpb, 2 * e * chunk)) ||
        (s = ensure(c, Lb, 2 * e * chunk)) || (s = ensure(c, pb, sizeof(double) * 5 * nch))) {
        cleanup();
        return s;
    }
    std::vector<float> lpf;
    std::vector<double> lpd;
    hipError_t err = hipSuccess;
    for (int64_t ci = 0; ci < nch && err == hipSuccess; ++ci) {
        const int64_t start = ci * chunk, n = std::min(chunk, total - start);
        err = nq_launch_iota_keys(P<uint64_t>(kb), n, (uint64_t)start, c->stream);
        if (err != hipSuccess) break;
        if (is_rbm(c)) {
            rbm_args a = make_rbm(c, P<uint64_t>(kb), n);
            a.do_conn = 1;
            a.lp = lpb.p;
            a.Lloc = Lb.p;
            err = nq_launch_rbm_estimate(a, fp64_net(c), est_grid_for(n), c->stream);
        } else {
            est_args a = make_est(c, P<uint64_t>(kb), n);
            a.do_conn = 1;
            a.lp = lpb.p;
            a.Lloc = Lb.p;
            err = nq_launch_estimate(a, fp64_net(c), est_grid_for(n), c->stream);
        }
        if (err == hipSuccess)
            err = nq_launch_exact_partial(lpb.p, Lb.p, fp64_net(c), n, P<double>(pb) + 5 * ci, c->stream);
        if (err == hipSuccess && logp_out) {
            if (fp64_net(c)) {
                lpd.resize(2 * n);
                err = hipMemcpyAsync(lpd.data(), lpb.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->stream);
                if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
                for (int64_t i = 0; i < n; ++i) logp_out[start + i] = 2.0 * lpd[2 * i];
            } else {
                lpf.resize(2 * n);
                err = hipMemcpyAsync(lpf.data(), lpb.p, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, c->stream);
                if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
                for (int64_t i = 0; i < n; ++i) logp_out[start + i] = 2.0 * (double)lpf[2 * i];
            }
        }
    }
    std::vector<double> part(5 * nch);
    if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
    if (err == hipSuccess) err = copy_sync(c->stream, part.data(), pb.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost);
    cleanup();
    if (err != hipSuccess) return hip_fail(c, err, "nqh_exact");
    double M = -INFINITY;
    for (int64_t ci = 0; ci < nch; ++ci) M = std::max(M, part[5 * ci]);
    double Z = 0.0, ZC = 0.0, ZR = 0.0, ZI = 0.0;
    for (int64_t ci = 0; ci < nch; ++ci) {
        const double f = std::exp(part[5 * ci] - M);
        Z += f * part[5 * ci + 1];
        ZC += f * part[5 * ci + 2];
        ZR += f * part[5 * ci + 3];
        ZI += f * part[5 * ci + 4];
    }
    out[0] = ZC / Z;
    out[1] = ZR / Z;
    out[2] = ZI / Z;
    out[3] = M + std::log(Z);
    if (logp_out)
        for (int64_t i = 0; i < total; ++i) logp_out[i] -= out[3];
    if (!std::isfinite(out[0])) return fail(c, NQH_E_NAN, "non-finite exact sum");
    return NQH_OK;
}

// ---- checkpoint of the state the reference does not keep (it restarts its
// chains every iteration, Metropolis.jl:106, and has no learned proposal):
// a little-endian blob
//   header: int64 [magic "NQHSTAT1", chains, chains_initialised, iter,
//                  made_P, made_t, made_calls, ring_cap, ring_fill, ring_pos]
//   chains x uint64 chain configurations
//   learned proposal (made_P > 0): made_P x float weights, Adam m, Adam v,
//                  ring_cap x uint64 ring of recent samples
namespace {
constexpr int64_t kStateMagic = 0x315441545348514eLL;   // "NQHSTAT1"
constexpr int kStateHdr = 10;
}

static int64_t state_bytes(const nqh_ctx* c) {
    int64_t b = kStateHdr * 8 + (c->has_sampler ? c->chains * 8 : 0);
    if (c->made_ready) b += 3 * (int64_t)c->made_P * 4 + c->ring_cap * 8;
    return b;
}

int64_t nqh_state_bytes(const nqh_ctx* c) { return c ? state_bytes(c) : -1; }

nqh_status nqh_save_state(nqh_ctx* c, void* buf, int64_t bytes) {
    if (!c || !buf) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    if (!c->has_sampler || !c->chain_state.p) return fail(c, NQH_E_STATE, "sampler not set (nqh_set_sampler)");
    if (bytes != state_bytes(c)) return fail(c, NQH_E_ARG, "buffer size must be nqh_state_bytes(ctx)");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int64_t* h = static_cast<int64_t*>(buf);
    const int64_t hdr[kStateHdr] = {kStateMagic, c->chains, c->chains_initialised ? 1 : 0, (int64_t)c->iter,
                                    c->made_ready ? c->made_P : 0, c->made_t, c->made_calls,
                                    c->made_ready ? c->ring_cap : 0, c->ring_fill, c->ring_pos};
    std::memcpy(h, hdr, sizeof(hdr));
    unsigned char* q = static_cast<unsigned char*>(buf) + sizeof(hdr);
    HIPCHK(c, copy_sync(c->stream, q, c->chain_state.p, sizeof(uint64_t) * c->chains, hipMemcpyDeviceToHost));
    q += sizeof(uint64_t) * c->chains;
    if (c->made_ready) {
        const size_t fb = sizeof(float) * c->made_P;
        HIPCHK(c, copy_sync(c->stream, q, c->made_th.p, fb, hipMemcpyDeviceToHost));
        HIPCHK(c, copy_sync(c->stream, q + fb, c->made_m.p, fb, hipMemcpyDeviceToHost));
        HIPCHK(c, copy_sync(c->stream, q + 2 * fb, c->made_v.p, fb, hipMemcpyDeviceToHost));
        q += 3 * fb;
        if (c->ring_cap > 0) HIPCHK(c, copy_sync(c->stream, q, c->ring.p, sizeof(uint64_t) * c->ring_cap, hipMemcpyDeviceToHost));
    }
    return NQH_OK;
}

nqh_status nqh_load_state(nqh_ctx* c, const void* buf, int64_t bytes) {
    if (!c || !buf) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    if (!c->has_sampler || !c->chain_state.p) return fail(c, NQH_E_STATE, "sampler not set (nqh_set_sampler)");
    int64_t hdr[kStateHdr];
    if (bytes < (int64_t)sizeof(hdr)) return fail(c, NQH_E_ARG, "state buffer too short");
    std::memcpy(hdr, buf, sizeof(hdr));
    if (hdr[0] != kStateMagic) return fail(c, NQH_E_ARG, "not a libnqhip state (bad magic)");
    if (hdr[1] != c->chains) return fail(c, NQH_E_ARG, "the state holds another chain count");
    const int64_t P = hdr[4];
    if (P != (c->made_ready ? c->made_P : 0) || (P > 0 && hdr[7] != c->ring_cap))
        return fail(c, NQH_E_ARG, "the state's learned proposal does not match this context's (nqh_made_setup)");
    if (bytes != state_bytes(c)) return fail(c, NQH_E_ARG, "buffer size must be nqh_state_bytes(ctx)");
    if (hdr[8] < 0 || hdr[8] > hdr[7] || hdr[9] < 0 || (hdr[7] > 0 && hdr[9] >= hdr[7]))
        return fail(c, NQH_E_ARG, "inconsistent ring position in the state");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graph(c);
    const unsigned char* q = static_cast<const unsigned char*>(buf) + sizeof(hdr);
    HIPCHK(c, copy_sync(c->stream, c->chain_state.p, q, sizeof(uint64_t) * c->chains, hipMemcpyHostToDevice));
    q += sizeof(uint64_t) * c->chains;
    c->chains_initialised = hdr[2] != 0;
    c->iter = (uint32_t)hdr[3];
    if (P > 0) {
        const size_t fb = sizeof(float) * P;
        HIPCHK(c, copy_sync(c->stream, c->made_th.p, q, fb, hipMemcpyHostToDevice));
        HIPCHK(c, copy_sync(c->stream, c->made_m.p, q + fb, fb, hipMemcpyHostToDevice));
        HIPCHK(c, copy_sync(c->stream, c->made_v.p, q + 2 * fb, fb, hipMemcpyHostToDevice));
        q += 3 * fb;
        c->made_t = (int)hdr[5];
        c->made_calls = hdr[6];
        c->ring_fill = hdr[8];
        c->ring_pos = hdr[9];
        if (c->ring_cap > 0) {
            HIPCHK(c, copy_sync(c->stream, c->ring.p, q, sizeof(uint64_t) * c->ring_cap, hipMemcpyHostToDevice));
            const int64_t pos[2] = {c->ring_pos, c->ring_fill};
            HIPCHK(c, copy_sync(c->stream, c->ring_posd.p, pos, sizeof(pos), hipMemcpyHostToDevice));
        }
    }
    return NQH_OK;
}

// ---- exact sums without device allocation.  Every buffer the kernels use
// beyond the caller's outputs (the keys of a chunk, its log rho and L_loc,
// its log p, the chunks' partial sums and the combined result) lives in one
// file-scope device array, `scratch`; each GPU holds its own copy at its own
// address, taken once per device with hipGetSymbolAddress and kept in
// g_scratch.  Calls on one GPU share that copy: calls that overlap on one GPU
// race.  Nothing in it needs clearing (every region is written before it is
// read within a call).
namespace {
constexpr int64_t kXChunk = (int64_t)1 << 16;            // configurations per chunk
constexpr int64_t kXMaxChunks = (int64_t)1 << 14;        // 2^30 configurations
constexpr size_t kXKeysOff = 0;
constexpr size_t kXLpOff = kXKeysOff + sizeof(uint64_t) * kXChunk;
constexpr size_t kXLOff = kXLpOff + 2 * sizeof(double) * kXChunk;
constexpr size_t kXLogpOff = kXLOff + 2 * sizeof(double) * kXChunk;
constexpr size_t kXPartOff = kXLogpOff + sizeof(double) * kXChunk;
constexpr size_t kXOutOff = kXPartOff + 5 * sizeof(double) * kXMaxChunks;
constexpr size_t kXBytes = kXOutOff + 256;               // 3,801,344 bytes
constexpr int kXMaxDevices = 64;
void* g_scratch[kXMaxDevices] = {};
std::mutex g_scratch_mu;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kXBytes];

// the device address of `scratch` on the context's device (which must be
// current), taken once per device
static nqh_status scratch_setup(nqh_ctx* c) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (c->device < 0 || c->device >= kXMaxDevices) return fail(c, NQH_E_ARG, "device index beyond 63");
    if (g_scratch[c->device]) return NQH_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return hip_fail(c, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch[c->device] = ptr;
    return NQH_OK;
}

nqh_status nqh_exact_noalloc(nqh_ctx* c, double* out, double* logp_out) {
    if (!c || !out) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_net(c)) || (s = need_problem(c))) return s;
    const int bits = is_ket(c) ? net_N(c) : 2 * net_N(c);
    if (bits > 30)
        return fail(c, NQH_E_ARG, "nqh_exact_noalloc: at most 2^30 configurations (2^16 per chunk, 2^14 chunks "
                                  "in its scratch memory)");
    if ((s = scratch_setup(c))) return s;
    unsigned char* x = static_cast<unsigned char*>(g_scratch[c->device]);
    uint64_t* kb = reinterpret_cast<uint64_t*>(x + kXKeysOff);
    void* lpb = x + kXLpOff;
    void* Lb = x + kXLOff;
    double* logp = reinterpret_cast<double*>(x + kXLogpOff);
    double* part = reinterpret_cast<double*>(x + kXPartOff);
    double* res = reinterpret_cast<double*>(x + kXOutOff);
    const int64_t total = (int64_t)1 << bits;
    const int64_t chunk = std::min<int64_t>(total, kXChunk);
    const int64_t nch = (total + chunk - 1) / chunk;
    for (int64_t ci = 0; ci < nch; ++ci) {
        const int64_t start = ci * chunk, n = std::min(chunk, total - start);
        HIPCHK(c, nq_launch_iota_keys(kb, n, (uint64_t)start, c->stream));
        if (is_rbm(c)) {
            rbm_args a = make_rbm(c, kb, n);
            a.do_conn = 1;
            a.lp = lpb;
            a.Lloc = Lb;
            HIPCHK(c, nq_launch_rbm_estimate(a, fp64_net(c), est_grid_for(n), c->stream));
        } else {
            est_args a = make_est(c, kb, n);
            a.do_conn = 1;
            a.lp = lpb;
            a.Lloc = Lb;
            HIPCHK(c, nq_launch_estimate(a, fp64_net(c), est_grid_for(n), c->stream));
        }
        HIPCHK(c, nq_launch_exact_partial(lpb, Lb, fp64_net(c), n, part + 5 * ci, c->stream));
        if (logp_out) {
            HIPCHK(c, nq_launch_logp_of(lpb, fp64_net(c), n, logp, c->stream));
            HIPCHK(c, hipMemcpyAsync(logp_out + start, logp, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));   // the chunk's buffers are reused by the next chunk
        }
    }
    HIPCHK(c, nq_launch_exact_combine(part, nch, res, c->stream));
    double r[4];
    HIPCHK(c, hipMemcpyAsync(r, res, sizeof(r), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 4; ++k) out[k] = r[k];
    if (logp_out)
        for (int64_t i = 0; i < total; ++i) logp_out[i] -= r[3];
    if (!std::isfinite(out[0])) return fail(c, NQH_E_NAN, "non-finite exact sum");
    return NQH_OK;
}

// compute_observables (BatchedObsDMSampler.jl:59-105) for one ket operator
nqh_status nqh_observable(nqh_ctx* c, const nqh_optable* op, int64_t chains, int32_t chain_length, int32_t passes,
                          int32_t burn, double* stats_out) {
    if (!c || !op || !stats_out) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_net(c))) return s;
    if (is_ket(c)) return fail(c, NQH_E_ARG, "observables need a density-matrix network (NDM or RBMSplit)");
    if (chains < 2 || chain_length < 2 || passes < 1 || burn < 0)
        return fail(c, NQH_E_ARG, "need chains >= 2, chain_length >= 2, passes >= 1, burn >= 0");
    if (op->n_sites != net_N(c)) return fail(c, NQH_E_ARG, "operator and network have different numbers of sites");
    for (int q = 0; q < op->n_terms; ++q)
        if (op->term_kind[q] != 3) return fail(c, NQH_E_ARG, "the observable must be a ket operator (kind 3 terms)");
    int maxc = 0, nk = 0;
    if ((s = upload_table(c, op, true, c->obs_table_mem, c->obs_tab, &maxc, &nk))) return s;
    const int64_t Ns = chains * chain_length;
    const size_t e = net_elt(c);
    if ((s = ensure(c, c->obs_keys, sizeof(uint64_t) * Ns)) || (s = ensure(c, c->obs_state, sizeof(uint64_t) * chains)) ||
        (s = ensure(c, c->obs_lp, 2 * e * Ns)) || (s = ensure(c, c->obs_L, 2 * e * Ns)) ||
        (s = ensure(c, c->obs_stats, sizeof(double) * 8)))
        return s;
    // the diagonal chain, restarted at every call (_sample_state!), its own random stream
    samp_args a{};
    a.d = c->d;
    a.p = fp64_net(c) ? c->p64.p : c->p32.p;
    a.chains = chains;
    a.passes = passes;
    a.burn = burn;
    a.stored = chain_length;
    a.restart = 1;
    a.diag = 1;
    a.tag = 1;
    a.lpc = is_rbm(c) ? 1 : nq_sampler_lanes_per_chain(c->d);
    a.seed = c->seed;
    a.iter = c->iter;
    a.chain_state = P<uint64_t>(c->obs_state);
    a.samples = P<uint64_t>(c->obs_keys);
    HIPCHK(c, launch_metropolis(c, a));
    // O_loc(sigma) = sum <sigma|O|sigma'> rho(sigma', sigma) / rho(sigma, sigma): the
    // operator's changes act on the row index of the doubled configuration
    if (is_rbm(c)) {
        rbm_args r = make_rbm(c, P<uint64_t>(c->obs_keys), Ns);
        r.tab = c->obs_tab;
        r.max_conn = std::max(maxc, 1);
        r.tab_lds = c->obs_tab.bytes <= kTableLdsMax ? 1 : 0;
        r.do_conn = 1;
        r.lp = c->obs_lp.p;
        r.Lloc = c->obs_L.p;
        HIPCHK(c, nq_launch_rbm_estimate(r, fp64_net(c), est_grid_for(Ns), c->stream));
    } else {
        est_args r = make_est(c, P<uint64_t>(c->obs_keys), Ns);
        r.tab = c->obs_tab;
        r.max_conn = std::max(maxc, 1);
        r.tab_lds = c->obs_tab.bytes <= kTableLdsMax ? 1 : 0;
        r.do_conn = 1;
        r.lp = c->obs_lp.p;
        r.Lloc = c->obs_L.p;
        HIPCHK(c, nq_launch_estimate(r, fp64_net(c), est_grid_for(Ns), c->stream));
    }
    HIPCHK(c, nq_launch_stats_cplx(c->obs_L.p, fp64_net(c), chains, chain_length, P<double>(c->obs_stats), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, copy_sync(c->stream, stats_out, c->obs_stats.p, sizeof(double) * 6, hipMemcpyDeviceToHost));
    return NQH_OK;
}

// the last observable's configurations (doubled keys, chain_length x chains) and local values
nqh_status nqh_get_observable_samples(nqh_ctx* c, uint64_t* keys_out, double* local_out, int64_t n) {
    if (!c || n < 0) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    if ((size_t)n * sizeof(uint64_t) > c->obs_keys.bytes) return fail(c, NQH_E_ARG, "n exceeds the last observable's samples");
    if (keys_out) HIPCHK(c, copy_sync(c->stream, keys_out, c->obs_keys.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    if (local_out) return download_cplx(c, c->obs_L, local_out, 2 * n);
    return NQH_OK;
}

nqh_status nqh_get_stats(nqh_ctx* c, double* out, int32_t n) {
    if (!c || !out || n < 0 || n > 8) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    if (!c->grad_ready) return fail(c, NQH_E_STATE, "no statistics yet");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, copy_sync(c->stream, out, c->stats.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return NQH_OK;
}

nqh_status nqh_get_local_values(nqh_ctx* c, double* lloc_out, int64_t n) {
    if (!c || !lloc_out) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    if (!c->grad_ready) return fail(c, NQH_E_STATE, "no local values yet");
    if (n != c->last_ns) return fail(c, NQH_E_ARG, "n must equal the number of samples of the last gradient");
    return download_cplx(c, c->Lloc, lloc_out, 2 * n);
}

// ---------------------------------------------------------------------------
// learned proposal

int64_t nqh_made_num_params(int32_t D) { return (D >= 2 && D <= 32) ? nq_made_param_count(D) : -1; }

static nqh_status made_chain_bufs(nqh_ctx* c) {
    if (!c->made_ready || !c->has_sampler) return NQH_OK;
    nqh_status s;
    const int64_t n2 = 2 * c->chains;
    if ((s = ensure(c, c->made_keys, sizeof(uint64_t) * n2))) return s;
    if ((s = ensure(c, c->made_lq, sizeof(float) * n2))) return s;
    if ((s = ensure(c, c->made_lp, 2 * sizeof(double) * n2))) return s;
    return NQH_OK;
}

static nqh_status need_made(nqh_ctx* c) {
    if (!c->made_ready) return fail(c, NQH_E_STATE, "learned proposal not set up (nqh_made_setup)");
    return NQH_OK;
}

// masked Gaussian initialisation, weights N(0, 1/fan_in), biases zero
static void made_init_host(int D, uint64_t seed, std::vector<float>& th) {
    const int H = 128;
    th.assign(nq_made_param_count(D), 0.f);
    uint64_t st = seed ^ 0x9e3779b97f4a7c15ull;
    auto next = [&st]() {   // splitmix64
        uint64_t z = (st += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    auto normal = [&]() {   // Box-Muller
        const double u1 = ((next() >> 11) + 1.0) * (1.0 / 9007199254740993.0);
        const double u2 = (next() >> 11) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    auto deg = [D](int k) { return (k % (D - 1)) + 1; };
    float* W1 = th.data();
    float* W2 = W1 + H * D + H;
    float* W3 = W2 + H * H + H;
    for (int k = 0; k < H; ++k)
        for (int d = 0; d < D; ++d) W1[k * D + d] = deg(k) >= d + 1 ? (float)(normal() / std::sqrt((double)D)) : 0.f;
    for (int k2 = 0; k2 < H; ++k2)
        for (int k = 0; k < H; ++k) W2[k2 * H + k] = deg(k2) >= deg(k) ? (float)(normal() / std::sqrt((double)H)) : 0.f;
    for (int d = 0; d < D; ++d)
        for (int k = 0; k < H; ++k) W3[d * H + k] = d + 1 > deg(k) ? (float)(normal() / std::sqrt((double)H)) : 0.f;
}

nqh_status nqh_made_setup(nqh_ctx* c, double mix, int32_t retrain_every, int32_t train_steps, int32_t batch,
                          double lr, int64_t buffer_cap) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_net(c))) return s;
    if (!(mix >= 0.0 && mix <= 1.0) || retrain_every < 0 || train_steps < 0 || batch < 1 || !(lr > 0.0) ||
        buffer_cap < 0)
        return fail(c, NQH_E_ARG, "need mix in [0,1], retrain_every >= 0, train_steps >= 0, batch >= 1, lr > 0, "
                                  "buffer_cap >= 0");
    const int D = is_ket(c) ? net_N(c) : 2 * net_N(c);
    if (D < 2 || D > 32) return fail(c, NQH_E_ARG, "the learned proposal needs 2 <= 2N <= 32");
    if (buffer_cap > 0 && buffer_cap < batch) return fail(c, NQH_E_ARG, "buffer_cap must be 0 or >= batch");
    drop_graph(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int P = nq_made_param_count(D);
    const size_t parts = nq_made_grad_parts(batch);
    if ((s = ensure(c, c->made_th, sizeof(float) * P))) return s;
    if ((s = ensure(c, c->made_m, sizeof(float) * P))) return s;
    if ((s = ensure(c, c->made_v, sizeof(float) * P))) return s;
    if ((s = ensure(c, c->made_g, sizeof(float) * P))) return s;
    if ((s = ensure(c, c->made_gpart, sizeof(float) * P * parts))) return s;
    if ((s = ensure(c, c->made_lpart, sizeof(float) * parts))) return s;
    if ((s = ensure(c, c->made_nll, sizeof(float) * 4))) return s;
    if ((s = ensure(c, c->ring_posd, sizeof(int64_t) * 2))) return s;
    if (buffer_cap > 0 && (s = ensure(c, c->ring, sizeof(uint64_t) * buffer_cap))) return s;
    std::vector<float> th;
    made_init_host(D, c->seed, th);
    HIPCHK(c, copy_sync(c->stream, c->made_th.p, th.data(), sizeof(float) * P, hipMemcpyHostToDevice));
    HIPCHK(c, zero_sync(c->stream, c->made_m.p, sizeof(float) * P));
    HIPCHK(c, zero_sync(c->stream, c->made_v.p, sizeof(float) * P));
    HIPCHK(c, zero_sync(c->stream, c->ring_posd.p, sizeof(int64_t) * 2));
    HIPCHK(c, nq_made_prepare());
    c->made_D = D;
    c->made_P = P;
    c->made_t = 0;
    c->made_mix = mix;
    c->made_every = retrain_every;
    c->made_steps = train_steps;
    c->made_batch = batch;
    c->made_lr = (float)lr;
    c->ring_cap = buffer_cap;
    c->ring_fill = c->ring_pos = 0;
    c->made_calls = 0;
    c->made_ready = true;
    return made_chain_bufs(c);
}

nqh_status nqh_made_set_params(nqh_ctx* c, const float* theta, int64_t n) {
    if (!c || !theta) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_made(c))) return s;
    if (n != c->made_P) return fail(c, NQH_E_ARG, "parameter count mismatch (nqh_made_num_params)");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, copy_sync(c->stream, c->made_th.p, theta, sizeof(float) * n, hipMemcpyHostToDevice));
    HIPCHK(c, zero_sync(c->stream, c->made_m.p, sizeof(float) * n));
    HIPCHK(c, zero_sync(c->stream, c->made_v.p, sizeof(float) * n));
    c->made_t = 0;
    return NQH_OK;
}

nqh_status nqh_made_get_params(nqh_ctx* c, float* theta, int64_t n) {
    if (!c || !theta) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_made(c))) return s;
    if (n != c->made_P) return fail(c, NQH_E_ARG, "parameter count mismatch (nqh_made_num_params)");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, copy_sync(c->stream, theta, c->made_th.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return NQH_OK;
}

static nqh_status made_keys_to_device(nqh_ctx* c, const uint64_t* keys, int64_t n) {
    nqh_status s;
    if ((s = ensure(c, c->made_tmp, sizeof(uint64_t) * n))) return s;
    HIPCHK(c, hipMemcpyAsync(c->made_tmp.p, keys, sizeof(uint64_t) * n, hipMemcpyHostToDevice, c->stream));
    return NQH_OK;
}

static nqh_status made_check_keys(nqh_ctx* c, const uint64_t* keys, int64_t n) {
    const uint64_t m = (c->made_D >= 64) ? ~0ull : ((1ull << c->made_D) - 1ull);
    for (int64_t i = 0; i < n; ++i)
        if (keys[i] & ~m) return fail(c, NQH_E_ARG, "configuration key has bits above 2N");
    return NQH_OK;
}

nqh_status nqh_made_grad(nqh_ctx* c, const uint64_t* keys, int64_t B, float* grad, double* nll) {
    if (!c || !keys || !grad || B < 1 || B > (int64_t)1 << 30) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_made(c)) || (s = made_check_keys(c, keys, B))) return s;
    const size_t parts = nq_made_grad_parts(B);
    if ((s = ensure(c, c->made_gpart, sizeof(float) * c->made_P * parts))) return s;
    if ((s = ensure(c, c->made_lpart, sizeof(float) * parts))) return s;
    if ((s = made_keys_to_device(c, keys, B))) return s;
    HIPCHK(c, nq_launch_made_grad(P<float>(c->made_th), c->made_D, P<uint64_t>(c->made_tmp), (int)B,
                                  P<float>(c->made_gpart), P<float>(c->made_lpart), c->stream));
    HIPCHK(c, nq_launch_made_sum_parts(P<float>(c->made_gpart), (int)parts, c->made_D, P<float>(c->made_g),
                                       P<float>(c->made_lpart), P<float>(c->made_nll), B, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, copy_sync(c->stream, grad, c->made_g.p, sizeof(float) * c->made_P, hipMemcpyDeviceToHost));
    if (nll) {
        float v = 0.f;
        HIPCHK(c, copy_sync(c->stream, &v, c->made_nll.p, sizeof(float), hipMemcpyDeviceToHost));
        *nll = v;
    }
    return NQH_OK;
}

// `steps` Adam steps; step t uses the contiguous chunk (t mod (n / batch)) of src
static nqh_status made_train_async(nqh_ctx* c, const uint64_t* src, int64_t n, int steps, int batch, float lr) {
    const int64_t chunks = n / batch;
    const size_t parts = nq_made_grad_parts(batch);
    nqh_status s;
    if ((s = ensure(c, c->made_gpart, sizeof(float) * c->made_P * parts))) return s;
    if ((s = ensure(c, c->made_lpart, sizeof(float) * parts))) return s;
    for (int t = 0; t < steps; ++t) {
        const uint64_t* kb = src + (int64_t)(c->made_t % chunks) * batch;
        HIPCHK(c, nq_launch_made_grad(P<float>(c->made_th), c->made_D, kb, batch, P<float>(c->made_gpart),
                                      P<float>(c->made_lpart), c->stream));
        if (t == steps - 1)
            HIPCHK(c, nq_launch_made_sum_parts(P<float>(c->made_gpart), (int)parts, c->made_D, P<float>(c->made_g),
                                               P<float>(c->made_lpart), P<float>(c->made_nll), batch, c->stream));
        c->made_t += 1;
        HIPCHK(c, nq_launch_made_adam(P<float>(c->made_th), P<float>(c->made_m), P<float>(c->made_v),
                                      P<float>(c->made_gpart), (int)parts, c->made_D, lr, c->made_t, c->stream));
    }
    return NQH_OK;
}

nqh_status nqh_made_train(nqh_ctx* c, const uint64_t* keys, int64_t n, int32_t steps, int32_t batch, double lr,
                          double* nll_last) {
    if (!c || steps < 0 || batch < 1 || !(lr > 0.0)) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_made(c))) return s;
    if (steps == 0) return NQH_OK;
    const uint64_t* src;
    if (keys) {
        if (n < batch) return fail(c, NQH_E_ARG, "need n >= batch");
        if ((s = made_check_keys(c, keys, n)) || (s = made_keys_to_device(c, keys, n))) return s;
        src = P<uint64_t>(c->made_tmp);
    } else {
        n = c->ring_fill;
        if (n < batch) return fail(c, NQH_E_STATE, "the sample buffer holds fewer than batch configurations");
        src = P<uint64_t>(c->ring);
    }
    if ((s = made_train_async(c, src, n, steps, batch, (float)lr))) return s;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (nll_last) {
        float v = 0.f;
        HIPCHK(c, copy_sync(c->stream, &v, c->made_nll.p, sizeof(float), hipMemcpyDeviceToHost));
        *nll_last = v;
    }
    return NQH_OK;
}

nqh_status nqh_made_logq(nqh_ctx* c, const uint64_t* keys, int64_t n, float* logq) {
    if (!c || !keys || !logq || n < 0) return NQH_E_ARG;
    if (n == 0) return NQH_OK;
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_made(c)) || (s = made_check_keys(c, keys, n))) return s;
    if ((s = made_keys_to_device(c, keys, n))) return s;
    dbuf out;
    if ((s = ensure(c, out, sizeof(float) * n))) return s;
    hipError_t e = nq_launch_made_logq(P<float>(c->made_th), c->made_D, P<uint64_t>(c->made_tmp), n, P<float>(out),
                                       c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = copy_sync(c->stream, logq, out.p, sizeof(float) * n, hipMemcpyDeviceToHost);
    release(out);
    if (e != hipSuccess) return hip_fail(c, e, "nqh_made_logq");
    return NQH_OK;
}

nqh_status nqh_made_sample(nqh_ctx* c, int64_t n, uint32_t iter, uint32_t tag, uint64_t* keys, float* logq) {
    if (!c || !keys || n < 0) return NQH_E_ARG;
    if (n == 0) return NQH_OK;
    (void)hipSetDevice(c->device);
    nqh_status s;
    if ((s = need_made(c))) return s;
    dbuf kd, lq;
    if ((s = ensure(c, kd, sizeof(uint64_t) * n))) return s;
    if ((s = ensure(c, lq, sizeof(float) * n))) { release(kd); return s; }
    hipError_t e = nq_launch_made_sample(P<float>(c->made_th), c->made_D, n, c->seed, iter, nullptr, tag,
                                         P<uint64_t>(kd), P<float>(lq), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = copy_sync(c->stream, keys, kd.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess && logq) e = copy_sync(c->stream, logq, lq.p, sizeof(float) * n, hipMemcpyDeviceToHost);
    release(kd);
    release(lq);
    if (e != hipSuccess) return hip_fail(c, e, "nqh_made_sample");
    return NQH_OK;
}

nqh_status nqh_made_stats(nqh_ctx* c, int64_t* proposed, int64_t* accepted, int64_t* buffer_fill) {
    if (!c) return NQH_E_ARG;
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (proposed) *proposed = c->made_proposed;
    if (accepted) {
        unsigned long long v[2] = {0, 0};
        if (c->acc.p) HIPCHK(c, copy_sync(c->stream, v, c->acc.p, sizeof(v), hipMemcpyDeviceToHost));
        *accepted = (int64_t)v[1];
    }
    if (buffer_fill) *buffer_fill = c->ring_fill;
    return NQH_OK;
}

}  // extern "C"
