/*
 * nqhip.h -- C ABI of libnqhip, the MI355X (gfx950) backend for the
 * variational Monte Carlo steady-state loop of NeuralQuantum.jl.
 *
 * The reference's boundary for this path is Julia method dispatch, not an
 * FFI; every entry point below names the reference method it stands in for
 * (file:line in NeuralQuantum.jl).  A Julia wrapper `ccall`s these symbols
 * (neuralquantum.jl_amd/julia/NeuralQuantumHIP.jl); in this repository they
 * are exercised from Python through ctypes.
 *
 * Conventions
 *   - status codes only, nothing is thrown across the ABI; on error the
 *     context keeps a message, nqh_last_error(ctx), valid until the next call;
 *   - host buffers are caller-owned, read or written only during the call and
 *     never retained; every call is synchronous with respect to them;
 *   - a context is not thread-safe: one context per host task;
 *   - configurations are packed keys: bit i (0-based) of the row half is site
 *     i+1 of sigma, bit N+i is site i+1 of eta, a set bit means v = +1; the
 *     reference's doubled index is toint((sigma, eta)) = key + 1
 *     (src/Hilbert/SuperHilbert.jl:50-56);
 *   - parameters are real FP64 in the reference's flat order
 *     (NDM: b_mu h_mu w_mu u_mu b_la h_la d_la w_la u_la, matrices
 *     column-major; src/Networks/MixedDensityMatrix/NDM.jl:17-18);
 *   - complex outputs are interleaved (re, im) doubles;
 *   - batches of samples are ordered s = chain + slice * chains, the
 *     reference's (batch, chain-length) layout (BatchedGradSampler.jl:88-110).
 */
#ifndef NQHIP_H
#define NQHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NQH_ABI_VERSION 1

typedef int32_t nqh_status;
#define NQH_OK          0
#define NQH_E_ARG      -1  /* invalid argument / unsupported shape */
#define NQH_E_HIP      -2  /* HIP runtime error */
#define NQH_E_NOTPD    -3  /* S + shift not positive definite (PosDefException, SRDirect.jl:79) */
#define NQH_E_NAN      -4  /* non-finite loss, force or update detected */
#define NQH_E_STATE    -5  /* call order: problem / network / sampler not set */
#define NQH_E_NOMEM    -6  /* device allocation failed */
#define NQH_E_CAPACITY -7  /* output capacity too small */

/* precision of the network evaluation / of the S build */
#define NQH_PREC_FP64 0
#define NQH_PREC_FP32 1
#define NQH_PREC_BF16 2    /* S build only: BF16 inputs, FP32 accumulation */

/* SR diagonal shift: additive eps*I (sr_shift, SRDirect.jl:60-65) or eps*diag(S) */
#define NQH_SHIFT_ADD 0
#define NQH_SHIFT_DIAG 1

/* sampler flags */
#define NQH_SAMPLER_COMPAT_STALE_LAST_SLICE 1  /* reproduce the reference's never-sampled last slice (defect D1) */
#define NQH_SAMPLER_PERSISTENT 2                /* keep chains across iterations instead of restarting */

typedef struct nqh_ctx nqh_ctx;

/*
 * Flat connection table of a KLocalLiouvillian (or of a KLocalOperator sum
 * for kind 3), built on the host by the operator algebra in the order the
 * reference enumerates connections (KLocalLiouvillian.jl:46-52,
 * KLocalOperatorTensor.jl:129-156):
 *   term t: kind 0 = row side only (HnH_l), 1 = column side only (HnH_r),
 *           2 = row (x) column nested loops (LLdag), 3 = ket operator;
 *           local operators term_op_l[t] / term_op_r[t] (-1 = identity);
 *   local operator o acts on sites op_sites[op_site_off[o] .. op_site_off[o+1])
 *           (0-based; local bit m <-> m-th listed site, HomogeneousSpin.jl:171-179)
 *           and owns rows op_row_off[o] .. op_row_off[o+1]-1 (2^k rows);
 *   row q owns entries row_ent_off[q] .. row_ent_off[q+1]-1: a complex mel
 *           and the mask of physical sites it flips (diagonal first, then
 *           columns ascending, |mel| >= 1e-5; KLocalOperator.jl:67-109).
 */
typedef struct nqh_optable {
    int32_t n_sites;
    int32_t n_terms;
    int32_t n_ops;
    int32_t n_rows;
    int32_t n_entries;
    const int32_t* term_kind;
    const int32_t* term_op_l;
    const int32_t* term_op_r;
    const int32_t* op_site_off;
    const int32_t* op_sites;
    const int32_t* op_row_off;
    const int32_t* row_ent_off;
    const double* ent_mel;
    const uint64_t* ent_flip;
} nqh_optable;

int32_t nqh_abi_version(void);

/* context on HIP device `device`; `seed` keys the Philox streams of the sampler
 * (replaces MetropolisSampler(...; seed), Metropolis.jl:26 / rng.jl:9-11) */
nqh_status nqh_create(nqh_ctx** out, int32_t device, uint64_t seed);
void nqh_destroy(nqh_ctx* ctx);
const char* nqh_last_error(const nqh_ctx* ctx);

/* the problem: KLocalLiouvillian connections (kind 0); replaces the Liouvillian
 * argument of BatchedSampler(net, sampler, prob, algo) (build_Batched.jl:1-14) */
nqh_status nqh_set_problem(nqh_ctx* ctx, const nqh_optable* table);

/* NDM activation f (src/Networks/activation.jl) */
#define NQH_ACT_SOFTPLUS 0   /* af_softplus = logℒ, f' = 1/(1+e^-x) (:14-18), the NDM default */
#define NQH_ACT_LOGCOSH  1   /* af_logcosh = logℒ2, f' = tanh (:5-10, :22-26) */

/* NDM(Float64, N, alpha_h, alpha_a, f) with nh = alpha_h N hidden and
 * na = alpha_a N ancilla units (NDM.jl:50-63); act = NQH_ACT_SOFTPLUS or
 * NQH_ACT_LOGCOSH.  N in 1..16 (doubled keys of 32 bits), nh and na in 1..64 */
nqh_status nqh_set_ndm(nqh_ctx* ctx, int32_t N, int32_t nh, int32_t na, int32_t act);

/* RBMSplit(T, N, alpha) with nh = alpha N hidden units (RBMSplit.jl:33-40):
 * complex parameters ar[N] ac[N] b[nh] Wr[nh x N] Wc[nh x N] (functor order,
 * column-major), passed to nqh_set_params interleaved (re, im).  The SR system
 * is complex Hermitian (SRDirect.jl:35-45 with a complex parameter type):
 * nqh_precondition returns Delta interleaved; nqh_grad_from_samples returns F
 * and Obar interleaved and S as its real embedding [[Re S, -Im S], [Im S, Re S]]
 * (2 npc x 2 npc). */
nqh_status nqh_set_rbmsplit(nqh_ctx* ctx, int32_t N, int32_t nh);

/* RBM(T, N, alpha) on kets (src/Networks/ClosedSystems/RBM.jl:28-61,
 * af_softplus): complex a[N] b[nh] W[nh x N], interleaved, for a Hamiltonian
 * problem (a table whose terms are all kets, kind 3).  One iteration is then
 * the energy path of BatchedValSampler.sample! (BatchedValSampler.jl:117-139):
 * E_loc = sum <s|H|s'> psi(s')/psi(s) (AccumulatorObsScalar), statistics of
 * E_loc, grad C = sum_s E_loc(s) conj(Oc(s)) / Ns, and the complex SR system
 * as for RBMSplit; configurations are keys of N bits. */
nqh_status nqh_set_rbm(nqh_ctx* ctx, int32_t N, int32_t nh);

/* doubles in the flat parameter vector (NDM: Np; RBMSplit, RBM: 2 x complex count) */
int64_t nqh_num_params(const nqh_ctx* ctx);
/* complex entries of one O row (NDM: Np; RBMSplit: complex parameter count) */
int64_t nqh_num_derivs(const nqh_ctx* ctx);

/* parameter upload / download (the Julia side keeps the canonical arrays and
 * re-uploads after Optimisers.update!, or reads back after nqh_update) */
nqh_status nqh_set_params(nqh_ctx* ctx, const double* flat, int64_t np);
nqh_status nqh_get_params(const nqh_ctx* ctx, double* flat, int64_t np);

/* precision of the network kernels (FP64 | FP32) and of the S build (FP64 | FP32 | BF16) */
nqh_status nqh_set_precision(nqh_ctx* ctx, int32_t net_prec, int32_t sbuild_prec);

/* log rho for B host-fed configurations: logpsi!(out, net, cache, sr, sc)
 * (NDMBatched.jl:94-175); out: 2B doubles */
nqh_status nqh_logpsi(nqh_ctx* ctx, const uint64_t* keys, int64_t B, double* logpsi);

/* log rho and the Np complex log-derivatives O (reference order):
 * logpsi_and_gradlogpsi!(grad, out, net, cache, sr, sc) (NDMBatched.jl:177-280);
 * O: B x Np complex, sample-major */
nqh_status nqh_logpsi_grad(nqh_ctx* ctx, const uint64_t* keys, int64_t B, double* logpsi, double* O);

/* connections of B configurations as the accumulator sees them (zero mels
 * skipped): accumulate_connections!(acc, liouv, v) (KLocalLiouvillian.jl:46-52);
 * counts[b] = number of connections (also when it exceeds cap: then
 * NQH_E_CAPACITY), keys_out / mel_out: B x cap */
nqh_status nqh_connections(nqh_ctx* ctx, const uint64_t* keys, int64_t B, int32_t cap, int32_t* counts,
                           uint64_t* keys_out, double* mel_out);

/* local estimator for B host-fed configurations: init! / accumulate_connections!
 * / finalize! of AccumulatorObsGrad (AccumulatorObsGrad.jl:39-127):
 * Lloc: 2B doubles; dLloc (nullable): B x Np complex, reference order */
nqh_status nqh_local_estimator(nqh_ctx* ctx, const uint64_t* keys, int64_t B, double* Lloc, double* dLloc);

/* sample!(is; sample=false) on host-fed samples (BatchedGradSampler.jl:77-123 +
 * setup_algorithm!, SRDirect.jl:26-49): chains x chain_length keys ->
 * stats = {mean, error, variance, tau, R} of |Lloc|^2 (stats.jl:52-77),
 * S = Re(Oc Oc^H)/Ns (Np x Np column-major, reference order, nullable),
 * F = Re(grad C) (Np, nullable), Obar (Np complex, nullable) */
nqh_status nqh_grad_from_samples(nqh_ctx* ctx, const uint64_t* keys, int64_t chains, int64_t chain_length,
                                 double* stats, double* S, double* F, double* Obar);

/* MetropolisSampler(LocalRule(), chain_length, passes; burn) with `chains`
 * chains (batch_sz) (Metropolis.jl:26-40); passes is used as given (the
 * reference bumps even values to odd at construction) */
nqh_status nqh_set_sampler(nqh_ctx* ctx, int64_t chains, int32_t chain_length, int32_t passes, int32_t burn,
                           int32_t flags);
/* iteration index that keys the Philox counters of the next sampling call */
nqh_status nqh_set_iteration(nqh_ctx* ctx, uint32_t iter);

/* one sampling pass (_sample_state!, BaseIterativeSampler.jl:5-17); keys_out
 * (nullable): chain_length x chains; accepted (nullable): accepted proposals */
nqh_status nqh_sample(nqh_ctx* ctx, uint64_t* keys_out, int64_t* accepted);

/* accepted moves per chain in the last sampling call (local proposals and,
 * with the learned proposal, global ones), the acceptance counter of
 * Metropolis.jl:57-59,162-163 split by chain: a chain at 0 has not moved
 * (stuck).  accepted_out: `chains` entries.  Valid after nqh_sample,
 * nqh_sample_grad or an SR iteration. */
nqh_status nqh_get_chain_acceptance(nqh_ctx* ctx, uint32_t* accepted_out, int64_t chains);

/* sample!(is) (BatchedGradSampler.jl:77-123): sample, log rho, O, L_loc,
 * statistics, centring, force and the S build; stats: 5 doubles (nullable) */
nqh_status nqh_sample_grad(nqh_ctx* ctx, double* stats);

/* precondition!(cache, SR, iter) (SRDirect.jl:51-90): S + shift, Cholesky,
 * Delta = S^-1 F; `refine` extra iterative-refinement steps in FP64;
 * delta_out (nullable): Np, reference order */
nqh_status nqh_precondition(nqh_ctx* ctx, double eps, int32_t shift_kind, int32_t refine, double* delta_out);

/* SR solver, the reference's SR(algorithm=...) (SR.jl:37-57):
 *   NQH_SOLVER_CHOLESKY  sr_cholesky (SRDirect.jl:26-90): S built on the
 *                        matrix cores and factored (default);
 *   NQH_SOLVER_CG        sr_cg through SRIterative.jl:71-152 with the
 *                        matrix-free S of SrMatrixR (SR_notfull.jl:84-148):
 *                        S is never formed (no Np^2 memory); conjugate
 *                        gradients on S v = Re(Oc^H Oc v)/Ns + shift, two
 *                        passes over the O rows per product, until
 *                        ||r|| <= tol ||F|| or maxiter products (0: 10 Np,
 *                        the reference's limit); without convergence the
 *                        step is zero (SRIterative.jl:134-150).  NDM networks
 *                        only; the solve synchronises with the host once per
 *                        product, so iterations with it are not graph-captured
 *                        and are not batched;
 *   NQH_SOLVER_MINRES    sr_minres (SRIterative.jl:101-108, IterativeSolvers.jl
 *                        0.8.1 `minres`) on the same matrix-free S: Lanczos
 *                        steps with Givens rotations, until the recurrence's
 *                        residual norm <= tol ||F|| or maxiter products; the
 *                        same zero step without convergence, host
 *                        synchronisation and limits as NQH_SOLVER_CG;
 *   NQH_SOLVER_MINRES_QLP sr_qlp (SRIterative.jl:92-100, the reference's
 *                        minresqlp.jl) on the same matrix-free S: the same
 *                        Lanczos vectors, the file's rotations, norm
 *                        estimates and stopping flags (relres = ||r|| /
 *                        (||A|| ||x|| + ||F||) <= tol, relAres <= tol, maxiter,
 *                        ...); as in that file a run that reaches maxiter
 *                        counts as converged and keeps its x;
 *   NQH_SOLVER_LSQR      sr_lsq (SRIterative.jl:109-116, IterativeSolvers.jl
 *                        0.8.1 `lsqr`) on the same matrix-free S: two products
 *                        per step, the package's default tolerances (atol =
 *                        btol = sqrt(eps), conlim = 1/sqrt(eps); tol is not
 *                        used, as the reference passes none); zero step when
 *                        it stops on the condition estimate or maxiter.
 * Takes effect from the next gradient (nqh_sample_grad / grad_from_samples). */
#define NQH_SOLVER_CHOLESKY 0
#define NQH_SOLVER_CG 1
#define NQH_SOLVER_MINRES 2
#define NQH_SOLVER_MINRES_QLP 3
#define NQH_SOLVER_LSQR 4
nqh_status nqh_set_sr_solver(nqh_ctx* ctx, int32_t solver, double tol, int32_t maxiter);
/* steps (one product each, two for LSQR) and final relative residual of the
 * last iterative solve (CG:
 * ||r|| / ||F|| of the recurrence's r; MINRES: its residual-norm estimate over
 * ||F||; MINRES-QLP: relres = ||r|| / (||A|| ||x|| + ||F||) of its estimates;
 * LSQR: its ||r|| estimate over ||F||) */
nqh_status nqh_get_sr_solver_info(nqh_ctx* ctx, int32_t* iters, double* rel_resid);

/* Optimisers.update!(Descent(lr), net, Delta) (apply.jl:25-31, rules.jl:15-17) */
nqh_status nqh_update(nqh_ctx* ctx, double lr);

/* one SR iteration = nqh_sample_grad + nqh_precondition + nqh_update */
nqh_status nqh_sr_iteration(nqh_ctx* ctx, double eps, int32_t shift_kind, int32_t refine, double lr,
                            double* stats);

/* the same iteration enqueued on the context's stream without waiting
 * (returns once the work is queued), and the wait for it (statistics, solve
 * and finiteness checks as nqh_sr_iteration).  Many contexts on one device,
 * each with its own stream, then run their iterations concurrently: a batch
 * of independent problems (for example a sweep of instances of uneven size).
 * With several iterations enqueued before one nqh_wait, the wait reports the
 * last of them.  No reference counterpart (execution strategy). */
nqh_status nqh_sr_iteration_async(nqh_ctx* ctx, double eps, int32_t shift_kind, int32_t refine, double lr);
nqh_status nqh_wait(nqh_ctx* ctx, double* stats);

/* A batch of independent problems, e.g. many Liouvillians of uneven size:
 * every member's SR iteration is captured, each on its member's stream,
 * forked from and joined into the batch's own stream, as ONE hipGraph that
 * is launched once per iteration (members with persistent chains run their
 * first, restart + burn-in, iteration eagerly and concurrently).  Members
 * are fully set-up contexts on one device (problem, network, sampler; no
 * learned proposal) and must not be destroyed while the batch exists.  A
 * member reconfigured between batch iterations (nqh_set_problem,
 * nqh_set_ndm / rbmsplit / rbm, nqh_set_sampler, nqh_set_precision,
 * nqh_set_iteration, or a buffer regrown by its own calls) makes the next
 * batch iteration capture the graph again with the member's new buffers,
 * shapes and iteration number; reading parameters or samples between
 * iterations is fine, and a member's own calls stay ordered with the
 * batch's iterations.  stats: n x 5, member i at 5 i in nqh_sr_iteration's
 * layout; on a member's failure the message names the member.  No reference
 * counterpart (the reference runs one problem per process). */
typedef struct nqh_batch nqh_batch;
nqh_status nqh_batch_create(nqh_batch** out, nqh_ctx* const* members, int32_t n);
void nqh_batch_destroy(nqh_batch* batch);
const char* nqh_batch_last_error(const nqh_batch* batch);
nqh_status nqh_batch_sr_iteration(nqh_batch* batch, double eps, int32_t shift_kind, int32_t refine, double lr,
                                  double* stats);

/* How a batch runs its iterations.  NQH_BATCH_PACKED (the default): when
 * every member qualifies (NDM network, Cholesky solver with refine = 0, BF16
 * S build, nh <= 32 and na <= 32, no learned proposal), each kernel of the
 * iteration is launched ONCE for all members: one grid holds every member's
 * workgroups, each running exactly the blocks of the member's own launch
 * with the member's arguments from a device descriptor, so a member's
 * results are bit-identical to its own run; the launches (about 30 plus the
 * longest member's Cholesky panels and substitution groups) form the
 * batch's graph.  Otherwise, and with NQH_BATCH_GRAPH, every member's own
 * launches are captured.  nqh_batch_launches: kernel launches per iteration
 * of the packed graph, 0 for the per-member graph, -1 before the first
 * captured iteration.  No reference counterpart. */
#define NQH_BATCH_PACKED 0
#define NQH_BATCH_GRAPH 1
nqh_status nqh_batch_set_mode(nqh_batch* batch, int32_t mode);
int32_t nqh_batch_launches(const nqh_batch* batch);

/* device time per phase, measured with HIP events on the context's stream
 * while profiling is on (accumulated, in ms; counts = timed occurrences):
 * [0] sampler  [1] estimator (log rho, O, L_loc, force)  [2] means, statistics,
 * force finish  [3] S build  [4] shift + Cholesky + solves  [5] update
 * [6] the Cholesky factorization alone (inside [4]).
 * No reference counterpart (the reference has no timers). */
#define NQH_NUM_PHASES 7
nqh_status nqh_set_profiling(nqh_ctx* ctx, int32_t on);
nqh_status nqh_get_phase_times(nqh_ctx* ctx, double* ms_out, int64_t* counts_out);

/* capture nqh_sr_iteration into a hipGraph and replay it (on = 1); the graph
 * is rebuilt when eps, the shift kind, refine or lr change, and dropped by
 * every setter.  With persistent chains the restart iteration runs eagerly.
 * No reference counterpart (execution strategy of sample! + precondition! +
 * update!). */
nqh_status nqh_set_graph(nqh_ctx* ctx, int32_t on);

/* launch shapes (0 = chosen from the problem): estimator samples per
 * workgroup (1..4096; default 16), estimator waves per workgroup (4 or 8;
 * default 8 while two workgroups fit a CU's LDS), BF16 Gram split-K chunks
 * (1..64; default about three tiles per CU with the partial sums at most
 * twice the input) and sampler waves per workgroup (1, 2 or 4; default 4).
 * They change how work is spread over the device, not what is computed; the
 * estimator's shape fixes the order of the force sum (deterministic for a
 * given shape).  Picked per workload by nqhip.autotune and stored in
 * nqhip/launch_shapes.json.  nqh_get_launch_shape: the four values as set.
 * No reference counterpart. */
nqh_status nqh_set_launch_shape(nqh_ctx* ctx, int32_t est_samples_per_block, int32_t est_waves, int32_t gram_splits,
                                int32_t sampler_waves);
nqh_status nqh_get_launch_shape(const nqh_ctx* ctx, int32_t* out4);

/* exact sums over all 2^(2N) doubled configurations (2^N for kets), the
 * batched ExactSampler (Exact.jl:110-181) with the accumulator on top:
 * out[0] = sum p |L_loc|^2 (= ||L rho||^2 / ||rho||^2 for a Liouvillian),
 * out[1], out[2] = Re, Im of sum p L_loc (the energy <H> for kets),
 * out[3] = log sum |psi|^2; p = |psi|^2 / sum |psi|^2.  logp_out (nullable,
 * 2^(2N) or 2^N doubles): log p of every configuration, indexed by key.
 * At most 2^30 configurations. */
nqh_status nqh_exact(nqh_ctx* ctx, double* out, double* logp_out);

/* Checkpoint and resume of the state this backend keeps beyond the
 * parameters (which nqh_get_params / nqh_set_params cover, the reference's
 * load_from_dict! hook, src/utils/loading.jl:8-14): the persistent chains'
 * configurations, the iteration index that keys the Philox streams, and with
 * the learned proposal its weights, Adam moments and step, its retraining
 * counter and the ring of recent samples it trains on.  The reference keeps
 * none of it (it restarts its chains every iteration, Metropolis.jl:106).
 * nqh_save_state writes nqh_state_bytes(ctx) bytes into buf; nqh_load_state
 * restores them into a context set up the same way (problem, network,
 * sampler and nqh_made_setup), after which its iterations continue those of
 * the saved context bit for bit. */
int64_t nqh_state_bytes(const nqh_ctx* ctx);
nqh_status nqh_save_state(nqh_ctx* ctx, void* buf, int64_t bytes);
nqh_status nqh_load_state(nqh_ctx* ctx, const void* buf, int64_t bytes);

/* nqh_exact without any device allocation: the same sums (out[0..3], logp_out
 * nullable), computed in chunks of 2^16 configurations with every buffer
 * beyond the caller's outputs in a file-scope device array of 3,801,344
 * bytes that each GPU holds from first use until the process ends; at most
 * 2^30 configurations (NQH_E_ARG beyond).  Calls that overlap on one GPU
 * share that array and race.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under hipGraph capture, serving
 * loops, harnesses that forbid hipMalloc). */
nqh_status nqh_exact_noalloc(nqh_ctx* ctx, double* out, double* logp_out);

/* compute_observables(is) for one observable (BatchedObsDMSampler.jl:59-105):
 * a Metropolis chain on the diagonal (sigma, sigma) of the density matrix,
 * restarted at every call and drawn from its own random stream (counter word
 * 3 = 16 + {0, 1}), `chains` x `chain_length` samples after `burn` calls of
 * `passes` proposals, each flipping one site of sigma and eta together; then
 * O_loc(sigma) = sum <sigma|O|sigma'> rho(sigma', sigma) / rho(sigma, sigma) for
 * the ket operator `op` (kind 3 terms) and the statistics of O_loc:
 * stats_out = [Re mean, error, variance, tau, R, Im mean]; the mean estimates
 * Tr(O rho) / Tr(rho).  Two reference behaviours are corrected: the chain uses
 * log p = Re log rho(sigma, sigma), not abs(log rho) (base_batched_networks.jl:
 * 249-253), and the network is evaluated at (sigma', sigma), not at the
 * diagonal (sigma', sigma') the accumulator feeds it (AccumulatorLogPsi.jl:85-91). */
nqh_status nqh_observable(nqh_ctx* ctx, const nqh_optable* op, int64_t chains, int32_t chain_length,
                          int32_t passes, int32_t burn, double* stats_out);
/* the last observable's doubled keys (nullable) and local values (nullable,
 * 2 n doubles), n <= chains x chain_length */
nqh_status nqh_get_observable_samples(nqh_ctx* ctx, uint64_t* keys_out, double* local_out, int64_t n);

/* statistics of the last gradient: [mean, error, variance, tau, R] of |L_loc|^2
 * (Liouvillian) or of E_loc (Hamiltonian: mean real part; out[5] = imaginary
 * part of the mean); n <= 8 values */
nqh_status nqh_get_stats(nqh_ctx* ctx, double* out, int32_t n);

/* last samples (chain_length x chains keys) / local values (2 Ns doubles) */
nqh_status nqh_get_samples(nqh_ctx* ctx, uint64_t* keys_out, int64_t n);
nqh_status nqh_get_local_values(nqh_ctx* ctx, double* lloc_out, int64_t n);

/* ---- learned proposal ---------------------------------------------------
 * A MADE over the D = 2N doubled bits (D -> 128 -> 128 -> D, ReLU, sigmoid
 * Bernoulli conditionals; FP32 parameters, BF16 MFMA products with FP32
 * accumulation), used as an independence proposal inside the Metropolis
 * chains with the exact Hastings ratio.  The reference has no learned
 * proposal; this plugs into the sampler where its `log_prob_bias` hook sits
 * (Metropolis.jl:54-67,149; precedent OperatorRule.jl:37,57): a global move
 * y ~ q is accepted iff u - exp(log p(y) - log p(x) + log q(x) - log q(y)) >= 0
 * is false.  Parameter layout (row-major): W1[128][D] b1[128] W2[128][128]
 * b2[128] W3[D][128] b3[D]. */
int64_t nqh_made_num_params(int32_t D);
/* enable with: `mix` = fraction of stored slices preceded by one global move
 * (slice i iff floor((i+1) mix) > floor(i mix)); every `retrain_every` sampler
 * calls, `train_steps` Adam steps (lr, batch) on a ring buffer of the last
 * `buffer_cap` samples (0: no buffer, no automatic training).  Parameters are
 * initialised (masked N(0, 1/fan_in), seeded by the context seed).  Needs
 * nqh_set_ndm; mix = 0 keeps the parameters but turns the global moves off. */
nqh_status nqh_made_setup(nqh_ctx* ctx, double mix, int32_t retrain_every, int32_t train_steps, int32_t batch,
                          double lr, int64_t buffer_cap);
nqh_status nqh_made_set_params(nqh_ctx* ctx, const float* theta, int64_t n);   /* resets Adam */
nqh_status nqh_made_get_params(nqh_ctx* ctx, float* theta, int64_t n);
/* gradient of the mean negative log-likelihood over keys (masked), and the NLL */
nqh_status nqh_made_grad(nqh_ctx* ctx, const uint64_t* keys, int64_t B, float* grad, double* nll);
/* `steps` Adam steps (Kingma & Ba 2015, beta 0.9/0.999, eps 1e-8); step t uses
 * the contiguous chunk (t mod (n / batch)) of keys, or of the sample buffer
 * when keys is NULL; nll_last: mean NLL of the last step's batch */
nqh_status nqh_made_train(nqh_ctx* ctx, const uint64_t* keys, int64_t n, int32_t steps, int32_t batch, double lr,
                          double* nll_last);
nqh_status nqh_made_logq(nqh_ctx* ctx, const uint64_t* keys, int64_t n, float* logq);
/* n ancestral samples; Philox counter (sample, bit, iter, 3 + 16 tag); logq nullable */
nqh_status nqh_made_sample(nqh_ctx* ctx, int64_t n, uint32_t iter, uint32_t tag, uint64_t* keys, float* logq);
/* global moves proposed / accepted by the last sampler call; configurations in the buffer */
nqh_status nqh_made_stats(nqh_ctx* ctx, int64_t* proposed, int64_t* accepted, int64_t* buffer_fill);

#ifdef __cplusplus
}
#endif

#endif /* NQHIP_H */
