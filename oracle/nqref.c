/*
 * nqref -- ORACLE, test infrastructure only.
 *
 * FP64 CPU restatement of NeuralQuantum.jl's steady-state loop, used by the
 * parity tests as the checker and by bench.py as the CPU baseline ("port").
 * Nothing in the product links or calls this file.  It follows the reference
 * line by line in structure (full network recompute per Metropolis proposal,
 * full Np-gradient per connected configuration), not the device design.
 *
 *   NDM log rho / O ........ src/Networks/MixedDensityMatrix/NDM.jl:76-96,
 *                            NDMBatched.jl:94-280, activation.jl:14-18
 *   connection walk ........ KLocalLiouvillian.jl:46-52,
 *                            KLocalOperatorTensor.jl:129-156 (on the flat table)
 *   local estimator ........ AccumulatorObsGrad.jl:39-127
 *   loss / gradient ........ BatchedGradSampler.jl:99-118, stats.jl:52-77,
 *                            BaseIterativeSampler.jl:19-26, SRDirect.jl:26-49
 *   Metropolis + LocalRule . Metropolis.jl:101-167, LocalRule.jl:30-38,
 *                            BaseIterativeSampler.jl:5-17 (schedule, defect D1)
 *   RNG .................... Philox4x32-10 (Salmon et al., SC'11), the stream
 *                            the device sampler uses, so trajectories compare
 *                            bit for bit (the reference's MersenneTwister
 *                            stream itself is not reproducible without Julia).
 */
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ------------------------------------------------------------------ RNG */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void nqref_philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* stream layout shared with the device sampler:
 *   initial configuration of chain c in iteration it: ctr = (0, it, c, 0),
 *     key bits = (x0 | x1 << 32) masked to 2N bits;
 *   proposal t (0-based, per chain, per iteration): ctr = (t, it, c, 1),
 *     doubled site = (x0 * 2N) >> 32, u = ((x1 >> 5) 2^26 + (x2 >> 6)) / 2^53. */
static void draw(uint64_t seed, uint32_t t, uint32_t it, uint32_t chain, uint32_t dom, uint32_t o[4]) {
    uint32_t ctr[4] = {t, it, chain, dom};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    nqref_philox(ctr, key, o);
}

/* ------------------------------------------------------------ activations */
static inline double sp_r(double x) { return log1p(exp(x)); }          /* logL, activation.jl:18 */
static inline double dsp_r(double x) { return 1.0 / (1.0 + exp(-x)); } /* dlogL, activation.jl:15 */
static inline cplx sp_c(cplx z) {
    double x = creal(z), y = cimag(z), ex = exp(x);
    double re = 0.5 * log1p(ex * (2.0 * cos(y) + ex));
    double im = atan2(ex * sin(y), 1.0 + ex * cos(y));
    return re + I * im;
}
static inline cplx dsp_c(cplx z) { return 1.0 / (1.0 + cexp(-z)); }

/* logL2 = log cosh (af_logcosh), as activation.jl:5-10 writes it:
 * log(cosh x) for |x| <= 12, |x| - log 2 beyond; complex: logL2(Re z) +
 * log(cos(Im z) + i tanh(Re z) sin(Im z)); derivative tanh (:6, :26) */
static inline double lc_r(double x) { return fabs(x) <= 12.0 ? log(cosh(x)) : fabs(x) - log(2.0); }
static inline double dlc_r(double x) { return tanh(x); }
static inline cplx lc_c(cplx z) {
    double a = creal(z), b = cimag(z);
    return lc_r(a) + clog(cos(b) + I * tanh(a) * sin(b));
}
static inline cplx dlc_c(cplx z) { return ctanh(z); }

/* the NDM's activation f (NDM.jl:50): 0 = softplus (default), 1 = logcosh */
static int g_act = 0;
void nqref_set_activation(int act) { g_act = act; }
static inline double f_r(double x) { return g_act ? lc_r(x) : sp_r(x); }
static inline double df_r(double x) { return g_act ? dlc_r(x) : dsp_r(x); }
static inline cplx f_c(cplx z) { return g_act ? lc_c(z) : sp_c(z); }
static inline cplx df_c(cplx z) { return g_act ? dlc_c(z) : dsp_c(z); }

/* ------------------------------------------------------------ NDM layout */
typedef struct { int N, nh, na; int64_t bmu, hmu, wmu, umu, bla, hla, dla, wla, ula, np; } ndm_off;

static ndm_off ndm_layout(int N, int nh, int na) {
    ndm_off o; o.N = N; o.nh = nh; o.na = na;
    int64_t p = 0;
    o.bmu = p; p += N;                 /* functor order, NDM.jl:17-18 */
    o.hmu = p; p += nh;
    o.wmu = p; p += (int64_t)nh * N;
    o.umu = p; p += (int64_t)na * N;
    o.bla = p; p += N;
    o.hla = p; p += nh;
    o.dla = p; p += na;
    o.wla = p; p += (int64_t)nh * N;
    o.ula = p; p += (int64_t)na * N;
    o.np = p;
    return o;
}

int64_t nqref_ndm_np(int N, int nh, int na) { return ndm_layout(N, nh, na).np; }

static void unpack(uint64_t key, int N, double *s, double *e) {
    for (int i = 0; i < N; ++i) {
        s[i] = ((key >> i) & 1) ? 1.0 : -1.0;
        e[i] = ((key >> (i + N)) & 1) ? 1.0 : -1.0;
    }
}

/* log rho(s, e) and, if O != NULL, the Np complex log-derivatives in the
 * reference's flat order (NDMBatched.jl:177-280). */
static cplx ndm_eval(const ndm_off *L, const double *p, const double *s, const double *e, cplx *O) {
    const int N = L->N, nh = L->nh, na = L->na;
    double sl_s = 0, sl_e = 0, sm_s = 0, sm_e = 0, bls = 0, bms = 0;
    for (int i = 0; i < N; ++i) {
        bls += p[L->bla + i] * (s[i] + e[i]);
        bms += p[L->bmu + i] * (s[i] - e[i]);
    }
    for (int h = 0; h < nh; ++h) {
        double tls = p[L->hla + h], tle = p[L->hla + h], tms = p[L->hmu + h], tme = p[L->hmu + h];
        for (int i = 0; i < N; ++i) {
            double wl = p[L->wla + h + (int64_t)nh * i], wm = p[L->wmu + h + (int64_t)nh * i];
            tls += wl * s[i]; tle += wl * e[i];
            tms += wm * s[i]; tme += wm * e[i];
        }
        sl_s += f_r(tls); sl_e += f_r(tle);
        sm_s += f_r(tms); sm_e += f_r(tme);
        if (O) {
            double als = df_r(tls), ale = df_r(tle), ams = df_r(tms), ame = df_r(tme);
            O[L->hla + h] = 0.5 * (als + ale);
            O[L->hmu + h] = 0.5 * I * (ams - ame);
            for (int i = 0; i < N; ++i) {
                O[L->wla + h + (int64_t)nh * i] = 0.5 * (als * s[i] + ale * e[i]);
                O[L->wmu + h + (int64_t)nh * i] = 0.5 * I * (ams * s[i] - ame * e[i]);
            }
        }
    }
    cplx pi = 0;
    for (int a = 0; a < na; ++a) {
        double ul = 0, um = 0;
        for (int i = 0; i < N; ++i) {
            ul += p[L->ula + a + (int64_t)na * i] * (s[i] + e[i]);
            um += p[L->umu + a + (int64_t)na * i] * (s[i] - e[i]);
        }
        cplx arg = (0.5 * ul + p[L->dla + a]) + I * (0.5 * um);
        pi += f_c(arg);
        if (O) {
            cplx d = df_c(arg);
            O[L->dla + a] = d;
            for (int i = 0; i < N; ++i) {
                O[L->ula + a + (int64_t)na * i] = 0.5 * d * (s[i] + e[i]);
                O[L->umu + a + (int64_t)na * i] = 0.5 * I * d * (s[i] - e[i]);
            }
        }
    }
    if (O) {
        for (int i = 0; i < N; ++i) {
            O[L->bla + i] = 0.5 * (s[i] + e[i]);
            O[L->bmu + i] = 0.5 * I * (s[i] - e[i]);
        }
    }
    double gl = 0.5 * (sl_s + sl_e + bls);
    double gm = 0.5 * (sm_s - sm_e + bms);
    return gl + I * gm + pi;
}

/* threads of the loops over samples and chains below (OpenMP; the
 * reference's @MT mode runs one sampler per thread, threads_wrappers.jl:14-31) */
void nqref_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

void nqref_ndm_logpsi(int N, int nh, int na, const double *p, int64_t B, const uint64_t *keys, double *out) {
    ndm_off L = ndm_layout(N, nh, na);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < B; ++b) {
        double s[64], e[64];
        unpack(keys[b], N, s, e);
        cplx v = ndm_eval(&L, p, s, e, NULL);
        out[2 * b] = creal(v); out[2 * b + 1] = cimag(v);
    }
}

/* O: B x Np complex, interleaved (re, im), sample-major. */
void nqref_ndm_logpsi_grad(int N, int nh, int na, const double *p, int64_t B, const uint64_t *keys,
                           double *out, double *O) {
    ndm_off L = ndm_layout(N, nh, na);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < B; ++b) {
        double s[64], e[64];
        unpack(keys[b], N, s, e);
        cplx v = ndm_eval(&L, p, s, e, (cplx *)(O + 2 * b * L.np));
        out[2 * b] = creal(v); out[2 * b + 1] = cimag(v);
    }
}

/* --------------------------------------------------------- connection walk */
typedef struct {
    int32_t n_sites, n_terms;
    const int32_t *term_kind, *term_op_l, *term_op_r;
    const int32_t *op_site_off, *op_sites, *op_row_off, *row_ent_off;
    const double *ent_mel;      /* 2 * n_entries, interleaved */
    const uint64_t *ent_flip;
} nqref_table;

static int local_row(const nqref_table *t, int o, uint64_t bits) {
    int r = 0;
    for (int m = t->op_site_off[o]; m < t->op_site_off[o + 1]; ++m)
        r |= (int)((bits >> t->op_sites[m]) & 1) << (m - t->op_site_off[o]);
    return r;
}

typedef void (*conn_cb)(void *ctx, cplx mel, uint64_t new_key);

/* Connections of `key` in the reference's order; zero mels skipped when asked. */
static void walk(const nqref_table *t, uint64_t key, int skip_zero, conn_cb cb, void *ctx) {
    const int N = t->n_sites;
    const uint64_t pm = (N >= 64) ? ~0ull : ((1ull << N) - 1);
    const uint64_t ks = key & pm, ke = key >> N;
    for (int q = 0; q < t->n_terms; ++q) {
        int kind = t->term_kind[q];
        if (kind == 0 || kind == 3) {
            int o = t->term_op_l[q];
            int row = t->op_row_off[o] + local_row(t, o, ks);
            for (int k = t->row_ent_off[row]; k < t->row_ent_off[row + 1]; ++k) {
                cplx mel = t->ent_mel[2 * k] + I * t->ent_mel[2 * k + 1];
                if (skip_zero && mel == 0) continue;
                cb(ctx, mel, key ^ t->ent_flip[k]);
            }
        } else if (kind == 1) {
            int o = t->term_op_r[q];
            int row = t->op_row_off[o] + local_row(t, o, ke);
            for (int k = t->row_ent_off[row]; k < t->row_ent_off[row + 1]; ++k) {
                cplx mel = t->ent_mel[2 * k] + I * t->ent_mel[2 * k + 1];
                if (skip_zero && mel == 0) continue;
                cb(ctx, mel, key ^ (t->ent_flip[k] << N));
            }
        } else {
            int ol = t->term_op_l[q], orr = t->term_op_r[q];
            int rl = t->op_row_off[ol] + local_row(t, ol, ks);
            int rr = t->op_row_off[orr] + local_row(t, orr, ke);
            for (int k = t->row_ent_off[rl]; k < t->row_ent_off[rl + 1]; ++k) {
                cplx ml = t->ent_mel[2 * k] + I * t->ent_mel[2 * k + 1];
                for (int k2 = t->row_ent_off[rr]; k2 < t->row_ent_off[rr + 1]; ++k2) {
                    cplx mr = t->ent_mel[2 * k2] + I * t->ent_mel[2 * k2 + 1];
                    cplx mel = ml * mr;
                    if (skip_zero && mel == 0) continue;
                    cb(ctx, mel, key ^ t->ent_flip[k] ^ (t->ent_flip[k2] << N));
                }
            }
        }
    }
}

typedef struct { int cap, n; uint64_t *keys; double *mel; } list_ctx;
static void list_cb(void *c, cplx mel, uint64_t k) {
    list_ctx *l = (list_ctx *)c;
    if (l->n < l->cap) {
        l->keys[l->n] = k;
        l->mel[2 * l->n] = creal(mel);
        l->mel[2 * l->n + 1] = cimag(mel);
    }
    l->n++;
}

/* Returns the number of connections (may exceed cap; only cap are written). */
int nqref_connections(const nqref_table *t, uint64_t key, int skip_zero, int cap, uint64_t *keys_out,
                      double *mel_out) {
    list_ctx l = {cap, 0, keys_out, mel_out};
    walk(t, key, skip_zero, list_cb, &l);
    return l.n;
}

/* ------------------------------------------------------- local estimator */
typedef struct {
    const ndm_off *L; const double *p; int N;
    cplx logpsi_s; cplx res; cplx *dres; cplx *Obuf;
    double s[64], e[64];
} est_ctx;

static void est_cb(void *c, cplx mel, uint64_t k) {
    est_ctx *x = (est_ctx *)c;
    unpack(k, x->N, x->s, x->e);
    /* the reference evaluates every connected configuration with its full
     * gradient (AccumulatorObsGrad.jl:102-122); no shortcut for unchanged ones */
    cplx lp = ndm_eval(x->L, x->p, x->s, x->e, x->dres ? x->Obuf : NULL);
    cplx w = mel * cexp(lp - x->logpsi_s);
    x->res += w;
    if (x->dres)
        for (int64_t q = 0; q < x->L->np; ++q) x->dres[q] += w * x->Obuf[q];
}

/* Lloc: B complex (interleaved); dLloc: B x Np complex or NULL. */
void nqref_local_estimator(const nqref_table *t, int nh, int na, const double *p, int64_t B,
                           const uint64_t *keys, double *Lloc, double *dLloc) {
    const int N = t->n_sites;
    ndm_off L = ndm_layout(N, nh, na);
#pragma omp parallel
    {
        cplx *Obuf = (cplx *)malloc(sizeof(cplx) * L.np);
        est_ctx x;
        x.L = &L; x.p = p; x.N = N; x.Obuf = Obuf;
#pragma omp for schedule(dynamic, 16)
        for (int64_t b = 0; b < B; ++b) {
            double s[64], e[64];
            unpack(keys[b], N, s, e);
            x.logpsi_s = ndm_eval(&L, p, s, e, NULL);
            x.res = 0;
            x.dres = dLloc ? (cplx *)(dLloc + 2 * b * L.np) : NULL;
            if (x.dres) memset(x.dres, 0, sizeof(cplx) * L.np);
            walk(t, keys[b], 1, est_cb, &x);
            Lloc[2 * b] = creal(x.res); Lloc[2 * b + 1] = cimag(x.res);
        }
        free(Obuf);
    }
}

/* ----------------------------------------------- loss and gradient (SR) */
/* stat_analysis of vals[j + i*B] (B chains x Lc slices), stats.jl:52-77:
 * out = {mean, error, variance, tau, R}. */
void nqref_stat_analysis(int64_t B, int64_t Lc, const double *vals, double *out) {
    double *mu_ch = (double *)malloc(sizeof(double) * B);
    double mu = 0;
    for (int64_t j = 0; j < B; ++j) {
        double m = 0;
        for (int64_t i = 0; i < Lc; ++i) m += vals[j + i * B];
        mu_ch[j] = m / (double)Lc;
    }
    for (int64_t j = 0; j < B; ++j) mu += mu_ch[j];
    mu /= (double)B;
    double var_ch_mean = 0;
    for (int64_t j = 0; j < B; ++j) {
        double v = 0;
        for (int64_t i = 0; i < Lc; ++i) { double d = vals[j + i * B] - mu_ch[j]; v += d * d; }
        var_ch_mean += v / (double)(Lc - 1);
    }
    var_ch_mean /= (double)B;
    double mean_mu = 0, var_mu_ch = 0;
    for (int64_t j = 0; j < B; ++j) mean_mu += mu_ch[j];
    mean_mu /= (double)B;
    for (int64_t j = 0; j < B; ++j) { double d = mu_ch[j] - mean_mu; var_mu_ch += d * d; }
    var_mu_ch /= (double)(B - 1);
    double var_mu = 0;
    for (int64_t q = 0; q < B * Lc; ++q) { double d = vals[q] - mu; var_mu += d * d; }
    var_mu /= (double)(B * Lc - 1);
    double t = var_mu_ch / var_mu;
    out[0] = mu;
    out[1] = sqrt(var_mu_ch / (double)B);
    out[2] = var_ch_mean;
    out[3] = fmax(0.0, 0.5 * (t * (double)Lc - 1.0));
    out[4] = sqrt((double)(Lc - 1) / (double)Lc + t);
    free(mu_ch);
}

/* Post-processing of BatchedGradSampler.sample! + setup_algorithm!:
 * O (Ns x Np complex), Lloc (Ns complex), dLloc (Ns x Np complex) ->
 * S (Np x Np real, column-major), F (Np), Obar (Np complex), stats[5].
 * Samples are ordered s = j + i*B (chain j fastest). */
void nqref_grad_step(int64_t Np, int64_t B, int64_t Lc, const double *O, const double *Lloc,
                     const double *dLloc, double *S, double *F, double *Obar, double *stats) {
    const int64_t Ns = B * Lc;
    const cplx *Oc = (const cplx *)O;
    const cplx *Ll = (const cplx *)Lloc;
    const cplx *dL = (const cplx *)dLloc;
    double *c2 = (double *)malloc(sizeof(double) * Ns);
    for (int64_t s = 0; s < Ns; ++s) c2[s] = creal(Ll[s]) * creal(Ll[s]) + cimag(Ll[s]) * cimag(Ll[s]);
    nqref_stat_analysis(B, Lc, c2, stats);
    const double C = stats[0];
    cplx *ob = (cplx *)Obar;
    for (int64_t k = 0; k < Np; ++k) {
        cplx m = 0;
        for (int64_t s = 0; s < Ns; ++s) m += Oc[s * Np + k];
        ob[k] = m / (double)Ns;
    }
    for (int64_t k = 0; k < Np; ++k) {
        cplx g = 0;
        for (int64_t s = 0; s < Ns; ++s) g += conj(Ll[s]) * dL[s * Np + k];
        g = g / (double)Ns - C * ob[k];
        F[k] = creal(g);
    }
    /* S = Re(Oc Oc^H) / Ns */
    cplx *X = (cplx *)malloc(sizeof(cplx) * Ns * Np);
    for (int64_t s = 0; s < Ns; ++s)
        for (int64_t k = 0; k < Np; ++k) X[s * Np + k] = Oc[s * Np + k] - ob[k];
    for (int64_t k = 0; k < Np; ++k)
        for (int64_t l = 0; l <= k; ++l) {
            double acc = 0;
            for (int64_t s = 0; s < Ns; ++s)
                acc += creal(X[s * Np + k]) * creal(X[s * Np + l]) + cimag(X[s * Np + k]) * cimag(X[s * Np + l]);
            S[k + l * Np] = S[l + k * Np] = acc / (double)Ns;
        }
    free(X);
    free(c2);
}

/* ------------------------------------------------------------ Metropolis */
static double ndm_logp(const ndm_off *L, const double *p, uint64_t key) {
    double s[64], e[64];
    unpack(key, L->N, s, e);
    return 2.0 * creal(ndm_eval(L, p, s, e, NULL));     /* base_batched_networks.jl:255-258 */
}

/* Reference-structured Metropolis with LocalRule on the doubled space.
 * passes is used as given (the reference bumps even values to odd at
 * construction, Metropolis.jl:30-38; the caller does that).
 * samples: Lc x B keys, s = j + i*B.  compat_stale_last_slice reproduces
 * defect D1 (slice Lc never sampled, left at the all -1 fill = key 0).
 * Returns the number of accepted proposals. */
int64_t nqref_metropolis(int N, int nh, int na, const double *p, int64_t B, int64_t Lc, int passes,
                         int burn, uint64_t seed, uint32_t iter, int compat_stale_last_slice,
                         uint64_t *samples) {
    ndm_off L = ndm_layout(N, nh, na);
    const uint64_t mask2 = (2 * N >= 64) ? ~0ull : ((1ull << (2 * N)) - 1);
    int64_t acc = 0;
    const int64_t stored = compat_stale_last_slice ? Lc - 1 : Lc;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : acc)
    for (int64_t j = 0; j < B; ++j) {
        uint32_t o[4];
        draw(seed, 0, iter, (uint32_t)j, 0, o);
        uint64_t key = ((uint64_t)o[0] | ((uint64_t)o[1] << 32)) & mask2;
        uint32_t t = 0;
        double lp = ndm_logp(&L, p, key);
        for (int64_t call = 0; call < burn + stored; ++call) {
            /* one samplenext! call: `passes` proposals (Metropolis.jl:146-167) */
            lp = ndm_logp(&L, p, key);
            for (int q = 0; q < passes; ++q, ++t) {
                draw(seed, t, iter, (uint32_t)j, 1, o);
                int site = (int)(((uint64_t)o[0] * (uint64_t)(2 * N)) >> 32);
                double u = ((double)(o[1] >> 5) * 67108864.0 + (double)(o[2] >> 6)) * (1.0 / 9007199254740992.0);
                uint64_t nk = key ^ (1ull << site);
                double lpn = ndm_logp(&L, p, nk);
                double prob = u - exp(lpn - lp);
                if (!(prob >= 0)) { key = nk; lp = lpn; ++acc; }
            }
            if (call >= burn) samples[j + (call - burn) * B] = key;
        }
        if (compat_stale_last_slice) samples[j + (Lc - 1) * B] = 0;
    }
    return acc;
}
