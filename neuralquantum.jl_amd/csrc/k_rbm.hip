// RBMSplit density-matrix network (src/Networks/MixedDensityMatrix/RBMSplit.jl,
// RBMSplitBatched.jl) on the steady-state path:
//
//   log rho(sr, sc) = ar . sr + ac . sc + sum_h logL(theta_h),
//   theta = b + Wr sr + Wc sc,                            (RBMSplit.jl:45-50,84-102)
//
// complex parameters in the reference's order ar[N] ac[N] b[nh] Wr[nh x N]
// Wc[nh x N] (column-major), stored interleaved (re, im) in precision T.  The
// network is analytic, so O is the holomorphic derivative (RBMSplit.jl:104-123)
// and the SR system is complex Hermitian (SRDirect.jl:26-46 with a complex
// parameter type): S_jk = sum_s conj(Oc_js) Oc_ks / Ns, F = conj(L)^T dL / Ns
// - C Obar.  It is solved as the real symmetric system
//   [[A, -B], [B, A]] [x; y] = [Re F; Im F],   S = A + i B,   Delta = x + i y
// whose matrix is assembled from the Gram matrix of the rows [Re Oc | Im Oc].
//
// Kernels
//   k_rbm_estimate   one wavefront per sample, lane h owns hidden unit h: log
//                    rho, the O row [Re O | Im O], L_loc, and the force
//                    sum_s conj(L_loc)/Ns sum_c W_c O(c) accumulated per unit
//                    (dense part x_s e_h plus sparse flipped-site terms) in an
//                    exact fixed-point LDS accumulator (order independent);
//                    or per-sample dL_loc (parity mode, one wave per block)
//   k_rbm_metropolis LocalRule Metropolis on |rho|^2, the same random streams
//                    and schedule as the NDM sampler; a group of lanes per chain
//                    keeps theta in registers
//   k_embed_complex  in-place [[G_rr, G_ri], [G_ir, G_ii]] -> [[A, -B], [B, A]]
//   k_force_all      F[k] = sum of partials - C mean[k] for every column
//   k_update_complex p -= lr (x + i y)
#include "nq_device.h"
#include "nq_kernels.h"
#include "nq_table.h"

namespace {

template <typename T>
__device__ __forceinline__ T xv(uint32_t bits, int i) { return ((bits >> i) & 1u) ? T(1) : T(-1); }
template <typename T>
__device__ __forceinline__ T cflip(uint32_t bits, int i) { return ((bits >> i) & 1u) ? T(-2) : T(2); }

template <typename T>
__device__ __forceinline__ cplx<T> cexp_(cplx<T> z) {
    T s, c;
    nq_sincos(z.im, &s, &c);
    const T e = nq_exp(z.re);
    return cmk(e * c, e * s);
}

template <typename T, int NMAX, int HS>
__global__ __launch_bounds__(256, 2) void k_rbm_estimate(rbm_args a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = a.N, nh = a.nh, npc = a.npc;
    cplx<T>* sp = reinterpret_cast<cplx<T>*>(smem);
    // exact fixed-point force accumulator (fx_add): entries [Re (npc) | Im (npc)], then the poison word
    unsigned long long* sF = reinterpret_cast<unsigned long long*>(smem + a.lds_param_bytes);
    int* poison = reinterpret_cast<int*>(sF + 4 * npc);
    const bool useF = a.Fpart != nullptr || a.G != nullptr;
    cplx<T>* sW = reinterpret_cast<cplx<T>*>(smem + a.lds_param_bytes + (useF ? fx_bytes(2 * npc) : 0));
    unsigned char* sT = smem + ((reinterpret_cast<unsigned char*>(sW + (blockDim.x >> 6) * a.max_conn) - smem + 15) & ~15);
    const cplx<T>* gp = reinterpret_cast<const cplx<T>*>(a.p);
    for (int k = threadIdx.x; k < npc; k += blockDim.x) sp[k] = gp[k];
    if (useF) {
        for (int k = threadIdx.x; k < 4 * npc; k += blockDim.x) sF[k] = 0ull;
        if (threadIdx.x == 0) *poison = 0;
    }
    const dev_table tab = a.tab_lds ? table_to_lds(a.tab, sT) : a.tab;
    __syncthreads();
    // integer sums: the same total whatever order the waves add in
    auto fadd = [&](int k, cplx<T> v) {
        fx_add(sF, k, (double)v.re, poison);
        fx_add(sF, npc + k, (double)v.im, poison);
    };
    // ket (RBM.jl): a[N] b[nh] W[nh x N] on the row configuration only
    const bool ket = a.ket != 0;
    const int oar = 0, oac = N, ob = ket ? N : 2 * N, owr = ob + nh, owc = owr + nh * N;

    const int lane0 = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wpb = blockDim.x >> 6;
    const int64_t gw = (int64_t)blockIdx.x * wpb + wave;
    const int64_t nw = (int64_t)gridDim.x * wpb;
    const uint32_t pmask = (N >= 32) ? 0xffffffffu : ((1u << N) - 1u);
    cplx<T>* wbuf = sW + wave * a.max_conn;
    const int R = 2 * npc;

    for (int64_t s = gw; s < a.Ns; s += nw) {
        int lane = lane0;
        asm volatile("" : "+v"(lane));
        const uint64_t key = a.keys[s];
        const uint32_t ks = (uint32_t)key & pmask;
        const uint32_t ke = (uint32_t)(key >> N) & pmask;

        cplx<T> th[HS], f[HS], fp[HS];
#pragma unroll
        for (int r = 0; r < HS; ++r) {
            const int h = lane + 64 * r;
            th[r] = f[r] = fp[r] = cmk(T(0), T(0));
            if (h < nh) {
                cplx<T> t = sp[ob + h];
#pragma unroll
                for (int i = 0; i < NMAX; ++i)
                    if (i < N) {
                        cfma_r(t, sp[owr + h + nh * i], xv<T>(ks, i));
                        if (!ket) cfma_r(t, sp[owc + h + nh * i], xv<T>(ke, i));
                    }
                th[r] = t;
                csoftplus_and_sigmoid(t, &f[r], &fp[r]);
            }
        }
        // log rho
        {
            cplx<T> acc = cmk(T(0), T(0));
#pragma unroll
            for (int r = 0; r < HS; ++r) acc = cadd(acc, f[r]);
            if (lane < N) {
                cfma_r(acc, sp[oar + lane], xv<T>(ks, lane));
                if (!ket) cfma_r(acc, sp[oac + lane], xv<T>(ke, lane));
            }
            acc.re = wave_total(acc.re);
            acc.im = wave_total(acc.im);
            if (lane == 0 && a.lp) {
                T* lp = reinterpret_cast<T*>(a.lp);
                lp[2 * s] = acc.re;
                lp[2 * s + 1] = acc.im;
            }
        }
        // O row: [Re O | Im O], reference order
        if (a.Y) {
            T* y = reinterpret_cast<T*>(a.Y) + (int64_t)s * R;
#pragma unroll
            for (int r = 0; r < HS; ++r) {
                const int h = lane + 64 * r;
                if (h < nh) {
                    y[ob + h] = fp[r].re;
                    y[npc + ob + h] = fp[r].im;
#pragma unroll
                    for (int i = 0; i < NMAX; ++i)
                        if (i < N) {
                            const T xs = xv<T>(ks, i), xe = xv<T>(ke, i);
                            y[owr + h + nh * i] = fp[r].re * xs;
                            y[npc + owr + h + nh * i] = fp[r].im * xs;
                            if (!ket) {
                                y[owc + h + nh * i] = fp[r].re * xe;
                                y[npc + owc + h + nh * i] = fp[r].im * xe;
                            }
                        }
                }
            }
            if (lane < N) {
                y[oar + lane] = xv<T>(ks, lane);
                y[npc + oar + lane] = T(0);
                if (!ket) {
                    y[oac + lane] = xv<T>(ke, lane);
                    y[npc + oac + lane] = T(0);
                }
            }
        }
        if (!a.do_conn) continue;

        // ---- pass 1: W_c = mel exp(log rho_c - log rho_s), L_loc
        cplx<T> Lloc = cmk(T(0), T(0));
        int nc = 0;
        table_walk<T>(tab, ks, ke, [&](cplx<T> mel, uint32_t ms, uint32_t me) {
            if ((ms | me) == 0u) {
                Lloc = cadd(Lloc, mel);
                if (nc < a.max_conn) wbuf[nc] = mel;
                ++nc;
                return;
            }
            cplx<T> dl = cmk(T(0), T(0));
#pragma unroll
            for (int r = 0; r < HS; ++r) {
                const int h = lane + 64 * r;
                if (h < nh) {
                    cplx<T> t = th[r];
                    for (uint32_t m = ms; m; m &= m - 1) {
                        const int i = __builtin_ctz(m);
                        cfma_r(t, sp[owr + h + nh * i], cflip<T>(ks, i));
                    }
                    for (uint32_t m = me; m; m &= m - 1) {
                        const int i = __builtin_ctz(m);
                        cfma_r(t, sp[owc + h + nh * i], cflip<T>(ke, i));
                    }
                    cplx<T> f2, p2;
                    csoftplus_and_sigmoid(t, &f2, &p2);
                    dl = cadd(dl, csub(f2, f[r]));
                }
            }
            dl.re = wave_total(dl.re);
            dl.im = wave_total(dl.im);
            for (uint32_t m = ms; m; m &= m - 1) {
                const int i = __builtin_ctz(m);
                cfma_r(dl, sp[oar + i], cflip<T>(ks, i));
            }
            for (uint32_t m = me; m; m &= m - 1) {
                const int i = __builtin_ctz(m);
                cfma_r(dl, sp[oac + i], cflip<T>(ke, i));
            }
            const cplx<T> W = cmul(mel, cexp_(dl));
            Lloc = cadd(Lloc, W);
            if (nc < a.max_conn) wbuf[nc] = W;
            ++nc;
        });
        if (lane == 0 && a.Lloc) {
            T* lo = reinterpret_cast<T*>(a.Lloc);
            lo[2 * s] = Lloc.re;
            lo[2 * s + 1] = Lloc.im;
        }
        if (!useF) continue;

        // ---- pass 2: sum_c V W_c O(c), V = conj(L_loc) / Ns (or 1 for dL_loc)
        const cplx<T> V = a.G ? cmk(T(1), T(0)) : cmk((T)((double)Lloc.re * a.invNs), (T)(-(double)Lloc.im * a.invNs));
        cplx<T> e[HS];
#pragma unroll
        for (int r = 0; r < HS; ++r) e[r] = cmk(T(0), T(0));
        cplx<T> sumVW = cmk(T(0), T(0));
        int c = 0;
        table_walk<T>(tab, ks, ke, [&](cplx<T> mel, uint32_t ms, uint32_t me) {
            (void)mel;
            const cplx<T> VW = cmul(V, c < a.max_conn ? wbuf[c] : cmk(T(0), T(0)));
            ++c;
            sumVW = cadd(sumVW, VW);
            if ((ms | me) == 0u) {
#pragma unroll
                for (int r = 0; r < HS; ++r) e[r] = cadd(e[r], cmul(VW, fp[r]));
                return;
            }
#pragma unroll
            for (int r = 0; r < HS; ++r) {
                const int h = lane + 64 * r;
                if (h < nh) {
                    cplx<T> t = th[r];
                    for (uint32_t m = ms; m; m &= m - 1) {
                        const int i = __builtin_ctz(m);
                        cfma_r(t, sp[owr + h + nh * i], cflip<T>(ks, i));
                    }
                    for (uint32_t m = me; m; m &= m - 1) {
                        const int i = __builtin_ctz(m);
                        cfma_r(t, sp[owc + h + nh * i], cflip<T>(ke, i));
                    }
                    cplx<T> f2, p2;
                    csoftplus_and_sigmoid(t, &f2, &p2);
                    const cplx<T> g = cmul(VW, p2);
                    e[r] = cadd(e[r], g);
                    // x_c = x_s + Delta on the flipped sites
                    for (uint32_t m = ms; m; m &= m - 1) {
                        const int i = __builtin_ctz(m);
                        fadd(owr + h + nh * i, cscale(g, cflip<T>(ks, i)));
                    }
                    for (uint32_t m = me; m; m &= m - 1) {
                        const int i = __builtin_ctz(m);
                        fadd(owc + h + nh * i, cscale(g, cflip<T>(ke, i)));
                    }
                }
            }
            if (lane < N) {
                if ((ms >> lane) & 1u) fadd(oar + lane, cscale(VW, cflip<T>(ks, lane)));
                if ((me >> lane) & 1u) fadd(oac + lane, cscale(VW, cflip<T>(ke, lane)));
            }
        });
        // dense parts
#pragma unroll
        for (int r = 0; r < HS; ++r) {
            const int h = lane + 64 * r;
            if (h < nh) {
                fadd(ob + h, e[r]);
#pragma unroll
                for (int i = 0; i < NMAX; ++i)
                    if (i < N) {
                        fadd(owr + h + nh * i, cscale(e[r], xv<T>(ks, i)));
                        fadd(owc + h + nh * i, cscale(e[r], xv<T>(ke, i)));
                    }
            }
        }
        if (lane < N) {
            fadd(oar + lane, cscale(sumVW, xv<T>(ks, lane)));
            fadd(oac + lane, cscale(sumVW, xv<T>(ke, lane)));
        }
        if (a.G) {   // one wave per block: sF belongs to this sample
            __syncthreads();
            double* g = a.G + (int64_t)s * 2 * npc;
            const bool bad = *poison != 0;
            for (int k = lane; k < npc; k += 64) {
                g[2 * k] = bad ? __builtin_nan("") : fx_value(sF, k);
                g[2 * k + 1] = bad ? __builtin_nan("") : fx_value(sF, npc + k);
                sF[2 * k] = sF[2 * k + 1] = 0ull;
                sF[2 * (npc + k)] = sF[2 * (npc + k) + 1] = 0ull;
            }
            __syncthreads();
            if (lane == 0) *poison = 0;
            __syncthreads();
        }
    }
    if (a.Fpart && !a.G) {
        __syncthreads();
        double* out = a.Fpart + (int64_t)blockIdx.x * 2 * npc;
        const bool bad = *poison != 0;
        for (int k = threadIdx.x; k < 2 * npc; k += blockDim.x) out[k] = bad ? __builtin_nan("") : fx_value(sF, k);
    }
}

// ---- Metropolis (LocalRule) on |rho|^2; log p = 2 Re log rho.  Lane g of a
// group of lpc lanes owns hidden units h = g + lpc r.
template <typename T, int US>
__global__ __launch_bounds__(256) void k_rbm_metropolis(rbm_samp_args a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = a.N, nh = a.nh, npc = a.npc;
    cplx<T>* sp = reinterpret_cast<cplx<T>*>(smem);
    const cplx<T>* gp = reinterpret_cast<const cplx<T>*>(a.p);
    for (int k = threadIdx.x; k < npc; k += blockDim.x) sp[k] = gp[k];
    __syncthreads();
    const bool ket = a.ket != 0;
    const int oar = 0, oac = N, ob = ket ? N : 2 * N, owr = ob + nh, owc = owr + nh * N;
    const bool diag = a.diag != 0;      // density matrix on its diagonal: log p = Re log rho(sigma, sigma)
    const int nsite = (ket || diag) ? N : 2 * N;   // sites a proposal picks from
    const uint32_t ctr3 = 16u * a.tag;

    const int lpc = a.lpc;
    const int lane = threadIdx.x & 63;
    const int g = lane & (lpc - 1);
    const int gbase = lane - g;
    const int cpw = 64 / lpc;
    const int64_t wave_glob = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t chain = wave_glob * cpw + lane / lpc;
    const bool live = chain < a.chains;
    const uint32_t pmask = (N >= 32) ? 0xffffffffu : ((1u << N) - 1u);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const uint32_t iter = a.iter_dev ? *a.iter_dev : a.iter;

    uint64_t key = 0;
    if (live) {
        if (a.restart) {
            philox_out o = philox4x32_10(0u, iter, (uint32_t)chain, ctr3, k0, k1);
            const uint64_t m2 = (nsite >= 64) ? ~0ull : ((1ull << nsite) - 1ull);
            key = ((uint64_t)o.x[0] | ((uint64_t)o.x[1] << 32)) & m2;
            if (diag) key |= key << N;
        } else {
            key = a.chain_state[chain];
        }
    }
    cplx<T> th[US];
    T fr[US];
    uint64_t acc = 0;
    uint32_t t = a.t0;
    uint32_t rx0 = 0, rx1 = 0, rx2 = 0;
    const int calls = a.burn + a.stored;
    for (int call = 0; call < calls; ++call) {
        const uint32_t ks = (uint32_t)key & pmask, ke = (uint32_t)(key >> N) & pmask;
#pragma unroll
        for (int r = 0; r < US; ++r) {
            const int h = g + lpc * r;
            th[r] = cmk(T(0), T(0));
            fr[r] = T(0);
            if (h < nh) {
                cplx<T> z = sp[ob + h];
                for (int i = 0; i < N; ++i) {
                    cfma_r(z, sp[owr + h + nh * i], xv<T>(ks, i));
                    if (!ket) cfma_r(z, sp[owc + h + nh * i], xv<T>(ke, i));
                }
                th[r] = z;
                fr[r] = csoftplus_re(z);
            }
        }
        for (int q = 0; q < a.passes; ++q, ++t) {
            const int slot = (int)((t - a.t0) % (uint32_t)lpc);
            if (slot == 0) {
                philox_out o = philox4x32_10(t + (uint32_t)g, iter, (uint32_t)chain, ctr3 + 1u, k0, k1);
                rx0 = o.x[0]; rx1 = o.x[1]; rx2 = o.x[2];
            }
            const uint32_t x0 = __shfl(rx0, gbase + slot, 64);
            const uint32_t x1 = __shfl(rx1, gbase + slot, 64);
            const uint32_t x2 = __shfl(rx2, gbase + slot, 64);
            const int site = (int)(((uint64_t)x0 * (uint64_t)nsite) >> 32);
            const bool col = site >= N;
            const int j = col ? site - N : site;
            const uint32_t bits = col ? (uint32_t)(key >> N) : (uint32_t)key;
            const T c = ((bits >> j) & 1u) ? T(-2) : T(2);
            const int ow = col ? owc : owr;
            T delta = T(0);
            cplx<T> nth[US];
            T nfr[US];
#pragma unroll
            for (int r = 0; r < US; ++r) {
                const int h = g + lpc * r;
                nth[r] = th[r];
                nfr[r] = fr[r];
                if (h < nh) {
                    cfma_r(nth[r], sp[ow + h + nh * j], c);
                    if (diag) cfma_r(nth[r], sp[owc + h + nh * j], c);
                    nfr[r] = csoftplus_re(nth[r]);
                    delta += nfr[r] - fr[r];
                }
            }
            delta = group_total(delta, lpc);
            delta += c * sp[(col ? oac : oar) + j].re;
            if (diag) delta += c * sp[oac + j].re;
            else delta *= T(2);
            bool accept;
            if constexpr (sizeof(T) == 8) {
                const double u = u01_53(x1, x2);
                accept = !(u - exp((double)delta) >= 0.0);
            } else {
                const float u = u01_24(x1);
                accept = !(u - expf((float)delta) >= 0.0f);
            }
            if (accept && live) {
                key ^= diag ? ((1ull << site) | (1ull << (site + N))) : (1ull << site);
                ++acc;
#pragma unroll
                for (int r = 0; r < US; ++r) { th[r] = nth[r]; fr[r] = nfr[r]; }
            }
        }
        if (call >= a.burn && live && g == 0 && a.samples) {
            const int slice = a.slice0 + call - a.burn;
            a.samples[chain + (int64_t)slice * a.chains] = key;
        }
    }
    if (live && g == 0) {
        if (a.chain_state) a.chain_state[chain] = key;
        if (a.compat_stale_last_slice && a.samples)
            a.samples[chain + (int64_t)(a.slice0 + a.stored) * a.chains] = 0ull;
        if (a.accepted) atomicAdd(a.accepted, (unsigned long long)acc);
        if (a.chain_acc) a.chain_acc[chain] += (uint32_t)acc;   // one lane per chain: no race
    }
}

// [[G_rr, G_ri], [G_ir, G_ii]] (n = npc blocks of a 2n x 2n symmetric G, column-major,
// leading dimension ld) -> [[A, -B], [B, A]], A = G_rr + G_ii, B = G_ri - G_ir.
// Thread (j, k) owns the four positions (j,k), (n+j,n+k), (n+j,k), (j,n+k).
__global__ void k_embed_complex(double* S, int n, int ld) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (j >= n || k >= n) return;
    const int64_t L = ld;
    const double grr = S[j + k * L], gii = S[(n + j) + (n + k) * L];
    const double gri = S[j + (n + k) * L], gir = S[(n + j) + k * L];
    const double A = grr + gii, B = gri - gir;
    S[j + k * L] = A;
    S[(n + j) + (n + k) * L] = A;
    S[(n + j) + k * L] = B;
    S[j + (n + k) * L] = -B;
}

// ---- energy path (BatchedValSampler.jl:70-139): statistics of the complex
// local energies, and grad C = sum_s E_s conj(Oc_s) / Ns.

__device__ double blk_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < nw; ++k) t += red[k];
    return t;
}

// stat_analysis (stats.jl:52-77) of complex values E[s], s = chain + slice B:
// out = [Re mean, error, mean chain variance, tau, R, Im mean], variances
// of complex values as mean |x - m|^2 with the n - 1 normalisation
template <typename T>
__global__ __launch_bounds__(256) void k_stats_cplx(const T* E, int64_t B, int64_t Lc, double* out) {
    __shared__ double red[16];
    const double invL = 1.0 / (double)Lc;
    double sr = 0.0, si = 0.0;
    for (int64_t j = threadIdx.x; j < B; j += blockDim.x)
        for (int64_t i = 0; i < Lc; ++i) {
            sr += (double)E[2 * (j + i * B)] * invL;
            si += (double)E[2 * (j + i * B) + 1] * invL;
        }
    const double mr = blk_sum(sr, red) / (double)B, mi = blk_sum(si, red) / (double)B;
    double vch = 0.0, vmu = 0.0, vall = 0.0;
    for (int64_t j = threadIdx.x; j < B; j += blockDim.x) {
        double cr = 0.0, ci = 0.0;
        for (int64_t i = 0; i < Lc; ++i) {
            cr += (double)E[2 * (j + i * B)];
            ci += (double)E[2 * (j + i * B) + 1];
        }
        cr *= invL;
        ci *= invL;
        double v = 0.0;
        for (int64_t i = 0; i < Lc; ++i) {
            const double xr = (double)E[2 * (j + i * B)], xi = (double)E[2 * (j + i * B) + 1];
            v += (xr - cr) * (xr - cr) + (xi - ci) * (xi - ci);
            vall += (xr - mr) * (xr - mr) + (xi - mi) * (xi - mi);
        }
        vch += v / (double)(Lc - 1);
        vmu += (cr - mr) * (cr - mr) + (ci - mi) * (ci - mi);
    }
    vch = blk_sum(vch, red);
    vmu = blk_sum(vmu, red);
    vall = blk_sum(vall, red);
    if (threadIdx.x == 0) {
        const double var_mu_ch = vmu / (double)(B - 1);
        const double var_mu = vall / (double)(B * Lc - 1);
        const double t = var_mu_ch / var_mu;
        out[0] = mr;
        out[1] = sqrt(var_mu_ch / (double)B);
        out[2] = vch / (double)B;
        out[3] = fmax(0.0, 0.5 * (t * (double)Lc - 1.0));
        out[4] = sqrt((double)(Lc - 1) / (double)Lc + t);
        out[5] = mi;
    }
}

// part[chunk][k] (+ [npc + k]) = Re (Im) sum over the chunk's samples of
// E_s conj(Y_s,k + i Y_s,npc+k - mean); Y rows [Re O | Im O]
template <typename T>
__global__ __launch_bounds__(256) void k_energy_force_part(const T* Y, const T* E, const double* mu, int64_t Ns,
                                                           int npc, int chunks, double* part) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int ch = blockIdx.y;
    if (k >= npc) return;
    const int64_t per = (Ns + chunks - 1) / chunks;
    const int64_t s0 = (int64_t)ch * per, s1 = (s0 + per < Ns) ? s0 + per : Ns;
    const double mr = mu[k], mi = mu[npc + k];
    double ar = 0.0, ai = 0.0;
    for (int64_t s = s0; s < s1; ++s) {
        const double er = (double)E[2 * s], ei = (double)E[2 * s + 1];
        const double orr = (double)Y[s * 2 * npc + k] - mr, oi = (double)Y[s * 2 * npc + npc + k] - mi;
        ar += er * orr + ei * oi;
        ai += ei * orr - er * oi;
    }
    part[(int64_t)ch * 2 * npc + k] = ar;
    part[(int64_t)ch * 2 * npc + npc + k] = ai;
}

// F[k] = sum_g part2[g][k] - C mean[k] for every column k < n
__global__ void k_force_all(const double* part2, int ngroups, const double* mu, const double* stats, int n,
                            double* F) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double acc = 0.0;
    for (int g = 0; g < ngroups; ++g) acc += part2[(int64_t)g * n + k];
    F[k] = acc - stats[0] * mu[k];
}

// p[k] -= lr (delta[k] + i delta[npc + k]) on the interleaved complex parameters
__global__ void k_update_complex(double* p64, float* p32, const double* delta, int npc, double lr, nq_step_guard g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npc) return;
    // no step after a failed solve, a non-finite step or a non-finite loss
    if ((g.info && *g.info != 0) || (g.flag && *g.flag != 0) || (g.stats && !isfinite(g.stats[0]))) return;
    const double re = p64[2 * k] - lr * delta[k];
    const double im = p64[2 * k + 1] - lr * delta[npc + k];
    p64[2 * k] = re;
    p64[2 * k + 1] = im;
    p32[2 * k] = (float)re;
    p32[2 * k + 1] = (float)im;
}

template <typename T, int NMAX, int HS>
hipError_t est_t(const rbm_args& a, int grid, hipStream_t st) {
    hipLaunchKernelGGL((k_rbm_estimate<T, NMAX, HS>), dim3(grid), dim3(a.G ? 64 : 256), rbm_est_lds_bytes(a, sizeof(T)),
                       st, a);
    return hipGetLastError();
}

template <typename T, int NMAX>
hipError_t est_n(const rbm_args& a, int grid, hipStream_t st) {
    const int hs = (a.nh + 63) / 64;
    if (hs <= 1) return est_t<T, NMAX, 1>(a, grid, st);
    if (hs <= 2) return est_t<T, NMAX, 2>(a, grid, st);
    return hipErrorInvalidValue;
}

template <typename T, int US>
hipError_t samp_t(const rbm_samp_args& a, hipStream_t st) {
    const int cpw = 64 / a.lpc;
    const int64_t waves = (a.chains + cpw - 1) / cpw;
    const int64_t grid = (waves + 3) / 4;
    hipLaunchKernelGGL((k_rbm_metropolis<T, US>), dim3((unsigned)grid), dim3(256), 2 * sizeof(T) * (size_t)a.npc, st, a);
    return hipGetLastError();
}

}  // namespace

size_t rbm_est_lds_bytes(const rbm_args& a, size_t elt) {
    const size_t b = (size_t)a.lds_param_bytes + ((a.Fpart || a.G) ? (size_t)fx_bytes(2 * a.npc) : 0) +
                     4 * (size_t)a.max_conn * 2 * elt;
    return ((b + 15) & ~(size_t)15) + (a.tab_lds ? (size_t)a.tab.bytes : 0);
}

hipError_t nq_launch_rbm_estimate(const rbm_args& a, bool fp64, int grid, hipStream_t st) {
    if (a.N > 16 || a.nh < 1) return hipErrorInvalidValue;
    if (rbm_est_lds_bytes(a, fp64 ? 8 : 4) > 160 * 1024) return hipErrorInvalidValue;
    if (fp64) return a.N <= 8 ? est_n<double, 8>(a, grid, st) : est_n<double, 16>(a, grid, st);
    return a.N <= 8 ? est_n<float, 8>(a, grid, st) : est_n<float, 16>(a, grid, st);
}

hipError_t nq_launch_rbm_metropolis(const rbm_samp_args& a, bool fp64, hipStream_t st) {
    const int us = (a.nh + a.lpc - 1) / a.lpc;
    if (a.N > 32 || a.lpc < 1 || a.lpc > 64 || (a.lpc & (a.lpc - 1))) return hipErrorInvalidValue;
    if (fp64) {
        if (us <= 1) return samp_t<double, 1>(a, st);
        if (us <= 2) return samp_t<double, 2>(a, st);
        if (us <= 4) return samp_t<double, 4>(a, st);
    } else {
        if (us <= 1) return samp_t<float, 1>(a, st);
        if (us <= 2) return samp_t<float, 2>(a, st);
        if (us <= 4) return samp_t<float, 4>(a, st);
    }
    return hipErrorInvalidValue;
}

hipError_t nq_launch_embed_complex(double* S, int npc, int ld, hipStream_t st) {
    hipLaunchKernelGGL(k_embed_complex, dim3((npc + 255) / 256, npc), dim3(256), 0, st, S, npc, ld);
    return hipGetLastError();
}

hipError_t nq_launch_force_all(const double* Fpart, int nblk, double* part2, const double* mu, const double* stats,
                               int n, double* F, hipStream_t st) {
    hipError_t e = nq_launch_force_part(Fpart, nblk, n, part2, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_force_all, dim3((n + 255) / 256), dim3(256), 0, st, part2, NQ_FORCE_GROUPS, mu, stats, n,
                       F);
    return hipGetLastError();
}

hipError_t nq_launch_stats_cplx(const void* E, int fp64, int64_t B, int64_t Lc, double* out, hipStream_t st) {
    if (fp64)
        hipLaunchKernelGGL(k_stats_cplx<double>, dim3(1), dim3(256), 0, st, (const double*)E, B, Lc, out);
    else
        hipLaunchKernelGGL(k_stats_cplx<float>, dim3(1), dim3(256), 0, st, (const float*)E, B, Lc, out);
    return hipGetLastError();
}

hipError_t nq_launch_energy_force(const void* Y, const void* E, int fp64, const double* mu, int64_t Ns, int npc,
                                  double* part, int chunks, double* F, hipStream_t st) {
    dim3 grid((npc + 255) / 256, chunks);
    if (fp64)
        hipLaunchKernelGGL(k_energy_force_part<double>, grid, dim3(256), 0, st, (const double*)Y, (const double*)E, mu,
                           Ns, npc, chunks, part);
    else
        hipLaunchKernelGGL(k_energy_force_part<float>, grid, dim3(256), 0, st, (const float*)Y, (const float*)E, mu,
                           Ns, npc, chunks, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return nq_launch_colsum_finish(part, chunks, 2 * npc, 1.0 / (double)Ns, F, st);
}

hipError_t nq_launch_update_complex(double* p64, float* p32, const double* delta, int npc, double lr,
                                    nq_step_guard g, hipStream_t st) {
    hipLaunchKernelGGL(k_update_complex, dim3((npc + 255) / 256), dim3(256), 0, st, p64, p32, delta, npc, lr, g);
    return hipGetLastError();
}

// ---- exact full sums over every configuration (ExactSampler, Exact.jl:110-181)
namespace {
__global__ void k_iota_keys(uint64_t* keys, int64_t n, uint64_t start) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = start + (uint64_t)i;
}

// one block: out = {max of log p, sum w, sum w |L|^2, sum w Re L, sum w Im L},
// w = exp(log p - max), log p = 2 Re log psi
template <typename T>
__global__ __launch_bounds__(1024) void k_exact_partial(const T* lp, const T* L, int64_t n, double* out) {
    __shared__ double red[16];
    double m = -INFINITY;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, 2.0 * (double)lp[2 * i]);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) red[w] = m;
    __syncthreads();
    m = red[0];
    for (int k = 1; k < nw; ++k) m = fmax(m, red[k]);
    __syncthreads();
    double z = 0.0, zc = 0.0, zr = 0.0, zi = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double wt = exp(2.0 * (double)lp[2 * i] - m);
        const double lr = (double)L[2 * i], li = (double)L[2 * i + 1];
        z += wt;
        zc += wt * (lr * lr + li * li);
        zr += wt * lr;
        zi += wt * li;
    }
    double v[4] = {z, zc, zr, zi};
    for (int q = 0; q < 4; ++q) {
        double x = wave_sum(v[q]);
        __syncthreads();
        if (lane == 0) red[w] = x;
        __syncthreads();
        double t = 0.0;
        for (int k = 0; k < nw; ++k) t += red[k];
        v[q] = t;
    }
    if (threadIdx.x == 0) {
        out[0] = m;
        out[1] = v[0];
        out[2] = v[1];
        out[3] = v[2];
        out[4] = v[3];
    }
}
}  // namespace

hipError_t nq_launch_iota_keys(uint64_t* keys, int64_t n, uint64_t start, hipStream_t st) {
    hipLaunchKernelGGL(k_iota_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys, n, start);
    return hipGetLastError();
}

namespace {
// log p = 2 Re log rho of n configurations, as doubles
template <typename T>
__global__ void k_logp_of(const T* lp, int64_t n, double* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = 2.0 * (double)lp[2 * i];
}

// the chunks' partial sums (max log p, sum w, sum w |L|^2, sum w Re L,
// sum w Im L, each relative to its chunk's max) combined in chunk order:
// out = [sum p |L|^2, Re sum p L, Im sum p L, log Z]
__global__ void k_exact_combine(const double* part, int64_t nch, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double M = -INFINITY;
    for (int64_t ci = 0; ci < nch; ++ci) M = fmax(M, part[5 * ci]);
    double Z = 0.0, ZC = 0.0, ZR = 0.0, ZI = 0.0;
    for (int64_t ci = 0; ci < nch; ++ci) {
        const double f = exp(part[5 * ci] - M);
        Z += f * part[5 * ci + 1];
        ZC += f * part[5 * ci + 2];
        ZR += f * part[5 * ci + 3];
        ZI += f * part[5 * ci + 4];
    }
    out[0] = ZC / Z;
    out[1] = ZR / Z;
    out[2] = ZI / Z;
    out[3] = M + log(Z);
}
}  // namespace

hipError_t nq_launch_logp_of(const void* lp, int fp64, int64_t n, double* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (fp64)
        hipLaunchKernelGGL(k_logp_of<double>, dim3(g), dim3(256), 0, st, (const double*)lp, n, out);
    else
        hipLaunchKernelGGL(k_logp_of<float>, dim3(g), dim3(256), 0, st, (const float*)lp, n, out);
    return hipGetLastError();
}

hipError_t nq_launch_exact_combine(const double* part, int64_t nch, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_exact_combine, dim3(1), dim3(64), 0, st, part, nch, out);
    return hipGetLastError();
}

hipError_t nq_launch_exact_partial(const void* lp, const void* L, int fp64, int64_t n, double* out, hipStream_t st) {
    if (fp64)
        hipLaunchKernelGGL(k_exact_partial<double>, dim3(1), dim3(1024), 0, st, (const double*)lp, (const double*)L, n,
                           out);
    else
        hipLaunchKernelGGL(k_exact_partial<float>, dim3(1), dim3(1024), 0, st, (const float*)lp, (const float*)L, n,
                           out);
    return hipGetLastError();
}
