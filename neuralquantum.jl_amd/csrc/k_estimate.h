// Per-sample kernels of the steady-state loop: log rho(s), the compact
// log-derivative row O(s), the local estimator L_loc(s) = sum_c mel_c
// exp(log rho_c - log rho_s) over the Liouvillian connections, and the force
// sum_s conj(L_loc(s)) dL_loc(s) / Ns contracted on the fly, never
// materialising the Np-vector dL_loc(s).
//
// Reference semantics: NDMBatched.jl:94-280 (log rho, O), AccumulatorObsGrad.jl
// :39-127 (L_loc, dL_loc = sum_c mel_c exp(..) O(c), no -O(s) term),
// BatchedGradSampler.jl:99-118 (force = conj(L_loc) dL_loc / Ns), with the
// NDM's activation f (softplus or logcosh, activation.jl:5-18) a template
// parameter.
//
// Mapping: one wavefront per sample (the table walk is wave-uniform), lanes
// over hidden units.  Lane l owns real unit u = l + 64 r (lambda for u < nh,
// mu otherwise, row and column side) and Pi half-unit q = l + 64 r (q < na:
// row half of Pi unit q, otherwise column half of unit q - na).  The
// effective angles theta = h + W x of the sample live in registers; a
// connection flipping k sites costs k FMAs per unit.
//
// Three kinds of launch (template parameter KIND):
//   EK_WALK   log rho, O rows, and pass 1 over the connections: softplus
//             (or logcosh) of the changed angles + one wave reduction -> W_c
//             = mel_c exp(log rho_c - log rho_s), L_loc = sum_c W_c; with
//             rec_mode 1 it writes per sample the changed connections (W_c,
//             flip masks), their count and the unchanged connections' mel sum.
//   EK_FORCE  reads those records and L_loc and adds the force
//             Re(V W_c O_k(c)), V = conj(L_loc)/Ns.  FP32 network: the
//             products on the matrix cores (force_mma_body, k_force_mma,
//             below).  FP64: per unit a dense part (x_s times a per-unit
//             scalar) plus a sparse part on the flipped sites only (x_c =
//             x_s + Delta_c); each lane sums its units' entries of one sample
//             in registers (the flipped site indexes a register vector with a
//             wave-uniform index: VGPR index mode), then adds them once per
//             sample into the workgroup's force accumulator in LDS, held as
//             exact 64-bit fixed-point integers (fx_add), so the sum is the
//             same whatever order the waves arrive in.  The per-workgroup
//             partials are then summed in a fixed order.
//   EK_GRAD   the parity mode: everything for one sample per wavefront and the
//             per-sample dL_loc(s) written out (the force pass run with V = 1
//             and V = -i for its real and imaginary parts).
#pragma once
#include "nq_device.h"
#include "nq_table.h"
#include "nq_kernels.h"

namespace nq_est {

enum { EK_WALK = 0, EK_FORCE = 1, EK_GRAD = 2 };

template <typename T>
__device__ __forceinline__ T xv(uint32_t bits, int i) { return ((bits >> i) & 1u) ? T(1) : T(-1); }
template <typename T>
__device__ __forceinline__ T cflip(uint32_t bits, int i) { return ((bits >> i) & 1u) ? T(-2) : T(2); }

// P-order index -> reference flat index (the parity mode stores reference order)
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

// changed connections evaluated per transposed reduction in pass 1 (fewer
// in FP64, where the shares take twice the registers); lanes per connection
// after the reduction
template <typename T> constexpr int est_batch() { return sizeof(T) == 8 ? 4 : 8; }

// TAB_LDS: the connection table is staged in LDS (a compile-time choice, so
// every table access is a known LDS or global access, never a flat one)
// the block body; bx = the workgroup's number in its problem's grid of gx
template <typename T, int NMAX, int RS, int PS, bool TAB_LDS, int KIND, int ACT>
__device__ __forceinline__ void estimate_body(const est_args& a, int bx, int gx) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr bool WALK = KIND == EK_WALK, FORCE = KIND == EK_FORCE, GRAD = KIND == EK_GRAD;
    using vN = T __attribute__((ext_vector_type(NMAX)));
    const ndm_dims d = a.d;
    const int N = d.N, nh = d.nh, na = d.na;
    T* sp = reinterpret_cast<T*>(smem);
    // force accumulator: two 64-bit words per entry (exact fixed point), then the poison word
    unsigned long long* sF = reinterpret_cast<unsigned long long*>(smem + a.lds_param_bytes);
    int* poison = reinterpret_cast<int*>(sF + 2 * d.np);
    const int fbytes = WALK ? 0 : fx_bytes(d.np);
    cplx<T>* sW = reinterpret_cast<cplx<T>*>(smem + a.lds_param_bytes + fbytes);
    const int wpb0 = blockDim.x >> 6;
    uint2* sM = reinterpret_cast<uint2*>(sW + wpb0 * a.max_conn);          // flip masks of the cached connections
    unsigned char* sT = smem + ((reinterpret_cast<unsigned char*>(sM + wpb0 * a.max_conn) - smem + 15) & ~15);
    const T* gp = reinterpret_cast<const T*>(a.p);
    const dev_table tab = TAB_LDS ? table_to_lds(a.tab, sT) : a.tab;
    for (int k = threadIdx.x; k < d.np; k += blockDim.x) sp[k] = gp[k];
    if constexpr (!WALK) {
        for (int k = threadIdx.x; k < 2 * d.np; k += blockDim.x) sF[k] = 0ull;
        if (threadIdx.x == 0) *poison = 0;
    }
    __syncthreads();

    const int lane0 = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wpb = blockDim.x >> 6;
    const int64_t gw = (int64_t)bx * wpb + wave;
    const int64_t nw = (int64_t)gx * wpb;
    const uint32_t pmask = (N >= 32) ? 0xffffffffu : ((1u << N) - 1u);
    cplx<T>* wbuf = sW + wave * a.max_conn;
    uint2* mbuf = sM + wave * a.max_conn;

    // P-order offsets of the force vector (compact layout, see ndm_dims)
    const int PB_bla = 0, PB_hla = N, PB_wla = N + nh;
    const int PB_d = d.nlam, PB_ula = d.nlam + na, PB_umu = d.nlam + na + na * N;
    const int PB_bmu = d.nlam + d.npi, PB_hmu = PB_bmu + N, PB_wmu = PB_hmu + nh;
    const int npi = d.npi, R = d.R;
    const T half = T(0.5);

    for (int64_t s = gw; s < a.Ns; s += nw) {
        // an opaque copy of the lane index per sample keeps the compiler from
        // hoisting (and spilling) every lane-dependent address out of the loop
        int lane = lane0;
        asm volatile("" : "+v"(lane));
        const uint64_t key = a.keys[s];
        const uint32_t ks = (uint32_t)key & pmask;
        const uint32_t ke = (uint32_t)(key >> N) & pmask;

        // ---- effective angles of the sample (the forces need only f')
        T th_s[RS], th_e[RS], f_s[RS], f_e[RS], a_s[RS], a_e[RS];
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            const int u = lane + 64 * r;
            th_s[r] = th_e[r] = f_s[r] = f_e[r] = a_s[r] = a_e[r] = T(0);
            if (u < 2 * nh) {
                const bool lam = u < nh;
                const T* w = lam ? sp + d.wla + u : sp + d.wmu + (u - nh);
                const T h0 = lam ? sp[d.hla + u] : sp[d.hmu + u - nh];
                T ts = h0, te = h0;
#pragma unroll
                for (int i = 0; i < NMAX; ++i) {
                    if (i < N) {
                        const T wi = w[nh * i];
                        ts += wi * xv<T>(ks, i);
                        te += wi * xv<T>(ke, i);
                    }
                }
                th_s[r] = ts; th_e[r] = te;
                if constexpr (FORCE) {
                    a_s[r] = act_df<ACT>(ts);
                    a_e[r] = act_df<ACT>(te);
                } else {
                    act_f_df<ACT>(ts, &f_s[r], &a_s[r]);
                    act_f_df<ACT>(te, &f_e[r], &a_e[r]);
                }
            }
        }
        cplx<T> pi_z[PS], pi_f[PS], pi_p[PS];
#pragma unroll
        for (int r = 0; r < PS; ++r) {
            const int q = lane + 64 * r;
            pi_z[r] = pi_f[r] = pi_p[r] = cmk(T(0), T(0));
            if (q < 2 * na) {
                const int au = q < na ? q : q - na;
                T zr = sp[d.dla + au], zi = T(0);
#pragma unroll
                for (int i = 0; i < NMAX; ++i) {
                    if (i < N) {
                        const T xs = xv<T>(ks, i), xe = xv<T>(ke, i);
                        zr += half * sp[d.ula + au + na * i] * (xs + xe);
                        zi += half * sp[d.umu + au + na * i] * (xs - xe);
                    }
                }
                pi_z[r] = cmk(zr, zi);
                if constexpr (FORCE) pi_p[r] = cact_df<ACT>(pi_z[r]);
                else cact_f_df<ACT>(pi_z[r], &pi_f[r], &pi_p[r]);
            }
        }

        if constexpr (!FORCE) {
            // ---- log rho(s) = Gamma_la + i Gamma_mu + Pi
            {
                T re = T(0), im = T(0);
#pragma unroll
                for (int r = 0; r < RS; ++r) {
                    const int u = lane + 64 * r;
                    if (u < nh) re += half * (f_s[r] + f_e[r]);
                    else if (u < 2 * nh) im += half * (f_s[r] - f_e[r]);
                }
#pragma unroll
                for (int r = 0; r < PS; ++r) {
                    const int q = lane + 64 * r;
                    if (q < na) { re += pi_f[r].re; im += pi_f[r].im; }
                }
                if (lane < N) {
                    const T xs = xv<T>(ks, lane), xe = xv<T>(ke, lane);
                    re += half * sp[d.bla + lane] * (xs + xe);
                    im += half * sp[d.bmu + lane] * (xs - xe);
                }
                re = wave_total(re);
                im = wave_total(im);
                if (lane == 0 && a.lp) {
                    T* lp = reinterpret_cast<T*>(a.lp);
                    lp[2 * s] = re;
                    lp[2 * s + 1] = im;
                }
            }

            // ---- compact O(s) row
            if (a.Y) {
                T* y = reinterpret_cast<T*>(a.Y) + (int64_t)s * R;
#pragma unroll
                for (int r = 0; r < RS; ++r) {
                    const int u = lane + 64 * r;
                    if (u < nh) {
                        y[PB_hla + u] = half * (a_s[r] + a_e[r]);
#pragma unroll
                        for (int i = 0; i < NMAX; ++i)
                            if (i < N) y[PB_wla + u + nh * i] = half * (a_s[r] * xv<T>(ks, i) + a_e[r] * xv<T>(ke, i));
                    } else if (u < 2 * nh) {
                        const int h = u - nh;
                        y[npi + PB_hmu + h] = half * (a_s[r] - a_e[r]);
#pragma unroll
                        for (int i = 0; i < NMAX; ++i)
                            if (i < N) y[npi + PB_wmu + h + nh * i] = half * (a_s[r] * xv<T>(ks, i) - a_e[r] * xv<T>(ke, i));
                    }
                }
#pragma unroll
                for (int r = 0; r < PS; ++r) {
                    const int q = lane + 64 * r;
                    if (q < na) {
                        const cplx<T> p = pi_p[r];
                        y[PB_d + q] = p.re;
                        y[npi + PB_d + q] = p.im;
#pragma unroll
                        for (int i = 0; i < NMAX; ++i)
                            if (i < N) {
                                const T sx = half * (xv<T>(ks, i) + xv<T>(ke, i));
                                y[PB_ula + q + na * i] = p.re * sx;
                                y[npi + PB_ula + q + na * i] = p.im * sx;
                            }
                    } else if (q < 2 * na) {
                        const int au = q - na;
                        const cplx<T> p = pi_p[r];
#pragma unroll
                        for (int i = 0; i < NMAX; ++i)
                            if (i < N) {
                                const T dx = half * (xv<T>(ks, i) - xv<T>(ke, i));
                                y[PB_umu + au + na * i] = -p.im * dx;
                                y[npi + PB_umu + au + na * i] = p.re * dx;
                            }
                    }
                }
                if (lane < N) {
                    const T xs = xv<T>(ks, lane), xe = xv<T>(ke, lane);
                    y[PB_bla + lane] = half * (xs + xe);
                    y[npi + PB_bmu + lane] = half * (xs - xe);
                }
            }
        }

        if (!a.do_conn) continue;

        int ncc;
        cplx<T> Wdiag, Lloc;
        if constexpr (FORCE) {
            // the connections of pass 1 from the walk launch
            const int64_t rb = s * (int64_t)a.max_conn;
            ncc = a.cnum[s];
            const T* wd = reinterpret_cast<const T*>(a.cwd) + 2 * s;
            const T* ll = reinterpret_cast<const T*>(a.Lloc) + 2 * s;
            Wdiag = cmk(wd[0], wd[1]);
            Lloc = cmk(ll[0], ll[1]);
            const cplx<T>* rw = reinterpret_cast<const cplx<T>*>(a.crw) + rb;
            for (int c = lane0; c < ncc; c += 64) {
                wbuf[c] = rw[c];
                mbuf[c] = a.crm[rb + c];
            }
            __builtin_amdgcn_wave_barrier();
        } else {
            // ---- pass 1: W_c and L_loc.  The walk (terms spread over the
            // lanes) sums the unchanged connections (W = mel) per lane and
            // caches the changed ones (mel, flip masks) in LDS; those are then
            // evaluated EB at a time: every lane adds its units' share of
            // dlog rho for each of them, one transposed reduction sums all
            // 2 EB shares, and a group of lanes per connection finishes it
            // (bias terms, e^dlog rho, W = mel e^dlog rho), all EB in
            // parallel; W replaces mel in the cache for the force pass
            cplx<T> Lacc = cmk(T(0), T(0));   // this lane's share of L_loc = sum_c W_c
            cplx<T> Wd = cmk(T(0), T(0));     // and of the unchanged connections' sum
            const int nc = table_walk_lanes<T>(
                tab, ks, ke, lane0,
                [&](cplx<T> mel) {
                    // unchanged configuration: exp(0) = 1, W = mel exactly (the
                    // reference evaluates the network anyway, AccumulatorObsGrad.jl:72)
                    Lacc = cadd(Lacc, mel);
                    Wd = cadd(Wd, mel);
                },
                [&](int slot, cplx<T> mel, uint32_t ms, uint32_t me) {
                    if (slot < a.max_conn) {
                        wbuf[slot] = mel;
                        mbuf[slot] = make_uint2(ms, me);
                    }
                });
            Wdiag = cmk(wave_total(Wd.re), wave_total(Wd.im));
            __builtin_amdgcn_wave_barrier();
            ncc = nc < a.max_conn ? nc : a.max_conn;
            if constexpr (WALK && sizeof(T) == 4) {
                // FP32 walk: four changed connections per round, lane i + 16 k
                // taking connection k of the four, its units 16 t + i (the
                // flipped sites' FMAs per lane, no scalar bit walk) and its Pi
                // units 16 t + i; the 16 lanes of a connection then sum its
                // share of dlog rho (the visible biases of a flipped site on
                // that site's lane) and one of them finishes W
                constexpr int TW = 4 * RS, TP = 2 * PS;
                const int il = lane0 & 15, kg = lane0 >> 4;
                T pths[TW], pthe[TW], pfs[TW], pfe[TW];
                int wofs[TW];
#pragma unroll
                for (int t = 0; t < TW; ++t) {
                    const int src = 16 * (t & 3) + il, r = t >> 2, u = 16 * t + il;
                    pths[t] = __shfl(th_s[r], src, 64);
                    pthe[t] = __shfl(th_e[r], src, 64);
                    pfs[t] = __shfl(f_s[r], src, 64);
                    pfe[t] = __shfl(f_e[r], src, 64);
                    wofs[t] = u < nh ? d.wla + u : (u < 2 * nh ? d.wmu + u - nh : d.wla);
                }
                cplx<T> ppz[TP], ppf[TP];
                int aofs[TP];
#pragma unroll
                for (int t = 0; t < TP; ++t) {
                    const int src = 16 * t + il;
                    ppz[t] = cmk(__shfl(pi_z[0].re, src, 64), __shfl(pi_z[0].im, src, 64));
                    ppf[t] = cmk(__shfl(pi_f[0].re, src, 64), __shfl(pi_f[0].im, src, 64));
                    aofs[t] = src < na ? src : na - 1;
                }
                const T bl = il < N ? sp[d.bla + il] : T(0), bm = il < N ? sp[d.bmu + il] : T(0);
                const T cfs = half * cflip<T>(ks, il), cfe = half * cflip<T>(ke, il);
                for (int c0 = 0; c0 < ncc; c0 += 4) {
                    const int c = c0 + kg;
                    uint32_t ms = 0u, me = 0u;
                    if (c < ncc) {
                        const uint2 mm = mbuf[c];
                        ms = mm.x;
                        me = mm.y;
                    }
                    // the flipped sites of each side, in one pass per side:
                    // the lambda / mu units' angles and the Pi units' angles
                    T ths2[TW], the2[TW];
                    cplx<T> z[TP];
#pragma unroll
                    for (int t = 0; t < TW; ++t) { ths2[t] = pths[t]; the2[t] = pthe[t]; }
#pragma unroll
                    for (int t = 0; t < TP; ++t) z[t] = ppz[t];
                    for (uint32_t m = ms; m; m &= m - 1) {
                        const int i = __builtin_ctz(m);
                        const T cf = cflip<T>(ks, i), cc = half * cf;
#pragma unroll
                        for (int t = 0; t < TW; ++t) ths2[t] += cf * sp[wofs[t] + nh * i];
#pragma unroll
                        for (int t = 0; t < TP; ++t) {
                            z[t].re += cc * sp[d.ula + aofs[t] + na * i];
                            z[t].im += cc * sp[d.umu + aofs[t] + na * i];
                        }
                    }
                    for (uint32_t m = me; m; m &= m - 1) {
                        const int i = __builtin_ctz(m);
                        const T cf = cflip<T>(ke, i), cc = half * cf;
#pragma unroll
                        for (int t = 0; t < TW; ++t) the2[t] += cf * sp[wofs[t] + nh * i];
#pragma unroll
                        for (int t = 0; t < TP; ++t) {
                            z[t].re += cc * sp[d.ula + aofs[t] + na * i];
                            z[t].im -= cc * sp[d.umu + aofs[t] + na * i];
                        }
                    }
                    T dre = T(0), dim = T(0);
                    if (ms) {
#pragma unroll
                        for (int t = 0; t < TW; ++t) {
                            if (16 * t < 2 * nh) {
                                const int u = 16 * t + il;
                                const T df = half * (act_f<ACT>(ths2[t]) - pfs[t]);
                                dre += u < nh ? df : T(0);
                                dim += (u >= nh && u < 2 * nh) ? df : T(0);
                            }
                        }
                    }
                    if (me) {
#pragma unroll
                        for (int t = 0; t < TW; ++t) {
                            if (16 * t < 2 * nh) {
                                const int u = 16 * t + il;
                                const T df = half * (act_f<ACT>(the2[t]) - pfe[t]);
                                dre += u < nh ? df : T(0);
                                dim -= (u >= nh && u < 2 * nh) ? df : T(0);
                            }
                        }
                    }
#pragma unroll
                    for (int t = 0; t < TP; ++t) {
                        if (16 * t < na) {
                            const cplx<T> f2 = cact_f<ACT>(z[t]);
                            const bool ok = 16 * t + il < na;
                            dre += ok ? f2.re - ppf[t].re : T(0);
                            dim += ok ? f2.im - ppf[t].im : T(0);
                        }
                    }
                    if ((ms >> il) & 1u) {
                        dre += cfs * bl;
                        dim += cfs * bm;
                    }
                    if ((me >> il) & 1u) {
                        dre += cfe * bl;
                        dim -= cfe * bm;
                    }
                    dre = group_total(dre, 16);
                    dim = group_total(dim, 16);
                    if (c < ncc && il == 0) {
                        T sn, cs;
                        nq_sincos(dim, &sn, &cs);
                        const T er = nq_exp(dre);
                        const cplx<T> W = cmul(wbuf[c], cmk(er * cs, er * sn));
                        wbuf[c] = W;
                        Lacc = cadd(Lacc, W);
                    }
                }
            } else {
                constexpr int EB = est_batch<T>(), LG = 32 / EB;   // LG lanes hold one connection's sums
                // FP32 with one Pi unit per lane (na <= 32): two connections' Pi
                // units per pass (in FP64 the second set of operands would spill)
                constexpr bool PAIR_PI = PS == 1 && sizeof(T) == 4;
                for (int c0 = 0; c0 < ncc; c0 += EB) {
                    T part[2 * EB];
    #pragma unroll
                    for (int j = 0; j < EB; ++j) {
                        T dre = T(0), dim = T(0);
                        if (c0 + j < ncc) {
                            const uint2 mm = mbuf[c0 + j];
                            const uint32_t ms = __builtin_amdgcn_readfirstlane(mm.x), me = __builtin_amdgcn_readfirstlane(mm.y);
    #pragma unroll
                            for (int r = 0; r < RS; ++r) {
                                const int u = lane + 64 * r;
                                if (u < 2 * nh) {
                                    const bool lam = u < nh;
                                    const T* w = lam ? sp + d.wla + u : sp + d.wmu + (u - nh);
                                    T dfs = T(0), dfe = T(0);
                                    if (ms) {
                                        T t = th_s[r];
                                        for (uint32_t m = ms; m; m &= m - 1) {
                                            const int i = __builtin_ctz(m);
                                            t += cflip<T>(ks, i) * w[nh * i];
                                        }
                                        dfs = act_f<ACT>(t) - f_s[r];
                                    }
                                    if (me) {
                                        T t = th_e[r];
                                        for (uint32_t m = me; m; m &= m - 1) {
                                            const int i = __builtin_ctz(m);
                                            t += cflip<T>(ke, i) * w[nh * i];
                                        }
                                        dfe = act_f<ACT>(t) - f_e[r];
                                    }
                                    if (lam) dre += half * (dfs + dfe);
                                    else dim += half * (dfs - dfe);
                                }
                            }
                            if (!PAIR_PI) {
    #pragma unroll
                                for (int r = 0; r < PS; ++r) {
                                    const int q = lane + 64 * r;
                                    if (q < na) {
                                        cplx<T> z = pi_z[r];
                                        for (uint32_t m = ms; m; m &= m - 1) {
                                            const int i = __builtin_ctz(m);
                                            const T c = cflip<T>(ks, i) * half;
                                            z.re += c * sp[d.ula + q + na * i];
                                            z.im += c * sp[d.umu + q + na * i];
                                        }
                                        for (uint32_t m = me; m; m &= m - 1) {
                                            const int i = __builtin_ctz(m);
                                            const T c = cflip<T>(ke, i) * half;
                                            z.re += c * sp[d.ula + q + na * i];
                                            z.im -= c * sp[d.umu + q + na * i];
                                        }
                                        const cplx<T> f2 = cact_f<ACT>(z);
                                        dre += f2.re - pi_f[r].re;
                                        dim += f2.im - pi_f[r].im;
                                    }
                                }
                            }
                        }
                        part[j] = dre;
                        part[EB + j] = dim;
                        if (PAIR_PI && (j & 1)) {
                            // the Pi units of connections j - 1 and j at once: lanes
                            // [0, na) take j - 1 and lanes [na, 2 na), which hold the
                            // same units' angles, take j (flip masks per lane)
                            const bool hi = lane0 >= na;
                            const int cj = c0 + (hi ? j : j - 1);
                            if (lane0 < 2 * na && cj < ncc) {
                                const int q = hi ? lane0 - na : lane0;
                                const uint2 mm = mbuf[cj];
                                cplx<T> z = pi_z[0];
                                for (uint32_t m = mm.x; m; m &= m - 1) {
                                    const int i = __builtin_ctz(m);
                                    const T c = cflip<T>(ks, i) * half;
                                    z.re += c * sp[d.ula + q + na * i];
                                    z.im += c * sp[d.umu + q + na * i];
                                }
                                for (uint32_t m = mm.y; m; m &= m - 1) {
                                    const int i = __builtin_ctz(m);
                                    const T c = cflip<T>(ke, i) * half;
                                    z.re += c * sp[d.ula + q + na * i];
                                    z.im -= c * sp[d.umu + q + na * i];
                                }
                                const cplx<T> f2 = cact_f<ACT>(z);
                                const T vr = f2.re - pi_f[0].re, vi = f2.im - pi_f[0].im;
                                part[j - 1] += hi ? T(0) : vr;
                                part[EB + j - 1] += hi ? T(0) : vi;
                                part[j] += hi ? vr : T(0);
                                part[EB + j] += hi ? vi : T(0);
                            }
                        }
                    }
                    // lane L < 32: sum of dre of connection c0 + L/LG; lane 32 + L: its dim
                    T dre = wave_totals<2 * EB>(part, lane0), dim = dre;
                    plswap<true>(dre, dim);   // lanes 0-31: dre in `dre`, dim in `dim`
                    const int c = c0 + lane0 / LG;
                    if (lane0 < 32 && c < ncc) {
                        const uint2 mm = mbuf[c];
                        for (uint32_t m = mm.x; m; m &= m - 1) {
                            const int i = __builtin_ctz(m);
                            const T cc = cflip<T>(ks, i) * half;
                            dre += cc * sp[d.bla + i];
                            dim += cc * sp[d.bmu + i];
                        }
                        for (uint32_t m = mm.y; m; m &= m - 1) {
                            const int i = __builtin_ctz(m);
                            const T cc = cflip<T>(ke, i) * half;
                            dre += cc * sp[d.bla + i];
                            dim -= cc * sp[d.bmu + i];
                        }
                        T sn, cs;
                        nq_sincos(dim, &sn, &cs);
                        const T er = nq_exp(dre);
                        const cplx<T> W = cmul(wbuf[c], cmk(er * cs, er * sn));
                        if (lane0 % LG == 0) {
                            wbuf[c] = W;
                            Lacc = cadd(Lacc, W);
                        }
                    }
                }
            }
            Lloc = cmk(wave_total(Lacc.re), wave_total(Lacc.im));
            __builtin_amdgcn_wave_barrier();

            if (lane == 0 && a.Lloc) {
                T* lo = reinterpret_cast<T*>(a.Lloc);
                lo[2 * s] = Lloc.re;
                lo[2 * s + 1] = Lloc.im;
            }
            if (WALK && a.rec_mode == 1) {
                // keep the changed connections (W, flip masks) for the forces launch
                const int64_t rb = s * (int64_t)a.max_conn;
                cplx<T>* rw = reinterpret_cast<cplx<T>*>(a.crw) + rb;
                for (int c = lane0; c < ncc; c += 64) {
                    rw[c] = wbuf[c];
                    a.crm[rb + c] = mbuf[c];
                }
                if (lane0 == 0) {
                    a.cnum[s] = ncc;
                    T* wd = reinterpret_cast<T*>(a.cwd) + 2 * s;
                    wd[0] = Wdiag.re;
                    wd[1] = Wdiag.im;
                }
            }
        }
        if constexpr (!WALK) {
            // ---- force pass: Re(V W_c O_k(c)) summed over the connections of
            // this sample in registers, then added to the workgroup's exact
            // accumulator once
            constexpr bool PAIR_PI = PS == 1 && sizeof(T) == 4;
            constexpr int NV = GRAD ? 2 : 1;
            for (int vi = 0; vi < NV; ++vi) {
                // Re(1 G) = Re G, Re(-i G) = Im G in the parity mode
                const cplx<double> Vd = GRAD ? (vi == 0 ? cmk(1.0, 0.0) : cmk(0.0, -1.0))
                                             : cmk((double)Lloc.re * a.invNs, -(double)Lloc.im * a.invNs);
                const cplx<T> V = cmk((T)Vd.re, (T)Vd.im);
                // per unit: weights of the dense part (e_s, e_e, rho1, rho2) and the
                // flipped-site shares (fw: w entries of the lambda / mu unit; fur,
                // fui: u_la, u_mu entries of the Pi unit, this half's sites)
                T e_s[RS], e_e[RS];
                T rho1[PS], rho2[PS];
                vN fw[RS], fur[PS], fui[PS];
#pragma unroll
                for (int r = 0; r < RS; ++r) {
                    e_s[r] = e_e[r] = T(0);
                    fw[r] = vN(T(0));
                }
#pragma unroll
                for (int r = 0; r < PS; ++r) {
                    rho1[r] = rho2[r] = T(0);
                    fur[r] = fui[r] = vN(T(0));
                }
                T fbl = T(0), fbm = T(0);   // b_la, b_mu of site `lane` (lane < N)
                // unchanged configurations contribute V W = V mel with O(c) = O(s):
                // only the dense parts, added once below
                const cplx<T> VWdiag = cmul(V, Wdiag);
                // one changed connection's terms; p2in: its Pi units' f'(z) when
                // already known (paired evaluation), else recomputed here
                auto conn = [&](int c, bool have_p2, cplx<T> p2in) {
                    const cplx<T> VW = cmul(V, wbuf[c]);
                    const uint2 mm = mbuf[c];
                    const uint32_t ms = __builtin_amdgcn_readfirstlane(mm.x), me = __builtin_amdgcn_readfirstlane(mm.y);
#pragma unroll
                    for (int r = 0; r < RS; ++r) {
                        const int u = lane + 64 * r;
                        if (u < 2 * nh) {
                            const bool lam = u < nh;
                            const int h = lam ? u : u - nh;
                            const T* w = lam ? sp + d.wla + u : sp + d.wmu + h;
                            const T gam = lam ? VW.re : -VW.im;
                            const T seta = lam ? T(1) : T(-1);
                            T as2 = a_s[r], ae2 = a_e[r];
                            if (ms) {
                                T t = th_s[r];
                                for (uint32_t m = ms; m; m &= m - 1) {
                                    const int i = __builtin_ctz(m);
                                    t += cflip<T>(ks, i) * w[nh * i];
                                }
                                as2 = act_df<ACT>(t);
                            }
                            if (me) {
                                T t = th_e[r];
                                for (uint32_t m = me; m; m &= m - 1) {
                                    const int i = __builtin_ctz(m);
                                    t += cflip<T>(ke, i) * w[nh * i];
                                }
                                ae2 = act_df<ACT>(t);
                            }
                            e_s[r] += gam * as2;
                            e_e[r] += seta * gam * ae2;
                            const T gs = half * gam * as2, ge = half * seta * gam * ae2;
                            for (uint32_t m = ms; m; m &= m - 1) {
                                const int i = __builtin_ctz(m);
                                fw[r][i] += gs * cflip<T>(ks, i);
                            }
                            for (uint32_t m = me; m; m &= m - 1) {
                                const int i = __builtin_ctz(m);
                                fw[r][i] += ge * cflip<T>(ke, i);
                            }
                        }
                    }
#pragma unroll
                    for (int r = 0; r < PS; ++r) {
                        const int q = lane + 64 * r;
                        if (q < 2 * na) {
                            const bool rowhalf = q < na;
                            const int au = rowhalf ? q : q - na;
                            cplx<T> p2 = have_p2 ? p2in : pi_p[r];
                            if (!have_p2 && (ms | me)) {
                                cplx<T> z = pi_z[r];
                                for (uint32_t m = ms; m; m &= m - 1) {
                                    const int i = __builtin_ctz(m);
                                    const T cc = cflip<T>(ks, i) * half;
                                    z.re += cc * sp[d.ula + au + na * i];
                                    z.im += cc * sp[d.umu + au + na * i];
                                }
                                for (uint32_t m = me; m; m &= m - 1) {
                                    const int i = __builtin_ctz(m);
                                    const T cc = cflip<T>(ke, i) * half;
                                    z.re += cc * sp[d.ula + au + na * i];
                                    z.im -= cc * sp[d.umu + au + na * i];
                                }
                                p2 = cact_df<ACT>(z);
                            }
                            const cplx<T> vp = cmul(VW, p2);
                            if (rowhalf) {
                                rho1[r] += vp.re;
                                rho2[r] += vp.im;
                            }
                            // the row half takes the row flips (u_mu share -cc Im),
                            // the column half the column flips (+cc Im)
                            const uint32_t mq = rowhalf ? ms : me, kq = rowhalf ? ks : ke;
                            const T vim = rowhalf ? -vp.im : vp.im;
                            for (uint32_t m = ms | me; m; m &= m - 1) {
                                const int i = __builtin_ctz(m);
                                const T cc = ((mq >> i) & 1u) ? half * cflip<T>(kq, i) : T(0);
                                fur[r][i] += cc * vp.re;
                                fui[r][i] += cc * vim;
                            }
                        }
                    }
                    if (lane < N) {
                        if ((ms >> lane) & 1u) {
                            const T cc = half * cflip<T>(ks, lane);
                            fbl += cc * VW.re;
                            fbm -= cc * VW.im;
                        }
                        if ((me >> lane) & 1u) {
                            const T cc = half * cflip<T>(ke, lane);
                            fbl += cc * VW.re;
                            fbm += cc * VW.im;
                        }
                    }
                };
                if (PAIR_PI) {
                    // connections in pairs: lanes [0, na) compute f'(z) of c's Pi
                    // units, lanes [na, 2 na) (the same units) that of c + 1, then
                    // each half takes the other's values by a lane exchange
                    for (int c = 0; c < ncc; c += 2) {
                        const bool hi = lane0 >= na;
                        const int cj = c + (hi ? 1 : 0);
                        cplx<T> p2 = pi_p[0];
                        if (lane0 < 2 * na && cj < ncc) {
                            const uint2 mm = mbuf[cj];
                            if (mm.x | mm.y) {
                                const int au = hi ? lane0 - na : lane0;
                                cplx<T> z = pi_z[0];
                                for (uint32_t m = mm.x; m; m &= m - 1) {
                                    const int i = __builtin_ctz(m);
                                    const T cc = cflip<T>(ks, i) * half;
                                    z.re += cc * sp[d.ula + au + na * i];
                                    z.im += cc * sp[d.umu + au + na * i];
                                }
                                for (uint32_t m = mm.y; m; m &= m - 1) {
                                    const int i = __builtin_ctz(m);
                                    const T cc = cflip<T>(ke, i) * half;
                                    z.re += cc * sp[d.ula + au + na * i];
                                    z.im -= cc * sp[d.umu + au + na * i];
                                }
                                p2 = cact_df<ACT>(z);
                            }
                        }
                        const int src = lane0 < 2 * na ? (hi ? lane0 - na : lane0 + na) : lane0;
                        const cplx<T> po = cmk(__shfl(p2.re, src, 64), __shfl(p2.im, src, 64));
                        conn(c, true, hi ? po : p2);
                        if (c + 1 < ncc) conn(c + 1, true, hi ? p2 : po);
                    }
                } else {
                    for (int c = 0; c < ncc; ++c) conn(c, false, pi_p[0]);
                }

                // unchanged configurations contribute O(s) with weight sum VWdiag
#pragma unroll
                for (int r = 0; r < RS; ++r) {
                    const int u = lane + 64 * r;
                    if (u < 2 * nh) {
                        const bool lam = u < nh;
                        const T gam = lam ? VWdiag.re : -VWdiag.im;
                        e_s[r] += gam * a_s[r];
                        e_e[r] += (lam ? gam : -gam) * a_e[r];
                    }
                }
#pragma unroll
                for (int r = 0; r < PS; ++r) {
                    const int q = lane + 64 * r;
                    if (q < na) {
                        const cplx<T> vp = cmul(VWdiag, pi_p[r]);
                        rho1[r] += vp.re;
                        rho2[r] += vp.im;
                    }
                }
                if (lane < N) {
                    const cplx<T> VL = cmul(V, Lloc);
                    const T xs = xv<T>(ks, lane), xe = xv<T>(ke, lane);
                    fbl += half * VL.re * (xs + xe);
                    fbm -= half * VL.im * (xs - xe);
                }

                // this sample's entries: dense parts (x_s times the per-unit
                // sums) plus the flipped-site shares, into the exact accumulator
                // (the opaque copy of the unit index keeps the addresses from
                // being hoisted out of the sample loop and spilled)
#pragma unroll
                for (int r = 0; r < RS; ++r) {
                    int u = lane + 64 * r;
                    asm volatile("" : "+v"(u));
                    if (u < 2 * nh) {
                        const bool lam = u < nh;
                        const int h = lam ? u : u - nh;
                        fx_add(sF, (lam ? PB_hla : PB_hmu) + h, (double)(half * (e_s[r] + e_e[r])), poison);
                        const int pb = lam ? PB_wla : PB_wmu;
#pragma unroll
                        for (int i = 0; i < NMAX; ++i)
                            if (i < N)
                                fx_add(sF, pb + h + nh * i,
                                       (double)(fw[r][i] + half * (e_s[r] * xv<T>(ks, i) + e_e[r] * xv<T>(ke, i))), poison);
                    }
                }
#pragma unroll
                for (int r = 0; r < PS; ++r) {
                    int q = lane + 64 * r;
                    asm volatile("" : "+v"(q));
                    if (q < 2 * na) {
                        const bool rowhalf = q < na;
                        const int au = rowhalf ? q : q - na;
                        if (rowhalf) fx_add(sF, PB_d + au, (double)rho1[r], poison);
#pragma unroll
                        for (int i = 0; i < NMAX; ++i)
                            if (i < N) {
                                T ur = fur[r][i], ui = fui[r][i];
                                if (rowhalf) {
                                    const T xs = xv<T>(ks, i), xe = xv<T>(ke, i);
                                    ur += half * rho1[r] * (xs + xe);
                                    ui -= half * rho2[r] * (xs - xe);
                                }
                                fx_add(sF, PB_ula + au + na * i, (double)ur, poison);
                                fx_add(sF, PB_umu + au + na * i, (double)ui, poison);
                            }
                    }
                }
                if (lane < N) {
                    fx_add(sF, PB_bla + lane, (double)fbl, poison);
                    fx_add(sF, PB_bmu + lane, (double)fbm, poison);
                }
                if constexpr (GRAD) {   // one wave per block in this mode: sF belongs to this sample
                    __syncthreads();
                    double* g = reinterpret_cast<double*>(a.G) + (int64_t)s * 2 * d.np;
                    const bool bad = *poison != 0;
                    for (int k = lane; k < d.np; k += 64) {
                        g[2 * p2ref(k, d) + vi] = bad ? __builtin_nan("") : fx_value(sF, k);
                        sF[2 * k] = 0ull;
                        sF[2 * k + 1] = 0ull;
                    }
                    __syncthreads();
                    if (lane == 0) *poison = 0;
                    __syncthreads();
                }
            }
        }
    }

    if constexpr (FORCE) {
        __syncthreads();
        double* out = a.Fpart + (int64_t)bx * d.np;
        const bool bad = *poison != 0;
        for (int k = threadIdx.x; k < d.np; k += blockDim.x) out[k] = bad ? __builtin_nan("") : fx_value(sF, k);
    }
}

// ---- the forces launch of the FP32 network on the matrix cores
// Every term of the force is an outer product over (unit, site).  A changed
// connection c (and the unchanged configurations together, as one more with
// W = their mel sum and no flips) adds, for each side (row configuration s,
// column configuration e, x_c its +-1 configuration after the flips):
//   lambda / mu weights  A_c(u) x_c(i):  A_c(u) = (1/2) g_u f'(theta_u(c)),
//                        g_u = Re(V W_c) (lambda), -Im(V W_c) (mu), sign -
//                        on the mu units' column side
//   Pi weights           Re(V W_c f'(z_a(c))) (x^s_c + x^e_c)(i) / 2 (u_la)
//                        -Im(V W_c f'(z_a(c))) (x^s_c - x^e_c)(i) / 2 (u_mu)
// and the biases the same sums with x = 1 (hidden units, d) or with the unit
// factor 1 (visible biases).  So the weight forces are products [units x
// terms] [terms x sites] over all terms of a wave's samples, on
// v_mfma_f32_16x16x4f32: four connections at a time, lane l = i + 16 k
// taking connection k of the four, the units 16 t + i of it (operand A of
// tile t: unit row i, term k) and site i of its configuration (operand B:
// term k, site column i).  Each lane re-evaluates the changed angles of its
// units for its own connection (one FMA per unit and flipped site, no
// register indexing), the bias sums stay in per-lane registers, and the
// products accumulate in FP32 over at most kForceFlush samples before they
// go into the workgroup's exact accumulator (fx_add), so the result does not
// depend on the order in which waves arrive.
//
// The tiles come in three sizes (class 0, 1, 2: TW = 2, 4, 8 tiles of
// lambda / mu units, TP = 1, 2, 4 of Pi units), the smallest that holds 2 nh
// and na.  The launch stages the parameters it needs in its own LDS layout,
// zero padded to whole tiles (SW = 16 TW units, SP = 16 TP Pi units), so no
// lane ever needs a bounds test on its units:
//   HW [N + 1][SW]  row 0 the hidden biases (lambda units, then mu), row
//                   1 + i the weights of site i
//   PU [N][SP][2]   (u_la, u_mu) of Pi unit a, site i
//   PD [SP]         d_la
typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int kForceFlush = 8;

// tile class of a network (-1: more than 64 hidden or Pi units)
__host__ __device__ inline int force_class(const ndm_dims& d) {
    const int k = max((2 * d.nh + 31) / 32, (d.na + 15) / 16);
    return k <= 1 ? 0 : (k <= 2 ? 1 : (k <= 4 ? 2 : -1));
}
// floats of the staged parameters of the FP32 forces launch
__host__ __device__ inline int force_stage_floats(const ndm_dims& d) {
    const int c = force_class(d) < 0 ? 2 : force_class(d);
    return (32 << c) * (d.N + 1) + (16 << c) * (2 * d.N + 1);
}

// f' of a real unit: the logistic function as 1 / (1 + e^-x) (the limits 0
// and 1 come out of the reciprocal of inf and of 1, no select), tanh for logcosh
template <int ACT>
__device__ __forceinline__ float force_df(float x) {
    if constexpr (ACT == NQ_ACT_SOFTPLUS) return nq_rcp(1.0f + nq_exp(-x));
    else return act_df<ACT>(x);
}

template <int NMAX, int CLS, int ACT>
__device__ __forceinline__ void force_mma_body(const est_args& a, int bx, int gx) {
    using T = float;
    static_assert(NMAX <= 16, "the sites fill one 16-column tile");
    constexpr int TW = 2 << CLS, TP = 1 << CLS;   // 16-unit tiles: lambda / mu units (2 nh), Pi units (na)
    constexpr int SW = 16 * TW, SP = 16 * TP, RS = (SW + 63) / 64;
    extern __shared__ __align__(16) unsigned char smem[];
    const ndm_dims d = a.d;
    const int N = d.N, nh = d.nh, na = d.na;
    const int pbytes = max(a.lds_param_bytes, (4 * force_stage_floats(d) + 15) & ~15);
    T* HW = reinterpret_cast<T*>(smem);
    T* PU = HW + SW * (N + 1);
    T* PD = PU + 2 * SP * N;
    unsigned long long* sF = reinterpret_cast<unsigned long long*>(smem + pbytes);
    int* poison = reinterpret_cast<int*>(sF + 2 * d.np);
    cplx<T>* sW = reinterpret_cast<cplx<T>*>(smem + pbytes + fx_bytes(d.np));
    const int wpb = blockDim.x >> 6;
    uint2* sM = reinterpret_cast<uint2*>(sW + wpb * a.max_conn);
    const T* gp = reinterpret_cast<const T*>(a.p);
    for (int k = threadIdx.x; k < SW * (N + 1); k += blockDim.x) {
        const int r = k / SW, u = k - r * SW;
        const int src = u < nh ? (r == 0 ? d.hla + u : d.wla + u + nh * (r - 1))
                               : (r == 0 ? d.hmu + u - nh : d.wmu + u - nh + nh * (r - 1));
        HW[k] = u < 2 * nh ? gp[src] : T(0);
    }
    for (int k = threadIdx.x; k < SP * N; k += blockDim.x) {
        const int i = k / SP, au = k - i * SP;
        const bool ok = au < na;
        PU[2 * k] = ok ? gp[d.ula + au + na * i] : T(0);
        PU[2 * k + 1] = ok ? gp[d.umu + au + na * i] : T(0);
    }
    for (int k = threadIdx.x; k < SP; k += blockDim.x) PD[k] = k < na ? gp[d.dla + k] : T(0);
    for (int k = threadIdx.x; k < 2 * d.np; k += blockDim.x) sF[k] = 0ull;
    if (threadIdx.x == 0) *poison = 0;
    __syncthreads();

    const int lane0 = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t gw = (int64_t)bx * wpb + wave, nw = (int64_t)gx * wpb;
    const uint32_t pmask = (N >= 32) ? 0xffffffffu : ((1u << N) - 1u);
    cplx<T>* wbuf = sW + wave * a.max_conn;
    uint2* mbuf = sM + wave * a.max_conn;
    const T half = T(0.5);
    const int il = lane0 & 15, kg = lane0 >> 4;

    f4v Fw[TW], Fr[TP], Fi[TP];
    T hb[TW], db[TP], vbl = T(0), vbm = T(0);
#pragma unroll
    for (int t = 0; t < TW; ++t) { Fw[t] = f4v{0.f, 0.f, 0.f, 0.f}; hb[t] = T(0); }
#pragma unroll
    for (int t = 0; t < TP; ++t) { Fr[t] = Fi[t] = f4v{0.f, 0.f, 0.f, 0.f}; db[t] = T(0); }

    int nacc = 0;
    for (int64_t s = gw; s < a.Ns; s += nw) {
        // an opaque copy of the lane index per sample keeps the compiler from
        // hoisting (and spilling) every lane-dependent address out of the loop
        int lane = lane0;
        asm volatile("" : "+v"(lane));
        const uint64_t key = a.keys[s];
        const uint32_t ks = (uint32_t)key & pmask;
        const uint32_t ke = (uint32_t)(key >> N) & pmask;

        // angles and f' of the sample, one unit per lane (u = lane + 64 r; Pi
        // unit lane), then moved into the product layout (units 16 t + i);
        // padding units have zero angles
        T ts1[RS], te1[RS], as1[RS], ae1[RS];
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            const int u = lane + 64 * r;
            T ts = HW[u], te = ts;
#pragma unroll
            for (int i = 0; i < NMAX; ++i) {
                if (i < N) {
                    const T wi = HW[SW * (1 + i) + u];
                    ts += wi * xv<T>(ks, i);
                    te += wi * xv<T>(ke, i);
                }
            }
            ts1[r] = ts;
            te1[r] = te;
            as1[r] = force_df<ACT>(ts);
            ae1[r] = force_df<ACT>(te);
        }
        cplx<T> pz1 = cmk(T(0), T(0));
        if (lane < SP) {
            T zr = PD[lane], zi = T(0);
#pragma unroll
            for (int i = 0; i < NMAX; ++i) {
                if (i < N) {
                    const T xs = xv<T>(ks, i), xe = xv<T>(ke, i);
                    const float2 uu = *reinterpret_cast<const float2*>(PU + 2 * (SP * i + lane));
                    zr += half * uu.x * (xs + xe);
                    zi += half * uu.y * (xs - xe);
                }
            }
            pz1 = cmk(zr, zi);
        }
        T ths[TW], the[TW], as[TW], ae[TW];
#pragma unroll
        for (int t = 0; t < TW; ++t) {
            const int src = 16 * (t & 3) + il, r = t >> 2;
            ths[t] = __shfl(ts1[r], src, 64);
            the[t] = __shfl(te1[r], src, 64);
            as[t] = __shfl(as1[r], src, 64);
            ae[t] = __shfl(ae1[r], src, 64);
        }
        cplx<T> pz[TP];
#pragma unroll
        for (int t = 0; t < TP; ++t) pz[t] = cmk(__shfl(pz1.re, 16 * t + il, 64), __shfl(pz1.im, 16 * t + il, 64));

        if (a.do_conn) {
            // the connections of pass 1 from the walk launch
            const int64_t rb = s * (int64_t)a.max_conn;
            const int ncc = a.cnum[s];
            const T* wd = reinterpret_cast<const T*>(a.cwd) + 2 * s;
            const T* ll = reinterpret_cast<const T*>(a.Lloc) + 2 * s;
            const cplx<T> Wdiag = cmk(wd[0], wd[1]);
            const cplx<T> V = cmk((T)((double)ll[0] * a.invNs), (T)(-(double)ll[1] * a.invNs));
            const cplx<T>* rw = reinterpret_cast<const cplx<T>*>(a.crw) + rb;
            for (int c = lane0; c < ncc; c += 64) {
                wbuf[c] = rw[c];
                mbuf[c] = a.crm[rb + c];
            }
            __builtin_amdgcn_wave_barrier();

            // four connections per round; index ncc is the unchanged configurations
            for (int c0 = 0; c0 <= ncc; c0 += 4) {
                const int c = c0 + kg;
                cplx<T> W = cmk(T(0), T(0));
                uint32_t ms = 0u, me = 0u;
                if (c < ncc) {
                    W = wbuf[c];
                    const uint2 mm = mbuf[c];
                    ms = mm.x;
                    me = mm.y;
                } else if (c == ncc) {
                    W = Wdiag;
                }
                const cplx<T> VW = cmul(V, W);

                // the flipped sites of each side, in one pass per side: the
                // lambda / mu units' angles and the Pi units' angles
                T ts[TW], te[TW];
                cplx<T> z[TP];
#pragma unroll
                for (int t = 0; t < TW; ++t) { ts[t] = ths[t]; te[t] = the[t]; }
#pragma unroll
                for (int t = 0; t < TP; ++t) z[t] = pz[t];
                for (uint32_t m = ms; m; m &= m - 1) {
                    const int i = __builtin_ctz(m);
                    const T cf = cflip<T>(ks, i), cc = half * cf;
                    const T* w = HW + SW * (1 + i) + il;
                    const float2* pu = reinterpret_cast<const float2*>(PU) + SP * i + il;
#pragma unroll
                    for (int t = 0; t < TW; ++t) ts[t] += cf * w[16 * t];
#pragma unroll
                    for (int t = 0; t < TP; ++t) {
                        const float2 uu = pu[16 * t];
                        z[t].re += cc * uu.x;
                        z[t].im += cc * uu.y;
                    }
                }
                for (uint32_t m = me; m; m &= m - 1) {
                    const int i = __builtin_ctz(m);
                    const T cf = cflip<T>(ke, i), cc = half * cf;
                    const T* w = HW + SW * (1 + i) + il;
                    const float2* pu = reinterpret_cast<const float2*>(PU) + SP * i + il;
#pragma unroll
                    for (int t = 0; t < TW; ++t) te[t] += cf * w[16 * t];
#pragma unroll
                    for (int t = 0; t < TP; ++t) {
                        const float2 uu = pu[16 * t];
                        z[t].re += cc * uu.x;
                        z[t].im -= cc * uu.y;
                    }
                }
                // f' at the connection's angles (the sample's own on an unflipped side)
                T As[TW], Ae[TW];
                if (ms) {
#pragma unroll
                    for (int t = 0; t < TW; ++t) As[t] = force_df<ACT>(ts[t]);
                } else {
#pragma unroll
                    for (int t = 0; t < TW; ++t) As[t] = as[t];
                }
                if (me) {
#pragma unroll
                    for (int t = 0; t < TW; ++t) Ae[t] = force_df<ACT>(te[t]);
                } else {
#pragma unroll
                    for (int t = 0; t < TW; ++t) Ae[t] = ae[t];
                }
                const T gl = half * VW.re, gm = -half * VW.im;
                const bool sok = il < N;
                const T Bs = sok ? xv<T>(ks ^ ms, il) : T(0), Be = sok ? xv<T>(ke ^ me, il) : T(0);
#pragma unroll
                for (int t = 0; t < TW; ++t) {
                    const int u = 16 * t + il;
                    const T gs = u < nh ? gl : (u < 2 * nh ? gm : T(0));
                    const T ge = u < nh ? gl : -gs;
                    const T xa = gs * As[t], ya = ge * Ae[t];
                    hb[t] += xa + ya;
                    Fw[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, Bs, Fw[t], 0, 0, 0);
                    Fw[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ya, Be, Fw[t], 0, 0, 0);
                }

                // Pi units: f'(z) at the connection's angle (for the unchanged
                // configurations the sample's own)
                const T Bp = half * (Bs + Be), Bm = half * (Bs - Be);
#pragma unroll
                for (int t = 0; t < TP; ++t) {
                    const cplx<T> vp = cmul(VW, cact_df<ACT>(z[t]));
                    const bool ok = 16 * t + il < na;
                    const T ar = ok ? vp.re : T(0), ai = ok ? -vp.im : T(0);
                    db[t] += ar;
                    Fr[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, Bp, Fr[t], 0, 0, 0);
                    Fi[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, Bm, Fi[t], 0, 0, 0);
                }
                vbl += VW.re * Bp;
                vbm -= VW.im * Bm;
            }
        }

        // every kForceFlush samples and after the wave's last one: the terms
        // into the exact accumulator, product entry (row 4 (lane / 16) + v,
        // column lane % 16) of each tile and the bias sums over the four
        // connection groups (lanes i + 16 k, k = 0..3)
        if (++nacc == kForceFlush || s + nw >= a.Ns) {
            nacc = 0;
            // the entry indexes from the opaque lane copy, so that they are
            // not hoisted out of the sample loop and spilled
            const int il = lane & 15, kg = lane >> 4;
            const int PB_bla = 0, PB_hla = N, PB_wla = N + nh;
            const int PB_d = d.nlam, PB_ula = d.nlam + na, PB_umu = d.nlam + na + na * N;
            const int PB_bmu = d.nlam + d.npi, PB_hmu = PB_bmu + N, PB_wmu = PB_hmu + nh;
#pragma unroll
            for (int t = 0; t < TW; ++t) {
                if (16 * t < 2 * nh) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int u = 16 * t + 4 * kg + v;
                        if (u < 2 * nh && il < N)
                            fx_add(sF, (u < nh ? PB_wla + u : PB_wmu + u - nh) + nh * il, (double)Fw[t][v], poison);
                    }
                    Fw[t] = f4v{0.f, 0.f, 0.f, 0.f};
                    T h = hb[t];
                    h += __shfl_xor(h, 16, 64);
                    h += __shfl_xor(h, 32, 64);
                    const int u = 16 * t + il;
                    if (kg == 0 && u < 2 * nh) fx_add(sF, u < nh ? PB_hla + u : PB_hmu + u - nh, (double)h, poison);
                    hb[t] = T(0);
                }
            }
#pragma unroll
            for (int t = 0; t < TP; ++t) {
                if (16 * t < na) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int au = 16 * t + 4 * kg + v;
                        if (au < na && il < N) {
                            fx_add(sF, PB_ula + au + na * il, (double)Fr[t][v], poison);
                            fx_add(sF, PB_umu + au + na * il, (double)Fi[t][v], poison);
                        }
                    }
                    Fr[t] = Fi[t] = f4v{0.f, 0.f, 0.f, 0.f};
                    T h = db[t];
                    h += __shfl_xor(h, 16, 64);
                    h += __shfl_xor(h, 32, 64);
                    if (kg == 0 && 16 * t + il < na) fx_add(sF, PB_d + 16 * t + il, (double)h, poison);
                    db[t] = T(0);
                }
            }
            T bl = vbl, bm = vbm;
            bl += __shfl_xor(bl, 16, 64);
            bl += __shfl_xor(bl, 32, 64);
            bm += __shfl_xor(bm, 16, 64);
            bm += __shfl_xor(bm, 32, 64);
            if (kg == 0 && il < N) {
                fx_add(sF, PB_bla + il, (double)bl, poison);
                fx_add(sF, PB_bmu + il, (double)bm, poison);
            }
            vbl = vbm = T(0);
        }
    }

    __syncthreads();
    double* out = a.Fpart + (int64_t)bx * d.np;
    const bool bad = *poison != 0;
    for (int k = threadIdx.x; k < d.np; k += blockDim.x) out[k] = bad ? __builtin_nan("") : fx_value(sF, k);
}

// the FP32 forces launch of one tile class (and its packed form)
template <int NMAX, int CLS, int ACT>
__global__ __launch_bounds__(512, CLS < 2 ? 4 : 2) void k_force_mma(est_args a) {
    force_mma_body<NMAX, CLS, ACT>(a, blockIdx.x, gridDim.x);
}
template <int NMAX, int CLS, int ACT>
__global__ __launch_bounds__(512, CLS < 2 ? 4 : 2) void k_force_mma_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = M[b.x & 0xffff];
    force_mma_body<NMAX, CLS, ACT>(m.ef, (int)(b.y & 0xffff), m.force_grid);
}

#define NQ_EST_BOUNDS __launch_bounds__(512, (sizeof(T) == 4 && RS == 1 && KIND != EK_GRAD) ? 4 : 2)

template <typename T, int NMAX, int RS, int PS, bool TAB_LDS, int KIND, int ACT>
__global__ NQ_EST_BOUNDS void k_estimate(est_args a) {
    estimate_body<T, NMAX, RS, PS, TAB_LDS, KIND, ACT>(a, blockIdx.x, gridDim.x);
}

// packed: the members' walk (KIND = EK_WALK) or force launches in one grid
template <typename T, int NMAX, int RS, int PS, bool TAB_LDS, int KIND, int ACT>
__global__ NQ_EST_BOUNDS void k_estimate_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const nq_pk_member& m = M[b.x & 0xffff];
    const est_args a = KIND == EK_FORCE ? m.ef : m.ew;
    estimate_body<T, NMAX, RS, PS, TAB_LDS, KIND, ACT>(a, (int)(b.y & 0xffff),
                                                      KIND == EK_FORCE ? m.force_grid : m.est_grid);
}
#undef NQ_EST_BOUNDS

// waves per workgroup of a launch: 8 share one copy of the parameters, the
// force accumulator and the table (two blocks per CU while that fits in half
// the LDS: 4 waves per SIMD); otherwise 4.  The parity mode (per-sample
// gradients) runs one wave per block.
inline int est_waves(const est_args& a, size_t elt) {
    if (a.G) return 1;
    if (a.waves == 4) return 4;
    if (a.waves == 8 && est_lds_bytes(a, elt, 8) <= 160 * 1024) return 8;
    return est_lds_bytes(a, elt, 8) > 80 * 1024 ? 4 : 8;
}

template <typename T, int NMAX, int RS, int PS, int ACT>
hipError_t launch_t(const est_args& a0, int grid, hipStream_t st) {
    est_args a = a0;
    const int kind = a.G ? EK_GRAD : (a.rec_mode == 2 ? EK_FORCE : EK_WALK);
    if (kind == EK_WALK && a.Fpart) return hipErrorInvalidValue;   // forces take a launch of their own
    if (kind == EK_FORCE) a.tab_lds = 0;                             // the records replace the walk
    const int wpb = est_waves(a, sizeof(T));
    const size_t lds = est_lds_bytes(a, sizeof(T), wpb);
    if (kind == EK_FORCE) {
        if constexpr (sizeof(T) == 4) {
            const int cls = force_class(a.d);
            if (cls == 0)
                hipLaunchKernelGGL((k_force_mma<NMAX, 0, ACT>), dim3(grid), dim3(64 * wpb), lds, st, a);
            else if (cls == 1)
                hipLaunchKernelGGL((k_force_mma<NMAX, 1, ACT>), dim3(grid), dim3(64 * wpb), lds, st, a);
            else if (cls == 2)
                hipLaunchKernelGGL((k_force_mma<NMAX, 2, ACT>), dim3(grid), dim3(64 * wpb), lds, st, a);
            else
                return hipErrorInvalidValue;
        } else {
            hipLaunchKernelGGL((k_estimate<T, NMAX, RS, PS, false, EK_FORCE, ACT>), dim3(grid), dim3(64 * wpb), lds, st,
                               a);
        }
    } else if (kind == EK_GRAD) {
        if (a.tab_lds)
            hipLaunchKernelGGL((k_estimate<T, NMAX, RS, PS, true, EK_GRAD, ACT>), dim3(grid), dim3(64), lds, st, a);
        else
            hipLaunchKernelGGL((k_estimate<T, NMAX, RS, PS, false, EK_GRAD, ACT>), dim3(grid), dim3(64), lds, st, a);
    } else {
        if (a.tab_lds)
            hipLaunchKernelGGL((k_estimate<T, NMAX, RS, PS, true, EK_WALK, ACT>), dim3(grid), dim3(64 * wpb), lds, st, a);
        else
            hipLaunchKernelGGL((k_estimate<T, NMAX, RS, PS, false, EK_WALK, ACT>), dim3(grid), dim3(64 * wpb), lds, st, a);
    }
    return hipGetLastError();
}

template <typename T, int NMAX, int ACT>
hipError_t launch_n(const est_args& a, int grid, hipStream_t st) {
    const int rs = (2 * a.d.nh + 63) / 64, ps = (2 * a.d.na + 63) / 64;
    if (rs <= 1 && ps <= 1) return launch_t<T, NMAX, 1, 1, ACT>(a, grid, st);
    if (rs <= 2 && ps <= 2) return launch_t<T, NMAX, 2, 2, ACT>(a, grid, st);
    return hipErrorInvalidValue;
}

// every launch of one precision
template <typename T>
hipError_t launch_prec(const est_args& a, int grid, hipStream_t st) {
    if (a.d.act == NQ_ACT_LOGCOSH) {
        if (a.d.N <= 8) return launch_n<T, 8, NQ_ACT_LOGCOSH>(a, grid, st);
        return launch_n<T, 16, NQ_ACT_LOGCOSH>(a, grid, st);
    }
    if (a.d.N <= 8) return launch_n<T, 8, NQ_ACT_SOFTPLUS>(a, grid, st);
    return launch_n<T, 16, NQ_ACT_SOFTPLUS>(a, grid, st);
}

// packed launches of one precision; variant bits: 0 force (else walk), 1 N > 8
// (NMAX 16), 2 table in LDS, 3 logcosh, 5 tile class 0 of the FP32 forces
// (the walk and force kinds, one unit row and one Pi row per lane only)
template <typename T, int NMAX, int ACT>
hipError_t pk_launch_n(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st) {
    const dim3 g((unsigned)L.nblk), b((unsigned)L.threads);
    const bool force = (L.variant & 1) != 0, tl = (L.variant & 4) != 0;
    if (force) {
        if constexpr (sizeof(T) == 4) {
            if (L.variant & 32)
                hipLaunchKernelGGL((k_force_mma_pk<NMAX, 0, ACT>), g, b, L.lds, st, M, L.blk);
            else
                hipLaunchKernelGGL((k_force_mma_pk<NMAX, 1, ACT>), g, b, L.lds, st, M, L.blk);
        } else {
            hipLaunchKernelGGL((k_estimate_pk<T, NMAX, 1, 1, false, EK_FORCE, ACT>), g, b, L.lds, st, M, L.blk);
        }
    } else if (tl)
        hipLaunchKernelGGL((k_estimate_pk<T, NMAX, 1, 1, true, EK_WALK, ACT>), g, b, L.lds, st, M, L.blk);
    else
        hipLaunchKernelGGL((k_estimate_pk<T, NMAX, 1, 1, false, EK_WALK, ACT>), g, b, L.lds, st, M, L.blk);
    return hipGetLastError();
}

template <typename T>
hipError_t pk_launch_prec(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st) {
    if (L.nblk <= 0) return hipSuccess;
    const bool n16 = (L.variant & 2) != 0, lc = (L.variant & 8) != 0;
    if (lc) return n16 ? pk_launch_n<T, 16, NQ_ACT_LOGCOSH>(M, L, st) : pk_launch_n<T, 8, NQ_ACT_LOGCOSH>(M, L, st);
    return n16 ? pk_launch_n<T, 16, NQ_ACT_SOFTPLUS>(M, L, st) : pk_launch_n<T, 8, NQ_ACT_SOFTPLUS>(M, L, st);
}

}  // namespace nq_est
