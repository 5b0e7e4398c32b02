// Fused per-sample kernel of the steady-state loop: log rho(s), the compact
// log-derivative row O(s), the local estimator L_loc(s) = sum_c mel_c
// exp(log rho_c - log rho_s) over the Liouvillian connections, and the force
// sum_s conj(L_loc(s)) dL_loc(s) / Ns contracted on the fly, never
// materialising the Np-vector dL_loc(s).
//
// Reference semantics: NDMBatched.jl:94-280 (log rho, O), AccumulatorObsGrad.jl
// :39-127 (L_loc, dL_loc = sum_c mel_c exp(..) O(c), no -O(s) term),
// BatchedGradSampler.jl:99-118 (force = conj(L_loc) dL_loc / Ns).
//
// Mapping: one wavefront per sample (the table walk is wave-uniform and runs
// in scalar registers), lanes over hidden units.  Lane l owns real unit
// u = l + 64 r (lambda for u < nh, mu otherwise, row and column side) and Pi
// half-unit q = l + 64 r (q < na: row half of Pi unit q, otherwise column half
// of unit q - na).  The effective angles theta = h + W x of the sample live in
// registers; a connection flipping k sites costs k FMAs per unit.
//
// Two passes over the connections of a sample:
//   1. softplus of the changed angles + one wave reduction -> W_c = mel_c
//      exp(log rho_c - log rho_s), L_loc = sum_c W_c (W_c cached in LDS);
//   2. with V = conj(L_loc)/Ns known, sigmoids of the changed angles; the force
//      Re(V W_c O_k(c)) is accumulated per unit as a dense part (x_s times a
//      per-unit scalar, added once per sample) plus a sparse part on the
//      flipped sites only (x_c = x_s + Delta_c), into a per-workgroup FP64
//      force vector in LDS.
// The parity mode writes the per-sample dL_loc(s) instead (pass 2 run with
// V = 1 and V = -i for its real and imaginary parts).
#include "nq_device.h"
#include "nq_table.h"
#include "nq_kernels.h"

namespace {

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
template <typename T, int NMAX, int RS, int PS, bool TAB_LDS>
__global__ __launch_bounds__(512, (sizeof(T) == 4 && RS == 1) ? 4 : 2) void k_estimate(est_args a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const ndm_dims d = a.d;
    const int N = d.N, nh = d.nh, na = d.na;
    T* sp = reinterpret_cast<T*>(smem);
    double* sF = reinterpret_cast<double*>(smem + a.lds_param_bytes);
    const int fbytes = (a.Fpart || a.G) ? d.np * (int)sizeof(double) : 0;
    cplx<T>* sW = reinterpret_cast<cplx<T>*>(smem + a.lds_param_bytes + fbytes);
    const int wpb0 = blockDim.x >> 6;
    uint2* sM = reinterpret_cast<uint2*>(sW + wpb0 * a.max_conn);          // flip masks of the cached connections
    unsigned char* sT = smem + ((reinterpret_cast<unsigned char*>(sM + wpb0 * a.max_conn) - smem + 15) & ~15);
    const T* gp = reinterpret_cast<const T*>(a.p);
    const bool useF = a.Fpart != nullptr || a.G != nullptr;
    const dev_table tab = TAB_LDS ? table_to_lds(a.tab, sT) : a.tab;
    for (int k = threadIdx.x; k < d.np; k += blockDim.x) sp[k] = gp[k];
    if (useF)
        for (int k = threadIdx.x; k < d.np; k += blockDim.x) sF[k] = 0.0;
    __syncthreads();
    auto fadd = [&](int pidx, double v) { atomicAdd(&sF[pidx], v); };

    const int lane0 = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wpb = blockDim.x >> 6;
    const int64_t gw = (int64_t)blockIdx.x * wpb + wave;
    const int64_t nw = (int64_t)gridDim.x * wpb;
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

        // ---- effective angles of the sample
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
                softplus_and_sigmoid(ts, &f_s[r], &a_s[r]);
                softplus_and_sigmoid(te, &f_e[r], &a_e[r]);
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
                csoftplus_and_sigmoid(pi_z[r], &pi_f[r], &pi_p[r]);
            }
        }

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

        if (!a.do_conn) continue;

        // ---- pass 1: W_c and L_loc.  The walk (terms spread over the lanes)
        // sums the unchanged connections (W = mel) per lane and caches the
        // changed ones (mel, flip masks) in LDS; those are then evaluated EB at
        // a time: every lane adds its units' share of dlog rho for each of
        // them, one transposed reduction sums all 2 EB shares, and a group of
        // lanes per connection finishes it (bias terms, e^dlog rho, W = mel
        // e^dlog rho), all EB in parallel; W replaces mel in the cache for pass 2
        int ncc;
        cplx<T> Wdiag, Lloc;
        if (a.rec_mode == 2) {
            // forces only: the connections of pass 1 from the earlier launch
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
        constexpr int EB = est_batch<T>(), LG = 32 / EB;   // LG lanes hold one connection's sums
        // FP32 with one Pi unit per lane (na <= 32): two connections' Pi units
        // per pass (in FP64 the second set of operands would spill)
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
                                dfs = softplus(t) - f_s[r];
                            }
                            if (me) {
                                T t = th_e[r];
                                for (uint32_t m = me; m; m &= m - 1) {
                                    const int i = __builtin_ctz(m);
                                    t += cflip<T>(ke, i) * w[nh * i];
                                }
                                dfe = softplus(t) - f_e[r];
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
                            cplx<T> f2, p2;
                            csoftplus_and_sigmoid(z, &f2, &p2);
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
                        cplx<T> f2, p2;
                        csoftplus_and_sigmoid(z, &f2, &p2);
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
        Lloc = cmk(wave_total(Lacc.re), wave_total(Lacc.im));
        __builtin_amdgcn_wave_barrier();

        if (lane == 0 && a.Lloc) {
            T* lo = reinterpret_cast<T*>(a.Lloc);
            lo[2 * s] = Lloc.re;
            lo[2 * s + 1] = Lloc.im;
        }
        if (a.rec_mode == 1) {
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
        }   // rec_mode != 2
        if (!a.Fpart && !a.G) continue;

        // ---- pass 2: force contributions Re(V W_c O_k(c))
        constexpr bool PAIR_PI = PS == 1 && sizeof(T) == 4;
        const int nv = a.G ? 2 : 1;
        for (int vi = 0; vi < nv; ++vi) {
            // Re(1 G) = Re G, Re(-i G) = Im G in the parity mode
            const cplx<double> Vd = a.G ? (vi == 0 ? cmk(1.0, 0.0) : cmk(0.0, -1.0))
                                        : cmk((double)Lloc.re * a.invNs, -(double)Lloc.im * a.invNs);
            const cplx<T> V = cmk((T)Vd.re, (T)Vd.im);
            T e_s[RS], e_e[RS];
            T rho1[PS], rho2[PS];
#pragma unroll
            for (int r = 0; r < RS; ++r) e_s[r] = e_e[r] = T(0);
#pragma unroll
            for (int r = 0; r < PS; ++r) rho1[r] = rho2[r] = T(0);
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
                            as2 = sigmoid(t);
                        }
                        if (me) {
                            T t = th_e[r];
                            for (uint32_t m = me; m; m &= m - 1) {
                                const int i = __builtin_ctz(m);
                                t += cflip<T>(ke, i) * w[nh * i];
                            }
                            ae2 = sigmoid(t);
                        }
                        e_s[r] += gam * as2;
                        e_e[r] += seta * gam * ae2;
                        const int pb = lam ? PB_wla : PB_wmu;
                        for (uint32_t m = ms; m; m &= m - 1) {
                            const int i = __builtin_ctz(m);
                            fadd(pb + h + nh * i, (double)(half * gam * as2 * cflip<T>(ks, i)));
                        }
                        for (uint32_t m = me; m; m &= m - 1) {
                            const int i = __builtin_ctz(m);
                            fadd(pb + h + nh * i, (double)(half * seta * gam * ae2 * cflip<T>(ke, i)));
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
                            // f'(z) = w / (1 + w), w = e^z
                            T sn, cs;
                            nq_sincos(z.im, &sn, &cs);
                            const T ex = nq_exp(z.re);
                            const T wr = ex * cs, wi = ex * sn, dr = T(1) + wr;
                            const T den = nq_rcp(dr * dr + wi * wi);
                            p2 = cmk((wr * dr + wi * wi) * den, wi * den);
                        }
                        const cplx<T> vp = cmul(VW, p2);
                        if (rowhalf) {
                            rho1[r] += vp.re;
                            rho2[r] += vp.im;
                            for (uint32_t m = ms; m; m &= m - 1) {
                                const int i = __builtin_ctz(m);
                                const T cc = half * cflip<T>(ks, i);
                                fadd(PB_ula + au + na * i, (double)(cc * vp.re));
                                fadd(PB_umu + au + na * i, (double)(-cc * vp.im));
                            }
                        } else {
                            for (uint32_t m = me; m; m &= m - 1) {
                                const int i = __builtin_ctz(m);
                                const T cc = half * cflip<T>(ke, i);
                                fadd(PB_ula + au + na * i, (double)(cc * vp.re));
                                fadd(PB_umu + au + na * i, (double)(cc * vp.im));
                            }
                        }
                    }
                }
                if (lane < N) {
                    if ((ms >> lane) & 1u) {
                        const T cc = half * cflip<T>(ks, lane);
                        fadd(PB_bla + lane, (double)(cc * VW.re));
                        fadd(PB_bmu + lane, (double)(-cc * VW.im));
                    }
                    if ((me >> lane) & 1u) {
                        const T cc = half * cflip<T>(ke, lane);
                        fadd(PB_bla + lane, (double)(cc * VW.re));
                        fadd(PB_bmu + lane, (double)(cc * VW.im));
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
                            T sn, cs;
                            nq_sincos(z.im, &sn, &cs);
                            const T ex = nq_exp(z.re);
                            const T wr = ex * cs, wi = ex * sn, dr = T(1) + wr;
                            const T den = nq_rcp(dr * dr + wi * wi);
                            p2 = cmk((wr * dr + wi * wi) * den, wi * den);
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
            // dense parts: x_s times the per-unit sums (the opaque copy of the
            // unit index keeps the 16 addresses per unit from being hoisted
            // out of the connection loop and spilled)
#pragma unroll
            for (int r = 0; r < RS; ++r) {
                int u = lane + 64 * r;
                asm volatile("" : "+v"(u));
                if (u < 2 * nh) {
                    const bool lam = u < nh;
                    const int h = lam ? u : u - nh;
                    fadd(lam ? PB_hla + h : PB_hmu + h, (double)(half * (e_s[r] + e_e[r])));
                    const int pb = lam ? PB_wla : PB_wmu;
#pragma unroll
                    for (int i = 0; i < NMAX; ++i)
                        if (i < N)
                            fadd(pb + h + nh * i, (double)(half * (e_s[r] * xv<T>(ks, i) + e_e[r] * xv<T>(ke, i))));
                }
            }
#pragma unroll
            for (int r = 0; r < PS; ++r) {
                int q = lane + 64 * r;
                asm volatile("" : "+v"(q));
                if (q < na) {
                    fadd(PB_d + q, (double)rho1[r]);
#pragma unroll
                    for (int i = 0; i < NMAX; ++i)
                        if (i < N) {
                            const T xs = xv<T>(ks, i), xe = xv<T>(ke, i);
                            fadd(PB_ula + q + na * i, (double)(half * rho1[r] * (xs + xe)));
                            fadd(PB_umu + q + na * i, (double)(-half * rho2[r] * (xs - xe)));
                        }
                }
            }
            if (lane < N) {
                const cplx<T> VL = cmul(V, Lloc);
                const T xs = xv<T>(ks, lane), xe = xv<T>(ke, lane);
                fadd(PB_bla + lane, (double)(half * VL.re * (xs + xe)));
                fadd(PB_bmu + lane, (double)(-half * VL.im * (xs - xe)));
            }
            if (a.G) {   // one wave per block in this mode: sF belongs to this sample
                __syncthreads();
                double* g = reinterpret_cast<double*>(a.G) + (int64_t)s * 2 * d.np;
                for (int k = lane; k < d.np; k += 64) {
                    g[2 * p2ref(k, d) + vi] = sF[k];
                    sF[k] = 0.0;
                }
                __syncthreads();
            }
        }
    }

    if (a.Fpart && !a.G) {
        __syncthreads();
        double* out = a.Fpart + (int64_t)blockIdx.x * d.np;
        for (int k = threadIdx.x; k < d.np; k += blockDim.x) out[k] = sF[k];
    }
}

template <typename T, int NMAX, int RS, int PS>
hipError_t launch_t(const est_args& a, int grid, hipStream_t st) {
    // 8 waves per block share one copy of the parameters, the force
    // accumulator and the table (two blocks per CU while that fits in half
    // the LDS: 4 waves per SIMD); otherwise 4.  The parity mode (per-sample
    // gradients) runs one wave per block.
    int wpb = a.G ? 1 : 8;
    if (!a.G && est_lds_bytes(a, sizeof(T), wpb) > 80 * 1024) wpb = 4;
    if (a.tab_lds)
        hipLaunchKernelGGL((k_estimate<T, NMAX, RS, PS, true>), dim3(grid), dim3(64 * wpb),
                           est_lds_bytes(a, sizeof(T), wpb), st, a);
    else
        hipLaunchKernelGGL((k_estimate<T, NMAX, RS, PS, false>), dim3(grid), dim3(64 * wpb),
                           est_lds_bytes(a, sizeof(T), wpb), st, a);
    return hipGetLastError();
}

template <typename T, int NMAX>
hipError_t launch_n(const est_args& a, int grid, hipStream_t st) {
    const int rs = (2 * a.d.nh + 63) / 64, ps = (2 * a.d.na + 63) / 64;
    if (rs <= 1 && ps <= 1) return launch_t<T, NMAX, 1, 1>(a, grid, st);
    if (rs <= 2 && ps <= 2) return launch_t<T, NMAX, 2, 2>(a, grid, st);
    return hipErrorInvalidValue;
}

}  // namespace

size_t est_lds_bytes(const est_args& a, size_t elt, int waves) {
    // params | FP64 force | waves x (W cache + flip masks) | table (16-aligned)
    const size_t b = (size_t)a.lds_param_bytes + ((a.Fpart || a.G) ? sizeof(double) * (size_t)a.d.np : 0) +
                     (size_t)waves * (size_t)a.max_conn * (2 * elt + 8);
    return ((b + 15) & ~(size_t)15) + (a.tab_lds ? (size_t)a.tab.bytes : 0);
}

hipError_t nq_launch_estimate(const est_args& a, bool fp64, int grid, hipStream_t st) {
    if (a.d.N > 16) return hipErrorInvalidValue;
    if (est_lds_bytes(a, fp64 ? 8 : 4, a.G ? 1 : 4) > 160 * 1024) return hipErrorInvalidValue;
    if (fp64) {
        if (a.d.N <= 8) return launch_n<double, 8>(a, grid, st);
        return launch_n<double, 16>(a, grid, st);
    }
    if (a.d.N <= 8) return launch_n<float, 8>(a, grid, st);
    return launch_n<float, 16>(a, grid, st);
}
