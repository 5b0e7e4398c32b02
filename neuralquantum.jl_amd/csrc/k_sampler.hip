// Metropolis-Hastings with LocalRule on the doubled (row, column) space,
// p(s, e) ∝ |rho(s, e)|^2, log p = 2 Re log rho (base_batched_networks.jl:255-258).
//
// Schedule (Metropolis.jl:101-167, BaseIterativeSampler.jl:5-17): optional
// restart from a uniform configuration, `burn` calls of `passes` proposals,
// then one stored slice after each further call of `passes` proposals.  A
// proposal flips one uniformly chosen doubled site (LocalRule.jl:30-38) and is
// rejected iff u - exp(log p' - log p) >= 0 (Metropolis.jl:148-154).
//
// Mapping: a group of `lpc` lanes (power of two) runs one chain, 64/lpc chains
// per wavefront.  Lane g of a group owns lambda units h = g + lpc r (both
// sides) and Pi units a = g + lpc r; only the lambda and Pi blocks enter
// Re log rho, the mu block does not.  The effective angles stay in registers;
// a proposal costs one FMA + one f (softplus or logcosh, the NDM's activation)
// per lambda unit of the flipped side and one complex update + Re f per Pi
// unit, then a group reduction.
// Angles are recomputed from the configuration at the start of every call so
// incremental rounding never accumulates over more than `passes` updates.
// Random numbers: Philox4x32-10, counter (t, iteration, chain, 1) for proposal
// t; each lane of the group draws for one future proposal and the draws are
// shuffled to the group when used.
#include "nq_device.h"
#include "nq_kernels.h"

namespace {

template <typename T>
__device__ __forceinline__ T xv(uint32_t bits, int i) { return ((bits >> i) & 1u) ? T(1) : T(-1); }

// the block body; bx = the workgroup's number in its problem's grid
template <typename T, int US, int ACT>
__device__ __forceinline__ void metropolis_body(const samp_args& a, int bx) {
    extern __shared__ __align__(16) unsigned char smem[];
    const ndm_dims d = a.d;
    const int N = d.N, nh = d.nh, na = d.na;
    T* sp = reinterpret_cast<T*>(smem);
    const T* gp = reinterpret_cast<const T*>(a.p);
    for (int k = threadIdx.x; k < d.np; k += blockDim.x) sp[k] = gp[k];
    __syncthreads();

    const int lpc = a.lpc;
    const int lane = threadIdx.x & 63;
    const int g = lane & (lpc - 1);
    const int gbase = lane - g;
    const int cpw = 64 / lpc;
    const int64_t wave_glob = (int64_t)bx * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t chain = wave_glob * cpw + lane / lpc;
    const bool live = chain < a.chains;
    const uint32_t pmask = (N >= 32) ? 0xffffffffu : ((1u << N) - 1u);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const uint32_t iter = a.iter_dev ? *a.iter_dev : a.iter;
    const T half = T(0.5);

    const bool diag = a.diag != 0;
    const uint32_t ctr3 = 16u * a.tag;
    uint64_t key = 0;
    if (live) {
        if (a.restart) {
            philox_out o = philox4x32_10(0u, iter, (uint32_t)chain, ctr3, k0, k1);
            if (diag) {
                const uint64_t s = ((uint64_t)o.x[0] | ((uint64_t)o.x[1] << 32)) & (uint64_t)pmask;
                key = s | (s << N);
            } else {
                const uint64_t m2 = (2 * N >= 64) ? ~0ull : ((1ull << (2 * N)) - 1ull);
                key = ((uint64_t)o.x[0] | ((uint64_t)o.x[1] << 32)) & m2;
            }
        } else {
            key = a.chain_state[chain];
        }
    }

    T th_s[US], th_e[US], f_s[US], f_e[US], fpi[US];
    cplx<T> z[US];
    uint64_t acc = 0;
    uint32_t t = a.t0;
    uint32_t rx0 = 0, rx1 = 0, rx2 = 0;   // this lane's pre-drawn proposal
    const int calls = a.burn + a.stored;

    for (int call = 0; call < calls; ++call) {
        // refresh the angles from the configuration
        const uint32_t ks = (uint32_t)key & pmask, ke = (uint32_t)(key >> N) & pmask;
#pragma unroll
        for (int r = 0; r < US; ++r) {
            const int h = g + lpc * r;
            th_s[r] = th_e[r] = f_s[r] = f_e[r] = fpi[r] = T(0);
            z[r] = cmk(T(0), T(0));
            if (h < nh) {
                T ts = sp[d.hla + h], te = ts;
                for (int i = 0; i < N; ++i) {
                    const T w = sp[d.wla + h + nh * i];
                    ts += w * xv<T>(ks, i);
                    te += w * xv<T>(ke, i);
                }
                th_s[r] = ts; th_e[r] = te;
                f_s[r] = act_f<ACT>(ts); f_e[r] = act_f<ACT>(te);
            }
            if (h < na) {
                T zr = sp[d.dla + h], zi = T(0);
                for (int i = 0; i < N; ++i) {
                    const T xs = xv<T>(ks, i), xe = xv<T>(ke, i);
                    zr += half * sp[d.ula + h + na * i] * (xs + xe);
                    zi += half * sp[d.umu + h + na * i] * (xs - xe);
                }
                z[r] = cmk(zr, zi);
                fpi[r] = cact_re<ACT>(z[r]);
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
            // diagonal chain: site j flipped in sigma and eta, log p = Re log rho(sigma, sigma)
            const int site = (int)(((uint64_t)x0 * (uint64_t)(diag ? N : 2 * N)) >> 32);
            const bool col = site >= N;
            const int j = col ? site - N : site;
            const uint32_t bits = col ? (uint32_t)(key >> N) : (uint32_t)key;
            const T c = ((bits >> j) & 1u) ? T(-2) : T(2);

            T delta = T(0);
            T nth[US], nf[US], nfpi[US];
            cplx<T> nz[US];
#pragma unroll
            for (int r = 0; r < US; ++r) {
                const int h = g + lpc * r;
                nth[r] = nf[r] = nfpi[r] = T(0);
                nz[r] = z[r];
                if (h < nh) {
                    const T th = col ? th_e[r] : th_s[r];
                    nth[r] = th + c * sp[d.wla + h + nh * j];
                    nf[r] = act_f<ACT>(nth[r]);
                    delta += nf[r] - (col ? f_e[r] : f_s[r]);
                }
                if (h < na) {
                    if (diag) {
                        nz[r].re += c * sp[d.ula + h + na * j];
                        nfpi[r] = cact_re<ACT>(nz[r]);
                        delta += nfpi[r] - fpi[r];
                    } else {
                        const T cu = c * half;
                        nz[r].re += cu * sp[d.ula + h + na * j];
                        nz[r].im += (col ? -cu : cu) * sp[d.umu + h + na * j];
                        nfpi[r] = cact_re<ACT>(nz[r]);
                        delta += T(2) * (nfpi[r] - fpi[r]);
                    }
                }
            }
            delta = group_total(delta, lpc);
            delta += c * sp[d.bla + j];

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
                for (int r = 0; r < US; ++r) {
                    const int h = g + lpc * r;
                    if (h < nh) {
                        if (diag) { th_s[r] = th_e[r] = nth[r]; f_s[r] = f_e[r] = nf[r]; }
                        else if (col) { th_e[r] = nth[r]; f_e[r] = nf[r]; }
                        else { th_s[r] = nth[r]; f_s[r] = nf[r]; }
                    }
                    if (h < na) { z[r] = nz[r]; fpi[r] = nfpi[r]; }
                }
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
            a.samples[chain + (int64_t)(a.stored) * a.chains] = 0ull;   // defect D1: slice L stays at the fill
        if (a.accepted) atomicAdd(a.accepted, (unsigned long long)acc);
        if (a.chain_acc) a.chain_acc[chain] += (uint32_t)acc;   // one lane per chain: no race
    }
}

template <typename T, int US, int ACT>
__global__ __launch_bounds__(256) void k_metropolis(samp_args a) {
    metropolis_body<T, US, ACT>(a, blockIdx.x);
}

// packed: the members' sampler launches in one grid (block records as in nq_pk_member)
template <typename T, int US, int ACT>
__global__ __launch_bounds__(256) void k_metropolis_pk(const nq_pk_member* M, const uint2* blk) {
    const uint2 b = blk[blockIdx.x];
    const samp_args a = M[b.x & 0xffff].sa;
    metropolis_body<T, US, ACT>(a, (int)(b.y & 0xffff));
}

constexpr int kSampWaves = 4;   // wavefronts per workgroup (default)

int samp_waves(const samp_args& a) { return (a.waves == 1 || a.waves == 2 || a.waves == 4) ? a.waves : kSampWaves; }

int64_t samp_grid(const samp_args& a) {
    const int cpw = 64 / a.lpc;
    const int64_t waves = (a.chains + cpw - 1) / cpw;
    return (waves + samp_waves(a) - 1) / samp_waves(a);
}

template <typename T, int US>
hipError_t launch_t(const samp_args& a, hipStream_t st) {
    const int wpb = samp_waves(a);
    const int64_t grid = samp_grid(a);
    size_t lds = sizeof(T) * (size_t)a.d.np;
    if (a.d.act == NQ_ACT_LOGCOSH)
        hipLaunchKernelGGL((k_metropolis<T, US, NQ_ACT_LOGCOSH>), dim3((unsigned)grid), dim3(64 * wpb), lds, st, a);
    else
        hipLaunchKernelGGL((k_metropolis<T, US, NQ_ACT_SOFTPLUS>), dim3((unsigned)grid), dim3(64 * wpb), lds, st, a);
    return hipGetLastError();
}

}  // namespace

int nq_sampler_lanes_per_chain(const ndm_dims& d) {
    int m = d.nh > d.na ? d.nh : d.na;
    int l = 1;
    while (l < m && l < 64) l <<= 1;
    return l;
}

hipError_t nq_launch_metropolis(const samp_args& a, bool fp64, hipStream_t st) {
    const int m = a.d.nh > a.d.na ? a.d.nh : a.d.na;
    const int us = (m + a.lpc - 1) / a.lpc;
    if (a.d.N > 32 || a.lpc < 1 || a.lpc > 64 || (a.lpc & (a.lpc - 1))) return hipErrorInvalidValue;
    if (fp64) {
        if (us <= 1) return launch_t<double, 1>(a, st);
        if (us <= 2) return launch_t<double, 2>(a, st);
        if (us <= 4) return launch_t<double, 4>(a, st);
    } else {
        if (us <= 1) return launch_t<float, 1>(a, st);
        if (us <= 2) return launch_t<float, 2>(a, st);
        if (us <= 4) return launch_t<float, 4>(a, st);
    }
    return hipErrorInvalidValue;
}

// packed launches: one lane group per chain covers every unit (lanes per
// chain >= max(nh, na), so one unit slot per lane); variant = fp64 | logcosh << 1
int nq_sampler_variant(const samp_args& a, bool fp64, int* threads, size_t* lds, int* grid) {
    const int m = a.d.nh > a.d.na ? a.d.nh : a.d.na;
    if (a.d.N > 32 || a.lpc < 1 || a.lpc > 64 || (a.lpc & (a.lpc - 1)) || m > a.lpc) return -1;
    const int64_t g = samp_grid(a);
    if (g > 0xffff) return -1;
    *threads = 64 * samp_waves(a);
    *lds = (fp64 ? sizeof(double) : sizeof(float)) * (size_t)a.d.np;
    *grid = (int)g;
    return (fp64 ? 1 : 0) | (a.d.act == NQ_ACT_LOGCOSH ? 2 : 0);
}

hipError_t nq_pk_sampler(const nq_pk_member* M, const nq_pk_launch& L, hipStream_t st) {
    if (L.nblk <= 0) return hipSuccess;
    const dim3 g((unsigned)L.nblk), b((unsigned)L.threads);
    switch (L.variant) {
        case 0: hipLaunchKernelGGL((k_metropolis_pk<float, 1, NQ_ACT_SOFTPLUS>), g, b, L.lds, st, M, L.blk); break;
        case 1: hipLaunchKernelGGL((k_metropolis_pk<double, 1, NQ_ACT_SOFTPLUS>), g, b, L.lds, st, M, L.blk); break;
        case 2: hipLaunchKernelGGL((k_metropolis_pk<float, 1, NQ_ACT_LOGCOSH>), g, b, L.lds, st, M, L.blk); break;
        case 3: hipLaunchKernelGGL((k_metropolis_pk<double, 1, NQ_ACT_LOGCOSH>), g, b, L.lds, st, M, L.blk); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
