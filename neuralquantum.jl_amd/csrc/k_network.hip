// Connection listing for host-fed configurations: the parity entry point for
// the connection indexing (targets, mels and order of KLocalLiouvillian.jl:46-52
// through AccumulatorObsGrad.jl:63-80).  One thread per configuration walks the
// same table and the same table_walk() the estimator uses.
#include "nq_kernels.h"

namespace {

__global__ __launch_bounds__(256) void k_connections(dev_table tab, int N, const uint64_t* keys, int64_t B,
                                                      int cap, int* counts, uint64_t* keys_out,
                                                      double* mel_out) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const uint64_t key = keys[b];
    const uint32_t pm = (N >= 32) ? 0xffffffffu : ((1u << N) - 1u);
    const uint32_t ks = (uint32_t)key & pm, ke = (uint32_t)(key >> N) & pm;
    int n = 0;
    table_walk<double>(tab, ks, ke, [&](cplx<double> mel, uint32_t ms, uint32_t me) {
        if (n < cap) {
            keys_out[b * cap + n] = key ^ (uint64_t)ms ^ ((uint64_t)me << N);
            mel_out[2 * (b * cap + n)] = mel.re;
            mel_out[2 * (b * cap + n) + 1] = mel.im;
        }
        ++n;
    });
    counts[b] = n;
}

}  // namespace

hipError_t nq_launch_connections(dev_table tab, int N, const uint64_t* keys, int64_t B, int cap, int* counts,
                                 uint64_t* keys_out, double* mel_out, hipStream_t st) {
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_connections, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, st, tab, N, keys, B,
                       cap, counts, keys_out, mel_out);
    return hipGetLastError();
}
