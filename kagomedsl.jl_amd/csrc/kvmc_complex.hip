// Launchers of the sampler kernels (kvmc_kernels.hpp) for complex orbitals
// (flux phases or a magnetic field: the spinon Hamiltonian and U are complex,
// src/Hamiltonian.jl:201, 325, 382-393, and so is W, src/MonteCarlo.jl:
// 181-182, 290): W in complex float (FP32 form) or complex double (FP64).
// The ratio's modulus squared, the delayed terms and the local energy are
// the complex counterparts of the real kernels.  gfx950 only.
#include "kvmc_kernels.hpp"

namespace kvmc {

hipError_t launch_sweep_c(int prec, const SweepArgs& a, int nw, int threads, size_t lds, hipStream_t s) {
    // 16-byte vectors: two complex floats or one complex double
    if (prec == 0) return launch_sweep_v<cf64, 1>(a, nw, threads, lds, s);
    return launch_sweep_v<cf32, 2>(a, nw, threads, lds, s);
}

hipError_t launch_flush_c(int prec, const SweepArgs& a, int nw, hipStream_t s) {
    return prec == 0 ? launch_flush_t<cf64, 1>(a, nw, s) : launch_flush_t<cf32, 2>(a, nw, s);
}

hipError_t launch_eloc_c(int prec, const ElocArgs& a, hipStream_t s) {
    return prec == 0 ? launch_eloc_t<cf64, false>(a, s) : launch_eloc_t<cf32, true>(a, s);
}

hipError_t launch_grad_c(int prec, const ElocArgs& a, hipStream_t s) {
    return prec == 0 ? launch_grad_t<cf64>(a, s) : launch_grad_t<cf32>(a, s);
}

hipError_t launch_q_build_c(const double* dU, const double* U, int ns, int n, int nparams, double* Q,
                            hipStream_t s) {
    return launch_q_build_t<cf64>(dU, U, ns, n, nparams, Q, s);
}

// kappa extended to the 2 ns sites of the real form of a complex refresh.
__global__ void k_kappa_real_form(const int* __restrict__ kap, int* __restrict__ out, int ns, int n, int nw) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nw * ns) return;
    const long long w = idx / ns;
    const int s = (int)(idx % ns);
    const int l = kap[idx];
    out[w * 2 * ns + s] = l;
    out[w * 2 * ns + ns + s] = l ? l + n : 0;
}

hipError_t launch_kappa_real_form(const int* kap, int* out, int ns, int n, int nw, hipStream_t s) {
    const long long total = (long long)nw * ns;
    hipLaunchKernelGGL(k_kappa_real_form, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, kap, out, ns, n,
                       nw);
    return hipGetLastError();
}

}  // namespace kvmc
