// Launchers of the sampler kernels (kvmc_kernels.hpp) for real orbitals:
// W in float (FP32 form) or double (FP64 form).  gfx950 only.
#include "kvmc_kernels.hpp"

namespace kvmc {

bool sweep_resident(size_t esize, int ns, int nup, int ndn) {
    return w_stride(ns, nup, ndn) * esize <= kResidentMaxBytes;
}

size_t sweep_lds_bytes(size_t esize, int ns, int nup, int ndn, int delay, int nb) {
    const int nmax = nup > ndn ? nup : ndn;
    return sweep_lds_base(esize, ns, w_nsa(ns, nup, ndn), nmax, delay, nb) +
           (sweep_resident(esize, ns, nup, ndn) ? w_stride(ns, nup, ndn) * esize : 0);
}

hipError_t launch_sweep(int prec, const SweepArgs& a, int nw, int threads, size_t lds, hipStream_t s) {
    // rows of W are padded to a multiple of 4: always whole 16-byte vectors
    if (prec == 0) return launch_sweep_v<double, 2>(a, nw, threads, lds, s);
    // FP32 W streamed from HBM with columns of at most 256 rows: pending rows
    // kept implicit (no row reads at an accept)
    if (sweep_ib_supported(prec, 0, a.ns, a.nup, a.ndn, a.delay)) return launch_sweep_ib(a, nw, s);
    return launch_sweep_v<float, 4>(a, nw, threads, lds, s);
}

hipError_t launch_flush(int prec, const SweepArgs& a, int nw, hipStream_t s) {
    return prec == 0 ? launch_flush_t<double, 2>(a, nw, s) : launch_flush_t<float, 4>(a, nw, s);
}

hipError_t launch_eloc(int prec, const ElocArgs& a, hipStream_t s) {
    return prec == 0 ? launch_eloc_t<double, false>(a, s) : launch_eloc_t<float, true>(a, s);
}

hipError_t launch_grad(int prec, const ElocArgs& a, hipStream_t s) {
    return prec == 0 ? launch_grad_t<double>(a, s) : launch_grad_t<float>(a, s);
}

hipError_t launch_q_build(const double* dU, const double* U, int ns, int n, int nparams, double* Q,
                          hipStream_t s) {
    return launch_q_build_t<double>(dU, U, ns, n, nparams, Q, s);
}

}  // namespace kvmc
