// Internal declarations shared by the device code (kvmc_device.hip) and the
// C-ABI host layer (kvmc_api.hip).  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kvmc {

// Per-sweep outcome codes (mirror of KVMC_OUT_* in kagome_vmc.h).
enum : int { OUT_PRE_REJECT = 0, OUT_NO_MOVE = 1, OUT_REJECT = 2, OUT_ACCEPT = 3 };
enum : int { KERNEL_REFERENCE = 0, KERNEL_UNBIASED = 1 };
enum : int { POLICY_REFERENCE = 0, POLICY_EVERY_K = 1, POLICY_AT_END = 2 };

// Device-wide counters (u64 slots).
enum : int { CNT_MOVES = 0, CNT_ACCEPTED = 1, CNT_FLUSHES = 2, CNT_SWEEPS = 3, CNT_NAP = 4,
             CNT_PAUSES = 5, CNT_SLOTS = 8 };

// Largest number of pending rank-1 updates kept per spin before a flush.
constexpr int kMaxDelay = 16;

// Complex W element (complex orbitals: flux phases or a magnetic field make
// the spinon Hamiltonian complex, src/Hamiltonian.jl:201, 325, 382-393).
// Stored interleaved (re, im) like Julia's ComplexF64 / ComplexF32.
template <typename R>
struct __align__(2 * sizeof(R)) cplx {
    R re, im;
    cplx() = default;
    __host__ __device__ constexpr cplx(R r) : re(r), im(R(0)) {}
    __host__ __device__ constexpr cplx(R r, R i) : re(r), im(i) {}
    template <typename S>
    __host__ __device__ explicit constexpr cplx(const cplx<S>& o) : re((R)o.re), im((R)o.im) {}
};
template <typename R> __host__ __device__ inline cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> __host__ __device__ inline cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> __host__ __device__ inline cplx<R> operator-(cplx<R> a) { return {-a.re, -a.im}; }
template <typename R> __host__ __device__ inline cplx<R> operator*(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> __host__ __device__ inline cplx<R> operator*(R a, cplx<R> b) { return {a * b.re, a * b.im}; }
template <typename R> __host__ __device__ inline cplx<R> operator/(cplx<R> a, cplx<R> b) {
    const R d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
template <typename R> __host__ __device__ inline cplx<R>& operator+=(cplx<R>& a, cplx<R> b) { a = a + b; return a; }
template <typename R> __host__ __device__ inline cplx<R>& operator-=(cplx<R>& a, cplx<R> b) { a = a - b; return a; }
typedef cplx<float> cf32;
typedef cplx<double> cf64;

template <typename T> struct is_cplx { static constexpr bool value = false; };
template <typename R> struct is_cplx<cplx<R>> { static constexpr bool value = true; };
// Real scalar of a storage type (float for cf32, double for cf64).
template <typename T> struct real_of { using type = T; };
template <typename R> struct real_of<cplx<R>> { using type = R; };

// W of one spin keeps only the rows of the empty sites ("slots"): the row of
// an occupied site is the unit row e_label by construction (W = U tildeU^-1,
// W[R_l, :] = e_l).  Per walker: W_up (ldw_up x nup) then W_dn (ldw_dn x
// ndn), column-major, ldw = rup4(ns - N); rows ns - N .. ldw - 1 are zero
// padding.  slot[w][spin][site] is the row of an empty site, -1 for an
// occupied one.  After a refresh the slots follow the empty sites in site
// order; an accepted move hands the slot of the newly occupied site to the
// site the electron left.
__host__ __device__ inline int w_ld(int ns, int n) { return (ns - n + 3) / 4 * 4; }
__host__ __device__ inline size_t w_off_dn(int ns, int nup) { return (size_t)w_ld(ns, nup) * nup; }
// Pending rank-1 terms in HBM: A as [walker][spin][row < w_nsa][pend_ld],
// B as [walker][spin][label < nmax][pend_ld], so the terms of one row (or
// label) are contiguous and 16-byte aligned.
__host__ __device__ inline int pend_ld(int delay) { return (delay + 3) / 4 * 4; }
// Number of rows of the pending columns A_q (the rows of the taller W).
__host__ __device__ inline int w_nsa(int ns, int nup, int ndn) {
    const int a = w_ld(ns, nup), b = w_ld(ns, ndn);
    return (a > b ? a : b) > 4 ? (a > b ? a : b) : 4;
}
__host__ __device__ inline size_t w_stride(int ns, int nup, int ndn) {
    return w_off_dn(ns, nup) + (size_t)w_ld(ns, ndn) * ndn;
}

// Philox4x32-10 (Salmon et al., SC'11): counter-based, so every
// (walker, sweep) draw is addressable without carried state.
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Sweep draws of (seed, walker, sweep counter): u in [0,1) from 53 bits and
// a bond index in [0, nb) by 32x32 multiply-high.
__host__ __device__ inline void sweep_draws(uint64_t seed, uint32_t walker, int64_t sweep, int nb,
                                            double* u, int* bond) {
    uint32_t x[4];
    philox4x32_10((uint32_t)sweep, (uint32_t)((uint64_t)sweep >> 32), walker, 0u, (uint32_t)seed,
                  (uint32_t)(seed >> 32), x);
    *u = ((double)(x[0] >> 5) * 67108864.0 + (double)(x[1] >> 6)) * (1.0 / 9007199254740992.0);
    *bond = (int)(((uint64_t)x[2] * (uint64_t)(uint32_t)nb) >> 32);
}

// Arguments of the sweep kernel.  Every pointer is device memory.
struct SweepArgs {
    int ns, nup, ndn, nb, nocc;     // nocc = min(nup, ndn)
    const int2* bonds;              // nb pairs, 0-based
    const int* nbr_off;             // ns+1 offsets into nbr
    const int* nbr;                 // bond ids touching each site
    int* kup;                       // [nw][ns] labels 1-based, 0 empty
    int* kdn;
    void* W;                        // [nw][w_stride] empty-site rows, column-major per spin
    int* slot;                      // [nw][2][ns] site -> row of W (-1: occupied)
    void* A;                        // [nw][2][w_nsa][pend_ld]  pending columns (over rows of W)
    void* Bm;                       // [nw][2][nmax][pend_ld] pending scaled rows
    int* npend;                     // [nw] pending count
    int delay;                      // 1 = immediate rank-1 update
    int nmax;                       // max(nup, ndn)
    long long* sweeps;              // [nw] Carlo ctx.sweeps
    const long long* stop;          // [nw] sweep count at which this call ends
    const long long* tape_base;     // [nw] sweep count at the start of the call
    int* zmu;                       // [nw] antiparallel bond count
    int* acc_since;                 // [nw] accepted moves since last refresh
    int* refresh_list;              // walkers paused for a refresh
    int* refresh_count;
    unsigned long long* counters;   // CNT_*
    uint64_t seed;
    const double* tape_u;           // [nw][tape_len] or null
    const int* tape_bond;
    long long tape_len;
    signed char* outcomes;          // [nw][tape_len] or null
    int kernel_mode, policy, refresh_k;
    int walker0;                    // first walker of this launch
    // FP32 sweep with implicit pending rows (k_sweep_ib): the row slot K_q
    // and label l_q of each pending term, [nw][2][kMaxDelay][2]
    int* pterm;
};

struct RefreshArgs {
    int ns, nup, ndn;
    const double* Uup;              // ns x nup column-major
    const double* Udn;
    const double* UTup;             // ns x nup row-major copy of Uup
    const double* UTdn;
    const int* kup;                 // [nw][ns]
    const int* kdn;
    void* W;                        // [nw][w_stride]
    int* slot;                      // [nw][2][ns] rewritten with W (write_w), read in drift mode
    const void* A;                  // pending (drift mode reads W + A B^T)
    const void* Bm;
    int* npend;                     // cleared for refreshed walkers (write_w)
    int* acc_since;                 // cleared for refreshed walkers (write_w)
    int delay, nmax;
    const int* jobs;                // walker ids; job j -> walker jobs[j/2], spin j%2
    int njobs;
    const int* count_dev;           // when set: only the first 2 * (*count_dev - walker_off) jobs run
    int walker_off;
    double* ws;                     // [njobs][refresh_ws_doubles_per_job] FP64 work matrices
    double* logdet;                 // [nw][2]
    int* status;                    // [nw] 0 ok, 1 singular
    int* failed;                    // smallest failing walker (atomicMin)
    int write_w;                    // 1: replace W; 0: measure drift only
    double pivot_tol;               // reject when min|pivot| < pivot_tol * max|pivot|
    double* drift;                  // [nw][2] per walker: max |W_new - W_cur|, max |W_new| (ordered bits)
    // Storage shape of W, the slot map and the pending terms.  Real orbitals:
    // the same as the factorised problem.  Complex orbitals (cplx_out = 1):
    // the refresh factors the real 2N x 2N form [[Re, -Im], [Im, Re]] of
    // tildeU with rows [U_re, -U_im ; U_im, U_re] of the 2 ns x 2 N real
    // form of U and kappa extended to 2 ns sites (site ns + s holds label
    // N + l when s holds l), so ns = 2 ns_out and N = 2 N_out; the rows of
    // the first ns_out sites give W = X_re + i X_im with X = [X_re, -X_im].
    int ns_out, nup_out, ndn_out, nmax_out;
    int cplx_out;
};

struct ElocArgs {
    int ns, nup, ndn, nb;
    const int2* bonds;
    const int* kup;
    const int* kdn;
    const void* W;
    const int* slot;
    const void* A;
    const void* Bm;
    const int* npend;
    int delay, nmax;
    int nw;
    double* eloc;                   // [nw]
    double* eloc_im;                // [nw] imaginary part (complex W only; may be null)
    unsigned long long* counters;   // CNT_NAP accumulates antiparallel counts
    // log-derivatives: Q_k = dU_k U^H (ns x ns row-major per parameter and
    // spin; complex interleaved for complex W)
    int nparams;
    const double* Qup;              // [nparams][ns][ns]
    const double* Qdn;
    double* O;                      // [nw][nparams] (complex: [nw][nparams][2])
};

// Storage type of W: `prec` 0 = FP64, 1 = FP32; `cplx` 0 = real orbitals,
// 1 = complex orbitals.
inline size_t w_elem_bytes(int prec, int cplx) { return (prec == 0 ? 8 : 4) * (cplx ? 2 : 1); }

// Launchers.  Real W: kvmc_device.hip; complex W: kvmc_complex.hip.
hipError_t launch_sweep(int prec, const SweepArgs& a, int nwalkers, int threads, size_t lds,
                        hipStream_t s);
hipError_t launch_sweep_c(int prec, const SweepArgs& a, int nwalkers, int threads, size_t lds,
                          hipStream_t s);
// FP32 sweep with implicit pending rows (kvmc_sweep_ib.hip): real W streamed
// from HBM, at most 256 rows per spin, 1 < delay <= kMaxDelay.
bool sweep_ib_supported(int prec, int cplx, int ns, int nup, int ndn, int delay);
hipError_t launch_sweep_ib(const SweepArgs& a, int nwalkers, hipStream_t s);
size_t sweep_lds_bytes(size_t esize, int ns, int nup, int ndn, int delay, int nb);
// Both W of a walker at most this large: the sweep kernel keeps them in LDS.
constexpr size_t kResidentMaxBytes = 64 * 1024;
bool sweep_resident(size_t esize, int ns, int nup, int ndn);
hipError_t launch_refresh(int prec, const RefreshArgs& a, hipStream_t s);
size_t refresh_ws_doubles_per_job(int ns, int nmax);
bool refresh_supported(int ns, int nup, int ndn);
hipError_t launch_eloc(int prec, const ElocArgs& a, hipStream_t s);
hipError_t launch_eloc_c(int prec, const ElocArgs& a, hipStream_t s);
hipError_t launch_grad(int prec, const ElocArgs& a, hipStream_t s);
hipError_t launch_grad_c(int prec, const ElocArgs& a, hipStream_t s);
hipError_t launch_flush(int prec, const SweepArgs& a, int nwalkers, hipStream_t s);
hipError_t launch_flush_c(int prec, const SweepArgs& a, int nwalkers, hipStream_t s);
hipError_t launch_q_build(const double* dU, const double* U, int ns, int n, int nparams,
                          double* Q, hipStream_t s);
hipError_t launch_q_build_c(const double* dU, const double* U, int ns, int n, int nparams,
                            double* Q, hipStream_t s);
// kappa of every walker extended to the real form of a complex refresh:
// out[w][s] = kap[w][s], out[w][ns + s] = kap[w][s] ? kap[w][s] + n : 0.
hipError_t launch_kappa_real_form(const int* kap, int* out, int ns, int n, int nw, hipStream_t s);

}  // namespace kvmc
