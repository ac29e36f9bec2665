// C ABI of the Kagome VMC sampler (declared in include/kagome_vmc.h).
//
// The handle owns every device buffer of a walker batch.  Host entry points
// validate their arguments, enqueue kernels on the handle's stream and
// synchronise only where host data is returned or a refresh decision has to
// be read back.
#include "../../include/kagome_vmc.h"
#include "kvmc_internal.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace kvmc;

// Device scratch of kvmc_sweep_noalloc: the tape (u as double, bond as
// int32) and the outcome record (int8) of one call, each segment starting
// on a 256-byte boundary.  Every GPU the library runs on holds its own copy
// for the life of the process.  In host code `scratch` names a host-side
// stand-in: its device address comes from hipGetSymbolAddress, once per
// device (scratch_address below).
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr size_t kScratchMaxEntries = (kScratchBytes - 512) / 13;   // walker-sweeps per call
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_err;

// Random starts whose pivots span more than this ratio are redrawn: such a
// tildeU is numerically singular even when no pivot is exactly zero.  With W
// stored in float the bar is higher: a start whose pivots span more than 1e4
// would carry a float W with fewer than ~10 significant bits in its ratios.
// (At L = 12 the pivot ratios of uniform starts are bimodal, 1e-15 or
// above 1e-3, so the stricter bar costs about 2 % more redraws.)
constexpr double kStartPivotTol = 1e-10;
constexpr double kStartPivotTolF32 = 1e-4;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(KVMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(KVMC_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                        " bytes failed: " + hipGetErrorString(e));
    }
    return KVMC_OK;
}

// Device buffer owned by a scope: freed on every exit path.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    int alloc(size_t count) { reset(); return dalloc(&p, count); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    T* release() { T* q = p; p = nullptr; return q; }
};

// Seeded uniform Sz=0 configuration (Fisher-Yates over Philox words keyed
// (seed, walker, attempt)).
void random_config(uint64_t seed, uint32_t walker, uint32_t attempt, int ns, int nup, int ndn,
                   int* kup, int* kdn) {
    std::vector<int> perm(ns);
    for (int i = 0; i < ns; ++i) perm[i] = i;
    uint32_t words[4];
    int widx = 4;
    uint32_t call = 0;
    for (int i = ns - 1; i >= 1; --i) {
        if (widx == 4) {
            philox4x32_10(call++, attempt, walker, 1u, (uint32_t)seed, (uint32_t)(seed >> 32), words);
            widx = 0;
        }
        int j = (int)(((uint64_t)words[widx++] * (uint64_t)(i + 1)) >> 32);
        std::swap(perm[i], perm[j]);
    }
    std::fill(kup, kup + ns, 0);
    std::fill(kdn, kdn + ns, 0);
    for (int a = 0; a < nup; ++a) kup[perm[a]] = a + 1;
    for (int b = 0; b < ndn; ++b) kdn[perm[nup + b]] = b + 1;
}

__global__ void k_set_stop(const long long* sweeps, long long* stop, long long* base,
                           long long nsweeps, int nw) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nw) {
        base[w] = sweeps[w];
        stop[w] = sweeps[w] + nsweeps;
    }
}

__global__ void k_fill_ll(long long* p, long long v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i(int* p, int v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_ull(unsigned long long* p, unsigned long long v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i8(signed char* p, signed char v, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Device address of `scratch` on each device, looked up on the first call
// made with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// Relative drift of each listed walker (max |W - W_fresh| / max |W_fresh|
// over both spins, from the per-walker maxima of a drift pass) folded into
// stats[0] (maximum, ordered bits), stats[1] (sum) and stats[2] (count).
__global__ void k_drift_fold(const int* list, int count, const double* drift, double* stats) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int w = list[i];
    const double d = drift[(size_t)w * 2], v = drift[(size_t)w * 2 + 1];
    const double rel = v > 0.0 ? d / v : 0.0;
    atomicMax(reinterpret_cast<unsigned long long*>(stats), (unsigned long long)__double_as_longlong(rel));
    atomicAdd(&stats[1], rel);
    atomicAdd(&stats[2], 1.0);
}

// Clears the per-walker drift maxima of the listed walkers.
__global__ void k_drift_clear(const int* list, int count, double* drift) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    drift[(size_t)list[i] * 2] = 0.0;
    drift[(size_t)list[i] * 2 + 1] = 0.0;
}

// Sweep launches per replay of a captured sweep sequence: sweep, refresh
// the walkers that paused, sweep again, ...; walkers still paused after the
// last round are finished by the host loop.
constexpr int kGraphRounds = 3;

}  // namespace

struct kvmc_handle {
    int ns = 0, nup = 0, ndn = 0, nb = 0, nmax = 0, nsp = 0;
    int nw = 0;
    int prec = 0, kernel = 0, policy = 0, refresh_k = 0, delay = 1, threads = 0;
    int cplx = 0;                   // complex orbitals: W in complex float / double
    uint64_t seed = 0;
    size_t esize = 8;
    bool ready = false;
    hipStream_t stream = nullptr;
    int device = 0;

    int2* bonds = nullptr;
    int* nbr_off = nullptr;
    int* nbr = nullptr;
    // U for the refresh (column-major) and its row-major copy; complex
    // orbitals: the 2 ns x 2 N real form [U_re, -U_im ; U_im, U_re]
    double* Uup = nullptr;
    double* Udn = nullptr;
    double* UTup = nullptr;   // row-major copies for the refresh gathers
    double* UTdn = nullptr;
    double* Ucup = nullptr;   // complex orbitals: U itself (interleaved), for Q = dU U^H
    double* Ucdn = nullptr;
    int* kRup = nullptr;      // complex orbitals: kappa on the 2 ns sites of the real form
    int* kRdn = nullptr;
    double* eloc_im = nullptr;
    int* kup = nullptr;
    int* kdn = nullptr;
    void* W = nullptr;              // empty-site rows per spin (w_stride per walker)
    int* slot = nullptr;            // [nw][2][ns] site -> row of W, -1 when occupied
    void* A = nullptr;
    void* Bm = nullptr;
    int* npend = nullptr;
    int* pterm = nullptr;           // [nw][2][kMaxDelay][2] rows and labels of the pending terms
    int* zmu = nullptr;
    int* acc_since = nullptr;
    int* status = nullptr;
    long long* sweeps = nullptr;
    long long* stop = nullptr;
    long long* tbase = nullptr;
    int* rlist = nullptr;
    int* rcount = nullptr;
    int* failed = nullptr;
    int* all_walkers = nullptr;
    unsigned long long* counters = nullptr;
    double* logdet = nullptr;
    double* eloc = nullptr;
    double* drift = nullptr;        // [nw][2] per-walker drift maxima of a drift pass
    double* dstats = nullptr;       // [3] drift before refresh: max, sum, count
    bool track_drift = false;       // drift pass before every policy refresh
    double* ws = nullptr;
    int ws_jobs = 0;
    int nparams = 0;
    double* Qup = nullptr;
    double* Qdn = nullptr;
    double* O = nullptr;
    int64_t failed_walker = -1;
    bool profiling = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // captured sweep sequence (kvmc_enable_graph)
    bool graph = false;
    hipGraphExec_t gexec = nullptr;
    int64_t g_nsweeps = -1;
    kvmc_profile prof{};
};

// Several handles on one GPU swept by one graph launch (kvmc_group_*).
struct kvmc_group {
    std::vector<kvmc_handle*> hs;
    hipStream_t s0 = nullptr;          // the stream the graph is captured on and launched on
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraphExec_t gexec = nullptr;
    std::vector<int64_t> g_nsweeps;    // the sweep counts the graph was captured for
};

namespace {

void free_all(kvmc_handle* h) {
    void* ptrs[] = {h->bonds, h->nbr_off, h->nbr, h->Uup, h->Udn, h->UTup, h->UTdn, h->Ucup, h->Ucdn, h->kRup,
                    h->kRdn, h->eloc_im, h->kup, h->kdn, h->W, h->slot, h->A,
                    h->Bm, h->npend, h->pterm, h->zmu, h->acc_since, h->status, h->sweeps, h->stop,
                    h->tbase, h->rlist, h->rcount, h->failed, h->all_walkers, h->counters,
                    h->logdet, h->eloc, h->drift, h->dstats, h->ws, h->Qup, h->Qdn, h->O};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

// Event pair around a launch sequence on the handle's stream; the elapsed
// time is read after the stream has been synchronised by the caller.
void prof_begin(kvmc_handle* h) {
    if (h->profiling) (void)hipEventRecord(h->ev0, h->stream);
}
void prof_end(kvmc_handle* h) {
    if (h->profiling) (void)hipEventRecord(h->ev1, h->stream);
}
float prof_elapsed(kvmc_handle* h) {
    float ms = 0.f;
    if (h->profiling && hipEventSynchronize(h->ev1) == hipSuccess)
        (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
    return ms;
}

SweepArgs sweep_args(kvmc_handle* h) {
    SweepArgs a{};
    a.ns = h->ns; a.nup = h->nup; a.ndn = h->ndn; a.nb = h->nb;
    a.nocc = std::min(h->nup, h->ndn);
    a.bonds = h->bonds; a.nbr_off = h->nbr_off; a.nbr = h->nbr;
    a.kup = h->kup; a.kdn = h->kdn; a.W = h->W; a.slot = h->slot; a.A = h->A; a.Bm = h->Bm;
    a.npend = h->npend; a.delay = h->delay; a.nmax = h->nmax;
    a.pterm = h->pterm;
    a.sweeps = h->sweeps; a.stop = h->stop; a.tape_base = h->tbase;
    a.zmu = h->zmu; a.acc_since = h->acc_since;
    a.refresh_list = h->rlist; a.refresh_count = h->rcount;
    a.counters = h->counters; a.seed = h->seed;
    a.kernel_mode = h->kernel; a.policy = h->policy; a.refresh_k = h->refresh_k;
    return a;
}

// Refreshes the walkers listed in d_list[0..count) (device array).  When
// write_w is 0 only the drift against the current W is measured.
RefreshArgs refresh_args(const kvmc_handle* h, int write_w, double pivot_tol) {
    RefreshArgs r{};
    const int f = h->cplx ? 2 : 1;   // complex orbitals: the real form's doubled sizes
    r.ns = f * h->ns; r.nup = f * h->nup; r.ndn = f * h->ndn; r.Uup = h->Uup; r.Udn = h->Udn;
    r.UTup = h->UTup; r.UTdn = h->UTdn;
    r.kup = h->cplx ? h->kRup : h->kup; r.kdn = h->cplx ? h->kRdn : h->kdn;
    r.W = h->W; r.slot = h->slot; r.A = h->A; r.Bm = h->Bm;
    r.npend = h->npend; r.acc_since = h->acc_since; r.delay = h->delay; r.nmax = f * h->nmax;
    r.ns_out = h->ns; r.nup_out = h->nup; r.ndn_out = h->ndn; r.nmax_out = h->nmax; r.cplx_out = h->cplx;
    r.ws = h->ws; r.logdet = h->logdet; r.status = h->status; r.failed = h->failed;
    r.write_w = write_w; r.drift = h->drift; r.pivot_tol = pivot_tol;
    return r;
}

// Complex orbitals: the refresh reads kappa on the 2 ns sites of the real form.
hipError_t prepare_refresh(kvmc_handle* h, hipStream_t s) {
    if (!h->cplx) return hipSuccess;
    hipError_t e = launch_kappa_real_form(h->kup, h->kRup, h->ns, h->nup, h->nw, s);
    return e != hipSuccess ? e : launch_kappa_real_form(h->kdn, h->kRdn, h->ns, h->ndn, h->nw, s);
}

int refresh_list(kvmc_handle* h, const int* d_list, int count, int write_w,
                 double pivot_tol = 0.0, bool drift_first = false) {
    if (count <= 0) return KVMC_OK;
    static const int kNone = INT_MAX;
    HIP_TRY(hipMemcpyAsync(h->failed, &kNone, sizeof(int), hipMemcpyHostToDevice, h->stream));
    const int batch_walkers = std::max(1, h->ws_jobs / 2);
    const int fb = (count + 255) / 256;
    HIP_TRY(prepare_refresh(h, h->stream));
    if (drift_first || !write_w) {
        hipLaunchKernelGGL(k_drift_clear, dim3(fb), dim3(256), 0, h->stream, d_list, count, h->drift);
        HIP_TRY(hipGetLastError());
    }
    if (drift_first) {
        // the carried W against a fresh FP64 refresh, just before it is replaced
        RefreshArgs d = refresh_args(h, 0, pivot_tol);
        for (int b0 = 0; b0 < count; b0 += batch_walkers) {
            const int nwb = std::min(batch_walkers, count - b0);
            d.jobs = d_list + b0;
            d.njobs = 2 * nwb;
            HIP_TRY(launch_refresh(h->prec, d, h->stream));
        }
        hipLaunchKernelGGL(k_drift_fold, dim3(fb), dim3(256), 0, h->stream, d_list, count, h->drift, h->dstats);
        HIP_TRY(hipGetLastError());
    }
    RefreshArgs r = refresh_args(h, write_w, pivot_tol);
    prof_begin(h);
    for (int b0 = 0; b0 < count; b0 += batch_walkers) {
        const int nwb = std::min(batch_walkers, count - b0);
        r.jobs = d_list + b0;
        r.njobs = 2 * nwb;
        HIP_TRY(launch_refresh(h->prec, r, h->stream));
        h->prof.refresh_launches += h->profiling;
    }
    prof_end(h);
    int failed = INT_MAX;
    HIP_TRY(hipMemcpyAsync(&failed, h->failed, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->prof.refresh_ms += prof_elapsed(h);
    if (failed != INT_MAX) {
        h->failed_walker = failed;
        return fail(KVMC_ERR_SINGULAR, "tildeU is singular for walker " + std::to_string(failed) +
                                           " (lu factorization failed)");
    }
    return KVMC_OK;
}

int refresh_all(kvmc_handle* h, int write_w) {
    HIP_TRY(hipMemsetAsync(h->status, 0, sizeof(int) * h->nw, h->stream));
    return refresh_list(h, h->all_walkers, h->nw, write_w);
}

int validate_configs(const kvmc_handle* h, const int32_t* ku, const int32_t* kd, int64_t w) {
    std::vector<char> seen_u(h->nup + 1, 0), seen_d(h->ndn + 1, 0);
    for (int s = 0; s < h->ns; ++s) {
        int lu = ku[s], ld = kd[s];
        if ((lu != 0) == (ld != 0))
            return fail(KVMC_ERR_INVALID, "walker " + std::to_string(w) + ": site " +
                                              std::to_string(s + 1) +
                                              ((lu != 0) ? " is doubly occupied" : " is unoccupied"));
        if (lu < 0 || lu > h->nup || ld < 0 || ld > h->ndn)
            return fail(KVMC_ERR_INVALID, "walker " + std::to_string(w) + ": label out of bounds");
        if (lu) { if (seen_u[lu]) return fail(KVMC_ERR_INVALID, "walker " + std::to_string(w) + ": repeated up label"); seen_u[lu] = 1; }
        if (ld) { if (seen_d[ld]) return fail(KVMC_ERR_INVALID, "walker " + std::to_string(w) + ": repeated down label"); seen_d[ld] = 1; }
    }
    for (int l = 1; l <= h->nup; ++l)
        if (!seen_u[l]) return fail(KVMC_ERR_INVALID, "walker " + std::to_string(w) + ": kappa_up is not valid");
    for (int l = 1; l <= h->ndn; ++l)
        if (!seen_d[l]) return fail(KVMC_ERR_INVALID, "walker " + std::to_string(w) + ": kappa_down is not valid");
    return KVMC_OK;
}

int upload_configs(kvmc_handle* h, const int32_t* ku, const int32_t* kd) {
    const size_t n = (size_t)h->nw * h->ns;
    HIP_TRY(hipMemcpyAsync(h->kup, ku, n * sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->kdn, kd, n * sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemsetAsync(h->npend, 0, sizeof(int) * h->nw, h->stream));
    HIP_TRY(hipMemsetAsync(h->acc_since, 0, sizeof(int) * h->nw, h->stream));
    return KVMC_OK;
}

int sweep_threads(const kvmc_handle* h) {
    const int V = (int)(16 / h->esize);   // W elements per 16-byte vector
    const int nvec = (h->ns % V == 0) ? h->ns / V : h->ns;
    // default: one wavefront when a column fits in it (more walkers per CU
    // hide the proposal latency), else 256 threads
    int t = h->threads > 0 ? h->threads : (nvec <= 64 ? 64 : 256);
    t = std::max(t, (nvec + 63) / 64 * 64);
    return std::min(t, 512);
}

// Real or complex instantiation of each kernel family.
hipError_t do_sweep(const kvmc_handle* h, const SweepArgs& a, int nw, int threads, size_t lds, hipStream_t s) {
    return h->cplx ? launch_sweep_c(h->prec, a, nw, threads, lds, s) : launch_sweep(h->prec, a, nw, threads, lds, s);
}
hipError_t do_flush(const kvmc_handle* h, const SweepArgs& a, int nw, hipStream_t s) {
    return h->cplx ? launch_flush_c(h->prec, a, nw, s) : launch_flush(h->prec, a, nw, s);
}
hipError_t do_eloc(const kvmc_handle* h, const ElocArgs& a, hipStream_t s) {
    return h->cplx ? launch_eloc_c(h->prec, a, s) : launch_eloc(h->prec, a, s);
}
hipError_t do_grad(const kvmc_handle* h, const ElocArgs& a, hipStream_t s) {
    return h->cplx ? launch_grad_c(h->prec, a, s) : launch_grad(h->prec, a, s);
}

}  // namespace

extern "C" {

const char* kvmc_last_error(void) { return g_err.c_str(); }

const char* kvmc_version(void) { return "kagome_vmc 0.1 gfx950"; }

int64_t kvmc_failed_walker(kvmc_handle* h) { return h ? h->failed_walker : -1; }

int kvmc_create(const kvmc_desc* d, kvmc_handle** out) {
    g_err.clear();
    if (!d || !out) return fail(KVMC_ERR_INVALID, "null descriptor or output");
    *out = nullptr;
    if (d->ns <= 0 || d->n_up <= 0 || d->n_down <= 0 || d->n_up + d->n_down != d->ns)
        return fail(KVMC_ERR_INVALID, "need n_up + n_down == ns (one spinon per site)");
    if (d->ns > 32767 || d->nb > 32767)
        return fail(KVMC_ERR_INVALID, "more than 32767 sites or bonds (16-bit site, label and bond tables)");
    if (d->nb <= 0 || !d->bonds || !d->u_up || !d->u_down)
        return fail(KVMC_ERR_INVALID, "bonds and orbitals are required");
    if (d->nwalkers <= 0 || d->nwalkers > INT_MAX / 2)
        return fail(KVMC_ERR_INVALID, "nwalkers out of range");
    if (d->precision != KVMC_PREC_FP64 && d->precision != KVMC_PREC_FP32)
        return fail(KVMC_ERR_INVALID, "unknown precision");
    if (d->kernel != KVMC_KERNEL_REFERENCE && d->kernel != KVMC_KERNEL_UNBIASED)
        return fail(KVMC_ERR_INVALID, "unknown transition kernel");
    if (d->refresh_policy != KVMC_REFRESH_REFERENCE && d->refresh_policy != KVMC_REFRESH_EVERY_K &&
        d->refresh_policy != KVMC_REFRESH_AT_END)
        return fail(KVMC_ERR_INVALID, "unknown refresh policy");
    if (d->refresh_policy != KVMC_REFRESH_REFERENCE && d->refresh_k <= 0)
        return fail(KVMC_ERR_INVALID, "refresh_k must be positive");
    const int delay = d->delay <= 1 ? 1 : d->delay;
    if (delay > kMaxDelay)
        return fail(KVMC_ERR_INVALID, "delay must be <= " + std::to_string(kMaxDelay));
    if (d->orbitals != KVMC_ORB_REAL && d->orbitals != KVMC_ORB_COMPLEX)
        return fail(KVMC_ERR_INVALID, "unknown orbital type");
    const int cf = d->orbitals == KVMC_ORB_COMPLEX ? 2 : 1;   // the refresh factors a 2N x 2N real form
    if (d->ns > 4096) return fail(KVMC_ERR_INVALID, "ns > 4096 is not supported");
    if (!refresh_supported(cf * d->ns, cf * d->n_up, cf * d->n_down))
        return fail(KVMC_ERR_INVALID, "lattice too large for the refresh kernel (LDS, or more than 768 "
                                      "electrons per spin; 384 with complex orbitals)");
    for (int b = 0; b < d->nb; ++b) {
        int i = d->bonds[2 * b], j = d->bonds[2 * b + 1];
        if (i < 1 || j < 1 || i > d->ns || j > d->ns || i > j)
            return fail(KVMC_ERR_INVALID, "bond " + std::to_string(b + 1) + " out of range");
    }
    kvmc_handle* h = new kvmc_handle();
    h->ns = d->ns; h->nup = d->n_up; h->ndn = d->n_down; h->nb = d->nb;
    h->nmax = std::max(h->nup, h->ndn);
    h->nsp = (h->ns + 3) / 4 * 4;
    h->nw = (int)d->nwalkers;
    h->prec = d->precision; h->kernel = d->kernel; h->policy = d->refresh_policy;
    h->refresh_k = d->refresh_k; h->delay = delay; h->threads = d->threads; h->seed = d->seed;
    h->cplx = d->orbitals == KVMC_ORB_COMPLEX ? 1 : 0;
    h->esize = w_elem_bytes(h->prec, h->cplx);
    (void)hipGetDevice(&h->device);
    int rc = KVMC_OK;
#define ALLOC(p, n) do { if ((rc = dalloc(&(p), (n))) != KVMC_OK) { free_all(h); delete h; return rc; } } while (0)
    hipError_t e = hipStreamCreate(&h->stream);
    if (e != hipSuccess) { delete h; return fail(KVMC_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
    const size_t nw = (size_t)h->nw, ns = (size_t)h->ns;
    ALLOC(h->bonds, (size_t)h->nb);
    ALLOC(h->nbr_off, ns + 1);
    ALLOC(h->nbr, (size_t)2 * h->nb);
    const size_t f2 = (size_t)cf * cf;   // the real form has cf^2 times the entries
    ALLOC(h->Uup, f2 * ns * h->nup);
    ALLOC(h->Udn, f2 * ns * h->ndn);
    ALLOC(h->UTup, f2 * ns * h->nup);
    ALLOC(h->UTdn, f2 * ns * h->ndn);
    if (h->cplx) {
        ALLOC(h->Ucup, 2 * ns * h->nup);
        ALLOC(h->Ucdn, 2 * ns * h->ndn);
        ALLOC(h->kRup, 2 * nw * ns);
        ALLOC(h->kRdn, 2 * nw * ns);
        ALLOC(h->eloc_im, nw);
    }
    ALLOC(h->kup, nw * ns);
    ALLOC(h->kdn, nw * ns);
    ALLOC(h->slot, nw * 2 * ns);
    // zero until the first refresh writes it: every entry stays a row of W or -1
    if (hipMemset(h->slot, 0, nw * 2 * ns * sizeof(int)) != hipSuccess) {
        free_all(h); delete h; return fail(KVMC_ERR_HIP, "hipMemset of the slot map failed");
    }
    {
        unsigned char* p = nullptr;
        const size_t wbytes = nw * w_stride(h->ns, h->nup, h->ndn) * h->esize;
        ALLOC(p, wbytes); h->W = p;
        // zero once: the padding rows of W must stay zero (no pending term touches them)
        if (hipMemset(p, 0, wbytes) != hipSuccess) { free_all(h); delete h; return fail(KVMC_ERR_HIP, "hipMemset of W failed"); }
        ALLOC(p, nw * 2 * pend_ld(delay) * (size_t)w_nsa(h->ns, h->nup, h->ndn) * h->esize); h->A = p;
        ALLOC(p, nw * 2 * pend_ld(delay) * h->nmax * h->esize); h->Bm = p;
    }
    ALLOC(h->npend, nw); ALLOC(h->pterm, nw * 2 * kMaxDelay * 2); ALLOC(h->zmu, nw); ALLOC(h->acc_since, nw); ALLOC(h->status, nw);
    ALLOC(h->sweeps, nw); ALLOC(h->stop, nw); ALLOC(h->tbase, nw);
    ALLOC(h->rlist, nw); ALLOC(h->rcount, 1); ALLOC(h->failed, 1); ALLOC(h->all_walkers, nw);
    ALLOC(h->counters, (size_t)CNT_SLOTS);
    ALLOC(h->logdet, 2 * nw); ALLOC(h->eloc, nw); ALLOC(h->drift, 2 * nw); ALLOC(h->dstats, 3);
    // refresh workspace: up to 16 GiB of FP64 work matrices per batch (about
    // 10000 walkers at 432 sites, so one launch pair covers even a refresh of
    // every walker)
    {
        const size_t per_job = refresh_ws_doubles_per_job(cf * h->ns, cf * h->nmax) * sizeof(double);
        size_t jobs = std::max<size_t>(2, ((size_t)16 << 30) / per_job);
        jobs = std::min(jobs, 2 * nw);
        h->ws_jobs = (int)(jobs & ~(size_t)1);
        if (h->ws_jobs < 2) h->ws_jobs = 2;
        ALLOC(h->ws, (size_t)h->ws_jobs * refresh_ws_doubles_per_job(cf * h->ns, cf * h->nmax));
    }
#undef ALLOC
    // bonds (0-based) and site -> bond adjacency
    std::vector<int2> hb(h->nb);
    std::vector<std::vector<int>> adj(ns);
    for (int b = 0; b < h->nb; ++b) {
        hb[b] = make_int2(d->bonds[2 * b] - 1, d->bonds[2 * b + 1] - 1);
        adj[hb[b].x].push_back(b);
        if (hb[b].y != hb[b].x) adj[hb[b].y].push_back(b);
    }
    std::vector<int> off(ns + 1, 0), lst;
    for (size_t s = 0; s < ns; ++s) {
        off[s + 1] = off[s] + (int)adj[s].size();
        lst.insert(lst.end(), adj[s].begin(), adj[s].end());
    }
    // U of the refresh (column-major) and its row-major copy.  Complex
    // orbitals (interleaved ns x N): the 2 ns x 2 N real form.
    auto real_form = [&](const double* u, int n, std::vector<double>& cm, std::vector<double>& rm) {
        const size_t R = (size_t)cf * ns, C = (size_t)cf * n;
        cm.assign(R * C, 0.0);
        for (size_t c = 0; c < (size_t)n; ++c)
            for (size_t s2 = 0; s2 < ns; ++s2) {
                if (!h->cplx) {
                    cm[c * R + s2] = u[c * ns + s2];
                } else {
                    const double re = u[2 * (c * ns + s2)], im = u[2 * (c * ns + s2) + 1];
                    cm[c * R + s2] = re;                // [U_re, -U_im ; U_im, U_re]
                    cm[(c + n) * R + s2] = -im;
                    cm[c * R + ns + s2] = im;
                    cm[(c + n) * R + ns + s2] = re;
                }
            }
        rm.resize(R * C);
        for (size_t r2 = 0; r2 < R; ++r2)
            for (size_t c = 0; c < C; ++c) rm[r2 * C + c] = cm[c * R + r2];
    };
    std::vector<double> ucu, ucd, utu, utd;
    real_form(d->u_up, h->nup, ucu, utu);
    real_form(d->u_down, h->ndn, ucd, utd);
    std::vector<int> allw(nw);
    for (size_t w = 0; w < nw; ++w) allw[w] = (int)w;
    auto up = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream);
    };
    if (up(h->bonds, hb.data(), hb.size() * sizeof(int2)) != hipSuccess ||
        up(h->nbr_off, off.data(), off.size() * sizeof(int)) != hipSuccess ||
        (lst.size() && up(h->nbr, lst.data(), lst.size() * sizeof(int)) != hipSuccess) ||
        up(h->Uup, ucu.data(), ucu.size() * sizeof(double)) != hipSuccess ||
        up(h->Udn, ucd.data(), ucd.size() * sizeof(double)) != hipSuccess ||
        (h->cplx && up(h->Ucup, d->u_up, 2 * ns * h->nup * sizeof(double)) != hipSuccess) ||
        (h->cplx && up(h->Ucdn, d->u_down, 2 * ns * h->ndn * sizeof(double)) != hipSuccess) ||
        up(h->UTup, utu.data(), utu.size() * sizeof(double)) != hipSuccess ||
        up(h->UTdn, utd.data(), utd.size() * sizeof(double)) != hipSuccess ||
        up(h->all_walkers, allw.data(), nw * sizeof(int)) != hipSuccess ||
        hipMemsetAsync(h->sweeps, 0, nw * sizeof(long long), h->stream) != hipSuccess ||
        hipMemsetAsync(h->npend, 0, nw * sizeof(int), h->stream) != hipSuccess ||
        hipMemsetAsync(h->acc_since, 0, nw * sizeof(int), h->stream) != hipSuccess ||
        hipMemsetAsync(h->logdet, 0, 2 * nw * sizeof(double), h->stream) != hipSuccess ||
        hipMemsetAsync(h->dstats, 0, 3 * sizeof(double), h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
        free_all(h);
        delete h;
        return fail(KVMC_ERR_HIP, "uploading lattice data failed");
    }
    *out = h;
    return KVMC_OK;
}

int kvmc_destroy(kvmc_handle* h) {
    if (!h) return KVMC_OK;
    (void)hipStreamSynchronize(h->stream);
    free_all(h);
    delete h;
    return KVMC_OK;
}

int kvmc_set_configs(kvmc_handle* h, const int32_t* ku, const int32_t* kd) {
    g_err.clear();
    if (!h || !ku || !kd) return fail(KVMC_ERR_INVALID, "null argument");
    for (int w = 0; w < h->nw; ++w) {
        int rc = validate_configs(h, ku + (size_t)w * h->ns, kd + (size_t)w * h->ns, w);
        if (rc) return rc;
    }
    int rc = upload_configs(h, ku, kd);
    if (rc) return rc;
    h->failed_walker = -1;
    rc = refresh_all(h, 1);
    h->ready = rc == KVMC_OK;
    return rc;
}

int kvmc_get_configs(kvmc_handle* h, int32_t* ku, int32_t* kd) {
    g_err.clear();
    if (!h || !ku || !kd) return fail(KVMC_ERR_INVALID, "null argument");
    const size_t n = (size_t)h->nw * h->ns;
    HIP_TRY(hipMemcpyAsync(ku, h->kup, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(kd, h->kdn, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return KVMC_OK;
}

int kvmc_init_random(kvmc_handle* h, uint64_t seed, int32_t max_attempts, int64_t* redraws) {
    g_err.clear();
    if (!h) return fail(KVMC_ERR_INVALID, "null handle");
    if (max_attempts < 1) max_attempts = 1;
    const size_t ns = h->ns;
    std::vector<int> ku((size_t)h->nw * ns), kd((size_t)h->nw * ns);
    for (int w = 0; w < h->nw; ++w)
        random_config(seed, (uint32_t)w, 0u, h->ns, h->nup, h->ndn, &ku[w * ns], &kd[w * ns]);
    int rc = upload_configs(h, ku.data(), kd.data());
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(h->status, 0, sizeof(int) * h->nw, h->stream));
    const double start_tol = h->prec == KVMC_PREC_FP32 ? kStartPivotTolF32 : kStartPivotTol;
    rc = refresh_list(h, h->all_walkers, h->nw, 1, start_tol);
    if (rc != KVMC_OK && rc != KVMC_ERR_SINGULAR) return rc;
    int64_t nredraw = 0;
    std::vector<int> status(h->nw);
    for (int attempt = 1; rc == KVMC_ERR_SINGULAR; ++attempt) {
        HIP_TRY(hipMemcpyAsync(status.data(), h->status, sizeof(int) * h->nw, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        std::vector<int> bad;
        for (int w = 0; w < h->nw; ++w)
            if (status[w]) bad.push_back(w);
        if (bad.empty()) { rc = KVMC_OK; break; }
        if (attempt >= max_attempts)
            return fail(KVMC_ERR_SINGULAR, std::to_string(bad.size()) +
                                               " walkers still singular after " +
                                               std::to_string(max_attempts) + " draws");
        for (int w : bad)
            random_config(seed, (uint32_t)w, (uint32_t)attempt, h->ns, h->nup, h->ndn, &ku[w * ns], &kd[w * ns]);
        HIP_TRY(hipMemcpyAsync(h->kup, ku.data(), ku.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipMemcpyAsync(h->kdn, kd.data(), kd.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
        nredraw += (int64_t)bad.size();
        DevBuf<int> dbad;
        if ((rc = dbad.alloc(bad.size())) != KVMC_OK) return rc;
        HIP_TRY(hipMemcpyAsync(dbad.p, bad.data(), bad.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
        for (int w : bad) status[w] = 0;
        HIP_TRY(hipMemcpyAsync(h->status, status.data(), sizeof(int) * h->nw, hipMemcpyHostToDevice, h->stream));
        rc = refresh_list(h, dbad.p, (int)bad.size(), 1, start_tol);
        HIP_TRY(hipStreamSynchronize(h->stream));
        if (rc != KVMC_OK && rc != KVMC_ERR_SINGULAR) return rc;
    }
    if (redraws) *redraws = nredraw;
    h->failed_walker = -1;
    h->ready = true;
    g_err.clear();
    return KVMC_OK;
}

int kvmc_refresh(kvmc_handle* h) {
    g_err.clear();
    if (!h) return fail(KVMC_ERR_INVALID, "null handle");
    if (!h->ready) return fail(KVMC_ERR_NOT_READY, "configurations not set");
    return refresh_all(h, 1);
}

}  // extern "C"

namespace {

// Argument checks shared by the sweep entry points.
int sweep_check(kvmc_handle* h, int64_t nsweeps, const double* tape_u, const int32_t* tape_bond) {
    if (!h) return fail(KVMC_ERR_INVALID, "null handle");
    if (!h->ready) return fail(KVMC_ERR_NOT_READY, "configurations not set");
    if (nsweeps < 0) return fail(KVMC_ERR_INVALID, "nsweeps < 0");
    if ((tape_u == nullptr) != (tape_bond == nullptr))
        return fail(KVMC_ERR_INVALID, "tape_u and tape_bond go together");
    return KVMC_OK;
}

int tape_check(const kvmc_handle* h, size_t tape_n, const double* tape_u, const int32_t* tape_bond) {
    if (tape_u)
        for (size_t q = 0; q < tape_n; ++q)
            if (tape_bond[q] < 0 || tape_bond[q] >= h->nb || !(tape_u[q] >= 0.0 && tape_u[q] < 1.0))
                return fail(KVMC_ERR_INVALID, "tape entry " + std::to_string(q) + " out of range");
    return KVMC_OK;
}

// Enqueues one captured sweep sequence of handle h on stream s: stop counts,
// counter and flag resets (fill kernels, no memset nodes), then
// kGraphRounds rounds of the sweep kernel, each but the last followed by
// the refresh of the walkers it paused, their count read on the device.
hipError_t enqueue_rounds(kvmc_handle* h, const SweepArgs& a, int threads, size_t lds, int64_t nsweeps,
                          hipStream_t s) {
    const int fb = (h->nw + 255) / 256;
    hipLaunchKernelGGL(k_set_stop, dim3(fb), dim3(256), 0, s, h->sweeps, h->stop, h->tbase, (long long)nsweeps,
                       h->nw);
    hipLaunchKernelGGL(k_fill_ull, dim3(1), dim3(64), 0, s, h->counters, 0ull, (int)CNT_SLOTS);
    hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, s, h->failed, INT_MAX, 1);
    hipError_t ce = hipGetLastError();
    for (int r = 0; r < kGraphRounds && ce == hipSuccess; ++r) {
        hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, s, h->rcount, 0, 1);
        ce = do_sweep(h, a, h->nw, threads, lds, s);
        if (r + 1 == kGraphRounds || ce != hipSuccess) break;
        RefreshArgs ra = refresh_args(h, 1, 0.0);
        ce = prepare_refresh(h, s);
        const int batch = std::max(1, h->ws_jobs / 2);
        for (int b0 = 0; b0 < h->nw && ce == hipSuccess; b0 += batch) {
            ra.jobs = h->rlist + b0;
            ra.njobs = 2 * std::min(batch, h->nw - b0);
            ra.count_dev = h->rcount;
            ra.walker_off = b0;
            ce = launch_refresh(h->prec, ra, s);
        }
    }
    return ce;
}

// After a replay of enqueue_rounds: reports a singular refresh, refreshes
// the walkers the last round paused; *finished when nothing is left to run.
int after_replay(kvmc_handle* h, bool* finished) {
    hipStream_t s = h->stream;
    int count = 0, failed = INT_MAX;
    HIP_TRY(hipMemcpyAsync(&count, h->rcount, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&failed, h->failed, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (failed != INT_MAX) {
        h->failed_walker = failed;
        return fail(KVMC_ERR_SINGULAR, "tildeU is singular for walker " + std::to_string(failed) +
                                           " (lu factorization failed)");
    }
    *finished = count == 0;
    if (count > 0) {
        int rc = refresh_list(h, h->rlist, count, 1, 0.0, h->track_drift);
        if (rc) return rc;
        *finished = h->policy == KVMC_REFRESH_AT_END;
    }
    return KVMC_OK;
}

// The ordinary launch loop (sweep launch, refresh of the paused walkers,
// again until no walker pauses), then the outcome copy and the counters.
int sweep_tail(kvmc_handle* h, const SweepArgs& a, int threads, size_t lds, bool finished, int64_t launches,
               int8_t* outcomes, const signed char* d_out, size_t tape_n, kvmc_counters* counters) {
    hipStream_t s = h->stream;
    while (!finished) {
        int count = 0;
        if (hipMemsetAsync(h->rcount, 0, sizeof(int), s) != hipSuccess) return fail(KVMC_ERR_HIP, "hipMemsetAsync failed");
        prof_begin(h);
        hipError_t e = do_sweep(h, a, h->nw, threads, lds, s);
        if (e != hipSuccess) return fail(KVMC_ERR_HIP, std::string("sweep launch: ") + hipGetErrorString(e));
        prof_end(h);
        ++launches;
        if (hipMemcpyAsync(&count, h->rcount, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return fail(KVMC_ERR_HIP, std::string("sweep kernel: ") + hipGetErrorString(hipGetLastError()));
        h->prof.sweep_ms += prof_elapsed(h);
        h->prof.sweep_launches += h->profiling;
        if (count == 0) break;
        int rc = refresh_list(h, h->rlist, count, 1, 0.0, h->track_drift);
        if (rc) return rc;
        if (h->policy == KVMC_REFRESH_AT_END) break;   // listed walkers had finished
    }
    if (outcomes && hipMemcpyAsync(outcomes, d_out, tape_n, hipMemcpyDeviceToHost, s) != hipSuccess)
        return fail(KVMC_ERR_HIP, "outcome copy failed");
    unsigned long long cnt[CNT_SLOTS] = {};
    if (hipMemcpyAsync(cnt, h->counters, sizeof(cnt), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(KVMC_ERR_HIP, "counter copy failed");
    if (counters) {
        counters->sweeps = (int64_t)cnt[CNT_SWEEPS];
        counters->moves = (int64_t)cnt[CNT_MOVES];
        counters->accepted = (int64_t)cnt[CNT_ACCEPTED];
        counters->refreshes = (int64_t)cnt[CNT_PAUSES];
        counters->flushes = (int64_t)cnt[CNT_FLUSHES];
        counters->launches = launches;
    }
    return KVMC_OK;
}

// The sweep calls with the tape and the outcome record staged in the given
// device buffers (each used only when its host pointer is set): uploads the
// tape, runs the launch / refresh loop (or the captured graph when
// allow_graph), copies the outcomes back and reads the counters.  Allocates
// no device memory.
int sweep_staged(kvmc_handle* h, int64_t nsweeps, const double* tape_u, const int32_t* tape_bond,
                 int8_t* outcomes, kvmc_counters* counters, double* d_tu, int* d_tb, signed char* d_out,
                 bool allow_graph) {
    const size_t tape_n = (size_t)h->nw * (size_t)nsweeps;
    if (tape_u) {
        if (hipMemcpyAsync(d_tu, tape_u, tape_n * sizeof(double), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
            hipMemcpyAsync(d_tb, tape_bond, tape_n * sizeof(int), hipMemcpyHostToDevice, h->stream) != hipSuccess)
            return fail(KVMC_ERR_HIP, "tape upload failed");
    }
    SweepArgs a = sweep_args(h);
    a.tape_u = tape_u ? d_tu : nullptr; a.tape_bond = tape_u ? d_tb : nullptr; a.tape_len = nsweeps;
    a.outcomes = outcomes ? d_out : nullptr;
    a.walker0 = 0;
    const int threads = sweep_threads(h);
    const size_t lds = sweep_lds_bytes(h->esize, h->ns, h->nup, h->ndn, h->delay, h->nb);
    if (lds > 160 * 1024) return fail(KVMC_ERR_INVALID, "LDS footprint exceeds 160 KiB; lower delay");
    int64_t launches = 0;
    hipStream_t s = h->stream;
    const bool use_graph = allow_graph && h->graph && !tape_u && !outcomes && !h->profiling && !h->track_drift;
    bool finished = false;   // the replay left no walker paused
    if (use_graph) {
        if (!h->gexec || h->g_nsweeps != nsweeps) {
            if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
            hipError_t ce = enqueue_rounds(h, a, threads, lds, nsweeps, s);
            hipError_t ee = hipStreamEndCapture(s, &graph);
            if (ce != hipSuccess || ee != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return fail(KVMC_ERR_HIP, std::string("sweep graph capture: ") +
                                              hipGetErrorString(ce != hipSuccess ? ce : ee));
            }
            hipError_t ie = hipGraphInstantiate(&h->gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) { h->gexec = nullptr; return fail(KVMC_ERR_HIP, "sweep graph instantiate failed"); }
            h->g_nsweeps = nsweeps;
        }
        HIP_TRY(hipGraphLaunch(h->gexec, s));
        launches += kGraphRounds;
        int rc = after_replay(h, &finished);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL(k_set_stop, dim3((h->nw + 255) / 256), dim3(256), 0, s,
                           h->sweeps, h->stop, h->tbase, (long long)nsweeps, h->nw);
        if (hipMemsetAsync(h->counters, 0, sizeof(unsigned long long) * CNT_SLOTS, s) != hipSuccess)
            return fail(KVMC_ERR_HIP, "counter reset failed");
    }
    return sweep_tail(h, a, threads, lds, finished, launches, outcomes, d_out, tape_n, counters);
}

int scratch_address(const kvmc_handle* h, unsigned char** out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != h->device) return fail(KVMC_ERR_INVALID, "the handle's device must be current");
    if (dev < 0 || dev >= kMaxDevices) return fail(KVMC_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(KVMC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_addr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_addr[dev]);
    return KVMC_OK;
}

}  // namespace

extern "C" {

int kvmc_sweep_noalloc(kvmc_handle* h, int64_t nsweeps, const double* tape_u, const int32_t* tape_bond,
                       int8_t* outcomes, kvmc_counters* counters) {
    g_err.clear();
    int rc = sweep_check(h, nsweeps, tape_u, tape_bond);
    if (rc) return rc;
    if (nsweeps == 0) { if (counters) *counters = kvmc_counters{}; return KVMC_OK; }
    const size_t tape_n = (size_t)h->nw * (size_t)nsweeps;
    if ((tape_u || outcomes) && tape_n > kScratchMaxEntries)
        return fail(KVMC_ERR_INVALID, "nwalkers * nsweeps = " + std::to_string(tape_n) +
                                          " exceeds the " + std::to_string(kScratchMaxEntries) +
                                          " entries of the 32 MiB device scratch");
    if ((rc = tape_check(h, tape_n, tape_u, tape_bond))) return rc;
    unsigned char* base = nullptr;
    if ((rc = scratch_address(h, &base))) return rc;
    auto up256 = [](size_t b) { return (b + 255) / 256 * 256; };
    double* d_tu = reinterpret_cast<double*>(base);
    int* d_tb = reinterpret_cast<int*>(base + up256(tape_n * sizeof(double)));
    signed char* d_out = reinterpret_cast<signed char*>(base + up256(tape_n * sizeof(double)) +
                                                        up256(tape_n * sizeof(int)));
    if (outcomes) {
        // the scratch keeps the previous call's contents: clear the record
        hipLaunchKernelGGL(k_fill_i8, dim3((unsigned)((tape_n + 255) / 256)), dim3(256), 0, h->stream, d_out,
                           (signed char)0, (long long)tape_n);
        HIP_TRY(hipGetLastError());
    }
    // graph replay stays off: instantiating a graph may allocate
    return sweep_staged(h, nsweeps, tape_u, tape_bond, outcomes, counters, d_tu, d_tb, d_out, false);
}

int kvmc_group_create(kvmc_handle* const* handles, int32_t n, kvmc_group** out) {
    g_err.clear();
    if (!handles || !out || n < 1) return fail(KVMC_ERR_INVALID, "bad group arguments");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!handles[i]) return fail(KVMC_ERR_INVALID, "null handle in group");
        if (handles[i]->device != handles[0]->device) return fail(KVMC_ERR_INVALID, "group handles on different devices");
        for (int j = 0; j < i; ++j)
            if (handles[j] == handles[i]) return fail(KVMC_ERR_INVALID, "handle listed twice in group");
    }
    kvmc_group* g = new kvmc_group();
    g->hs.assign(handles, handles + n);
    g->join.assign(n, nullptr);
    bool ok = hipStreamCreate(&g->s0) == hipSuccess &&
              hipEventCreateWithFlags(&g->fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < n; ++i) ok = hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        kvmc_group_destroy(g);
        return fail(KVMC_ERR_HIP, "group stream or event creation failed");
    }
    *out = g;
    return KVMC_OK;
}

int kvmc_group_destroy(kvmc_group* g) {
    if (!g) return KVMC_OK;
    if (g->s0) (void)hipStreamSynchronize(g->s0);
    if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
    for (hipEvent_t e : g->join)
        if (e) (void)hipEventDestroy(e);
    if (g->fork) (void)hipEventDestroy(g->fork);
    if (g->s0) (void)hipStreamDestroy(g->s0);
    delete g;
    return KVMC_OK;
}

int kvmc_group_sweep(kvmc_group* g, const int64_t* nsweeps, kvmc_counters* counters) {
    g_err.clear();
    if (!g || !nsweeps) return fail(KVMC_ERR_INVALID, "null group or sweep counts");
    const int n = (int)g->hs.size();
    bool plain = false;   // profiling or drift tracking: each handle sweeps on its own
    for (int i = 0; i < n; ++i) {
        int rc = sweep_check(g->hs[i], nsweeps[i], nullptr, nullptr);
        if (rc) return rc;
        plain = plain || g->hs[i]->profiling || g->hs[i]->track_drift;
    }
    if (plain) {
        for (int i = 0; i < n; ++i) {
            int rc = nsweeps[i] ? sweep_staged(g->hs[i], nsweeps[i], nullptr, nullptr, nullptr,
                                               counters ? counters + i : nullptr, nullptr, nullptr, nullptr, false)
                                : KVMC_OK;
            if (rc) return rc;
            if (!nsweeps[i] && counters) counters[i] = kvmc_counters{};
        }
        return KVMC_OK;
    }
    std::vector<SweepArgs> args(n);
    std::vector<int> threads(n);
    std::vector<size_t> lds(n);
    for (int i = 0; i < n; ++i) {
        kvmc_handle* h = g->hs[i];
        args[i] = sweep_args(h);
        threads[i] = sweep_threads(h);
        lds[i] = sweep_lds_bytes(h->esize, h->ns, h->nup, h->ndn, h->delay, h->nb);
        if (lds[i] > 160 * 1024) return fail(KVMC_ERR_INVALID, "LDS footprint exceeds 160 KiB; lower delay");
    }
    if (!g->gexec || g->g_nsweeps != std::vector<int64_t>(nsweeps, nsweeps + n)) {
        if (g->gexec) { (void)hipGraphExecDestroy(g->gexec); g->gexec = nullptr; }
        // fork: every handle's stream joins the capture of s0, enqueues its
        // rounds, and s0 joins back on its event
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(g->s0, hipStreamCaptureModeRelaxed));
        hipError_t ce = hipEventRecord(g->fork, g->s0);
        for (int i = 0; i < n && ce == hipSuccess; ++i) {
            if (!nsweeps[i]) continue;
            kvmc_handle* h = g->hs[i];
            ce = hipStreamWaitEvent(h->stream, g->fork, 0);
            if (ce == hipSuccess) ce = enqueue_rounds(h, args[i], threads[i], lds[i], nsweeps[i], h->stream);
            if (ce == hipSuccess) ce = hipEventRecord(g->join[i], h->stream);
            if (ce == hipSuccess) ce = hipStreamWaitEvent(g->s0, g->join[i], 0);
        }
        hipError_t ee = hipStreamEndCapture(g->s0, &graph);
        if (ce != hipSuccess || ee != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(KVMC_ERR_HIP, std::string("group graph capture: ") +
                                          hipGetErrorString(ce != hipSuccess ? ce : ee));
        }
        hipError_t ie = hipGraphInstantiate(&g->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) { g->gexec = nullptr; return fail(KVMC_ERR_HIP, "group graph instantiate failed"); }
        g->g_nsweeps.assign(nsweeps, nsweeps + n);
    }
    HIP_TRY(hipGraphLaunch(g->gexec, g->s0));
    HIP_TRY(hipStreamSynchronize(g->s0));
    // walkers the last round paused finish on their own handle's stream
    for (int i = 0; i < n; ++i) {
        kvmc_handle* h = g->hs[i];
        if (!nsweeps[i]) { if (counters) counters[i] = kvmc_counters{}; continue; }
        bool finished = false;
        int rc = after_replay(h, &finished);
        if (rc) return rc;
        rc = sweep_tail(h, args[i], threads[i], lds[i], finished, kGraphRounds, nullptr, nullptr, 0,
                        counters ? counters + i : nullptr);
        if (rc) return rc;
    }
    return KVMC_OK;
}

int kvmc_sweep(kvmc_handle* h, int64_t nsweeps, const double* tape_u, const int32_t* tape_bond,
               int8_t* outcomes, kvmc_counters* counters) {
    g_err.clear();
    int rc = sweep_check(h, nsweeps, tape_u, tape_bond);
    if (rc) return rc;
    if (nsweeps == 0) { if (counters) *counters = kvmc_counters{}; return KVMC_OK; }
    const size_t tape_n = (size_t)h->nw * (size_t)nsweeps;
    if ((rc = tape_check(h, tape_n, tape_u, tape_bond))) return rc;
    DevBuf<double> d_tu;
    DevBuf<int> d_tb;
    DevBuf<signed char> d_out;
    if (tape_u && ((rc = d_tu.alloc(tape_n)) || (rc = d_tb.alloc(tape_n)))) return rc;
    if (outcomes && (rc = d_out.alloc(tape_n))) return rc;
    return sweep_staged(h, nsweeps, tape_u, tape_bond, outcomes, counters, d_tu.p, d_tb.p, d_out.p, true);
}

static ElocArgs eloc_args(kvmc_handle* h) {
    ElocArgs e{};
    e.ns = h->ns; e.nup = h->nup; e.ndn = h->ndn; e.nb = h->nb;
    e.bonds = h->bonds; e.kup = h->kup; e.kdn = h->kdn; e.W = h->W; e.slot = h->slot; e.A = h->A; e.Bm = h->Bm;
    e.npend = h->npend; e.delay = h->delay; e.nmax = h->nmax; e.nw = h->nw;
    e.eloc = h->eloc; e.eloc_im = h->eloc_im; e.counters = h->counters;
    e.nparams = h->nparams; e.Qup = h->Qup; e.Qdn = h->Qdn; e.O = h->O;
    return e;
}

int kvmc_measure(kvmc_handle* h, double* eloc, double* sum, int64_t* n_ap) {
    g_err.clear();
    if (!h) return fail(KVMC_ERR_INVALID, "null handle");
    if (!h->ready) return fail(KVMC_ERR_NOT_READY, "configurations not set");
    hipStream_t s = h->stream;
    HIP_TRY(hipMemsetAsync(h->counters + CNT_NAP, 0, sizeof(unsigned long long), s));
    prof_begin(h);
    HIP_TRY(do_eloc(h, eloc_args(h), s));
    prof_end(h);
    std::vector<double> tmp;
    double* dst = eloc;
    if (!dst && sum) { tmp.resize(h->nw); dst = tmp.data(); }
    if (dst) HIP_TRY(hipMemcpyAsync(dst, h->eloc, sizeof(double) * h->nw, hipMemcpyDeviceToHost, s));
    unsigned long long nap = 0;
    HIP_TRY(hipMemcpyAsync(&nap, h->counters + CNT_NAP, sizeof(nap), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    h->prof.measure_ms += prof_elapsed(h);
    h->prof.measure_launches += h->profiling;
    if (sum) {
        double acc = 0.0;
        for (int w = 0; w < h->nw; ++w) acc += dst[w];
        *sum = acc;
    }
    if (n_ap) *n_ap = (int64_t)nap;
    return KVMC_OK;
}

int kvmc_set_orbital_derivs(kvmc_handle* h, int32_t nparams, const double* du_up,
                            const double* du_dn) {
    g_err.clear();
    if (!h || nparams < 0 || (nparams > 0 && (!du_up || !du_dn)))
        return fail(KVMC_ERR_INVALID, "bad orbital-derivative arguments");
    if (h->Qup) { (void)hipFree(h->Qup); h->Qup = nullptr; }
    if (h->Qdn) { (void)hipFree(h->Qdn); h->Qdn = nullptr; }
    if (h->O) { (void)hipFree(h->O); h->O = nullptr; }
    h->nparams = 0;
    if (nparams == 0) return KVMC_OK;
    const size_t ns = h->ns;
    const size_t cw = h->cplx ? 2 : 1;   // doubles per value (complex: interleaved)
    int rc;
    // built in scope-owned buffers; the handle takes them only on success
    DevBuf<double> qu, qd, o, dU;
    if ((rc = qu.alloc(cw * nparams * ns * ns)) || (rc = qd.alloc(cw * nparams * ns * ns)) ||
        (rc = o.alloc(cw * nparams * h->nw)) || (rc = dU.alloc(cw * nparams * ns * h->nmax)))
        return rc;
    hipStream_t s = h->stream;
    auto q_build = [&](const double* U, int n, double* Q) {
        return h->cplx ? launch_q_build_c(dU.p, U, h->ns, n, nparams, Q, s)
                       : launch_q_build(dU.p, U, h->ns, n, nparams, Q, s);
    };
    HIP_TRY(hipMemcpyAsync(dU.p, du_up, cw * nparams * ns * h->nup * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(q_build(h->cplx ? h->Ucup : h->Uup, h->nup, qu.p));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipMemcpyAsync(dU.p, du_dn, cw * nparams * ns * h->ndn * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(q_build(h->cplx ? h->Ucdn : h->Udn, h->ndn, qd.p));
    HIP_TRY(hipStreamSynchronize(s));
    h->Qup = qu.release();
    h->Qdn = qd.release();
    h->O = o.release();
    h->nparams = nparams;
    return KVMC_OK;
}

int kvmc_measure_grad(kvmc_handle* h, double* eloc, double* o) {
    g_err.clear();
    if (!h) return fail(KVMC_ERR_INVALID, "null handle");
    if (!h->ready) return fail(KVMC_ERR_NOT_READY, "configurations not set");
    if (h->nparams <= 0) return fail(KVMC_ERR_INVALID, "orbital derivatives not set");
    hipStream_t s = h->stream;
    SweepArgs a = sweep_args(h);
    HIP_TRY(do_flush(h, a, h->nw, s));
    HIP_TRY(do_eloc(h, eloc_args(h), s));
    HIP_TRY(do_grad(h, eloc_args(h), s));
    const size_t cw = h->cplx ? 2 : 1;
    std::vector<double> re, im;
    if (eloc && h->cplx) { re.resize(h->nw); im.resize(h->nw); }
    if (eloc) HIP_TRY(hipMemcpyAsync(h->cplx ? re.data() : eloc, h->eloc, sizeof(double) * h->nw, hipMemcpyDeviceToHost, s));
    if (eloc && h->cplx) HIP_TRY(hipMemcpyAsync(im.data(), h->eloc_im, sizeof(double) * h->nw, hipMemcpyDeviceToHost, s));
    if (o) HIP_TRY(hipMemcpyAsync(o, h->O, cw * sizeof(double) * h->nw * h->nparams, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (eloc && h->cplx)
        for (int w = 0; w < h->nw; ++w) { eloc[2 * w] = re[w]; eloc[2 * w + 1] = im[w]; }
    return KVMC_OK;
}

int kvmc_get_logdet(kvmc_handle* h, double* logdet) {
    g_err.clear();
    if (!h || !logdet) return fail(KVMC_ERR_INVALID, "null argument");
    std::vector<double> two((size_t)2 * h->nw);
    HIP_TRY(hipMemcpyAsync(two.data(), h->logdet, two.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    // complex orbitals: the real form's determinant is |det tildeU|^2
    const double f = h->cplx ? 0.5 : 1.0;
    for (int w = 0; w < h->nw; ++w) logdet[w] = f * (two[2 * w] + two[2 * w + 1]);
    return KVMC_OK;
}

int kvmc_get_w(kvmc_handle* h, int64_t walker, int32_t spin, double* out) {
    g_err.clear();
    if (!h || !out || walker < 0 || walker >= h->nw || spin < 0 || spin > 1)
        return fail(KVMC_ERR_INVALID, "bad walker/spin");
    hipStream_t s = h->stream;
    SweepArgs a = sweep_args(h);
    a.walker0 = (int)walker;
    HIP_TRY(do_flush(h, a, 1, s));
    // W keeps the empty-site rows; the occupied rows are the unit rows e_label
    const size_t ns = h->ns, n = spin ? h->ndn : h->nup, ld = w_ld(h->ns, (int)n);
    const size_t cw = h->cplx ? 2 : 1;   // complex: (re, im) pairs
    const size_t off = cw * ((size_t)walker * w_stride(h->ns, h->nup, h->ndn) + (spin ? w_off_dn(h->ns, h->nup) : 0));
    std::vector<double> rows(cw * ld * n);
    std::vector<int> sl(ns), kap(ns);
    if (h->prec == KVMC_PREC_FP64) {
        HIP_TRY(hipMemcpyAsync(rows.data(), (double*)h->W + off, rows.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    } else {
        std::vector<float> tmp(rows.size());
        HIP_TRY(hipMemcpyAsync(tmp.data(), (float*)h->W + off, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (size_t q = 0; q < rows.size(); ++q) rows[q] = tmp[q];
    }
    HIP_TRY(hipMemcpyAsync(sl.data(), h->slot + ((size_t)walker * 2 + spin) * ns, ns * sizeof(int),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(kap.data(), (spin ? h->kdn : h->kup) + (size_t)walker * ns, ns * sizeof(int),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (size_t l = 0; l < n; ++l)
        for (size_t r = 0; r < ns; ++r)
            for (size_t p = 0; p < cw; ++p)
                out[cw * (l * ns + r) + p] = (sl[r] >= 0 && (size_t)sl[r] < ld)
                                                 ? rows[cw * (l * ld + sl[r]) + p]
                                                 : (p == 0 && kap[r] == (int)l + 1 ? 1.0 : 0.0);
    return KVMC_OK;
}

int kvmc_set_sweep_count(kvmc_handle* h, int64_t sweeps) {
    g_err.clear();
    if (!h || sweeps < 0) return fail(KVMC_ERR_INVALID, "bad argument");
    hipLaunchKernelGGL(k_fill_ll, dim3((h->nw + 255) / 256), dim3(256), 0, h->stream, h->sweeps,
                       (long long)sweeps, h->nw);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    return KVMC_OK;
}

int kvmc_enable_graph(kvmc_handle* h, int32_t enable) {
    g_err.clear();
    if (!h) return fail(KVMC_ERR_INVALID, "null handle");
    h->graph = enable != 0;
    if (!h->graph && h->gexec) {
        HIP_TRY(hipGraphExecDestroy(h->gexec));
        h->gexec = nullptr;
        h->g_nsweeps = -1;
    }
    return KVMC_OK;
}

int kvmc_stagger_refresh(kvmc_handle* h, uint64_t seed) {
    g_err.clear();
    if (!h) return fail(KVMC_ERR_INVALID, "null handle");
    if (h->policy == KVMC_REFRESH_REFERENCE) return KVMC_OK;
    std::vector<int> since(h->nw);
    for (int w = 0; w < h->nw; ++w) {
        uint32_t x[4];
        philox4x32_10((uint32_t)w, 0u, 0u, 2u, (uint32_t)seed, (uint32_t)(seed >> 32), x);
        since[w] = (int)(((uint64_t)x[0] * (uint64_t)(uint32_t)h->refresh_k) >> 32);
    }
    HIP_TRY(hipMemcpyAsync(h->acc_since, since.data(), sizeof(int) * h->nw, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return KVMC_OK;
}

int kvmc_get_sweep_count(kvmc_handle* h, int64_t* sweeps) {
    g_err.clear();
    if (!h || !sweeps) return fail(KVMC_ERR_INVALID, "bad argument");
    HIP_TRY(hipMemcpyAsync(sweeps, h->sweeps, sizeof(long long) * h->nw, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return KVMC_OK;
}

int kvmc_profile_enable(kvmc_handle* h, int32_t enable) {
    g_err.clear();
    if (!h) return fail(KVMC_ERR_INVALID, "null handle");
    if (enable && !h->ev0) {
        HIP_TRY(hipEventCreate(&h->ev0));
        HIP_TRY(hipEventCreate(&h->ev1));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->profiling = enable != 0;
    h->prof = kvmc_profile{};
    return KVMC_OK;
}

int kvmc_get_profile(kvmc_handle* h, kvmc_profile* out) {
    g_err.clear();
    if (!h || !out) return fail(KVMC_ERR_INVALID, "null argument");
    *out = h->prof;
    return KVMC_OK;
}

int kvmc_w_drift(kvmc_handle* h, double* max_rel) {
    g_err.clear();
    if (!h || !max_rel) return fail(KVMC_ERR_INVALID, "null argument");
    if (!h->ready) return fail(KVMC_ERR_NOT_READY, "configurations not set");
    int rc = refresh_all(h, 0);
    if (rc) return rc;
    std::vector<double> dr((size_t)2 * h->nw);
    HIP_TRY(hipMemcpyAsync(dr.data(), h->drift, dr.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    double m = 0.0;
    for (int w = 0; w < h->nw; ++w)
        if (dr[2 * w + 1] > 0.0) m = std::max(m, dr[2 * w] / dr[2 * w + 1]);
    *max_rel = m;
    return KVMC_OK;
}

int kvmc_track_drift(kvmc_handle* h, int32_t enable) {
    g_err.clear();
    if (!h) return fail(KVMC_ERR_INVALID, "null handle");
    h->track_drift = enable != 0;
    if (h->gexec) {   // a captured sequence would skip the drift passes
        HIP_TRY(hipGraphExecDestroy(h->gexec));
        h->gexec = nullptr;
        h->g_nsweeps = -1;
    }
    HIP_TRY(hipMemsetAsync(h->dstats, 0, 3 * sizeof(double), h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return KVMC_OK;
}

int kvmc_get_drift_stats(kvmc_handle* h, double* max_rel, double* mean_rel, int64_t* count) {
    g_err.clear();
    if (!h) return fail(KVMC_ERR_INVALID, "null handle");
    double st[3] = {0, 0, 0};
    HIP_TRY(hipMemcpyAsync(st, h->dstats, sizeof(st), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (max_rel) *max_rel = st[0];
    if (mean_rel) *mean_rel = st[2] > 0 ? st[1] / st[2] : 0.0;
    if (count) *count = (int64_t)st[2];
    return KVMC_OK;
}

}  // extern "C"
