/*
 * kagome_vmc.h -- C ABI of the MI355X (gfx950) Gutzwiller-projected
 * determinant VMC sampler for the Kagome Heisenberg model.
 *
 * One handle holds a batch of independent walkers (Markov chains) on one GPU:
 * per walker the configurations kappa_up/kappa_down and the matrices
 * W_sigma = U_sigma * tildeU_sigma^-1.  The entry points replace the
 * reference's Carlo.jl plugin hooks on KagomeDSL.MC (hz-xiaxz/KagomeDSL.jl);
 * each declaration names the reference function it stands in for.
 *
 * Conventions (same as the reference's Julia data):
 *   - sites and labels are 1-based; kappa[s] = 0 marks an empty site;
 *   - U matrices are column-major ns x N (Julia Matrix layout), FP64;
 *   - per-walker arrays are walker-major: walker w's kappa is
 *     kappa[w*ns .. w*ns+ns-1];
 *   - bonds are 1-based (i, j) pairs with i <= j in get_nn order
 *     (src/Hamiltonian.jl:447-451).
 * Host pointers are borrowed for the duration of a call.  All work runs on
 * the handle's HIP stream; calls that return host data synchronise.
 * Every call returns KVMC_OK (0) or an error code; kvmc_last_error() gives
 * the message for the calling thread.
 */
#ifndef KAGOME_VMC_H
#define KAGOME_VMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KVMC_OK 0
#define KVMC_ERR_INVALID 1   /* bad argument (ArgumentError / DimensionMismatch) */
#define KVMC_ERR_SINGULAR 2  /* singular tildeU (SingularException) */
#define KVMC_ERR_HIP 3       /* HIP runtime failure */
#define KVMC_ERR_NOMEM 4     /* device allocation failed */
#define KVMC_ERR_NOT_READY 5 /* W not initialised (set configs first) */

/* W storage precision. FP64: ratios, rank-1 updates and W in double.
 * FP32: W, ratios and rank-1 (or delayed rank-k) updates in float; local
 * energies accumulate in double; refresh LU and solve in double. */
#define KVMC_PREC_FP64 0
#define KVMC_PREC_FP32 1

/* Orbital type.  REAL: u_up / u_down hold ns x N doubles (the pi-flux ansatz
 * at B = 0 is real, and so are U and W).  COMPLEX: they hold ns x N complex
 * doubles, interleaved (re, im) like Julia's Matrix{ComplexF64}: flux phases
 * or a magnetic field make H complex (src/Hamiltonian.jl:201, 325), U comes
 * from eigen(Hermitian) (:382-393) and W has eltype(U) (src/MonteCarlo.jl:
 * 181-182), updated by a complex geru! (:290).  W is then stored as complex
 * float (FP32) or complex double (FP64); the acceptance uses |ratio|^2. */
#define KVMC_ORB_REAL 0
#define KVMC_ORB_COMPLEX 1

/* Transition kernel.  REFERENCE: the reference's shared-uniform kernel
 * (src/MonteCarlo.jl:538-607), stationary law |psi|^2 / Z_mu.
 * UNBIASED: uniform bond, accept with min(1, |ratio|^2), law |psi|^2. */
#define KVMC_KERNEL_REFERENCE 0
#define KVMC_KERNEL_UNBIASED 1

/* Refresh policy. REFERENCE: after a sweep that passed both early returns
 * when sweeps % min(N_up, N_down) == 0 (src/MonteCarlo.jl:593-604).
 * EVERY_K: after every refresh_k accepted moves of a walker. */
#define KVMC_REFRESH_REFERENCE 0
#define KVMC_REFRESH_EVERY_K 1
/* AT_END (no reference counterpart): a walker that has made refresh_k
 * accepted moves since its last refresh finishes the sweeps of the current
 * kvmc_sweep call and is refreshed at its end (at most one call's accepts
 * late).  Saves the second sweep launch that EVERY_K needs for walkers
 * paused mid-call. */
#define KVMC_REFRESH_AT_END 2

/* Per-sweep outcome codes written by kvmc_sweep (optional record). */
#define KVMC_OUT_PRE_REJECT 0 /* r > Z_mu / Z_max (src/MonteCarlo.jl:547-550) */
#define KVMC_OUT_NO_MOVE 1    /* drawn bond not antiparallel (:564-567) */
#define KVMC_OUT_REJECT 2     /* Metropolis reject */
#define KVMC_OUT_ACCEPT 3     /* accepted move (:581-591) */

typedef struct kvmc_handle kvmc_handle;
typedef struct kvmc_group kvmc_group;

typedef struct kvmc_desc {
    int32_t ns;             /* sites */
    int32_t n_up, n_down;   /* particles per spin (columns of U) */
    int32_t nb;             /* bonds */
    const int32_t* bonds;   /* 2*nb, 1-based, get_nn order */
    const double* u_up;     /* ns x n_up column-major */
    const double* u_down;   /* ns x n_down column-major */
    int64_t nwalkers;
    int32_t precision;      /* KVMC_PREC_* */
    int32_t kernel;         /* KVMC_KERNEL_* */
    int32_t refresh_policy; /* KVMC_REFRESH_* */
    int32_t refresh_k;      /* accepted moves per refresh for EVERY_K */
    int32_t delay;          /* pending rank-1 updates kept before one rank-`delay`
                               flush of W (<= 1: immediate, as the reference);
                               FP32 and FP64 alike */
    int32_t threads;        /* threads per workgroup of the sweep kernel (0: auto) */
    uint64_t seed;          /* Philox key of the device RNG */
    int32_t orbitals;       /* KVMC_ORB_* */
} kvmc_desc;

typedef struct kvmc_counters {
    int64_t sweeps;      /* walker-sweeps executed */
    int64_t moves;       /* proposals that reached the ratio (passed both early returns) */
    int64_t accepted;    /* accepted moves */
    int64_t refreshes;   /* walker refreshes (both spins) */
    int64_t flushes;     /* delayed-update flushes */
    int64_t launches;    /* sweep-kernel launches */
} kvmc_counters;

/* MC(params) + the W/kappa allocation of the low-level constructor
 * (src/MonteCarlo.jl:165-185, 209-234).  Copies U and the bonds to the
 * device and allocates per-walker state for desc->nwalkers walkers.
 * KVMC_ERR_INVALID (with the reason) for sizes the refresh does not take:
 * more than 768 electrons per spin (384 with complex orbitals), or an LDS
 * footprint beyond 160 KiB. */
int kvmc_create(const kvmc_desc* desc, kvmc_handle** out);

/* Frees all device memory of the handle. */
int kvmc_destroy(kvmc_handle* h);

/* Sets kappa for all walkers (nwalkers*ns each), zeroes pending updates and
 * recomputes W by refresh (find_initial_configuration! / read_checkpoint!,
 * src/MonteCarlo.jl:382-410, 751-755; unlike the reference, W is rebuilt on
 * restore).  Returns KVMC_ERR_SINGULAR if any tildeU is singular; the first
 * failing walker is reported by kvmc_last_error and kvmc_failed_walker. */
int kvmc_set_configs(kvmc_handle* h, const int32_t* kappa_up, const int32_t* kappa_down);

/* Copies kappa of all walkers to the host (write_checkpoint, :715-719). */
int kvmc_get_configs(kvmc_handle* h, int32_t* kappa_up, int32_t* kappa_down);

/* Seeded uniform Sz=0 start for every walker (replaces the QR start of
 * Carlo.init!, src/MonteCarlo.jl:435-441): Philox Fisher-Yates permutation
 * keyed (seed, walker, attempt); walkers whose tildeU is singular are
 * redrawn up to max_attempts times.  Singular here means pivots of the FP64
 * LU spanning more than 1e10 (FP64 handles) or 1e4 (FP32 handles, whose
 * float W could not resolve the ratios).  *redraws receives the number of
 * redraws. */
int kvmc_init_random(kvmc_handle* h, uint64_t seed, int32_t max_attempts, int64_t* redraws);

/* reevaluateW! for every walker (src/MonteCarlo.jl:55-66): FP64 pivoted LU
 * of tildeU and W = U * tildeU^-1, converted to the storage precision. */
int kvmc_refresh(kvmc_handle* h);

/* Carlo.sweep! (src/MonteCarlo.jl:538-607), nsweeps times on every walker,
 * including the refreshes the policy triggers.  Draws: Philox keyed
 * (desc.seed, walker, sweep counter) when tape_u is NULL, otherwise the tape
 * (nwalkers*nsweeps entries each, walker-major: u in [0,1) and 0-based bond
 * index).  outcomes (optional, nwalkers*nsweeps int8) receives KVMC_OUT_*
 * per sweep.  counters (optional) receives this call's totals. */
int kvmc_sweep(kvmc_handle* h, int64_t nsweeps, const double* tape_u, const int32_t* tape_bond,
               int8_t* outcomes, kvmc_counters* counters);

/* No counterpart in the reference.  kvmc_sweep for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid allocation):
 * the same sweeps, draws, outcomes and counters, with the tape and the
 * outcome record staged in a fixed 32 MiB device array (one per GPU, held
 * until the process ends) instead of per-call buffers.  Limit: nwalkers *
 * nsweeps <= 2581071 when a tape or an outcome record is passed (larger
 * calls return KVMC_ERR_INVALID); graph replay is not used.  The handle's
 * device must be current.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (two handles or threads at once) race. */
int kvmc_sweep_noalloc(kvmc_handle* h, int64_t nsweeps, const double* tape_u,
                       const int32_t* tape_bond, int8_t* outcomes, kvmc_counters* counters);

/* No counterpart in the reference, whose runs are one lattice each (Carlo
 * farms separate runs).  A group of handles on one GPU -- e.g. lattices of
 * different sizes, one handle each -- whose sweeps run as one graph launch:
 * kvmc_group_sweep captures every handle's sweep sequence (as replayed by
 * kvmc_enable_graph) as a parallel branch of one hipGraph, on the handles'
 * own streams, so each size keeps its own kernel shapes and LDS footprint,
 * then replays it with one launch per call.  Walkers still paused after the
 * captured rounds are finished by each handle's ordinary launch loop.
 * Results equal kvmc_sweep on each handle.  A handle that is profiling or
 * tracking drift makes the call sweep the handles one after another.
 * The group borrows the handles: destroy it before them. */
int kvmc_group_create(kvmc_handle* const* handles, int32_t n, kvmc_group** out);
/* nsweeps: one count per handle (0 skips it); counters (optional): one per handle. */
int kvmc_group_sweep(kvmc_group* g, const int64_t* nsweeps, kvmc_counters* counters);
int kvmc_group_destroy(kvmc_group* g);

/* getOL for every walker (src/Hamiltonian.jl:762-778): Heisenberg J1 local
 * energy real(<x|H|psi>/<x|psi>), accumulated in FP64.  eloc: nwalkers
 * doubles (may be NULL); sum (optional) receives sum over walkers and
 * n_antiparallel (optional) the total antiparallel-bond count. */
int kvmc_measure(kvmc_handle* h, double* eloc, double* sum, int64_t* n_antiparallel);

/* Sets ns x N column-major derivatives dU_k/dtheta_k of the occupied
 * orbitals for nparams variational parameters (nparams*ns*N values per
 * spin, parameter-major; complex handles: interleaved complex).  Enables
 * kvmc_measure_grad.  No reference counterpart: the reference has no
 * differentiation. */
int kvmc_set_orbital_derivs(kvmc_handle* h, int32_t nparams, const double* du_up,
                            const double* du_down);

/* Local energy and log-derivatives O_k(x) = d log det tildeU_up/dtheta_k +
 * d log det tildeU_dn/dtheta_k for every walker.  Real handles: eloc holds
 * nwalkers doubles, o nwalkers*nparams doubles (walker-major).  Complex
 * handles: both are complex, interleaved (re, im): eloc 2*nwalkers doubles
 * (the local energy <x|H|psi>/<x|psi>, whose real part is getOL's), o
 * 2*nwalkers*nparams doubles. */
int kvmc_measure_grad(kvmc_handle* h, double* eloc, double* o);

/* log|det tildeU_up| + log|det tildeU_dn| per walker from the last refresh
 * (complex handles too: the modulus). */
int kvmc_get_logdet(kvmc_handle* h, double* logdet);

/* Copies W_sigma (spin 0 = up, 1 = down) of one walker as FP64, column-major
 * ns x N (complex handles: ns x N complex, interleaved), with pending delayed
 * updates applied.  The device keeps only the rows of the empty sites; the
 * rows of occupied sites are returned as the unit rows e_label they are by
 * construction. */
int kvmc_get_w(kvmc_handle* h, int64_t walker, int32_t spin, double* out);

/* Sets / reads ctx.sweeps of every walker (Carlo's sweep counter). */
int kvmc_set_sweep_count(kvmc_handle* h, int64_t sweeps);
int kvmc_get_sweep_count(kvmc_handle* h, int64_t* sweeps);

/* Enables (1) or disables (0) graph replay for kvmc_sweep calls without a
 * tape, outcome record or profiling.  The first such call for a given
 * nsweeps captures on the handle's stream: stop counts, counter resets (by
 * fill kernels), then 3 rounds of the sweep kernel and the refresh of the
 * walkers it paused, their count read on the device.  Later calls replay
 * it, and walkers still paused after the last round are finished by the
 * ordinary launch loop.  Results are identical to the launch loop.  No
 * reference counterpart: it removes per-launch host round trips for small
 * clusters. */
int kvmc_enable_graph(kvmc_handle* h, int32_t enable);

/* EVERY_K and AT_END policies (no effect otherwise): sets each walker's count of
 * accepted moves since its last refresh to a seeded uniform draw in
 * [0, refresh_k), so that refreshes are spread over time as in a long run
 * instead of arriving together after a common start.  Decisions of the
 * chains are unchanged; only when each W is rebuilt moves.  No reference
 * counterpart. */
int kvmc_stagger_refresh(kvmc_handle* h, uint64_t seed);

/* Largest max|W - W_refreshed| / max|W_refreshed| over walkers (each
 * walker's maxima over both spins), measured by recomputing W in FP64
 * without replacing it (drift of the rank-1 updates).  Stands beside
 * reevaluateW! (src/MonteCarlo.jl:55-66), which replaces W without
 * measuring what it replaced. */
int kvmc_w_drift(kvmc_handle* h, double* max_rel_drift);

/* Enables (1) or disables (0) a drift pass before every refresh that the
 * refresh policy triggers inside kvmc_sweep (src/MonteCarlo.jl:593-604):
 * each refreshed walker's max|W - W_fresh| / max|W_fresh| is recorded just
 * before W is replaced.  Costs one extra refresh computation per refreshed
 * walker; graph replay is bypassed while enabled.  Resets the statistics. */
int kvmc_track_drift(kvmc_handle* h, int32_t enable);

/* Statistics of the drift passes since kvmc_track_drift: the largest and
 * the mean relative drift over refreshed walkers, and their number.  Any
 * output may be NULL. */
int kvmc_get_drift_stats(kvmc_handle* h, double* max_rel, double* mean_rel, int64_t* count);

/* Kernel timing measured with HIP events on the handle's stream (no
 * reference counterpart; used by benchmarks).  Enabling resets the sums. */
typedef struct kvmc_profile {
    double sweep_ms;         /* sweep-kernel time */
    double refresh_ms;       /* refresh-kernel time */
    double measure_ms;       /* local-energy kernel time */
    int64_t sweep_launches;
    int64_t refresh_launches;
    int64_t measure_launches;
} kvmc_profile;
int kvmc_profile_enable(kvmc_handle* h, int32_t enable);
int kvmc_get_profile(kvmc_handle* h, kvmc_profile* out);

/* Walker index of the last singular refresh (-1 if none). */
int64_t kvmc_failed_walker(kvmc_handle* h);

/* Message of the last error on the calling thread ("" if none). */
const char* kvmc_last_error(void);

/* Library version string. */
const char* kvmc_version(void);

#ifdef __cplusplus
}
#endif

#endif /* KAGOME_VMC_H */
