// Device kernels of the Kagome VMC sampler, written for gfx950 (CDNA4),
// generic over the storage type T of W: float or double (real orbitals),
// cf32 or cf64 (complex orbitals).  kvmc_device.hip instantiates the real
// types, kvmc_complex.hip the complex ones.
//
//   k_sweep    Metropolis spin-exchange chain per walker (one workgroup per
//              walker, many sweeps per launch) fused with the Sherman-Morrison
//              W update.  The update is kept as `delay` pending rank-1 terms
//              (columns A_q, scaled rows B_q in LDS) and applied to W in HBM
//              by one streaming rank-`delay` pass; delay = 1 is the
//              reference's immediate geru (src/MonteCarlo.jl:279-292).
//   (the FP64 W refresh lives in kvmc_refresh.hip)
//   k_eloc     Heisenberg J1 local energy (src/Hamiltonian.jl:762-778), one
//              wavefront per walker, FP64 accumulation.
//   k_eloc_wg  the same, one workgroup per walker, FP32 W with pending terms.
//   k_grad     log-derivatives O_k = d log det tildeU / d theta_k.
//   k_flush    applies pending delayed updates to W.
//
// Layout in HBM (per walker, walker-major): kappa_up[ns], kappa_dn[ns]
// (int32, 1-based labels, 0 = empty); W_up (ldw_up x nup) then W_dn (ldw_dn x
// ndn), the rows of the empty sites only (the occupied rows are unit rows),
// column-major, so column l (a label) is contiguous; slot[2][ns] maps a site
// to its row.  For small clusters (both W of a walker within
// kResidentMaxBytes) the sweep kernel copies W into LDS at launch and back at
// exit, so sweeps and flushes never touch HBM.
#pragma once

#include "kvmc_internal.hpp"

#include <float.h>
#include <math.h>

namespace kvmc {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// 16-byte vector of V elements of T (the streaming pass's unit).
template <typename T, int V> struct VecT;
template <> struct VecT<float, 4> { using type = f32x4; };
template <> struct VecT<double, 2> { using type = f64x2; };
template <> struct VecT<cf32, 2> { using type = f32x4; };
template <> struct VecT<cf64, 1> { using type = f64x2; };

// w += a * b for a 16-byte vector a of W elements and one element b.
__device__ __forceinline__ void vmac(f32x4& w, const f32x4& a, float b) { w += a * b; }
__device__ __forceinline__ void vmac(f64x2& w, const f64x2& a, double b) { w += a * b; }
__device__ __forceinline__ void vmac(f32x4& w, const f32x4& a, cf32 b) {
    w.x += a.x * b.re - a.y * b.im;
    w.y += a.y * b.re + a.x * b.im;
    w.z += a.z * b.re - a.w * b.im;
    w.w += a.w * b.re + a.z * b.im;
}
__device__ __forceinline__ void vmac(f64x2& w, const f64x2& a, cf64 b) {
    w.x += a.x * b.re - a.y * b.im;
    w.y += a.y * b.re + a.x * b.im;
}

// |x|^2 in double.
__device__ __forceinline__ double abs2d(float x) { return (double)x * (double)x; }
__device__ __forceinline__ double abs2d(double x) { return x * x; }
__device__ __forceinline__ double abs2d(cf32 x) { return (double)x.re * (double)x.re + (double)x.im * (double)x.im; }
__device__ __forceinline__ double abs2d(cf64 x) { return x.re * x.re + x.im * x.im; }

// Value of lane l (uniform l) as a scalar, for a W element of any storage type.
__device__ __forceinline__ float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((unsigned long long)b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float lane_t(float v, int l) { return readlane_f(v, l); }
__device__ __forceinline__ double lane_t(double v, int l) { return readlane_d(v, l); }
__device__ __forceinline__ cf32 lane_t(cf32 v, int l) { return cf32(readlane_f(v.re, l), readlane_f(v.im, l)); }
__device__ __forceinline__ cf64 lane_t(cf64 v, int l) { return cf64(readlane_d(v.re, l), readlane_d(v.im, l)); }

// FP64 accumulator of a W element: double, or cf64 for complex W.
template <typename T> struct AccT { using type = double; };
template <typename R> struct AccT<cplx<R>> { using type = cf64; };
__device__ __forceinline__ double to_acc(float x) { return (double)x; }
__device__ __forceinline__ double to_acc(double x) { return x; }
__device__ __forceinline__ cf64 to_acc(cf32 x) { return cf64((double)x.re, (double)x.im); }
__device__ __forceinline__ cf64 to_acc(cf64 x) { return x; }
// Real and imaginary part of an accumulator.
__device__ __forceinline__ double re_of(double x) { return x; }
__device__ __forceinline__ double re_of(cf64 x) { return x.re; }
__device__ __forceinline__ double im_of(double) { return 0.0; }
__device__ __forceinline__ double im_of(cf64 x) { return x.im; }

__device__ inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// ---------------------------------------------------------------------------
// Streaming rank-c update of one spin's W (ld rows, column stride ld):
// W[:, j] += sum_{q<c} A_q * B_q[j].  Thread t owns the row vector
// iv = t % nvec of every column j == lane (mod lanes); its A_q slices stay in
// registers for the whole pass, B_q[j] is an LDS broadcast.  Requires
// blockDim.x >= ld / V (checked on the host).
// ---------------------------------------------------------------------------
template <typename T, int V, int DMAX, int UNR = 4>
__device__ __forceinline__ void stream_update(T* __restrict__ W0, const T* __restrict__ sA,
                                              const T* __restrict__ sB, int ld, int nsa,
                                              int nmax, int ncols, int c) {
    using VT = typename VecT<T, V>::type;   // UNR: columns in flight per thread
    const int nvec = ld / V;
    if (nvec == 0) return;
    const int lanes = (int)blockDim.x / nvec;
    const int t = (int)threadIdx.x;
    const int iv = t % nvec, lane = t / nvec;
    if (lane >= lanes) return;
    VT a[DMAX];
#pragma unroll
    for (int q = 0; q < DMAX; ++q)
        a[q] = (q < c) ? *reinterpret_cast<const VT*>(sA + (size_t)q * nsa + (size_t)iv * V) : VT(0);
    T* __restrict__ base = W0 + (size_t)iv * V;
    int j = lane;
    for (; j + (UNR - 1) * lanes < ncols; j += UNR * lanes) {
        VT w[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) w[u] = *reinterpret_cast<const VT*>(base + (size_t)(j + u * lanes) * ld);
#pragma unroll
        for (int q = 0; q < DMAX; ++q) {
            if (q < c) {
#pragma unroll
                for (int u = 0; u < UNR; ++u) vmac(w[u], a[q], sB[q * nmax + j + u * lanes]);
            }
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) *reinterpret_cast<VT*>(base + (size_t)(j + u * lanes) * ld) = w[u];
    }
#pragma nounroll
    for (; j < ncols; j += lanes) {
        VT w0 = *reinterpret_cast<const VT*>(base + (size_t)j * ld);
#pragma unroll
        for (int q = 0; q < DMAX; ++q)
            if (q < c) vmac(w0, a[q], sB[q * nmax + j]);
        *reinterpret_cast<VT*>(base + (size_t)j * ld) = w0;
    }
}

// Current W element of row e (the slot of an empty site) and label l:
// W0[e, l] + sum_{q<c} A_q[e] B_q[l].
template <typename T>
__device__ __forceinline__ T w_cur(const T* __restrict__ W0, const T* sA, const T* sB, int ld,
                                   int nsa, int nmax, int K, int l, int c) {
    T v = W0[(size_t)l * ld + K];
    for (int q = 0; q < c; ++q) v += sA[q * nsa + K] * sB[q * nmax + l];
    return v;
}

// Appends both spins' rank-1 terms of an accepted move (target rows K, i.e.
// the slots of the sites the electrons move to, labels l, 0-based):
// B_c = alpha (W[K, :] - e_l), alpha = -1/W[K, l], and A_c = W[:, l] + e_K,
// the reference's update_W! factors (src/MonteCarlo.jl:279-292) on the
// empty-site rows.  The + e_K turns row K, which now stands for the site the
// electron left (unit row e_l before the move), into that site's new row
// e_l + B_c.  The column and row elements of both W are loaded before any
// correction or LDS store, so the accept costs one memory round trip.
template <typename T>
__device__ __forceinline__ void append_terms(const T* __restrict__ Wup, const T* __restrict__ Wdn,
                                             T* sAu, T* sBu, T* sAd, T* sBd, int ldu, int ldd, int nsa,
                                             int nmax, int nup, int ndn, int Kup, int lup, int Kdn,
                                             int ldn, int c, T wu, T wd) {
    constexpr int CH = 4;
    const int bd = (int)blockDim.x, t0 = (int)threadIdx.x;
    const int nld = ldu > ldd ? ldu : ldd;
    const int nmx = nld > nmax ? nld : nmax;
    for (int base = 0; base < nmx; base += CH * bd) {
        T cu[CH], cd[CH], ru[CH], rd[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const int t = base + t0 + u * bd;
            cu[u] = t < ldu ? Wup[(size_t)lup * ldu + t] : T(0);
            cd[u] = t < ldd ? Wdn[(size_t)ldn * ldd + t] : T(0);
            ru[u] = t < nup ? Wup[(size_t)t * ldu + Kup] : T(0);
            rd[u] = t < ndn ? Wdn[(size_t)t * ldd + Kdn] : T(0);
        }
        const T au = T(-1) / wu, ad = T(-1) / wd;
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const int t = base + t0 + u * bd;
            if (t < ldu) {
                T vu = cu[u];
                for (int q = 0; q < c; ++q) vu += sAu[q * nsa + t] * sBu[q * nmax + lup];
                sAu[c * nsa + t] = vu + (t == Kup ? T(1) : T(0));
            }
            if (t < ldd) {
                T vd = cd[u];
                for (int q = 0; q < c; ++q) vd += sAd[q * nsa + t] * sBd[q * nmax + ldn];
                sAd[c * nsa + t] = vd + (t == Kdn ? T(1) : T(0));
            }
            if (t < nup) {
                T v = ru[u];
                for (int q = 0; q < c; ++q) v += sAu[q * nsa + Kup] * sBu[q * nmax + t];
                if (t == lup) v -= T(1);
                sBu[c * nmax + t] = au * v;
            }
            if (t < ndn) {
                T v = rd[u];
                for (int q = 0; q < c; ++q) v += sAd[q * nsa + Kdn] * sBd[q * nmax + t];
                if (t == ldn) v -= T(1);
                sBd[c * nmax + t] = ad * v;
            }
        }
    }
}

// Pending terms between HBM ([spin][row][dp], one row's terms contiguous)
// and LDS ([spin][q][row], one term's column contiguous); only the c live
// terms move.  One thread per row of A or label of B.
template <typename T>
__device__ __forceinline__ void pend_load(T* s_A, T* s_B, const T* gA, const T* gB, int nsa,
                                          int nmax, int delay, int dp, int c) {
    for (int t = (int)threadIdx.x; t < 2 * nsa; t += (int)blockDim.x) {
        const int sp = t >= nsa ? 1 : 0;
        T* d = s_A + sp * delay * nsa + (t - sp * nsa);
        const T* g = gA + (size_t)t * dp;
        for (int q = 0; q < c; ++q) d[q * nsa] = g[q];
    }
    for (int t = (int)threadIdx.x; t < 2 * nmax; t += (int)blockDim.x) {
        const int sp = t >= nmax ? 1 : 0;
        T* d = s_B + sp * delay * nmax + (t - sp * nmax);
        const T* g = gB + (size_t)t * dp;
        for (int q = 0; q < c; ++q) d[q * nmax] = g[q];
    }
}

template <typename T>
__device__ __forceinline__ void pend_store(const T* s_A, const T* s_B, T* gA, T* gB, int nsa,
                                           int nmax, int delay, int dp, int c) {
    for (int t = (int)threadIdx.x; t < 2 * nsa; t += (int)blockDim.x) {
        const int sp = t >= nsa ? 1 : 0;
        const T* d = s_A + sp * delay * nsa + (t - sp * nsa);
        T* g = gA + (size_t)t * dp;
        for (int q = 0; q < c; ++q) g[q] = d[q * nsa];
    }
    for (int t = (int)threadIdx.x; t < 2 * nmax; t += (int)blockDim.x) {
        const int sp = t >= nmax ? 1 : 0;
        const T* d = s_B + sp * delay * nmax + (t - sp * nmax);
        T* g = gB + (size_t)t * dp;
        for (int q = 0; q < c; ++q) g[q] = d[q * nmax];
    }
}

// Change of Z (src/MonteCarlo.jl:460-474) when the electrons on the
// antiparallel pair (i, j) are exchanged: at half filling every site holds
// one electron, so each bond touching exactly one of i and j flips its
// antiparallel status, while bonds between i and j and on-site bonds keep
// theirs.  dZ = (flipping bonds) - 2 (those antiparallel before).
// Bond tables are LDS-resident as 16-bit site and bond ids: bond b joins
// sites bonds[b] & 0xffff and bonds[b] >> 16; nbr[nbr_off[x] .. nbr_off[x+1])
// lists the bonds touching site x.
__device__ __forceinline__ int z_delta(const uint32_t* bonds, const uint16_t* nbr_off,
                                       const uint16_t* nbr, const int16_t* s_kup, const int16_t* s_kdn,
                                       int i, int j) {
    int flip = 0, ap = 0;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int x = side ? j : i, other = side ? i : j;
        for (int p = nbr_off[x]; p < nbr_off[x + 1]; ++p) {
            const uint32_t bp = bonds[nbr[p]];
            const int bx = (int)(bp & 0xffffu), by = (int)(bp >> 16);
            const int y = bx == x ? by : bx;
            if (bx == by || y == other) continue;
            ++flip;
            ap += (s_kup[bx] != 0 && s_kdn[by] != 0) || (s_kup[by] != 0 && s_kdn[bx] != 0);
        }
    }
    return flip - 2 * ap;
}

__device__ inline double block_sum_double(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int wave = (int)threadIdx.x / 64, lane = (int)threadIdx.x % 64;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = ((int)blockDim.x + 63) / 64;
    for (int q = 0; q < nw; ++q) s += red[q];
    return s;
}

// ---------------------------------------------------------------------------
// Sweep kernel: Carlo.sweep! (src/MonteCarlo.jl:538-607) repeated until the
// walker reaches its stop count or a refresh is due.  Every thread walks the
// same (uniform) control flow; proposals are evaluated redundantly so no
// barrier is needed on rejected sweeps.
// ---------------------------------------------------------------------------
__host__ __device__ inline size_t sweep_lds_base(size_t es, int ns, int nsa, int nmax, int delay, int nb) {
    const size_t nsp = (size_t)(ns + 3) / 4 * 4;
    const size_t b = 4 * nsp * sizeof(int16_t) + 2 * (size_t)delay * ((size_t)nsa + nmax) * es + 16 * sizeof(double) +
                     (size_t)nb * sizeof(uint32_t) + (size_t)(ns + 1 + 2 * nb) * sizeof(uint16_t);
    return (b + 15) / 16 * 16;
}

template <typename T, int V, int DMAX, bool RES>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(3))) k_sweep(SweepArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int w = a.walker0 + (int)blockIdx.x;
    const int ns = a.ns, nup = a.nup, ndn = a.ndn, nb = a.nb, nmax = a.nmax;
    const int nsp = round_up(ns, 4);
    const int nsa = w_nsa(ns, nup, ndn);                 // stride of the pending columns
    const int delay = a.delay;
    // labels and rows fit 16 bits (ns <= 65535 is checked on the host)
    int16_t* s_kup = reinterpret_cast<int16_t*>(lds);
    int16_t* s_kdn = s_kup + nsp;
    int16_t* s_slu = s_kdn + nsp;                         // site -> row of W_up (-1: occupied)
    int16_t* s_sld = s_slu + nsp;
    T* s_A = reinterpret_cast<T*>(s_sld + nsp);           // [2][delay][nsa] (8 nsp bytes: 16-byte aligned)
    T* s_B = s_A + 2 * delay * nsa;                       // [2][delay][nmax]
    double* s_red = reinterpret_cast<double*>(s_B + 2 * delay * nmax);
    uint32_t* s_bond = reinterpret_cast<uint32_t*>(s_red + 16);   // [nb] packed bond table
    uint16_t* s_noff = reinterpret_cast<uint16_t*>(s_bond + nb);    // [ns + 1]
    uint16_t* s_nbr = s_noff + ns + 1;                              // [<= 2 nb]

    const size_t wstride = w_stride(ns, nup, ndn);
    const int ldu = w_ld(ns, nup), ldd = w_ld(ns, ndn);
    T* const Wg = reinterpret_cast<T*>(a.W) + (size_t)w * wstride;
    // W of this walker: in HBM, or (RES) an LDS copy for the whole launch
    T* Wup = RES ? reinterpret_cast<T*>(lds + sweep_lds_base(sizeof(T), ns, nsa, nmax, delay, nb)) : Wg;
    T* Wdn = Wup + w_off_dn(ns, nup);
    int* const slot_g = a.slot + (size_t)w * 2 * ns;
    const int dp = pend_ld(delay);
    T* gA = reinterpret_cast<T*>(a.A) + (size_t)w * 2 * nsa * dp;
    T* gB = reinterpret_cast<T*>(a.Bm) + (size_t)w * 2 * nmax * dp;
    T* sAu = s_A;
    T* sAd = s_A + delay * nsa;
    T* sBu = s_B;
    T* sBd = s_B + delay * nmax;

    long long sweeps = a.sweeps[w];
    const long long stop = a.stop[w];
    if (sweeps >= stop) return;
    int c = a.npend[w];
    int acc_since = a.acc_since[w];

    for (int t = (int)threadIdx.x; t < ns; t += (int)blockDim.x) {
        s_kup[t] = (int16_t)a.kup[(size_t)w * ns + t];
        s_kdn[t] = (int16_t)a.kdn[(size_t)w * ns + t];
        s_slu[t] = (int16_t)slot_g[t];
        s_sld[t] = (int16_t)slot_g[ns + t];
    }
    if (c > 0) pend_load(s_A, s_B, gA, gB, nsa, nmax, delay, dp, c);
    if (RES) {
        for (size_t t = threadIdx.x; t < wstride; t += blockDim.x) Wup[t] = Wg[t];
    }
    for (int b = (int)threadIdx.x; b < nb; b += (int)blockDim.x) {
        const int2 bd = a.bonds[b];
        s_bond[b] = (uint32_t)bd.x | ((uint32_t)bd.y << 16);
    }
    for (int t = (int)threadIdx.x; t <= ns; t += (int)blockDim.x) s_noff[t] = (uint16_t)a.nbr_off[t];
    for (int t = (int)threadIdx.x; t < a.nbr_off[ns]; t += (int)blockDim.x) s_nbr[t] = (uint16_t)a.nbr[t];
    __syncthreads();
    // Z_mu: full recount once per launch, then maintained incrementally.
    int zmu;
    {
        double part = 0.0;
        for (int b = (int)threadIdx.x; b < nb; b += (int)blockDim.x) {
            int2 bd = a.bonds[b];
            part += (s_kup[bd.x] != 0 && s_kdn[bd.y] != 0) ||
                    (s_kup[bd.y] != 0 && s_kdn[bd.x] != 0);
        }
        zmu = (int)block_sum_double(part, s_red);
    }
    const double zmax = (double)nb;
    const long long tbase = a.tape_base[w];
    unsigned long long n_moves = 0, n_acc = 0, n_flush = 0, n_sweeps = 0;
    bool pause = false;
    const int lane = (int)threadIdx.x & 63;

    // Between two accepted moves the walker's state (kappa, Z_mu, W and the
    // pending terms) does not change, so every proposal up to and including
    // the next accept is a function of that state and its own draws.  Each
    // wavefront therefore evaluates a window of 64 consecutive sweeps, one
    // per lane (all W gathers of the window in flight together), and the
    // walker advances to the first sweep that accepts or triggers a refresh.
    // Every wavefront evaluates the same window, so the result is uniform
    // without a barrier.  Decisions equal those of one-at-a-time evaluation.
    while (sweeps < stop && !pause) {
        const long long sw = sweeps + lane;
        const bool valid = sw < stop;
        double r = 1.0;
        int bidx = 0;
        if (valid) {
            if (a.tape_u) {
                const long long tidx = sw - tbase;
                r = a.tape_u[(size_t)w * a.tape_len + tidx];
                bidx = a.tape_bond[(size_t)w * a.tape_len + tidx];
            } else {
                sweep_draws(a.seed, (uint32_t)w, sw, nb, &r, &bidx);
            }
        }
        const double zr = (double)zmu / zmax;
        int oc = OUT_PRE_REJECT;
        bool mv = false, acc = false;
        int i = 0, site = 0, lup = 0, ldn = 0, Kup = 0, Kdn = 0, dz = 0;   // K: target sites
        bool can1 = false;
        T wu = T(0), wd = T(0);
        if (valid && !(a.kernel_mode == KERNEL_REFERENCE && r > zr)) {
            const uint32_t bp = s_bond[bidx];
            i = (int)(bp & 0xffffu);
            site = (int)(bp >> 16);
            const int iu = s_kup[i], su = s_kup[site], id = s_kdn[i], sd = s_kdn[site];
            can1 = iu != 0 && sd != 0;
            const bool can2 = su != 0 && id != 0;
            if (!can1 && !can2) {
                oc = OUT_NO_MOVE;
            } else {
                mv = true;
                // flag = 1: up moves i -> site, down moves site -> i
                lup = (can1 ? iu : su) - 1;
                ldn = (can1 ? sd : id) - 1;
                Kup = can1 ? site : i;
                Kdn = can1 ? i : site;
                // targets are empty for their spin; max() keeps a walker with an
                // inconsistent state (failed refresh) inside W
                wu = w_cur(Wup, sAu, sBu, ldu, nsa, nmax, max(s_slu[Kup], 0), lup, c);
                wd = w_cur(Wdn, sAd, sBd, ldd, nsa, nmax, max(s_sld[Kdn], 0), ldn, c);
                dz = z_delta(s_bond, s_noff, s_nbr, s_kup, s_kdn, i, site);
                const double acc_prob = abs2d(wu * wd);
                if (a.kernel_mode == KERNEL_REFERENCE)
                    acc = (acc_prob >= 1.0 && r < zr) || (acc_prob < 1.0 && r < zr * acc_prob);
                else
                    acc = r < acc_prob;
                oc = acc ? OUT_ACCEPT : OUT_REJECT;
            }
        }
        // terminal sweep: an accept, or a move on which the reference cadence refreshes
        const bool term = acc || (mv && a.policy == POLICY_REFERENCE && (sw % a.nocc) == 0);
        const unsigned long long tmask = __ballot(valid && term);
        const int nvalid = __popcll(__ballot(valid));
        const int ncommit = tmask ? __ffsll((long long)tmask) : nvalid;   // sweeps consumed
        if (a.outcomes && threadIdx.x < 64 && lane < ncommit)
            a.outcomes[(size_t)w * a.tape_len + (sw - tbase)] = (signed char)oc;
        n_moves += __popcll(__ballot(mv && lane < ncommit));
        sweeps += ncommit;
        n_sweeps += ncommit;
        if (!tmask) continue;
        const int tl = ncommit - 1;   // lane of the terminal sweep
        // readlane (not a shuffle) keeps the terminal sweep's values scalar,
        // so the term count and everything branching on it stay uniform
        const bool t_acc = __builtin_amdgcn_readlane((int)acc, tl) != 0;
        const long long t_sw = sweeps - 1;
        bool refresh_due;
        if (t_acc) {
            const int ti = __builtin_amdgcn_readlane(i, tl), tsite = __builtin_amdgcn_readlane(site, tl);
            const bool tcan1 = __builtin_amdgcn_readlane((int)can1, tl) != 0;
            const int tlup = __builtin_amdgcn_readlane(lup, tl), tldn = __builtin_amdgcn_readlane(ldn, tl);
            const int tKup = __builtin_amdgcn_readlane(Kup, tl), tKdn = __builtin_amdgcn_readlane(Kdn, tl);
            const T twu = lane_t(wu, tl), twd = lane_t(wd, tl);
            const int tdz = __builtin_amdgcn_readlane(dz, tl);
            ++n_acc;
            ++acc_since;
            const int eu = max(s_slu[tKup], 0), ed = max(s_sld[tKdn], 0);   // rows of the target sites
            append_terms(Wup, Wdn, sAu, sBu, sAd, sBd, ldu, ldd, nsa, nmax, nup, ndn, eu, tlup, ed,
                         tldn, c, twu, twd);
            __syncthreads();
            if (threadIdx.x == 0) {
                // each electron's row goes to the site it left
                const int fu = tcan1 ? ti : tsite, fd = tcan1 ? tsite : ti;   // source sites
                s_slu[fu] = (int16_t)eu; s_slu[tKup] = -1;
                s_sld[fd] = (int16_t)ed; s_sld[tKdn] = -1;
                if (tcan1) {
                    s_kup[ti] = 0; s_kup[tsite] = (int16_t)(tlup + 1);
                    s_kdn[ti] = (int16_t)(tldn + 1); s_kdn[tsite] = 0;
                } else {
                    s_kup[ti] = (int16_t)(tlup + 1); s_kup[tsite] = 0;
                    s_kdn[ti] = 0; s_kdn[tsite] = (int16_t)(tldn + 1);
                }
            }
            ++c;
            if (c == delay) {
                stream_update<T, V, DMAX>(Wup, sAu, sBu, ldu, nsa, nmax, nup, c);
                stream_update<T, V, DMAX>(Wdn, sAd, sBd, ldd, nsa, nmax, ndn, c);
                c = 0;
                ++n_flush;
            }
            __syncthreads();
            zmu += tdz;
            refresh_due = a.policy == POLICY_REFERENCE ? (t_sw % a.nocc) == 0
                        : a.policy == POLICY_EVERY_K ? acc_since >= a.refresh_k
                                                     : false;   // AT_END: listed at exit
        } else {
            refresh_due = true;   // rejected move on a reference-cadence refresh sweep
        }
        if (refresh_due) pause = true;
    }
    __syncthreads();
    for (int t = (int)threadIdx.x; t < ns; t += (int)blockDim.x) {
        a.kup[(size_t)w * ns + t] = s_kup[t];
        a.kdn[(size_t)w * ns + t] = s_kdn[t];
        slot_g[t] = s_slu[t];
        slot_g[ns + t] = s_sld[t];
    }
    if (RES) {
        for (size_t t = threadIdx.x; t < wstride; t += blockDim.x) Wg[t] = Wup[t];
    }
    if (c > 0) pend_store(s_A, s_B, gA, gB, nsa, nmax, delay, dp, c);
    if (threadIdx.x == 0) {
        a.sweeps[w] = sweeps;
        a.npend[w] = c;
        a.acc_since[w] = acc_since;
        a.zmu[w] = zmu;
        // AT_END: a walker that reached refresh_k accepts during this call
        // finishes its sweeps and is refreshed after the launch
        if (pause || (a.policy == POLICY_AT_END && acc_since >= a.refresh_k)) {
            int slot = atomicAdd(a.refresh_count, 1);
            a.refresh_list[slot] = w;
            atomicAdd(&a.counters[CNT_PAUSES], 1ull);
        }
        atomicAdd(&a.counters[CNT_MOVES], n_moves);
        atomicAdd(&a.counters[CNT_ACCEPTED], n_acc);
        atomicAdd(&a.counters[CNT_FLUSHES], n_flush);
        atomicAdd(&a.counters[CNT_SWEEPS], n_sweeps);
    }
}

template <typename T, int V, int D>
hipError_t launch_sweep_d(const SweepArgs& a, int nw, int threads, size_t lds, hipStream_t s) {
    const bool res = sweep_resident(sizeof(T), a.ns, a.nup, a.ndn);
    const void* fn = res ? reinterpret_cast<const void*>(&k_sweep<T, V, D, true>)
                         : reinterpret_cast<const void*>(&k_sweep<T, V, D, false>);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (res) hipLaunchKernelGGL((k_sweep<T, V, D, true>), dim3(nw), dim3(threads), lds, s, a);
    else hipLaunchKernelGGL((k_sweep<T, V, D, false>), dim3(nw), dim3(threads), lds, s, a);
    return hipGetLastError();
}

template <typename T, int V>
hipError_t launch_sweep_v(const SweepArgs& a, int nw, int threads, size_t lds, hipStream_t s) {
    if (a.delay <= 1) return launch_sweep_d<T, V, 1>(a, nw, threads, lds, s);
    if (a.delay <= 4) return launch_sweep_d<T, V, 4>(a, nw, threads, lds, s);
    if (a.delay <= 8) return launch_sweep_d<T, V, 8>(a, nw, threads, lds, s);
    if (a.delay <= 12) return launch_sweep_d<T, V, 12>(a, nw, threads, lds, s);
    return launch_sweep_d<T, V, kMaxDelay>(a, nw, threads, lds, s);
}

// ---------------------------------------------------------------------------
// Flush: applies each walker's pending rank-c update to W and clears it.
// ---------------------------------------------------------------------------
template <typename T, int V>
__global__ void __launch_bounds__(1024) k_flush(SweepArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int w = a.walker0 + (int)blockIdx.x;
    const int c = a.npend[w];
    if (c == 0) return;
    const int ns = a.ns, nsa = w_nsa(ns, a.nup, a.ndn), nmax = a.nmax, delay = a.delay;
    T* s_A = reinterpret_cast<T*>(lds);
    T* s_B = s_A + 2 * delay * nsa;
    const int dp = pend_ld(delay);
    const T* gA = reinterpret_cast<const T*>(a.A) + (size_t)w * 2 * nsa * dp;
    const T* gB = reinterpret_cast<const T*>(a.Bm) + (size_t)w * 2 * nmax * dp;
    pend_load(s_A, s_B, gA, gB, nsa, nmax, delay, dp, c);
    __syncthreads();
    T* Wup = reinterpret_cast<T*>(a.W) + (size_t)w * w_stride(ns, a.nup, a.ndn);
    T* Wdn = Wup + w_off_dn(ns, a.nup);
    stream_update<T, V, kMaxDelay, 2>(Wup, s_A, s_B, w_ld(ns, a.nup), nsa, nmax, a.nup, c);
    stream_update<T, V, kMaxDelay, 2>(Wdn, s_A + delay * nsa, s_B + delay * nmax, w_ld(ns, a.ndn), nsa,
                                      nmax, a.ndn, c);
    __syncthreads();
    if (threadIdx.x == 0) a.npend[w] = 0;
}

template <typename T, int V>
hipError_t launch_flush_t(const SweepArgs& a, int nw, hipStream_t s) {
    const int nsa = w_nsa(a.ns, a.nup, a.ndn);
    const size_t lds = (size_t)2 * a.delay * (nsa + a.nmax) * sizeof(T);
    int threads = 256;
    const int ldm = w_ld(a.ns, a.nup) > w_ld(a.ns, a.ndn) ? w_ld(a.ns, a.nup) : w_ld(a.ns, a.ndn);
    const int nvec = ldm / V;
    while (threads < nvec) threads += 64;
    dim3 grid(nw), block(threads);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_flush<T, V>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_flush<T, V>), grid, block, lds, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Local energy (getOL, src/Hamiltonian.jl:762-778 with spinInteraction!
// :636-672 and SzInteraction! :593-604): one wavefront per walker, lanes
// stride over the bond list, FP64 accumulation, shuffle reduction.  The
// reference keeps the real part; complex W also yields the imaginary part
// (needed by the energy gradient of a complex wave function).
// ---------------------------------------------------------------------------
// The pending terms' part of W_cur[K][l], added to v: the c terms of row K
// and of label l are contiguous (pend_ld apart per row).
template <typename T>
__device__ __forceinline__ typename AccT<T>::type w_terms(typename AccT<T>::type v, const T* gA, const T* gB,
                                                          int dp, int K, int l, int c) {
    const T* ra = gA + (size_t)K * dp;
    const T* rb = gB + (size_t)l * dp;
    if constexpr (!is_cplx<T>::value) {
        // real: the terms as 16-byte vectors
        constexpr int NV = 16 / (int)sizeof(T);
        using VT = typename VecT<T, NV>::type;
        for (int q0 = 0; q0 < c; q0 += NV) {
            const VT x = *reinterpret_cast<const VT*>(ra + q0);
            const VT y = *reinterpret_cast<const VT*>(rb + q0);
#pragma unroll
            for (int u = 0; u < NV; ++u)
                if (q0 + u < c) v += (double)(x[u] * y[u]);
        }
    } else {
        for (int q = 0; q < c; ++q) v += to_acc(ra[q] * rb[q]);
    }
    return v;
}

// W_cur[K][l]: the flushed W0 (label-major, row stride ld) plus the terms.
template <typename T>
__device__ __forceinline__ typename AccT<T>::type w_elem(const T* W0, const T* gA, const T* gB, int ld, int dp,
                                                         int K, int l, int c) {
    return w_terms<T>(to_acc(W0[(size_t)l * ld + K]), gA, gB, dp, K, l, c);
}

// off -= 1/2 x y for the exchange term (real part), offi likewise (imaginary).
__device__ __forceinline__ void exch_acc(double& off, double&, double x, double y) { off += -0.5 * x * y; }
__device__ __forceinline__ void exch_acc(double& off, double& offi, cf64 x, cf64 y) {
    const cf64 p = x * y;
    off += -0.5 * p.re;
    offi += -0.5 * p.im;
}

// STAGE: each wavefront first copies its walker's labels and row maps into
// LDS (16-bit, coalesced), so the bond loop goes to memory only for W.
template <typename T, bool STAGE>
__global__ void __launch_bounds__(256) k_eloc(ElocArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int wave = (int)threadIdx.x / 64, lane = (int)threadIdx.x % 64;
    const int w = (int)blockIdx.x * ((int)blockDim.x / 64) + wave;
    if (w >= a.nw) return;   // no block-wide barrier below: each wavefront works alone
    const int ns = a.ns, nmax = a.nmax, nsp = round_up(ns, 4);
    const int ldu = w_ld(ns, a.nup), ldd = w_ld(ns, a.ndn);
    const T* Wup = reinterpret_cast<const T*>(a.W) + (size_t)w * w_stride(ns, a.nup, a.ndn);
    const T* Wdn = Wup + w_off_dn(ns, a.nup);
    const int* gku = a.kup + (size_t)w * ns;
    const int* gkd = a.kdn + (size_t)w * ns;
    const int* gsu = a.slot + (size_t)w * 2 * ns;     // rows of the (spin-)empty sites
    const int* gsd = gsu + ns;
    int16_t* sku = reinterpret_cast<int16_t*>(lds) + (size_t)wave * 4 * nsp;
    int16_t* skd = sku + nsp;
    int16_t* ssu = skd + nsp;
    int16_t* ssd = ssu + nsp;
    if (STAGE) {
        for (int t = lane; t < ns; t += 64) {
            sku[t] = (int16_t)gku[t];
            skd[t] = (int16_t)gkd[t];
            ssu[t] = (int16_t)gsu[t];
            ssd[t] = (int16_t)gsd[t];
        }
        // the wavefront reads what its own lanes wrote
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    auto ku = [&](int x) -> int { return STAGE ? (int)sku[x] : gku[x]; };
    auto kd = [&](int x) -> int { return STAGE ? (int)skd[x] : gkd[x]; };
    auto su = [&](int x) -> int { return STAGE ? (int)ssu[x] : gsu[x]; };
    auto sd = [&](int x) -> int { return STAGE ? (int)ssd[x] : gsd[x]; };
    const int c = a.npend[w];
    const int nsa = w_nsa(ns, a.nup, a.ndn);
    const int dp = pend_ld(a.delay);
    const T* gAu = reinterpret_cast<const T*>(a.A) + (size_t)w * 2 * nsa * dp;
    const T* gAd = gAu + (size_t)nsa * dp;
    const T* gBu = reinterpret_cast<const T*>(a.Bm) + (size_t)w * 2 * nmax * dp;
    const T* gBd = gBu + (size_t)nmax * dp;
    double diag = 0.0, off = 0.0, offi = 0.0;
    unsigned long long nap = 0;
#pragma unroll 2
    for (int b = lane; b < a.nb; b += 64) {
        const int2 bd = a.bonds[b];
        const int i = bd.x, j = bd.y;
        const int iu = ku(i), ju = ku(j), id = kd(i), jd = kd(j);
        const bool t1 = ju != 0 && id != 0;   // S+_i S-_j
        const bool t2 = iu != 0 && jd != 0;   // S-_i S+_j
        // the target sites are empty for the moving spin (no site holds both)
        if (t1)
            exch_acc(off, offi, w_elem(Wup, gAu, gBu, ldu, dp, max(su(i), 0), ju - 1, c),
                     w_elem(Wdn, gAd, gBd, ldd, dp, max(sd(j), 0), id - 1, c));
        if (t2)
            exch_acc(off, offi, w_elem(Wup, gAu, gBu, ldu, dp, max(su(j), 0), iu - 1, c),
                     w_elem(Wdn, gAd, gBd, ldd, dp, max(sd(i), 0), jd - 1, c));
        nap += (iu != 0 && jd != 0) || (ju != 0 && id != 0);
        const double szi = iu != 0 ? 0.5 : -0.5, szj = ju != 0 ? 0.5 : -0.5;
        diag += szi * szj;
    }
    double v = diag + off;
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, 64);
        if constexpr (is_cplx<T>::value) offi += __shfl_xor(offi, o, 64);
        nap += __shfl_xor(nap, o, 64);
    }
    if (lane == 0) {
        a.eloc[w] = v;
        if (is_cplx<T>::value && a.eloc_im) a.eloc_im[w] = offi;
        atomicAdd(&a.counters[CNT_NAP], nap);
    }
}

// FP32 with delayed updates: one workgroup per walker.  The labels, row
// maps and the walker's pending terms (A as [spin][row][dp], B as
// [spin][label][dp], both contiguous in HBM) are copied to LDS once, with
// 16-byte loads, so the bond loop gathers only W from memory; the 256 lanes
// split the bonds and a block reduction forms E_L.
template <typename T>
size_t eloc_wg_lds(const ElocArgs& a) {
    const size_t nsp = (size_t)(a.ns + 3) / 4 * 4, dp = (size_t)pend_ld(a.delay);
    return 8 * nsp + (2 * (size_t)w_nsa(a.ns, a.nup, a.ndn) + 2 * (size_t)a.nmax) * dp * sizeof(T) +
           32 * sizeof(double);
}

template <typename T>
__global__ void __launch_bounds__(256) k_eloc_wg(ElocArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int w = (int)blockIdx.x;
    const int ns = a.ns, nmax = a.nmax, nsp = round_up(ns, 4);
    const int nsa = w_nsa(ns, a.nup, a.ndn), dp = pend_ld(a.delay);
    const int ldu = w_ld(ns, a.nup), ldd = w_ld(ns, a.ndn);
    const T* Wup = reinterpret_cast<const T*>(a.W) + (size_t)w * w_stride(ns, a.nup, a.ndn);
    const T* Wdn = Wup + w_off_dn(ns, a.nup);
    int16_t* sku = reinterpret_cast<int16_t*>(lds);
    int16_t* skd = sku + nsp;
    int16_t* ssu = skd + nsp;
    int16_t* ssd = ssu + nsp;
    T* sAu = reinterpret_cast<T*>(ssd + nsp);             // [2][nsa][dp] (8 nsp bytes: 16-byte aligned)
    T* sAd = sAu + (size_t)nsa * dp;
    T* sBu = sAu + (size_t)2 * nsa * dp;                  // [2][nmax][dp]
    T* sBd = sBu + (size_t)nmax * dp;
    double* red = reinterpret_cast<double*>(sBu + (size_t)2 * nmax * dp);
    const int t0 = (int)threadIdx.x, bd = (int)blockDim.x;
    // copies unrolled so that several loads are in flight before the first
    // LDS store waits on them
#pragma unroll 2
    for (int t = t0; t < ns; t += bd) {
        sku[t] = (int16_t)a.kup[(size_t)w * ns + t];
        skd[t] = (int16_t)a.kdn[(size_t)w * ns + t];
        ssu[t] = (int16_t)a.slot[(size_t)w * 2 * ns + t];
        ssd[t] = (int16_t)a.slot[(size_t)w * 2 * ns + ns + t];
    }
    const int c = a.npend[w];
    if (c > 0) {
        // dp is a multiple of 4 elements: whole 16-byte vectors
        const size_t na = 2 * (size_t)nsa * dp * sizeof(T) / 16, nbv = 2 * (size_t)nmax * dp * sizeof(T) / 16;
        const f32x4* gA = reinterpret_cast<const f32x4*>(reinterpret_cast<const T*>(a.A) + (size_t)w * 2 * nsa * dp);
        const f32x4* gB = reinterpret_cast<const f32x4*>(reinterpret_cast<const T*>(a.Bm) + (size_t)w * 2 * nmax * dp);
        f32x4* dA = reinterpret_cast<f32x4*>(sAu);
        f32x4* dB = reinterpret_cast<f32x4*>(sBu);
#pragma unroll 4
        for (size_t t = t0; t < na; t += bd) dA[t] = gA[t];
#pragma unroll 4
        for (size_t t = t0; t < nbv; t += bd) dB[t] = gB[t];
    }
    __syncthreads();
    double diag = 0.0, off = 0.0, offi = 0.0, nap = 0.0;
    // A thread's bonds (BPT per pass) issue all their W gathers before any
    // is used: a bond with no move gathers W[0][0] and discards it.  Every
    // site holds one electron, so at most one of the two exchanges
    // (S+_i S-_j: i down, j up; S-_i S+_j: i up, j down) applies, and its
    // target sites are empty for the moving spin.
    constexpr int BPT = 4;
    for (int b0 = t0; b0 < a.nb; b0 += BPT * bd) {
        int Ku[BPT], Lu[BPT], Kd[BPT], Ld[BPT];
        bool mv[BPT];
        T w0u[BPT], w0d[BPT];
#pragma unroll
        for (int u = 0; u < BPT; ++u) {
            const int b = b0 + u * bd;
            const bool in = b < a.nb;
            const int2 bb = in ? a.bonds[b] : make_int2(0, 0);
            const int i = bb.x, j = bb.y;
            const int iu = sku[i], ju = sku[j], id = skd[i], jd = skd[j];
            const bool t1 = ju != 0 && id != 0, t2 = iu != 0 && jd != 0;
            mv[u] = in && (t1 || t2);
            Lu[u] = mv[u] ? (t1 ? ju : iu) - 1 : 0;
            Ku[u] = mv[u] ? max((int)(t1 ? ssu[i] : ssu[j]), 0) : 0;
            Ld[u] = mv[u] ? (t1 ? id : jd) - 1 : 0;
            Kd[u] = mv[u] ? max((int)(t1 ? ssd[j] : ssd[i]), 0) : 0;
            w0u[u] = Wup[(size_t)Lu[u] * ldu + Ku[u]];
            w0d[u] = Wdn[(size_t)Ld[u] * ldd + Kd[u]];
            if (in) {
                nap += (t1 || t2) ? 1.0 : 0.0;
                const double szi = iu != 0 ? 0.5 : -0.5, szj = ju != 0 ? 0.5 : -0.5;
                diag += szi * szj;
            }
        }
#pragma unroll
        for (int u = 0; u < BPT; ++u) {
            if (mv[u])
                exch_acc(off, offi, w_terms(to_acc(w0u[u]), sAu, sBu, dp, Ku[u], Lu[u], c),
                         w_terms(to_acc(w0d[u]), sAd, sBd, dp, Kd[u], Ld[u], c));
        }
    }
    const double v = block_sum_double(diag + off, red);
    const double napt = block_sum_double(nap, red + 8);
    double vi = 0.0;
    if constexpr (is_cplx<T>::value) vi = block_sum_double(offi, red + 16);
    if (t0 == 0) {
        a.eloc[w] = v;
        if (is_cplx<T>::value && a.eloc_im) a.eloc_im[w] = vi;
        atomicAdd(&a.counters[CNT_NAP], (unsigned long long)napt);
    }
}

// FP32 W with delayed updates and its pending terms within 64 KiB of LDS:
// one workgroup per walker; otherwise one wavefront per walker.
template <typename T, bool WG>
hipError_t launch_eloc_t(const ElocArgs& a, hipStream_t s) {
    if constexpr (WG) {
        if (a.delay > 1 && eloc_wg_lds<T>(a) <= 64 * 1024) {
            hipLaunchKernelGGL(k_eloc_wg<T>, dim3(a.nw), dim3(256), eloc_wg_lds<T>(a), s, a);
            return hipGetLastError();
        }
    }
    dim3 grid((a.nw + 3) / 4), block(256);
    const size_t lds = (size_t)4 * 4 * ((a.ns + 3) / 4 * 4) * sizeof(int16_t);   // 4 wavefronts x 4 arrays
    if (lds <= 64 * 1024) hipLaunchKernelGGL((k_eloc<T, true>), grid, block, lds, s, a);
    else hipLaunchKernelGGL((k_eloc<T, false>), grid, block, 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Log-derivatives.  With orthonormal orbital columns (U^H U = I, true for
// eigenvectors of a Hermitian H), tildeU^-1 = U^H W, so
//   O_k = tr(tildeU^-1 d tildeU) = sum_l sum_r W[r, l] Q_k[R_l, r],
//   Q_k = dU_k U^H (ns x ns, built once per ansatz).
// Occupied rows of W are unit rows, so the occupied sites contribute
// Q_k[R_l, R_l] and the empty ones W[slot(r), l] Q_k[R_l, r].
// One workgroup per walker; wavefronts take labels, lanes stride over r.
// Requires W without pending terms (the host flushes first).  Complex W:
// Q and O are complex (interleaved).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_grad(ElocArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    using A = typename AccT<T>::type;
    constexpr int CW = is_cplx<T>::value ? 2 : 1;      // doubles per value
    const int w = (int)blockIdx.x;
    const int ns = a.ns, nmax = a.nmax, P = a.nparams;
    int* Rsite = reinterpret_cast<int*>(lds);           // [2][nmax]
    double* acc = reinterpret_cast<double*>(Rsite + 2 * nmax + 2);  // [P][CW] (aligned enough)
    const int tid = (int)threadIdx.x;
    for (int s = tid; s < ns; s += (int)blockDim.x) {
        const int lu = a.kup[(size_t)w * ns + s], ld = a.kdn[(size_t)w * ns + s];
        if (lu) Rsite[lu - 1] = s;
        if (ld) Rsite[nmax + ld - 1] = s;
    }
    for (int p = tid; p < CW * P; p += (int)blockDim.x) acc[p] = 0.0;
    __syncthreads();
    const int wave = tid / 64, lane = tid % 64, nwave = (int)blockDim.x / 64;
    for (int spin = 0; spin < 2; ++spin) {
        const int N = spin ? a.ndn : a.nup;
        const int ld = w_ld(ns, N);
        const T* W = reinterpret_cast<const T*>(a.W) + (size_t)w * w_stride(ns, a.nup, a.ndn) +
                     (spin ? w_off_dn(ns, a.nup) : 0);
        const int* sl = a.slot + ((size_t)w * 2 + spin) * ns;
        const double* Q = spin ? a.Qdn : a.Qup;
        for (int p = 0; p < P; ++p) {
            const A* Qp = reinterpret_cast<const A*>(Q) + (size_t)p * ns * ns;
            A part = A(0.0);
            for (int l = wave; l < N; l += nwave) {
                const int R = Rsite[spin * nmax + l];
                const A* qrow = Qp + (size_t)R * ns;
                const T* wcol = W + (size_t)l * ld;
                if (lane == 0) part += qrow[R];
                for (int r = lane; r < ns; r += 64) {
                    const int e = sl[r];
                    if (e >= 0) part += to_acc(wcol[e]) * qrow[r];
                }
            }
            double pr = re_of(part), pi = im_of(part);
            for (int o = 32; o > 0; o >>= 1) {
                pr += __shfl_xor(pr, o, 64);
                if constexpr (CW == 2) pi += __shfl_xor(pi, o, 64);
            }
            if (lane == 0) {
                atomicAdd(&acc[CW * p], pr);
                if constexpr (CW == 2) atomicAdd(&acc[CW * p + 1], pi);
            }
        }
    }
    __syncthreads();
    for (int p = tid; p < CW * P; p += (int)blockDim.x) a.O[(size_t)w * CW * P + p] = acc[p];
}

// The same sums taken site by site, for up to PM parameters: every site s
// holds one electron (label l of the spin at s), so
//   O_k = sum_s ( Q_k[s, s] + sum_r W[slot(r), l] Q_k[s, r] )
// over the empty sites r of that spin.  A workgroup takes GW walkers with
// four wavefronts each; wavefront (walker g, quarter q) sums the sites of
// quarter q of walker g.  The GW wavefronts of a quarter walk its sites in
// the same order, so each Q row is fetched from L2 once per workgroup and
// served to the other walkers from L1.  Each W element is read once, for all
// parameters at a time.  Labels and slot maps are staged in LDS (16-bit), so
// a site costs one memory round trip (its W and Q loads together).
constexpr int kGradWalkers = 4;   // walkers per workgroup of k_grad_sites
template <typename T, int PM>
__global__ void __launch_bounds__(1024) k_grad_sites(ElocArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    using A = typename AccT<T>::type;
    constexpr int CW = is_cplx<T>::value ? 2 : 1;   // doubles per value
    constexpr int GW = kGradWalkers;
    const int wave = (int)threadIdx.x / 64, lane = (int)threadIdx.x % 64;
    const int g = wave % GW, q = wave / GW;          // walker slot, quarter of the sites
    const int w = (int)blockIdx.x * GW + g;
    const bool live = w < a.nw;                      // every wavefront reaches both barriers
    const int ns = a.ns, nsp = round_up(ns, 4), P = a.nparams;
    double* red = reinterpret_cast<double*>(lds);    // [GW][4][PM][CW]
    int16_t* mk = reinterpret_cast<int16_t*>(red + GW * 4 * PM * CW) + (size_t)g * 4 * nsp;   // kup, kdn, slots
    if (live) {
        const int* ku = a.kup + (size_t)w * ns;
        const int* kd = a.kdn + (size_t)w * ns;
        const int* sg = a.slot + (size_t)w * 2 * ns;
        for (int t = lane + 64 * q; t < ns; t += 256) {
            mk[t] = (int16_t)ku[t];
            mk[nsp + t] = (int16_t)kd[t];
            mk[2 * nsp + t] = (int16_t)sg[t];
            mk[3 * nsp + t] = (int16_t)sg[ns + t];
        }
    }
    __syncthreads();
    A part[PM];
#pragma unroll
    for (int p = 0; p < PM; ++p) part[p] = A(0.0);
    if (live) {
        const int ldu = w_ld(ns, a.nup), ldd = w_ld(ns, a.ndn);
        const T* Wup = reinterpret_cast<const T*>(a.W) + (size_t)w * w_stride(ns, a.nup, a.ndn);
        const T* Wdn = Wup + w_off_dn(ns, a.nup);
        const A* Qu = reinterpret_cast<const A*>(a.Qup);
        const A* Qd = reinterpret_cast<const A*>(a.Qdn);
        const size_t qstride = (size_t)ns * ns;      // one parameter's Q
        const int s0 = (int)(((long long)ns * q) / 4), s1 = (int)(((long long)ns * (q + 1)) / 4);
        for (int s = s0; s < s1; ++s) {
            const int lu = mk[s], lw = mk[nsp + s];
            if (lu == 0 && lw == 0) continue;        // not a valid kappa: nothing to add
            const bool up = lu != 0;
            const int l = (up ? lu : lw) - 1;
            const T* wcol = up ? Wup + (size_t)l * ldu : Wdn + (size_t)l * ldd;
            const int16_t* sl = mk + (up ? 2 : 3) * nsp;
            const A* qs = (up ? Qu : Qd) + (size_t)s * ns;
            if (lane == 0) {
#pragma unroll
                for (int p = 0; p < PM; ++p)
                    if (p < P) part[p] += qs[p * qstride + s];
            }
            for (int r = lane; r < ns; r += 64) {
                const int e = sl[r];
                if (e >= 0) {
                    const A wv = to_acc(wcol[e]);
#pragma unroll
                    for (int p = 0; p < PM; ++p)
                        if (p < P) part[p] += wv * qs[p * qstride + r];
                }
            }
        }
    }
#pragma unroll
    for (int p = 0; p < PM; ++p) {
        double pr = re_of(part[p]), pi = im_of(part[p]);
        for (int o = 32; o > 0; o >>= 1) {
            pr += __shfl_xor(pr, o, 64);
            if constexpr (CW == 2) pi += __shfl_xor(pi, o, 64);
        }
        if (lane == 0) {
            red[((g * 4 + q) * PM + p) * CW] = pr;
            if constexpr (CW == 2) red[((g * 4 + q) * PM + p) * CW + 1] = pi;
        }
    }
    __syncthreads();
    if (live && q == 0 && lane < CW * PM) {
        const int p = lane / CW, c = lane % CW;
        if (p < P) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) v += red[((g * 4 + k) * PM + p) * CW + c];
            a.O[(size_t)w * CW * P + CW * p + c] = v;
        }
    }
}

template <typename T>
hipError_t launch_grad_t(const ElocArgs& a, hipStream_t s) {
    constexpr int CW = is_cplx<T>::value ? 2 : 1;
    if (a.nparams <= 4) {
        constexpr int GW = kGradWalkers;
        const size_t lds = (size_t)GW * 4 * 4 * CW * sizeof(double) +                 // partial sums
                           (size_t)GW * 4 * ((a.ns + 3) / 4 * 4) * sizeof(int16_t);    // maps per walker
        hipLaunchKernelGGL((k_grad_sites<T, 4>), dim3((unsigned)((a.nw + GW - 1) / GW)), dim3(64 * 4 * GW), lds, s,
                           a);
        return hipGetLastError();
    }
    const size_t lds = (size_t)(2 * a.nmax + 2) * sizeof(int) + (size_t)CW * a.nparams * sizeof(double) + 16;
    hipLaunchKernelGGL((k_grad<T>), dim3(a.nw), dim3(256), lds, s, a);
    return hipGetLastError();
}

// Q_k[s][r] = sum_m dU_k[s, m] conj(U[r, m])  (row-major ns x ns per parameter;
// dU and U column-major ns x n; real or complex interleaved).
template <typename A>
__global__ void k_q_build(const A* __restrict__ dU, const A* __restrict__ U, int ns, int n, int nparams,
                          A* __restrict__ Q) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)nparams * ns * ns;
    if (idx >= total) return;
    const int p = (int)(idx / ((long long)ns * ns));
    const int rem = (int)(idx % ((long long)ns * ns));
    const int s = rem / ns, r = rem % ns;
    const A* d = dU + (size_t)p * ns * n;
    A v = A(0.0);
    for (int m = 0; m < n; ++m) {
        if constexpr (is_cplx<A>::value) {
            const A u = U[(size_t)m * ns + r];
            v += d[(size_t)m * ns + s] * A(u.re, -u.im);
        } else {
            v += d[(size_t)m * ns + s] * U[(size_t)m * ns + r];
        }
    }
    Q[idx] = v;
}

template <typename A>
hipError_t launch_q_build_t(const double* dU, const double* U, int ns, int n, int nparams, double* Q,
                            hipStream_t s) {
    const long long total = (long long)nparams * ns * ns;
    dim3 grid((unsigned)((total + 255) / 256)), block(256);
    hipLaunchKernelGGL((k_q_build<A>), grid, block, 0, s, reinterpret_cast<const A*>(dU),
                       reinterpret_cast<const A*>(U), ns, n, nparams, reinterpret_cast<A*>(Q));
    return hipGetLastError();
}

}  // namespace kvmc
