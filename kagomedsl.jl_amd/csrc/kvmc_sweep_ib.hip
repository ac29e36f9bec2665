// FP32 sweep kernel with implicit pending rows (gfx950).
//
// Carlo.sweep! (src/MonteCarlo.jl:538-607) with delayed Sherman-Morrison
// updates (update_W!, :279-292) for FP32 W streamed from HBM, when a column
// of W fits one wavefront (at most 256 rows: 4 per lane).
//
// The delayed update of k_sweep keeps, per accepted move q, a column A_q and
// a row B_q = alpha_q (W_cur[K_q, :] - e_l), and building B_q reads the row
// K_q of W: N single elements at a stride of ld, one cache line each.  Here
// the rows are never read at an accept.  With g_r = W0[K_r, :] - e_{l_r} (W0:
// W in HBM since the last flush) every row term is a combination of the g_r,
// B_q = sum_{r<=q} T[q][r] g_r, so
//
//   W_cur = W0 + sum_r At_r g_r^T,    At_r = sum_{q>=r} A_q T[q][r],
//
// and an accept updates the columns At in LDS only:
//   A_c   = W_cur[:, l] + e_K          (column l of W0 plus the terms),
//   At_r += A_c * alpha At_r[K]        for r < c,
//   At_c  = alpha A_c,                 alpha = -1 / W_cur[K, l].
// The flush applies W0 += sum_r At_r g_r^T with one wavefront per column of
// W: lane i holds rows 4i..4i+3, and g_r[j] = W0[K_r, j] - delta(j, l_r) is
// read from the column itself (through the wavefront's LDS slot), so the
// rows cost no traffic.
// A proposal needs W_cur[K, l] = W0[K, l] + sum_r At_r[K] (W0[K_r, l] -
// delta(l, l_r)): c + 1 elements of column l, all issued before the first is
// used.  Proposal windows are WIN lanes wide so that few proposals past the
// accept are evaluated (16 lanes measured fastest among 8, 16, 32 and 64).
//
// Between launches the pending terms are kept in the layout of k_sweep (A as
// [row][term], B as [label][term]) with A = At and B = g, so the local
// energy, the flush kernel, get_w and the refresh drift check read them
// unchanged; the rows g are gathered once at launch exit.  The term rows and
// labels (K_r, l_r) are kept beside them for the next launch.
#include "kvmc_kernels.hpp"

namespace kvmc {

__host__ __device__ inline size_t sweep_ib_lds(int ns, int nsa, int delay, int nb) {
    const size_t nsp = (size_t)(ns + 3) / 4 * 4;
    const size_t b = 4 * nsp * sizeof(int16_t)                  // kappa and slot maps
                     + 2 * (size_t)delay * nsa * sizeof(float)    // At columns
                     + 16 * sizeof(double)                        // reductions
                     + 4 * (size_t)delay * sizeof(int)            // term rows and labels
                     + 2 * (size_t)delay * sizeof(float)          // an accept's coefficients
                     + 4 * (size_t)2 * 256 * sizeof(float)   // flush staging, 4 wavefronts (also the accept's columns)
                     + (size_t)nb * sizeof(uint32_t) + (size_t)(ns + 1 + 2 * nb) * sizeof(uint16_t);
    return (b + 15) / 16 * 16;
}

__device__ __forceinline__ float lane_elem(const f32x4& v, int comp) {
    return comp == 0 ? v.x : comp == 1 ? v.y : comp == 2 ? v.z : v.w;
}

// W0 += sum_{r<c} At_r g_r^T for one spin (ld rows, N columns).  Wavefront
// `wave` of `nwave` takes columns wave, wave + nwave, ...; lane i holds
// rows 4i .. 4i+3 (ld <= 256).  K and L (uniform) are the terms' rows and
// labels.  Each loaded column is also written to the wavefront's own LDS
// slot (stg: UNR columns of 256 floats), and g_r[j] = W0[K_r, j] -
// delta(j, l_r) is read back as an LDS broadcast before the column is
// updated.  A wavefront reads only what it wrote, in order, so no barrier.
template <int DMAX, int UNR>
__device__ __forceinline__ void flush_ib(float* __restrict__ W0, const float* __restrict__ sAt, const int* sK,
                                         const int* sL, int ld, int nsa, int N, int c, float* s_stage) {
    const int lane = (int)threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int nwave = (int)blockDim.x >> 6;
    const bool act = 4 * lane < ld;
    float* stg = s_stage + wave * UNR * 256;
    f32x4 at[DMAX];
    int K[DMAX], L[DMAX];
#pragma unroll
    for (int q = 0; q < DMAX; ++q) {
        at[q] = (q < c && act) ? *reinterpret_cast<const f32x4*>(sAt + (size_t)q * nsa + 4 * lane) : f32x4(0.f);
        K[q] = q < c ? __builtin_amdgcn_readfirstlane(sK[q]) : 0;
        L[q] = q < c ? __builtin_amdgcn_readfirstlane(sL[q]) : -1;
    }
    float* __restrict__ base = W0 + 4 * lane;
    for (int j0 = wave; j0 < N; j0 += UNR * nwave) {
        f32x4 v[UNR], d[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int j = j0 + u * nwave;
            v[u] = (act && j < N) ? *reinterpret_cast<const f32x4*>(base + (size_t)j * ld) : f32x4(0.f);
            d[u] = f32x4(0.f);
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) *reinterpret_cast<f32x4*>(stg + u * 256 + 4 * lane) = v[u];
#pragma unroll
        for (int q = 0; q < DMAX; ++q) {
            if (q < c) {
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    const int j = j0 + u * nwave;
                    const float g = stg[u * 256 + K[q]] - (j == L[q] ? 1.f : 0.f);
                    d[u] += at[q] * g;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int j = j0 + u * nwave;
            if (act && j < N) *reinterpret_cast<f32x4*>(base + (size_t)j * ld) = v[u] + d[u];
        }
    }
}

// Columns of W in flight per wavefront in the flush.  Two keep the LDS of a
// walker at L=12 and delay 12 within 40 KB, so four walkers share a CU.
template <int DMAX> struct FlushUnroll { static constexpr int value = 2; };

// Four waves per SIMD (128 VGPRs) up to delay 12: four walkers per CU hide
// more of the proposal and accept latency (8.8 % less sweep time at L=12 than
// three); the delay-16 register set needs three.
template <int DMAX, int WIN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DMAX <= 12 ? 4 : 3))) k_sweep_ib(SweepArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int w = a.walker0 + (int)blockIdx.x;
    const int ns = a.ns, nup = a.nup, ndn = a.ndn, nb = a.nb, nmax = a.nmax;
    const int nsp = round_up(ns, 4);
    const int nsa = w_nsa(ns, nup, ndn);
    const int delay = a.delay;
    int16_t* s_kup = reinterpret_cast<int16_t*>(lds);
    int16_t* s_kdn = s_kup + nsp;
    int16_t* s_slu = s_kdn + nsp;                    // site -> row of W_up (-1: occupied)
    int16_t* s_sld = s_slu + nsp;
    float* s_At = reinterpret_cast<float*>(s_sld + nsp);   // [2][delay][nsa]
    constexpr int UNR = FlushUnroll<DMAX>::value;
    float* s_stage = s_At + 2 * delay * nsa;               // [4 wavefronts][UNR][256] flush staging (16-byte aligned)
    float* s_col = s_stage;                                // [2][nsa] an accept's columns (not live during a flush)
    double* s_red = reinterpret_cast<double*>(s_stage + 4 * UNR * 256);
    int* s_K = reinterpret_cast<int*>(s_red + 16);         // [2][delay] term rows
    int* s_L = s_K + 2 * delay;                            // [2][delay] term labels
    float* s_tc = reinterpret_cast<float*>(s_L + 2 * delay);   // [2][delay] alpha At_q[K] of an accept
    uint32_t* s_bond = reinterpret_cast<uint32_t*>(s_tc + 2 * delay);
    uint16_t* s_noff = reinterpret_cast<uint16_t*>(s_bond + nb);
    uint16_t* s_nbr = s_noff + ns + 1;

    const size_t wstride = w_stride(ns, nup, ndn);
    const int ldu = w_ld(ns, nup), ldd = w_ld(ns, ndn);
    float* const Wup = reinterpret_cast<float*>(a.W) + (size_t)w * wstride;
    float* const Wdn = Wup + w_off_dn(ns, nup);
    int* const slot_g = a.slot + (size_t)w * 2 * ns;
    const int dp = pend_ld(delay);
    float* gA = reinterpret_cast<float*>(a.A) + (size_t)w * 2 * nsa * dp;
    float* gB = reinterpret_cast<float*>(a.Bm) + (size_t)w * 2 * nmax * dp;
    int* pt = a.pterm + (size_t)w * 2 * kMaxDelay * 2;
    float* sAu = s_At;
    float* sAd = s_At + delay * nsa;

    long long sweeps = a.sweeps[w];
    const long long stop = a.stop[w];
    if (sweeps >= stop) return;
    int c = a.npend[w];
    int acc_since = a.acc_since[w];
    const int tid = (int)threadIdx.x, bd = (int)blockDim.x;

    // The copies into LDS are unrolled so that several loads are in flight
    // before the first store waits on them.
#pragma unroll 2
    for (int t = tid; t < ns; t += bd) {
        s_kup[t] = (int16_t)a.kup[(size_t)w * ns + t];
        s_kdn[t] = (int16_t)a.kdn[(size_t)w * ns + t];
        s_slu[t] = (int16_t)slot_g[t];
        s_sld[t] = (int16_t)slot_g[ns + t];
    }
    if (c > 0) {
        // the At columns ([spin][row][dp] in HBM, dp a multiple of 4: 16-byte
        // loads) and the term rows and labels
#pragma unroll 2
        for (int t = tid; t < 2 * nsa; t += bd) {
            const int sp = t >= nsa ? 1 : 0;
            float* d = s_At + sp * delay * nsa + (t - sp * nsa);
            const f32x4* g = reinterpret_cast<const f32x4*>(gA + (size_t)t * dp);
            f32x4 v[(DMAX + 3) / 4];
#pragma unroll
            for (int q4 = 0; q4 < (DMAX + 3) / 4; ++q4)
                if (4 * q4 < c) v[q4] = g[q4];
#pragma unroll
            for (int q4 = 0; q4 < (DMAX + 3) / 4; ++q4) {
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (4 * q4 + u < c) d[(4 * q4 + u) * nsa] = v[q4][u];
            }
        }
        for (int t = tid; t < 2 * c; t += bd) {
            const int sp = t / c, q = t % c;
            s_K[sp * delay + q] = pt[(sp * kMaxDelay + q) * 2];
            s_L[sp * delay + q] = pt[(sp * kMaxDelay + q) * 2 + 1];
        }
    }
#pragma unroll 4
    for (int b = tid; b < nb; b += bd) {
        const int2 bb = a.bonds[b];
        s_bond[b] = (uint32_t)bb.x | ((uint32_t)bb.y << 16);
    }
#pragma unroll 2
    for (int t = tid; t <= ns; t += bd) s_noff[t] = (uint16_t)a.nbr_off[t];
    const int nnbr = a.nbr_off[ns];
#pragma unroll 4
    for (int t = tid; t < nnbr; t += bd) s_nbr[t] = (uint16_t)a.nbr[t];
    __syncthreads();
    int zmu;
    {
        double part = 0.0;
        for (int b = tid; b < nb; b += bd) {
            const uint32_t bp = s_bond[b];
            const int i = (int)(bp & 0xffffu), j = (int)(bp >> 16);
            part += (s_kup[i] != 0 && s_kdn[j] != 0) || (s_kup[j] != 0 && s_kdn[i] != 0);
        }
        zmu = (int)block_sum_double(part, s_red);
    }
    const double zmax = (double)nb;
    const long long tbase = a.tape_base[w];
    unsigned long long n_moves = 0, n_acc = 0, n_flush = 0, n_sweeps = 0;
    bool pause = false;
    const int lane = tid & 63;

    // Proposal windows of WIN consecutive sweeps, one per lane (as in
    // k_sweep): every wavefront evaluates the same window, and the walker
    // advances to the first sweep that accepts or triggers a refresh.
    while (sweeps < stop && !pause) {
        const long long sw = sweeps + lane;
        const bool valid = lane < WIN && sw < stop;
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
        int i = 0, site = 0, lup = 0, ldn = 0, Kup = 0, Kdn = 0, dz = 0;
        bool can1 = false;
        float wu = 0.f, wd = 0.f;
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
                lup = (can1 ? iu : su) - 1;
                ldn = (can1 ? sd : id) - 1;
                Kup = can1 ? site : i;
                Kdn = can1 ? i : site;
                const int eu = max((int)s_slu[Kup], 0), ed = max((int)s_sld[Kdn], 0);
                const float* cu = Wup + (size_t)lup * ldu;
                const float* cd = Wdn + (size_t)ldn * ldd;
                // every gather is issued before any is used (one memory round
                // trip): slots past the live terms re-read row e, in bounds.
                // (Skipping the groups of 4 past c with a uniform branch kept
                // more registers live: 35 spills instead of 12, 11 % slower.)
                float gu[DMAX], gd[DMAX];
#pragma unroll
                for (int q = 0; q < DMAX; ++q) {
                    gu[q] = cu[q < c ? (int)s_K[q] : eu];
                    gd[q] = cd[q < c ? (int)s_K[delay + q] : ed];
                }
                float vu = cu[eu], vd = cd[ed];
#pragma unroll
                for (int q = 0; q < DMAX; ++q) {
                    if (q < c) {
                        vu += sAu[q * nsa + eu] * (gu[q] - (lup == s_L[q] ? 1.f : 0.f));
                        vd += sAd[q * nsa + ed] * (gd[q] - (ldn == s_L[delay + q] ? 1.f : 0.f));
                    }
                }
                wu = vu;
                wd = vd;
                dz = z_delta(s_bond, s_noff, s_nbr, s_kup, s_kdn, i, site);
                const float ratio = wu * wd;
                const double acc_prob = (double)ratio * (double)ratio;
                if (a.kernel_mode == KERNEL_REFERENCE)
                    acc = (acc_prob >= 1.0 && r < zr) || (acc_prob < 1.0 && r < zr * acc_prob);
                else
                    acc = r < acc_prob;
                oc = acc ? OUT_ACCEPT : OUT_REJECT;
            }
        }
        const bool term = acc || (mv && a.policy == POLICY_REFERENCE && (sw % a.nocc) == 0);
        const unsigned long long tmask = __ballot(valid && term);
        const int nvalid = __popcll(__ballot(valid));
        const int ncommit = tmask ? __ffsll((long long)tmask) : nvalid;
        if (a.outcomes && tid < 64 && lane < ncommit)
            a.outcomes[(size_t)w * a.tape_len + (sw - tbase)] = (signed char)oc;
        n_moves += __popcll(__ballot(mv && lane < ncommit));
        sweeps += ncommit;
        n_sweeps += ncommit;
        if (!tmask) continue;
        const int tl = ncommit - 1;
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
            const float twu = readlane_f(wu, tl), twd = readlane_f(wd, tl);
            const int tdz = __builtin_amdgcn_readlane(dz, tl);
            ++n_acc;
            ++acc_since;
            const int eu = max((int)s_slu[tKup], 0), ed = max((int)s_sld[tKdn], 0);
            const float au = -1.f / twu, ad = -1.f / twd;
            // the columns of the moving labels, as stored (one element per
            // thread and spin: blockDim.x >= nsa), and the new term's
            // coefficients on the old terms, alpha At_q[K].  At is unchanged
            // until the barrier: every wavefront has left its window.
            {
                const float cu = tid < ldu ? Wup[(size_t)tlup * ldu + tid] : 0.f;
                const float cd = tid < ldd ? Wdn[(size_t)tldn * ldd + tid] : 0.f;
                if (tid < c) {
                    s_tc[tid] = au * sAu[tid * nsa + eu];
                    s_tc[delay + tid] = ad * sAd[tid * nsa + ed];
                }
                if (tid < nsa) { s_col[tid] = cu; s_col[nsa + tid] = cd; }
            }
            __syncthreads();
            // A_c = W_cur[:, l] + e_K, one row per thread and spin; each thread
            // then updates only its own row of At, so no barrier is needed
            // between reading and writing At
            float Au = 0.f, Ad = 0.f;
            if (tid < ldu) {
                float v = s_col[tid];
                for (int q = 0; q < c; ++q)
                    v += sAu[q * nsa + tid] * (s_col[s_K[q]] - (tlup == s_L[q] ? 1.f : 0.f));
                Au = v + (tid == eu ? 1.f : 0.f);
            }
            if (tid < ldd) {
                float v = s_col[nsa + tid];
                for (int q = 0; q < c; ++q)
                    v += sAd[q * nsa + tid] * (s_col[nsa + s_K[delay + q]] - (tldn == s_L[delay + q] ? 1.f : 0.f));
                Ad = v + (tid == ed ? 1.f : 0.f);
            }
            if (tid < nsa) {
                const bool ru = tid < ldu, rd = tid < ldd;
                for (int q = 0; q < c; ++q) {
                    if (ru) sAu[q * nsa + tid] += Au * s_tc[q];
                    if (rd) sAd[q * nsa + tid] += Ad * s_tc[delay + q];
                }
                sAu[c * nsa + tid] = ru ? au * Au : 0.f;
                sAd[c * nsa + tid] = rd ? ad * Ad : 0.f;
            }
            if (tid == 0) {
                s_K[c] = eu; s_L[c] = tlup;
                s_K[delay + c] = ed; s_L[delay + c] = tldn;
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
            __syncthreads();
            if (c == delay) {
                flush_ib<DMAX, UNR>(Wup, sAu, s_K, s_L, ldu, nsa, nup, c, s_stage);
                flush_ib<DMAX, UNR>(Wdn, sAd, s_K + delay, s_L + delay, ldd, nsa, ndn, c, s_stage);
                c = 0;
                ++n_flush;
                __syncthreads();
            }
            zmu += tdz;
            refresh_due = a.policy == POLICY_REFERENCE ? (t_sw % a.nocc) == 0
                        : a.policy == POLICY_EVERY_K ? acc_since >= a.refresh_k
                                                     : false;
        } else {
            refresh_due = true;
        }
        if (refresh_due) pause = true;
    }
    __syncthreads();
    for (int t = tid; t < ns; t += bd) {
        a.kup[(size_t)w * ns + t] = s_kup[t];
        a.kdn[(size_t)w * ns + t] = s_kdn[t];
        slot_g[t] = s_slu[t];
        slot_g[ns + t] = s_sld[t];
    }
    // the terms are kept also for a walker due a refresh: the refresh may
    // measure the drift of the carried W first, and a walker whose refresh
    // fails keeps a W consistent with its configuration
    const bool listed = pause || (a.policy == POLICY_AT_END && acc_since >= a.refresh_k);
    if (c > 0) {
        // written as 16-byte vectors (slots c.. of the last one get zeros;
        // readers use the first c); all of a label's gathers are issued
        // before the first is stored
        constexpr int NV4 = (DMAX + 3) / 4;
        for (int t = tid; t < 2 * nsa; t += bd) {
            const int sp = t >= nsa ? 1 : 0;
            const float* d = s_At + sp * delay * nsa + (t - sp * nsa);
            f32x4* g = reinterpret_cast<f32x4*>(gA + (size_t)t * dp);
#pragma unroll
            for (int q4 = 0; q4 < NV4; ++q4) {
                if (4 * q4 < c) {
                    f32x4 v;
#pragma unroll
                    for (int u = 0; u < 4; ++u) v[u] = 4 * q4 + u < c ? d[(4 * q4 + u) * nsa] : 0.f;
                    g[q4] = v;
                }
            }
        }
        // the rows g_q = W0[K_q, :] - e_{l_q}, one label per thread
        for (int t = tid; t < 2 * nmax; t += bd) {
            const int sp = t >= nmax ? 1 : 0, l = t - sp * nmax;
            if (l >= (sp ? ndn : nup)) continue;
            const float* col = (sp ? Wdn : Wup) + (size_t)l * (sp ? ldd : ldu);
            float v[4 * NV4];
#pragma unroll
            for (int q = 0; q < 4 * NV4; ++q)
                v[q] = q < c ? col[s_K[sp * delay + q]] - (l == s_L[sp * delay + q] ? 1.f : 0.f) : 0.f;
            f32x4* g = reinterpret_cast<f32x4*>(gB + (size_t)t * dp);
#pragma unroll
            for (int q4 = 0; q4 < NV4; ++q4)
                if (4 * q4 < c) g[q4] = f32x4{v[4 * q4], v[4 * q4 + 1], v[4 * q4 + 2], v[4 * q4 + 3]};
        }
        for (int t = tid; t < 2 * c; t += bd) {
            const int sp = t / c, q = t % c;
            pt[(sp * kMaxDelay + q) * 2] = s_K[sp * delay + q];
            pt[(sp * kMaxDelay + q) * 2 + 1] = s_L[sp * delay + q];
        }
    }
    if (tid == 0) {
        a.sweeps[w] = sweeps;
        a.npend[w] = c;
        a.acc_since[w] = acc_since;
        a.zmu[w] = zmu;
        if (listed) {
            const int s = atomicAdd(a.refresh_count, 1);
            a.refresh_list[s] = w;
            atomicAdd(&a.counters[CNT_PAUSES], 1ull);
        }
        atomicAdd(&a.counters[CNT_MOVES], n_moves);
        atomicAdd(&a.counters[CNT_ACCEPTED], n_acc);
        atomicAdd(&a.counters[CNT_FLUSHES], n_flush);
        atomicAdd(&a.counters[CNT_SWEEPS], n_sweeps);
    }
}

bool sweep_ib_supported(int prec, int cplx, int ns, int nup, int ndn, int delay) {
    const int ld = w_ld(ns, nup) > w_ld(ns, ndn) ? w_ld(ns, nup) : w_ld(ns, ndn);
    return prec == 1 && !cplx && delay > 1 && delay <= kMaxDelay && ld <= 256 &&
           !sweep_resident(sizeof(float), ns, nup, ndn);
}

size_t sweep_ib_lds_bytes(int ns, int nup, int ndn, int delay, int nb) {
    return sweep_ib_lds(ns, w_nsa(ns, nup, ndn), delay, nb);
}

int sweep_ib_threads(int ns, int nup, int ndn) {
    const int nsa = w_nsa(ns, nup, ndn);
    return nsa <= 64 ? 64 : nsa <= 128 ? 128 : nsa <= 192 ? 192 : 256;
}

template <int D>
static hipError_t launch_ib_d(const SweepArgs& a, int nw, hipStream_t s) {
    constexpr int WIN = 16;
    const size_t lds = sweep_ib_lds_bytes(a.ns, a.nup, a.ndn, a.delay, a.nb);
    const void* fn = reinterpret_cast<const void*>(&k_sweep_ib<D, WIN>);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_sweep_ib<D, WIN>), dim3(nw), dim3(sweep_ib_threads(a.ns, a.nup, a.ndn)), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_sweep_ib(const SweepArgs& a, int nw, hipStream_t s) {
    if (a.delay <= 8) return launch_ib_d<8>(a, nw, s);
    if (a.delay <= 12) return launch_ib_d<12>(a, nw, s);
    return launch_ib_d<kMaxDelay>(a, nw, s);
}

}  // namespace kvmc
