// W refresh (reevaluateW!, src/MonteCarlo.jl:55-66) as a blocked FP64
// factorisation on gfx950.
//
// Work matrix (row-major, zero padded): M = [ tildeU ; U_empty ], ns x N.
// Physical rows 0..N-1 are U at the occupied sites ordered by label (tildeU,
// as in tilde_U, src/MonteCarlo.jl:92-115); rows N..ns-1 are U at the empty
// sites.  Row interchanges are implicit: phys[] maps the logical (pivoted)
// row order onto physical rows, so no row of M ever moves.
//
// Two paths share the LU code (k_refresh_blocked, one 256-thread workgroup
// per (walker, spin)):
//
// N <= 256: two kernels.
//   k_refresh_blocked<FULL = false> factors tildeU alone: P tildeU = L1 U1,
//   right-looking, panels of RB = 32 columns; the panel lives in registers
//   (one row per thread) and one barrier per pivot column publishes each
//   wavefront's best candidate row.  U12 = L11^-1 A12 and the trailing
//   update run on FP64 MFMA (v_mfma_f64_16x16x4f64).  It saves L1, U1, the
//   32x32 inverses of the diagonal blocks of L1 and U1 and the final row map.
//   k_refresh_solve then forms the empty-site rows of W: every wavefront
//   holds 16 empty sites, i.e. their rows b of U transposed, as 16x16 MFMA
//   accumulator tiles in registers, and solves y U1 = b and x L1 = y in place
//   (block rows of the factors staged through LDS, shared by the four
//   wavefronts).  W[s][phys[c]] = x[c]; occupied rows of W are unit rows.
//
// N > 256 (register tiles would not fit): one kernel, k_refresh_blocked<FULL
// = true>, factors the tall matrix [tildeU ; U_empty] (rows >= N end as
// L2 = U_empty U1^-1), solves X L1 = L2 panel by panel on MFMA and writes
// W_empty = X P through LDS.
#include "kvmc_internal.hpp"

#include <math.h>
#include <type_traits>

namespace kvmc {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4f64 operand layout (lane l of a wavefront):
//   A[i][k] with i = l % 16, k = l / 16;  B[k][j] with k = l / 16, j = l % 16;
//   D register g holds D[l / 16 + 4 g][l % 16].
__device__ __forceinline__ f64x4 mfma16(double a, double b, f64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int RB = 32;    // panel width
constexpr int RT = 256;   // threads per workgroup
constexpr int NWAVE = RT / 64;
constexpr int LB = RB + 1;  // LDS row stride of 32-wide blocks
constexpr int LU_WAVES = 3;  // waves per SIMD of the LU-only kernel
constexpr int kStoreBlock = 8;   // sites per LDS block of the single kernel's W store

__host__ __device__ inline int rup_i(int x, int m) { return (x + m - 1) / m * m; }
__host__ __device__ inline int refresh_nrows(int ns) { return rup_i(ns, 16) + 16; }
// Register tiles per wavefront of k_refresh_solve for N <= 256: 4, 8, 14 or
// 16 tiles of 16 labels (0: N > 256).
__host__ __device__ inline int solve_nt(int nmax) {
    const int nt = (nmax + 15) / 16;
    return nt <= 4 ? 4 : nt <= 8 ? 8 : nt <= 14 ? 14 : nt <= 16 ? 16 : 0;
}
// Row stride of M: whole 16 * NT rows for the solve kernel's row loads.
__host__ __device__ inline int refresh_ldm(int nmax) { return solve_nt(nmax) ? 16 * solve_nt(nmax) : rup_i(nmax, 32); }
__host__ __device__ inline int refresh_npanel(int nmax) { return (nmax + RB - 1) / RB; }
// Work doubles per job: the padded tall matrix, the saved 32x32 inverses of
// the diagonal blocks of L1 and of U1, then ints: [0] factorisation ok flag,
// [1..nmax] logical row -> label.  Rounded to 256 bytes.
__host__ __device__ inline size_t refresh_ws_job(int ns, int nmax) {
    const size_t d = (size_t)refresh_nrows(ns) * refresh_ldm(nmax) + 2 * (size_t)refresh_npanel(nmax) * RB * RB +
                     (size_t)(nmax + 2) / 2 + 1;
    return (d + 31) / 32 * 32;
}

__device__ inline void atomic_max_nonneg2(double* addr, double v) {
    atomicMax(reinterpret_cast<unsigned long long*>(addr),
              (unsigned long long)__double_as_longlong(v));
}

// Inverse of the unit-lower jb x jb block L (row stride LB; the diagonal and
// the entries above it are not read) into Linv (row stride LB), identity
// beyond jb.  Lanes 0..RB-1 of wavefront `wv` each hold one column of the
// inverse in registers and run its forward substitution fully unrolled (four
// partial sums per row to shorten the dependency chain); L is read as LDS
// broadcasts.
__device__ __attribute__((noinline)) void unit_lower_inverse(const double* L, double* Linv, int jb, int wv) {
    if ((int)threadIdx.x / 64 != wv) return;
    const int j = (int)threadIdx.x & (RB - 1);
    double x[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        double v = (i == j) ? 1.0 : 0.0;
        if (i > j && i < jb) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
            for (int k = 0; k < i; k += 4) {
                s0 += L[i * LB + k] * x[k];
                if (k + 1 < i) s1 += L[i * LB + k + 1] * x[k + 1];
                if (k + 2 < i) s2 += L[i * LB + k + 2] * x[k + 2];
                if (k + 3 < i) s3 += L[i * LB + k + 3] * x[k + 3];
            }
            v = -((s0 + s1) + (s2 + s3));
        }
        x[i] = v;
    }
    if ((threadIdx.x & 63) < RB) {
#pragma unroll
        for (int i = 0; i < RB; ++i) Linv[i * LB + j] = x[i];
    }
}

// 16 x 16 diagonal blocks at (off, off) of the same inverses, for the
// two-kernel path, one block per wavefront: lanes 0..15 of wavefront `wv`
// each hold one column.  Rows >= jb of the block are unit rows.
__device__ __attribute__((noinline)) void unit_lower_inverse16(const double* L, double* Linv, int off, int jb, int wv) {
    if ((int)threadIdx.x / 64 != wv) return;
    const int j = (int)threadIdx.x & 15;
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        double v = (i == j) ? 1.0 : 0.0;
        if (i > j && off + i < jb) {
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int k = 0; k < i; k += 2) {
                s0 += L[(off + i) * LB + off + k] * x[k];
                if (k + 1 < i) s1 += L[(off + i) * LB + off + k + 1] * x[k + 1];
            }
            v = -(s0 + s1);
        }
        x[i] = v;
    }
    if ((threadIdx.x & 63) < 16) {
#pragma unroll
        for (int i = 0; i < 16; ++i) Linv[(off + i) * LB + off + j] = x[i];
    }
}

__device__ __forceinline__ unsigned long long readlane64(unsigned long long k, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(k >> 32), l);
    return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
    return __longlong_as_double((long long)readlane64((unsigned long long)__double_as_longlong(v), l));
}

// Inverse of the upper-triangular block D (row stride LB, entries below the
// diagonal ignored; rows >= jb must be unit rows) into Uinv (row stride ld,
// LDS or global memory).  Lanes 0..RB-1 of wavefront `wv` each back-substitute
// one column j in registers: x_i = (delta_ij - sum_{k>i} D[i][k] x_k) /
// D[i][i]; entries below the diagonal come out zero without special cases.
// Lane i computes 1/D[i][i] once and the others read it back through readlane.
__device__ __attribute__((noinline)) void upper_inverse(const double* D, double* Uinv, int ld, int wv) {
    if ((int)threadIdx.x / 64 != wv) return;
    const int lane = (int)threadIdx.x & 63;
    const int j = lane & (RB - 1);
    const double dinv = 1.0 / D[j * LB + j];
    double x[RB];
#pragma unroll
    for (int i = RB - 1; i >= 0; --i) {
        double s0 = (i == j) ? 1.0 : 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
        for (int k = i + 1; k < RB; k += 4) {
            s0 -= D[i * LB + k] * x[k];
            if (k + 1 < RB) s1 -= D[i * LB + k + 1] * x[k + 1];
            if (k + 2 < RB) s2 -= D[i * LB + k + 2] * x[k + 2];
            if (k + 3 < RB) s3 -= D[i * LB + k + 3] * x[k + 3];
        }
        x[i] = ((s0 + s1) + (s2 + s3)) * readlane_f64(dinv, i);
    }
    if (lane < RB) {
#pragma unroll
        for (int i = 0; i < RB; ++i) Uinv[i * ld + j] = x[i];
    }
}

__device__ __attribute__((noinline)) void upper_inverse16(const double* D, double* Uinv, int ld, int off, int wv) {
    if ((int)threadIdx.x / 64 != wv) return;
    const int lane = (int)threadIdx.x & 63;
    const int j = lane & 15;
    const double dinv = 1.0 / D[(off + j) * LB + off + j];
    double x[16];
#pragma unroll
    for (int i = 15; i >= 0; --i) {
        double s0 = (i == j) ? 1.0 : 0.0, s1 = 0.0;
#pragma unroll
        for (int k = i + 1; k < 16; k += 2) {
            s0 -= D[(off + i) * LB + off + k] * x[k];
            if (k + 1 < 16) s1 -= D[(off + i) * LB + off + k + 1] * x[k + 1];
        }
        x[i] = (s0 + s1) * readlane_f64(dinv, i);
    }
    if (lane < 16) {
#pragma unroll
        for (int i = 0; i < 16; ++i) Uinv[(off + i) * ld + off + j] = x[i];
    }
}

// One DPP step of a wavefront max over 64-bit keys.
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_max_step(unsigned long long k) {
    const int lo = (int)(uint32_t)k, hi = (int)(uint32_t)(k >> 32);
    const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xf, 0xf, false);
    const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xf, 0xf, false);
    const unsigned long long o = ((unsigned long long)(uint32_t)hi2 << 32) | (uint32_t)lo2;
    return o > k ? o : k;
}

// Maximum of a 64-bit key over the wavefront, uniform result: DPP
// butterflies inside each 16-lane row (quad_perm xor 1, xor 2, half mirror,
// mirror), then the four row maxima through readlane.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long k) {
    k = dpp_max_step<0xB1>(k);
    k = dpp_max_step<0x4E>(k);
    k = dpp_max_step<0x141>(k);
    k = dpp_max_step<0x140>(k);
    unsigned long long m = readlane64(k, 0);
    const unsigned long long m1 = readlane64(k, 16), m2 = readlane64(k, 32), m3 = readlane64(k, 48);
    m = m1 > m ? m1 : m;
    m = m2 > m ? m2 : m;
    m = m3 > m ? m3 : m;
    return m;
}

// Pivot key of a candidate row: |v| as double bits (monotone for |v| >= 0)
// with the low 12 mantissa bits replaced by 4095 - row, so the largest key
// is the largest |v|, candidates within 2^-40 relative counting as equal and
// the lowest row winning.  Non-candidates have key 0.
__device__ __forceinline__ unsigned long long pivot_key(double v, int row) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(v));
    return (b & ~0xFFFull) | (unsigned long long)(4095 - row);
}

// 1/x to about one ulp: hardware reciprocal and two Newton steps.
__device__ __forceinline__ double rcp_f64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

template <typename T, int RPT, bool FULL>
__global__ void __launch_bounds__(RT) __attribute__((amdgpu_waves_per_eu(FULL ? 2 : LU_WAVES))) k_refresh_blocked(RefreshArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int job = (int)blockIdx.x;
    if (job >= a.njobs) return;
    if (a.count_dev && job >= 2 * (*a.count_dev - a.walker_off)) return;
    const int w = a.jobs[job >> 1];
    const int spin = job & 1;
    const int ns = a.ns, N = spin ? a.ndn : a.nup, nmax = a.nmax;
    const int nrows = refresh_nrows(ns), ldm = refresh_ldm(nmax);
    // Empty-site rows the single kernel carries: all of them, or, for the
    // real form of complex orbitals (2 ns_out sites, site ns_out + s a copy
    // of site s), those of the first ns_out sites only, the rows W is made
    // of.  The pivots come from the rows of tildeU, so the others are not
    // needed.
    const int nkeep = (FULL && a.cplx_out) ? a.ns_out : ns;
    const int nE = (FULL && a.cplx_out) ? a.ns_out - (spin ? a.ndn_out : a.nup_out) : ns - N;
    const int nr16 = rup_i(N + nE, 16);  // logical rows that carry data
    const int nc16 = rup_i(N, 16);       // columns that carry data
    const int m16 = rup_i(nE, 16);       // empty-site rows, padded
    const double* __restrict__ UT = spin ? a.UTdn : a.UTup;   // ns x N row-major
    const int* __restrict__ kap = (spin ? a.kdn : a.kup) + (size_t)w * ns;
    double* __restrict__ M = a.ws + (size_t)job * refresh_ws_job(ns, nmax);
    double* __restrict__ Lsave = M + (size_t)nrows * ldm;   // [npanel][RB][RB] inverses of L11
    double* __restrict__ Usave = Lsave + (size_t)refresh_npanel(nmax) * RB * RB;   // inverses of U11
    int* __restrict__ jint = reinterpret_cast<int*>(Usave + (size_t)refresh_npanel(nmax) * RB * RB);
    const int t = (int)threadIdx.x, wave = t / 64, lane = t % 64;
    if (!FULL && t == 0) jint[0] = 0;                       // set to 1 once the factors are complete
    // logical rows carrying data: all of M, or only tildeU
    const int nrl = FULL ? nr16 : rup_i(N, 16);

    // ---- LDS carve-up ----------------------------------------------------
    double* cand = reinterpret_cast<double*>(lds);          // [2][NWAVE][RB]
    unsigned long long* candk =
        reinterpret_cast<unsigned long long*>(cand + 2 * NWAVE * RB);   // [2][NWAVE]
    double* Dblk = reinterpret_cast<double*>(candk + 2 * NWAVE);    // [RB][LB] factored diagonal block
    double* Linv = Dblk + RB * LB;                          // [RB][LB] U11^-1, then L11^-1
    double* stage = Linv + RB * LB;                         // W scatter block (single kernel)
    const int stage_len = FULL ? kStoreBlock * (nc16 | 1) : 0;
    double* pivabs = stage + stage_len;                     // [nmax] |pivot|
    double* redd = pivabs + nmax;                           // [3 * NWAVE]
    int* wsum = reinterpret_cast<int*>(redd + 3 * NWAVE);   // [NWAVE]
    int* phys = wsum + NWAVE;                               // [nrows] logical -> physical row
    int* rowsite = phys + nrows;                            // [ns] physical row -> site
    int* rowof = rowsite + ns;                              // [ns] site -> physical row
    int* pinv = rowof + ns;                                 // [nmax] label -> logical row
    __shared__ int s_bad;

    // ---- rows of M from kappa (ballot prefix sum over empty sites) --------
    if (t == 0) s_bad = 0;
    for (int k = t; k < N; k += RT) pinv[k] = 0;
    for (int r = t; r < nrows; r += RT) phys[r] = r;
    __syncthreads();
    int base = 0;
    for (int s0 = 0; s0 < ns; s0 += RT) {
        const int s = s0 + t;
        const int l = s < ns ? kap[s] : -1;
        const bool empty = (l == 0) && s < nkeep;
        const unsigned long long m = __ballot(empty);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int q = 0; q < wave; ++q) off += wsum[q];
        int tot = 0;
        for (int q = 0; q < NWAVE; ++q) tot += wsum[q];
        if (s < ns) {
            if (empty) {
                const int row = N + off + before;
                if (row < N + nE) { rowsite[row] = s; rowof[s] = row; }
                else s_bad = 1;
            } else if (l == 0) {
                // an empty site of the copied half (real form): no row
            } else if (l < 1 || l > N) {
                s_bad = 1;
            } else {
                rowsite[l - 1] = s;
                rowof[s] = l - 1;
                if (atomicAdd(&pinv[l - 1], 1) != 0) s_bad = 1;   // pinv doubles as a label count
            }
        }
        base += tot;
        __syncthreads();
    }
    if (t == 0 && base != nE) s_bad = 1;
    __syncthreads();
    if (s_bad) {
        if (t == 0) { a.status[w] = 2; atomicMin(a.failed, w); }
        return;
    }
    if (!FULL) {
        // columns nc16..ldm-1 of M are read (as zeros) by the solve kernel
        const int nz = ldm - nc16;
        for (int idx = t; idx < nrl * nz; idx += RT) M[(size_t)(idx / nz) * ldm + nc16 + idx % nz] = 0.0;
    }
    // M is never filled from U up front: the first panel (j0 = 0) reads its
    // operands straight from U (L2-resident, shared by every job) and every
    // later read of M finds values written by an earlier step.
    auto u_at = [&](int prow, int col) -> double {
        return (prow < (FULL ? N + nE : N) && col < N) ? UT[(size_t)rowsite[prow] * N + col] : 0.0;
    };

    // ---- 1. blocked LU ---------------------------------------------------
    bool singular = false;
    for (int j0 = 0; j0 < N && !singular; j0 += RB) {
        const int jb = min(RB, N - j0);
        // Registers hold the panel slice of the physical rows myp; myr is the
        // logical row they currently stand for.
        double x[RPT][RB];
        int myr[RPT], myp[RPT];
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int lr = j0 + t + q * RT;
            myr[q] = lr;
            myp[q] = lr < N ? phys[lr] : -1;
#pragma unroll
            for (int c = 0; c < RB; ++c)
                x[q][c] = (myp[q] < 0) ? 0.0 : (j0 == 0 ? u_at(myp[q], c) : M[(size_t)myp[q] * ldm + j0 + c]);
        }
#pragma clang loop unroll(full)
        for (int jj = 0; jj < RB; ++jj) {
            if (jj < jb && !singular) {
                const int kk = j0 + jj;
                const int par = jj & 1;
                unsigned long long key = 0ull;
                int bq = -1;
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    if (myr[q] >= kk && myr[q] < N) {
                        const unsigned long long kq = pivot_key(x[q][jj], myr[q]);
                        if (kq > key) { key = kq; bq = q; }
                    }
                }
                const unsigned long long wkey = wave_max_u64(key);
                // the lane owning the wavefront's best row publishes its
                // entries jj.. (the ones the update reads), 16 bytes a store
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    if (bq == q && key == wkey && wkey != 0ull) {
#pragma unroll
                        for (int c = 0; c < RB; c += 2) {
                            if (c >= (jj & ~1))
                                *reinterpret_cast<double2*>(&cand[(par * NWAVE + wave) * RB + c]) =
                                    make_double2(x[q][c], x[q][c + 1]);
                        }
                    }
                }
                if (lane == 0) candk[par * NWAVE + wave] = wkey;
                __syncthreads();
                // every wavefront's key and candidate pivot in one LDS round
                // (a losing wavefront's entry may be stale; it is never used)
                unsigned long long gk = candk[par * NWAVE];
                double gv = cand[(par * NWAVE) * RB + jj];
                int gw = 0;
#pragma unroll
                for (int q = 1; q < NWAVE; ++q) {
                    const unsigned long long k2 = candk[par * NWAVE + q];
                    const double v2 = cand[(par * NWAVE + q) * RB + jj];
                    if (k2 > gk) { gk = k2; gv = v2; gw = q; }
                }
                const int p = gk ? 4095 - (int)(gk & 0xFFFull) : 0x7fffffff;
                // the pivot row, read by every lane at the same LDS address
                const double* prow = cand + (par * NWAVE + gw) * RB;
                const double piv = (p < N) ? gv : 0.0;
                if (!(fabs(piv) > 0.0) || !isfinite(piv)) {
                    singular = true;
                } else {
                    if (t == 0) {
                        const int tmp = phys[kk]; phys[kk] = phys[p]; phys[p] = tmp;
                        pivabs[kk] = fabs(piv);
                    }
                    const double rp = rcp_f64(piv);
#pragma unroll
                    for (int q = 0; q < RPT; ++q) {
                        const int r = myr[q];
                        myr[q] = (r == p) ? kk : (r == kk ? p : r);
                        if (myr[q] > kk && myp[q] >= 0) {
                            const double l = x[q][jj] * rp;
                            x[q][jj] = l;
                            // constant bounds, so that x stays in registers
                            // once the jj loop is unrolled
#pragma unroll
                            for (int c = 1; c < RB; ++c)
                                if (c > jj) x[q][c] -= l * prow[c];
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (singular) break;
        // panel back to its physical rows; the factored diagonal block to LDS
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            if (myp[q] >= 0) {
#pragma unroll
                for (int c = 0; c < RB; ++c) M[(size_t)myp[q] * ldm + j0 + c] = x[q][c];
            }
            const int i = myr[q] - j0;
            if (myp[q] >= 0 && i >= 0 && i < RB) {
#pragma unroll
                for (int c = 0; c < RB; ++c) Dblk[i * LB + c] = x[q][c];
            }
        }
        for (int i = jb + t; i < RB; i += RT) {
            for (int c = 0; c < RB; ++c) Dblk[i * LB + c] = (c == i) ? 1.0 : 0.0;
        }
        if (!FULL) {
            // padding rows N..nrl-1 (phys[r] = r) enter the trailing update
            // as zero rows; nothing reads their results
            for (int r = N + t; r < nrl; r += RT) {
                for (int c = 0; c < RB; ++c) M[(size_t)r * ldm + j0 + c] = 0.0;
            }
        }
        __syncthreads();
        // U11^-1: kept for the solve kernel (computed by wavefront 0 while
        // wavefront 1 forms L11^-1), or applied here to the empty-site rows,
        // whose panel slice solves x U11 = a  ->  x = a U11^-1
        if constexpr (!FULL) {
            // the 16x16 diagonal blocks of U11^-1 (all the solve kernel reads)
            // and of L11^-1, one per wavefront; then the lower-left block of
            // L11^-1, -Lb^-1 Lc La^-1, on MFMA (Lc: rows 16.., columns ..15 of
            // L11); the upper-right block is zero
            double* Us = Usave + (size_t)(j0 / RB) * RB * RB;
            upper_inverse16(Dblk, Us, RB, 0, 0);
            upper_inverse16(Dblk, Us, RB, 16, 1);
            unit_lower_inverse16(Dblk, Linv, 0, jb, 2);
            unit_lower_inverse16(Dblk, Linv, 16, jb, 3);
            __syncthreads();
            if (wave == 0) {
                f64x4 tq = {0, 0, 0, 0}, xq = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int k = 4 * s + (lane >> 4);
                    tq = mfma16(Dblk[(16 + (lane & 15)) * LB + k], Linv[k * LB + (lane & 15)], tq);
                }
                // the product's accumulator layout is the B-operand layout
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    xq = mfma16(Linv[(16 + (lane & 15)) * LB + 16 + 4 * s + (lane >> 4)], tq[s], xq);
#pragma unroll
                for (int g = 0; g < 4; ++g) Linv[(16 + (lane >> 4) + 4 * g) * LB + (lane & 15)] = -xq[g];
            } else if (wave == 1) {
                for (int idx = lane; idx < 256; idx += 64) Linv[(idx / 16) * LB + 16 + idx % 16] = 0.0;
            }
            __syncthreads();
        } else {
            upper_inverse(Dblk, Linv, LB, 0);
            __syncthreads();
            const int ntr = m16 / 16;
            for (int ti = wave; ti < ntr; ti += NWAVE) {
                const int ra = N + 16 * ti + (lane & 15);
                f64x4 y0 = {0, 0, 0, 0}, y1 = {0, 0, 0, 0};
#pragma unroll
                for (int s2 = 0; s2 < RB / 4; ++s2) {
                    const int k = 4 * s2 + (lane >> 4);
                    const double av = j0 == 0 ? u_at(ra, k) : M[(size_t)ra * ldm + j0 + k];
                    y0 = mfma16(av, Linv[k * LB + (lane & 15)], y0);
                    y1 = mfma16(av, Linv[k * LB + 16 + (lane & 15)], y1);
                }
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const size_t row = (size_t)(N + 16 * ti + (lane >> 4) + 4 * g) * ldm + j0;
                    M[row + (lane & 15)] = y0[g];
                    M[row + 16 + (lane & 15)] = y1[g];
                }
            }
            __syncthreads();
            unit_lower_inverse(Dblk, Linv, jb, 0);
            __syncthreads();
        }
        for (int idx = t; idx < RB * RB; idx += RT)
            Lsave[(size_t)(j0 / RB) * RB * RB + idx] = Linv[(idx / RB) * LB + idx % RB];
        const int c0 = j0 + RB;
        if (c0 >= nc16) continue;
        // U12 = L11^-1 A12 on MFMA: one 16-column block per wavefront at a
        // time.  The single kernel at one or two panel rows per thread (two
        // waves per SIMD) loads the next block's A12 operands while the
        // current one is formed (bofs: the rows' offsets in U in the first
        // panel, -1 for padding rows, in M afterwards); elsewhere that
        // spills, and the operands are loaded per block.
        const int ntc = (nc16 - c0) / 16;
        if constexpr (FULL && RPT <= 2) {
            int bofs[RB / 4];
#pragma unroll
            for (int s = 0; s < RB / 4; ++s) {
                const int pr = phys[j0 + 4 * s + (lane >> 4)];
                bofs[s] = j0 == 0 ? (pr < N + nE ? rowsite[pr] * N : -1) : pr * ldm + c0 + (lane & 15);
            }
            auto bval = [&](int s, int tj) -> double {
                const int col = c0 + 16 * tj + (lane & 15);
                if (j0 == 0) return (bofs[s] >= 0 && col < N) ? UT[(size_t)bofs[s] + col] : 0.0;
                return M[(size_t)bofs[s] + 16 * tj];
            };
            double bcur[RB / 4], bnext[RB / 4];
            if (wave < ntc) {
#pragma unroll
                for (int s = 0; s < RB / 4; ++s) bcur[s] = bval(s, wave);
            }
            for (int tj = wave; tj < ntc; tj += NWAVE) {
                const int cc = 16 * tj + (lane & 15);
                if (tj + NWAVE < ntc) {
#pragma unroll
                    for (int s = 0; s < RB / 4; ++s) bnext[s] = bval(s, tj + NWAVE);
                }
                f64x4 u0 = {0, 0, 0, 0}, u1 = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < RB / 4; ++s) {
                    const int k = 4 * s + (lane >> 4);
                    u0 = mfma16(Linv[(lane & 15) * LB + k], bcur[s], u0);
                    u1 = mfma16(Linv[(16 + (lane & 15)) * LB + k], bcur[s], u1);
                }
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int r0 = (lane >> 4) + 4 * g, r1 = r0 + 16;
                    M[(size_t)phys[j0 + r0] * ldm + c0 + cc] = u0[g];
                    M[(size_t)phys[j0 + r1] * ldm + c0 + cc] = u1[g];
                }
#pragma unroll
                for (int s = 0; s < RB / 4; ++s) bcur[s] = bnext[s];
            }
        } else {
            for (int tj = wave; tj < ntc; tj += NWAVE) {
                const int cc = 16 * tj + (lane & 15);
                f64x4 u0 = {0, 0, 0, 0}, u1 = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < RB / 4; ++s) {
                    const int k = 4 * s + (lane >> 4);
                    const double b = j0 == 0 ? u_at(phys[k], c0 + cc) : M[(size_t)phys[j0 + k] * ldm + c0 + cc];
                    u0 = mfma16(Linv[(lane & 15) * LB + k], b, u0);
                    u1 = mfma16(Linv[(16 + (lane & 15)) * LB + k], b, u1);
                }
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int r0 = (lane >> 4) + 4 * g, r1 = r0 + 16;
                    M[(size_t)phys[j0 + r0] * ldm + c0 + cc] = u0[g];
                    M[(size_t)phys[j0 + r1] * ldm + c0 + cc] = u1[g];
                }
            }
        }
        __syncthreads();
        // trailing update: M[r][c] -= L21[r][:] U12[:][c].  A wavefront
        // updates TG row tiles at a time: each load of the U12 operands
        // (read back from M through L2; staged in LDS they took 102 KB at
        // N = 432, one workgroup per CU) feeds TG MFMA chains, and the next
        // column tile's operands are loaded while the current one is
        // updated.  TG = 2 needs about 200 VGPRs: it fits the single kernel at
        // two panel rows per thread (two waves per SIMD, under the panel's
        // register peak) and spills elsewhere.  The C operands come from U in
        // the first panel and from M afterwards, a loop for each, so that
        // only one set of row offsets is live.
        constexpr int TG = (FULL && RPT == 2) ? 2 : 1;
        const int r0 = c0;
        const int ntr = (nrl - r0) / 16;
        int uofs[RB / 4];   // U12 operand rows: offsets in M
#pragma unroll
        for (int s = 0; s < RB / 4; ++s) uofs[s] = phys[j0 + 4 * s + (lane >> 4)] * ldm + c0 + (lane & 15);
        auto tiles = [&](auto first_panel) {
            constexpr bool FIRST = decltype(first_panel)::value;
            for (int tb = wave; tb < ntr; tb += TG * NWAVE) {
                double af[TG][RB / 4];
                int mofs[TG][4];   // C rows: offsets in M
                int cofs[TG][4];   // first panel: the rows' offsets in U, -1 for padding rows
#pragma unroll
                for (int k = 0; k < TG; ++k) {
                    // a tile past ntr repeats the last one and is never stored
                    const int ti = min(tb + k * NWAVE, ntr - 1);
                    const int pa = phys[r0 + 16 * ti + (lane & 15)];
#pragma unroll
                    for (int s = 0; s < RB / 4; ++s) af[k][s] = -M[(size_t)pa * ldm + j0 + 4 * s + (lane >> 4)];
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const int pr = phys[r0 + 16 * ti + (lane >> 4) + 4 * g];
                        mofs[k][g] = pr * ldm + c0 + (lane & 15);
                        if constexpr (FIRST) cofs[k][g] = pr < (FULL ? N + nE : N) ? rowsite[pr] * N : -1;
                    }
                }
                auto cval = [&](int k, int g, int tj) -> double {
                    if constexpr (FIRST) {
                        const int col = c0 + 16 * tj + (lane & 15);
                        return (cofs[k][g] >= 0 && col < N) ? UT[(size_t)cofs[k][g] + col] : 0.0;
                    } else {
                        return M[(size_t)mofs[k][g] + 16 * tj];
                    }
                };
                f64x4 acc[TG], nxt[TG];
                double ub[RB / 4], un[RB / 4];
#pragma unroll
                for (int k = 0; k < TG; ++k) {
#pragma unroll
                    for (int g = 0; g < 4; ++g) acc[k][g] = cval(k, g, 0);
                }
#pragma unroll
                for (int s = 0; s < RB / 4; ++s) ub[s] = M[uofs[s]];
                for (int tj = 0; tj < ntc; ++tj) {
                    if (tj + 1 < ntc) {
#pragma unroll
                        for (int k = 0; k < TG; ++k) {
#pragma unroll
                            for (int g = 0; g < 4; ++g) nxt[k][g] = cval(k, g, tj + 1);
                        }
#pragma unroll
                        for (int s = 0; s < RB / 4; ++s) un[s] = M[uofs[s] + 16 * (tj + 1)];
                    }
#pragma unroll
                    for (int s = 0; s < RB / 4; ++s) {
#pragma unroll
                        for (int k = 0; k < TG; ++k) acc[k] = mfma16(af[k][s], ub[s], acc[k]);
                    }
#pragma unroll
                    for (int k = 0; k < TG; ++k) {
                        if (tb + k * NWAVE < ntr) {
#pragma unroll
                            for (int g = 0; g < 4; ++g) M[(size_t)mofs[k][g] + 16 * tj] = acc[k][g];
                        }
                    }
#pragma unroll
                    for (int k = 0; k < TG; ++k) acc[k] = nxt[k];
#pragma unroll
                    for (int s = 0; s < RB / 4; ++s) ub[s] = un[s];
                }
            }
        };
        if (j0 == 0) tiles(std::true_type{});
        else tiles(std::false_type{});
        __syncthreads();
    }
    double logdet = 0.0, pmin = INFINITY, pmax = 0.0;
    if (!singular) {
        double ls = 0.0, mn = INFINITY, mx = 0.0;
        for (int k = t; k < N; k += RT) {
            const double v = pivabs[k];
            ls += log(v);
            mn = fmin(mn, v);
            mx = fmax(mx, v);
        }
        for (int o = 32; o > 0; o >>= 1) {
            ls += __shfl_xor(ls, o, 64);
            mn = fmin(mn, __shfl_xor(mn, o, 64));
            mx = fmax(mx, __shfl_xor(mx, o, 64));
        }
        if (lane == 0) { redd[wave] = ls; redd[NWAVE + wave] = mn; redd[2 * NWAVE + wave] = mx; }
        __syncthreads();
        for (int q = 0; q < NWAVE; ++q) {
            logdet += redd[q];
            pmin = fmin(pmin, redd[NWAVE + q]);
            pmax = fmax(pmax, redd[2 * NWAVE + q]);
        }
    }
    if (!singular && pmin < a.pivot_tol * pmax) singular = true;
    if (singular) {
        if (t == 0) { a.status[w] = 1; atomicMin(a.failed, w); }
        return;
    }
    if (t == 0) {
        a.logdet[(size_t)w * 2 + spin] = logdet;
        if (a.write_w && spin == 0) { a.npend[w] = 0; a.acc_since[w] = 0; }
    }
    if constexpr (!FULL) {
        // logical row c < N holds label phys[c]; W is formed by k_refresh_solve
        for (int c = t; c < N; c += RT) jint[1 + c] = phys[c];
        if (t == 0) jint[0] = 1;
        return;
    }
    // logical row c < N holds label phys[c]
    for (int c = t; c < N; c += RT) pinv[phys[c]] = c;

    // ---- 2. X L1 = L2 for the empty-site rows (physical = logical rows >= N)
    const int npanel = (N + RB - 1) / RB;
    for (int pq = npanel - 1; pq >= 0; --pq) {
        const int j0 = pq * RB, jb = min(RB, N - j0);
        for (int idx = t; idx < RB * RB; idx += RT)
            Linv[(idx / RB) * LB + idx % RB] = Lsave[(size_t)pq * RB * RB + idx];
        __syncthreads();
        // X_J = B_J L1_JJ^-1, one 16-row tile per wavefront (both column
        // tiles), the next tile's B operands loaded while this one is formed
        const int ntr = m16 / 16;
        {
            double bcur[RB / 4], bnext[RB / 4];
            auto bload = [&](int ti, double (&v)[RB / 4]) {
                const int ra = N + 16 * ti + (lane & 15);
#pragma unroll
                for (int s = 0; s < RB / 4; ++s) v[s] = M[(size_t)ra * ldm + j0 + 4 * s + (lane >> 4)];
            };
            if (wave < ntr) bload(wave, bcur);
            for (int ti = wave; ti < ntr; ti += NWAVE) {
                if (ti + NWAVE < ntr) bload(ti + NWAVE, bnext);
                f64x4 x0 = {0, 0, 0, 0}, x1 = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < RB / 4; ++s) {
                    const int k = 4 * s + (lane >> 4);
                    x0 = mfma16(bcur[s], Linv[k * LB + (lane & 15)], x0);
                    x1 = mfma16(bcur[s], Linv[k * LB + 16 + (lane & 15)], x1);
                }
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int rr = 16 * ti + (lane >> 4) + 4 * g;
                    M[(size_t)(N + rr) * ldm + j0 + (lane & 15)] = x0[g];
                    M[(size_t)(N + rr) * ldm + j0 + 16 + (lane & 15)] = x1[g];
                }
#pragma unroll
                for (int s = 0; s < RB / 4; ++s) bcur[s] = bnext[s];
            }
        }
        __syncthreads();
        // M[N.., 0:j0] -= X_J L1[J, 0:j0]: a wavefront takes two row tiles at
        // a time, which share each load of the L1 operand rows, and loads the
        // next column tile's operands while the current one is updated
        constexpr int TS = 2;
        const int ntc = j0 / 16;
        int pb[RB / 4];
#pragma unroll
        for (int s = 0; s < RB / 4; ++s) {
            const int k = 4 * s + (lane >> 4);
            pb[s] = k < jb ? phys[j0 + k] : -1;
        }
        for (int tb = wave; tb < ntr; tb += TS * NWAVE) {
            double af[TS][RB / 4];
            int cofs[TS];   // offset in M of the tile row this lane updates first
#pragma unroll
            for (int q = 0; q < TS; ++q) {
                // a tile past ntr repeats the last one and is never stored
                const int ti = min(tb + q * NWAVE, ntr - 1);
#pragma unroll
                for (int s = 0; s < RB / 4; ++s)
                    af[q][s] = -M[(size_t)(N + 16 * ti + (lane & 15)) * ldm + j0 + 4 * s + (lane >> 4)];
                cofs[q] = (N + 16 * ti + (lane >> 4)) * ldm + (lane & 15);
            }
            f64x4 acc[TS], nxt[TS];
            double bf[RB / 4], bn[RB / 4];
            if (ntc > 0) {
#pragma unroll
                for (int q = 0; q < TS; ++q) {
#pragma unroll
                    for (int g = 0; g < 4; ++g) acc[q][g] = M[(size_t)cofs[q] + (size_t)4 * g * ldm];
                }
#pragma unroll
                for (int s = 0; s < RB / 4; ++s) bf[s] = pb[s] >= 0 ? M[(size_t)pb[s] * ldm + (lane & 15)] : 0.0;
            }
            for (int tj = 0; tj < ntc; ++tj) {
                if (tj + 1 < ntc) {
                    const int cn = 16 * (tj + 1);
#pragma unroll
                    for (int q = 0; q < TS; ++q) {
#pragma unroll
                        for (int g = 0; g < 4; ++g) nxt[q][g] = M[(size_t)cofs[q] + (size_t)4 * g * ldm + cn];
                    }
#pragma unroll
                    for (int s = 0; s < RB / 4; ++s)
                        bn[s] = pb[s] >= 0 ? M[(size_t)pb[s] * ldm + cn + (lane & 15)] : 0.0;
                }
#pragma unroll
                for (int s = 0; s < RB / 4; ++s) {
#pragma unroll
                    for (int q = 0; q < TS; ++q) acc[q] = mfma16(af[q][s], bf[s], acc[q]);
                }
#pragma unroll
                for (int q = 0; q < TS; ++q) {
                    if (tb + q * NWAVE < ntr) {
#pragma unroll
                        for (int g = 0; g < 4; ++g) M[(size_t)cofs[q] + (size_t)4 * g * ldm + 16 * tj] = acc[q][g];
                    }
                }
#pragma unroll
                for (int q = 0; q < TS; ++q) acc[q] = nxt[q];
#pragma unroll
                for (int s = 0; s < RB / 4; ++s) bf[s] = bn[s];
            }
        }
        __syncthreads();
    }

    // ---- 3. W from X, label column by label column ------------------------
    // Empty site s takes row rowof[s] - N (site order); occupied rows are unit
    // rows and are not stored.  Drift mode reads the current rows through the
    // current slot map.  Complex orbitals: only the first ns_out sites are
    // stored; label l < N_out gives Re W[., l], label N_out + l gives -Im.
    const int nso = a.ns_out, NO = spin ? a.ndn_out : a.nup_out, cw = a.cplx_out ? 2 : 1;
    T* __restrict__ Wo = reinterpret_cast<T*>(a.W) +
                         ((size_t)w * w_stride(nso, a.nup_out, a.ndn_out) + (spin ? w_off_dn(nso, a.nup_out) : 0)) * cw;
    const int ld = w_ld(nso, NO);
    int* __restrict__ sl = a.slot + ((size_t)w * 2 + spin) * nso;
    // Blocks of SB consecutive sites: the solved rows of the block's empty
    // sites are staged in LDS (contiguous row reads, odd row stride against
    // bank conflicts), then every label column of W gets the block's entries
    // written as one contiguous run.
    constexpr int SB = kStoreBlock;
    const int xs = (nc16 | 1);                       // LDS row stride (odd)
    double* xb = stage;                              // [SB][xs]
    const int nsa = w_nsa(nso, a.nup_out, a.ndn_out);
    const int cpend = a.write_w ? 0 : a.npend[w];
    const int dp = pend_ld(a.delay);
    const T* gA = reinterpret_cast<const T*>(a.A) + ((size_t)w * 2 + spin) * nsa * dp * cw;
    const T* gB = reinterpret_cast<const T*>(a.Bm) + ((size_t)w * 2 + spin) * a.nmax_out * dp * cw;
    double dmax = 0.0, vmax = N > 0 ? 1.0 : 0.0;   // occupied rows: unit rows, exact
    // Site s < nso is empty exactly when its row is one of X's (rowof[s] >=
    // N).  A wavefront holds rows wave and wave + 4 of a block in registers
    // (N <= RPT * RT columns): the next block's are loaded while the current
    // block is written.
    constexpr int RPW = (SB + NWAVE - 1) / NWAVE, CPL = RPT * RT / 64;
    double pre[RPW][CPL];
    auto load_rows = [&](int sb) {
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int i = wave + r * NWAVE, s = sb + i;
            const bool ok = i < SB && s < nso && rowof[s] >= N;
            const double* __restrict__ src = M + (size_t)(ok ? rowof[s] : 0) * ldm;
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const int col = lane + 64 * c;
                pre[r][c] = (ok && col < N) ? src[col] : 0.0;
            }
        }
    };
    load_rows(0);
    for (int sb = 0; sb < nso; sb += SB) {
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int i = wave + r * NWAVE, s = sb + i;
            if (i < SB && s < nso && rowof[s] >= N) {
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const int col = lane + 64 * c;
                    if (col < N) xb[i * xs + col] = pre[r][c];
                }
            }
        }
        __syncthreads();
        if (sb + SB < nso) load_rows(sb + SB);
        for (int idx = t; idx < N * SB; idx += RT) {
            const int l = idx / SB, i = idx % SB, s = sb + i;
            if (s >= nso || rowof[s] < N) continue;
            const int part = l >= NO ? 1 : 0, lo = l - part * NO;
            const double v = part ? -xb[i * xs + pinv[l]] : xb[i * xs + pinv[l]];
            const int row = a.write_w ? rowof[s] - N : max(sl[s], 0);
            const size_t at = (size_t)lo * ld + row;
            if (a.write_w) {
                Wo[at * cw + part] = (T)v;
            } else {
                double cur = (double)Wo[at * cw + part];
                for (int c = 0; c < cpend; ++c) {
                    const size_t ia = ((size_t)row * dp + c) * cw, ib = ((size_t)lo * dp + c) * cw;
                    if (cw == 1) {
                        cur += (double)gA[ia] * (double)gB[ib];
                    } else {   // component `part` of the complex product
                        const double ar = gA[ia], ai = gA[ia + 1], br = gB[ib], bi = gB[ib + 1];
                        cur += part ? ar * bi + ai * br : ar * br - ai * bi;
                    }
                }
                dmax = fmax(dmax, fabs(v - cur));
                vmax = fmax(vmax, fabs(v));
            }
        }
        __syncthreads();
    }
    if (a.write_w) {
        for (int s = t; s < nso; s += RT) sl[s] = rowof[s] >= N ? rowof[s] - N : -1;
    } else {
        atomic_max_nonneg2(&a.drift[(size_t)w * 2], dmax);
        atomic_max_nonneg2(&a.drift[(size_t)w * 2 + 1], vmax);
    }
}

typedef __attribute__((address_space(3))) void lds_void;

// Stages the 16x16 tiles K in [k0, k1) of row block J of the factors
// (logical rows 16J..16J+15 of M) into buf, tile-major (tile K at buf +
// 256 K, row-major inside), and the 16x16 diagonal block of the saved 32x32
// inverse `inv` into dbuf, with direct global-to-LDS loads
// (global_load_lds_dwordx4: lane l's 16 bytes land at dst + 16 l).  A
// 1 KiB chunk is half a tile: lane l reads columns 2 (l % 8) and the next
// of tile row 8 h + l / 8.  Only the tiles a step's update reads are
// staged: K > J in the first solve, K < J in the second.  Chunk c goes to
// wavefront c % 4, so a wavefront always loads the same half h = wave % 2
// of the tiles, from one row of M.  The two chunks of the diagonal inverse
// follow.  The loads are tracked by vmcnt, so the barrier that ends a
// block's step also completes the next block's staging.
__device__ __forceinline__ void solve_stage(const double* M, int ldm, const int* lab, const double* inv, int J,
                                            int k0, int k1, double* buf, double* dbuf) {
    const int wave = (int)threadIdx.x / 64, lane = (int)threadIdx.x % 64;
    const int nch = k1 > k0 ? 2 * (k1 - k0) : 0;
    const int h = wave & 1;
    const double* row = M + (size_t)lab[16 * J + 8 * h + (lane >> 3)] * ldm + 2 * (lane & 7);
    for (int c = wave; c < nch; c += NWAVE) {
        const int K = k0 + (c >> 1);
        __builtin_amdgcn_global_load_lds(row + 16 * K, (lds_void*)(buf + K * 256 + h * 128), 16, 0, 0);
    }
    const int d = (wave - nch) & (NWAVE - 1);   // chunk nch + d of this wavefront
    if (d < 2) {
        const int o = d * 128 + lane * 2;
        const int hb = (J & 1) * 16;
        const double* src = inv + (size_t)(J >> 1) * RB * RB + (hb + o / 16) * RB + hb + o % 16;
        __builtin_amdgcn_global_load_lds(src, (lds_void*)(dbuf + d * 128), 16, 0, 0);
    }
}

// The empty-site rows of W from the factors of k_refresh_blocked<FULL =
// false>: W[s][:] = U[s][:] tildeU^-1 = ((U[s][:] U1^-1) L1^-1) P.
//
// Each wavefront takes 16 empty sites and keeps the transposed rows as
// 16x16 tiles T_J (labels 16J.. of the 16 sites) in MFMA accumulator layout:
// lane l, register g holds T_J[l / 16 + 4 g][l % 16].  That layout is also
// the B-operand layout of the four k-slices of v_mfma_f64_16x16x4f64
// (slice s = register s), so a solved tile feeds the next products without
// leaving registers:
//   y U1 = b  <=>  U1^T y^T = b^T, J ascending:
//     T_J = (U1_JJ^-1)^T T_J,  then T_K -= U1[J,K]^T T_J for K > J;
//   x L1 = y  <=>  L1^T x^T = y^T, J descending:
//     T_J = (L1_JJ^-1)^T T_J,  then T_K -= L1[J,K]^T T_J for K < J.
// The tiles of row block J that the update reads and the 16x16 diagonal
// inverse are staged in LDS, double buffered: the next block streams in
// while the current one is used, one barrier per block.  Step q of the 2 nt steps
// reads buffer q & 1; q = J in the first solve and 2 nt - 1 - J in the
// second, so the buffer is (J + 1) & 1 there, known at compile time.
//
// Grid: 8 * groups * ceil(njobs / 8) workgroups of 256 threads; a
// workgroup takes 64 consecutive empty sites (in site order) of one job,
// and the groups of one job share blockIdx % 8, i.e. one XCD and its L2.
// It also writes the unit rows of the occupied sites between its first
// empty site and the next group's, so every entry of W is written once.
template <typename T, int NT>
__global__ void __launch_bounds__(RT) __attribute__((amdgpu_waves_per_eu(2)))
k_refresh_solve(RefreshArgs a, int groups) {
    __shared__ __align__(16) double sb0[NT * 256], sb1[NT * 256];   // tile-major row blocks
    __shared__ __align__(16) double sd0[256], sd1[256];
    extern __shared__ __align__(16) unsigned char lds[];
    const int b = (int)blockIdx.x;
    const int job = ((b >> 3) / groups) * 8 + (b & 7);
    const int grp = (b >> 3) % groups;
    if (job >= a.njobs) return;
    const int spin = job & 1;
    const int ns = a.ns, N = spin ? a.ndn : a.nup, nmax = a.nmax;
    const int nE = ns - N;
    // rows stored: complex orbitals keep the first ns_out sites only
    const int nso = a.ns_out, NO = spin ? a.ndn_out : a.nup_out, cw = a.cplx_out ? 2 : 1;
    const int nEo = nso - NO;
    if (grp > 0 && 64 * grp >= nEo) return;
    const int nt = (N + 15) / 16;
    const double* __restrict__ M = a.ws + (size_t)job * refresh_ws_job(ns, nmax);
    const int ldm = refresh_ldm(nmax);
    const double* __restrict__ Lsave = M + (size_t)refresh_nrows(ns) * ldm;
    const double* __restrict__ Usave = Lsave + (size_t)refresh_npanel(nmax) * RB * RB;
    const int* __restrict__ jint = reinterpret_cast<const int*>(Usave + (size_t)refresh_npanel(nmax) * RB * RB);
    // the job's count, walker and factorisation flag in one memory round
    // trip (every address is in bounds for job < njobs)
    const int cnt = a.count_dev ? *a.count_dev : 0;
    const int w = a.jobs[job >> 1];
    const int fact = jint[0];
    if (a.count_dev && job >= 2 * (cnt - a.walker_off)) return;
    if (fact != 1) return;                        // factorisation rejected (status already set)
    const int* __restrict__ kap = (spin ? a.kdn : a.kup) + (size_t)w * ns;
    // ns x N row-major: a lane's four labels 16J + kr + 4g of one site lie in
    // one 128-byte line, shared by the kr lanes of that site
    const double* __restrict__ UT = spin ? a.UTdn : a.UTup;
    const int t = (int)threadIdx.x, wave = t / 64, lane = t % 64;

    int* lab = reinterpret_cast<int*>(lds);       // [NT*16] logical row -> physical row of M
    int* wsum = lab + NT * 16;                    // [NWAVE]
    int* esite = wsum + NWAVE;                    // [ns] empty sites in order

    // rows N..16 nt - 1 are the zeroed padding rows of M; the labels and
    // the first sites' kappa are loaded together, and each pass of the
    // ballot loads the next pass's kappa before its barriers
    int labv[(NT * 16 + RT - 1) / RT];
#pragma unroll
    for (int i = 0; i < (NT * 16 + RT - 1) / RT; ++i) {
        const int c = t + i * RT;
        labv[i] = c < N ? jint[1 + c] : c;
    }
    int lnext = t < ns ? kap[t] : -1;
#pragma unroll
    for (int i = 0; i < (NT * 16 + RT - 1) / RT; ++i) {
        const int c = t + i * RT;
        if (c < NT * 16) lab[c] = labv[i];
    }
    int base = 0;
    for (int s0 = 0; s0 < ns; s0 += RT) {
        const int s = s0 + t;
        const int l = lnext;
        lnext = s + RT < ns ? kap[s + RT] : -1;
        const bool empty = s < ns && l == 0;
        const unsigned long long m = __ballot(empty);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        int off = base, tot = 0;
        for (int q = 0; q < NWAVE; ++q) {
            off += q < wave ? wsum[q] : 0;
            tot += wsum[q];
        }
        if (empty && off + before < nE) esite[off + before] = s;   // count checked by the LU kernel
        base += tot;
        __syncthreads();
    }

    // the sites whose slot map this workgroup rewrites (write mode): their
    // first pass of kappa is loaded now, with the tiles
    int s_lo = 0, s_hi = 0, kfirst = 1;
    if (a.write_w) {
        s_lo = grp == 0 ? 0 : esite[64 * grp];
        s_hi = 64 * (grp + 1) < nEo ? esite[64 * (grp + 1)] : nso;
        kfirst = s_lo + t < s_hi ? kap[s_lo + t] : 1;
    }
    const int e = 64 * grp + 16 * wave + (lane & 15);
    const int site = e < nEo ? esite[e] : -1;
    const bool active = 64 * grp + 16 * wave < nEo;   // wavefront holds at least one site
    const int kr = lane >> 4, kc = lane & 15;

    // The tiles hold the accumulators negated: T_K = -acc_K.  The diagonal
    // step then reads the inverse negated (4 values per step) and the
    // updates T_K += F[J,K]^T y_J need no sign change per block.
    f64x4 tile[NT];
#pragma unroll
    for (int J = 0; J < NT; ++J) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int c = 16 * J + kr + 4 * g;
            tile[J][g] = (site >= 0 && c < N) ? -UT[(size_t)site * N + c] : 0.0;
        }
    }
    // A operands of tile K of the staged row block (64 consecutive doubles
    // per k-slice: no bank conflicts)
    auto lds_a = [&](const double* bb, int K, double (&v)[4]) {
#pragma unroll
        for (int s = 0; s < 4; ++s) v[s] = bb[K * 256 + (4 * s + kr) * 16 + kc];
    };
    // T_J = y = (-inv)^T T_J from the staged diagonal inverse
    auto diag = [&](const double* db, int J) -> f64x4 {
        double d[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) d[s] = -db[(4 * s + kr) * 16 + kc];
        f64x4 y = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) y = mfma16(d[s], tile[J][s], y);
        return y;
    };
    // T_K += F[J,K]^T y for K in [k0, k1), two blocks at a time so that
    // consecutive products are independent; operands one pair ahead
    auto update = [&](const double* bb, int k0, int k1, const f64x4& y) {
        double oa[2][4], ob[2][4];
        if (k0 < k1) lds_a(bb, k0, oa[0]);
        if (k0 + 1 < k1) lds_a(bb, k0 + 1, ob[0]);
#pragma unroll
        for (int P = 0; k0 + 2 * P < k1; ++P) {
            const int K = k0 + 2 * P, pp = P & 1;
            if (K + 2 < k1) lds_a(bb, K + 2, oa[pp ^ 1]);
            if (K + 3 < k1) lds_a(bb, K + 3, ob[pp ^ 1]);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                tile[K] = mfma16(oa[pp][s], y[s], tile[K]);
                if (K + 1 < k1) tile[K + 1] = mfma16(ob[pp][s], y[s], tile[K + 1]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };

    solve_stage(M, ldm, lab, Usave, 0, 1, NT, sb0, sd0);
    __syncthreads();
    // y U1 = b, block rows ascending (step J reads buffer J & 1).  Blocks
    // K >= nt are zero padding, so the updates run to NT without guards.
#pragma unroll
    for (int J = 0; J < NT; ++J) {
        if (J < nt) {
            double* nb = (J & 1) ? sb0 : sb1;
            double* nd = (J & 1) ? sd0 : sd1;
            if (J + 1 < nt) solve_stage(M, ldm, lab, Usave, J + 1, J + 2, NT, nb, nd);
            else solve_stage(M, ldm, lab, Lsave, nt - 1, 0, nt - 1, nb, nd);
            const double* bb = (J & 1) ? sb1 : sb0;
            const double* db = (J & 1) ? sd1 : sd0;
            if (active) {
                const f64x4 y = diag(db, J);
                tile[J] = y;
                update(bb, J + 1, NT, y);
            }
            __syncthreads();
        }
    }
    // the second solve starts from T_K = -y_K
#pragma unroll
    for (int J = 0; J < NT; ++J) tile[J] = -tile[J];
    // x L1 = y, block rows descending (step reads buffer (J + 1) & 1)
#pragma unroll
    for (int J = NT - 1; J >= 0; --J) {
        if (J < nt) {
            if (J > 0) solve_stage(M, ldm, lab, Lsave, J - 1, 0, J - 1, (J & 1) ? sb1 : sb0, (J & 1) ? sd1 : sd0);
            const double* bb = (J & 1) ? sb0 : sb1;
            const double* db = (J & 1) ? sd0 : sd1;
            if (active) {
                const f64x4 x = diag(db, J);
                tile[J] = x;
                update(bb, 0, J, x);
            }
            __syncthreads();
        }
    }

    // ---- W: the solved tiles, one row per empty site; the slot map --------
    // Write mode: the empty sites take rows 0..nE-1 in site order and the
    // map is rewritten for this workgroup's span of sites.  Drift mode reads
    // the current rows through the current map.
    T* __restrict__ Wo = reinterpret_cast<T*>(a.W) +
                         ((size_t)w * w_stride(nso, a.nup_out, a.ndn_out) + (spin ? w_off_dn(nso, a.nup_out) : 0)) * cw;
    const int ld = w_ld(nso, NO);
    int* __restrict__ sl = a.slot + ((size_t)w * 2 + spin) * nso;
    const int nsa = w_nsa(nso, a.nup_out, a.ndn_out);
    const int cpend = a.write_w ? 0 : a.npend[w];
    const int dp = pend_ld(a.delay);
    const T* gA = reinterpret_cast<const T*>(a.A) + ((size_t)w * 2 + spin) * nsa * dp * cw;
    const T* gB = reinterpret_cast<const T*>(a.Bm) + ((size_t)w * 2 + spin) * a.nmax_out * dp * cw;
    double dmax = 0.0, vmax = N > 0 ? 1.0 : 0.0;   // occupied rows: unit rows, exact
    if (site >= 0) {
        const int row = a.write_w ? e : max(sl[site], 0);
#pragma unroll
        for (int J = 0; J < NT; ++J) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int c = 16 * J + kr + 4 * g;
                if (c < N) {
                    // complex orbitals: label l < N_out is Re W[., l], N_out + l is -Im
                    const int l = lab[c];
                    const int part = l >= NO ? 1 : 0, lo = l - part * NO;
                    const double v = part ? -tile[J][g] : tile[J][g];
                    const size_t at = (size_t)lo * ld + row;
                    if (a.write_w) {
                        Wo[at * cw + part] = (T)v;
                    } else {
                        double cur = (double)Wo[at * cw + part];
                        for (int q = 0; q < cpend; ++q) {
                            const size_t ia = ((size_t)row * dp + q) * cw, ib = ((size_t)lo * dp + q) * cw;
                            if (cw == 1) {
                                cur += (double)gA[ia] * (double)gB[ib];
                            } else {
                                const double ar = gA[ia], ai = gA[ia + 1], br = gB[ib], bi = gB[ib + 1];
                                cur += part ? ar * bi + ai * br : ar * br - ai * bi;
                            }
                        }
                        dmax = fmax(dmax, fabs(v - cur));
                        vmax = fmax(vmax, fabs(v));
                    }
                }
            }
        }
    }
    if (a.write_w) {
        if (s_lo + t < s_hi && kfirst != 0) sl[s_lo + t] = -1;
        for (int s = s_lo + t + RT; s < s_hi; s += RT)
            if (kap[s] != 0) sl[s] = -1;
        if (t < 64 && 64 * grp + t < nEo) sl[esite[64 * grp + t]] = 64 * grp + t;
    } else {
        atomic_max_nonneg2(&a.drift[(size_t)w * 2], dmax);
        atomic_max_nonneg2(&a.drift[(size_t)w * 2 + 1], vmax);
    }
}

// LDS of k_refresh_solve: static row-block buffers plus the dynamic part.
static size_t solve_static_lds(int nt) { return (size_t)(2 * nt * 256 + 2 * 256) * sizeof(double); }
static size_t solve_lds_bytes(int ns, int nt) { return (size_t)(nt * 16 + NWAVE + ns) * sizeof(int); }

size_t refresh_ws_doubles_per_job(int ns, int nmax) { return refresh_ws_job(ns, nmax); }

static size_t refresh_lds_bytes(int ns, int nup, int ndn, bool full = true) {
    const int nmax = nup > ndn ? nup : ndn;
    const int stage = !full ? 0 : kStoreBlock * (rup_i(nmax, 16) | 1);
    const size_t dbl = (size_t)(2 * NWAVE * RB + 2 * NWAVE + 2 * RB * LB + stage + nmax + 3 * NWAVE) * sizeof(double);
    const size_t ints = (size_t)(NWAVE + refresh_nrows(ns) + 2 * ns + nmax) * sizeof(int);
    return dbl + ints + 16;
}

static int refresh_rows_per_thread(int nmax) {
    return (nmax + RT - 1) / RT;
}

static hipError_t set_lds_limit(const void* fn, size_t lds) {
    if (lds <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

template <typename T, int RPT, bool FULL>
static hipError_t launch_lu(const RefreshArgs& a, size_t lds, hipStream_t s) {
    hipError_t e = set_lds_limit(reinterpret_cast<const void*>(&k_refresh_blocked<T, RPT, FULL>), lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_refresh_blocked<T, RPT, FULL>), dim3(a.njobs), dim3(RT), lds, s, a);
    return hipGetLastError();
}

template <typename T, int NT>
static hipError_t launch_solve(const RefreshArgs& a, hipStream_t s) {
    // workgroups of 64 stored empty-site rows per job
    const int nmin = a.nup_out < a.ndn_out ? a.nup_out : a.ndn_out;
    const int groups = (a.ns_out - nmin + 63) / 64 > 1 ? (a.ns_out - nmin + 63) / 64 : 1;
    const size_t lds = solve_lds_bytes(a.ns, NT);
    if (solve_static_lds(NT) + lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_refresh_solve<T, NT>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int grid = 8 * groups * ((a.njobs + 7) / 8);
    hipLaunchKernelGGL((k_refresh_solve<T, NT>), dim3(grid), dim3(RT), lds, s, a, groups);
    return hipGetLastError();
}

// Register tiles per wavefront of the solve kernel, or 0 for the
// single-kernel path (N > 256, or LDS beyond 160 KiB).
static int solve_tiles(int ns, int nmax) {
    const int NT = solve_nt(nmax);
    if (NT == 0 || solve_static_lds(NT) + solve_lds_bytes(ns, NT) > 160 * 1024) return 0;
    return NT;
}

template <typename T>
static hipError_t launch_refresh_t(const RefreshArgs& a, size_t lds, hipStream_t s) {
    switch (solve_tiles(a.ns, a.nmax)) {
        case 4: {
            hipError_t e = launch_lu<T, 1, false>(a, refresh_lds_bytes(a.ns, a.nup, a.ndn, false), s);
            return e != hipSuccess ? e : launch_solve<T, 4>(a, s);
        }
        case 8: {
            hipError_t e = launch_lu<T, 1, false>(a, refresh_lds_bytes(a.ns, a.nup, a.ndn, false), s);
            return e != hipSuccess ? e : launch_solve<T, 8>(a, s);
        }
        case 14: {
            hipError_t e = launch_lu<T, 1, false>(a, refresh_lds_bytes(a.ns, a.nup, a.ndn, false), s);
            return e != hipSuccess ? e : launch_solve<T, 14>(a, s);
        }
        case 16: {
            hipError_t e = launch_lu<T, 1, false>(a, refresh_lds_bytes(a.ns, a.nup, a.ndn, false), s);
            return e != hipSuccess ? e : launch_solve<T, 16>(a, s);
        }
        default:
            break;
    }
    switch (refresh_rows_per_thread(a.nmax)) {
        case 1: return launch_lu<T, 1, true>(a, lds, s);
        case 2: return launch_lu<T, 2, true>(a, lds, s);
        case 3: return launch_lu<T, 3, true>(a, lds, s);
        default: return hipErrorInvalidValue;   // rejected by refresh_supported
    }
}

// At most three panel rows per thread (N <= 768, i.e. 2N <= 768 for the real
// form of complex orbitals).  A four-row instance (about 700 spilled VGPRs)
// reported a zero pivot for a complex L = 18 start with cond(tildeU) < 1e6
// that the oracle factors; the cause was not found, so such sizes are
// rejected when the handle is created.
bool refresh_supported(int ns, int nup, int ndn) {
    return refresh_lds_bytes(ns, nup, ndn) <= 160 * 1024 && refresh_rows_per_thread(nup > ndn ? nup : ndn) <= 3;
}

hipError_t launch_refresh(int prec, const RefreshArgs& a, hipStream_t s) {
    const size_t lds = refresh_lds_bytes(a.ns, a.nup, a.ndn);
    if (!refresh_supported(a.ns, a.nup, a.ndn)) return hipErrorInvalidValue;
    return prec == 0 ? launch_refresh_t<double>(a, lds, s) : launch_refresh_t<float>(a, lds, s);
}

}  // namespace kvmc
