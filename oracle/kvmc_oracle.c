/*
 * kvmc_oracle.c -- CPU FP64 restatement of the KagomeDSL.jl VMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (kagomedsl.jl_amd/) never links, imports or calls it.
 *
 * Every function restates one reference function; the citation is given as
 * path:line into hz-xiaxz/KagomeDSL.jl.  Arithmetic follows the reference:
 * W is complex double (eltype(U) = ComplexF64, src/MonteCarlo.jl:181-182),
 * the rank-1 update is BLAS geru (src/MonteCarlo.jl:279-292), the refresh is
 * LU + inverse + GEMM (src/MonteCarlo.jl:55-66), and the sweep keeps the
 * reference's shared-uniform transition kernel (src/MonteCarlo.jl:538-607).
 *
 * Conventions: matrices are column-major (Julia layout), site and label
 * indices at this interface are 1-based as in the reference, kappa[s] = 0
 * marks an empty site.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  The sampler's device  */
/* RNG; restated here so device chains can be replayed on the CPU.           */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void kvo_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Draws of one sweep: u in [0,1) with 53 random bits, bond index in [0,nb). */
static void sweep_draws(uint64_t seed, uint64_t walker, int64_t sweep, int nb,
                        double* u, int* bond) {
    uint32_t ctr[4] = {(uint32_t)sweep, (uint32_t)((uint64_t)sweep >> 32), (uint32_t)walker, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t x[4];
    kvo_philox4x32_10(ctr, key, x);
    *u = ((double)(x[0] >> 5) * 67108864.0 + (double)(x[1] >> 6)) * (1.0 / 9007199254740992.0);
    *bond = (int)(((uint64_t)x[2] * (uint64_t)nb) >> 32);
}

void kvo_sweep_draws(uint64_t seed, uint64_t walker, int64_t sweep, int nb, double* u, int* bond) {
    sweep_draws(seed, walker, sweep, nb, u, bond);
}

/* Seeded uniform Sz=0 start: Fisher-Yates permutation of the sites, first
 * n_up sites up (labels 1..n_up in permutation order), next n_dn down. */
void kvo_random_config(uint64_t seed, uint64_t walker, uint32_t attempt, int ns, int n_up,
                       int n_dn, int* kup, int* kdn) {
    int* perm = (int*)malloc(sizeof(int) * (size_t)ns);
    for (int i = 0; i < ns; ++i) perm[i] = i;
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t words[4];
    int widx = 4;
    uint32_t call = 0;
    for (int i = ns - 1; i >= 1; --i) {
        if (widx == 4) {
            uint32_t ctr[4] = {call++, attempt, (uint32_t)walker, 1u};
            kvo_philox4x32_10(ctr, key, words);
            widx = 0;
        }
        int j = (int)(((uint64_t)words[widx++] * (uint64_t)(i + 1)) >> 32);
        int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
    }
    memset(kup, 0, sizeof(int) * (size_t)ns);
    memset(kdn, 0, sizeof(int) * (size_t)ns);
    for (int a = 0; a < n_up; ++a) kup[perm[a]] = a + 1;
    for (int b = 0; b < n_dn; ++b) kdn[perm[n_up + b]] = b + 1;
    free(perm);
}

/* ------------------------------------------------------------------------ */
/* Lattice geometry: src/Lattice.jl:68-97, src/Hamiltonian.jl:28-76,279-283 */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n1, n2;
    int pbc[2], apbc[2];
    double t;
    double a1[2], a2[2];
    double r[6][2];
} kvo_lattice;

/* DoubleKagome(t, n1, n2, PBC; antiPBC): src/Lattice.jl:68-92.
 * Returns 0, or 1 for antiPBC without PBC (ArgumentError, :3-15), 2 for odd n1 (:77). */
int kvo_lattice_init(kvo_lattice* lat, double t, int n1, int n2, int pbc1, int pbc2, int apbc1,
                     int apbc2) {
    if ((apbc1 && !pbc1) || (apbc2 && !pbc2)) return 1;
    if (n1 % 2 != 0) return 2;
    double a = 2.0 * t;
    lat->n1 = n1; lat->n2 = n2; lat->t = t;
    lat->pbc[0] = pbc1; lat->pbc[1] = pbc2; lat->apbc[0] = apbc1; lat->apbc[1] = apbc2;
    lat->a1[0] = 2.0 * a; lat->a1[1] = 0.0;
    lat->a2[0] = 0.5 * a; lat->a2[1] = 0.5 * sqrt(3.0) * a;
    lat->r[0][0] = 0.0; lat->r[0][1] = 0.0;
    lat->r[1][0] = 0.25 * lat->a1[0]; lat->r[1][1] = 0.25 * lat->a1[1];
    lat->r[2][0] = 0.5 * lat->a2[0]; lat->r[2][1] = 0.5 * lat->a2[1];
    lat->r[3][0] = 0.5 * lat->a1[0]; lat->r[3][1] = 0.5 * lat->a1[1];
    lat->r[4][0] = 0.75 * lat->a1[0]; lat->r[4][1] = 0.75 * lat->a1[1];
    lat->r[5][0] = 2.5 * t; lat->r[5][1] = 0.5 * sqrt(3.0) * t;
    return 0;
}

/* unitcell_coord: src/Hamiltonian.jl:28-38 (s is 1-based) */
void kvo_unitcell_coord(const kvo_lattice* lat, int s, double out[2]) {
    int n1 = lat->n1 / 2;
    int cell = (s - 1) / 6;
    double c1 = (double)(cell % n1), c2 = (double)(cell / n1);
    out[0] = c1 * lat->a1[0] + c2 * lat->a2[0];
    out[1] = c1 * lat->a1[1] + c2 * lat->a2[1];
}

/* unitcell_diff: src/Hamiltonian.jl:56-76 (Julia round = ties to even) */
void kvo_unitcell_diff(const kvo_lattice* lat, const double c1[2], const double c2[2], int* dx,
                       int* dy) {
    double d0 = c1[0] - c2[0], d1 = c1[1] - c2[1];
    const double* a1 = lat->a1;
    const double* a2 = lat->a2;
    double det = a1[0] * a2[1] - a1[1] * a2[0];
    *dx = (int)nearbyint((a2[1] * d0 - a2[0] * d1) / det);
    *dy = (int)nearbyint((-a1[1] * d0 + a1[0] * d1) / det);
}

/* get_site_coord: src/Hamiltonian.jl:279-283 */
void kvo_site_coord(const kvo_lattice* lat, int s, double out[2]) {
    int label = (s - 1) % 6;
    kvo_unitcell_coord(lat, s, out);
    out[0] += lat->r[label][0];
    out[1] += lat->r[label][1];
}

/* get_boundary_shifts: src/Hamiltonian.jl:99-153.  Writes up to 9 (dx,dy,sign)
 * triples after `unique`; returns their count. */
int kvo_boundary_shifts(const kvo_lattice* lat, int s1, int s2, int* dxs, int* dys,
                        double* signs) {
    int n1 = lat->n1 / 2, n2 = lat->n2;
    double u1[2], u2[2];
    kvo_unitcell_coord(lat, s1, u1);
    kvo_unitcell_coord(lat, s2, u2);
    int dx, dy;
    kvo_unitcell_diff(lat, u2, u1, &dx, &dy);
    int cnt = 0;
    dxs[cnt] = dx; dys[cnt] = dy; signs[cnt] = 1.0; ++cnt;
    if (!lat->pbc[0] && !lat->pbc[1]) return cnt;
    int xs[3], ys[3], nx = 0, ny = 0;
    if (lat->pbc[0]) { xs[0] = -n1; xs[1] = 0; xs[2] = n1; nx = 3; } else { xs[0] = 0; nx = 1; }
    if (lat->pbc[1]) { ys[0] = -n2; ys[1] = 0; ys[2] = n2; ny = 3; } else { ys[0] = 0; ny = 1; }
    for (int a = 0; a < nx; ++a) {
        for (int b = 0; b < ny; ++b) {
            int sh1 = xs[a], sh2 = ys[b];
            if (sh1 == 0 && sh2 == 0) continue;
            double sg = 1.0;
            if (lat->apbc[0] && sh1 != 0) sg *= -1.0;
            if (lat->apbc[1] && sh2 != 0) sg *= -1.0;
            int ndx = dx + sh1, ndy = dy + sh2;
            int dup = 0;
            for (int q = 0; q < cnt; ++q)
                if (dxs[q] == ndx && dys[q] == ndy && signs[q] == sg) dup = 1;
            if (dup) continue;
            dxs[cnt] = ndx; dys[cnt] = ndy; signs[cnt] = sg; ++cnt;
        }
    }
    return cnt;
}

/* Hmat: src/Hamiltonian.jl:308-354 with apply_boundary_conditions! (:177-207).
 * link_in: n_in entries (l1, l2, value); link_inter: n_inter entries
 * (l1, l2, dx, dy, value); values are complex (the reference's Dict values
 * multiply a ComplexF64 matrix).  H is ns x ns, column-major, H = -(T + T').
 * Returns 0, or 3 if T is not upper triangular (the reference's error). */
int kvo_hmat(const kvo_lattice* lat, int n_in, const int* in_l1, const int* in_l2,
             const cplx* in_val, int n_inter, const int* it_l1, const int* it_l2,
             const int* it_dx, const int* it_dy, const cplx* it_val, double B, cplx* H) {
    int n1 = lat->n1, n2 = lat->n2;
    int ns = n1 * n2 * 3;
    int ncell = n1 * n2 / 2;
    cplx* T = (cplx*)calloc((size_t)ns * ns, sizeof(cplx));
#define TIDX(i, j) ((size_t)((j) - 1) * ns + (size_t)((i) - 1))
    /* in-cell bonds */
    for (int cell = 1; cell <= ncell; ++cell) {
        for (int s1 = (cell - 1) * 6 + 1; s1 <= cell * 6; ++s1) {
            for (int s2 = (cell - 1) * 6 + 1; s2 <= cell * 6; ++s2) {
                if (s1 >= s2) continue;
                int l1 = (s1 - 1) % 6 + 1, l2 = (s2 - 1) % 6 + 1;
                for (int q = 0; q < n_in; ++q) {
                    if (in_l1[q] == l1 && in_l2[q] == l2) {
                        double r1[2], r2[2];
                        kvo_site_coord(lat, s1, r1);
                        kvo_site_coord(lat, s2, r2);
                        double ph = (B / 2.0) * (r1[0] + r2[0]) * (r2[1] - r1[1]);
                        T[TIDX(s1, s2)] = in_val[q] * cexp(I * ph);
                        break;
                    }
                }
            }
        }
    }
    /* inter-cell bonds */
    int dxs[9], dys[9];
    double sgs[9];
    for (int c1 = 1; c1 <= ncell; ++c1) {
        for (int c2 = 1; c2 <= ncell; ++c2) {
            if (c1 == c2) continue;
            for (int s1 = (c1 - 1) * 6 + 1; s1 <= c1 * 6; ++s1) {
                for (int s2 = (c2 - 1) * 6 + 1; s2 <= c2 * 6; ++s2) {
                    if (s1 >= s2) continue;
                    int l1 = (s1 - 1) % 6 + 1, l2 = (s2 - 1) % 6 + 1;
                    int nsh = kvo_boundary_shifts(lat, s1, s2, dxs, dys, sgs);
                    double r1[2], ruc1[2], rs2[2], ruc2[2], dr2[2];
                    kvo_site_coord(lat, s1, r1);
                    kvo_unitcell_coord(lat, s1, ruc1);
                    kvo_site_coord(lat, s2, rs2);
                    kvo_unitcell_coord(lat, s2, ruc2);
                    dr2[0] = rs2[0] - ruc2[0];
                    dr2[1] = rs2[1] - ruc2[1];
                    for (int k = 0; k < nsh; ++k) {
                        for (int q = 0; q < n_inter; ++q) {
                            if (it_l1[q] == l1 && it_l2[q] == l2 && it_dx[q] == dxs[k] &&
                                it_dy[q] == dys[k]) {
                                double ru[2] = {ruc1[0] + dxs[k] * lat->a1[0] + dys[k] * lat->a2[0],
                                                ruc1[1] + dxs[k] * lat->a1[1] + dys[k] * lat->a2[1]};
                                double r2[2] = {ru[0] + dr2[0], ru[1] + dr2[1]};
                                double ph = (B / 2.0) * (r1[0] + r2[0]) * (r2[1] - r1[1]);
                                T[TIDX(s1, s2)] += sgs[k] * it_val[q] * cexp(I * ph);
                                break;
                            }
                        }
                    }
                }
            }
        }
    }
    int bad = 0;
    for (int i = 1; i <= ns; ++i)
        for (int j = 1; j < i; ++j)
            if (T[TIDX(i, j)] != 0) bad = 1;
    for (int j = 1; j <= ns; ++j)
        for (int i = 1; i <= ns; ++i)
            H[TIDX(i, j)] = -(T[TIDX(i, j)] + conj(T[TIDX(j, i)]));
#undef TIDX
    free(T);
    return bad ? 3 : 0;
}

/* get_nn: src/Hamiltonian.jl:447-451.  findall over UpperTriangular(H) is
 * column-major: j outer, i inner, i <= j.  bonds receives 1-based pairs;
 * returns nb. */
int kvo_get_nn(int ns, const cplx* H, int* bonds) {
    int nb = 0;
    for (int j = 0; j < ns; ++j)
        for (int i = 0; i <= j; ++i)
            if (H[(size_t)j * ns + i] != 0) {
                bonds[2 * nb] = i + 1;
                bonds[2 * nb + 1] = j + 1;
                ++nb;
            }
    return nb;
}

/* ------------------------------------------------------------------------ */
/* Sampler: src/MonteCarlo.jl                                               */
/* ------------------------------------------------------------------------ */

/* tilde_U: src/MonteCarlo.jl:92-115.  Returns 0, 1 (kappa count != m),
 * 2 (label out of range). */
int kvo_tilde_u(int ns, int m, const cplx* U, const int* kappa, cplx* tU) {
    int cnt = 0;
    for (int s = 0; s < ns; ++s) cnt += kappa[s] != 0;
    if (cnt != m) return 1;
    memset(tU, 0, sizeof(cplx) * (size_t)m * m);
    for (int s = 0; s < ns; ++s) {
        int l = kappa[s];
        if (l == 0) continue;
        if (l < 1 || l > m) return 2;
        for (int c = 0; c < m; ++c) tU[(size_t)c * m + (l - 1)] = U[(size_t)c * ns + s];
    }
    return 0;
}

static double cabs1(cplx z) { return fabs(creal(z)) + fabs(cimag(z)); }

/* In-place LU with partial pivoting (LAPACK getrf semantics: pivot = first
 * max of |re|+|im|).  A is n x n column-major.  Returns 0 or k>0 when
 * U(k,k) is exactly zero (SingularException). */
int kvo_getrf(int n, cplx* A, int* piv) {
    int info = 0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = cabs1(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            double v = cabs1(A[(size_t)k * n + i]);
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (A[(size_t)k * n + p] == 0) {
            if (!info) info = k + 1;
            continue;
        }
        if (p != k)
            for (int c = 0; c < n; ++c) {
                cplx t = A[(size_t)c * n + k];
                A[(size_t)c * n + k] = A[(size_t)c * n + p];
                A[(size_t)c * n + p] = t;
            }
        cplx rp = 1.0 / A[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) A[(size_t)k * n + i] *= rp;
        for (int c = k + 1; c < n; ++c) {
            cplx akc = A[(size_t)c * n + k];
            if (akc == 0) continue;
            for (int i = k + 1; i < n; ++i) A[(size_t)c * n + i] -= A[(size_t)k * n + i] * akc;
        }
    }
    return info;
}

/* inv(A) from its LU factors (solve A X = I column by column). */
static void lu_inverse(int n, const cplx* LU, const int* piv, cplx* X) {
    for (int c = 0; c < n; ++c) {
        cplx* x = X + (size_t)c * n;
        for (int i = 0; i < n; ++i) x[i] = (i == c) ? 1.0 : 0.0;
        for (int k = 0; k < n; ++k)
            if (piv[k] != k) { cplx t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
        for (int k = 0; k < n; ++k) {
            cplx xk = x[k];
            if (xk == 0) continue;
            for (int i = k + 1; i < n; ++i) x[i] -= xk * LU[(size_t)k * n + i];
        }
        for (int k = n - 1; k >= 0; --k) {
            x[k] /= LU[(size_t)k * n + k];
            cplx xk = x[k];
            for (int i = 0; i < k; ++i) x[i] -= xk * LU[(size_t)k * n + i];
        }
    }
}

/* reevaluateW!: src/MonteCarlo.jl:55-66 for one spin:
 * W = U * (tilde_U(U, kappa) \ I).  Returns 0, or a nonzero code when
 * tilde_U is invalid (1,2) or singular (3).  logdet (optional) receives
 * log|det tilde_U|. */
int kvo_refresh_w(int ns, int m, const cplx* U, const int* kappa, cplx* W, double* logdet) {
    cplx* tU = (cplx*)malloc(sizeof(cplx) * (size_t)m * m);
    cplx* inv = (cplx*)malloc(sizeof(cplx) * (size_t)m * m);
    int* piv = (int*)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    int rc = kvo_tilde_u(ns, m, U, kappa, tU);
    if (rc == 0 && kvo_getrf(m, tU, piv) != 0) rc = 3;
    if (rc == 0) {
        if (logdet) {
            double ld = 0.0;
            for (int k = 0; k < m; ++k) ld += log(cabs(tU[(size_t)k * m + k]));
            *logdet = ld;
        }
        lu_inverse(m, tU, piv, inv);
        for (int c = 0; c < m; ++c) {
            cplx* w = W + (size_t)c * ns;
            for (int i = 0; i < ns; ++i) w[i] = 0;
            for (int k = 0; k < m; ++k) {
                cplx b = inv[(size_t)c * m + k];
                const cplx* u = U + (size_t)k * ns;
                for (int i = 0; i < ns; ++i) w[i] += u[i] * b;
            }
        }
    }
    free(tU); free(inv); free(piv);
    return rc;
}

/* update_W!: src/MonteCarlo.jl:279-292 (l, K are 1-based).  geru as in the
 * reference BLAS: temp = alpha*y[j]; A[i,j] += x[i]*temp. */
void kvo_update_w(int ns, int m, cplx* W, int l, int K) {
    cplx* row = (cplx*)malloc(sizeof(cplx) * (size_t)m);
    cplx* col = (cplx*)malloc(sizeof(cplx) * (size_t)ns);
    for (int j = 0; j < m; ++j) row[j] = W[(size_t)j * ns + (K - 1)];
    row[l - 1] -= 1.0;
    for (int i = 0; i < ns; ++i) col[i] = W[(size_t)(l - 1) * ns + i];
    cplx alpha = -1.0 / W[(size_t)(l - 1) * ns + (K - 1)];
    for (int j = 0; j < m; ++j) {
        if (row[j] == 0) continue;
        cplx temp = alpha * row[j];
        cplx* w = W + (size_t)j * ns;
        for (int i = 0; i < ns; ++i) w[i] += col[i] * temp;
    }
    free(row); free(col);
}

/* Z: src/MonteCarlo.jl:460-474 (bonds 1-based pairs). */
int kvo_z(int nb, const int* bonds, const int* kup, const int* kdn) {
    int count = 0;
    for (int b = 0; b < nb; ++b) {
        int s1 = bonds[2 * b] - 1, s2 = bonds[2 * b + 1] - 1;
        if (kup[s1] != 0 && kdn[s2] != 0) ++count;
        else if (kup[s2] != 0 && kdn[s1] != 0) ++count;
    }
    return count;
}

/* getOL: src/Hamiltonian.jl:762-778 with getxprime (:711-720),
 * spinInteraction! (:636-672), SzInteraction! (:593-604) and Sz (:531-565).
 * Returns the complex sum of the diagonal and exchange terms; the reference's
 * OL is its real part.  *err = 1 if a bond touches an empty or doubly
 * occupied site (Sz's ArgumentError). */
static cplx getol_sum(int ns, int nb, const int* bonds, const int* kup, const int* kdn, const cplx* Wup,
                      const cplx* Wdn, int* err) {
    double diag = 0.0;
    cplx off = 0.0;
    *err = 0;
    for (int b = 0; b < nb; ++b) {
        int i = bonds[2 * b] - 1, j = bonds[2 * b + 1] - 1;
        int iu = kup[i], ju = kup[j], id = kdn[i], jd = kdn[j];
        if (ju != 0 && id != 0)
            off += -0.5 * Wup[(size_t)(ju - 1) * ns + i] * Wdn[(size_t)(id - 1) * ns + j];
        if (iu != 0 && jd != 0)
            off += -0.5 * Wup[(size_t)(iu - 1) * ns + j] * Wdn[(size_t)(jd - 1) * ns + i];
        int ui = iu != 0, di = id != 0, uj = ju != 0, dj = jd != 0;
        if (ui == di || uj == dj) { *err = 1; continue; }
        double szi = ui ? 0.5 : -0.5, szj = uj ? 0.5 : -0.5;
        diag += szi * szj;
    }
    return diag + off;
}

double kvo_getol(int ns, int nup, int ndn, int nb, const int* bonds, const int* kup,
                 const int* kdn, const cplx* Wup, const cplx* Wdn, int* err) {
    (void)nup; (void)ndn;
    return creal(getol_sum(ns, nb, bonds, kup, kdn, Wup, Wdn, err));
}

/* The complex local energy <x|H|psi>/<x|psi> (real part: getOL); its
 * imaginary part enters the energy gradient of a complex wave function. */
void kvo_getol_c(int ns, int nb, const int* bonds, const int* kup, const int* kdn, const cplx* Wup,
                 const cplx* Wdn, double* re, double* im, int* err) {
    const cplx v = getol_sum(ns, nb, bonds, kup, kdn, Wup, Wdn, err);
    *re = creal(v);
    *im = cimag(v);
}

/* One chain (one walker). */
typedef struct {
    int ns, nup, ndn, nb;
    const int* bonds;
    const cplx* Uup;
    const cplx* Udn;
    int* kup;
    int* kdn;
    cplx* Wup;
    cplx* Wdn;
} kvo_chain;

enum { KVO_KERNEL_REFERENCE = 0, KVO_KERNEL_UNBIASED = 1 };
enum { KVO_REFRESH_REFERENCE = 0, KVO_REFRESH_EVERY_K = 1 };

/* update_configurations!: src/MonteCarlo.jl:497-511 (i, site 1-based) */
static void update_configurations(kvo_chain* c, int flag, int i, int site, int l_up, int l_dn) {
    if (flag == 1) {
        kvo_update_w(c->ns, c->nup, c->Wup, l_up, site);
        kvo_update_w(c->ns, c->ndn, c->Wdn, l_dn, i);
        c->kup[i - 1] = 0; c->kup[site - 1] = l_up;
        c->kdn[i - 1] = l_dn; c->kdn[site - 1] = 0;
    } else {
        kvo_update_w(c->ns, c->nup, c->Wup, l_up, i);
        kvo_update_w(c->ns, c->ndn, c->Wdn, l_dn, site);
        c->kup[i - 1] = l_up; c->kup[site - 1] = 0;
        c->kdn[i - 1] = 0; c->kdn[site - 1] = l_dn;
    }
}

/* Outcome codes of one sweep. */
enum {
    KVO_PRE_REJECT = 0,   /* r > Zmu/Zmax: early return, :547-550 */
    KVO_NO_MOVE = 1,      /* bond not antiparallel: early return, :564-567 */
    KVO_REJECT = 2,       /* Metropolis reject */
    KVO_ACCEPT = 3        /* accepted, :581-591 */
};

/* Carlo.sweep!: src/MonteCarlo.jl:538-607, with the draws (r, bond) supplied
 * by the caller.  sweeps_before is ctx.sweeps as seen by sweep!.  The
 * refresh follows the policy: reference cadence (ctx.sweeps % n_occupied
 * == 0 on sweeps that pass both early returns, :593-604) or every k
 * accepted moves (acc_since counts accepted moves since the last refresh).
 * *refreshed = 1 when a refresh ran; returns the outcome code, or -1 on a
 * singular refresh. */
int kvo_sweep_one(kvo_chain* c, double r, int bond_idx, int64_t sweeps_before, int kernel,
                  int policy, int k, int* acc_since, int* refreshed, double* ratio_out) {
    *refreshed = 0;
    int Zmu = kvo_z(c->nb, c->bonds, c->kup, c->kdn);
    int Zmax = c->nb;
    double zr = (double)Zmu / (double)Zmax;
    if (kernel == KVO_KERNEL_REFERENCE && r > zr) return KVO_PRE_REJECT;
    int i = c->bonds[2 * bond_idx], site = c->bonds[2 * bond_idx + 1];
    int can1 = c->kup[i - 1] != 0 && c->kdn[site - 1] != 0;
    int can2 = c->kup[site - 1] != 0 && c->kdn[i - 1] != 0;
    if (!can1 && !can2) return KVO_NO_MOVE;
    /* flag = sample(rng, maybe_update): the valid set has one element when
     * every site is singly occupied */
    int flag = can1 ? 1 : 2;
    int l_up = flag == 1 ? c->kup[i - 1] : c->kup[site - 1];
    int l_dn = flag == 1 ? c->kdn[site - 1] : c->kdn[i - 1];
    cplx ratio;
    if (flag == 1)
        ratio = c->Wup[(size_t)(l_up - 1) * c->ns + (site - 1)] *
                c->Wdn[(size_t)(l_dn - 1) * c->ns + (i - 1)];
    else
        ratio = c->Wup[(size_t)(l_up - 1) * c->ns + (i - 1)] *
                c->Wdn[(size_t)(l_dn - 1) * c->ns + (site - 1)];
    double acc_prob = creal(ratio) * creal(ratio) + cimag(ratio) * cimag(ratio);
    if (ratio_out) *ratio_out = acc_prob;
    int accept;
    if (kernel == KVO_KERNEL_REFERENCE)
        accept = (acc_prob >= 1 && r < zr) || (acc_prob < 1 && r < zr * acc_prob);
    else
        accept = r < acc_prob;
    int outcome = KVO_REJECT;
    if (accept) {
        update_configurations(c, flag, i, site, l_up, l_dn);
        outcome = KVO_ACCEPT;
        ++*acc_since;
    }
    int do_refresh;
    if (policy == KVO_REFRESH_REFERENCE) {
        int n_occ = c->nup < c->ndn ? c->nup : c->ndn;
        do_refresh = (sweeps_before % n_occ) == 0;
    } else {
        do_refresh = accept && *acc_since >= k;
    }
    if (do_refresh) {
        if (kvo_refresh_w(c->ns, c->nup, c->Uup, c->kup, c->Wup, NULL) != 0) return -1;
        if (kvo_refresh_w(c->ns, c->ndn, c->Udn, c->kdn, c->Wdn, NULL) != 0) return -1;
        *refreshed = 1;
        *acc_since = 0;
    }
    return outcome;
}

/* Runs nsweeps sweeps of one chain from sweep counter sweep0.  Draws come
 * from the tape (tape_u/tape_bond, one entry per sweep) when tape_u is not
 * NULL, else from Philox keyed (seed, walker).  outcomes (optional) gets one
 * code per sweep.  When measure_every > 0 the local energy is taken after
 * each sweep whose post-increment counter is a multiple of measure_every
 * and exceeds `thermalization` (Carlo's measure! gating, src/MonteCarlo.jl:
 * 628-634); eloc receives them in order and *n_eloc their count.  Returns
 * the number of accepted moves, or -1 on a singular refresh. */
int64_t kvo_run_chain(int ns, int nup, int ndn, int nb, const int* bonds, const cplx* Uup,
                      const cplx* Udn, int* kup, int* kdn, cplx* Wup, cplx* Wdn, int64_t sweep0,
                      int64_t nsweeps, const double* tape_u, const int* tape_bond, uint64_t seed,
                      uint64_t walker, int kernel, int policy, int k, int* acc_since,
                      int64_t measure_every, int64_t thermalization, double* eloc,
                      int64_t* n_eloc, int8_t* outcomes, int64_t* n_refresh) {
    kvo_chain c = {ns, nup, ndn, nb, bonds, Uup, Udn, kup, kdn, Wup, Wdn};
    int64_t acc = 0, ne = 0, nr = 0;
    for (int64_t s = 0; s < nsweeps; ++s) {
        int64_t sweeps_before = sweep0 + s;
        double r;
        int bidx;
        if (tape_u) {
            r = tape_u[s];
            bidx = tape_bond[s];
        } else {
            sweep_draws(seed, walker, sweeps_before, nb, &r, &bidx);
        }
        int refreshed = 0;
        int oc = kvo_sweep_one(&c, r, bidx, sweeps_before, kernel, policy, k, acc_since,
                               &refreshed, NULL);
        if (oc < 0) return -1;
        nr += refreshed;
        if (outcomes) outcomes[s] = (int8_t)oc;
        acc += oc == KVO_ACCEPT;
        int64_t after = sweeps_before + 1;
        if (measure_every > 0 && after > thermalization && after % measure_every == 0) {
            int err;
            double e = kvo_getol(ns, nup, ndn, nb, bonds, kup, kdn, Wup, Wdn, &err);
            if (eloc) eloc[ne] = e;
            ++ne;
        }
    }
    if (n_eloc) *n_eloc = ne;
    if (n_refresh) *n_refresh = nr;
    return acc;
}

/* getxprime: src/Hamiltonian.jl:711-720.  Writes the diagonal coefficient
 * (key (-1,-1,-1,-1)) to *diag and the off-diagonal keys (1-based
 * (K_up, l_up, K_dn, l_dn)) with their merged coefficients to keys/coef;
 * returns the number of off-diagonal keys, or -1 on an invalid site (Sz's
 * ArgumentError).  keys must hold 4*2*nb ints. */
int kvo_getxprime(int nb, const int* bonds, const int* kup, const int* kdn, double* diag,
                  int* keys, double* coef) {
    int nk = 0;
    *diag = 0.0;
    for (int b = 0; b < nb; ++b) {
        int i = bonds[2 * b], j = bonds[2 * b + 1];
        int iu = kup[i - 1], ju = kup[j - 1], id = kdn[i - 1], jd = kdn[j - 1];
        int cand[2][4] = {{i, ju, j, id}, {j, iu, i, jd}};
        int ok[2] = {ju != 0 && id != 0, iu != 0 && jd != 0};
        for (int c = 0; c < 2; ++c) {
            if (!ok[c]) continue;
            int q;
            for (q = 0; q < nk; ++q)
                if (keys[4 * q] == cand[c][0] && keys[4 * q + 1] == cand[c][1] &&
                    keys[4 * q + 2] == cand[c][2] && keys[4 * q + 3] == cand[c][3])
                    break;
            if (q == nk) {
                memcpy(keys + 4 * nk, cand[c], sizeof(int) * 4);
                coef[nk] = 0.0;
                ++nk;
            }
            coef[q] -= 0.5;
        }
        int ui = iu != 0, di = id != 0, uj = ju != 0, dj = jd != 0;
        if (ui == di || uj == dj) return -1;
        *diag += (ui ? 0.5 : -0.5) * (uj ? 0.5 : -0.5);
    }
    return nk;
}
