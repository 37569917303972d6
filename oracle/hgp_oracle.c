/*
 * hgp_oracle.c -- CPU restatement of MolecularIntegrals.jl's HGP ERI path.
 *
 * TEST INFRASTRUCTURE ONLY (see hgp_oracle.h): the checker for the HIP path and
 * the CPU baseline timed beside it.  Never linked into the product library.
 *
 * Citations are to files of rpmuller/MolecularIntegrals.jl.
 * Parity status of each part (how it is pinned) is listed in the header
 * comment of tests/test_oracle.py and tests/test_oracle_tho.py.
 */
#include "hgp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* AO tables: src/Basis.jl:269-385.                                          */
/* shell_indices order for L=0..4 is the reference's literal table          */
/* (Basis.jl:269-279); note L=3 lists xyy, xzz before xyz.  L>=5 (only used  */
/* for HRR intermediates, which the reference cannot reach) uses the usual   */
/* descending-I, descending-J order, which the L=0,1,2,4 rows also follow.   */
/* ------------------------------------------------------------------------ */
#define OR_LMAX 8
#define OR_NAO_MAX (((OR_LMAX + 1) * (OR_LMAX + 2) * (OR_LMAX + 3)) / 6)

static int g_init = 0;
static int g_ao2m[OR_NAO_MAX][3];
static int g_shift_index[OR_NAO_MAX][3];
static int g_shift_index_plus[OR_NAO_MAX][3];
static int g_shift_direction[OR_NAO_MAX];
static int g_shell_number[OR_NAO_MAX];

int oracle_lmax(void) { return OR_LMAX; }

int oracle_nao(int l) { return l < 0 ? 0 : ((l + 1) * (l + 2) * (l + 3)) / 6; }

static int ncart(int l) { return ((l + 1) * (l + 2)) / 2; }

static void shell_components(int L, int comps[][3]) {
    int n = 0;
    if (L == 3) {
        /* Basis.jl:273-275 */
        static const int f[10][3] = {{3, 0, 0}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {1, 0, 2},
                                     {1, 1, 1}, {0, 3, 0}, {0, 2, 1}, {0, 1, 2}, {0, 0, 3}};
        memcpy(comps, f, sizeof(f));
        return;
    }
    for (int I = L; I >= 0; --I)
        for (int J = L - I; J >= 0; --J) {
            comps[n][0] = I;
            comps[n][1] = J;
            comps[n][2] = L - I - J;
            ++n;
        }
}

static int m2ao_raw(int I, int J, int K) {
    int L = I + J + K;
    int comps[64][3];
    shell_components(L, comps);
    for (int m = 0; m < ncart(L); ++m)
        if (comps[m][0] == I && comps[m][1] == J && comps[m][2] == K) return oracle_nao(L - 1) + m;
    return -1;
}

static void init_tables(void) {
    if (g_init) return;
    int a = 0;
    for (int L = 0; L <= OR_LMAX; ++L) {
        int comps[64][3];
        shell_components(L, comps);
        for (int m = 0; m < ncart(L); ++m, ++a) {
            g_ao2m[a][0] = comps[m][0];
            g_ao2m[a][1] = comps[m][1];
            g_ao2m[a][2] = comps[m][2];
        }
    }
    for (a = 0; a < OR_NAO_MAX; ++a) {
        const int *m = g_ao2m[a];
        int L = m[0] + m[1] + m[2];
        g_shell_number[a] = L; /* Basis.jl:356-364 */
        /* Julia argmax returns the first maximal index (Basis.jl:346-354) */
        int best = 0;
        for (int i = 1; i < 3; ++i)
            if (m[i] > m[best]) best = i;
        g_shift_direction[a] = best;
        for (int i = 0; i < 3; ++i) {
            int mm[3] = {m[0], m[1], m[2]};
            if (m[i] == 0) {
                g_shift_index[a][i] = -1; /* Basis.jl:326-343 */
            } else {
                mm[i] -= 1;
                g_shift_index[a][i] = m2ao_raw(mm[0], mm[1], mm[2]);
            }
            mm[0] = m[0];
            mm[1] = m[1];
            mm[2] = m[2];
            if (L == OR_LMAX) {
                g_shift_index_plus[a][i] = -1; /* Basis.jl:366-385 */
            } else {
                mm[i] += 1;
                g_shift_index_plus[a][i] = m2ao_raw(mm[0], mm[1], mm[2]);
            }
        }
    }
    g_init = 1;
}

void oracle_ao2m(int a, int32_t m[3]) {
    init_tables();
    m[0] = g_ao2m[a][0];
    m[1] = g_ao2m[a][1];
    m[2] = g_ao2m[a][2];
}
int oracle_m2ao(int I, int J, int K) { return m2ao_raw(I, J, K); }
int oracle_shift_index(int a, int i) { init_tables(); return g_shift_index[a][i]; }
int oracle_shift_direction(int a) { init_tables(); return g_shift_direction[a]; }
int oracle_shell_number(int a) { init_tables(); return g_shell_number[a]; }
int oracle_shift_index_plus(int a, int i) { init_tables(); return g_shift_index_plus[a][i]; }

/* ------------------------------------------------------------------------ */
/* Small math helpers.                                                      */
/* ------------------------------------------------------------------------ */
/* factorial2(n) = prod(n:-2:1), so factorial2(-1) = 1 (src/Utils.jl:5). */
static double factorial2(int n) {
    double p = 1.0;
    for (int k = n; k >= 1; k -= 2) p *= k;
    return p;
}
static double ipow(double x, int n) {
    double r = 1.0; /* 0.0^0 == 1.0, as in Julia */
    for (int k = 0; k < n; ++k) r *= x;
    return r;
}
static double factorial(int n) {
    double p = 1.0;
    for (int k = 2; k <= n; ++k) p *= k;
    return p;
}
static double binomial(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}
static double dist2(const double *a, const double *b) {
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

/* ------------------------------------------------------------------------ */
/* Boys function: src/Boys.jl.                                              */
/* ------------------------------------------------------------------------ */
/* Fm(m,T) = gamma(m+1/2) * gamma_inc(m+1/2,T,0)[1] / (2 T^(m+1/2)), with     */
/* T < 1e-18 -> 1/(2m+1)  (Boys.jl:29-34).  The reference gets the lower     */
/* regularised incomplete gamma from SpecialFunctions.jl 2.3.0; here the     */
/* product gamma(a)P(a,T) = gamma_lower(a,T) is evaluated by its power        */
/* series (DLMF 8.7.1 / Numerical Recipes "gser"):                           */
/*   gamma_lower(a,x) = x^a e^{-x} sum_{n>=0} x^n / (a (a+1) ... (a+n)),      */
/* so Fm(T) = e^{-T}/2 * sum_n T^n / prod_{k=0..n} (m+1/2+k).  All terms are   */
/* positive (no cancellation); the sum runs in long double.                  */
double oracle_Fm(int m, double T) {
    const double SMALL = 1e-18;
    if (T < SMALL) return 1.0 / (2 * m + 1);
    long double a = (long double)m + 0.5L;
    long double term = 1.0L / a;
    long double sum = term;
    long double x = (long double)T;
    for (int n = 1; n < 10000; ++n) {
        term *= x / (a + n);
        sum += term;
        if (term < sum * 1e-21L) break;
    }
    return (double)(0.5L * expl(-x) * sum);
}

/* farray(mmax,T): T < Tmax=30 -> top m from Fm then downward recursion      */
/* F_m = (2T F_{m+1} + e^{-T})/(2m+1) (Boys.jl:9-17); T >= 30 -> asymptotic  */
/* Fm_asymp (Boys.jl:36) then F_m = T F_{m+1}/(m+1/2), no e^{-T} term        */
/* (Boys.jl:21-26).  Fms[k] holds F_k (0-based).                              */
void oracle_farray(int mmax, double T, double *Fms) {
    const double Tmax = 30.0;
    if (T < Tmax) {
        Fms[mmax - 1] = oracle_Fm(mmax - 1, T);
        double emt = exp(-T);
        for (int m = mmax - 2; m >= 0; --m) Fms[m] = (2 * T * Fms[m + 1] + emt) / (2 * m + 1);
        return;
    }
    int mt = mmax - 1;
    Fms[mt] = sqrt(M_PI / 2) * factorial2(2 * mt - 1) / pow(2 * T, mt + 0.5);
    for (int m = mt - 1; m >= 0; --m) Fms[m] = T * Fms[m + 1] / (m + 0.5);
}

/* ------------------------------------------------------------------------ */
/* vrr! -- src/HGP.jl:135-242 (HGP eq. 6, b = d = 0).                        */
/* ------------------------------------------------------------------------ */
#define VIDX(m, a, c) (((size_t)(m) * na + (a)) * nc + (c))

void oracle_vrr(double *vrrs, int amax, int cmax, double aexpn, double bexpn,
                double cexpn, double dexpn, const double *A, const double *B,
                const double *C, const double *D, int literal) {
    init_tables();
    const int mmax = amax + cmax + 1;
    const int na = oracle_nao(amax), nc = oracle_nao(cmax);
    memset(vrrs, 0, sizeof(double) * (size_t)mmax * na * nc);

    double zeta = aexpn + bexpn, eta = cexpn + dexpn, ze = zeta + eta;
    double ooz = 1 / zeta, ooe = 1 / eta, ooze = 1 / ze;
    double oortze = sqrt(ooze);
    double P[3], Q[3], W[3], PA[3], WP[3], QC[3], WQ[3];
    for (int k = 0; k < 3; ++k) {
        P[k] = (aexpn * A[k] + bexpn * B[k]) * ooz;
        Q[k] = (cexpn * C[k] + dexpn * D[k]) * ooe;
    }
    for (int k = 0; k < 3; ++k) {
        W[k] = (zeta * P[k] + eta * Q[k]) * ooze;
        PA[k] = P[k] - A[k];
        WP[k] = W[k] - P[k];
        QC[k] = Q[k] - C[k];
        WQ[k] = W[k] - Q[k];
    }
    double rab2 = dist2(A, B), rcd2 = dist2(C, D), rpq2 = dist2(P, Q);
    double T = zeta * eta * rpq2 * ooze;
    double expo = exp(-aexpn * bexpn * rab2 * ooz - cexpn * dexpn * rcd2 * ooe);
    double pref;
    if (literal & ORACLE_LIT_D1)
        pref = 2 * M_PI * M_PI * sqrt(M_PI) * ooze * oortze * expo; /* HGP.jl:150-151 */
    else
        pref = 2 * M_PI * M_PI * sqrt(M_PI) * ooz * ooe * oortze * expo; /* ERI.jl:45 */

    double Fms[64];
    oracle_farray(mmax, T, Fms); /* HGP.jl:168 */
    for (int m = 0; m < mmax; ++m) vrrs[VIDX(m, 0, 0)] = pref * Fms[m];
    if (mmax == 1) return;

    /* a-build, HGP.jl:180-197 */
    for (int aplus = 1; aplus < na; ++aplus) {
        int ashell = g_shell_number[aplus];
        int i = g_shift_direction[aplus];
        int a = g_shift_index[aplus][i];
        int lim = mmax - ashell;
        int aminus = g_shift_index[a][i];
        if (aminus >= 0) {
            double a_i = 0.5 * ooz * g_ao2m[a][i];
            for (int m = 0; m < lim; ++m)
                vrrs[VIDX(m, aplus, 0)] = PA[i] * vrrs[VIDX(m, a, 0)] + WP[i] * vrrs[VIDX(m + 1, a, 0)] +
                                          a_i * (vrrs[VIDX(m, aminus, 0)] - eta * ooze * vrrs[VIDX(m + 1, aminus, 0)]);
        } else {
            for (int m = 0; m < lim; ++m)
                vrrs[VIDX(m, aplus, 0)] = PA[i] * vrrs[VIDX(m, a, 0)] + WP[i] * vrrs[VIDX(m + 1, a, 0)];
        }
    }

    /* c-build, HGP.jl:204-240 */
    for (int cplus = 1; cplus < nc; ++cplus) {
        int cshell = g_shell_number[cplus];
        int i = g_shift_direction[cplus];
        int c = g_shift_index[cplus][i];
        int cminus = g_shift_index[c][i];
        for (int a = 0; a < na; ++a) {
            int ashell = g_shell_number[a];
            int lim = mmax - cshell - ashell;
            int aminus = g_shift_index[a][i];
            if (cminus >= 0) {
                double c_i = 0.5 * ooe * g_ao2m[c][i];
                if (aminus >= 0) {
                    double a_i = 0.5 * ooze * g_ao2m[a][i];
                    for (int m = 0; m < lim; ++m)
                        vrrs[VIDX(m, a, cplus)] = QC[i] * vrrs[VIDX(m, a, c)] + WQ[i] * vrrs[VIDX(m + 1, a, c)] +
                                                  a_i * vrrs[VIDX(m + 1, aminus, c)] +
                                                  c_i * (vrrs[VIDX(m, a, cminus)] - zeta * ooze * vrrs[VIDX(m + 1, a, cminus)]);
                } else {
                    for (int m = 0; m < lim; ++m)
                        vrrs[VIDX(m, a, cplus)] = QC[i] * vrrs[VIDX(m, a, c)] + WQ[i] * vrrs[VIDX(m + 1, a, c)] +
                                                  c_i * (vrrs[VIDX(m, a, cminus)] - zeta * ooze * vrrs[VIDX(m + 1, a, cminus)]);
                }
            } else if (aminus >= 0) {
                double a_i = 0.5 * ooze * g_ao2m[a][i];
                for (int m = 0; m < lim; ++m)
                    vrrs[VIDX(m, a, cplus)] = QC[i] * vrrs[VIDX(m, a, c)] + WQ[i] * vrrs[VIDX(m + 1, a, c)] +
                                              a_i * vrrs[VIDX(m + 1, aminus, c)];
            } else {
                for (int m = 0; m < lim; ++m)
                    vrrs[VIDX(m, a, cplus)] = QC[i] * vrrs[VIDX(m, a, c)] + WQ[i] * vrrs[VIDX(m + 1, a, c)];
            }
        }
    }
}
#undef VIDX

/* ------------------------------------------------------------------------ */
/* HRR on a [nao(la+lb)][nao(lc+ld)] contracted (e0|f0) block:              */
/* src/HGP.jl:87-114 (chrr) == :395-422 (hrr).                              */
/* ------------------------------------------------------------------------ */
#define HIDX(a, b, c, d) ((((size_t)(a) * nb + (b)) * nc + (c)) * nd + (d))
static void hrr_from_vrr(double *hrrs, const double *e0f0, int la, int lb, int lc, int ld,
                         const double *A, const double *B, const double *C, const double *D) {
    init_tables();
    const int na = oracle_nao(la + lb), nb = oracle_nao(lb), nc = oracle_nao(lc + ld), nd = oracle_nao(ld);
    memset(hrrs, 0, sizeof(double) * (size_t)na * nb * nc * nd);
    for (int c = 0; c < nc; ++c)
        for (int a = 0; a < na; ++a) hrrs[HIDX(a, 0, c, 0)] = e0f0[(size_t)a * nc + c];
    /* bra: (a, b+1_j | = (a+1_j, b| + (A_j - B_j)(a, b|  (HGP.jl:93-101) */
    for (int c = 0; c < nc; ++c)
        for (int bplus = 1; bplus < nb; ++bplus) {
            int j = g_shift_direction[bplus];
            int b = g_shift_index[bplus][j];
            int bs = g_shell_number[b];
            int alim = oracle_nao(la + lb - bs - 1);
            for (int a = 0; a < alim; ++a) {
                int aplus = g_shift_index_plus[a][j];
                hrrs[HIDX(a, bplus, c, 0)] = hrrs[HIDX(aplus, b, c, 0)] + (A[j] - B[j]) * hrrs[HIDX(a, b, c, 0)];
            }
        }
    /* ket: (ab| c, d+1_j) = (ab| c+1_j, d) + (C_j - D_j)(ab| c, d)  (HGP.jl:104-114) */
    for (int dplus = 1; dplus < nd; ++dplus) {
        int j = g_shift_direction[dplus];
        int d = g_shift_index[dplus][j];
        int ds = g_shell_number[d];
        int clim = oracle_nao(lc + ld - ds - 1);
        for (int c = 0; c < clim; ++c) {
            int cplus = g_shift_index_plus[c][j];
            for (int b = 0; b < nb; ++b)
                for (int a = 0; a < oracle_nao(la); ++a)
                    hrrs[HIDX(a, b, c, dplus)] = hrrs[HIDX(a, b, cplus, d)] + (C[j] - D[j]) * hrrs[HIDX(a, b, c, d)];
        }
    }
}

void oracle_hrr(double *hrrs, int la, int lb, int lc, int ld, double aexpn, double bexpn,
                double cexpn, double dexpn, const double *A, const double *B, const double *C,
                const double *D, int literal) {
    /* hrr(): vrr then HRR on the [1,:,:] slice, no contraction (HGP.jl:392-424) */
    int amax = la + lb, cmax = lc + ld, mmax = amax + cmax + 1;
    int na = oracle_nao(amax), nc = oracle_nao(cmax);
    double *v = (double *)malloc(sizeof(double) * (size_t)mmax * na * nc);
    oracle_vrr(v, amax, cmax, aexpn, bexpn, cexpn, dexpn, A, B, C, D, literal);
    hrr_from_vrr(hrrs, v, la, lb, lc, ld, A, B, C, D); /* m=0 slice is the first na*nc */
    free(v);
}

/* ------------------------------------------------------------------------ */
/* cvrr + chrr on shells -- src/HGP.jl:50-68 and :83-116.                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int nshell, nbf;
    const double *xyz;
    const int32_t *L, *poff;
    const double *pexp, *pcoef, *pnorm, *bfscale;
    int *bfoff; /* first function of each shell */
} flat_basis;

static int check_basis(flat_basis *fb) {
    init_tables();
    int n = 0;
    for (int s = 0; s < fb->nshell; ++s) {
        if (fb->L[s] < 0 || fb->L[s] > 4) return 1; /* reference tables: la+lb <= 4 needs L<=2...4 */
        if (fb->poff[s + 1] <= fb->poff[s]) return 2;
        fb->bfoff[s] = n;
        n += ncart(fb->L[s]);
    }
    fb->bfoff[fb->nshell] = n;
    return n == fb->nbf ? 0 : 3;
}

/* Contracted (e0|f0) over the primitive quartets, loop order d, c, b, a
 * (HGP.jl:57-58); coefficient = raw coef (literal D2) or coef * radial norm. */
static void cvrr_shells(double *cv, double *work, const flat_basis *fb, int sa, int sb, int sc,
                        int sd, int literal) {
    int amax = fb->L[sa] + fb->L[sb], cmax = fb->L[sc] + fb->L[sd];
    int na = oracle_nao(amax), nc = oracle_nao(cmax);
    memset(cv, 0, sizeof(double) * (size_t)na * nc);
    const double *A = fb->xyz + 3 * sa, *B = fb->xyz + 3 * sb, *C = fb->xyz + 3 * sc, *D = fb->xyz + 3 * sd;
    int raw = (literal & ORACLE_LIT_D2) != 0;
    for (int pd = fb->poff[sd]; pd < fb->poff[sd + 1]; ++pd)
        for (int pc = fb->poff[sc]; pc < fb->poff[sc + 1]; ++pc)
            for (int pb = fb->poff[sb]; pb < fb->poff[sb + 1]; ++pb)
                for (int pa = fb->poff[sa]; pa < fb->poff[sa + 1]; ++pa) {
                    oracle_vrr(work, amax, cmax, fb->pexp[pa], fb->pexp[pb], fb->pexp[pc], fb->pexp[pd], A, B, C,
                               D, literal);
                    double w = fb->pcoef[pa] * fb->pcoef[pb] * fb->pcoef[pc] * fb->pcoef[pd];
                    if (!raw) w *= fb->pnorm[pa] * fb->pnorm[pb] * fb->pnorm[pc] * fb->pnorm[pd];
                    for (size_t x = 0; x < (size_t)na * nc; ++x) cv[x] += w * work[x];
                }
}

static int64_t tri1(int64_t i, int64_t j) { /* triangle(i,j), 1-based, Utils.jl:15 */
    return i < j ? (j - 1) * j / 2 + i : (i - 1) * i / 2 + j;
}

int64_t oracle_n_unique(int64_t n) { return n * (n + 1) * (n * n + n + 2) / 8; }

int64_t oracle_slot(int64_t nbf, int64_t i, int64_t j, int64_t k, int64_t l, int order) {
    int64_t ij = tri1(i, j), kl = tri1(k, l);
    if (order == ORACLE_ORDER_IINDEX) return tri1(ij, kl) - 1; /* Utils.jl:19 */
    int64_t M = nbf * (nbf + 1) / 2;
    if (ij > kl) { int64_t t = ij; ij = kl; kl = t; }
    return (ij - 1) * (M + 1) - ij * (ij - 1) / 2 + kl - ij; /* enumerate(iiterator) - 1 */
}

/* The reference's fetcher keys are (ishell[i],ishell[j],ishell[k],ishell[l])
 * of canonical function quartets (Basis.jl:201-212).  A shell quartet
 * (I>=J, K>=L) is a key iff some element has tri(i,j) <= tri(k,l). */
static int is_key(const flat_basis *fb, int I, int J, int K, int Lsh) {
    int64_t i0 = fb->bfoff[I] + 1, j0 = fb->bfoff[J] + 1;
    int64_t k1 = fb->bfoff[K + 1], l1 = fb->bfoff[Lsh + 1];
    int64_t lmax = (K == Lsh) ? k1 : l1; /* k >= l */
    return tri1(i0, j0) <= tri1(k1, lmax);
}

static void process_key(const flat_basis *fb, int I, int J, int K, int Lsh, int order, int literal,
                        double *cv, double *work, double *hr, double *out, double *checksum) {
    int la = fb->L[I], lb = fb->L[J], lc = fb->L[K], ld = fb->L[Lsh];
    cvrr_shells(cv, work, fb, I, J, K, Lsh, literal);
    hrr_from_vrr(hr, cv, la, lb, lc, ld, fb->xyz + 3 * I, fb->xyz + 3 * J, fb->xyz + 3 * K, fb->xyz + 3 * Lsh);
    const int nb = oracle_nao(lb), nc = oracle_nao(lc + ld), nd = oracle_nao(ld);
    int64_t n = fb->nbf;
    int raw = (literal & ORACLE_LIT_D2) != 0;
    int d3 = (literal & ORACLE_LIT_D3) != 0;
    for (int ma = 0; ma < ncart(la); ++ma)
        for (int mb = 0; mb < ncart(lb); ++mb)
            for (int mc = 0; mc < ncart(lc); ++mc)
                for (int md = 0; md < ncart(ld); ++md) {
                    int64_t i = fb->bfoff[I] + ma + 1, j = fb->bfoff[J] + mb + 1;
                    int64_t k = fb->bfoff[K] + mc + 1, l = fb->bfoff[Lsh] + md + 1;
                    if (i < j || k < l || tri1(i, j) > tri1(k, l)) continue;
                    int ha, hb, hc, hd;
                    if (d3) { /* HGP.jl:29-30 with mshell (1-based position in shell) */
                        ha = ma; hb = mb; hc = mc; hd = md;
                    } else {
                        ha = oracle_nao(la - 1) + ma;
                        hb = oracle_nao(lb - 1) + mb;
                        hc = oracle_nao(lc - 1) + mc;
                        hd = oracle_nao(ld - 1) + md;
                    }
                    double v = hr[(((size_t)ha * nb + hb) * nc + hc) * nd + hd];
                    if (!raw)
                        v *= fb->bfscale[i - 1] * fb->bfscale[j - 1] * fb->bfscale[k - 1] * fb->bfscale[l - 1];
                    if (out) out[oracle_slot(n, i, j, k, l, order)] = v;
                    if (checksum) *checksum += v;
                }
}

static size_t work_sizes(int maxl, size_t *cvn, size_t *wn, size_t *hn) {
    int amax = 2 * maxl, mmax = 4 * maxl + 1;
    size_t na = oracle_nao(amax), nb = oracle_nao(maxl);
    *cvn = na * na;
    *wn = (size_t)mmax * na * na;
    *hn = na * nb * na * nb;
    return *cvn + *wn + *hn;
}

int oracle_eri_unique_hgp(int nshell, const double *shell_xyz, const int32_t *shell_l,
                          const int32_t *shell_prim_off, const double *prim_exp,
                          const double *prim_coef, const double *prim_norm, int nbf,
                          const double *bf_scale, int order, int literal, int nthreads, double *out) {
    flat_basis fb = {nshell, nbf, shell_xyz, shell_l, shell_prim_off, prim_exp, prim_coef, prim_norm, bf_scale, NULL};
    fb.bfoff = (int *)malloc(sizeof(int) * (nshell + 1));
    int rc = check_basis(&fb);
    if (rc) { free(fb.bfoff); return rc; }
    int maxl = 0;
    for (int s = 0; s < nshell; ++s) maxl = shell_l[s] > maxl ? shell_l[s] : maxl;
    /* Key list (the reference materialises it through eri_fetcher). */
    int64_t npairs = (int64_t)nshell * (nshell + 1) / 2, nkeys = 0;
    int32_t *keys = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)npairs * npairs);
    for (int I = 0; I < nshell; ++I)
        for (int J = 0; J <= I; ++J)
            for (int K = 0; K < nshell; ++K)
                for (int Lsh = 0; Lsh <= K; ++Lsh)
                    if (is_key(&fb, I, J, K, Lsh)) {
                        int32_t *q = keys + 4 * nkeys++;
                        q[0] = I; q[1] = J; q[2] = K; q[3] = Lsh;
                    }
    memset(out, 0, sizeof(double) * (size_t)oracle_n_unique(nbf));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        size_t cvn, wn, hn;
        work_sizes(maxl, &cvn, &wn, &hn);
        double *buf = (double *)malloc(sizeof(double) * (cvn + wn + hn));
#pragma omp for schedule(dynamic, 16)
        for (int64_t q = 0; q < nkeys; ++q)
            process_key(&fb, keys[4 * q], keys[4 * q + 1], keys[4 * q + 2], keys[4 * q + 3], order, literal, buf,
                        buf + cvn, buf + cvn + wn, out, NULL);
        free(buf);
    }
    free(keys);
    free(fb.bfoff);
    return 0;
}

int64_t oracle_eri_hgp_quartet_sample(int nshell, const double *shell_xyz, const int32_t *shell_l,
                                      const int32_t *shell_prim_off, const double *prim_exp,
                                      const double *prim_coef, const double *prim_norm, int nbf,
                                      const double *bf_scale, int64_t nq, const int32_t *quartets,
                                      int nthreads, double *checksum) {
    flat_basis fb = {nshell, nbf, shell_xyz, shell_l, shell_prim_off, prim_exp, prim_coef, prim_norm, bf_scale, NULL};
    fb.bfoff = (int *)malloc(sizeof(int) * (nshell + 1));
    if (check_basis(&fb)) { free(fb.bfoff); return -1; }
    int maxl = 0;
    for (int s = 0; s < nshell; ++s) maxl = shell_l[s] > maxl ? shell_l[s] : maxl;
    int64_t count = 0;
    double total = 0.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : count, total)
    {
        size_t cvn, wn, hn;
        work_sizes(maxl, &cvn, &wn, &hn);
        double *buf = (double *)malloc(sizeof(double) * (cvn + wn + hn));
#pragma omp for schedule(dynamic, 16)
        for (int64_t q = 0; q < nq; ++q) {
            const int32_t *k = quartets + 4 * q;
            double cs = 0.0;
            process_key(&fb, k[0], k[1], k[2], k[3], ORACLE_ORDER_CHRR, 0, buf, buf + cvn, buf + cvn + wn, NULL, &cs);
            total += cs;
            count += (int64_t)ncart(shell_l[k[0]]) * ncart(shell_l[k[1]]) * ncart(shell_l[k[2]]) * ncart(shell_l[k[3]]);
        }
        free(buf);
    }
    if (checksum) *checksum = total;
    free(fb.bfoff);
    return count;
}

/* ------------------------------------------------------------------------ */
/* Overlap and normalisation -- src/OneInts.jl:11-94, src/Basis.jl:29-83.      */
/* ------------------------------------------------------------------------ */
static double binomial_prefactor(int s, int ia, int ib, double xpa, double xpb) {
    /* OneInts.jl:92-93 */
    double sum = 0.0;
    for (int t = 0; t <= s; ++t)
        if (s - ia <= t && t <= ib)
            sum += binomial(ia, s - t) * ipow(xpa, ia - (s - t)) * binomial(ib, t) * ipow(xpb, ib - t);
    return sum;
}
static double overlap1d(int la, int lb, double ax, double bx, double gamma) {
    double total = 0.0; /* OneInts.jl:84-90 */
    for (int i = 0; i <= (la + lb) / 2; ++i)
        total += binomial_prefactor(2 * i, la, lb, ax, bx) * factorial2(2 * i - 1) / ipow(2 * gamma, i);
    return total;
}
double oracle_overlap_prim(double aexpn, const double *A, int aI, int aJ, int aK, double bexpn,
                           const double *B, int bI, int bJ, int bK) {
    double gamma = aexpn + bexpn; /* OneInts.jl:41-52 */
    double P[3], pa[3], pb[3];
    for (int k = 0; k < 3; ++k) {
        P[k] = (aexpn / gamma) * A[k] + (bexpn / gamma) * B[k];
        pa[k] = P[k] - A[k];
        pb[k] = P[k] - B[k];
    }
    double pre = pow(M_PI / gamma, 1.5) * exp(-aexpn * bexpn * dist2(A, B) / gamma);
    return pre * overlap1d(aI, bI, pa[0], pb[0], gamma) * overlap1d(aJ, bJ, pa[1], pb[1], gamma) *
           overlap1d(aK, bK, pa[2], pb[2], gamma);
}
double oracle_pgbf_norm(double expn, int I, int J, int K) {
    const double O[3] = {0, 0, 0};
    return 1.0 / sqrt(oracle_overlap_prim(expn, O, I, J, K, expn, O, I, J, K));
}
double oracle_cgbf_norm(int nprim, const double *expn, const double *coef, int I, int J, int K) {
    const double O[3] = {0, 0, 0};
    double s = 0.0;
    for (int p = 0; p < nprim; ++p)
        for (int q = 0; q < nprim; ++q)
            s += coef[p] * coef[q] * oracle_pgbf_norm(expn[p], I, J, K) * oracle_pgbf_norm(expn[q], I, J, K) *
                 oracle_overlap_prim(expn[p], O, I, J, K, expn[q], O, I, J, K);
    return 1.0 / sqrt(s);
}

/* ------------------------------------------------------------------------ */
/* THO/Huzinaga primitive ERI -- src/ERI.jl:15-94.                          */
/* ------------------------------------------------------------------------ */
static double fact_ratio2(int a, int b) { /* prod((b+1):a)/factorial(a-2b), ERI.jl:65 */
    double p = 1.0;
    for (int k = b + 1; k <= a; ++k) p *= k;
    return p / factorial(a - 2 * b);
}
static double B0(int i, int r, double g) { return fact_ratio2(i, r) * pow(4 * g, r - i); }
static double fB(int i, int l1, int l2, double p, double a, double b, int r, double g) {
    return binomial_prefactor(i, l1, l2, p - a, p - b) * B0(i, r, g);
}
static double Bterm(int i1, int i2, int r1, int r2, int u, int l1, int l2, int l3, int l4, double Px, double Ax,
                    double Bx, double Qx, double Cx, double Dx, double g1, double g2, double delta) {
    int e = i1 + i2 - 2 * (r1 + r2); /* ERI.jl:67-75 */
    return fB(i1, l1, l2, Px, Ax, Bx, r1, g1) * ((i2 & 1) ? -1.0 : 1.0) * fB(i2, l3, l4, Qx, Cx, Dx, r2, g2) *
           ((u & 1) ? -1.0 : 1.0) * fact_ratio2(e, u) * ipow(Qx - Px, e - 2 * u) / ipow(delta, e - u);
}
static void Barray(double *B, int l1, int l2, int l3, int l4, double p, double a, double b, double q, double c,
                   double d, double g1, double g2, double delta) {
    int Imax = l1 + l2 + l3 + l4 + 1; /* ERI.jl:77-94 */
    for (int I = 0; I < Imax; ++I) B[I] = 0.0;
    for (int i1 = 0; i1 <= l1 + l2; ++i1)
        for (int i2 = 0; i2 <= l3 + l4; ++i2)
            for (int r1 = 0; r1 <= i1 / 2; ++r1)
                for (int r2 = 0; r2 <= i2 / 2; ++r2)
                    for (int u = 0; u <= (i1 + i2) / 2 - r1 - r2; ++u) {
                        int I = i1 + i2 - 2 * (r1 + r2) - u;
                        B[I] += Bterm(i1, i2, r1, r2, u, l1, l2, l3, l4, p, a, b, q, c, d, g1, g2, delta);
                    }
}
double oracle_coulomb_prim(double aexpn, const double *A, int aI, int aJ, int aK, double bexpn, const double *B,
                           int bI, int bJ, int bK, double cexpn, const double *C, int cI, int cJ, int cK,
                           double dexpn, const double *D, int dI, int dJ, int dK) {
    double rab2 = dist2(A, B), rcd2 = dist2(C, D);
    double g1 = aexpn + bexpn, g2 = cexpn + dexpn;
    double P[3], Q[3];
    for (int k = 0; k < 3; ++k) {
        P[k] = (aexpn / g1) * A[k] + (bexpn / g1) * B[k];
        Q[k] = (cexpn / g2) * C[k] + (dexpn / g2) * D[k];
    }
    double rpq2 = dist2(P, Q);
    double delta = 0.25 * (1 / g1 + 1 / g2);
    double Bx[32], By[32], Bz[32];
    Barray(Bx, aI, bI, cI, dI, P[0], A[0], B[0], Q[0], C[0], D[0], g1, g2, delta);
    Barray(By, aJ, bJ, cJ, dJ, P[1], A[1], B[1], Q[1], C[1], D[1], g1, g2, delta);
    Barray(Bz, aK, bK, cK, dK, P[2], A[2], B[2], Q[2], C[2], D[2], g1, g2, delta);
    double x = 0.25 * rpq2 / delta;
    double s = 0.0;
    for (int I = 0; I <= aI + bI + cI + dI; ++I)
        for (int J = 0; J <= aJ + bJ + cJ + dJ; ++J)
            for (int K = 0; K <= aK + bK + cK + dK; ++K) s += Bx[I] * By[J] * Bz[K] * oracle_Fm(I + J + K, x);
    return 2 * pow(M_PI, 2.5) / (g1 * g2 * sqrt(g1 + g2)) * exp(-aexpn * bexpn * rab2 / g1) *
           exp(-cexpn * dexpn * rcd2 / g2) * s;
}

int oracle_eri_unique_tho(int nshell, const double *shell_xyz, const int32_t *shell_l,
                          const int32_t *shell_prim_off, const double *prim_exp, const double *prim_coef,
                          const double *prim_norm, int nbf, const double *bf_scale, int order, double *out) {
    flat_basis fb = {nshell, nbf, shell_xyz, shell_l, shell_prim_off, prim_exp, prim_coef, prim_norm, bf_scale, NULL};
    fb.bfoff = (int *)malloc(sizeof(int) * (nshell + 1));
    int rc = check_basis(&fb);
    if (rc) { free(fb.bfoff); return rc; }
    int *bsh = (int *)malloc(sizeof(int) * nbf), *bm = (int *)malloc(sizeof(int) * nbf);
    for (int s = 0; s < nshell; ++s)
        for (int m = 0; m < ncart(shell_l[s]); ++m) {
            bsh[fb.bfoff[s] + m] = s;
            bm[fb.bfoff[s] + m] = oracle_nao(shell_l[s] - 1) + m;
        }
    /* all_twoe_ints: loop over iiterator (Utils.jl:17), coulomb(CGBF..) via
     * contract (Basis.jl:86-98) with per-primitive norms n_p (here
     * prim_norm * component factor, folded into bf_scale). */
    for (int64_t i = 1; i <= nbf; ++i)
        for (int64_t j = 1; j <= i; ++j)
            for (int64_t k = 1; k <= nbf; ++k)
                for (int64_t l = 1; l <= k; ++l) {
                    if (tri1(i, j) > tri1(k, l)) continue;
                    int64_t f[4] = {i - 1, j - 1, k - 1, l - 1};
                    int sh[4], I[4][3];
                    for (int t = 0; t < 4; ++t) {
                        sh[t] = bsh[f[t]];
                        int32_t m[3];
                        oracle_ao2m(bm[f[t]], m);
                        I[t][0] = m[0]; I[t][1] = m[1]; I[t][2] = m[2];
                    }
                    double s = 0.0;
                    for (int pa = shell_prim_off[sh[0]]; pa < shell_prim_off[sh[0] + 1]; ++pa)
                        for (int pb = shell_prim_off[sh[1]]; pb < shell_prim_off[sh[1] + 1]; ++pb)
                            for (int pc = shell_prim_off[sh[2]]; pc < shell_prim_off[sh[2] + 1]; ++pc)
                                for (int pd = shell_prim_off[sh[3]]; pd < shell_prim_off[sh[3] + 1]; ++pd) {
                                    double w = prim_coef[pa] * prim_coef[pb] * prim_coef[pc] * prim_coef[pd] *
                                               prim_norm[pa] * prim_norm[pb] * prim_norm[pc] * prim_norm[pd];
                                    s += w * oracle_coulomb_prim(
                                                 prim_exp[pa], shell_xyz + 3 * sh[0], I[0][0], I[0][1], I[0][2],
                                                 prim_exp[pb], shell_xyz + 3 * sh[1], I[1][0], I[1][1], I[1][2],
                                                 prim_exp[pc], shell_xyz + 3 * sh[2], I[2][0], I[2][1], I[2][2],
                                                 prim_exp[pd], shell_xyz + 3 * sh[3], I[3][0], I[3][1], I[3][2]);
                                }
                    s *= bf_scale[f[0]] * bf_scale[f[1]] * bf_scale[f[2]] * bf_scale[f[3]];
                    out[oracle_slot(nbf, i, j, k, l, order)] = s;
                }
    free(bsh);
    free(bm);
    free(fb.bfoff);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* J/K from the unique array -- index algebra of make2JmK, src/ERI.jl:114-128. */
/* ------------------------------------------------------------------------ */
void oracle_jk_from_unique(int nbf, const double *eri, const double *D, double *J, double *K) {
    int64_t n = nbf;
    for (int64_t i = 1; i <= n; ++i)
        for (int64_t j = 1; j <= n; ++j) {
            double sj = 0.0, sk = 0.0;
            for (int64_t k = 1; k <= n; ++k)
                for (int64_t l = 1; l <= n; ++l) {
                    double dkl = D[(k - 1) * n + (l - 1)];
                    sj += dkl * eri[tri1(tri1(i, j), tri1(k, l)) - 1];
                    sk += dkl * eri[tri1(tri1(i, k), tri1(j, l)) - 1];
                }
            J[(i - 1) * n + (j - 1)] = sj;
            K[(i - 1) * n + (j - 1)] = sk;
        }
}

int oracle_schwarz_pairs(int nshell, const double *shell_xyz, const int32_t *shell_l,
                         const int32_t *shell_prim_off, const double *prim_exp, const double *prim_coef,
                         const double *prim_norm, int nbf, const double *bf_scale, double *q_pair) {
    flat_basis fb = {nshell, nbf, shell_xyz, shell_l, shell_prim_off, prim_exp, prim_coef, prim_norm, bf_scale, NULL};
    fb.bfoff = (int *)malloc(sizeof(int) * (nshell + 1));
    int rc = check_basis(&fb);
    if (rc) { free(fb.bfoff); return rc; }
    int maxl = 0;
    for (int s = 0; s < nshell; ++s) maxl = shell_l[s] > maxl ? shell_l[s] : maxl;
    size_t cvn, wn, hn;
    work_sizes(maxl, &cvn, &wn, &hn);
    double *buf = (double *)malloc(sizeof(double) * (cvn + wn + hn));
    for (int I = 0; I < nshell; ++I)
        for (int J = 0; J <= I; ++J) {
            int la = shell_l[I], lb = shell_l[J];
            cvrr_shells(buf, buf + cvn, &fb, I, J, I, J, 0);
            double *hr = buf + cvn + wn;
            hrr_from_vrr(hr, buf, la, lb, la, lb, shell_xyz + 3 * I, shell_xyz + 3 * J, shell_xyz + 3 * I,
                         shell_xyz + 3 * J);
            int nb = oracle_nao(lb), nc = oracle_nao(la + lb), nd = oracle_nao(lb);
            double mx = 0.0;
            for (int ma = 0; ma < ncart(la); ++ma)
                for (int mb = 0; mb < ncart(lb); ++mb) {
                    int ha = oracle_nao(la - 1) + ma, hb = oracle_nao(lb - 1) + mb;
                    double v = fabs(hr[(((size_t)ha * nb + hb) * nc + ha) * nd + hb]) * bf_scale[fb.bfoff[I] + ma] *
                               bf_scale[fb.bfoff[I] + ma] * bf_scale[fb.bfoff[J] + mb] * bf_scale[fb.bfoff[J] + mb];
                    if (v > mx) mx = v;
                }
            q_pair[(int64_t)I * (I + 1) / 2 + J] = sqrt(mx);
        }
    free(buf);
    free(fb.bfoff);
    return 0;
}
