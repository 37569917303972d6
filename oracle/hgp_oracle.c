/*
 * hgp_oracle.c -- CPU restatement of MolecularIntegrals.jl's HGP ERI path.
 *
 * TEST INFRASTRUCTURE ONLY (see hgp_oracle.h): the checker for the HIP path and
 * the CPU baseline timed beside it.  Never linked into the product library.
 *
 * Citations are to files of rpmuller/MolecularIntegrals.jl.
 * Parity status of each part (how it is pinned) is listed in the header
 * comment of tests/test_oracle.py.
 */
#include "hgp_oracle.h"

#include <math.h>
#include <quadmath.h>
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
static long double boys_top_ld(int m, long double x) {
    const long double SMALL = 1e-18L;
    if (x < SMALL) return 1.0L / (2 * m + 1);
    long double a = (long double)m + 0.5L;
    if (x >= 200) { /* Gamma(m+1/2) / (2 x^{m+1/2}); the dropped tail is < e^{-200} x^{m-1/2}/Gamma(m+1/2) */
        long double g = sqrtl(3.141592653589793238462643383279502884L);
        for (int k = 1; k <= m; ++k) g *= ((long double)k - 0.5L);
        return g / (2 * powl(x, a));
    }
    long double term = 1.0L / a;
    long double sum = term;
    for (int n = 1; n < 10000; ++n) {
        term *= x / (a + n);
        sum += term;
        if (term < sum * 1e-21L) break;
    }
    return 0.5L * expl(-x) * sum;
}

double oracle_Fm(int m, double T) { return (double)boys_top_ld(m, (long double)T); }

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


/* The recurrences, instantiated in FP64 and in x87 extended precision. */
#define BOYS_TOP_(m, T) ((REAL)boys_top_ld((m), (long double)(T)))
#define REAL double
#define SFX(name) name##_d
#define RSQRT_(x) sqrt(x)
#define REXP_(x) exp(x)
#define RPOW_(x, y) pow(x, y)
#define RPI M_PI
#include "hgp_core.inc"
#undef REAL
#undef SFX
#undef RSQRT_
#undef REXP_
#undef RPOW_
#undef RPI
#undef BOYS_TOP_
#define BOYS_TOP_(m, T) ((REAL)boys_top_ld((m), (long double)(T)))
#define REAL long double
#define SFX(name) name##_x
#define RSQRT_(x) sqrtl(x)
#define REXP_(x) expl(x)
#define RPOW_(x, y) powl(x, y)
#define RPI 3.141592653589793238462643383279502884L
#include "hgp_core.inc"
#undef REAL
#undef SFX
#undef RSQRT_
#undef REXP_
#undef RPOW_
#undef RPI
#undef BOYS_TOP_

/* binary128 instantiation: the recurrences in quad precision.  Boys:
 * F_m(T) = e^{-T}/2 sum_n T^n / prod_{k=0..n} (m+1/2+k) (the series of
 * boys_top_ld, all terms positive) for T < 200; above, F_m(T) =
 * Gamma(m+1/2)/(2 T^{m+1/2}) - Gamma(m+1/2,T)/(2 T^{m+1/2}), whose second term is
 * below e^{-200} T^{m-1/2} / Gamma(m+1/2) relative and is dropped. */
static __float128 boys_top_q(int m, __float128 x) {
    if (x < 1e-18Q) return 1 / (__float128)(2 * m + 1);
    const __float128 a = (__float128)m + 0.5Q;
    if (x >= 200) {
        /* Gamma(m+1/2) = (2m-1)!! sqrt(pi) / 2^m */
        __float128 g = sqrtq(M_PIq);
        for (int k = 1; k <= m; ++k) g *= ((__float128)k - 0.5Q);
        return g / (2 * powq(x, a));
    }
    __float128 term = 1 / a, sum = term;
    for (int n = 1; n < 100000; ++n) {
        term *= x / (a + n);
        sum += term;
        if (term < sum * 1e-36Q) break;
    }
    return 0.5Q * expq(-x) * sum;
}
#define BOYS_TOP_(m, T) boys_top_q((m), (T))
#define REAL __float128
#define SFX(name) name##_q
#define RSQRT_(x) sqrtq(x)
#define REXP_(x) expq(x)
#define RPOW_(x, y) powq(x, y)
#define RPI M_PIq
#include "hgp_core.inc"
#undef REAL
#undef SFX
#undef RSQRT_
#undef REXP_
#undef RPOW_
#undef RPI
#undef BOYS_TOP_

double oracle_Fm_quad(int m, double T, int accurate) {
    __float128 F[64];
    farray_q(m + 1, (__float128)T, F, !accurate);
    return (double)F[m];
}

void oracle_farray(int mmax, double T, double *Fms) { farray_d(mmax, T, Fms, 1); }
void oracle_farray_accurate(int mmax, double T, double *Fms) { farray_d(mmax, T, Fms, 0); }

void oracle_vrr(double *vrrs, int amax, int cmax, double aexpn, double bexpn, double cexpn, double dexpn,
                const double *A, const double *B, const double *C, const double *D, int literal) {
    vrr_d(vrrs, amax, cmax, aexpn, bexpn, cexpn, dexpn, A, B, C, D, literal);
}

void oracle_hrr(double *hrrs, int la, int lb, int lc, int ld, double aexpn, double bexpn, double cexpn,
                double dexpn, const double *A, const double *B, const double *C, const double *D, int literal) {
    /* hrr(): vrr then HRR on the [1,:,:] slice, no contraction (HGP.jl:392-424) */
    int amax = la + lb, cmax = lc + ld, mmax = amax + cmax + 1;
    int na = oracle_nao(amax), nc = oracle_nao(cmax);
    double *v = (double *)malloc(sizeof(double) * (size_t)mmax * na * nc);
    vrr_d(v, amax, cmax, aexpn, bexpn, cexpn, dexpn, A, B, C, D, literal);
    hrr_from_vrr_d(hrrs, v, la, lb, lc, ld, A, B, C, D); /* m=0 slice is the first na*nc */
    free(v);
}

static size_t work_sizes(int maxl, size_t *cvn, size_t *wn, size_t *hn) {
    int amax = 2 * maxl, mmax = 4 * maxl + 1;
    size_t na = oracle_nao(amax), nb = oracle_nao(maxl);
    *cvn = na * na;
    *wn = (size_t)mmax * na * na;
    *hn = na * nb * na * nb;
    return *cvn + *wn + *hn;
}

static int eri_unique_hgp_impl(int nshell, const double *shell_xyz, const int32_t *shell_l,
                               const int32_t *shell_prim_off, const double *prim_exp,
                               const double *prim_coef, const double *prim_norm, int nbf,
                               const double *bf_scale, int order, int literal, int nthreads, double *out,
                               int extended) {
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
    if (extended == 2) literal |= ORACLE_LIT_D5;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        size_t cvn, wn, hn;
        work_sizes(maxl, &cvn, &wn, &hn);
        if (extended >= 2) {
            __float128 *buf = (__float128 *)malloc(sizeof(__float128) * (cvn + wn + hn));
#pragma omp for schedule(dynamic, 4)
            for (int64_t q = 0; q < nkeys; ++q)
                process_key_q(&fb, keys[4 * q], keys[4 * q + 1], keys[4 * q + 2], keys[4 * q + 3], order, literal,
                              buf, buf + cvn, buf + cvn + wn, out, NULL);
            free(buf);
        } else if (extended) {
            long double *buf = (long double *)malloc(sizeof(long double) * (cvn + wn + hn));
#pragma omp for schedule(dynamic, 16)
            for (int64_t q = 0; q < nkeys; ++q)
                process_key_x(&fb, keys[4 * q], keys[4 * q + 1], keys[4 * q + 2], keys[4 * q + 3], order, literal,
                              buf, buf + cvn, buf + cvn + wn, out, NULL);
            free(buf);
        } else {
            double *buf = (double *)malloc(sizeof(double) * (cvn + wn + hn));
#pragma omp for schedule(dynamic, 16)
            for (int64_t q = 0; q < nkeys; ++q)
                process_key_d(&fb, keys[4 * q], keys[4 * q + 1], keys[4 * q + 2], keys[4 * q + 3], order, literal,
                              buf, buf + cvn, buf + cvn + wn, out, NULL);
            free(buf);
        }
    }
    free(keys);
    free(fb.bfoff);
    return 0;
}

int oracle_eri_unique_hgp(int nshell, const double *shell_xyz, const int32_t *shell_l,
                          const int32_t *shell_prim_off, const double *prim_exp, const double *prim_coef,
                          const double *prim_norm, int nbf, const double *bf_scale, int order, int literal,
                          int nthreads, double *out) {
    return eri_unique_hgp_impl(nshell, shell_xyz, shell_l, shell_prim_off, prim_exp, prim_coef, prim_norm, nbf,
                               bf_scale, order, literal, nthreads, out, 0);
}

int oracle_eri_unique_hgp_ext(int nshell, const double *shell_xyz, const int32_t *shell_l,
                              const int32_t *shell_prim_off, const double *prim_exp, const double *prim_coef,
                              const double *prim_norm, int nbf, const double *bf_scale, int order, int nthreads,
                              double *out) {
    return eri_unique_hgp_impl(nshell, shell_xyz, shell_l, shell_prim_off, prim_exp, prim_coef, prim_norm, nbf,
                               bf_scale, order, 0, nthreads, out, 1);
}

int oracle_eri_unique_hgp_quad(int nshell, const double *shell_xyz, const int32_t *shell_l,
                               const int32_t *shell_prim_off, const double *prim_exp, const double *prim_coef,
                               const double *prim_norm, int nbf, const double *bf_scale, int order, int accurate,
                               int nthreads, double *out) {
    return eri_unique_hgp_impl(nshell, shell_xyz, shell_l, shell_prim_off, prim_exp, prim_coef, prim_norm, nbf,
                               bf_scale, order, 0, nthreads, out, accurate ? 3 : 2);
}

/* Selected slots of the unique array in binary128 (arbiter of GPU parity
 * tests where the FP64 oracle loses digits): each slot is decoded to its
 * function quartet (i >= j, k >= l, tri(i,j) <= tri(k,l)), whose shell
 * quartet block (cvrr + chrr, src/HGP.jl:50-116) is computed once per
 * distinct shell quartet. */
typedef struct { int64_t key, idx; } slot_job;
static int cmp_job(const void *a, const void *b) {
    int64_t x = ((const slot_job *)a)->key, y = ((const slot_job *)b)->key;
    return x < y ? -1 : (x > y ? 1 : 0);
}
static void slot_decode(int64_t nbf, int64_t slot, int order, int64_t f[4]) {
    int64_t M = nbf * (nbf + 1) / 2, ij, kl;
    if (order == ORACLE_ORDER_IINDEX) { /* slot + 1 = kl (kl - 1) / 2 + ij, ij <= kl */
        kl = (int64_t)((sqrt(8.0 * (double)slot + 1.0) + 1.0) / 2.0);
        while (kl * (kl - 1) / 2 > slot) --kl;
        while ((kl + 1) * kl / 2 <= slot) ++kl;
        ij = slot - kl * (kl - 1) / 2 + 1;
    } else { /* slot = (ij - 1)(M + 1) - ij (ij - 1) / 2 + kl - ij */
        int64_t lo = 1, hi = M;
        while (lo < hi) {
            int64_t mid = (lo + hi + 1) / 2;
            if ((mid - 1) * (M + 1) - mid * (mid - 1) / 2 <= slot) lo = mid; else hi = mid - 1;
        }
        ij = lo;
        kl = slot - ((ij - 1) * (M + 1) - ij * (ij - 1) / 2) + ij;
    }
    int64_t p[2] = {ij, kl};
    for (int t = 0; t < 2; ++t) {
        int64_t a = (int64_t)((sqrt(8.0 * (double)p[t]) + 1.0) / 2.0);
        while (a * (a - 1) / 2 >= p[t]) --a;
        while (a * (a + 1) / 2 < p[t]) ++a;
        f[2 * t] = a;
        f[2 * t + 1] = p[t] - a * (a - 1) / 2;
    }
}
static int shell_of(const flat_basis *fb, int64_t f1) { /* 1-based function -> shell */
    int lo = 0, hi = fb->nshell - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (fb->bfoff[mid] < f1) lo = mid; else hi = mid - 1;
    }
    return lo;
}
int oracle_eri_slots_quad(int nshell, const double *shell_xyz, const int32_t *shell_l,
                          const int32_t *shell_prim_off, const double *prim_exp, const double *prim_coef,
                          const double *prim_norm, int nbf, const double *bf_scale, int order, int accurate,
                          int64_t nslot, const int64_t *slots, int nthreads, double *out) {
    flat_basis fb = {nshell, nbf, shell_xyz, shell_l, shell_prim_off, prim_exp, prim_coef, prim_norm, bf_scale, NULL};
    fb.bfoff = (int *)malloc(sizeof(int) * (nshell + 1));
    int rc = check_basis(&fb);
    if (rc) { free(fb.bfoff); return rc; }
    int maxl = 0;
    for (int s2 = 0; s2 < nshell; ++s2) maxl = shell_l[s2] > maxl ? shell_l[s2] : maxl;
    slot_job *jobs = (slot_job *)malloc(sizeof(slot_job) * (size_t)(nslot > 0 ? nslot : 1));
    int64_t ns = nshell;
    for (int64_t q = 0; q < nslot; ++q) {
        int64_t f[4];
        if (slots[q] < 0 || slots[q] >= oracle_n_unique(nbf)) { free(jobs); free(fb.bfoff); return 3; }
        slot_decode(nbf, slots[q], order, f);
        int64_t key = 0;
        for (int t = 0; t < 4; ++t) key = key * ns + shell_of(&fb, f[t]);
        jobs[q].key = key;
        jobs[q].idx = q;
    }
    qsort(jobs, (size_t)nslot, sizeof(slot_job), cmp_job);
    int64_t ngroup = 0;
    int64_t *gstart = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nslot + 1));
    for (int64_t q = 0; q < nslot; ++q)
        if (q == 0 || jobs[q].key != jobs[q - 1].key) gstart[ngroup++] = q;
    gstart[ngroup] = nslot;
    const int literal = accurate ? 0 : ORACLE_LIT_D5;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        size_t cvn, wn, hn;
        work_sizes(maxl, &cvn, &wn, &hn);
        __float128 *buf = (__float128 *)malloc(sizeof(__float128) * (cvn + wn + hn));
#pragma omp for schedule(dynamic, 1)
        for (int64_t g = 0; g < ngroup; ++g) {
            int64_t key = jobs[gstart[g]].key;
            int sh[4];
            for (int t = 3; t >= 0; --t) { sh[t] = (int)(key % ns); key /= ns; }
            int la = shell_l[sh[0]], lb = shell_l[sh[1]], lc = shell_l[sh[2]], ld = shell_l[sh[3]];
            cvrr_shells_q(buf, buf + cvn, &fb, sh[0], sh[1], sh[2], sh[3], literal);
            __float128 A[3], B[3], C[3], D[3];
            for (int k = 0; k < 3; ++k) {
                A[k] = shell_xyz[3 * sh[0] + k];
                B[k] = shell_xyz[3 * sh[1] + k];
                C[k] = shell_xyz[3 * sh[2] + k];
                D[k] = shell_xyz[3 * sh[3] + k];
            }
            __float128 *hr = buf + cvn + wn;
            hrr_from_vrr_q(hr, buf, la, lb, lc, ld, A, B, C, D);
            const int nb = oracle_nao(lb), nc = oracle_nao(lc + ld), nd = oracle_nao(ld);
            for (int64_t q = gstart[g]; q < gstart[g + 1]; ++q) {
                int64_t f[4];
                slot_decode(nbf, slots[jobs[q].idx], order, f);
                int h[4];
                for (int t = 0; t < 4; ++t) h[t] = oracle_nao(shell_l[sh[t]] - 1) + (int)(f[t] - 1 - fb.bfoff[sh[t]]);
                __float128 v = hr[(((size_t)h[0] * nb + h[1]) * nc + h[2]) * nd + h[3]];
                v *= (__float128)bf_scale[f[0] - 1] * bf_scale[f[1] - 1] * bf_scale[f[2] - 1] * bf_scale[f[3] - 1];
                out[jobs[q].idx] = (double)v;
            }
        }
        free(buf);
    }
    free(gstart);
    free(jobs);
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
            process_key_d(&fb, k[0], k[1], k[2], k[3], ORACLE_ORDER_CHRR, 0, buf, buf + cvn, buf + cvn + wn, NULL, &cs);
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
/* Rys quadrature -- src/Rys.jl.                                              */
/* ------------------------------------------------------------------------ */
/* The n-point Rys rule at X (Roots(n, X, roots, weights), src/Rys.jl:161-174).
 * The reference evaluates fitted polynomials (Root1..Root5, :176-1480); here
 * the rule is computed from its definition, the Gauss rule of the weight
 * e^{-X x} x^{-1/2}/2 on x = t^2 in [0, 1], whose moments are F_k(X): in
 * binary128, moments by the accurate Boys function, the three-term
 * recurrence by Chebyshev's algorithm, then the Golub-Welsch eigenproblem of
 * the Jacobi matrix by cyclic Jacobi rotations (nodes = eigenvalues, weights
 * = F_0(X) (first eigenvector component)^2).  Roots are returned as the
 * reference's u = x / (1 - x) (RecurFactors, :141-151, takes u). */
#define ORACLE_RYS_NMAX 10
static void rys_rule_q(int n, __float128 X, __float128 *x, __float128 *w) {
    __float128 mu[2 * ORACLE_RYS_NMAX];
    farray_q(2 * n, X, mu, 0); /* F_0 .. F_{2n-1}, accurate for every X */
    __float128 alpha[ORACLE_RYS_NMAX], beta[ORACLE_RYS_NMAX];
    __float128 sm[2 * ORACLE_RYS_NMAX], s0[2 * ORACLE_RYS_NMAX], s1[2 * ORACLE_RYS_NMAX];
    for (int l = 0; l < 2 * n; ++l) { sm[l] = 0; s0[l] = mu[l]; }
    alpha[0] = mu[1] / mu[0];
    beta[0] = mu[0];
    for (int k = 1; k < n; ++k) {
        for (int l = k; l < 2 * n - k; ++l) s1[l] = s0[l + 1] - alpha[k - 1] * s0[l] - beta[k - 1] * sm[l];
        alpha[k] = s1[k + 1] / s1[k] - s0[k] / s0[k - 1];
        beta[k] = s1[k] / s0[k - 1];
        for (int l = 0; l < 2 * n; ++l) { sm[l] = s0[l]; s0[l] = s1[l]; }
    }
    /* symmetric Jacobi matrix J, eigenvectors V (columns) by cyclic Jacobi */
    __float128 A[ORACLE_RYS_NMAX][ORACLE_RYS_NMAX], V[ORACLE_RYS_NMAX][ORACLE_RYS_NMAX];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            A[i][j] = i == j ? alpha[i] : (j == i + 1 ? sqrtq(beta[j]) : (i == j + 1 ? sqrtq(beta[i]) : 0));
            V[i][j] = i == j;
        }
    for (int sweep = 0; sweep < 100; ++sweep) {
        __float128 off = 0, diag = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                if (i != j) off += A[i][j] * A[i][j];
                else diag += A[i][j] * A[i][j];
            }
        if (off <= 1e-66Q * diag) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                if (A[p][q] == 0) continue;
                const __float128 th = (A[q][q] - A[p][p]) / (2 * A[p][q]);
                const __float128 t = (th >= 0 ? 1 : -1) / (fabsq(th) + sqrtq(th * th + 1));
                const __float128 c = 1 / sqrtq(t * t + 1), sn = t * c;
                for (int k = 0; k < n; ++k) { /* A <- A R (columns p, q) */
                    const __float128 akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - sn * akq;
                    A[k][q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) { /* A <- R^T A (rows p, q) */
                    const __float128 apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - sn * aqk;
                    A[q][k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const __float128 vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - sn * vkq;
                    V[k][q] = sn * vkp + c * vkq;
                }
            }
    }
    /* ascending nodes */
    int idx[ORACLE_RYS_NMAX];
    for (int i = 0; i < n; ++i) idx[i] = i;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (A[idx[j]][idx[j]] < A[idx[i]][idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
    for (int r = 0; r < n; ++r) {
        x[r] = A[idx[r]][idx[r]];
        w[r] = mu[0] * V[0][idx[r]] * V[0][idx[r]];
    }
}

int oracle_rys_roots(int n, double X, double *roots, double *weights) {
    if (n < 1 || n > ORACLE_RYS_NMAX || !(X >= 0)) return 1;
    __float128 x[ORACLE_RYS_NMAX], w[ORACLE_RYS_NMAX];
    rys_rule_q(n, (__float128)X, x, w);
    for (int r = 0; r < n; ++r) {
        roots[r] = (double)(x[r] / (1 - x[r]));
        weights[r] = (double)w[r];
    }
    return 0;
}

/* RecurFactors (src/Rys.jl:141-151), t = u. */
static void rys_recur_factors(double t, double A, double B, double Px, double Qx, double xi, double xk,
                              double *C, double *Cp, double *B0, double *B1, double *B1p) {
    const double ooopt = 1 / (1 + t);
    const double fact = t * ooopt / (A + B);
    *B0 = 0.5 * fact;
    *B1 = 0.5 * ooopt / A + 0.5 * fact;
    *B1p = 0.5 * ooopt / B + 0.5 * fact;
    *C = (Px - xi) * ooopt + (B * (Qx - xi) + A * (Px - xi)) * fact;
    *Cp = (Qx - xk) * ooopt + (B * (Qx - xk) + A * (Px - xk)) * fact;
}

#define RYS_GDIM 13 /* i+j, k+l <= 12 */
/* Recur (src/Rys.jl:92-126): G[a][b] = I(a,0,b,0) for a <= i+j, b <= k+l. */
static void rys_recur(double G[RYS_GDIM][RYS_GDIM], double t, int i, int j, int k, int l, double xi, double xj,
                      double xk, double xl, double ai, double aj, double ak, double al) {
    const int n = i + j, m = k + l;
    const double A = ai + aj, B = ak + al;
    const double Px = (ai * xi + aj * xj) / A, Qx = (ak * xk + al * xl) / B;
    double C, Cp, B0, B1, B1p;
    rys_recur_factors(t, A, B, Px, Qx, xi, xk, &C, &Cp, &B0, &B1, &B1p);
    G[0][0] = M_PI * exp(-ai * aj * (xi - xj) * (xi - xj) / (ai + aj) - ak * al * (xk - xl) * (xk - xl) / (ak + al)) /
              sqrt(A * B);
    if (n > 0) G[1][0] = C * G[0][0];
    if (m > 0) G[0][1] = Cp * G[0][0];
    for (int a = 2; a <= n; ++a) G[a][0] = B1 * (a - 1) * G[a - 2][0] + C * G[a - 1][0];
    for (int b = 2; b <= m; ++b) G[0][b] = B1p * (b - 1) * G[0][b - 2] + Cp * G[0][b - 1];
    if (m == 0 || n == 0) return;
    for (int a = 1; a <= n; ++a) {
        G[a][1] = a * B0 * G[a - 1][0] + Cp * G[a][0];
        for (int b = 2; b <= m; ++b) G[a][b] = B1p * (b - 1) * G[a][b - 2] + a * B0 * G[a - 1][b - 1] + Cp * G[a][b - 1];
    }
}

/* Shift (src/Rys.jl:128-136): I(i,j,k,l) from G = I(i+j,0,k+l,0). */
static double rys_shift(double G[RYS_GDIM][RYS_GDIM], int i, int j, int k, int l, double xij, double xkl) {
    double ijkl = 0;
    for (int m = 0; m <= l; ++m) {
        double ijm0 = 0;
        for (int n = 0; n <= j; ++n) ijm0 += binomial(j, n) * ipow(xij, j - n) * G[n + i][m + k];
        ijkl += binomial(l, m) * ipow(xkl, l - m) * ijm0;
    }
    return ijkl;
}

/* coulomb_rys(a::PGBF, b, c, d) (src/Rys.jl:59-84) without the primitive
 * norms (the caller applies them, as for oracle_coulomb_prim). */
double oracle_coulomb_rys_prim(double aexpn, const double *A, int aI, int aJ, int aK, double bexpn, const double *B,
                               int bI, int bJ, int bK, double cexpn, const double *C, int cI, int cJ, int cK,
                               double dexpn, const double *D, int dI, int dJ, int dK) {
    const int norder = (aI + aJ + aK + bI + bJ + bK + cI + cJ + cK + dI + dJ + dK) / 2 + 1;
    const double Ae = aexpn + bexpn, Be = cexpn + dexpn;
    const double rho = Ae * Be / (Ae + Be);
    double P[3], Q[3], rpq2 = 0;
    for (int x = 0; x < 3; ++x) {
        P[x] = (aexpn * A[x] + bexpn * B[x]) / Ae;
        Q[x] = (cexpn * C[x] + dexpn * D[x]) / Be;
        rpq2 += (P[x] - Q[x]) * (P[x] - Q[x]);
    }
    const double X = rpq2 * rho;
    double roots[ORACLE_RYS_NMAX], weights[ORACLE_RYS_NMAX];
    if (norder > ORACLE_RYS_NMAX || oracle_rys_roots(norder, X, roots, weights)) return NAN;
    double G[RYS_GDIM][RYS_GDIM];
    double ijkl = 0;
    const int Ia[3] = {aI, aJ, aK}, Ib[3] = {bI, bJ, bK}, Ic[3] = {cI, cJ, cK}, Id[3] = {dI, dJ, dK};
    for (int r = 0; r < norder; ++r) {
        double I1[3];
        for (int x = 0; x < 3; ++x) { /* Int1d (:86-90) */
            rys_recur(G, roots[r], Ia[x], Ib[x], Ic[x], Id[x], A[x], B[x], C[x], D[x], aexpn, bexpn, cexpn, dexpn);
            I1[x] = rys_shift(G, Ia[x], Ib[x], Ic[x], Id[x], A[x] - B[x], C[x] - D[x]);
        }
        ijkl += I1[0] * I1[1] * I1[2] * weights[r];
    }
    return 2 * sqrt(rho / M_PI) * ijkl;
}

/* all_twoe_ints_rys (src/Rys.jl:16-25): every unique quartet (iiterator,
 * Utils.jl:17) by coulomb_rys on CGBFs (:27-44), contraction and norms as in
 * oracle_eri_unique_tho. */
int oracle_eri_unique_rys(int nshell, const double *shell_xyz, const int32_t *shell_l,
                          const int32_t *shell_prim_off, const double *prim_exp, const double *prim_coef,
                          const double *prim_norm, int nbf, const double *bf_scale, int order, int nthreads,
                          double *out) {
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
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
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
                                    s += w * oracle_coulomb_rys_prim(
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
            cvrr_shells_d(buf, buf + cvn, &fb, I, J, I, J, 0);
            double *hr = buf + cvn + wn;
            hrr_from_vrr_d(hr, buf, la, lb, la, lb, shell_xyz + 3 * I, shell_xyz + 3 * J, shell_xyz + 3 * I,
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
