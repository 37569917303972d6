/*
 * gen_rys.c -- writes the Rys quadrature tables of the GPU Rys kernels
 * (csrc/gen/rys_table.inc); run by build.py, with --check it reports the
 * fit errors instead.
 *
 * The n-point Rys rule for T >= 0 is the Gauss rule of the weight
 *   w_T(x) = e^{-T x} x^{-1/2} / 2   on x = t^2 in [0, 1],
 * whose moments are the Boys function values F_k(T) = int_0^1 t^{2k} e^{-T t^2} dt:
 *   sum_r w_r x_r^k = F_k(T)   for k = 0 .. 2n-1.
 * (src/Rys.jl:161-1480 evaluates the same rule from fitted polynomials,
 * Root1..Root5, for n <= 5; its roots are u = x/(1-x).)
 *
 * Here the rule is computed in binary128 from the moments: F_k(T) from the
 * all-positive series of F_{2n-1} and downward recursion, recurrence
 * coefficients by Chebyshev's algorithm (ordinary moments; binary128 leaves
 * ~20 digits after the Hankel conditioning of n <= 7), nodes by Sturm-sequence
 * bisection of the Jacobi matrix, weights as Christoffel numbers.
 *
 * Table: for each n = 1..RYS_NMAX and each interval [i H, (i+1) H) below
 * T_max(n), the 2n functions x_1..x_n, w_1..w_n as polynomials of degree
 * RYS_DEG in y = (T - mid) * 2 / H (Chebyshev interpolation, stored in the
 * power basis for Horner evaluation).  Above T_max(n) the kernels use the
 * large-T limit x_r = h_r / T, w_r = W_r / sqrt(T), the Gauss rule of
 * x^{-1/2} e^{-x} / 2 on [0, inf) (moments Gamma(k+1/2)/2), whose error there is
 * below 1e-17 relative (the weight's tail beyond x = 1 is e^{-T}-small).
 */
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef RYS_NMAX
#define RYS_NMAX 7
#endif
#ifndef RYS_DEG
#define RYS_DEG 11
#endif
#ifndef RYS_H
#define RYS_H 1.0
#endif
typedef __float128 q;

/* F_0..F_kmax at T, binary128 */
static void boys_all(int kmax, q T, q *F) {
    const q a = (q)kmax + 0.5Q;
    q top;
    if (T < 1e-30Q) {
        for (int k = 0; k <= kmax; ++k) F[k] = 1 / (q)(2 * k + 1);
        return;
    }
    q term = 1 / a, sum = term;
    for (int i = 1; i < 1000000; ++i) {
        term *= T / (a + i);
        sum += term;
        if (term < sum * 1e-36Q) break;
    }
    top = 0.5Q * expq(-T) * sum;
    F[kmax] = top;
    const q e = expq(-T);
    for (int k = kmax - 1; k >= 0; --k) F[k] = (2 * T * F[k + 1] + e) / (2 * k + 1);
}

/* Gauss rule of n points from 2n ordinary moments (Chebyshev's algorithm). */
static void gauss_from_moments(int n, const q *mu, q *x, q *w) {
    q alpha[RYS_NMAX], beta[RYS_NMAX];
    q sig_prev[2 * RYS_NMAX], sig[2 * RYS_NMAX], sig_new[2 * RYS_NMAX];
    for (int l = 0; l < 2 * n; ++l) {
        sig_prev[l] = 0;
        sig[l] = mu[l];
    }
    alpha[0] = mu[1] / mu[0];
    beta[0] = mu[0];
    for (int k = 1; k < n; ++k) {
        for (int l = k; l < 2 * n - k; ++l)
            sig_new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l];
        alpha[k] = sig_new[k + 1] / sig_new[k] - sig[k] / sig[k - 1];
        beta[k] = sig_new[k] / sig[k - 1];
        for (int l = 0; l < 2 * n; ++l) {
            sig_prev[l] = sig[l];
            sig[l] = sig_new[l];
        }
    }
    /* eigenvalues of the Jacobi matrix (diag alpha, off-diag sqrt beta[1..]):
       the number of eigenvalues below s is the number of negative terms of
       the Sturm sequence d_0 = alpha_0 - s, d_k = alpha_k - s - beta_k / d_{k-1} */
    q lo = 0, hi = 0;
    for (int k = 0; k < n; ++k) {
        q r = (k > 0 ? sqrtq(beta[k]) : 0) + (k + 1 < n ? sqrtq(beta[k + 1]) : 0);
        q a = alpha[k] - r, b = alpha[k] + r;
        if (k == 0 || a < lo) lo = a;
        if (k == 0 || b > hi) hi = b;
    }
    for (int r = 0; r < n; ++r) {
        q a = lo, b = hi;
        for (int it = 0; it < 400; ++it) {
            const q s = 0.5Q * (a + b);
            int cnt = 0;
            q d = 1;
            for (int k = 0; k < n; ++k) {
                d = alpha[k] - s - (k > 0 ? beta[k] / d : 0);
                if (d == 0) d = 1e-60Q;
                if (d < 0) ++cnt;
            }
            if (cnt > r) b = s; else a = s;
            if (b - a <= 1e-33Q * fabsq(b)) break;
        }
        x[r] = 0.5Q * (a + b);
        /* Christoffel number 1 / sum_k p_k(x)^2 (orthonormal p_k) */
        q pm = 0, p = 1 / sqrtq(beta[0]), s2 = p * p;
        for (int k = 0; k + 1 < n; ++k) {
            const q pn = ((x[r] - alpha[k]) * p - (k > 0 ? sqrtq(beta[k]) : 0) * pm) / sqrtq(beta[k + 1]);
            pm = p;
            p = pn;
            s2 += p * p;
        }
        w[r] = 1 / s2;
    }
}

static void rys_exact(int n, q T, q *x, q *w) {
    q F[2 * RYS_NMAX];
    boys_all(2 * n - 1, T, F);
    gauss_from_moments(n, F, x, w);
}

static void rys_limit(int n, q *h, q *W) {
    q mu[2 * RYS_NMAX];
    mu[0] = 0.5Q * sqrtq(M_PIq);
    for (int k = 1; k < 2 * n; ++k) mu[k] = mu[k - 1] * ((q)k - 0.5Q);
    gauss_from_moments(n, mu, h, W);
}

/* smallest multiple of H from which the large-T limit is within tol
   (relative) of the exact rule, checked on a fine grid up to 4x that point */
static double limit_start(int n, double tol) {
    q h[RYS_NMAX], W[RYS_NMAX], x[RYS_NMAX], w[RYS_NMAX];
    rys_limit(n, h, W);
    double start = 0;
    for (double T = 400.0; T >= RYS_H; T -= 0.25) {
        rys_exact(n, T, x, w);
        double e = 0;
        for (int r = 0; r < n; ++r) {
            double ex = (double)fabsq((h[r] / T - x[r]) / x[r]);
            double ew = (double)fabsq((W[r] / sqrtq(T) - w[r]) / w[r]);
            if (ex > e) e = ex;
            if (ew > e) e = ew;
        }
        if (e > tol) {
            start = T;
            break;
        }
    }
    double s = RYS_H;
    while (s <= start) s += RYS_H;
    return s;
}

/* power-basis coefficients (in y on [-1,1]) of the degree-RYS_DEG Chebyshev
   interpolant of the 2n functions on [a, a+H] */
static void fit_interval(int n, double a, q coef[][RYS_DEG + 1]) {
    const int N = RYS_DEG + 1;
    q fv[2 * RYS_NMAX][RYS_DEG + 1], yj[RYS_DEG + 1];
    for (int j = 0; j < N; ++j) {
        yj[j] = cosq(M_PIq * ((q)j + 0.5Q) / N);
        q x[RYS_NMAX], w[RYS_NMAX];
        rys_exact(n, (q)a + 0.5Q * RYS_H * (1 + yj[j]), x, w);
        for (int r = 0; r < n; ++r) {
            fv[r][j] = x[r];
            fv[n + r][j] = w[r];
        }
    }
    /* Chebyshev coefficients, then T_k -> power basis */
    q Tk[RYS_DEG + 1][RYS_DEG + 1];
    memset(Tk, 0, sizeof(Tk));
    Tk[0][0] = 1;
    if (RYS_DEG >= 1) Tk[1][1] = 1;
    for (int k = 2; k <= RYS_DEG; ++k)
        for (int i = 0; i <= k; ++i) Tk[k][i] = (i > 0 ? 2 * Tk[k - 1][i - 1] : 0) - Tk[k - 2][i];
    for (int f = 0; f < 2 * n; ++f) {
        q c[RYS_DEG + 1];
        for (int k = 0; k < N; ++k) {
            q s = 0;
            for (int j = 0; j < N; ++j) s += fv[f][j] * cosq(M_PIq * k * ((q)j + 0.5Q) / N);
            c[k] = s * 2 / N;
        }
        c[0] *= 0.5Q;
        for (int i = 0; i < N; ++i) {
            q s = 0;
            for (int k = i; k < N; ++k) s += c[k] * Tk[k][i];
            coef[f][i] = s;
        }
    }
}

static double horner_d(const double *c, double y) {
    double v = c[RYS_DEG];
    for (int i = RYS_DEG - 1; i >= 0; --i) v = __builtin_fma(v, y, c[i]);
    return v;
}

int main(int argc, char **argv) {
    const char *out = NULL;
    int check = 0;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--out") && i + 1 < argc) out = argv[++i];
        else if (!strcmp(argv[i], "--check")) check = 1;
    }
    double tmax[RYS_NMAX + 1];
    int nint[RYS_NMAX + 1], off[RYS_NMAX + 2];
    off[1] = 0;
    for (int n = 1; n <= RYS_NMAX; ++n) {
        tmax[n] = limit_start(n, 2e-17);
        nint[n] = (int)(tmax[n] / RYS_H + 0.5);
        off[n + 1] = off[n] + nint[n] * 2 * n * (RYS_DEG + 1);
    }
    double *tab = (double *)malloc(sizeof(double) * off[RYS_NMAX + 1]);
    for (int n = 1; n <= RYS_NMAX; ++n) {
        for (int i = 0; i < nint[n]; ++i) {
            q coef[2 * RYS_NMAX][RYS_DEG + 1];
            fit_interval(n, i * RYS_H, coef);
            for (int f = 0; f < 2 * n; ++f)
                for (int k = 0; k <= RYS_DEG; ++k)
                    tab[off[n] + ((size_t)i * 2 * n + f) * (RYS_DEG + 1) + k] = (double)coef[f][k];
        }
    }
    if (check) {
        /* FP64 Horner evaluation against the binary128 rule at points between
           the interpolation nodes, and the large-T limit just above T_max */
        srand(12345);
        for (int n = 1; n <= RYS_NMAX; ++n) {
            double ex = 0, ew = 0;
            for (int s = 0; s < 4000; ++s) {
                const double T = (tmax[n] + 2.0) * (double)rand() / RAND_MAX;
                q x[RYS_NMAX], w[RYS_NMAX];
                rys_exact(n, T, x, w);
                double gx[RYS_NMAX], gw[RYS_NMAX];
                if (T < tmax[n]) {
                    const int i = (int)(T / RYS_H);
                    const double y = (T - (i + 0.5) * RYS_H) * (2.0 / RYS_H);
                    for (int r = 0; r < n; ++r) {
                        gx[r] = horner_d(tab + off[n] + ((size_t)i * 2 * n + r) * (RYS_DEG + 1), y);
                        gw[r] = horner_d(tab + off[n] + ((size_t)i * 2 * n + n + r) * (RYS_DEG + 1), y);
                    }
                } else {
                    q h[RYS_NMAX], W[RYS_NMAX];
                    rys_limit(n, h, W);
                    for (int r = 0; r < n; ++r) {
                        gx[r] = (double)h[r] / T;
                        gw[r] = (double)W[r] / __builtin_sqrt(T);
                    }
                }
                for (int r = 0; r < n; ++r) {
                    double a = (double)fabsq(((q)gx[r] - x[r]) / x[r]);
                    double b = (double)fabsq(((q)gw[r] - w[r]) / w[r]);
                    if (a > ex) ex = a;
                    if (b > ew) ew = b;
                }
            }
            printf("n=%d  T_max=%5.1f intervals=%3d  max rel err: roots %.2e weights %.2e\n", n, tmax[n], nint[n], ex,
                   ew);
        }
        printf("table: %d doubles (%.1f KB)\n", off[RYS_NMAX + 1], off[RYS_NMAX + 1] * 8 / 1024.0);
    }
    if (out) {
        /* included twice by mi_rys.hip (host and device copies): the includer
           defines RYS_QUAL_SMALL / RYS_QUAL_BIG (storage qualifiers) and
           RYS_NAME(x) (name decoration) */
        FILE *f = fopen(out, "w");
        if (!f) return 1;
        fprintf(f, "// generated by gen_rys.c: Rys quadrature tables (do not edit)\n");
        fprintf(f, "#ifndef RYS_NMAX\n#define RYS_NMAX %d\n#define RYS_DEG %d\n#define RYS_H %.17g\n#define RYS_NCOEF %d\n#endif\n",
                RYS_NMAX, RYS_DEG, RYS_H, off[RYS_NMAX + 1]);
        fprintf(f, "// per n: T below which the table applies\n");
        fprintf(f, "RYS_QUAL_SMALL double RYS_NAME(kRysTmax)[RYS_NMAX + 1] = {0");
        for (int n = 1; n <= RYS_NMAX; ++n) fprintf(f, ", %.17g", tmax[n]);
        fprintf(f, "};\n// per n: first coefficient of its intervals\n");
        fprintf(f, "RYS_QUAL_SMALL int RYS_NAME(kRysOff)[RYS_NMAX + 2] = {0");
        for (int n = 1; n <= RYS_NMAX + 1; ++n) fprintf(f, ", %d", off[n]);
        fprintf(f, "};\n// large-T limit: x_r = h_r / T, w_r = W_r / sqrt(T); [n][0][r] = h_r, [n][1][r] = W_r\n");
        fprintf(f, "RYS_QUAL_SMALL double RYS_NAME(kRysLimit)[RYS_NMAX + 1][2][RYS_NMAX] = {{{0}}");
        for (int n = 1; n <= RYS_NMAX; ++n) {
            q h[RYS_NMAX], W[RYS_NMAX];
            rys_limit(n, h, W);
            fprintf(f, ",\n {{");
            for (int r = 0; r < RYS_NMAX; ++r) fprintf(f, "%s%.17g", r ? ", " : "", r < n ? (double)h[r] : 0.0);
            fprintf(f, "}, {");
            for (int r = 0; r < RYS_NMAX; ++r) fprintf(f, "%s%.17g", r ? ", " : "", r < n ? (double)W[r] : 0.0);
            fprintf(f, "}}");
        }
        fprintf(f, "};\n// [n][interval][x_1..x_n, w_1..w_n][power 0..RYS_DEG] from kRysOff[n]\n");
        fprintf(f, "RYS_QUAL_BIG double RYS_NAME(kRysCoef)[RYS_NCOEF] = {\n");
        for (int i = 0; i < off[RYS_NMAX + 1]; ++i) fprintf(f, "%.17g,%s", tab[i], (i % 4 == 3) ? "\n" : " ");
        fprintf(f, "};\n");
        fclose(f);
    }
    free(tab);
    return 0;
}
