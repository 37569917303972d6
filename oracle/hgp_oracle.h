/*
 * hgp_oracle.h -- CPU restatement of MolecularIntegrals.jl's Head-Gordon/Pople
 * two-electron integral path, in plain C (FP64).
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py load it, always as the checker (or as the
 * timed CPU baseline), never as the thing measured on the GPU.  The product
 * library (libmi_hgp.so) never links or calls anything declared here.
 *
 * Every function cites the reference file:line it restates.  "literal" flags
 * reproduce the defects of the reference's all_twoe_ints_chrr bug for bug:
 *   ORACLE_LIT_D1  wrong [00|00] prefactor 2 pi^{5/2} (zeta+eta)^{-3/2}   src/HGP.jl:150-151
 *   ORACLE_LIT_D2  no normalisation (raw Shell.coefs, no CGBF norms)        src/HGP.jl:57-63
 *   ORACLE_LIT_D3  gather at hrrs[mshell...] instead of the AO offsets    src/HGP.jl:29-30
 *   ORACLE_LIT_D5  farray's T >= 30 branch: asymptotic top order and no
 *                  e^{-T} in the recursion (src/Boys.jl:21-26), off by up to
 *                  1e-4 relative for F_12 at T = 30 and by up to 5e-10
 *                  relative on f-shell integrals
 * With flags == 0 the oracle computes the package's tested value definition,
 * coulomb(bfs[i],bfs[j],bfs[k],bfs[l]) (src/ERI.jl:102, whose Fgamma = Fm is
 * the Boys function for every T), by HGP recurrences.
 *
 * Flattened basis format (shared with the product C ABI, see include/mi_hgp.h):
 *   shells in Basis.shells order; per shell xyz (bohr), L, primitive range;
 *   prim_coef = raw Shell.coefs, prim_norm = PGBF norm of the shell's first
 *   Cartesian component; functions in Basis.cgbfs order (shell by shell, each
 *   shell in shell_indices[L] order, src/Basis.jl:127-136), bf_scale[i] =
 *   CGBF norm of function i times (PGBF norm of its component / PGBF norm of
 *   the shell's first component).
 */
#ifndef HGP_ORACLE_H
#define HGP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_LIT_D1 1
#define ORACLE_LIT_D2 2
#define ORACLE_LIT_D3 4
#define ORACLE_LIT_D5 8

#define ORACLE_ORDER_IINDEX 0 /* all_twoe_ints slot = iindex(i,j,k,l)  src/ERI.jl:109 */
#define ORACLE_ORDER_CHRR 1   /* enumeration order of iiterator           src/Basis.jl:203 */

/* AO tables (src/Basis.jl:264-385), 0-based, extended to oracle_lmax(). */
int oracle_lmax(void);
int oracle_nao(int l);                      /* cumulative count; nao(-1)=0 */
void oracle_ao2m(int a, int32_t m[3]);
int oracle_m2ao(int I, int J, int K);
int oracle_shift_index(int a, int i);       /* -1 when absent */
int oracle_shift_direction(int a);
int oracle_shell_number(int a);
int oracle_shift_index_plus(int a, int i);  /* -1 when absent */

/* Boys function (src/Boys.jl). */
double oracle_Fm(int m, double T);                        /* Boys.jl:29-34 */
void oracle_farray(int mmax, double T, double *Fms);      /* Boys.jl:19-27 */

/* Primitive recurrences (src/HGP.jl). vrrs is [mmax][nao(amax)][nao(cmax)]. */
void oracle_vrr(double *vrrs, int amax, int cmax, double aexpn, double bexpn,
                double cexpn, double dexpn, const double *A, const double *B,
                const double *C, const double *D, int literal);
/* hrrs is [nao(la+lb)][nao(lb)][nao(lc+ld)][nao(ld)] (HGP.jl:392-424). */
void oracle_hrr(double *hrrs, int la, int lb, int lc, int ld, double aexpn,
                double bexpn, double cexpn, double dexpn, const double *A,
                const double *B, const double *C, const double *D, int literal);

/* One-electron pieces needed for normalisation (src/OneInts.jl:11-94). */
double oracle_overlap_prim(double aexpn, const double *A, int aI, int aJ, int aK,
                           double bexpn, const double *B, int bI, int bJ, int bK);
double oracle_pgbf_norm(double expn, int I, int J, int K);        /* Basis.jl:42-44 */
double oracle_cgbf_norm(int nprim, const double *expn, const double *coef,
                        int I, int J, int K);                     /* Basis.jl:73-83 */

/* Primitive THO/Huzinaga ERI (src/ERI.jl:15-46), unnormalised. */
double oracle_coulomb_prim(double aexpn, const double *A, int aI, int aJ, int aK,
                           double bexpn, const double *B, int bI, int bJ, int bK,
                           double cexpn, const double *C, int cI, int cJ, int cK,
                           double dexpn, const double *D, int dI, int dJ, int dK);

/* Number of unique ERIs n(n+1)(n^2+n+2)/8 (src/ERI.jl:106). */
int64_t oracle_n_unique(int64_t nbf);
/* 0-based slot of canonical (i>=j, k>=l, tri(i,j)<=tri(k,l)) 1-based functions. */
int64_t oracle_slot(int64_t nbf, int64_t i, int64_t j, int64_t k, int64_t l, int order);

/* all_twoe_ints_chrr (src/HGP.jl:22-34) over a flattened basis.  Returns 0 on
 * success, nonzero on bad input.  nthreads<=0 uses the OpenMP default. */
int oracle_eri_unique_hgp(int nshell, const double *shell_xyz, const int32_t *shell_l,
                          const int32_t *shell_prim_off, const double *prim_exp,
                          const double *prim_coef, const double *prim_norm,
                          int nbf, const double *bf_scale, int order, int literal,
                          int nthreads, double *out);

/* The same algorithm in x87 80-bit arithmetic (inputs and output FP64).  Used
 * only to decide which side lost digits where the FP64 oracle and the GPU
 * disagree (HRR cancellation with tight/diffuse shell pairs). */
int oracle_eri_unique_hgp_ext(int nshell, const double *shell_xyz, const int32_t *shell_l,
                              const int32_t *shell_prim_off, const double *prim_exp,
                              const double *prim_coef, const double *prim_norm, int nbf,
                              const double *bf_scale, int order, int nthreads, double *out);

/* The same algorithm in binary128 (__float128; inputs and output FP64): the
 * recurrences' rounding is ~1e-34 relative, far below every parity bar, so
 * this is the exact value of the reference algorithm.  accurate = 0: Boys by
 * the reference's farray branches (T >= 30 asymptotic, e^-T dropped,
 * src/Boys.jl:19-27); accurate = 1: the Boys function itself for every T, as
 * coulomb()'s Fgamma (src/ERI.jl:41, src/Boys.jl:29-34). */
int oracle_eri_unique_hgp_quad(int nshell, const double *shell_xyz, const int32_t *shell_l,
                               const int32_t *shell_prim_off, const double *prim_exp,
                               const double *prim_coef, const double *prim_norm, int nbf,
                               const double *bf_scale, int order, int accurate, int nthreads,
                               double *out);
/* farray with the Boys function exact for every T (no asymptotic branch). */
void oracle_farray_accurate(int mmax, double T, double *Fms);

/* The binary128 values of nslot selected 0-based slots (in `order`) of the
 * unique array, without computing the rest. */
int oracle_eri_slots_quad(int nshell, const double *shell_xyz, const int32_t *shell_l,
                          const int32_t *shell_prim_off, const double *prim_exp,
                          const double *prim_coef, const double *prim_norm, int nbf,
                          const double *bf_scale, int order, int accurate, int64_t nslot,
                          const int64_t *slots, int nthreads, double *out);
/* F_m(T) through the binary128 farray (accurate as above). */
double oracle_Fm_quad(int m, double T, int accurate);

/* Same contract as above but only for the listed shell quartets (sampled CPU
 * baseline); returns number of ERIs computed (written nowhere). */
int64_t oracle_eri_hgp_quartet_sample(int nshell, const double *shell_xyz,
                                      const int32_t *shell_l, const int32_t *shell_prim_off,
                                      const double *prim_exp, const double *prim_coef,
                                      const double *prim_norm, int nbf, const double *bf_scale,
                                      int64_t nq, const int32_t *quartets, int nthreads,
                                      double *checksum);

/* all_twoe_ints(bfs, coulomb) (src/ERI.jl:104-112) via THO -- independent
 * cross-check of the HGP recurrences; O(n^4 K^4), tiny inputs only. */
int oracle_eri_unique_tho(int nshell, const double *shell_xyz, const int32_t *shell_l,
                          const int32_t *shell_prim_off, const double *prim_exp,
                          const double *prim_coef, const double *prim_norm,
                          int nbf, const double *bf_scale, int order, double *out);

/* Rys quadrature (src/Rys.jl).  oracle_rys_roots: the n-point Rys rule at X
 * (n <= 10) from its definition (binary128: moments F_k(X), Chebyshev's
 * algorithm, Golub-Welsch by Jacobi rotations), roots as the reference's
 * u = t^2/(1-t^2) (Roots, :161-174; the reference fits them for n <= 5).
 * oracle_coulomb_rys_prim: coulomb_rys on primitives without norms
 * (:59-84, Recur :92-126, Shift :128-136, RecurFactors :141-151), FP64.
 * oracle_eri_unique_rys: all_twoe_ints_rys (:16-25) over a flattened basis. */
int oracle_rys_roots(int n, double X, double *roots, double *weights);
double oracle_coulomb_rys_prim(double aexpn, const double *A, int aI, int aJ, int aK,
                               double bexpn, const double *B, int bI, int bJ, int bK,
                               double cexpn, const double *C, int cI, int cJ, int cK,
                               double dexpn, const double *D, int dI, int dJ, int dK);
int oracle_eri_unique_rys(int nshell, const double *shell_xyz, const int32_t *shell_l,
                          const int32_t *shell_prim_off, const double *prim_exp,
                          const double *prim_coef, const double *prim_norm,
                          int nbf, const double *bf_scale, int order, int nthreads, double *out);

/* Coulomb and exchange from an iindex-ordered unique array, following the
 * index algebra of make2JmK (src/ERI.jl:114-128): J_ij = sum_kl D_kl (ij|kl),
 * K_ij = sum_kl D_kl (ik|jl).  D, J, K are n x n row-major. */
void oracle_jk_from_unique(int nbf, const double *eri_iindex, const double *D,
                           double *J, double *K);

/* Schwarz diagonal max_{ab in pair} |(ab|ab)| per shell pair I>=J, tri order
 * of shell indices; computed with the corrected HGP path. */
int oracle_schwarz_pairs(int nshell, const double *shell_xyz, const int32_t *shell_l,
                         const int32_t *shell_prim_off, const double *prim_exp,
                         const double *prim_coef, const double *prim_norm,
                         int nbf, const double *bf_scale, double *q_pair);

#ifdef __cplusplus
}
#endif
#endif
