// mi_rys.hip -- Rys quadrature kernel (tuning variant 5): the two-electron
// integrals of MolecularIntegrals.jl's Rys path (coulomb_rys,
// src/Rys.jl:27-84; all_twoe_ints_rys, :16-25) for every quartet class.
//
// For a primitive quartet the integral over Cartesian Gaussians is
//   [e|f] = pref * sum_r w_r Gx_r[ex][fx] Gy_r[ey][fy] Gz_r[ez][fz]
// (ABD eq. 5 and 9; pref = K_ab K_cd / sqrt(zeta + eta), the prefactor the
// HGP kernels give [0|0]^(0) / F_0), where (x_r, w_r) is the n-point Rys rule
// at T = zeta eta |PQ|^2 / (zeta + eta), n = (L_a+L_b+L_c+L_d)/2 + 1, and the
// 2D integrals G_r[a][c] follow the recurrence of Recur (src/Rys.jl:92-126,
// factors RecurFactors :141-151 with the root as x = t^2 rather than
// u = t^2 / (1 - t^2)):
//   G[0][0] = 1, G[a+1][0] = C G[a][0] + a B10 G[a-1][0],
//   G[a][c+1] = C' G[a][c] + c B01 G[a][c-1] + a B00 G[a-1][c],
//   C = PA + x WP, C' = QC + x WQ, B00 = x / (2(zeta+eta)),
//   B10 = (1 - x eta/(zeta+eta)) / (2 zeta), B01 = (1 - x zeta/(zeta+eta)) / (2 eta).
// The kernel forms [e|f] for the e (L_a <= |e| <= L_a+L_b) and f
// (L_c <= |f| <= L_c+L_d) of the class -- the contracted (e0|f0) block the
// HGP kernels build by VRR -- sums it over the primitive quartets and
// transfers it to (ab|cd) with the HGP horizontal recurrence (the same op
// tables).  The reference transfers each primitive's 1D integrals instead
// (Shift, src/Rys.jl:128-136, I(i,j,k,l) from I(i+j,0,k+l,0)); the transfer
// is linear, so after the contraction it gives the same integrals with the
// transfer done once per shell quartet.
//
// Rys rule: Gauss rule of e^{-T x} x^{-1/2}/2 on x in [0, 1] (moments F_k(T)),
// from the tables of gen_rys.c: degree-11 polynomials on unit intervals of T
// below T_max(n), the large-T limit above (rule error ~2e-16 relative).
//
// Shape: persistent blocks of nw wavefronts, one shell quartet at a time; the
// primitive quartets go in batches of nbat: each (primitive, root, direction)
// item builds its 2D table in LDS, then each thread forms the products of the
// (e0|f0) elements it owns (ACC per thread, kept in registers across batches).
#include <cmath>
#include <vector>

#include "mi_device.h"

__device__ __forceinline__ int ry_ncart(int l) { return (l + 1) * (l + 2) / 2; }

#define RYS_QUAL_SMALL static __constant__ const
#define RYS_QUAL_BIG static __device__ const
#define RYS_NAME(x) x##Dev
#include "gen/rys_table.inc"
#undef RYS_QUAL_SMALL
#undef RYS_QUAL_BIG
#undef RYS_NAME
#define RYS_QUAL_SMALL static const
#define RYS_QUAL_BIG static const
#define RYS_NAME(x) x##Host
#include "gen/rys_table.inc"
#undef RYS_QUAL_SMALL
#undef RYS_QUAL_BIG
#undef RYS_NAME

#define RY_SCL 16  // LDS stride of the per-shell scale rows
static size_t g_ry_lds_limit = 65536;  // dynamic LDS a block may use (set_rys_lds_limit)

namespace mi {

// Root r (x = t^2) and weight of the n-point rule at T.
__device__ __forceinline__ void rys_rule_dev(int n, int r, double T, double& x, double& w) {
  if (T < kRysTmaxDev[n]) {
    const int i = (int)T;  // interval [i, i+1), RYS_H = 1
    const double y = 2.0 * (T - (double)i) - 1.0;
    const double* c = kRysCoefDev + kRysOffDev[n] + (size_t)(i * 2 * n + r) * (RYS_DEG + 1);
    const double* d = c + (size_t)n * (RYS_DEG + 1);
    double px = c[RYS_DEG], pw = d[RYS_DEG];
#pragma unroll
    for (int k = RYS_DEG - 1; k >= 0; --k) {
      px = fma(px, y, c[k]);
      pw = fma(pw, y, d[k]);
    }
    x = px;
    w = pw;
  } else {
    const double rs = rsqrt_nr(T);
    x = kRysLimitDev[n][0][r] * (rs * rs);
    w = kRysLimitDev[n][1][r] * rs;
  }
}

// One polynomial of the rule at T: root r (which = 0) or weight r (which = 1).
__device__ __forceinline__ double rys_poly_dev(int n, int r, int which, double T) {
  if (T < kRysTmaxDev[n]) {
    const int i = (int)T;
    const double y = 2.0 * (T - (double)i) - 1.0;
    const double* c = kRysCoefDev + kRysOffDev[n] + (size_t)(i * 2 * n + which * n + r) * (RYS_DEG + 1);
    double p = c[RYS_DEG];
#pragma unroll
    for (int k = RYS_DEG - 1; k >= 0; --k) p = fma(p, y, c[k]);
    return p;
  }
  const double rs = rsqrt_nr(T);
  return which ? kRysLimitDev[n][1][r] * rs : kRysLimitDev[n][0][r] * (rs * rs);
}

// Host evaluation of the same tables (C ABI mi_rys_rule, CPU tests).
void rys_rule_host(int n, double T, double* x, double* w) {
  for (int r = 0; r < n; ++r) {
    if (T < kRysTmaxHost[n]) {
      const int i = (int)T;
      const double y = 2.0 * (T - (double)i) - 1.0;
      const double* c = kRysCoefHost + kRysOffHost[n] + (size_t)(i * 2 * n + r) * (RYS_DEG + 1);
      const double* d = c + (size_t)n * (RYS_DEG + 1);
      double px = c[RYS_DEG], pw = d[RYS_DEG];
      for (int k = RYS_DEG - 1; k >= 0; --k) {
        px = std::fma(px, y, c[k]);
        pw = std::fma(pw, y, d[k]);
      }
      x[r] = px;
      w[r] = pw;
    } else {
      x[r] = kRysLimitHost[n][0][r] / T;
      w[r] = kRysLimitHost[n][1][r] / std::sqrt(T);
    }
  }
}
int rys_nmax() { return RYS_NMAX; }

// LDS: [scales: 4*RY_SCL][work][hrr: nhrr if HL]
// S = (amax+1)(cmax+1) entries per 2D table; work (>= nbat n 3 S doubles)
// holds the batch's 2D tables, then (J/K build, work >= nout) the scaled block.
template <int ACC, bool HL>
__global__ void __launch_bounds__(256) hgp_ry_kernel(const EriArgs args, const WClassDesc D,
                                                     const uint32_t* __restrict__ gblob, double* __restrict__ scratch,
                                                     int64_t nq_launch, int nbat, int work) {
  extern __shared__ double lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int n = (D.mmax + 1) / 2;
  const int amax = D.la + D.lb, cmax = D.lc + D.ld;
  const int c1 = cmax + 1, S = (amax + 1) * c1;
  double* scl = lds;
  double* G = lds + 4 * RY_SCL;
  double* hrr = HL ? G + work : scratch + (size_t)blockIdx.x * D.nhrr;
  const uint4* gH = reinterpret_cast<const uint4*>(gblob + D.gh);
  const uint2_t* gO = reinterpret_cast<const uint2_t*>(gblob + D.go);
  const uint2_t* gA = reinterpret_cast<const uint2_t*>(gblob + D.ga);
  const uint32_t* gR = gblob + D.gr;
  const int na = ry_ncart(D.la), nb = ry_ncart(D.lb), nc = ry_ncart(D.lc), nd = ry_ncart(D.ld);
  // offsets of the owned (e0|f0) elements in a (root) group of three 2D tables
  int ox[ACC], oy[ACC], oz[ACC];
#pragma unroll
  for (int k = 0; k < ACC; ++k) {
    const int x = tid + k * nthr;
    const uint32_t e = x < D.nacc ? gR[x] : 0u;
    ox[k] = (int)(e & 7) * c1 + (int)((e >> 9) & 7);
    oy[k] = S + (int)((e >> 3) & 7) * c1 + (int)((e >> 12) & 7);
    oz[k] = 2 * S + (int)((e >> 6) & 7) * c1 + (int)((e >> 15) & 7);
  }
  JkRun run{kNoRun, kNoRun, kNoRun, 0, 0, 0};
  int ti = 0;
  for (int64_t q = blockIdx.x; q < nq_launch; q += gridDim.x) {
    ti = advance_task(args, ti, q);
    const TaskDesc T = args.tasks[ti];
    QuartetCtx qc;
    decode_index(T, q - T.blockStart, qc);
    qc.G = 1;
    qc.g = 0;
    if (quartet_mode(args, qc) == 2) continue;  // screened (uniform per block): output pre-zeroed
    const PairTopo P = args.topo[qc.bra];
    const PairTopo Q = args.topo[qc.ket];
    if (tid < na) scl[tid] = args.bf_scale[P.bfA + tid];
    if (tid < nb) scl[RY_SCL + tid] = args.bf_scale[P.bfB + tid];
    if (tid < nc) scl[2 * RY_SCL + tid] = args.bf_scale[Q.bfA + tid];
    if (tid < nd) scl[3 * RY_SCL + tid] = args.bf_scale[Q.bfB + tid];
    double acc[ACC];
#pragma unroll
    for (int k = 0; k < ACC; ++k) acc[k] = 0.0;
    const int K = P.npp * Q.npp;
    const double rnv = rcp_int(Q.npp);
    for (int t0 = 0; t0 < K; t0 += nbat) {
      const int nbc = K - t0 < nbat ? K - t0 : nbat;
      // 2D tables of item (primitive p, root r, direction d)
      for (int it = tid; it < nbc * n * 3; it += nthr) {
        const int p = it / (3 * n), rd = it - p * 3 * n, r = rd / 3, d = rd - r * 3;
        int v;
        const int u = udiv_rcp(t0 + p, Q.npp, rnv, v);
        const PrimPair bp = args.pp[P.ppoff + u];
        const PrimPair kp = args.pp[Q.ppoff + v];
        const double zeta = bp.zeta, eta = kp.zeta;
        const double rs = rsqrt_nr(zeta + eta), inv_s = rs * rs;
        const double* A = args.shell_xyz + 3 * P.shA;
        const double* C = args.shell_xyz + 3 * Q.shA;
        const double PQx = (A[0] + bp.PAx) - (C[0] + kp.PAx), PQy = (A[1] + bp.PAy) - (C[1] + kp.PAy),
                     PQz = (A[2] + bp.PAz) - (C[2] + kp.PAz);
        const double Targ = zeta * eta * inv_s * (PQx * PQx + PQy * PQy + PQz * PQz);
        double x, w;
        rys_rule_dev(n, r, Targ, x, w);
        const double PQd = d == 0 ? PQx : (d == 1 ? PQy : PQz);
        const double PAd = d == 0 ? bp.PAx : (d == 1 ? bp.PAy : bp.PAz);
        const double QCd = d == 0 ? kp.PAx : (d == 1 ? kp.PAy : kp.PAz);
        const double eta_s = eta * inv_s, zeta_s = zeta * inv_s;
        const double Cf = PAd - x * eta_s * PQd;   // PA + x WP
        const double Cp = QCd + x * zeta_s * PQd;  // QC + x WQ
        const double B00 = 0.5 * x * inv_s;
        const double B10 = bp.oo2z * (1.0 - x * eta_s);
        const double B01 = kp.oo2z * (1.0 - x * zeta_s);
        double* g = G + (size_t)it * S;
        g[0] = d == 0 ? bp.K * kp.K * rs * w : 1.0;
        if (amax > 0) g[c1] = Cf * g[0];
        for (int a = 1; a < amax; ++a) g[(a + 1) * c1] = Cf * g[a * c1] + (double)a * B10 * g[(a - 1) * c1];
        for (int a = 0; a <= amax; ++a) {
          double* ga = g + a * c1;
          const double* gm = g + (a - 1) * c1;
          const double aB = (double)a * B00;
          if (cmax > 0) ga[1] = Cp * ga[0] + (a > 0 ? aB * gm[0] : 0.0);
          for (int c = 1; c < cmax; ++c)
            ga[c + 1] = Cp * ga[c] + (double)c * B01 * ga[c - 1] + (a > 0 ? aB * gm[c] : 0.0);
        }
      }
      __syncthreads();
      const int ng = nbc * n;
#pragma unroll
      for (int k = 0; k < ACC; ++k) {
        if (tid + k * nthr < D.nacc) {
          double s = 0.0;
          for (int j = 0; j < ng; ++j) {
            const double* g = G + (size_t)j * 3 * S;
            s = fma(g[ox[k]] * g[oy[k]], g[oz[k]], s);
          }
          acc[k] += s;
        }
      }
      __syncthreads();  // the next batch overwrites the tables
    }
#pragma unroll
    for (int k = 0; k < ACC; ++k) {
      const int x = tid + k * nthr;
      if (x < D.nacc) {
        hrr[x] = acc[k];
        if (args.mode == MI_MODE_SCHWARZ && gA[x].y) record_vmax(&args.geom[qc.bra], acc[k]);
      }
    }
    __syncthreads();
    // HRR levels: item k writes slot nacc + k (src/HGP.jl:83-116)
    const PairGeom PG = args.geom[qc.bra];
    const PairGeom QG = args.geom[qc.ket];
    for (int lv = 0; lv < D.nhlev; ++lv) {
      const int i1 = D.hlev[lv + 1];
      for (int it = D.hlev[lv] + tid; it < i1; it += nthr) {
        const uint4 o = gH[it];
        const int j = o.z & 3;
        const bool ket = (o.z >> 2) & 1;
        const double X = ket ? (j == 0 ? QG.ABx : (j == 1 ? QG.ABy : QG.ABz))
                             : (j == 0 ? PG.ABx : (j == 1 ? PG.ABy : PG.ABz));
        hrr[D.nacc + it] = hrr[o.x] + X * hrr[o.y];
      }
      __syncthreads();
    }
    const ScatterCtx sc = make_scatter(args, qc, P, Q);
    if (args.mode == MI_MODE_JK) {  // uniform per block
      for (int x = tid; x < D.nout; x += nthr) {
        const uint2_t o = gO[x];
        const int ma = o.y & 15, mb = (o.y >> 4) & 15, mc = (o.y >> 8) & 15, md = (o.y >> 12) & 15;
        G[x] = hrr[o.x] * (scl[ma] * scl[RY_SCL + mb]) * (scl[2 * RY_SCL + mc] * scl[3 * RY_SCL + md]);
      }
      __syncthreads();
      jk_block_lds(args, sc, G, na, nb, nc, nd, tid, nthr, nullptr, run);
      __syncthreads();
      continue;
    }
    for (int x = tid; x < D.nout; x += nthr) {
      const uint2_t o = gO[x];
      const int ma = o.y & 15, mb = (o.y >> 4) & 15, mc = (o.y >> 8) & 15, md = (o.y >> 12) & 15;
      scatter_elem(args, sc, qc.bra, sc.fa0 + ma, sc.fb0 + mb, sc.fc0 + mc, sc.fd0 + md,
                   hrr[o.x] * (scl[ma] * scl[RY_SCL + mb]) * (scl[2 * RY_SCL + mc] * scl[3 * RY_SCL + md]));
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Rys lane-group kernel (tuning variant 6): one-wavefront blocks whose lanes
// form 64/LG groups of LG lanes; each group takes its own shell quartets
// (persistent, strided over the launch), so a wavefront carries several
// quartets and no block-wide barrier waits on another wavefront.  Per
// primitive quartet the group's lanes evaluate its rule (one polynomial per
// lane), build its 3n 2D tables (one table per lane) in the group's LDS
// area, and each lane accumulates
// the (e0|f0) elements it owns (ACC per lane, in registers); after the last
// primitive quartet the group runs the HRR levels in the same area (the
// tables are dead then) and scatters.  All groups of a wavefront run the
// same number of quartet slots and, per slot, of primitive iterations (the
// most any of its groups needs), so the barriers stay in uniform control flow.
// LDS per group: 4*RY_SCL scales, then max(3n S + 2n, nhrr [+ nout, J/K]).
// AM, CM, NR >= 0: instantiation for classes with L_a+L_b = AM, L_c+L_d = CM
// and NR roots (recurrence loops and root sums unrolled at compile time);
// -1: read from the class descriptor.
// ---------------------------------------------------------------------------
template <int LG, int ACC, int AM = -1, int CM = -1, int NR = -1>
__global__ void __launch_bounds__(64) hgp_ryg_kernel(const EriArgs args, const WClassDesc D,
                                                     const uint32_t* __restrict__ gblob, int64_t nq_launch,
                                                     int area) {
  extern __shared__ double lds[];
  constexpr int NG = 64 / LG;
  const int lane = threadIdx.x, grp = lane / LG, lg = lane - grp * LG;
  double* scl = lds + (size_t)grp * area;
  double* W = scl + 4 * RY_SCL;
  const int n = NR > 0 ? NR : (D.mmax + 1) / 2;
  const int amax = AM >= 0 ? AM : D.la + D.lb, cmax = CM >= 0 ? CM : D.lc + D.ld;
  const int c1 = cmax + 1, S = (amax + 1) * c1, tab = 3 * n * S;
  const uint4* gH = reinterpret_cast<const uint4*>(gblob + D.gh);
  const uint2_t* gO = reinterpret_cast<const uint2_t*>(gblob + D.go);
  const uint2_t* gA = reinterpret_cast<const uint2_t*>(gblob + D.ga);
  const uint32_t* gR = gblob + D.gr;
  const int na = ry_ncart(D.la), nb = ry_ncart(D.lb), nc = ry_ncart(D.lc), nd = ry_ncart(D.ld);
  int ox[ACC], oy[ACC], oz[ACC];
#pragma unroll
  for (int k = 0; k < ACC; ++k) {
    const int x = lg + k * LG;
    const uint32_t e = x < D.nacc ? gR[x] : 0u;
    ox[k] = (int)(e & 7) * c1 + (int)((e >> 9) & 7);
    oy[k] = S + (int)((e >> 3) & 7) * c1 + (int)((e >> 12) & 7);
    oz[k] = 2 * S + (int)((e >> 6) & 7) * c1 + (int)((e >> 15) & 7);
  }
  JkRun run{kNoRun, kNoRun, kNoRun, 0, 0, 0};
  const int64_t ngroups = (int64_t)gridDim.x * NG;
  const int64_t gid = (int64_t)blockIdx.x * NG + grp;
  const int64_t slots = (nq_launch + ngroups - 1) / ngroups;  // uniform over the wavefront
  int ti = 0;
  for (int64_t it = 0; it < slots; ++it) {
    const int64_t q = gid + it * ngroups;
    bool valid = q < nq_launch;
    QuartetCtx qc{};
    PairTopo P{}, Q{};
    if (valid) {
      ti = advance_task(args, ti, q);
      const TaskDesc T = args.tasks[ti];
      decode_index(T, q - T.blockStart, qc);
      qc.G = 1;
      qc.g = 0;
      valid = quartet_mode(args, qc) != 2;  // screened: output pre-zeroed
      P = args.topo[qc.bra];
      Q = args.topo[qc.ket];
    }
    const int K = valid ? P.npp * Q.npp : 0;
    int Kmax = K;  // most primitive quartets of any group of the wavefront
#pragma unroll
    for (int sft = LG; sft < 64; sft <<= 1) Kmax = max(Kmax, __shfl_xor(Kmax, sft, 64));
    if (valid) {
      if (lg < na) scl[lg] = args.bf_scale[P.bfA + lg];
      if (lg < nb) scl[RY_SCL + lg] = args.bf_scale[P.bfB + lg];
      if (lg < nc) scl[2 * RY_SCL + lg] = args.bf_scale[Q.bfA + lg];
      if (lg < nd) scl[3 * RY_SCL + lg] = args.bf_scale[Q.bfB + lg];
    }
    const double* A = args.shell_xyz + 3 * (valid ? P.shA : 0);
    const double* C = args.shell_xyz + 3 * (valid ? Q.shA : 0);
    const double rnv = rcp_int(valid ? Q.npp : 1);
    double acc[ACC];
#pragma unroll
    for (int k = 0; k < ACC; ++k) acc[k] = 0.0;
    double* G = W;             // the 3n 2D tables of the current primitive quartet
    double* xw = W + tab;      // its roots x_r (first n) and weights w_r (next n)
    for (int t = 0; t < Kmax; ++t) {
      const bool act = t < K;
      double zeta = 0.0, eta = 0.0, rs = 0.0, PQx = 0.0, PQy = 0.0, PQz = 0.0;
      PrimPair bp{}, kp{};
      if (act) {
        int v;
        const int u = udiv_rcp(t, Q.npp, rnv, v);
        bp = args.pp[P.ppoff + u];
        kp = args.pp[Q.ppoff + v];
        zeta = bp.zeta;
        eta = kp.zeta;
        rs = rsqrt_nr(zeta + eta);
        PQx = (A[0] + bp.PAx) - (C[0] + kp.PAx);
        PQy = (A[1] + bp.PAy) - (C[1] + kp.PAy);
        PQz = (A[2] + bp.PAz) - (C[2] + kp.PAz);
        // the rule: one polynomial per lane (root or weight r of the n-point rule)
        const double inv_s = rs * rs;
        const double Targ = zeta * eta * inv_s * (PQx * PQx + PQy * PQy + PQz * PQz);
        for (int item = lg; item < 2 * n; item += LG)
          xw[item] = rys_poly_dev(n, item < n ? item : item - n, item < n ? 0 : 1, Targ);
      }
      __syncthreads();  // the group's rule is in LDS
      if (act) {
        const double inv_s = rs * rs, eta_s = eta * inv_s, zeta_s = zeta * inv_s;
        for (int item = lg; item < 3 * n; item += LG) {
          const int r = item / 3, d = item - r * 3;
          const double x = xw[r];
          const double PQd = d == 0 ? PQx : (d == 1 ? PQy : PQz);
          const double PAd = d == 0 ? bp.PAx : (d == 1 ? bp.PAy : bp.PAz);
          const double QCd = d == 0 ? kp.PAx : (d == 1 ? kp.PAy : kp.PAz);
          const double Cf = PAd - x * eta_s * PQd;
          const double Cp = QCd + x * zeta_s * PQd;
          const double B00 = 0.5 * x * inv_s;
          const double B10 = bp.oo2z * (1.0 - x * eta_s);
          const double B01 = kp.oo2z * (1.0 - x * zeta_s);
          double* g = G + item * S;
          g[0] = d == 0 ? bp.K * kp.K * rs * xw[n + r] : 1.0;
          if (amax > 0) g[c1] = Cf * g[0];
          for (int a = 1; a < amax; ++a) g[(a + 1) * c1] = Cf * g[a * c1] + (double)a * B10 * g[(a - 1) * c1];
          for (int a = 0; a <= amax; ++a) {
            double* ga = g + a * c1;
            const double* gm = g + (a - 1) * c1;
            const double aB = (double)a * B00;
            if (cmax > 0) ga[1] = Cp * ga[0] + (a > 0 ? aB * gm[0] : 0.0);
            for (int c = 1; c < cmax; ++c)
              ga[c + 1] = Cp * ga[c] + (double)c * B01 * ga[c - 1] + (a > 0 ? aB * gm[c] : 0.0);
          }
        }
      }
      __syncthreads();  // the group's tables are in LDS
      if (act) {
#pragma unroll
        for (int k = 0; k < ACC; ++k) {
          if (lg + k * LG < D.nacc) {
            double sum = 0.0;
            for (int r = 0; r < n; ++r) {
              const double* g = G + r * 3 * S;
              sum = fma(g[ox[k]] * g[oy[k]], g[oz[k]], sum);
            }
            acc[k] += sum;
          }
        }
      }
      // the next iteration overwrites the rule and the tables only after its
      // first barrier, when every lane is past these products
    }
    __syncthreads();  // tables dead: the area holds the HRR levels now
    double* hrr = W;
    if (valid) {
#pragma unroll
      for (int k = 0; k < ACC; ++k) {
        const int x = lg + k * LG;
        if (x < D.nacc) {
          hrr[x] = acc[k];
          if (args.mode == MI_MODE_SCHWARZ && gA[x].y) record_vmax(&args.geom[qc.bra], acc[k]);
        }
      }
    }
    __syncthreads();
    const PairGeom PG = args.geom[valid ? qc.bra : 0];
    const PairGeom QG = args.geom[valid ? qc.ket : 0];
    for (int lv = 0; lv < D.nhlev; ++lv) {
      if (valid) {
        const int i1 = D.hlev[lv + 1];
        for (int itm = D.hlev[lv] + lg; itm < i1; itm += LG) {
          const uint4 o = gH[itm];
          const int j = o.z & 3;
          const bool ket = (o.z >> 2) & 1;
          const double X = ket ? (j == 0 ? QG.ABx : (j == 1 ? QG.ABy : QG.ABz))
                               : (j == 0 ? PG.ABx : (j == 1 ? PG.ABy : PG.ABz));
          hrr[D.nacc + itm] = hrr[o.x] + X * hrr[o.y];
        }
      }
      __syncthreads();
    }
    if (args.mode == MI_MODE_JK) {  // uniform per launch
      double* V = hrr + D.nhrr;
      ScatterCtx sc{};
      if (valid) {
        sc = make_scatter(args, qc, P, Q);
        for (int x = lg; x < D.nout; x += LG) {
          const uint2_t o = gO[x];
          const int ma = o.y & 15, mb = (o.y >> 4) & 15, mc = (o.y >> 8) & 15, md = (o.y >> 12) & 15;
          V[x] = hrr[o.x] * (scl[ma] * scl[RY_SCL + mb]) * (scl[2 * RY_SCL + mc] * scl[3 * RY_SCL + md]);
        }
      }
      __syncthreads();
      if (valid) jk_block_lds(args, sc, V, na, nb, nc, nd, lg, LG, nullptr, run);
      __syncthreads();
      continue;
    }
    if (valid) {
      const ScatterCtx sc = make_scatter(args, qc, P, Q);
      for (int x = lg; x < D.nout; x += LG) {
        const uint2_t o = gO[x];
        const int ma = o.y & 15, mb = (o.y >> 4) & 15, mc = (o.y >> 8) & 15, md = (o.y >> 12) & 15;
        scatter_elem(args, sc, qc.bra, sc.fa0 + ma, sc.fb0 + mb, sc.fc0 + mc, sc.fd0 + md,
                     hrr[o.x] * (scl[ma] * scl[RY_SCL + mb]) * (scl[2 * RY_SCL + mc] * scl[3 * RY_SCL + md]));
      }
    }
    __syncthreads();  // the next slot overwrites the area
  }
}

template <int LG>
static const void* ryg_fn_lg(int acc) {
  switch (acc) {
    case 2: return (const void*)hgp_ryg_kernel<LG, 2>;
    case 4: return (const void*)hgp_ryg_kernel<LG, 4>;
    case 8: return (const void*)hgp_ryg_kernel<LG, 8>;
    default: return (const void*)hgp_ryg_kernel<LG, 16>;
  }
}
// Compile-time instantiations (L_a+L_b, L_c+L_d, roots, lanes per group,
// accumulators per lane = ceil(nacc / lanes)): every class up to d with 3 or
// more units of angular momentum at each lane count its accumulators allow,
// and the f-class shapes of the tuning table.  Exact accumulator counts keep
// the registers of unused accumulators and their table indices free.
#define RYG_SPECS(X) \
  X(2, 1, 2, 16, 2) X(2, 1, 2, 32, 1) X(2, 1, 2, 64, 1) X(2, 2, 3, 16, 3) X(2, 2, 3, 16, 4)    \
  X(2, 2, 3, 16, 6) X(2, 2, 3, 32, 2) X(2, 2, 3, 32, 3) X(2, 2, 3, 64, 1) X(2, 2, 3, 64, 2)    \
  X(3, 0, 2, 16, 1) X(3, 0, 2, 32, 1) X(3, 0, 2, 64, 1) X(3, 1, 3, 16, 3) X(3, 1, 3, 32, 2)    \
  X(3, 1, 3, 64, 1) X(3, 2, 3, 16, 6) X(3, 2, 3, 16, 9) X(3, 2, 3, 32, 3) X(3, 2, 3, 32, 5)    \
  X(3, 2, 3, 64, 2) X(3, 2, 3, 64, 3) X(3, 3, 4, 16, 7) X(3, 3, 4, 16, 10) X(3, 3, 4, 16, 16)  \
  X(3, 3, 4, 32, 5) X(3, 3, 4, 32, 8) X(3, 3, 4, 64, 4) X(3, 4, 4, 64, 5) X(4, 0, 3, 16, 2)    \
  X(4, 0, 3, 32, 1) X(4, 0, 3, 64, 1) X(4, 1, 3, 16, 6) X(4, 1, 3, 32, 3) X(4, 1, 3, 64, 2)    \
  X(4, 2, 4, 16, 10) X(4, 2, 4, 16, 12) X(4, 2, 4, 32, 5) X(4, 2, 4, 32, 6) X(4, 2, 4, 32, 9)  \
  X(4, 2, 4, 64, 3) X(4, 2, 4, 64, 4) X(4, 2, 4, 64, 5) X(4, 3, 4, 32, 16) X(4, 3, 4, 64, 4)   \
  X(4, 3, 4, 64, 7) X(4, 3, 4, 64, 8) X(4, 4, 5, 64, 16) X(5, 0, 3, 16, 4) X(5, 1, 4, 32, 5)   \
  X(5, 1, 4, 64, 3) X(5, 2, 4, 64, 5) X(5, 2, 4, 64, 8) X(5, 3, 5, 64, 8) X(5, 3, 5, 64, 9)    \
  X(5, 3, 5, 64, 13)
static const void* ryg_kernel_fn(const WClassDesc& D, int lg, int acc) {
  const int am = D.la + D.lb, cm = D.lc + D.ld, n = (D.mmax + 1) / 2, need = (D.nacc + lg - 1) / lg;
#define RYG_PICK(A, C, N, L, K) \
  if (am == A && cm == C && n == N && lg == L && need == K) return (const void*)hgp_ryg_kernel<L, K, A, C, N>;
  RYG_SPECS(RYG_PICK)
#undef RYG_PICK
  return lg == 16 ? ryg_fn_lg<16>(acc) : (lg == 32 ? ryg_fn_lg<32>(acc) : ryg_fn_lg<64>(acc));
}

// Lane-group geometry of class D with LG lanes per group: accumulators per
// lane, doubles per group area, dynamic LDS of the block.
struct RygGeom {
  int acc = 0, area = 0;
  size_t lds = 0;
};
static RygGeom ryg_geom(const WClassDesc& D, int lg, bool jk) {
  RygGeom g;
  const int n = (D.mmax + 1) / 2;
  const int S = (D.la + D.lb + 1) * (D.lc + D.ld + 1);
  const int need = (D.nacc + lg - 1) / lg;
  g.acc = need <= 2 ? 2 : (need <= 4 ? 4 : (need <= 8 ? 8 : (need <= 16 ? 16 : 0)));
  g.area = 4 * RY_SCL + std::max(3 * n * S + 2 * n, D.nhrr + (jk ? D.nout : 0));
  g.area = (g.area + 1) & ~1;  // 16-byte aligned group areas
  g.lds = sizeof(double) * (size_t)(64 / lg) * g.area;
  return g;
}

#define RYG_LDS_MAX (64 * 1024)  // LDS a lane-group block may take (several blocks stay resident per CU)

bool rygclass_fits(const WClassDesc& D, int lg) {
  if (lg != 16 && lg != 32 && lg != 64) return false;
  if ((D.mmax + 1) / 2 > RYS_NMAX) return false;
  const RygGeom g = ryg_geom(D, lg, true);
  return g.acc > 0 && g.lds <= std::min((size_t)RYG_LDS_MAX, g_ry_lds_limit);
}

int ryg_kernel_blocks_per_cu(const WClassDesc& D, int lg) {
  const RygGeom g = ryg_geom(D, lg, false);
  int nb = 0;
  if (g.acc == 0 ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ryg_kernel_fn(D, lg, g.acc), 64, g.lds) != hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

hipError_t launch_ryg(const WClassDesc& D, int lg, const EriArgs& args, int nblocks, int64_t nq_launch,
                      const uint32_t* gblob, hipStream_t stream) {
  const RygGeom g = ryg_geom(D, lg, args.mode == MI_MODE_JK);
  if (g.acc == 0 || g.lds > g_ry_lds_limit) return hipErrorInvalidValue;
  EriArgs a = args;
  a.jk_run = 0;
  WClassDesc d = D;
  int area = g.area;
  void* kargs[] = {&a, &d, const_cast<const uint32_t**>(&gblob), &nq_launch, &area};
  return hipLaunchKernel(ryg_kernel_fn(D, lg, g.acc), dim3(nblocks), dim3(64), kargs, g.lds, stream);
}

static const void* ry_kernel_fn(int acc, bool hl) {
  switch (acc) {
    case 1: return hl ? (const void*)hgp_ry_kernel<1, true> : (const void*)hgp_ry_kernel<1, false>;
    case 2: return hl ? (const void*)hgp_ry_kernel<2, true> : (const void*)hgp_ry_kernel<2, false>;
    case 4: return hl ? (const void*)hgp_ry_kernel<4, true> : (const void*)hgp_ry_kernel<4, false>;
    case 8: return hl ? (const void*)hgp_ry_kernel<8, true> : (const void*)hgp_ry_kernel<8, false>;
    default: return hl ? (const void*)hgp_ry_kernel<24, true> : (const void*)hgp_ry_kernel<24, false>;
  }
}

// Accumulators per thread for nacc elements over nthr threads (0: too many).
static int ry_acc(int nacc, int nthr) {
  const int need = (nacc + nthr - 1) / nthr;
  if (need <= 1) return 1;
  if (need <= 2) return 2;
  if (need <= 4) return 4;
  if (need <= 8) return 8;
  return need <= 24 ? 24 : 0;
}

#define RY_LDS_TARGET (64 * 1024)      // LDS per block the batch size and the HRR placement aim for

// Launch geometry of class D with nw wavefronts per block: primitive quartets
// per batch (enough (primitive, root, direction) items for about two rounds
// of the block's threads, fewer to keep the block within RY_LDS_TARGET), the
// work area (the batch's 2D tables; for the J/K build also the scaled block),
// whether the HRR area fits in LDS beside it, and the dynamic LDS bytes.
struct RyGeom {
  int acc = 0, nbat = 1, work = 0;
  bool hl = false;
  size_t lds = 0;
};
static RyGeom ry_geom(const WClassDesc& D, int nw, bool jk) {
  RyGeom g;
  const int nthr = 64 * nw;
  const int n = (D.mmax + 1) / 2;
  const size_t S = (size_t)(D.la + D.lb + 1) * (D.lc + D.ld + 1);
  g.acc = ry_acc(D.nacc, nthr);
  g.nbat = std::max(1, (2 * nthr) / (3 * n));
  auto work = [&](int nbat) { return std::max((size_t)nbat * n * 3 * S, jk ? (size_t)D.nout : (size_t)0); };
  auto bytes = [&](int nbat, bool hl) {
    return sizeof(double) * ((size_t)4 * RY_SCL + work(nbat) + (hl ? (size_t)D.nhrr : 0));
  };
  while (g.nbat > 1 && bytes(g.nbat, false) > RY_LDS_TARGET) --g.nbat;
  g.hl = bytes(g.nbat, true) <= RY_LDS_TARGET;
  g.work = (int)work(g.nbat);
  g.lds = bytes(g.nbat, g.hl);
  return g;
}

// Class D runs in the Rys kernel with nw wavefronts per block (unique array,
// Schwarz pass and J/K build alike).
bool ryclass_fits_nw(const WClassDesc& D, int nw, size_t lds_limit) {
  if (nw != 1 && nw != 2 && nw != 4) return false;
  if ((D.mmax + 1) / 2 > RYS_NMAX) return false;
  const RyGeom g = ry_geom(D, nw, true);
  return g.acc > 0 && g.lds <= std::min(lds_limit, g_ry_lds_limit);
}

size_t ry_scratch_doubles(const WClassDesc& D, int nw, int nblocks) {
  // the J/K geometry keeps the HRR area in LDS only where the unique-array one does
  return ry_geom(D, nw, false).hl && ry_geom(D, nw, true).hl ? 0 : (size_t)nblocks * D.nhrr;
}

int ry_kernel_blocks_per_cu(const WClassDesc& D, int nw) {
  const RyGeom g = ry_geom(D, nw, false);
  int nb = 0;
  if (g.acc == 0 ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ry_kernel_fn(g.acc, g.hl), 64 * nw, g.lds) != hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

hipError_t launch_rys(const WClassDesc& D, int nw, const EriArgs& args, int nblocks, int64_t nq_launch,
                      const uint32_t* gblob, double* scratch, hipStream_t stream) {
  const RyGeom g = ry_geom(D, nw, args.mode == MI_MODE_JK);
  if (g.acc == 0 || g.lds > g_ry_lds_limit) return hipErrorInvalidValue;
  EriArgs a = args;
  a.jk_run = 0;
  WClassDesc d = D;
  int nbat = g.nbat, work = g.work;
  void* kargs[] = {&a, &d, const_cast<const uint32_t**>(&gblob), &scratch, &nq_launch, &nbat, &work};
  return hipLaunchKernel(ry_kernel_fn(g.acc, g.hl), dim3(nblocks), dim3(64 * nw), kargs, g.lds, stream);
}

hipError_t set_rys_lds_limit(size_t bytes) {
  for (int acc : {1, 2, 4, 8, 24})
    for (bool hl : {false, true}) {
      hipError_t e = hipFuncSetAttribute(ry_kernel_fn(acc, hl), hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
      if (e != hipSuccess) return e;
    }
  std::vector<const void*> fns;
  for (int acc : {2, 4, 8, 16}) {
    fns.push_back((const void*)ryg_fn_lg<16>(acc));
    fns.push_back((const void*)ryg_fn_lg<32>(acc));
    fns.push_back((const void*)ryg_fn_lg<64>(acc));
  }
#define RYG_FN(A, C, N, L, K) fns.push_back((const void*)hgp_ryg_kernel<L, K, A, C, N>);
  RYG_SPECS(RYG_FN)
#undef RYG_FN
  for (const void* f : fns) {
    hipError_t e =
        hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)std::min(bytes, (size_t)RYG_LDS_MAX));
    if (e != hipSuccess) return e;
  }
  g_ry_lds_limit = bytes;
  return hipSuccess;
}

}  // namespace mi
