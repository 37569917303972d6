// mi_device.h -- data layout and device helpers of the HGP ERI engine (gfx950).
//
// HBM layout (one plan = one basis):
//   PairTopo[npair]   shell-pair records sorted by (pair class, contraction
//                     bucket); a pair's shell A has L_A >= L_B, except in
//                     the swapped pair classes (L_A < L_B, the tighter
//                     shell on A; see orient_pair in mi_host.cpp).
//   PairGeom[npair]   AB = A - B, written by the pair kernel.
//   PrimPair[npp]     per primitive pair: P, zeta, PA = P - A, the folded
//                     prefactor K and 1/(2 zeta); 96-byte rows so one lane
//                     gathers a whole record with six 16-byte loads.
//   boys[mtop][481][8] Boys grid: row k (T_k = k/16) holds F_{mtop+j}(T_k)/j!,
//                     j = 0..6, and exp(-T_k).
//   out[]             caller-owned unique ERI array (iindex or chrr order).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BOYS_NROW 481          // T_k = k/16, k = 0..480 covers T < 30
#define BOYS_TMAX 30.0         // src/Boys.jl:19 (Tmax = 30)
#define BOYS_MTOP_MAX 16       // top m of the largest class supported by the table
#define BOYS_MCOL 23           // F_m columns computed on the host (m = 0..22)

struct uint4_t { uint32_t x, y, z, w; };
struct uint2_t { uint32_t x, y; };

struct PairTopo {
  int32_t shA, shB;   // oriented shells (L_A >= L_B, or a swapped pair class)
  int32_t bfA, bfB;   // first function of each shell (0-based)
  int32_t ppoff, npp; // primitive-pair range [ppoff, ppoff+npp)
  int32_t unused;
  int32_t pidx;       // canonical pair index tri(I,J) of the original shells
};

struct PairGeom {
  double ABx, ABy, ABz;
  double Q;   // max |(ab|ab)| over the pair's functions (Schwarz pass), 0 until computed
  double V;   // max |(e0|e0)| over the pair's contracted VRR outputs: bounds the HRR inputs
  double pad;
};

// One primitive pair, one 64-byte line: P itself is not stored, kernels form
// P = A + PA with the pair's first centre A exactly as the pair kernel does.
struct __attribute__((aligned(64))) PrimPair {
  double PAx, PAy, PAz, zeta;
  double K;                  // K = sqrt(2) pi^{5/4} c'_a c'_b exp(-ab|AB|^2/zeta) / zeta
  double oo2z, pad0, pad1;
};

struct TaskDesc {
  int32_t braStart, nBra, ketStart, nKet;
  int32_t same;        // bra list == ket list: triangular quartet space
  int32_t G;           // lanes per quartet (power of two, <= 64)
  int32_t blockStart;  // register kernels: first wavefront of the task in its launch;
                       // wave kernels: first shell quartet (launch-flat index)
  int32_t cls;         // quartet class id
  int32_t variant;     // 1 register kernel, 0 LDS wave kernel
  int32_t bf0;         // first function of the task's molecule (pairs use global indices)
  int64_t nq;          // shell quartets in the task
  int64_t nprim;       // primitive quartets in the task
  int64_t obase;       // output offset of the task's molecule
  int64_t M;           // nbf(nbf+1)/2 of the task's molecule
  double rnKet;        // 1 / nKet (the rectangle decode divides by nKet)
};

// What a launch does with each computed block element.
#define MI_MODE_UNIQUE 0   // write the unique (ij|kl) array
#define MI_MODE_SCHWARZ 1  // diagonal (ab|ab): atomic max of |value| into PairGeom.Q
#define MI_MODE_JK 2       // accumulate Coulomb/exchange against a density
#ifndef MI_JK_RUN
#define MI_JK_RUN 32       // J/K, wave kernels: longest run of consecutive quartets per block (< 2: off)
#endif

struct EriArgs {
  const PairTopo* topo;
  PairGeom* geom;
  const PrimPair* pp;
  const double* shell_xyz;  // shell centres (P = A + PA of a primitive pair)
  const double* bf_scale;
  const double* boys;
  double* out;
  const TaskDesc* tasks;
  const int32_t* block_task;  // register kernels: task of each block's first wavefront
  int32_t order;       // 0 iindex, 1 chrr
  int32_t ntasks;
  int32_t mode;        // MI_MODE_*
  int32_t nbf;         // J/K: matrix dimension
  double tau2;         // FP32 VRR when Vmax_ab * Vmax_cd < tau2 (0: never)
  double cut2;         // skip when Qmax_ab * Qmax_cd < cut2 (0: never; the output is pre-zeroed)
  const double* D;     // J/K: density, nbf x nbf row-major
  double* Jacc;        // J/K: unsymmetrised accumulators (J = Jacc + Jacc^T)
  double* Kacc;
  int32_t jk_run;      // J/K, wave kernels: quartets per contiguous run (0: one atomic per contribution)
  unsigned long long* run_ctr;  // J/K runs: next run of this launch (zero before the launch)
};

// Fused launch: several class launches of one kernel family issued as one
// grid.  Segment s (one class launch) takes blocks [bstart[s], bstart[s+1])
// and keeps its own tasks (task_off, ntasks), register-kernel block -> task
// table (bt_off) and wave-kernel quartet count (nq); within the segment a
// block runs exactly as block (blockIdx.x - bstart[s]) of the class launch.
#define MI_FUSE_MAX 24
struct FusedSegs {
  int32_t nseg;
  int32_t bstart[MI_FUSE_MAX + 1];
  int32_t cls[MI_FUSE_MAX];
  int32_t task_off[MI_FUSE_MAX];
  int32_t ntasks[MI_FUSE_MAX];
  int64_t bt_off[MI_FUSE_MAX];
  int64_t nq[MI_FUSE_MAX];
};

// Segment of fused-launch block b (uniform per block; nseg <= MI_FUSE_MAX).
__device__ __forceinline__ int fused_segment(const FusedSegs& F, int b) {
  int s = 0;
  while (s + 1 < F.nseg && b >= F.bstart[s + 1]) ++s;
  return s;
}

struct QuartetCtx {
  int bra, ket;
  int G, g;
  bool same_pair;
  int bf0;
  int64_t obase, M;
};

struct ScatterCtx {
  double* out;
  int64_t M;
  int order;
  int fa0, fb0, fc0, fd0;
  bool braSame, ketSame, quartetSame;
  int gmask, g;
  int bra;
};

// Wave-cooperative class descriptor (tables generated by gen_hgp.py).
struct WClassDesc {
  int la, lb, lc, ld, mmax;
  int nvrr, nhrr, vt0;       // vt0: LDS slot written by VRR item 0
  int nvlev, nhlev, nacc, nout;
  int boff, bwords;          // LDS-kernel op-table blob: 16-byte offset (-1: none) and word count
  int gv, gh, go, ga;        // global-kernel tables: word offsets of VRR, HRR, out, acc sections
  int vlev[16];  // level starts (VRR items)
  int hlev[10];
  int is_register;
  int gr;        // global tables: word offset of the Rys kernel's per-(e0|f0) Cartesian powers
};

// 1/sqrt(x) for a positive normal x: the hardware estimate (v_rsq_f64)
// refined by two Newton steps y <- y + y (1/2 - (x/2) y^2), each of which at
// least doubles the correct bits, so the result is within about one ulp.
// Callers form 1/x as y*y (within about two ulp).  The arguments here
// (exponent sums, Boys arguments T >= 30) are positive and normal, so this
// replaces a correctly rounded division plus square root, whose scaling and
// special-case handling cost about 25 instructions, with eight.
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double h = 0.5 * x;
  double y = __builtin_amdgcn_rsq(x);
  y = fma(y, fma(-h * y, y, 0.5), y);
  y = fma(y, fma(-h * y, y, 0.5), y);
  return y;
}

// ---------------------------------------------------------------------------
// Boys function F_m(T), m = 0..MMAX-1 (reference: farray, src/Boys.jl:19-27).
// T < 30: the top order from a 7-term Taylor expansion about the nearest grid
// point T_k = k/16 (|T - T_k| <= 1/32, truncation < 6e-15 relative), exp(-T) =
// exp(-T_k) * exp(T_k - T) by a degree-8 series, then the downward recursion
// F_m = (2T F_{m+1} + e^{-T}) / (2m+1).  T >= 30: boys_large_T below (the
// Boys function itself; the reference's farray branch there drops e^{-T}).
// ---------------------------------------------------------------------------
// T >= 30: F_0 = sqrt(pi/(4T)) erf(sqrt T) with erfc(x) = e^{-x^2}/(x sqrt pi)
// (1 - 1/(2x^2) + 3/(4x^4) - ...) (three terms: erfc(sqrt 30) = 1.2e-14 is
// itself below 1e-13 of F_0, so the truncated series is exact to rounding),
// then the upward recursion F_{m+1} = ((2m+1) F_m - e^{-T}) / (2T), stable
// for T > m + 1/2.  The reference's T >= 30 branch (src/Boys.jl:21-26) keeps
// only the first term and drops e^{-T}: 1e-4 relative on F_12 at T = 30 and
// up to 5e-10 relative on f-shell integrals, so it is not reproduced.
// y = 1/sqrt(T), rT = y^2 = 1/T (rsqrt_nr below).
__device__ __forceinline__ double boys_large_T_f0(double y, double rT, double eT) {
  return 0.88622692545275801365 * y - 0.5 * eT * rT * fma(rT, fma(0.75, rT, -0.5), 1.0);
}

__device__ __forceinline__ double exp_small(double d) {
  // e^d for |d| <= 1/32: Taylor to d^8 (remainder < 1e-19)
  double r = 1.0 / 40320.0;
  r = fma(r, d, 1.0 / 5040.0);
  r = fma(r, d, 1.0 / 720.0);
  r = fma(r, d, 1.0 / 120.0);
  r = fma(r, d, 1.0 / 24.0);
  r = fma(r, d, 1.0 / 6.0);
  r = fma(r, d, 0.5);
  r = fma(r, d, 1.0);
  r = fma(r, d, 1.0);
  return r;
}

// e^{-T} for T >= 30 where it enters F_0..F_{MMAX-1} only as a correction:
// relative to the term it corrects (F_0's erfc tail, the upward recursion's
// (2m-1) F_{m-1}), e^{-T} weighs at most 1.4e-10 for orders up to 3 (at T =
// 30, less above), so the single-precision hardware exponential (relative
// error ~2e-6 with the rounding of T) changes F_m by at most ~3e-16 relative
// -- at FP64 rounding -- for about three instructions instead of the
// double-precision exp's twenty-five.  Higher orders weigh e^{-T} up to 7e-5
// (m = 12) and take the double-precision exp.
template <int MMAX>
__device__ __forceinline__ double exp_minus_large_T(double T) {
  if (MMAX <= 4) return (double)__expf(-(float)T);
  return exp(-T);
}

template <int MMAX>
__device__ __forceinline__ void boys_eval(double T, const double* __restrict__ tab, double (&F)[MMAX]) {
  if (T < BOYS_TMAX) {
    const int k = (int)(T * 16.0 + 0.5);
    const double d = (double)k * 0.0625 - T;
    const double* r = tab + k * 8;
    double f = r[6];
    f = fma(f, d, r[5]);
    f = fma(f, d, r[4]);
    f = fma(f, d, r[3]);
    f = fma(f, d, r[2]);
    f = fma(f, d, r[1]);
    f = fma(f, d, r[0]);
    F[MMAX - 1] = f;
    if (MMAX > 1) {
      const double eT = r[7] * exp_small(d);
      const double t2 = T + T;
#pragma unroll
      for (int m = MMAX - 2; m >= 0; --m) F[m] = fma(t2, F[m + 1], eT) * (1.0 / (2 * m + 1));
    }
  } else {
    const double eT = exp_minus_large_T<MMAX>(T);
    const double y = rsqrt_nr(T), rT = y * y;
    F[0] = boys_large_T_f0(y, rT, eT);
    const double o2T = 0.5 * rT;
#pragma unroll
    for (int m = 1; m < MMAX; ++m) F[m] = fma((double)(2 * m - 1), F[m - 1], -eT) * o2T;
  }
}

// Copy n 16-byte chunks global -> LDS with all threads of the block: the main
// loop issues eight unconditional loads per thread before its stores (plain
// registers, all eight in flight), the tail moves the remainder one by one.
__device__ __forceinline__ void stage_chunks(uint4* __restrict__ dst, const uint4* __restrict__ src, int n) {
  const int st = blockDim.x;
  int i = threadIdx.x;
  for (; i + 7 * st < n; i += 8 * st) {
    const uint4 r0 = src[i], r1 = src[i + st], r2 = src[i + 2 * st], r3 = src[i + 3 * st];
    const uint4 r4 = src[i + 4 * st], r5 = src[i + 5 * st], r6 = src[i + 6 * st], r7 = src[i + 7 * st];
    dst[i] = r0;
    dst[i + st] = r1;
    dst[i + 2 * st] = r2;
    dst[i + 3 * st] = r3;
    dst[i + 4 * st] = r4;
    dst[i + 5 * st] = r5;
    dst[i + 6 * st] = r6;
    dst[i + 7 * st] = r7;
  }
  for (; i < n; i += st) dst[i] = src[i];
}

// Copy the Boys slice of one top order into LDS (all threads of the block).
__device__ __forceinline__ void stage_boys_slice(double* lds, const double* __restrict__ boys, int mtop) {
  stage_chunks(reinterpret_cast<uint4*>(lds), reinterpret_cast<const uint4*>(boys + (size_t)mtop * BOYS_NROW * 8),
               BOYS_NROW * 4);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Task/quartet decoding.  A launch covers one quartet class; blocks are
// assigned to tasks by blockStart (tasks sorted by it).  A task pairs two
// contiguous pair lists; same==1 means the triangular space p <= q.
// ---------------------------------------------------------------------------
// Task holding launch-flat quartet index q (tasks sorted by blockStart).
__device__ __forceinline__ int find_task(const EriArgs& a, int64_t q) {
  int lo = 0, hi = a.ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (a.tasks[mid].blockStart <= q) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// The same, starting from a task ti known to start at or before q: gallop
// forward, then bisect (cost logarithmic in the distance, not in ntasks).
__device__ __forceinline__ int advance_task(const EriArgs& a, int ti, int64_t q) {
  int lo = ti, hi = ti + 1, step = 1;
  while (hi < a.ntasks && a.tasks[hi].blockStart <= q) {
    lo = hi;
    step <<= 1;
    hi = lo + step;
  }
  if (hi > a.ntasks) hi = a.ntasks;
  while (hi - lo > 1) {  // tasks[lo] <= q < tasks[hi]
    const int mid = (lo + hi) >> 1;
    if (a.tasks[mid].blockStart <= q) lo = mid; else hi = mid;
  }
  return lo;
}

// floor(a / b) for 0 <= a < 2^31 and 1 <= b < 2^31, given rb ~ 1/b (double,
// relative error below 2^-40): the product is within one of the quotient, so
// one correction each way makes it exact; r = a - b floor(a / b).  About ten
// instructions instead of an integer division's twenty-five or more.
__device__ __forceinline__ int udiv_rcp(int a, int b, double rb, int& r) {
  int q = (int)((double)a * rb);
  r = a - q * b;
  if (r < 0) {
    --q;
    r += b;
  }
  if (r >= b) {
    ++q;
    r -= b;
  }
  return q;
}
// 1/b for 1 <= b < 2^31: the hardware estimate plus one Newton step.
__device__ __forceinline__ double rcp_int(int b) {
  const double x = (double)b;
  const double y = __builtin_amdgcn_rcp(x);
  return fma(y, fma(-x, y, 1.0), y);
}

// qi: a quartet index within the task (tasks and launches hold fewer than
// 2^31 quartets, which the host checks).
__device__ __forceinline__ void decode_index(const TaskDesc& T, int64_t qi64, QuartetCtx& qc) {
  const int qi = (int)qi64;
  if (T.same == 2) {  // Schwarz pass: the diagonal quartets (p|p)
    qc.bra = qc.ket = T.braStart + qi;
  } else if (T.same) {
    // largest q with q(q+1)/2 <= qi: the float square root is within one of it
    int q = (int)((__builtin_sqrtf(8.0f * (float)qi + 1.0f) - 1.0f) * 0.5f);
    while ((int64_t)q * (q + 1) / 2 > qi) --q;
    while ((int64_t)(q + 1) * (q + 2) / 2 <= qi) ++q;
    const int p = qi - (int)((int64_t)q * (q + 1) / 2);
    qc.bra = T.braStart + p;
    qc.ket = T.braStart + q;
  } else {
    int q;
    const int p = udiv_rcp(qi, T.nKet, T.rnKet, q);
    qc.bra = T.braStart + p;
    qc.ket = T.ketStart + q;
  }
  qc.same_pair = (qc.bra == qc.ket);
  qc.bf0 = T.bf0;
  qc.obase = T.obase;
  qc.M = T.M;
}

// Lane-group decode for the register kernels: G lanes per quartet.  Each task
// of a launch owns ceil(nq G / 64) whole wavefronts (blockStart = its first
// wavefront), so a wavefront never spans two tasks, lane groups stay aligned
// whatever each task's G, and blocks stay full across task boundaries.
// bid: the block's index within its class launch (blockIdx.x, or the block's
// offset in its segment of a fused launch).
__device__ __forceinline__ bool decode_quartet(const EriArgs& a, QuartetCtx& qc, int bid) {
  const int64_t w0 = (int64_t)bid * (blockDim.x >> 6);
  const int64_t w = w0 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // the task of the block's first wavefront (host table, else a block-uniform search)
  const int t0 = a.block_task ? a.block_task[bid] : find_task(a, w0);
  const TaskDesc T = a.tasks[advance_task(a, t0, w)];
  const int G = T.G;  // a power of two
  const int64_t qi = (int64_t)((((uint32_t)(w - T.blockStart)) << 6 | (uint32_t)lane) >> __builtin_ctz((unsigned)G));
  qc.G = G;
  qc.g = lane & (G - 1);
  // a lane past the task's last quartet gets the task's first quartet (valid
  // indices; the caller drops it or, in the J/K build, contributes nothing)
  const bool live = qi < T.nq;
  decode_index(T, live ? qi : 0, qc);
  return live;
}

template <int NA, int NB, int NC, int ND>
__device__ __forceinline__ void load_scales(const EriArgs& a, const PairTopo& P, const PairTopo& Q,
                                            double (&sA)[NA], double (&sB)[NB], double (&sC)[NC], double (&sD)[ND]) {
#pragma unroll
  for (int i = 0; i < NA; ++i) sA[i] = a.bf_scale[P.bfA + i];
#pragma unroll
  for (int i = 0; i < NB; ++i) sB[i] = a.bf_scale[P.bfB + i];
#pragma unroll
  for (int i = 0; i < NC; ++i) sC[i] = a.bf_scale[Q.bfA + i];
#pragma unroll
  for (int i = 0; i < ND; ++i) sD[i] = a.bf_scale[Q.bfB + i];
}

__device__ __forceinline__ ScatterCtx make_scatter(const EriArgs& a, const QuartetCtx& qc, const PairTopo& P,
                                                   const PairTopo& Q) {
  ScatterCtx s;
  s.out = a.out + qc.obase;
  s.M = qc.M;
  s.order = a.order;
  s.fa0 = P.bfA - qc.bf0;
  s.fb0 = P.bfB - qc.bf0;
  s.fc0 = Q.bfA - qc.bf0;
  s.fd0 = Q.bfB - qc.bf0;
  s.braSame = (P.shA == P.shB);
  s.ketSame = (Q.shA == Q.shB);
  s.quartetSame = qc.same_pair;
  s.gmask = qc.G - 1;
  s.g = qc.g;
  s.bra = qc.bra;
  return s;
}

// Unique slot of the block element (fa fb|fc fd) or -1 when the element is a
// duplicate inside this shell quartet.  Slot algebra: src/Utils.jl:14-19 and
// the enumeration order of src/Basis.jl:203.
// Pair indices in 32 bits (molecules hold at most 65535 functions, so
// i(i-1) < 2^32 and M + 1 < 2^32), slot products as 32x32->64 multiplies.
__device__ __forceinline__ int64_t unique_slot(const ScatterCtx& s, int fa, int fb, int fc, int fd) {
  if (s.braSame && fa < fb) return -1;
  if (s.ketSame && fc < fd) return -1;
  const uint32_t i = (uint32_t)(fa > fb ? fa : fb) + 1u, j = (uint32_t)(fa > fb ? fb : fa) + 1u;
  const uint32_t k = (uint32_t)(fc > fd ? fc : fd) + 1u, l = (uint32_t)(fc > fd ? fd : fc) + 1u;
  const uint32_t ij = ((i * (i - 1u)) >> 1) + j, kl = ((k * (k - 1u)) >> 1) + l;
  if (s.quartetSame && ij > kl) return -1;
  const uint32_t lo = ij < kl ? ij : kl, hi = ij < kl ? kl : ij;
  if (s.order == 0) return (int64_t)(((uint64_t)hi * (uint64_t)(hi - 1u)) >> 1) + (int64_t)lo - 1;
  return (int64_t)((uint64_t)(lo - 1u) * (uint64_t)(s.M + 1)) - (int64_t)(((uint64_t)lo * (uint64_t)(lo - 1u)) >> 1) +
         (int64_t)(hi - lo);
}

// 0: FP64 VRR, 1: FP32 VRR, 2: screened.  Screening uses the Schwarz bound of
// the integrals, Q_ab Q_cd < cut; FP32 VRR uses the Schwarz bound of the VRR
// outputs (e0|f0) themselves, V_ab V_cd < tau (each (e0|f0) is an integral
// (e_A s_B|f_C s_D), so |(e0|f0)| <= sqrt(V_ab V_cd)): HRR cancellation can
// make (e0|f0) much larger than the final integral, so bounding only the
// final value would not bound the FP32 rounding.
__device__ __forceinline__ int quartet_mode(const EriArgs& a, const QuartetCtx& qc) {
  if (a.tau2 <= 0.0 && a.cut2 <= 0.0) return 0;
  const PairGeom& gb = a.geom[qc.bra];
  const PairGeom& gk = a.geom[qc.ket];
  if (gb.Q * gk.Q < a.cut2) return 2;
  return gb.V * gk.V < a.tau2 ? 1 : 0;
}

// Schwarz pass: fold |x| into a pair's V (non-negative doubles order like their bits).
__device__ __forceinline__ void record_vmax(PairGeom* g, double x) {
  atomicMax(reinterpret_cast<unsigned long long*>(&g->V), (unsigned long long)__double_as_longlong(fabs(x)));
}

// One element (fa fb|fc fd) of a shell-quartet block (molecule-local indices).
__device__ __forceinline__ void scatter_elem(const EriArgs& a, const ScatterCtx& s, int bra, int fa, int fb, int fc,
                                             int fd, double val) {
  if (a.mode == MI_MODE_UNIQUE) {
    const int64_t slot = unique_slot(s, fa, fb, fc, fd);
    if (slot >= 0) s.out[slot] = val;
    return;
  }
  if (a.mode == MI_MODE_SCHWARZ) {
    if (fa == fc && fb == fd)  // non-negative doubles order like their bit patterns
      atomicMax(reinterpret_cast<unsigned long long*>(&a.geom[bra].Q),
                (unsigned long long)__double_as_longlong(fabs(val)));
    return;
  }
  // Coulomb/exchange (index algebra of make2JmK, src/ERI.jl:114-128): each
  // unique (ij|kl) stands for its 8 permutations; with the degeneracy weight
  // w = v / (2^#coincidences) they contribute Jacc[i][j] += 2w D_kl,
  // Jacc[k][l] += 2w D_ij, Kacc[i][k] += w D_jl, Kacc[j][k] += w D_il,
  // Kacc[i][l] += w D_jk, Kacc[j][l] += w D_ik, and J = Jacc + Jacc^T,
  // K = Kacc + Kacc^T.
  if (unique_slot(s, fa, fb, fc, fd) < 0) return;
  int i = fa > fb ? fa : fb, j = fa > fb ? fb : fa, k = fc > fd ? fc : fd, l = fc > fd ? fd : fc;
  const int64_t ij = (int64_t)i * (i + 1) / 2 + j, kl = (int64_t)k * (k + 1) / 2 + l;
  if (ij > kl) {
    int t = i; i = k; k = t;
    t = j; j = l; l = t;
  }
  double w = val;
  if (i == j) w *= 0.5;
  if (k == l) w *= 0.5;
  if (i == k && j == l) w *= 0.5;
  const int n = a.nbf;
  const double* D = a.D;
  atomicAdd(&a.Jacc[(int64_t)i * n + j], 2.0 * w * D[(int64_t)k * n + l]);
  atomicAdd(&a.Jacc[(int64_t)k * n + l], 2.0 * w * D[(int64_t)i * n + j]);
  atomicAdd(&a.Kacc[(int64_t)i * n + k], w * D[(int64_t)j * n + l]);
  atomicAdd(&a.Kacc[(int64_t)j * n + k], w * D[(int64_t)i * n + l]);
  atomicAdd(&a.Kacc[(int64_t)i * n + l], w * D[(int64_t)j * n + k]);
  atomicAdd(&a.Kacc[(int64_t)j * n + l], w * D[(int64_t)i * n + k]);
}

// Coulomb/exchange from one whole shell-quartet block V[a][b][c][d]
// (na nb nc nd, scaled values).  Shell-level degeneracy: the block of a
// unique shell quartet carries every element of its diagonal sub-blocks, so
// f = 1/2 per coincidence (A = B, C = D, AB = CD) weights each of the 8
// permutations once (the element-level rule of make2JmK, src/ERI.jl:114-128,
// summed over the block):
//   Jacc[ab] += 2f sum_cd V D_cd    Jacc[cd] += 2f sum_ab V D_ab
//   Kacc[ac] += f sum_bd V D_bd     Kacc[ad] += f sum_bc V D_bc
//   Kacc[bc] += f sum_ad V D_ad     Kacc[bd] += f sum_ac V D_ac
// and J = Jacc + Jacc^T, K = Kacc + Kacc^T afterwards.  One atomic per
// output element of the six contractions instead of six per integral.
__device__ __forceinline__ double jk_factor(const ScatterCtx& s) {
  return (s.braSame ? 0.5 : 1.0) * (s.ketSame ? 0.5 : 1.0) * (s.quartetSame ? 0.5 : 1.0);
}

// Segments of a wavefront's quartet slots (G lanes each, slot = lane / G):
// maximal runs of consecutive slots with equal key.  start: first slot of
// this slot's run; tail: last slot of its run; merge: some run is longer than
// one slot (wavefront-uniform).
struct WaveSeg {
  int slot, nslot, G, start;
  bool tail, merge;
};

__device__ __forceinline__ WaveSeg wave_segments(int gmask, bool live, uint32_t key) {
  WaveSeg r;
  r.G = gmask + 1;
  r.slot = (threadIdx.x & 63) / r.G;
  r.nslot = 64 / r.G;
  if (!live) key = 0xFFFF0000u | (uint32_t)r.slot;  // never equal to a live key (offsets < 65535)
  const uint32_t prev = __shfl_up(key, r.G, 64), next = __shfl_down(key, r.G, 64);
  const bool head = r.slot == 0 || prev != key;
  r.tail = r.slot == r.nslot - 1 || next != key;
  r.merge = __any(!head);
  int st = head ? r.slot : 0;
  if (r.merge)
    for (int d = 1; d < r.nslot; d <<= 1) {
      const int o = __shfl_up(st, d * r.G, 64);
      if (r.slot >= d) st = max(st, o);
    }
  r.start = st;
  return r;
}

// Segmented inclusive sum over the slots of the run (lanes of equal g): the
// run's tail slot ends with the run's total.  Every lane of the wavefront calls it.
__device__ __forceinline__ double wave_segment_sum(const WaveSeg& r, double x) {
  if (!r.merge) return x;
  for (int d = 1; d < r.nslot; d <<= 1) {
    const double o = __shfl_up(x, d * r.G, 64);
    if (r.slot - d >= r.start) x += o;
  }
  return x;
}

// Register-resident block (generated kernels): outputs t = 0..T-1 of the six
// contractions are split over the G lanes of the quartet (t mod G).  The
// J_ab, K_ac and K_bc targets repeat over consecutive quartets of a task row
// (same bra pair, and the same ket shell C in plans laid out for J/K), so
// those contributions are first summed over each run of equal targets in the
// wavefront, and only the run's last quartet issues the atomic.  Every lane of
// the wavefront calls this; live = false marks a lane without a quartet (its
// block is zero and it issues no atomic).
template <int NA, int NB, int NC, int ND>
__device__ __forceinline__ void jk_block_regs(const EriArgs& a, const ScatterCtx& s, bool live,
                                              const double (&v)[NA * NB * NC * ND]) {
  const double f = jk_factor(s);
  const int64_t n = a.nbf;
  const double* __restrict__ D = a.D;
  const int64_t ra = s.fa0 * n, rb = s.fb0 * n, rc = s.fc0 * n;
  const WaveSeg rab = wave_segments(s.gmask, live, (uint32_t)s.fa0 << 16 | (uint32_t)s.fb0);
  const WaveSeg rac = wave_segments(s.gmask, live, (uint32_t)s.fa0 << 16 | (uint32_t)s.fc0);
  const WaveSeg rbc = wave_segments(s.gmask, live, (uint32_t)s.fb0 << 16 | (uint32_t)s.fc0);
#define V_(i, j, k, l) v[(((i) * NB + (j)) * NC + (k)) * ND + (l)]
  constexpr int T0 = NA * NB, T1 = T0 + NC * ND, T2 = T1 + NA * NC, T3 = T2 + NA * ND, T4 = T3 + NB * NC;
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const bool own = ((i * NB + j) & s.gmask) == s.g;
      double acc = 0.0;
      if (own) {
#pragma unroll
        for (int k = 0; k < NC; ++k)
#pragma unroll
          for (int l = 0; l < ND; ++l) acc += V_(i, j, k, l) * D[rc + k * n + s.fd0 + l];
      }
      acc = wave_segment_sum(rab, 2.0 * f * acc);
      if (own && live && rab.tail) atomicAdd(&a.Jacc[ra + i * n + s.fb0 + j], acc);
    }
#pragma unroll
  for (int k = 0; k < NC; ++k)
#pragma unroll
    for (int l = 0; l < ND; ++l)
      if (live && ((T0 + k * ND + l) & s.gmask) == s.g) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j) acc += V_(i, j, k, l) * D[ra + i * n + s.fb0 + j];
        atomicAdd(&a.Jacc[rc + k * n + s.fd0 + l], 2.0 * f * acc);
      }
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const bool own = ((T1 + i * NC + k) & s.gmask) == s.g;
      double acc = 0.0;
      if (own) {
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int l = 0; l < ND; ++l) acc += V_(i, j, k, l) * D[rb + j * n + s.fd0 + l];
      }
      acc = wave_segment_sum(rac, f * acc);
      if (own && live && rac.tail) atomicAdd(&a.Kacc[ra + i * n + s.fc0 + k], acc);
    }
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int l = 0; l < ND; ++l)
      if (live && ((T2 + i * ND + l) & s.gmask) == s.g) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int k = 0; k < NC; ++k) acc += V_(i, j, k, l) * D[rb + j * n + s.fc0 + k];
        atomicAdd(&a.Kacc[ra + i * n + s.fd0 + l], f * acc);
      }
#pragma unroll
  for (int j = 0; j < NB; ++j)
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const bool own = ((T3 + j * NC + k) & s.gmask) == s.g;
      double acc = 0.0;
      if (own) {
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
          for (int l = 0; l < ND; ++l) acc += V_(i, j, k, l) * D[ra + i * n + s.fd0 + l];
      }
      acc = wave_segment_sum(rbc, f * acc);
      if (own && live && rbc.tail) atomicAdd(&a.Kacc[rb + j * n + s.fc0 + k], acc);
    }
#pragma unroll
  for (int j = 0; j < NB; ++j)
#pragma unroll
    for (int l = 0; l < ND; ++l)
      if (live && ((T4 + j * ND + l) & s.gmask) == s.g) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
          for (int k = 0; k < NC; ++k) acc += V_(i, j, k, l) * D[ra + i * n + s.fc0 + k];
        atomicAdd(&a.Kacc[rb + j * n + s.fd0 + l], f * acc);
      }
#undef V_
}

// Block staged in LDS (wave kernels): thread tid of nthr takes outputs
// t = tid, tid + nthr, ... of the six contractions, ordered J_ab, K_ac, K_bc
// (the targets shared by consecutive quartets of a task row), then J_cd,
// K_ad, K_bd.  With racc (LDS, one double per output of the first three
// groups), the first three groups are summed there over a run of quartets
// and written when their target changes (jk_run_flush); output t keeps the
// same thread for every quartet, so racc[t] needs no synchronisation.
// Next run of jk_run quartets for the block (dynamic: runs differ in cost, so
// blocks take them from a launch-wide counter until the launch is done).
__device__ __forceinline__ int64_t next_run(const EriArgs& a, unsigned long long* s_run) {
  __syncthreads();  // every thread has read the previous run
  if (threadIdx.x == 0) *s_run = atomicAdd(a.run_ctr, 1ull);
  __syncthreads();
  return (int64_t)(*s_run) * a.jk_run;
}

struct JkRun {
  uint32_t kab, kac, kbc;  // keys of the targets held in racc (kNoRun: none)
  int fa0, fb0, fc0;
};
constexpr uint32_t kNoRun = 0xFFFFFFFFu;

__device__ __forceinline__ void jk_run_flush(const EriArgs& a, JkRun& r, double* racc, int na, int nb, int nc,
                                             bool fab, bool fac, bool fbc, int tid, int nthr) {
  const int64_t n = a.nbf;
  const int nab = na * nb, nac = na * nc, nbc = nb * nc;
  fab = fab && r.kab != kNoRun;
  fac = fac && r.kac != kNoRun;
  fbc = fbc && r.kbc != kNoRun;
  if (!(fab || fac || fbc)) return;
  for (int t = tid; t < nab + nac + nbc; t += nthr) {
    int u = t;
    if (u < nab) {
      if (!fab) continue;
      const int i = u / nb, j = u - i * nb;
      atomicAdd(&a.Jacc[(int64_t)(r.fa0 + i) * n + r.fb0 + j], racc[t]);
    } else if ((u -= nab) < nac) {
      if (!fac) continue;
      const int i = u / nc, k = u - i * nc;
      atomicAdd(&a.Kacc[(int64_t)(r.fa0 + i) * n + r.fc0 + k], racc[t]);
    } else {
      if (!fbc) continue;
      u -= nac;
      const int j = u / nc, k = u - j * nc;
      atomicAdd(&a.Kacc[(int64_t)(r.fb0 + j) * n + r.fc0 + k], racc[t]);
    }
    racc[t] = 0.0;
  }
  if (fab) r.kab = kNoRun;
  if (fac) r.kac = kNoRun;
  if (fbc) r.kbc = kNoRun;
}

__device__ __forceinline__ void jk_block_lds(const EriArgs& a, const ScatterCtx& s, const double* V, int na, int nb,
                                             int nc, int nd, int tid, int nthr, double* racc, JkRun& run) {
  const double f = jk_factor(s);
  const int64_t n = a.nbf;
  const double* __restrict__ D = a.D;
  const int nab = na * nb, ncd = nc * nd, nac = na * nc, nad = na * nd, nbc = nb * nc, nbd = nb * nd;
  const int total = nab + ncd + nac + nad + nbc + nbd;
  const int nrun = racc ? nab + nac + nbc : 0;
  if (racc) {
    const uint32_t kab = (uint32_t)s.fa0 << 16 | (uint32_t)s.fb0, kac = (uint32_t)s.fa0 << 16 | (uint32_t)s.fc0,
                   kbc = (uint32_t)s.fb0 << 16 | (uint32_t)s.fc0;
    jk_run_flush(a, run, racc, na, nb, nc, kab != run.kab, kac != run.kac, kbc != run.kbc, tid, nthr);
    run.kab = kab;
    run.kac = kac;
    run.kbc = kbc;
    run.fa0 = s.fa0;
    run.fb0 = s.fb0;
    run.fc0 = s.fc0;
  }
  for (int t = tid; t < total; t += nthr) {
    int u = t;
    double acc = 0.0, w = f;
    double* dst;
    if (u < nab) {
      const int i = u / nb, j = u - i * nb;
      for (int k = 0; k < nc; ++k)
        for (int l = 0; l < nd; ++l) acc += V[((i * nb + j) * nc + k) * nd + l] * D[(s.fc0 + k) * n + s.fd0 + l];
      dst = &a.Jacc[(s.fa0 + i) * n + s.fb0 + j];
      w = 2.0 * f;
    } else if ((u -= nab) < nac) {
      const int i = u / nc, k = u - i * nc;
      for (int j = 0; j < nb; ++j)
        for (int l = 0; l < nd; ++l) acc += V[((i * nb + j) * nc + k) * nd + l] * D[(s.fb0 + j) * n + s.fd0 + l];
      dst = &a.Kacc[(s.fa0 + i) * n + s.fc0 + k];
    } else if ((u -= nac) < nbc) {
      const int j = u / nc, k = u - j * nc;
      for (int i = 0; i < na; ++i)
        for (int l = 0; l < nd; ++l) acc += V[((i * nb + j) * nc + k) * nd + l] * D[(s.fa0 + i) * n + s.fd0 + l];
      dst = &a.Kacc[(s.fb0 + j) * n + s.fc0 + k];
    } else if ((u -= nbc) < ncd) {
      const int k = u / nd, l = u - k * nd;
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) acc += V[((i * nb + j) * nc + k) * nd + l] * D[(s.fa0 + i) * n + s.fb0 + j];
      dst = &a.Jacc[(s.fc0 + k) * n + s.fd0 + l];
      w = 2.0 * f;
    } else if ((u -= ncd) < nad) {
      const int i = u / nd, l = u - i * nd;
      for (int j = 0; j < nb; ++j)
        for (int k = 0; k < nc; ++k) acc += V[((i * nb + j) * nc + k) * nd + l] * D[(s.fb0 + j) * n + s.fc0 + k];
      dst = &a.Kacc[(s.fa0 + i) * n + s.fd0 + l];
    } else {
      u -= nad;
      const int j = u / nd, l = u - j * nd;
      for (int i = 0; i < na; ++i)
        for (int k = 0; k < nc; ++k) acc += V[((i * nb + j) * nc + k) * nd + l] * D[(s.fa0 + i) * n + s.fc0 + k];
      dst = &a.Kacc[(s.fb0 + j) * n + s.fd0 + l];
    }
    if (t < nrun)
      racc[t] += w * acc;
    else
      atomicAdd(dst, w * acc);
  }
}

// Element e of a block with NB, NC, ND functions on shells B, C, D.
template <int NB, int NC, int ND>
__device__ __forceinline__ void scatter_block_elem(const EriArgs& a, const ScatterCtx& s, int e, double v) {
  const int md = e % ND, mc = (e / ND) % NC, mb = (e / (ND * NC)) % NB, ma = e / (ND * NC * NB);
  scatter_elem(a, s, s.bra, s.fa0 + ma, s.fb0 + mb, s.fc0 + mc, s.fd0 + md, v);
}

__device__ __forceinline__ void scatter_one(const EriArgs& a, const ScatterCtx& s, int idx, int ma, int mb, int mc,
                                            int md, double val) {
  if ((idx & s.gmask) != s.g) return;
  scatter_elem(a, s, s.bra, s.fa0 + ma, s.fb0 + mb, s.fc0 + mc, s.fd0 + md, val);
}
