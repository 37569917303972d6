// mi_kernels.hip -- HGP ERI device kernels for gfx950 (CDNA4).
//
// Per execute:
//   1. pair_prims_kernel   one lane per primitive pair: P, zeta, PA, folded
//                          prefactor K, 1/(2 zeta); lane 0 of a pair writes AB.
//   2. per quartet class   a generated register kernel (hgp_r_<class>, lane
//                          group per shell quartet), the table-driven wave
//                          kernel hgp_w_kernel (1-4 wavefronts per shell
//                          quartet, tables and VRR/HRR in LDS), or, for the
//                          largest f classes, hgp_wg_kernel (a block per shell
//                          quartet, VRR in LDS, HRR in global scratch).
// MFMA is not used: HGP's VRR/HRR are short dependent FMA chains with
// per-quartet coefficients, not dense contractions.
#include <algorithm>

#include "mi_device.h"
#include "mi_kernels.h"

#include "gen/hgp_wtables.inc"
#include "gen/hgp_rkernels.inc"
#include "gen/hgp_rfused.inc"  // HGP_NRGROUP (no kernel without MI_FUSED_GROUP)

// ---------------------------------------------------------------------------
// Shell-pair precompute (reference: the pair part of vrr!'s prefactor block,
// src/HGP.jl:139-155, hoisted out of the quartet loop).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) pair_prims_kernel(int npp, const int32_t* __restrict__ pp_pair,
                                                         const int32_t* __restrict__ pp_ab,
                                                         const PairTopo* __restrict__ topo,
                                                         const double* __restrict__ shell_xyz,
                                                         const int32_t* __restrict__ shell_prim_off,
                                                         const double* __restrict__ prim_exp,
                                                         const double* __restrict__ prim_cn,
                                                         PrimPair* __restrict__ pp, PairGeom* __restrict__ geom,
                                                         const double2* __restrict__ bounds) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= npp) return;
  const int p = pp_pair[r];
  const PairTopo t = topo[p];
  const int local = r - t.ppoff;
  const int ia = pp_ab[r] & 0xFFFF, ib = pp_ab[r] >> 16;
  const int pa = shell_prim_off[t.shA] + ia, pb = shell_prim_off[t.shB] + ib;
  const double a = prim_exp[pa], b = prim_exp[pb];
  const double Ax = shell_xyz[3 * t.shA], Ay = shell_xyz[3 * t.shA + 1], Az = shell_xyz[3 * t.shA + 2];
  const double Bx = shell_xyz[3 * t.shB], By = shell_xyz[3 * t.shB + 1], Bz = shell_xyz[3 * t.shB + 2];
  const double zeta = a + b, ooz = 1.0 / zeta;
  const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;
  const double rab2 = ABx * ABx + ABy * ABy + ABz * ABz;
  // PA = P - A = -(b/zeta) AB and P = A + PA (formed by the kernels): exactly
  // A for one-centre pairs, so one-centre quartets keep PA = QC = W - P = 0
  // without rounding residue (the textbook (aA + bB)/zeta leaves ~1e-16|A|
  // that tight-exponent prefactors amplify past 1e-12).
  const double bz = b * ooz;
  const double PAx = -bz * ABx, PAy = -bz * ABy, PAz = -bz * ABz;
  // 2 pi^{5/2} / (zeta eta sqrt(zeta+eta)) exp(...) split as K_ab K_cd / sqrt(zeta+eta)
  const double kfac = 5.9149671727956129;  // sqrt(2) * pi^{5/4}
  PrimPair o;
  o.zeta = zeta;
  o.PAx = PAx; o.PAy = PAy; o.PAz = PAz;
  o.K = kfac * prim_cn[pa] * prim_cn[pb] * exp(-a * b * rab2 * ooz) * ooz;
  o.oo2z = 0.5 * ooz;
  o.pad0 = o.pad1 = 0.0;
  pp[r] = o;
  if (local == 0) {
    PairGeom g;
    g.ABx = ABx; g.ABy = ABy; g.ABz = ABz; g.pad = 0.0;
    // Schwarz bounds computed when the plan was built, or zero for a Schwarz pass to fill
    g.Q = bounds ? bounds[p].x : 0.0;
    g.V = bounds ? bounds[p].y : 0.0;
    geom[p] = g;
  }
}

// ---------------------------------------------------------------------------
// Wave-cooperative kernel: one 64-lane block per shell quartet.  Per primitive
// quartet the lanes fill the VRR levels of the class (items of one level are
// independent; a barrier separates levels), each lane keeps up to W_ACC_MAX
// contracted (e0|f0) values in registers; after the primitive loop the HRR
// levels run in LDS and the lanes scatter the block.  Item descriptors are
// loaded four per lane before use so their latencies overlap.
// ---------------------------------------------------------------------------
#define W_ACC_MAX 16   // contracted values per lane: nacc <= 1024 (classes up to (dd|dd))
#define W_SCL 16       // LDS stride of the per-shell scale rows
#define W_CF 48        // per-wavefront primitive coefficient block (indices: gen_hgp.py CF_*)

struct WTablesDev {
  const uint32_t* blob;  // kWBlob on the device
};

__constant__ double kInvOdd[32] = {
    1.0 / 1,  1.0 / 3,  1.0 / 5,  1.0 / 7,  1.0 / 9,  1.0 / 11, 1.0 / 13, 1.0 / 15, 1.0 / 17, 1.0 / 19, 1.0 / 21,
    1.0 / 23, 1.0 / 25, 1.0 / 27, 1.0 / 29, 1.0 / 31, 1.0 / 33, 1.0 / 35, 1.0 / 37, 1.0 / 39, 1.0 / 41, 1.0 / 43,
    1.0 / 45, 1.0 / 47, 1.0 / 49, 1.0 / 51, 1.0 / 53, 1.0 / 55, 1.0 / 57, 1.0 / 59, 1.0 / 61, 1.0 / 63};

__device__ __forceinline__ double sel3(int i, double x, double y, double z) { return i == 0 ? x : (i == 1 ? y : z); }

__device__ __forceinline__ int ncart_dev(int l) { return (l + 1) * (l + 2) / 2; }

// F_m(T) for one order m <= mtop: grid Taylor value of the top order, then
// downward recursion to m (T < 30); F_0 and the upward recursion of
// boys_large_T_f0 otherwise (src/Boys.jl:19-34).  btab is the grid of the top
// order mtop.
__device__ __forceinline__ double boys_lane(const double* __restrict__ btab, double T2, int mtop, int m) {
  double fm;
  if (T2 < BOYS_TMAX) {
    const int k = (int)(T2 * 16.0 + 0.5);
    const double d = (double)k * 0.0625 - T2;
    const double* r = btab + k * 8;
    double f = r[6];
    f = fma(f, d, r[5]);
    f = fma(f, d, r[4]);
    f = fma(f, d, r[3]);
    f = fma(f, d, r[2]);
    f = fma(f, d, r[1]);
    f = fma(f, d, r[0]);
    const double eT = r[7] * exp_small(d);
    const double t2 = T2 + T2;
    fm = f;
    for (int j = mtop - 1; j >= m; --j) fm = fma(t2, fm, eT) * kInvOdd[j];
  } else {
    const double eT = exp(-T2);
    const double y = rsqrt_nr(T2), rT = y * y;
    fm = boys_large_T_f0(y, rT, eT);
    const double o2T = 0.5 * rT;
    for (int j = 1; j <= m; ++j) fm = fma((double)(2 * j - 1), fm, -eT) * o2T;
  }
  return fm;
}

// LDS layout of a wave-kernel block: the class's op tables (staged once per
// block, then read at LDS latency), followed by the per-quartet work areas.
__host__ __device__ __forceinline__ int w_desc_doubles(const WClassDesc& D) { return D.bwords / 2; }

// Persistent: each block stages the tables, then walks quartets
// blockIdx.x, blockIdx.x + gridDim.x, ... of the launch (tasks hold the
// first launch-quartet index of each task in blockStart).  A block has nw
// wavefronts: wave w evaluates primitive quartets w, w+nw, ... in its own
// VRR area (levels separated by wave-local syncs) and keeps its share of the
// contracted (e0|f0) in registers; the shares are summed in wave order (a
// fixed order: results do not depend on scheduling), then all waves run the
// HRR levels and the scatter.  In the J/K build with args.jk_run > 0 a block
// takes runs of jk_run consecutive quartets instead (runs blockIdx.x,
// blockIdx.x + gridDim.x, ...), so the J_ab, K_ac and K_bc targets shared by
// a run are summed in LDS and written once per run or target change.
// LDS: [op tables][hrr: nhrr][scales: 4*W_SCL][coefficients: nw * W_CF][vrr: nw * nvrr]
//      [J/K run sums: na nb + na nc + nb nc][run index: 1], J/K build with jk_run > 0
// JR: instantiation with the J/K run code (the other keeps the ERI path's registers).
// bid / nblk: the block's index and the block count of its class launch
// (blockIdx.x / gridDim.x, or its segment's in a fused launch).
// TG: the op tables are read from the global blob (through L1/L2) instead of
// being staged into LDS; the block's LDS then starts with the HRR area, so
// more blocks stay resident per CU (large launches).
template <int ACC, bool JR, bool TG>
__device__ __forceinline__ void w_body(const EriArgs& args, const WClassDesc& D, const WTablesDev& W,
                                       int64_t nq_launch, int bid, int nblk, double* lds) {
  const int nv = D.vlev[D.nvlev], nh = D.hlev[D.nhlev];
  const uint32_t* tb = TG ? W.blob + 4 * (size_t)D.boff : reinterpret_cast<const uint32_t*>(lds);
  const uint4* sV = reinterpret_cast<const uint4*>(tb);
  const uint2_t* sH = reinterpret_cast<const uint2_t*>(tb) + 2 * nv;
  const uint32_t* sO = reinterpret_cast<const uint32_t*>(sH + nh);
  const uint32_t* sA = sO + D.nout;
  double* hrr = lds + (TG ? 0 : w_desc_doubles(D));
  double* scl = hrr + D.nhrr;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int nw = nthr >> 6, wv = tid >> 6, lane = tid & 63;
  double* cf = scl + 4 * W_SCL + wv * W_CF;
  double* vrr_all = scl + 4 * W_SCL + nw * W_CF;
  double* vrr = vrr_all + wv * D.nvrr;
  if (!TG) {
    stage_chunks(reinterpret_cast<uint4*>(lds), reinterpret_cast<const uint4*>(W.blob) + D.boff, D.bwords / 4);
    __syncthreads();
  }
  const int na = ncart_dev(D.la), nb = ncart_dev(D.lb), nc = ncart_dev(D.lc), nd = ncart_dev(D.ld);
  const int mtop = D.mmax - 1;
  const double* btab = args.boys + (size_t)mtop * BOYS_NROW * 8;
  int asrc[ACC];
#pragma unroll
  for (int k = 0; k < ACC; ++k) {
    const int x = k * 64 + lane;
    asrc[k] = x < D.nacc ? (int)(sA[x] & 0xFFFF) : 0;  // bit 31 marks a diagonal (e0|e0)
  }
  const bool jrun = JR && args.mode == MI_MODE_JK && args.jk_run > 0;
  double* racc = jrun ? vrr_all + nw * D.nvrr : nullptr;
  JkRun run{kNoRun, kNoRun, kNoRun, 0, 0, 0};
  if (jrun)
    for (int t = tid; t < na * nb + na * nc + nb * nc; t += nthr) racc[t] = 0.0;
  unsigned long long* s_run = jrun ? reinterpret_cast<unsigned long long*>(racc + na * nb + na * nc + nb * nc) : nullptr;
  const int64_t CH = jrun ? args.jk_run : 1;
  int ti = 0;
  int64_t c0 = jrun ? next_run(args, s_run) : (int64_t)bid;
  while (c0 < nq_launch) {
  const int64_t c1 = c0 + CH < nq_launch ? c0 + CH : nq_launch;
  for (int64_t q = c0; q < c1; ++q) {
    ti = advance_task(args, ti, q);
    const TaskDesc T = args.tasks[ti];
    QuartetCtx qc;
    decode_index(T, q - T.blockStart, qc);
    qc.G = 1;
    qc.g = 0;
    const int qmode = quartet_mode(args, qc);  // the wave kernel keeps FP64 VRR; 2 = screened
    if (qmode == 2) continue;                   // uniform per block (output pre-zeroed)
    const PairTopo P = args.topo[qc.bra];
    const PairTopo Q = args.topo[qc.ket];
    const double Ax = args.shell_xyz[3 * P.shA], Ay = args.shell_xyz[3 * P.shA + 1], Az = args.shell_xyz[3 * P.shA + 2];
    const double Cx = args.shell_xyz[3 * Q.shA], Cy = args.shell_xyz[3 * Q.shA + 1], Cz = args.shell_xyz[3 * Q.shA + 2];
    if (tid < na) scl[tid] = args.bf_scale[P.bfA + tid];
    if (tid < nb) scl[W_SCL + tid] = args.bf_scale[P.bfB + tid];
    if (tid < nc) scl[2 * W_SCL + tid] = args.bf_scale[Q.bfA + tid];
    if (tid < nd) scl[3 * W_SCL + tid] = args.bf_scale[Q.bfB + tid];
    double acc[ACC];
#pragma unroll
    for (int k = 0; k < ACC; ++k) acc[k] = 0.0;
    const int K = P.npp * Q.npp;
    const double rnv = rcp_int(Q.npp);
    for (int t = wv; t < K; t += nw) {
      int v;
      const int u = udiv_rcp(t, Q.npp, rnv, v);
      const PrimPair bp = args.pp[P.ppoff + u];
      const PrimPair kp = args.pp[Q.ppoff + v];
      const double zeta = bp.zeta, eta = kp.zeta;
      const double rs = rsqrt_nr(zeta + eta), inv_s = rs * rs;
      const double PQx = (Ax + bp.PAx) - (Cx + kp.PAx), PQy = (Ay + bp.PAy) - (Cy + kp.PAy),
                   PQz = (Az + bp.PAz) - (Cz + kp.PAz);
      const double T2 = zeta * eta * inv_s * (PQx * PQx + PQy * PQy + PQz * PQz);
      const double pref = bp.K * kp.K * rs;
      // Boys: lane m (< vt0) evaluates the top order and recurses down to m.
      if (lane < D.vt0) vrr[lane] = pref * boys_lane(btab, T2, mtop, lane);
      __syncwarp();
      // the primitive quartet's coefficient block (indices: gen_hgp.py CF_*)
      if (lane < 40) {
        const double eta_s = eta * inv_s, zeta_s = zeta * inv_s;
        double c = 0.0;
        if (lane < 3) c = sel3(lane, bp.PAx, bp.PAy, bp.PAz);
        else if (lane < 6) c = -eta_s * sel3(lane - 3, PQx, PQy, PQz);
        else if (lane < 9) c = sel3(lane - 6, kp.PAx, kp.PAy, kp.PAz);
        else if (lane < 12) c = zeta_s * sel3(lane - 9, PQx, PQy, PQz);
        else if (lane == 12) c = eta_s;
        else if (lane == 13) c = zeta_s;
        else if (lane >= 16) {
          const int grp = (lane - 16) >> 3;
          const double n = (double)((lane - 16) & 7);
          c = n * (grp == 0 ? bp.oo2z : (grp == 1 ? kp.oo2z : 0.5 * inv_s));
        }
        cf[lane] = c;
      }
      __syncwarp();
      const char* vb = reinterpret_cast<const char*>(vrr);
      for (int lv = 0; lv < D.nvlev; ++lv) {
        const int i1 = D.vlev[lv + 1];
        for (int it = D.vlev[lv] + lane; it < i1; it += 64) {
          const uint4 o = sV[it];
          const double v1 = *reinterpret_cast<const double*>(vb + (o.x & 0xFFFFF));
          const double v1p = *reinterpret_cast<const double*>(vb + (o.x & 0xFFFFF) + 8);
          const double v2 = *reinterpret_cast<const double*>(vb + (o.y & 0xFFFFF));
          const double v2p = *reinterpret_cast<const double*>(vb + (o.y & 0xFFFFF) + 8);
          const double v3p = *reinterpret_cast<const double*>(vb + (o.z & 0xFFFFF) + 8);
          double val = cf[(o.x >> 20) & 63] * v1 + cf[o.x >> 26] * v1p;
          val += cf[(o.y >> 20) & 63] * v3p;
          val += cf[o.y >> 26] * (v2 - cf[o.z >> 20] * v2p);
          vrr[D.vt0 + it] = val;
        }
        __syncwarp();
      }
#pragma unroll
      for (int k = 0; k < ACC; ++k)
        if (k * 64 + lane < D.nacc) acc[k] += vrr[asrc[k]];
      __syncwarp();
    }
    // per-wave shares -> own VRR area, then summed in wave order into hrr[0..nacc)
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ACC; ++k)
      if (k * 64 + lane < D.nacc) vrr[k * 64 + lane] = acc[k];
    __syncthreads();
    for (int x = tid; x < D.nacc; x += nthr) {
      double sum = vrr_all[x];
      for (int w = 1; w < nw; ++w) sum += vrr_all[w * D.nvrr + x];
      hrr[x] = sum;
      if (args.mode == MI_MODE_SCHWARZ && (sA[x] >> 31)) record_vmax(&args.geom[qc.bra], sum);
    }
    __syncthreads();
    // HRR levels: item k writes slot nacc + k.
    const PairGeom PG = args.geom[qc.bra];
    const PairGeom QG = args.geom[qc.ket];
    for (int lv = 0; lv < D.nhlev; ++lv) {
      const int i1 = D.hlev[lv + 1];
      for (int it = D.hlev[lv] + tid; it < i1; it += nthr) {
        const uint2_t o = sH[it];
        const int s1 = o.x & 0xFFFF, s2 = o.x >> 16;
        const int j = o.y & 3;
        const bool ket = (o.y >> 2) & 1;
        const double X = ket ? sel3(j, QG.ABx, QG.ABy, QG.ABz) : sel3(j, PG.ABx, PG.ABy, PG.ABz);
        hrr[D.nacc + it] = hrr[s1] + X * hrr[s2];
      }
      __syncthreads();
    }
    const ScatterCtx sc = make_scatter(args, qc, P, Q);
    if (args.mode == MI_MODE_JK) {  // uniform per block
      // scaled block into the (now free) VRR areas, then six contractions
      for (int x = tid; x < D.nout; x += nthr) {
        const uint32_t o = sO[x];
        const int ma = (o >> 16) & 15, mb = (o >> 20) & 15, mc = (o >> 24) & 15, md = (o >> 28) & 15;
        vrr_all[x] = hrr[o & 0xFFFF] * (scl[ma] * scl[W_SCL + mb]) * (scl[2 * W_SCL + mc] * scl[3 * W_SCL + md]);
      }
      __syncthreads();
      jk_block_lds(args, sc, vrr_all, na, nb, nc, nd, tid, nthr, racc, run);
      __syncthreads();
      continue;
    }
    for (int x = tid; x < D.nout; x += nthr) {
      const uint32_t o = sO[x];
      const int off = o & 0xFFFF;
      const int ma = (o >> 16) & 15, mb = (o >> 20) & 15, mc = (o >> 24) & 15, md = (o >> 28) & 15;
      scatter_elem(args, sc, qc.bra, sc.fa0 + ma, sc.fb0 + mb, sc.fc0 + mc, sc.fd0 + md,
                   hrr[off] * (scl[ma] * scl[W_SCL + mb]) * (scl[2 * W_SCL + mc] * scl[3 * W_SCL + md]));
    }
    __syncthreads();
  }
  if (jrun) jk_run_flush(args, run, racc, na, nb, nc, true, true, true, tid, nthr);
  c0 = jrun ? next_run(args, s_run) : c0 + nblk;
  }
}

template <int ACC, bool JR, bool TG>
__global__ void __launch_bounds__(256) hgp_w_kernel(const EriArgs args, const WClassDesc D, const WTablesDev W,
                                                    int64_t nq_launch) {
  extern __shared__ double lds[];
  w_body<ACC, JR, TG>(args, D, W, nq_launch, (int)blockIdx.x, (int)gridDim.x, lds);
}

// Fused launch over several wave-kernel classes of one wavefront count
// (unique-array build only), in the instantiation of their largest
// accumulator count; class descriptors come from kWClassDescDev, the dynamic
// LDS is sized for the segment that needs the most.
template <int ACC>
__global__ void __launch_bounds__(256) hgp_wf_kernel(const EriArgs args, const FusedSegs F, const WTablesDev W) {
  extern __shared__ double lds[];
  const int s = fused_segment(F, (int)blockIdx.x);
  EriArgs a = args;
  a.tasks = args.tasks + F.task_off[s];
  a.ntasks = F.ntasks[s];
  w_body<ACC, false, false>(a, kWClassDescDev[F.cls[s]], W, F.nq[s], (int)blockIdx.x - F.bstart[s],
                     F.bstart[s + 1] - F.bstart[s], lds);
}

// ---------------------------------------------------------------------------
// Global-table block kernel for the classes whose contracted block or HRR
// work area does not fit the wave kernel (f classes with nacc > 1024 or more
// than 65535 HRR slots).  All threads of a block cooperate on one primitive
// quartet at a time: the VRR work area of the class is in LDS (nvrr <= 17273
// doubles for (ff|ff)), the op tables are read from global memory with 32-bit
// slot fields, and the contracted (e0|f0) block plus the HRR levels live in a
// per-block slice of a global scratch buffer (nhrr doubles per block; each
// accumulator slot is owned by one thread across the primitive loop).  Blocks
// are persistent over the launch's quartets like the wave kernel.
// LDS: [scales: 4*W_SCL][vrr: nvrr]
// ---------------------------------------------------------------------------
// HL: the contracted block and HRR levels also fit in LDS (after the VRR
// area) -- then no global scratch is used.  TL: the class's op tables are
// staged into LDS too (once per persistent block), read there instead of
// through L1/L2.
// J/K build with args.jk_run > 0: runs of quartets as in the wave kernel.
// LDS: [scales: 4*W_SCL][vrr: nvrr][hrr: nhrr if HL][tables if TL]
//      [J/K run sums: na nb + na nc + nb nc][run index: 1], J/K build with jk_run > 0
// The global-table kernel keeps its HRR area in LDS too when the whole
// block's LDS stays within this (several blocks per CU remain resident).
#define WG_HRR_LDS_MAX (64 * 1024)

// bid / nblk: the block's index and the block count of its class launch
// (blockIdx.x / gridDim.x, or its segment's in a fused launch).
template <bool HL, bool TL, bool JR>
__device__ __forceinline__ void wg_body(const EriArgs& args, const WClassDesc& D, const uint32_t* __restrict__ gblob,
                                        double* __restrict__ scratch, int64_t nq_launch, int bid, int nblk,
                                        double* lds) {
  double* scl = lds;
  double* vrr = lds + 4 * W_SCL;
  double* hrr = HL ? vrr + D.nvrr : scratch + (size_t)bid * D.nhrr;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const uint32_t* tab = gblob;  // the class's tables start at word D.gv
  double* lds_end = vrr + D.nvrr + (HL ? D.nhrr : 0);
  if (TL) {
    // 16-byte aligned copy of words [gv, ga + 2 nacc) after the work areas
    double* t = lds_end;
    t += (reinterpret_cast<uintptr_t>(t) & 8) ? 1 : 0;
    const int nwords = (D.ga + 2 * D.nacc - D.gv + 3) & ~3;
    stage_chunks(reinterpret_cast<uint4*>(t), reinterpret_cast<const uint4*>(gblob + D.gv), nwords / 4);
    __syncthreads();
    tab = reinterpret_cast<const uint32_t*>(t) - D.gv;
    lds_end = t + nwords / 2;
  }
  const uint4* gV = reinterpret_cast<const uint4*>(tab + D.gv);
  const uint4* gH = reinterpret_cast<const uint4*>(tab + D.gh);
  const uint2_t* gO = reinterpret_cast<const uint2_t*>(tab + D.go);
  const uint2_t* gA = reinterpret_cast<const uint2_t*>(tab + D.ga);
  const int na = ncart_dev(D.la), nb = ncart_dev(D.lb), nc = ncart_dev(D.lc), nd = ncart_dev(D.ld);
  const int mtop = D.mmax - 1;
  const double* btab = args.boys + (size_t)mtop * BOYS_NROW * 8;
  const bool jrun = JR && args.mode == MI_MODE_JK && args.jk_run > 0;
  double* racc = jrun ? lds_end : nullptr;
  JkRun run{kNoRun, kNoRun, kNoRun, 0, 0, 0};
  if (jrun)
    for (int t = tid; t < na * nb + na * nc + nb * nc; t += nthr) racc[t] = 0.0;
  unsigned long long* s_run = jrun ? reinterpret_cast<unsigned long long*>(racc + na * nb + na * nc + nb * nc) : nullptr;
  const int64_t CH = jrun ? args.jk_run : 1;
  int ti = 0;
  int64_t c0 = jrun ? next_run(args, s_run) : (int64_t)bid;
  while (c0 < nq_launch) {
  const int64_t c1 = c0 + CH < nq_launch ? c0 + CH : nq_launch;
  for (int64_t q = c0; q < c1; ++q) {
    ti = advance_task(args, ti, q);
    const TaskDesc T = args.tasks[ti];
    QuartetCtx qc;
    decode_index(T, q - T.blockStart, qc);
    qc.G = 1;
    qc.g = 0;
    const int qmode = quartet_mode(args, qc);  // FP64 VRR throughout; 2 = screened (uniform per block)
    if (qmode == 2) continue;                   // output pre-zeroed
    const PairTopo P = args.topo[qc.bra];
    const PairTopo Q = args.topo[qc.ket];
    const double Ax = args.shell_xyz[3 * P.shA], Ay = args.shell_xyz[3 * P.shA + 1], Az = args.shell_xyz[3 * P.shA + 2];
    const double Cx = args.shell_xyz[3 * Q.shA], Cy = args.shell_xyz[3 * Q.shA + 1], Cz = args.shell_xyz[3 * Q.shA + 2];
    if (tid < na) scl[tid] = args.bf_scale[P.bfA + tid];
    if (tid < nb) scl[W_SCL + tid] = args.bf_scale[P.bfB + tid];
    if (tid < nc) scl[2 * W_SCL + tid] = args.bf_scale[Q.bfA + tid];
    if (tid < nd) scl[3 * W_SCL + tid] = args.bf_scale[Q.bfB + tid];
    const int K = P.npp * Q.npp;
    for (int t = 0, u = 0, v = 0; t < K; ++t, v = v + 1 == Q.npp ? 0 : v + 1, u += v == 0 ? 1 : 0) {
      const PrimPair bp = args.pp[P.ppoff + u];
      const PrimPair kp = args.pp[Q.ppoff + v];
      const double zeta = bp.zeta, eta = kp.zeta;
      const double rs = rsqrt_nr(zeta + eta), inv_s = rs * rs;
      const double PQx = (Ax + bp.PAx) - (Cx + kp.PAx), PQy = (Ay + bp.PAy) - (Cy + kp.PAy),
                   PQz = (Az + bp.PAz) - (Cz + kp.PAz);
      const double T2 = zeta * eta * inv_s * (PQx * PQx + PQy * PQy + PQz * PQz);
      const double pref = bp.K * kp.K * rs;
      if (tid < D.vt0) vrr[tid] = pref * boys_lane(btab, T2, mtop, tid);
      __syncthreads();
      const double oo2z = bp.oo2z, oo2e = kp.oo2z, oo2s = 0.5 * inv_s;
      const double eta_s = eta * inv_s, zeta_s = zeta * inv_s;
      const double WPx = -eta_s * PQx, WPy = -eta_s * PQy, WPz = -eta_s * PQz;
      const double WQx = zeta_s * PQx, WQy = zeta_s * PQy, WQz = zeta_s * PQz;
      for (int lv = 0; lv < D.nvlev; ++lv) {
        const int i1 = D.vlev[lv + 1];
        for (int it = D.vlev[lv] + tid; it < i1; it += nthr) {
          const uint4 o = gV[it];
          const int s1 = o.x, s2 = o.y, s3 = o.z;
          const int dir = o.w & 3;
          const double n1 = (double)((o.w >> 2) & 7), n2 = (double)((o.w >> 5) & 7);
          const bool cb = (o.w >> 8) & 1;
          const double X1 = cb ? sel3(dir, kp.PAx, kp.PAy, kp.PAz) : sel3(dir, bp.PAx, bp.PAy, bp.PAz);
          const double X2 = cb ? sel3(dir, WQx, WQy, WQz) : sel3(dir, WPx, WPy, WPz);
          const double c3 = cb ? n1 * oo2s : 0.0;
          const double c2 = cb ? n2 * oo2e : n1 * oo2z;
          const double r2 = cb ? zeta_s : eta_s;
          double val = X1 * vrr[s1] + X2 * vrr[s1 + 1];
          val += c3 * vrr[s3 + 1];
          val += c2 * (vrr[s2] - r2 * vrr[s2 + 1]);
          vrr[D.vt0 + it] = val;
        }
        __syncthreads();
      }
      // contraction: slot x is always added by the same thread
      for (int x = tid; x < D.nacc; x += nthr) {
        const double v0 = vrr[gA[x].x];
        hrr[x] = t == 0 ? v0 : hrr[x] + v0;
      }
      __syncthreads();  // the next primitive overwrites vrr
    }
    if (args.mode == MI_MODE_SCHWARZ)
      for (int x = tid; x < D.nacc; x += nthr)
        if (gA[x].y) record_vmax(&args.geom[qc.bra], hrr[x]);
    // HRR levels: item k writes slot nacc + k.  __syncthreads orders the
    // block's global-memory writes and reads (workgroup scope, one CU).
    const PairGeom PG = args.geom[qc.bra];
    const PairGeom QG = args.geom[qc.ket];
    for (int lv = 0; lv < D.nhlev; ++lv) {
      const int i1 = D.hlev[lv + 1];
      for (int it = D.hlev[lv] + tid; it < i1; it += nthr) {
        const uint4 o = gH[it];
        const int j = o.z & 3;
        const bool ket = (o.z >> 2) & 1;
        const double X = ket ? sel3(j, QG.ABx, QG.ABy, QG.ABz) : sel3(j, PG.ABx, PG.ABy, PG.ABz);
        hrr[D.nacc + it] = hrr[o.x] + X * hrr[o.y];
      }
      __syncthreads();
    }
    const ScatterCtx sc = make_scatter(args, qc, P, Q);
    if (args.mode == MI_MODE_JK) {  // uniform per block
      for (int x = tid; x < D.nout; x += nthr) {
        const uint2_t o = gO[x];
        const int ma = o.y & 15, mb = (o.y >> 4) & 15, mc = (o.y >> 8) & 15, md = (o.y >> 12) & 15;
        vrr[x] = hrr[o.x] * (scl[ma] * scl[W_SCL + mb]) * (scl[2 * W_SCL + mc] * scl[3 * W_SCL + md]);
      }
      __syncthreads();
      jk_block_lds(args, sc, vrr, na, nb, nc, nd, tid, nthr, racc, run);
      __syncthreads();
      continue;
    }
    for (int x = tid; x < D.nout; x += nthr) {
      const uint2_t o = gO[x];
      const int ma = o.y & 15, mb = (o.y >> 4) & 15, mc = (o.y >> 8) & 15, md = (o.y >> 12) & 15;
      scatter_elem(args, sc, qc.bra, sc.fa0 + ma, sc.fb0 + mb, sc.fc0 + mc, sc.fd0 + md,
                   hrr[o.x] * (scl[ma] * scl[W_SCL + mb]) * (scl[2 * W_SCL + mc] * scl[3 * W_SCL + md]));
    }
    __syncthreads();
  }
  if (jrun) jk_run_flush(args, run, racc, na, nb, nc, true, true, true, tid, nthr);
  c0 = jrun ? next_run(args, s_run) : c0 + nblk;
  }
}

template <bool HL, bool TL, bool JR>
__global__ void __launch_bounds__(512) hgp_wg_kernel(const EriArgs args, const WClassDesc D,
                                                     const uint32_t* __restrict__ gblob,
                                                     double* __restrict__ scratch, int64_t nq_launch) {
  extern __shared__ double lds[];
  wg_body<HL, TL, JR>(args, D, gblob, scratch, nq_launch, (int)blockIdx.x, (int)gridDim.x, lds);
}

// Fused launch over several global-table classes of one wavefront count
// (unique-array build only; op tables read through L1/L2); class
// descriptors come from kWClassDescDev, the dynamic LDS is sized for the
// segment that needs the most.  A class keeps its contracted block and HRR
// levels in LDS when they fit WG_HRR_LDS_MAX with the VRR area (as in its
// own launch), else in its launch's slice of the plan scratch, whose offset
// the segment carries in bt_off.
__global__ void __launch_bounds__(512) hgp_wgf_kernel(const EriArgs args, const FusedSegs F,
                                                      const uint32_t* __restrict__ gblob,
                                                      double* __restrict__ scratch) {
  extern __shared__ double lds[];
  const int s = fused_segment(F, (int)blockIdx.x);
  EriArgs a = args;
  a.tasks = args.tasks + F.task_off[s];
  a.ntasks = F.ntasks[s];
  const WClassDesc& D = kWClassDescDev[F.cls[s]];
  const int bid = (int)blockIdx.x - F.bstart[s], nblk = F.bstart[s + 1] - F.bstart[s];
  if (sizeof(double) * (size_t)(4 * W_SCL + D.nvrr + D.nhrr) <= WG_HRR_LDS_MAX)
    wg_body<true, false, false>(a, D, gblob, nullptr, F.nq[s], bid, nblk, lds);
  else
    wg_body<false, false, false>(a, D, gblob, scratch + F.bt_off[s], F.nq[s], bid, nblk, lds);
}

// ---------------------------------------------------------------------------
// Extended-precision table kernel (tuning variant 4): the global-table
// kernel's op tables evaluated in double-double arithmetic for the quartets
// of shell pairs whose HRR would amplify rounding beyond amp_limit (mi_host.cpp)
// in either orientation (contracted pairs mixing tight and diffuse primitives,
// several bohr apart: mi_host.cpp, pair orientation).  Inputs of the
// recurrences (PA, WP, QC, WQ, the n/2zeta factors, AB, CD, the [0|0]^(m)
// values from FP64 Boys and prefactors) stay FP64 -- they carry the problem's
// data, and their rounding changes results by ~1e-16 relative -- while every
// VRR entry, the contraction sums and the HRR levels carry ~106 bits, so the
// cancellation happens on numbers that are exact to far below the bar.
// Double-double arithmetic: Knuth's TwoSum, Dekker's FastTwoSum and the
// fma-based TwoProd (exact product), the usual "sloppy" pair addition plus
// one renormalisation.  One wavefront per block, one quartet per block at a
// time (persistent blocks), VRR and HRR areas in a global slice of the plan
// scratch per block (16 bytes per entry: a (ff|ff) VRR area is 276 KB,
// beyond LDS).
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)
struct ddv {
  double hi, lo;
};
__device__ __forceinline__ ddv dd_two_sum(double a, double b) {
  const double s = a + b, bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ ddv dd_fast_two_sum(double a, double b) {
  const double s = a + b;
  return {s, b - (s - a)};
}
__device__ __forceinline__ ddv dd_add(ddv x, ddv y) {
  ddv s = dd_two_sum(x.hi, y.hi);
  s.lo += x.lo + y.lo;
  return dd_fast_two_sum(s.hi, s.lo);
}
// x * b, b a double
__device__ __forceinline__ ddv dd_mul_d(ddv x, double b) {
  const double p = x.hi * b;
  double e = fma(x.hi, b, -p);
  e = fma(x.lo, b, e);
  return dd_fast_two_sum(p, e);
}
#pragma clang fp contract(on)

__device__ __forceinline__ ddv dd_load(const double2* a, int i) {
  const double2 v = a[i];
  return {v.x, v.y};
}
__device__ __forceinline__ void dd_store(double2* a, int i, ddv v) { a[i] = make_double2(v.hi, v.lo); }

__global__ void __launch_bounds__(64) hgp_wx_kernel(const EriArgs args, const WClassDesc D,
                                                     const uint32_t* __restrict__ gblob, double2* __restrict__ scratch,
                                                     int64_t nq_launch) {
  __shared__ double scl[4 * W_SCL];
  double2* vrr = scratch + (size_t)blockIdx.x * (D.nvrr + D.nhrr);
  double2* hrr = vrr + D.nvrr;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const uint4* gV = reinterpret_cast<const uint4*>(gblob + D.gv);
  const uint4* gH = reinterpret_cast<const uint4*>(gblob + D.gh);
  const uint2_t* gO = reinterpret_cast<const uint2_t*>(gblob + D.go);
  const uint2_t* gA = reinterpret_cast<const uint2_t*>(gblob + D.ga);
  const int na = ncart_dev(D.la), nb = ncart_dev(D.lb), nc = ncart_dev(D.lc), nd = ncart_dev(D.ld);
  const int mtop = D.mmax - 1;
  const double* btab = args.boys + (size_t)mtop * BOYS_NROW * 8;
  int ti = 0;
  for (int64_t q = blockIdx.x; q < nq_launch; q += gridDim.x) {
    ti = advance_task(args, ti, q);
    const TaskDesc T = args.tasks[ti];
    QuartetCtx qc;
    decode_index(T, q - T.blockStart, qc);
    qc.G = 1;
    qc.g = 0;
    if (quartet_mode(args, qc) == 2) continue;  // screened (uniform per block); output pre-zeroed
    const PairTopo P = args.topo[qc.bra];
    const PairTopo Q = args.topo[qc.ket];
    const double Ax = args.shell_xyz[3 * P.shA], Ay = args.shell_xyz[3 * P.shA + 1], Az = args.shell_xyz[3 * P.shA + 2];
    const double Cx = args.shell_xyz[3 * Q.shA], Cy = args.shell_xyz[3 * Q.shA + 1], Cz = args.shell_xyz[3 * Q.shA + 2];
    if (tid < na) scl[tid] = args.bf_scale[P.bfA + tid];
    if (tid < nb) scl[W_SCL + tid] = args.bf_scale[P.bfB + tid];
    if (tid < nc) scl[2 * W_SCL + tid] = args.bf_scale[Q.bfA + tid];
    if (tid < nd) scl[3 * W_SCL + tid] = args.bf_scale[Q.bfB + tid];
    const int K = P.npp * Q.npp;
    for (int t = 0, u = 0, v = 0; t < K; ++t, v = v + 1 == Q.npp ? 0 : v + 1, u += v == 0 ? 1 : 0) {
      const PrimPair bp = args.pp[P.ppoff + u];
      const PrimPair kp = args.pp[Q.ppoff + v];
      const double zeta = bp.zeta, eta = kp.zeta;
      const double rs = rsqrt_nr(zeta + eta), inv_s = rs * rs;
      const double PQx = (Ax + bp.PAx) - (Cx + kp.PAx), PQy = (Ay + bp.PAy) - (Cy + kp.PAy),
                   PQz = (Az + bp.PAz) - (Cz + kp.PAz);
      const double T2 = zeta * eta * inv_s * (PQx * PQx + PQy * PQy + PQz * PQz);
      const double pref = bp.K * kp.K * rs;
      if (tid < D.vt0) {
        const ddv b = dd_two_sum(0.0, boys_lane(btab, T2, mtop, tid));
        dd_store(vrr, tid, dd_mul_d(b, pref));
      }
      __syncthreads();
      const double oo2z = bp.oo2z, oo2e = kp.oo2z, oo2s = 0.5 * inv_s;
      const double eta_s = eta * inv_s, zeta_s = zeta * inv_s;
      const double WPx = -eta_s * PQx, WPy = -eta_s * PQy, WPz = -eta_s * PQz;
      const double WQx = zeta_s * PQx, WQy = zeta_s * PQy, WQz = zeta_s * PQz;
      for (int lv = 0; lv < D.nvlev; ++lv) {
        const int i1 = D.vlev[lv + 1];
        for (int it = D.vlev[lv] + tid; it < i1; it += nthr) {
          const uint4 o = gV[it];
          const int s1 = o.x, s2 = o.y, s3 = o.z;
          const int dir = o.w & 3;
          const double n1 = (double)((o.w >> 2) & 7), n2 = (double)((o.w >> 5) & 7);
          const bool cb = (o.w >> 8) & 1;
          const bool h2 = (o.w >> 9) & 1, h3 = (o.w >> 10) & 1;
          const double X1 = cb ? sel3(dir, kp.PAx, kp.PAy, kp.PAz) : sel3(dir, bp.PAx, bp.PAy, bp.PAz);
          const double X2 = cb ? sel3(dir, WQx, WQy, WQz) : sel3(dir, WPx, WPy, WPz);
          ddv val = dd_add(dd_mul_d(dd_load(vrr, s1), X1), dd_mul_d(dd_load(vrr, s1 + 1), X2));
          if (cb && h3) val = dd_add(val, dd_mul_d(dd_load(vrr, s3 + 1), n1 * oo2s));
          if (h2) {
            const double c2 = cb ? n2 * oo2e : n1 * oo2z;
            const double r2 = cb ? zeta_s : eta_s;
            const ddv d = dd_add(dd_load(vrr, s2), dd_mul_d(dd_load(vrr, s2 + 1), -r2));
            val = dd_add(val, dd_mul_d(d, c2));
          }
          dd_store(vrr, D.vt0 + it, val);
        }
        __syncthreads();
      }
      // contraction: slot x is always added by the same thread
      for (int x = tid; x < D.nacc; x += nthr) {
        const ddv v0 = dd_load(vrr, gA[x].x);
        dd_store(hrr, x, t == 0 ? v0 : dd_add(dd_load(hrr, x), v0));
      }
      __syncthreads();  // the next primitive overwrites vrr
    }
    if (args.mode == MI_MODE_SCHWARZ)
      for (int x = tid; x < D.nacc; x += nthr)
        if (gA[x].y) record_vmax(&args.geom[qc.bra], hrr[x].x + hrr[x].y);
    const PairGeom PG = args.geom[qc.bra];
    const PairGeom QG = args.geom[qc.ket];
    for (int lv = 0; lv < D.nhlev; ++lv) {
      const int i1 = D.hlev[lv + 1];
      for (int it = D.hlev[lv] + tid; it < i1; it += nthr) {
        const uint4 o = gH[it];
        const int j = o.z & 3;
        const bool ket = (o.z >> 2) & 1;
        const double X = ket ? sel3(j, QG.ABx, QG.ABy, QG.ABz) : sel3(j, PG.ABx, PG.ABy, PG.ABz);
        dd_store(hrr, D.nacc + it, dd_add(dd_load(hrr, o.x), dd_mul_d(dd_load(hrr, o.y), X)));
      }
      __syncthreads();
    }
    const ScatterCtx sc = make_scatter(args, qc, P, Q);
    if (args.mode == MI_MODE_JK) {  // uniform per block: the scaled block staged in the VRR area
      double* blk = reinterpret_cast<double*>(vrr);
      for (int x = tid; x < D.nout; x += nthr) {
        const uint2_t o = gO[x];
        const int ma = o.y & 15, mb = (o.y >> 4) & 15, mc = (o.y >> 8) & 15, md = (o.y >> 12) & 15;
        blk[x] = (hrr[o.x].x + hrr[o.x].y) * (scl[ma] * scl[W_SCL + mb]) * (scl[2 * W_SCL + mc] * scl[3 * W_SCL + md]);
      }
      __syncthreads();
      JkRun run{kNoRun, kNoRun, kNoRun, 0, 0, 0};
      jk_block_lds(args, sc, blk, na, nb, nc, nd, tid, nthr, nullptr, run);
      __syncthreads();
      continue;
    }
    for (int x = tid; x < D.nout; x += nthr) {
      const uint2_t o = gO[x];
      const int ma = o.y & 15, mb = (o.y >> 4) & 15, mc = (o.y >> 8) & 15, md = (o.y >> 12) & 15;
      scatter_elem(args, sc, qc.bra, sc.fa0 + ma, sc.fb0 + mb, sc.fc0 + mc, sc.fd0 + md,
                   (hrr[o.x].x + hrr[o.x].y) * (scl[ma] * scl[W_SCL + mb]) * (scl[2 * W_SCL + mc] * scl[3 * W_SCL + md]));
    }
    __syncthreads();
  }
}

// Zero a device array (a kernel, not a memset node, so captured graphs replay it safely).
// Plain 8-byte stores, 4096 x 256 lanes: 5.8 TB/s on 40 GiB, as fast as
// hipMemsetAsync (5.9); 16-byte, non-temporal and unrolled variants measured
// 5.0-5.8 TB/s.
__global__ void __launch_bounds__(256) zero_kernel(double* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

// J = Jacc + Jacc^T, K = Kacc + Kacc^T, in place (one lane per i <= j).
__global__ void __launch_bounds__(256) symmetrize_kernel(double* __restrict__ J, double* __restrict__ K, int n) {
  const int64_t npair = (int64_t)n * (n + 1) / 2;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < npair; t += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (i * (i + 1) / 2 > t) --i;
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    const int64_t j = t - i * (i + 1) / 2;  // j <= i
    const double jv = J[i * n + j] + J[j * n + i];
    const double kv = K[i * n + j] + K[j * n + i];
    J[i * n + j] = jv;
    J[j * n + i] = jv;
    K[i * n + j] = kv;
    K[j * n + i] = kv;
  }
}

// ---------------------------------------------------------------------------
// Host-side launch helpers (called from mi_host.cpp).
// ---------------------------------------------------------------------------
namespace mi {

int nclass() { return HGP_NCLASS; }
int lmax_shell() { return HGP_LMAX_SHELL; }
const WClassDesc& wclass(int c) { return kWClassDesc[c]; }
bool has_register_kernel(int c) { return kRegisterKernels[c] != nullptr; }

struct WTablesHost {
  uint32_t* blob = nullptr;   // kWBlob (LDS wave kernel)
  uint32_t* gblob = nullptr;  // kWGBlob (global-table kernel)
};

hipError_t upload_wtables(void** handle) {
  WTablesHost* t = new WTablesHost();
  hipError_t e;
  if ((e = hipMalloc(&t->blob, sizeof(kWBlob))) != hipSuccess) return e;
  if ((e = hipMemcpy(t->blob, kWBlob, sizeof(kWBlob), hipMemcpyHostToDevice)) != hipSuccess) return e;
  if ((e = hipMalloc(&t->gblob, sizeof(kWGBlob))) != hipSuccess) return e;
  if ((e = hipMemcpy(t->gblob, kWGBlob, sizeof(kWGBlob), hipMemcpyHostToDevice)) != hipSuccess) return e;
  *handle = t;
  return hipSuccess;
}

void free_wtables(void* handle) {
  WTablesHost* t = static_cast<WTablesHost*>(handle);
  if (!t) return;
  hipFree(t->blob);
  hipFree(t->gblob);
  delete t;
}

hipError_t launch_pair_prims(int npp, const int32_t* pp_pair, const int32_t* pp_ab, const PairTopo* topo,
                             const double* shell_xyz,
                             const int32_t* shell_prim_off, const double* prim_exp, const double* prim_cn,
                             PrimPair* pp, PairGeom* geom, const double2* bounds, hipStream_t stream) {
  if (npp == 0) return hipSuccess;
  dim3 grid((npp + 255) / 256), block(256);
  hipLaunchKernelGGL(pair_prims_kernel, grid, block, 0, stream, npp, pp_pair, pp_ab, topo, shell_xyz, shell_prim_off,
                     prim_exp, prim_cn, pp, geom, bounds);
  return hipGetLastError();
}

int register_waves_per_block(int c) { return kRegisterWavesPerBlock[c]; }
int fused_register_waves_per_block(int g) { return kFusedWavesPerBlock[g]; }

hipError_t launch_class_register(int c, const EriArgs& args, int nblocks, int wpb, hipStream_t stream) {
  const void* fn = kRegisterKernels[c];
  void* kargs[] = {const_cast<EriArgs*>(&args)};
  return hipLaunchKernel(fn, dim3(nblocks), dim3(64 * wpb), kargs, 0, stream);
}

size_t wclass_lds_bytes(int c, int nw, bool tg) {
  const WClassDesc& D = kWClassDesc[c];
  return sizeof(double) *
         (size_t)((tg ? 0 : w_desc_doubles(D)) + D.nhrr + 4 * W_SCL + (size_t)nw * (W_CF + D.nvrr));
}

// Instantiation holding the class's contracted block: ACC values per lane;
// JR: with the J/K run code; TG: op tables read from global memory.
template <bool JR, bool TG>
static const void* w_kernel_fn_t(int c) {
  const int nacc = kWClassDesc[c].nacc;
  if (nacc <= 64 * 2) return (const void*)hgp_w_kernel<2, JR, TG>;
  if (nacc <= 64 * 4) return (const void*)hgp_w_kernel<4, JR, TG>;
  if (nacc <= 64 * 8) return (const void*)hgp_w_kernel<8, JR, TG>;
  return (const void*)hgp_w_kernel<16, JR, TG>;
}

static const void* w_kernel_fn(int c, bool jr, bool tg) {
  return jr ? (tg ? w_kernel_fn_t<true, true>(c) : w_kernel_fn_t<true, false>(c))
            : (tg ? w_kernel_fn_t<false, true>(c) : w_kernel_fn_t<false, false>(c));
}

int wave_kernel_blocks_per_cu(int c, int nw, bool tg) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, w_kernel_fn(c, false, tg), 64 * nw,
                                                   wclass_lds_bytes(c, nw, tg)) != hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

bool wclass_fits(int c) { return kWClassDesc[c].boff >= 0 && kWClassDesc[c].nacc <= 64 * W_ACC_MAX; }

bool wclass_fits_nw(int c, int nw, size_t lds_limit, bool tg) {
  return wclass_fits(c) && nw >= 1 && nw <= 4 && wclass_lds_bytes(c, nw, tg) <= lds_limit;
}

// J/K build: LDS for the run sums of the wave kernels (one double per J_ab,
// K_ac, K_bc output, then the block's current run index), used when the
// block's LDS stays within the limit.
static size_t jk_run_bytes(int c) {
  const WClassDesc& D = kWClassDesc[c];
  const int na = (D.la + 1) * (D.la + 2) / 2, nb = (D.lb + 1) * (D.lb + 2) / 2, nc = (D.lc + 1) * (D.lc + 2) / 2;
  return sizeof(double) * (size_t)(na * nb + na * nc + nb * nc + 1);
}

static size_t g_lds_limit = 65536;  // dynamic LDS a wave-kernel block may use (set_wave_kernel_lds_limit)

// Run length: up to MI_JK_RUN quartets, but short enough that the launch
// holds about four runs per block (small launches stay spread out).
// The run sums take LDS: the persistent grid is cut to what stays resident
// with it (the blocks loop over the launch's quartets, any grid is correct).
static EriArgs with_jk_run(const EriArgs& args, int c, const void* fn, int nw, size_t& lds, int64_t nq_launch,
                           int& nblocks) {
  // fn: the kernel instantiation with the J/K run code
  EriArgs a = args;
  a.jk_run = 0;
  const int64_t run = std::min<int64_t>(MI_JK_RUN, nq_launch / (4 * (int64_t)std::max(nblocks, 1)));
  if (a.mode == MI_MODE_JK && a.run_ctr && run >= 2 && lds + jk_run_bytes(c) <= g_lds_limit) {
    a.jk_run = (int32_t)run;
    lds += jk_run_bytes(c);
    int dev = 0, ncu = 0, bpc = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, fn, 64 * nw, lds) == hipSuccess && ncu > 0 && bpc > 0)
      nblocks = std::min(nblocks, ncu * bpc);
  }
  return a;
}

hipError_t launch_class_wave(int c, int nw, bool tg, const EriArgs& args, int nblocks, int64_t nq_launch,
                             void* wtables, hipStream_t stream) {
  const WTablesHost* t = static_cast<const WTablesHost*>(wtables);
  WTablesDev W{t->blob};
  const WClassDesc D = kWClassDesc[c];
  size_t lds = wclass_lds_bytes(c, nw, tg);
  EriArgs a = with_jk_run(args, c, w_kernel_fn(c, true, tg), nw, lds, nq_launch, nblocks);
  void* kargs[] = {&a, const_cast<WClassDesc*>(&D), &W, &nq_launch};
  return hipLaunchKernel(w_kernel_fn(c, a.jk_run > 0, tg), dim3(nblocks), dim3(64 * nw), kargs, lds, stream);
}

hipError_t launch_fused_register_g0(const EriArgs&, const FusedSegs&, int, int, hipStream_t);
hipError_t launch_fused_register_g1(const EriArgs&, const FusedSegs&, int, int, hipStream_t);
hipError_t launch_fused_register_g2(const EriArgs&, const FusedSegs&, int, int, hipStream_t);
hipError_t launch_fused_register_g3(const EriArgs&, const FusedSegs&, int, int, hipStream_t);
hipError_t launch_fused_register_g4(const EriArgs&, const FusedSegs&, int, int, hipStream_t);
static_assert(HGP_NRGROUP == 5, "one launch_fused_register_g<k> per fusion group");

hipError_t launch_fused_register(int group, const EriArgs& args, const FusedSegs& F, int nblocks,
                                 hipStream_t stream) {
  const int wpb = kFusedWavesPerBlock[group];
  switch (group) {
    case 0: return launch_fused_register_g0(args, F, nblocks, wpb, stream);
    case 1: return launch_fused_register_g1(args, F, nblocks, wpb, stream);
    case 2: return launch_fused_register_g2(args, F, nblocks, wpb, stream);
    case 3: return launch_fused_register_g3(args, F, nblocks, wpb, stream);
    case 4: return launch_fused_register_g4(args, F, nblocks, wpb, stream);
    default: return hipErrorInvalidValue;
  }
}

int register_group(int c) { return kRegisterGroup[c]; }

int wave_kernel_acc(int c) {
  const int nacc = kWClassDesc[c].nacc;
  return nacc <= 64 * 2 ? 2 : (nacc <= 64 * 4 ? 4 : (nacc <= 64 * 8 ? 8 : 16));
}

hipError_t launch_fused_wave(int acc, int nw, const EriArgs& args, const FusedSegs& F, size_t lds, int nblocks,
                             void* wtables, hipStream_t stream) {
  const WTablesHost* t = static_cast<const WTablesHost*>(wtables);
  WTablesDev W{t->blob};
  const void* fn = acc <= 2 ? (const void*)hgp_wf_kernel<2>
                            : (acc <= 4 ? (const void*)hgp_wf_kernel<4>
                                        : (acc <= 8 ? (const void*)hgp_wf_kernel<8> : (const void*)hgp_wf_kernel<16>));
  void* kargs[] = {const_cast<EriArgs*>(&args), const_cast<FusedSegs*>(&F), &W};
  return hipLaunchKernel(fn, dim3(nblocks), dim3(64 * nw), kargs, lds, stream);
}


static size_t wg_table_bytes(int c) {
  const WClassDesc& D = kWClassDesc[c];
  return 4 * (size_t)((D.ga + 2 * D.nacc - D.gv + 3) & ~3) + 16;  // + alignment slack
}

static bool wg_hrr_in_lds(int c) {
  const WClassDesc& D = kWClassDesc[c];
  return sizeof(double) * (size_t)(4 * W_SCL + D.nvrr + D.nhrr) <= WG_HRR_LDS_MAX;
}

bool wg_tables_fit(int c) {
  const WClassDesc& D = kWClassDesc[c];
  return wg_hrr_in_lds(c) &&
         sizeof(double) * (size_t)(4 * W_SCL + D.nvrr + D.nhrr) + wg_table_bytes(c) <= WG_HRR_LDS_MAX;
}

template <bool JR>
static const void* wg_kernel_fn_t(int c, bool tl) {
  if (tl && wg_tables_fit(c)) return (const void*)hgp_wg_kernel<true, true, JR>;
  return wg_hrr_in_lds(c) ? (const void*)hgp_wg_kernel<true, false, JR> : (const void*)hgp_wg_kernel<false, false, JR>;
}

static const void* wg_kernel_fn(int c, bool tl, bool jr = false) {
  return jr ? wg_kernel_fn_t<true>(c, tl) : wg_kernel_fn_t<false>(c, tl);
}

size_t wgclass_lds_bytes(int c, bool tl) {
  const WClassDesc& D = kWClassDesc[c];
  return sizeof(double) * (size_t)(4 * W_SCL + D.nvrr + (wg_hrr_in_lds(c) ? D.nhrr : 0)) +
         (tl && wg_tables_fit(c) ? wg_table_bytes(c) : 0);
}

bool wgclass_fits_nw(int c, int nw, size_t lds_limit) {
  return (nw == 1 || nw == 2 || nw == 4 || nw == 8) && wgclass_lds_bytes(c, false) <= lds_limit;
}

int wg_kernel_blocks_per_cu(int c, int nw, bool tl) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, wg_kernel_fn(c, tl), 64 * nw,
                                                   wgclass_lds_bytes(c, tl)) != hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

hipError_t launch_fused_wave_global(int nw, const EriArgs& args, const FusedSegs& F, size_t lds, int nblocks,
                                    void* wtables, double* scratch, hipStream_t stream) {
  const WTablesHost* t = static_cast<const WTablesHost*>(wtables);
  const uint32_t* gb = t->gblob;
  void* kargs[] = {const_cast<EriArgs*>(&args), const_cast<FusedSegs*>(&F), &gb, &scratch};
  return hipLaunchKernel((const void*)hgp_wgf_kernel, dim3(nblocks), dim3(64 * nw), kargs, lds, stream);
}

size_t wx_scratch_doubles(int c, int nblocks) {
  return 2 * (size_t)nblocks * (kWClassDesc[c].nvrr + kWClassDesc[c].nhrr);
}

hipError_t launch_class_extended(int c, const EriArgs& args, int nblocks, int64_t nq_launch, void* wtables,
                                 double* scratch, hipStream_t stream) {
  const WTablesHost* t = static_cast<const WTablesHost*>(wtables);
  const WClassDesc D = kWClassDesc[c];
  const uint32_t* gb = t->gblob;
  double2* sc = reinterpret_cast<double2*>(scratch);
  void* kargs[] = {const_cast<EriArgs*>(&args), const_cast<WClassDesc*>(&D), &gb, &sc, &nq_launch};
  return hipLaunchKernel((const void*)hgp_wx_kernel, dim3(nblocks), dim3(64), kargs, 0, stream);
}

size_t wg_scratch_doubles(int c, int nblocks) { return wg_hrr_in_lds(c) ? 0 : (size_t)nblocks * kWClassDesc[c].nhrr; }

hipError_t launch_class_wave_global(int c, int nw, bool tl, const EriArgs& args, int nblocks, int64_t nq_launch,
                                    void* wtables, double* scratch, hipStream_t stream) {
  const WTablesHost* t = static_cast<const WTablesHost*>(wtables);
  const WClassDesc D = kWClassDesc[c];
  const uint32_t* gb = t->gblob;
  size_t lds = wgclass_lds_bytes(c, tl);
  EriArgs a = with_jk_run(args, c, wg_kernel_fn(c, tl, true), nw, lds, nq_launch, nblocks);
  void* kargs[] = {&a, const_cast<WClassDesc*>(&D), &gb, &scratch, &nq_launch};
  return hipLaunchKernel(wg_kernel_fn(c, tl, a.jk_run > 0), dim3(nblocks), dim3(64 * nw), kargs, lds, stream);
}

// Rys kernel (mi_rys.hip) on the class descriptors and global op tables
bool ryclass_fits_nw(const WClassDesc& D, int nw, size_t lds_limit);
size_t ry_scratch_doubles(const WClassDesc& D, int nw, int nblocks);
int ry_kernel_blocks_per_cu(const WClassDesc& D, int nw);
hipError_t launch_rys(const WClassDesc& D, int nw, const EriArgs& args, int nblocks, int64_t nq_launch,
                      const uint32_t* gblob, double* scratch, hipStream_t stream);
hipError_t set_rys_lds_limit(size_t bytes);
bool rygclass_fits(const WClassDesc& D, int lg);
int ryg_kernel_blocks_per_cu(const WClassDesc& D, int lg);
hipError_t launch_ryg(const WClassDesc& D, int lg, const EriArgs& args, int nblocks, int64_t nq_launch,
                      const uint32_t* gblob, hipStream_t stream);

bool rygclass_fits(int c, int lg) { return rygclass_fits(kWClassDesc[c], lg); }
int ryg_kernel_blocks_per_cu(int c, int lg) { return ryg_kernel_blocks_per_cu(kWClassDesc[c], lg); }
hipError_t launch_class_ryg(int c, int lg, const EriArgs& args, int nblocks, int64_t nq_launch, void* wtables,
                            hipStream_t stream) {
  const WTablesHost* t = static_cast<const WTablesHost*>(wtables);
  return launch_ryg(kWClassDesc[c], lg, args, nblocks, nq_launch, t->gblob, stream);
}

bool ryclass_fits_nw(int c, int nw, size_t lds_limit) { return ryclass_fits_nw(kWClassDesc[c], nw, lds_limit); }
size_t ry_scratch_doubles(int c, int nw, int nblocks) { return ry_scratch_doubles(kWClassDesc[c], nw, nblocks); }
int ry_kernel_blocks_per_cu(int c, int nw) { return ry_kernel_blocks_per_cu(kWClassDesc[c], nw); }
hipError_t launch_class_rys(int c, int nw, const EriArgs& args, int nblocks, int64_t nq_launch, void* wtables,
                            double* scratch, hipStream_t stream) {
  const WTablesHost* t = static_cast<const WTablesHost*>(wtables);
  return launch_rys(kWClassDesc[c], nw, args, nblocks, nq_launch, t->gblob, scratch, stream);
}

hipError_t launch_zero(double* x, int64_t n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const int64_t nb = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)nb), dim3(256), 0, stream, x, n);
  return hipGetLastError();
}

hipError_t launch_symmetrize(double* J, double* K, int n, hipStream_t stream) {
  const int64_t npair = (int64_t)n * (n + 1) / 2;
  const int64_t nb = std::min<int64_t>((npair + 255) / 256, 4096);
  hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)std::max<int64_t>(nb, 1)), dim3(256), 0, stream, J, K, n);
  return hipGetLastError();
}

hipError_t set_wave_kernel_lds_limit(size_t bytes) {
  const void* wk[] = {(const void*)hgp_w_kernel<2, false, false>, (const void*)hgp_w_kernel<4, false, false>,
                      (const void*)hgp_w_kernel<8, false, false>, (const void*)hgp_w_kernel<16, false, false>,
                      (const void*)hgp_w_kernel<2, true, false>,  (const void*)hgp_w_kernel<4, true, false>,
                      (const void*)hgp_w_kernel<8, true, false>,  (const void*)hgp_w_kernel<16, true, false>,
                      (const void*)hgp_w_kernel<2, false, true>,  (const void*)hgp_w_kernel<4, false, true>,
                      (const void*)hgp_w_kernel<8, false, true>,  (const void*)hgp_w_kernel<16, false, true>,
                      (const void*)hgp_w_kernel<2, true, true>,   (const void*)hgp_w_kernel<4, true, true>,
                      (const void*)hgp_w_kernel<8, true, true>,   (const void*)hgp_w_kernel<16, true, true>,
                      (const void*)hgp_wf_kernel<2>,        (const void*)hgp_wf_kernel<4>,
                      (const void*)hgp_wf_kernel<8>,        (const void*)hgp_wf_kernel<16>};
  for (const void* f : wk) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
  }
  const void* gk[] = {(const void*)hgp_wgf_kernel,
                      (const void*)hgp_wg_kernel<false, false, false>, (const void*)hgp_wg_kernel<true, false, false>,
                      (const void*)hgp_wg_kernel<true, true, false>,   (const void*)hgp_wg_kernel<false, false, true>,
                      (const void*)hgp_wg_kernel<true, false, true>,   (const void*)hgp_wg_kernel<true, true, true>};
  for (const void* f : gk) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
  }
  g_lds_limit = bytes;
  return set_rys_lds_limit(bytes);
}

}  // namespace mi
