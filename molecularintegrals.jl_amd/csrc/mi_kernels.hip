// mi_kernels.hip -- HGP ERI device kernels for gfx950 (CDNA4).
//
// Per execute:
//   1. pair_prims_kernel   one lane per primitive pair: P, zeta, PA, folded
//                          prefactor K, 1/(2 zeta); lane 0 of a pair writes AB.
//   2. per quartet class   either a generated register kernel (hgp_r_<class>,
//                          lane group per shell quartet) or the table-driven
//                          wave-cooperative kernel hgp_w_kernel (one wavefront
//                          per shell quartet, VRR/HRR staged in LDS).
// MFMA is not used: HGP's VRR/HRR are short dependent FMA chains with
// per-quartet coefficients, not dense contractions.
#include "mi_device.h"
#include "mi_kernels.h"

#include "gen/hgp_wtables.inc"
#include "gen/hgp_rkernels.inc"

// ---------------------------------------------------------------------------
// Shell-pair precompute (reference: the pair part of vrr!'s prefactor block,
// src/HGP.jl:139-155, hoisted out of the quartet loop).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) pair_prims_kernel(int npp, const int32_t* __restrict__ pp_pair,
                                                         const PairTopo* __restrict__ topo,
                                                         const double* __restrict__ shell_xyz,
                                                         const int32_t* __restrict__ shell_prim_off,
                                                         const double* __restrict__ prim_exp,
                                                         const double* __restrict__ prim_cn,
                                                         PrimPair* __restrict__ pp, PairGeom* __restrict__ geom) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= npp) return;
  const int p = pp_pair[r];
  const PairTopo t = topo[p];
  const int local = r - t.ppoff;
  const int ia = local / t.nprimB, ib = local - ia * t.nprimB;
  const int pa = shell_prim_off[t.shA] + ia, pb = shell_prim_off[t.shB] + ib;
  const double a = prim_exp[pa], b = prim_exp[pb];
  const double Ax = shell_xyz[3 * t.shA], Ay = shell_xyz[3 * t.shA + 1], Az = shell_xyz[3 * t.shA + 2];
  const double Bx = shell_xyz[3 * t.shB], By = shell_xyz[3 * t.shB + 1], Bz = shell_xyz[3 * t.shB + 2];
  const double zeta = a + b, ooz = 1.0 / zeta;
  const double Px = (a * Ax + b * Bx) * ooz, Py = (a * Ay + b * By) * ooz, Pz = (a * Az + b * Bz) * ooz;
  const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;
  const double rab2 = ABx * ABx + ABy * ABy + ABz * ABz;
  // 2 pi^{5/2} / (zeta eta sqrt(zeta+eta)) exp(...) split as K_ab K_cd / sqrt(zeta+eta)
  const double kfac = 5.9149671727956129;  // sqrt(2) * pi^{5/4}
  PrimPair o;
  o.Px = Px; o.Py = Py; o.Pz = Pz; o.zeta = zeta;
  o.PAx = Px - Ax; o.PAy = Py - Ay; o.PAz = Pz - Az;
  o.K = kfac * prim_cn[pa] * prim_cn[pb] * exp(-a * b * rab2 * ooz) * ooz;
  o.oo2z = 0.5 * ooz;
  o.pad0 = o.pad1 = o.pad2 = 0.0;
  pp[r] = o;
  if (local == 0) {
    PairGeom g;
    g.ABx = ABx; g.ABy = ABy; g.ABz = ABz; g.Q = 0.0;
    geom[p] = g;
  }
}

// ---------------------------------------------------------------------------
// Wave-cooperative kernel: one 64-lane block per shell quartet.
// ---------------------------------------------------------------------------
struct WTablesDev {
  const uint4_t* vitems;
  const uint2_t* accmap;
  const uint2_t* hitems;
  const uint32_t* outs;
};

__device__ __forceinline__ double sel3(int i, double x, double y, double z) { return i == 0 ? x : (i == 1 ? y : z); }

__global__ void __launch_bounds__(64) hgp_w_kernel(const EriArgs args, const WClassDesc D, const WTablesDev W) {
  extern __shared__ double lds[];
  double* vrr = lds;
  double* hrr = lds + D.nvrr;
  const int lane = threadIdx.x;
  const TaskDesc T = args.tasks[find_task(args)];
  const int64_t qi = (int64_t)(blockIdx.x - T.blockStart);
  if (qi >= T.nq) return;  // uniform per block
  QuartetCtx qc;
  decode_index(T, qi, qc);
  qc.G = 1;
  qc.g = 0;
  const PairTopo P = args.topo[qc.bra];
  const PairTopo Q = args.topo[qc.ket];
  for (int x = lane; x < D.nhrr; x += 64) hrr[x] = 0.0;
  const int mtop = D.mmax - 1;
  const double* btab = args.boys + (size_t)mtop * BOYS_NROW * 8;
  const uint4_t* VI = W.vitems + D.voff;
  const uint2_t* AM = W.accmap + D.aoff;
  __syncthreads();
  for (int u = 0; u < P.npp; ++u) {
    const PrimPair bp = args.pp[P.ppoff + u];
    for (int v = 0; v < Q.npp; ++v) {
      const PrimPair kp = args.pp[Q.ppoff + v];
      const double zeta = bp.zeta, eta = kp.zeta;
      const double inv_s = 1.0 / (zeta + eta);
      const double PQx = bp.Px - kp.Px, PQy = bp.Py - kp.Py, PQz = bp.Pz - kp.Pz;
      const double T2 = zeta * eta * inv_s * (PQx * PQx + PQy * PQy + PQz * PQz);
      const double pref = bp.K * kp.K * sqrt(inv_s);
      // Boys: every lane evaluates the top order and the recursion; lane m keeps F_m.
      if (lane < D.mmax) {
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
          for (int m = mtop - 1; m >= lane; --m) fm = fma(t2, fm, eT) / (double)(2 * m + 1);
        } else {
          const double rT = 1.0 / T2;
          fm = 0.88622692545275801365 * sqrt(rT);
          const double o2T = 0.5 * rT;
          for (int m = 1; m <= lane; ++m) fm *= (2 * m - 1) * o2T;
        }
        vrr[D.base00 + lane] = pref * fm;
      }
      __syncthreads();
      const double oo2z = bp.oo2z, oo2e = kp.oo2z, oo2s = 0.5 * inv_s;
      const double eta_s = eta * inv_s, zeta_s = zeta * inv_s;
      for (int lv = 0; lv < D.nvlev; ++lv) {
        const int i0 = D.vlev[lv], i1 = D.vlev[lv + 1];
        for (int it = i0 + lane; it < i1; it += 64) {
          const uint4_t o = VI[it];
          const int t = o.x & 0xFFFF, s1 = o.x >> 16;
          const int dir = o.z & 3;
          const int kind = (o.z >> 14) & 3;
          const bool has2 = (o.z >> 16) & 1;
          const double PQd = sel3(dir, PQx, PQy, PQz);
          double val;
          if (kind == 1) {
            const double PA = sel3(dir, bp.PAx, bp.PAy, bp.PAz);
            const double WP = -eta * inv_s * PQd;
            val = PA * vrr[s1] + WP * vrr[s1 + 1];
            if (has2) {
              const int s2 = o.y & 0xFFFF;
              const double n1 = (double)((o.z >> 2) & 63);
              val += (n1 * oo2z) * (vrr[s2] - eta_s * vrr[s2 + 1]);
            }
          } else {
            const double QC = sel3(dir, kp.PAx, kp.PAy, kp.PAz);
            const double WQ = zeta * inv_s * PQd;
            val = QC * vrr[s1] + WQ * vrr[s1 + 1];
            if ((o.z >> 17) & 1) {
              const int s3 = o.y >> 16;
              const double ne = (double)((o.z >> 2) & 63);
              val += (ne * oo2s) * vrr[s3 + 1];
            }
            if (has2) {
              const int s2 = o.y & 0xFFFF;
              const double nf = (double)((o.z >> 8) & 63);
              val += (nf * oo2e) * (vrr[s2] - zeta_s * vrr[s2 + 1]);
            }
          }
          vrr[t] = val;
        }
        __syncthreads();
      }
      for (int x = lane; x < D.nacc; x += 64) {
        const uint2_t m = AM[x];
        hrr[m.y] += vrr[m.x];
      }
      __syncthreads();
    }
  }
  // HRR levels
  const PairGeom PG = args.geom[qc.bra];
  const PairGeom QG = args.geom[qc.ket];
  const uint2_t* HI = W.hitems + D.hoff;
  for (int lv = 0; lv < D.nhlev; ++lv) {
    const int i0 = D.hlev[lv], i1 = D.hlev[lv + 1];
    for (int it = i0 + lane; it < i1; it += 64) {
      const uint2_t o = HI[it];
      const int t = o.x & 0xFFFF, s1 = o.x >> 16, s2 = o.y & 0xFFFF;
      const int j = (o.y >> 16) & 3;
      const bool ket = (o.y >> 18) & 1;
      const double X = ket ? sel3(j, QG.ABx, QG.ABy, QG.ABz) : sel3(j, PG.ABx, PG.ABy, PG.ABz);
      hrr[t] = hrr[s1] + X * hrr[s2];
    }
    __syncthreads();
  }
  const ScatterCtx sc = make_scatter(args, qc, P, Q);
  const uint32_t* OU = W.outs + D.ooff;
  for (int x = lane; x < D.nout; x += 64) {
    const uint32_t o = OU[x];
    const int off = o & 0xFFFF;
    const int ma = (o >> 16) & 15, mb = (o >> 20) & 15, mc = (o >> 24) & 15, md = (o >> 28) & 15;
    const int fa = sc.fa0 + ma, fb = sc.fb0 + mb, fc = sc.fc0 + mc, fd = sc.fd0 + md;
    const int64_t slot = unique_slot(sc, fa, fb, fc, fd);
    if (slot >= 0)
      sc.out[slot] = hrr[off] * (args.bf_scale[fa] * args.bf_scale[fb]) * (args.bf_scale[fc] * args.bf_scale[fd]);
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
  uint4_t* v = nullptr;
  uint2_t* a = nullptr;
  uint2_t* h = nullptr;
  uint32_t* o = nullptr;
};

hipError_t upload_wtables(void** handle) {
  WTablesHost* t = new WTablesHost();
  hipError_t e;
  if ((e = hipMalloc(&t->v, sizeof(kWVItems))) != hipSuccess) return e;
  if ((e = hipMalloc(&t->a, sizeof(kWAccMap))) != hipSuccess) return e;
  if ((e = hipMalloc(&t->h, sizeof(kWHItems))) != hipSuccess) return e;
  if ((e = hipMalloc(&t->o, sizeof(kWOut))) != hipSuccess) return e;
  if ((e = hipMemcpy(t->v, kWVItems, sizeof(kWVItems), hipMemcpyHostToDevice)) != hipSuccess) return e;
  if ((e = hipMemcpy(t->a, kWAccMap, sizeof(kWAccMap), hipMemcpyHostToDevice)) != hipSuccess) return e;
  if ((e = hipMemcpy(t->h, kWHItems, sizeof(kWHItems), hipMemcpyHostToDevice)) != hipSuccess) return e;
  if ((e = hipMemcpy(t->o, kWOut, sizeof(kWOut), hipMemcpyHostToDevice)) != hipSuccess) return e;
  *handle = t;
  return hipSuccess;
}

void free_wtables(void* handle) {
  WTablesHost* t = static_cast<WTablesHost*>(handle);
  if (!t) return;
  hipFree(t->v);
  hipFree(t->a);
  hipFree(t->h);
  hipFree(t->o);
  delete t;
}

hipError_t launch_pair_prims(int npp, const int32_t* pp_pair, const PairTopo* topo, const double* shell_xyz,
                             const int32_t* shell_prim_off, const double* prim_exp, const double* prim_cn,
                             PrimPair* pp, PairGeom* geom, hipStream_t stream) {
  if (npp == 0) return hipSuccess;
  dim3 grid((npp + 255) / 256), block(256);
  hipLaunchKernelGGL(pair_prims_kernel, grid, block, 0, stream, npp, pp_pair, topo, shell_xyz, shell_prim_off,
                     prim_exp, prim_cn, pp, geom);
  return hipGetLastError();
}

hipError_t launch_class_register(int c, const EriArgs& args, int nblocks, hipStream_t stream) {
  const void* fn = kRegisterKernels[c];
  void* kargs[] = {const_cast<EriArgs*>(&args)};
  return hipLaunchKernel(fn, dim3(nblocks), dim3(256), kargs, 0, stream);
}

size_t wclass_lds_bytes(int c) {
  const WClassDesc& D = kWClassDesc[c];
  return sizeof(double) * (size_t)(D.nvrr + D.nhrr);
}

hipError_t launch_class_wave(int c, const EriArgs& args, int nblocks, void* wtables, hipStream_t stream) {
  const WTablesHost* t = static_cast<const WTablesHost*>(wtables);
  WTablesDev W{t->v, t->a, t->h, t->o};
  const WClassDesc D = kWClassDesc[c];
  const size_t lds = wclass_lds_bytes(c);
  hipLaunchKernelGGL(hgp_w_kernel, dim3(nblocks), dim3(64), lds, stream, args, D, W);
  return hipGetLastError();
}

hipError_t set_wave_kernel_lds_limit(size_t bytes) {
  return hipFuncSetAttribute((const void*)hgp_w_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

}  // namespace mi
