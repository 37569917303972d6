// mi_kernels.h -- host-side entry points into mi_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "mi_device.h"

namespace mi {

int nclass();
int lmax_shell();
const WClassDesc& wclass(int c);
bool has_register_kernel(int c);

hipError_t upload_wtables(void** handle);
void free_wtables(void* handle);

hipError_t launch_pair_prims(int npp, const int32_t* pp_pair, const int32_t* pp_ab, const PairTopo* topo,
                             const double* shell_xyz,
                             const int32_t* shell_prim_off, const double* prim_exp, const double* prim_cn,
                             PrimPair* pp, PairGeom* geom, const double2* bounds, hipStream_t stream);
hipError_t launch_class_register(int c, const EriArgs& args, int nblocks, int wpb, hipStream_t stream);
// wavefronts per block of class c's register kernel in unique-array plans (1
// or 4; J/K plans use 4) and of fusion group g's fused kernel
int register_waves_per_block(int c);
int fused_register_waves_per_block(int g);
// LDS wave kernel; tg: op tables read from global memory (tuning variant 3)
// instead of staged into LDS (variant 0)
hipError_t launch_class_wave(int c, int nw, bool tg, const EriArgs& args, int nblocks, int64_t nq_launch,
                             void* wtables, hipStream_t stream);
int wave_kernel_blocks_per_cu(int c, int nw, bool tg = false);
bool wclass_fits_nw(int c, int nw, size_t lds_limit, bool tg = false);
size_t wclass_lds_bytes(int c, int nw, bool tg = false);
bool wclass_fits(int c);
hipError_t set_wave_kernel_lds_limit(size_t bytes);
// global-table kernel (variant 2): nw in {1, 2, 4, 8} wavefronts per block,
// scratch = wg_scratch_doubles(c, nblocks) doubles; tl: op tables staged in
// LDS (only where wg_tables_fit(c); more LDS per block, fewer blocks per CU)
bool wg_tables_fit(int c);
size_t wgclass_lds_bytes(int c, bool tl);
bool wgclass_fits_nw(int c, int nw, size_t lds_limit);
int wg_kernel_blocks_per_cu(int c, int nw, bool tl);
size_t wg_scratch_doubles(int c, int nblocks);
hipError_t launch_class_wave_global(int c, int nw, bool tl, const EriArgs& args, int nblocks, int64_t nq_launch,
                                    void* wtables, double* scratch, hipStream_t stream);
// extended-precision (double-double) table kernel, variant 4: one wavefront
// per block, one quartet per block at a time; scratch = wx_scratch_doubles(c,
// nblocks) doubles, 16-byte aligned
size_t wx_scratch_doubles(int c, int nblocks);
hipError_t launch_class_extended(int c, const EriArgs& args, int nblocks, int64_t nq_launch, void* wtables,
                                 double* scratch, hipStream_t stream);
// fused launches (segments in F, bases of tasks / block_task in args):
// register kernels (mi_fused.hip), wave kernels of accumulator count acc
// (wave_kernel_acc of each member class) with nw wavefronts per block
hipError_t launch_fused_register(int group, const EriArgs& args, const FusedSegs& F, int nblocks, hipStream_t stream);
int register_group(int c);  // fusion group of class c's register kernel, -1 if none
int wave_kernel_acc(int c);
hipError_t launch_fused_wave(int acc, int nw, const EriArgs& args, const FusedSegs& F, size_t lds, int nblocks,
                             void* wtables, hipStream_t stream);
// global-table kernels of nw wavefronts per block, fused; a segment whose
// HRR area does not fit in LDS uses scratch + F.bt_off[s]
hipError_t launch_fused_wave_global(int nw, const EriArgs& args, const FusedSegs& F, size_t lds, int nblocks,
                                    void* wtables, double* scratch, hipStream_t stream);
// Rys quadrature kernel, variant 5 (mi_rys.hip): nw in {1, 2, 4} wavefronts
// per block; scratch = ry_scratch_doubles(c, nw, nblocks) doubles (HRR areas
// that do not fit in LDS)
bool ryclass_fits_nw(int c, int nw, size_t lds_limit);
size_t ry_scratch_doubles(int c, int nw, int nblocks);
int ry_kernel_blocks_per_cu(int c, int nw);
hipError_t launch_class_rys(int c, int nw, const EriArgs& args, int nblocks, int64_t nq_launch, void* wtables,
                            double* scratch, hipStream_t stream);
// Rys lane-group kernel, variant 6 (mi_rys.hip): one-wavefront blocks of
// 64/lg groups of lg in {16, 32, 64} lanes, one shell quartet per group
bool rygclass_fits(int c, int lg);
int ryg_kernel_blocks_per_cu(int c, int lg);
hipError_t launch_class_ryg(int c, int lg, const EriArgs& args, int nblocks, int64_t nq_launch, void* wtables,
                            hipStream_t stream);
// host evaluation of the kernels' Rys rule: n roots x (= t^2) and weights at T
void rys_rule_host(int n, double T, double* x, double* w);
int rys_nmax();
hipError_t launch_zero(double* x, int64_t n, hipStream_t stream);
hipError_t launch_symmetrize(double* J, double* K, int n, hipStream_t stream);

}  // namespace mi
