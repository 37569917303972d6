// mi_internal.h -- library-internal types shared by mi_host.cpp (plans, C
// ABI) and mi_server.cpp (serving engine).  Not installed; the ABI is
// include/mi_hgp.h.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mi_hgp.h"
#include "mi_kernels.h"

#define MI_NBUCKET 5
#define MI_NREGIME 3             // tuning regimes: small and large class launches, J/K plans
#define MI_SMALL_LAUNCH 65536    // shell quartets below which a class launch is "small" (latency-bound)
#define MI_PRIM_PAIR_CUT 1e-30   // primitive pairs with |K| below this times their shell pair's largest are dropped
#define MI_MAX_STREAMS 16

// Kernel choice for one (quartet class, bra bucket, ket bucket): variant 1 =
// generated register kernel with G lanes per quartet, 0 = LDS wave kernel with
// G wavefronts per quartet, 2 = global-table kernel with G wavefronts per block,
// 3 = LDS wave kernel with its op tables read from global memory (less LDS per
// block, more blocks resident).
struct TuneEntry {
  int variant = -1;  // -1: not tuned, use the heuristic
  int G = 1;
};

struct mi_ctx {
  int device = 0;
  char* noalloc_scratch = nullptr;  // device address of the static scratch area on this device (or null)
  double* d_boys = nullptr;
  void* wtables = nullptr;
  std::string err;
  size_t lds_limit = 0;
  int ncu = 256;
  int nstreams = 4;                           // class launches spread over this many streams
  bool fuse = true;                           // fuse small class launches (mi_ctx_set_fusion)
  hipStream_t aux[MI_MAX_STREAMS - 1] = {};   // default-priority streams beside the caller's
  hipStream_t bstream = nullptr;              // Schwarz bounds passes at plan build
  std::vector<TuneEntry> tune;                // [regime][class][bra bucket][ket bucket]
};

struct Launch {
  int cls;
  int var;  // 1 register kernel, 0 LDS wave kernel, 2 global-table kernel, 3 wave kernel, global tables
  int nw;   // wave kernels: wavefronts per block
  bool tl;  // global-table kernel: op tables staged in LDS
  size_t scratch_off;  // global-table kernel: offset of its slice of the plan scratch (doubles)
  int wpb = 4;         // register kernel: wavefronts per block
  int64_t bt_off;      // register kernel: offset of its block -> task table in the plan's
  int task_off, ntasks, nblocks;
  int64_t nq_shell, nq_prim;
  double cost;  // relative device time estimate (longest launches go first)
  int lane;     // stream lane: 0 = caller's stream, 1.. = context aux streams
};

// Launch unit of a unique-array execute: one class launch (kind 0), or a
// fused launch of several (kind 1: register kernels of one fusion group,
// kind 2: LDS wave kernels of nw wavefronts per block, run in the
// instantiation of the members' largest accumulator count, kind 3:
// global-table kernels of nw wavefronts per block).  Class launches with
// fewer than MI_SMALL_LAUNCH shell quartets are latency-bound and leave most
// of the chip idle, so fusing them runs their classes side by side.
struct Unit {
  int kind = 0;
  std::vector<int> members;  // indices into mi_plan::launches, heaviest first
  int nblocks = 0, nw = 0;
  int acc = 0;               // kind 1: fusion group; kind 2: accumulator count of the instantiation
  size_t lds = 0;            // dynamic LDS (fused wave kernels)
  double cost = 0.0;
  int lane = 0;
  FusedSegs segs{};
};

struct mi_plan {
  mi_ctx* ctx = nullptr;
  int nshell = 0, nbf = 0, npair = 0, npp = 0;
  int64_t nq_shell = 0, nq_prim = 0, n_unique = 0;
  int order = MI_ORDER_CHRR;
  int nmol = 1;
  std::vector<int64_t> mol_off;  // output offset of each molecule, [nmol+1]
  // device buffers
  double* d_shell_xyz = nullptr;
  int32_t* d_shell_prim_off = nullptr;
  double* d_prim_exp = nullptr;
  double* d_prim_cn = nullptr;
  double* d_bf_scale = nullptr;
  PairTopo* d_topo = nullptr;
  PairGeom* d_geom = nullptr;
  PrimPair* d_pp = nullptr;
  int32_t* d_pp_pair = nullptr;
  int32_t* d_pp_ab = nullptr;       // (ia | ib << 16) of each kept primitive pair
  TaskDesc* d_tasks = nullptr;
  double* d_scratch = nullptr;     // HRR work areas of the global-table launches
  size_t scratch_doubles = 0;
  std::vector<Launch> launches;   // in issue order (cost descending)
  std::vector<Launch> slaunches;  // Schwarz pass (diagonal quartets), run before the main launches
  std::vector<Unit> units;        // unique-array executes issue these (J/K and profiles: launches)
  bool need_schwarz = false;
  std::vector<int32_t> pair_mol;  // molecule of each pair record
  double tau2 = 0.0, cut2 = 0.0;
  int nlanes = 1;
  hipGraphExec_t gexec = nullptr;  // captured execute (mi_plan_graph_capture)
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[MI_MAX_STREAMS] = {};
  // Stream-ordered allocation (serving): device buffers come from
  // hipMallocAsync on this stream and go back with hipFreeAsync, so building
  // and retiring a plan never waits for work on other streams.  The host
  // copies of uploaded data live as long as the plan.
  hipStream_t alloc_stream = nullptr;
  // Arena placement (the entry point that allocates nothing): device buffers
  // are carved from [arena, arena + arena_cap) and filled on arena_stream;
  // nothing is freed.
  char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  hipStream_t arena_stream = nullptr;
  bool bounds_only = false;         // Schwarz pass only (mi_plan_build with MI_BUILD_BOUNDS_ONLY)
  double2* d_bounds = nullptr;      // per pair (Q, V) computed when the plan was built, or null
  int32_t* d_block_task = nullptr;  // register launches: task of each block's first wavefront
  unsigned long long* d_run_ctr = nullptr;  // J/K: run counter of each launch (cleared every J/K execute)
  size_t device_bytes = 0;  // bytes of device memory the plan holds (or needs, before upload)
  // Per-molecule output buffers (mi_plan_set_mol_outputs): molecule m writes
  // its unique array at mol_out[m] instead of out + mol_off[m].
  std::vector<double*> mol_out;
  struct HostData {         // host images of the uploaded buffers
    std::vector<double> xyz, pexp, pcn, bfs;
    std::vector<int32_t> poff, pp_pair, pp_ab;
    std::vector<PairTopo> topo;
    std::vector<TaskDesc> tasks;
    std::vector<double2> bounds;
    std::vector<int32_t> block_task;
  } h;
};

struct JKArgs {
  const double* D;
  double* J;
  double* K;
};

// Error text into the context; returns code.
int mi_set_err(mi_ctx* c, int code, const char* fmt, ...);

// Plan over nmol molecules.  Flags:
//   MI_BUILD_DEFER_UPLOAD  compute everything on the host (P->device_bytes is
//                          then what mi_plan_upload will allocate) and leave
//                          the device buffers to mi_plan_upload, whose
//                          alloc_stream != nullptr selects stream-ordered
//                          allocation (hipMallocAsync / hipFreeAsync);
//   MI_BUILD_NO_BOUNDS     do not run the Schwarz bounds pass at build time
//                          (screening and FP32 selection then happen in the
//                          kernels on bounds from a per-execute pass);
//   MI_BUILD_BOUNDS_ONLY   a plan of the diagonal quartets alone, whose
//                          execute fills the per-pair bounds;
//   MI_BUILD_ONE_LANE      one stream lane (no events are created).
#define MI_BUILD_DEFER_UPLOAD 1
#define MI_BUILD_NO_BOUNDS 2
#define MI_BUILD_BOUNDS_ONLY 4
#define MI_BUILD_ONE_LANE 8      // every launch on the caller's stream (no fork/join events)
int mi_plan_build(mi_ctx* ctx, const mi_basis* bases, int nmol, const mi_opts* opts, int flags, mi_plan** out);
int mi_plan_upload(mi_plan* P, hipStream_t alloc_stream);
// Before the upload of a deferred build: molecule m's unique array goes to
// outs[m] (caller-owned device buffers of mi_n_unique(nbf_m) doubles, any
// addresses) instead of out + mol_off[m]; executes then take out = outs[0]
// and zero-fill each buffer separately where the plan screens.
int mi_plan_set_mol_outputs(mi_plan* P, double* const* outs);
// One execute on stream (ev: per-launch timing events or nullptr; jk: the
// Coulomb/exchange arrays or nullptr for the unique array).
// unit_events: with ev, time the launch units (fused launches as executed)
// instead of each class launch.
int mi_plan_run(mi_plan* P, double* out, int64_t out_len, void* stream, hipEvent_t* ev, const JKArgs* jk,
                bool unit_events = false);
