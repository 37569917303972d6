// mi_host.cpp -- host side of libmi_hgp.so: C ABI, plans, Boys grid.
//
// Replaces the host half of the reference's HGP driver:
//   all_twoe_ints_chrr  src/HGP.jl:22-34  (quartet loop, scatter)
//   eri_fetcher         src/Basis.jl:201-212 (quartet -> slots map; here the
//                       slot is a closed form evaluated in the kernels)
// A plan sorts the shell pairs by (pair class, contraction bucket) so that a
// "task" -- the cross product of two contiguous pair lists -- is one
// angular-momentum class with similar primitive counts; each class is one
// kernel launch over a table of tasks.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mi_internal.h"

int mi_set_err(mi_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

namespace {

#define set_err mi_set_err

#define HIP_TRY(ctx, expr)                                                                   \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_err(ctx, MI_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// Boys values F_m(T) = e^{-T}/2 sum_n T^n / prod_{k<=n} (m + 1/2 + k)
// (power series of the lower incomplete gamma function, DLMF 8.7.1), in
// long double; all terms are positive.
long double boys_series(int m, long double T) {
  long double a = m + 0.5L, term = 1.0L / a, sum = term;
  for (int n = 1; n < 20000; ++n) {
    term *= T / (a + n);
    sum += term;
    if (term < sum * 1e-22L) break;
  }
  return 0.5L * expl(-T) * sum;
}

// Device grid: for each top order mtop, rows k = 0..480 (T_k = k/16) of
// { F_{mtop+j}(T_k)/j!  (j = 0..6),  exp(-T_k) }.
std::vector<double> make_boys_table() {
  std::vector<long double> F((size_t)BOYS_MCOL * BOYS_NROW);
  for (int k = 0; k < BOYS_NROW; ++k)
    for (int m = 0; m < BOYS_MCOL; ++m) F[(size_t)m * BOYS_NROW + k] = boys_series(m, (long double)k / 16.0L);
  std::vector<double> tab((size_t)(BOYS_MTOP_MAX + 1) * BOYS_NROW * 8);
  for (int mt = 0; mt <= BOYS_MTOP_MAX; ++mt)
    for (int k = 0; k < BOYS_NROW; ++k) {
      double* row = &tab[((size_t)mt * BOYS_NROW + k) * 8];
      long double fact = 1.0L;
      for (int j = 0; j < 7; ++j) {
        if (j > 0) fact *= j;
        row[j] = (double)(F[(size_t)(mt + j) * BOYS_NROW + k] / fact);
      }
      row[7] = (double)expl(-(long double)k / 16.0L);
    }
  return tab;
}

// Pair classes: (L_A, L_B) with L_A >= L_B at L_A (L_A+1)/2 + L_B, then the
// swapped ones (L_A < L_B) after them (the order of the generated class tables).
int pair_class(int la, int lb) {
  if (la >= lb) return la * (la + 1) / 2 + lb;
  const int lm = mi::lmax_shell();
  return (lm + 1) * (lm + 2) / 2 + lb * (lb - 1) / 2 + la;
}

// HRR cancellation factor of a shell pair taken in the orientation (A, B).
// chrr builds (a b| from (e 0| = (a + k, 0| (k <= L_B) through
// (a, b+1_j| = (a+1_j, b| + AB_j (a, b|.  For a primitive pair the (e 0|
// terms grow like (|PA| + s)^e with s = 1/sqrt(2 zeta) the width of the
// product Gaussian, and the result like (|PA| + s)^L_A (|PB| + s)^L_B, so
// the HRR cancels about ((|PA| + s) / (|PB| + s))^L_B of its inputs' digits:
// harmless when the tighter shell is A (P near A), up to 1e9 and more when a
// diffuse A sits far from a tight B (P near B, |PA| ~ |AB|).  The maximum
// over the pair's primitive pairs (|PA| = b/zeta |AB|, |PB| = a/zeta |AB|).
static double hrr_amplification(const double* xyz, const std::vector<int32_t>& L, const std::vector<int32_t>& poff,
                                const std::vector<double>& pexp, int A, int B) {
  if (L[B] == 0) return 1.0;
  double r2 = 0.0;
  for (int k = 0; k < 3; ++k) r2 += (xyz[3 * A + k] - xyz[3 * B + k]) * (xyz[3 * A + k] - xyz[3 * B + k]);
  const double r = std::sqrt(r2);
  double rmax = 1.0;  // x^L_B rises with x > 0: the largest ratio gives the largest power
  for (int pa = poff[A]; pa < poff[A + 1]; ++pa)
    for (int pb = poff[B]; pb < poff[B + 1]; ++pb) {
      const double a = pexp[pa], b = pexp[pb], z = a + b, s = 1.0 / std::sqrt(2.0 * z);
      rmax = std::max(rmax, (b / z * r + s) / (a / z * r + s));
    }
  double amp = rmax;
  for (int e = 1; e < L[B]; ++e) amp *= rmax;
  return amp;
}

// Largest HRR amplification a pair with L_B transferred units may keep in
// FP64: 100 for d and f transfers (pairs of amplification 290-490 already
// miss the 1e-12/1e-10 bar), 1e4 for a p transfer, whose single HRR level
// stays below the bar up to that point on every sampled workload (the seeded
// 920-function cluster's 100,100 benchmark quartets, 165,000 quartets of the
// 512-molecule set).
static double amp_limit(int lb) { return lb == 1 ? 1.0e4 : 100.0; }

// Orientation (A, B) of the shell pair {I, J}: the canonical L_A >= L_B, with
// the tighter shell on A between shells of equal L (the smaller HRR
// amplification; plain FP64 HGP otherwise loses up to 1e-8 absolute on
// far-apart tight/diffuse pairs of f shells), and swapped to L_A < L_B when
// the canonical orientation amplifies beyond amp_limit and the swapped one
// less.  Swapped pairs form their own pair classes and run in the
// table-driven kernels; on the seeded cluster 0.1% of pairs swap.
// Double-double launches: one wavefront per block, as many persistent blocks
// as MI_WX_SCRATCH bytes of work areas allow (16 bytes per VRR/HRR entry).
#define MI_WX_SCRATCH ((size_t)16 << 20)
static void orient_pair(const double* xyz, const std::vector<int32_t>& L, const std::vector<int32_t>& poff,
                        const std::vector<double>& pexp, int I, int J, int& A, int& B) {
  A = I;
  B = J;
  if (L[J] > L[I]) std::swap(A, B);
  if (I == J) return;
  const double keep = hrr_amplification(xyz, L, poff, pexp, A, B);
  if (L[A] > L[B] && keep <= amp_limit(L[B])) return;
  if (hrr_amplification(xyz, L, poff, pexp, B, A) < keep) std::swap(A, B);
}
int quartet_class(int X, int Y) { return X >= Y ? X * (X + 1) / 2 + Y : Y * (Y + 1) / 2 + X; }

int kbucket(int K) {
  if (K <= 1) return 0;
  if (K <= 4) return 1;
  if (K <= 9) return 2;
  if (K <= 36) return 3;
  return 4;
}

// Finer contraction buckets for the pair lists of unique-array plans, nested
// in kbucket's (the tuning table stays indexed by kbucket): a task pairs two
// lists, and the lanes of a register-kernel wavefront loop as long as the
// deepest quartet among them, so lists of pairs of similar primitive-pair
// count K keep the lanes busy (quartet K = K_bra K_ket varies by up to 3.5x
// per pair within one kbucket).
int kbucket_fine(int K) {
  static const int ub[] = {1, 3, 4, 6, 9, 12, 18, 27, 36, 54, 81};
  int f = 0;
  while (f < (int)(sizeof(ub) / sizeof(ub[0])) && K > ub[f]) ++f;
  return f;
}

int pow2floor(double x) {
  int g = 1;
  while (g * 2 <= x && g < 64) g *= 2;
  return g;
}

size_t tune_index(int regime, int cls, int bb, int kb) {
  return (((size_t)regime * mi::nclass() + cls) * MI_NBUCKET + bb) * MI_NBUCKET + kb;
}

// Check one tuning row; regime -1 applies it to both regimes.
bool tuning_row_ok(const mi_ctx* c, int cls, int bb, int kb, int var, int G, int regime) {
  if (cls < 0 || cls >= mi::nclass() || bb < 0 || bb >= MI_NBUCKET || kb < 0 || kb >= MI_NBUCKET) return false;
  if (regime < -1 || regime >= MI_NREGIME) return false;
  if (G < 1 || G > 64 || (G & (G - 1))) return false;
  if (var < 0 || var > 6 || var == 4) return false;
  if (var == 5 && (G > 4 || !mi::ryclass_fits_nw(cls, G, c->lds_limit))) return false;
  if (var == 6 && !mi::rygclass_fits(cls, G)) return false;
  if (var == 1 && !mi::has_register_kernel(cls)) return false;
  if (var == 0 && !mi::wclass_fits_nw(cls, G, c->lds_limit)) return false;
  if (var == 3 && !mi::wclass_fits_nw(cls, G, c->lds_limit, true)) return false;
  if (var == 2 && !mi::wgclass_fits_nw(cls, G, c->lds_limit)) return false;
  return true;
}

void set_tuning_row(mi_ctx* c, int cls, int bb, int kb, int var, int G, int regime) {
  for (int r = 0; r < MI_NREGIME; ++r)
    if (regime < 0 || regime == r) c->tune[tune_index(r, cls, bb, kb)] = {var, G};
}

// Tuning file: lines "class bra_bucket ket_bucket variant G [regime]" ('#'
// comments; regime 0 = class launches of fewer than MI_SMALL_LAUNCH shell
// quartets, 1 = larger ones, 2 = launches of J/K plans (MI_ORDER_JK), absent
// = all), written by mi_amd/autotune.py.
// Default location: hgp_tuning.txt next to the library; MI_HGP_TUNING
// overrides, MI_HGP_TUNING=none disables.
void load_tuning_file(mi_ctx* c) {
  std::string path;
  const char* env = std::getenv("MI_HGP_TUNING");
  if (env && *env) {
    if (std::string(env) == "none") return;
    path = env;
  } else {
    Dl_info info;
    if (dladdr((void*)&load_tuning_file, &info) && info.dli_fname) {
      std::string so = info.dli_fname;
      size_t k = so.rfind('/');
      path = (k == std::string::npos ? std::string(".") : so.substr(0, k)) + "/hgp_tuning.txt";
    }
  }
  std::ifstream f(path);
  std::string line;
  while (f && std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    int cls, bb, kb, var, G, regime = -1;
    if (!(is >> cls >> bb >> kb >> var >> G)) continue;
    if (!(is >> regime)) regime = -1;
    if (!tuning_row_ok(c, cls, bb, kb, var, G, regime)) continue;
    set_tuning_row(c, cls, bb, kb, var, G, regime);
  }
}

}  // namespace

namespace {

int validate(mi_ctx* ctx, const mi_basis* b) {
  if (!b) return set_err(ctx, MI_EINVAL, "basis is NULL");
  if (b->nshell <= 0) return set_err(ctx, MI_EINVAL, "basis has no shells");
  if (!b->shell_xyz || !b->shell_l || !b->shell_prim_off || !b->prim_exp || !b->prim_coef || !b->prim_norm ||
      !b->bf_scale)
    return set_err(ctx, MI_EINVAL, "basis has a NULL array");
  if (b->shell_prim_off[0] != 0) return set_err(ctx, MI_EINVAL, "shell_prim_off[0] must be 0");
  int nbf = 0;
  for (int s = 0; s < b->nshell; ++s) {
    const int L = b->shell_l[s];
    if (L < 0) return set_err(ctx, MI_EINVAL, "shell %d has negative L", s);
    if (L > mi::lmax_shell())
      return set_err(ctx, MI_EUNSUPPORTED, "shell %d has L=%d; this build supports L<=%d", s, L, mi::lmax_shell());
    const int np = b->shell_prim_off[s + 1] - b->shell_prim_off[s];
    if (np <= 0) return set_err(ctx, MI_EINVAL, "shell %d has no primitives", s);
    for (int p = b->shell_prim_off[s]; p < b->shell_prim_off[s + 1]; ++p)
      if (!(b->prim_exp[p] > 0.0) || !std::isfinite(b->prim_exp[p]))
        return set_err(ctx, MI_EINVAL, "primitive %d has a non-positive exponent", p);
    nbf += ncart(L);
  }
  if (nbf != b->nbf) return set_err(ctx, MI_EINVAL, "nbf=%d but the shells hold %d functions", b->nbf, nbf);
  if (nbf > 65535) return set_err(ctx, MI_EUNSUPPORTED, "nbf=%d: at most 65535 functions per molecule", nbf);
  return MI_OK;
}

// Device buffer of n elements, filled from src when given.  With a plan
// alloc_stream the allocation and copy are ordered on that stream (src is
// host data the plan keeps until it is destroyed).
template <typename T>
int upload(mi_plan* P, T** dst, const T* src, size_t n) {
  mi_ctx* ctx = P->ctx;
  if (n == 0) n = 1;
  const size_t bytes = sizeof(T) * n;
  if (P->arena) {
    const size_t at = (P->arena_used + 255) & ~(size_t)255;
    if (at + bytes > P->arena_cap)
      return set_err(ctx, MI_ENOMEM, "the basis needs more device memory than the %zu-byte static scratch area of "
                     "the allocation-free entry point", P->arena_cap);
    *dst = reinterpret_cast<T*>(P->arena + at);
    P->arena_used = at + bytes;
    if (src) HIP_TRY(ctx, hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, P->arena_stream));
    P->device_bytes += bytes;
    return MI_OK;
  }
  if (P->alloc_stream) {
    HIP_TRY(ctx, hipMallocAsync((void**)dst, bytes, P->alloc_stream));
    if (src) HIP_TRY(ctx, hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, P->alloc_stream));
  } else {
    HIP_TRY(ctx, hipMalloc((void**)dst, bytes));
    if (src) HIP_TRY(ctx, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
  }
  P->device_bytes += bytes;
  return MI_OK;
}

void free_plan_buffers(mi_plan* p) {
  if (p->gexec) hipGraphExecDestroy(p->gexec);
  if (p->ev_fork) hipEventDestroy(p->ev_fork);
  for (auto& e : p->ev_join)
    if (e) hipEventDestroy(e);
  if (p->arena) return;  // buffers live in the static scratch area
  void* bufs[] = {p->d_shell_xyz, p->d_shell_prim_off, p->d_prim_exp, p->d_prim_cn, p->d_bf_scale, p->d_topo,
                  p->d_geom,      p->d_pp,             p->d_pp_pair,  p->d_tasks,  p->d_scratch,   p->d_bounds,
                  p->d_block_task, p->d_pp_ab, p->d_run_ctr};
  for (void* b : bufs) {
    if (!b) continue;
    if (p->alloc_stream)
      hipFreeAsync(b, p->alloc_stream);
    else
      hipFree(b);
  }
}

}  // namespace

extern "C" {

int32_t mi_lmax(void) { return mi::lmax_shell(); }

int mi_rys_rule(int32_t n, double T, double* x, double* w) {
  if (n < 1 || n > mi::rys_nmax() || !(T >= 0.0) || !x || !w) return MI_EINVAL;
  mi::rys_rule_host(n, T, x, w);
  return MI_OK;
}

int32_t mi_rys_nmax(void) { return mi::rys_nmax(); }

int64_t mi_n_unique(int32_t nbf) {
  const int64_t n = nbf;
  return n * (n + 1) * (n * n + n + 2) / 8;
}

const char* mi_last_error(const mi_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// Static scratch area of mi_eri_unique_noalloc: every device the library
// runs on holds its own copy for the life of the process.  Host code never
// uses the name `scratch` as an address (on the host it is a stand-in): each
// context reads the device address once, with its device current, through
// hipGetSymbolAddress into ctx->noalloc_scratch.
#define MI_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
static __device__ __align__(256) char scratch[MI_NOALLOC_SCRATCH_BYTES];

int mi_create(int device, mi_ctx** out) {
  if (!out) return MI_EINVAL;
  *out = nullptr;
  mi_ctx* c = new mi_ctx();
  c->device = device;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0 || device < 0 || device >= ndev) {
    delete c;
    return MI_EDEVICE;
  }
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return MI_EDEVICE;
  }
  // device address of this device's static scratch area (null if unavailable:
  // then only mi_eri_unique_noalloc fails)
  if (hipGetSymbolAddress((void**)&c->noalloc_scratch, HIP_SYMBOL(scratch)) != hipSuccess) {
    c->noalloc_scratch = nullptr;
    (void)hipGetLastError();
  }
  std::vector<double> tab = make_boys_table();
  if (hipMalloc(&c->d_boys, sizeof(double) * tab.size()) != hipSuccess ||
      hipMemcpy(c->d_boys, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess ||
      mi::upload_wtables(&c->wtables) != hipSuccess) {
    hipFree(c->d_boys);
    delete c;
    return MI_EDEVICE;
  }
  for (int k = 0; k < MI_MAX_STREAMS - 1; ++k)
    if (hipStreamCreateWithFlags(&c->aux[k], hipStreamNonBlocking) != hipSuccess) {
      c->nstreams = k + 1;
      break;
    }
  if (hipStreamCreateWithFlags(&c->bstream, hipStreamNonBlocking) != hipSuccess) {
    for (auto& st : c->aux)
      if (st) hipStreamDestroy(st);
    hipFree(c->d_boys);
    mi::free_wtables(c->wtables);
    delete c;
    return MI_EDEVICE;
  }
  hipDeviceProp_t prop;
  c->lds_limit = 65536;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->ncu = prop.multiProcessorCount;
    c->lds_limit = std::max<size_t>(prop.sharedMemPerBlock, prop.maxSharedMemoryPerMultiProcessor);
  }
  c->lds_limit = std::min<size_t>(c->lds_limit, 160 * 1024);  // CDNA4: 160 KiB LDS per CU
  if (c->lds_limit > 65536 && mi::set_wave_kernel_lds_limit(c->lds_limit) != hipSuccess) {
    (void)hipGetLastError();  // not sticky: later launches check hipGetLastError
    c->lds_limit = 65536;
  }
  c->tune.assign((size_t)MI_NREGIME * mi::nclass() * MI_NBUCKET * MI_NBUCKET, TuneEntry());
  load_tuning_file(c);
  *out = c;
  return MI_OK;
}

void mi_destroy(mi_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  for (auto& st : ctx->aux)
    if (st) hipStreamDestroy(st);
  if (ctx->bstream) hipStreamDestroy(ctx->bstream);
  hipFree(ctx->d_boys);
  mi::free_wtables(ctx->wtables);
  delete ctx;
}

}  // extern "C"

// Plan over one or more molecules: shells are concatenated, shell pairs and
// quartets stay inside a molecule, and every task carries its molecule's
// output offset, pair count and first function, so the slot algebra of the
// kernels is per molecule.
// Kernel choice for a task from the tuning table, else the heuristic:
// register kernel when one exists with lanes ~ primitive quartets / 4, else the
// wave kernel with as many wavefronts (<= 4) as the LDS budget allows, else the
// global-table kernel with 4 wavefronts per block.
// regime: 0/1 by the size of the class launch; J/K plans use the J/K regime
// (2) where it has an entry, else the size regime.
// rys: the plan's method is Rys quadrature -- every class in the Rys kernel
// (variant 5) with 1, 2 or 4 wavefronts per block by the size of its
// contracted block, unless the tuning table names a Rys choice for it
// (variant 5 or the lane-group kernel, variant 6).
static int choose_kernel(mi_ctx* ctx, TaskDesc& t, int bb, int kb, double kavg, int regime, bool jk, bool rys) {
  const TuneEntry* tj = jk ? &ctx->tune[tune_index(2, t.cls, bb, kb)] : nullptr;
  const TuneEntry& te = (tj && tj->variant >= 0) ? *tj : ctx->tune[tune_index(regime, t.cls, bb, kb)];
  int variant, G;
  if (rys && te.variant == 6) {  // the tuning table's Rys lane-group choice
    variant = 6;
    G = te.G;
  } else if (rys) {
    const int nacc = mi::wclass(t.cls).nacc;
    variant = 5;
    G = te.variant == 5 ? te.G : (nacc <= 256 ? 1 : (nacc <= 1024 ? 2 : 4));
    while (G < 4 && !mi::ryclass_fits_nw(t.cls, G, ctx->lds_limit)) G <<= 1;
    if (!mi::ryclass_fits_nw(t.cls, G, ctx->lds_limit))
      return set_err(ctx, MI_EUNSUPPORTED, "class %d has no Rys kernel in this build", t.cls);
  } else if (te.variant >= 0) {
    variant = te.variant;
    G = te.G;
  } else {
    variant = mi::has_register_kernel(t.cls) ? 1 : 0;
    if (variant) {
      G = pow2floor(std::max(1.0, kavg / 4.0));
      if (mi::wclass(t.cls).nacc > 16) G = std::min(G, 4);
    } else {
      G = 4;
      while (G > 1 && !mi::wclass_fits_nw(t.cls, G, ctx->lds_limit)) G >>= 1;
      if (!mi::wclass_fits_nw(t.cls, G, ctx->lds_limit)) {
        variant = 2;
        G = 4;
      }
    }
  }
  if ((variant == 0 && !mi::wclass_fits_nw(t.cls, G, ctx->lds_limit)) ||
      (variant == 3 && !mi::wclass_fits_nw(t.cls, G, ctx->lds_limit, true)) ||
      (variant == 2 && !mi::wgclass_fits_nw(t.cls, G, ctx->lds_limit)) ||
      (variant == 5 && !mi::ryclass_fits_nw(t.cls, G, ctx->lds_limit)) ||
      (variant == 6 && !mi::rygclass_fits(t.cls, G)))
    return set_err(ctx, MI_EUNSUPPORTED, "class %d has no kernel in this build", t.cls);
  t.G = G;
  t.variant = variant;
  return MI_OK;
}

// One launch per (class, kernel variant, wavefronts); tasks appended to
// all_tasks, launches get task ranges, grid sizes and a cost estimate.
static int group_launches(mi_ctx* ctx, mi_plan* P, std::vector<std::vector<TaskDesc>>& per_class,
                          std::vector<TaskDesc>& all_tasks, std::vector<Launch>& launches) {
  // One launch per (class, variant); within a launch, tasks with the most
  // work per block come first so the longest blocks start earliest.
  for (int c = 0; c < mi::nclass(); ++c) {
    // vk 0: register kernel (each task its own G); 1-3: wave kernel with
    // 1, 2, 4 wavefronts; 4-7: global-table kernel with 1, 2, 4, 8
    // wavefronts; 8-10: wave kernel with global op tables, 1, 2, 4 wavefronts;
    // 11: double-double table kernel (1 wavefront); 12-14: Rys kernel with
    // 1, 2, 4 wavefronts; 15-17: Rys lane-group kernel with 16, 32, 64 lanes
    // per group
    for (int vk = 0; vk < 18; ++vk) {
      const int var = vk == 0 ? 1 : (vk < 4 ? 0 : (vk < 8 ? 2 : (vk < 11 ? 3 : (vk < 12 ? 4 : (vk < 15 ? 5 : 6)))));
      const int nw = vk == 0 ? 0
                             : (vk < 4 ? (1 << (vk - 1))
                                       : (vk < 8 ? (1 << (vk - 4))
                                                 : (vk < 11 ? (1 << (vk - 8))
                                                            : (vk < 12 ? 1 : (vk < 15 ? (1 << (vk - 12)) : (16 << (vk - 15)))))));
      std::vector<TaskDesc> ts;
      for (const TaskDesc& t : per_class[c])
        if (t.variant == var && (var == 1 || t.G == nw)) ts.push_back(t);
      if (ts.empty()) continue;
      Launch L{};
      L.cls = c;
      L.var = var;
      L.nw = nw;
      const bool reg = var == 1;
      L.task_off = (int)all_tasks.size();
      L.ntasks = (int)ts.size();
      int64_t blocks = 0;
      int64_t qs = 0;
      int64_t waves = 0;
      for (TaskDesc& t : ts) {
        // register kernel: first wavefront of the task (a task takes
        // ceil(nq G / 64) whole wavefronts); wave kernels: first quartet
        t.blockStart = (int32_t)(reg ? waves : qs);
        t.rnKet = 1.0 / (double)std::max(1, t.nKet);
        waves += reg ? (t.nq * t.G + 63) / 64 : 0;
        qs += t.nq;
        all_tasks.push_back(t);
      }
      if (reg) {
        // block -> task of its first wavefront (one load in the kernel
        // instead of a search)
        // one-wavefront blocks where the class reads the Boys grid through
        // L1/L2 (each wavefront frees its slot when it ends), except in J/K
        // plans: their launches run beside other requests' in the serving
        // engine, and four-wavefront blocks leave those room between blocks
        const int wpb = P->order == MI_ORDER_JK ? 4 : mi::register_waves_per_block(c);
        L.wpb = wpb;
        blocks = (waves + wpb - 1) / wpb;
        L.bt_off = (int64_t)P->h.block_task.size();
        int ti = 0;
        for (int64_t b = 0; b < blocks; ++b) {
          const int64_t w0 = wpb * b;
          while (ti + 1 < L.ntasks && all_tasks[L.task_off + ti + 1].blockStart <= w0) ++ti;
          P->h.block_task.push_back(ti);
        }
      }
      if (var == 0 || var == 3)
        blocks = std::min<int64_t>(qs, (int64_t)ctx->ncu * mi::wave_kernel_blocks_per_cu(c, nw, var == 3));
      if (var == 5) blocks = std::min<int64_t>(qs, (int64_t)ctx->ncu * mi::ry_kernel_blocks_per_cu(c, nw));
      if (var == 6)  // nw = lanes per group; 64 / nw groups per one-wavefront block
        blocks = std::min<int64_t>((qs * nw + 63) / 64, (int64_t)ctx->ncu * mi::ryg_kernel_blocks_per_cu(c, nw));
      // double-double kernel: a few persistent blocks (its quartets are rare,
      // and each block holds a 16-byte-per-entry work area in the scratch)
      if (var == 4) {
        const int64_t per = (int64_t)(mi::wx_scratch_doubles(c, 1) * sizeof(double));
        blocks = std::min<int64_t>(qs, std::max<int64_t>(64, std::min<int64_t>(16384, (int64_t)MI_WX_SCRATCH / per)));
      }
      L.tl = false;
      if (var == 2) {
        // tables in LDS only where that costs no resident blocks: a launch
        // smaller than the chip's capacity at the larger LDS footprint
        blocks = std::min<int64_t>(qs, (int64_t)ctx->ncu * mi::wg_kernel_blocks_per_cu(c, nw, false));
        L.tl = mi::wg_tables_fit(c) && blocks <= (int64_t)ctx->ncu * mi::wg_kernel_blocks_per_cu(c, nw, true);
      }
      if (blocks > INT32_MAX || qs > INT32_MAX || waves > INT32_MAX) {
        return set_err(ctx, MI_ENOMEM, "class %d: %lld quartets exceed one launch", c, (long long)qs);
      }
      L.nblocks = (int)blocks;
      L.nq_shell = qs;
      L.nq_prim = 0;
      for (const TaskDesc& t : ts) L.nq_prim += t.nprim;
      L.cost = 0.0;
      L.scratch_off = 0;
      launches.push_back(L);
    }
  }
  for (Launch& L : launches) {
    const WClassDesc& D = mi::wclass(L.cls);
    L.cost = (double)L.nq_prim * (D.nvrr + D.nacc) + (double)L.nq_shell * (D.nhrr + D.nout) +
             2.0e4 * (D.nvlev + D.nhlev + 1);
    if (L.var == 4) L.cost *= 12.0;  // double-double arithmetic: ~12 FP64 operations per operation
  }
  (void)P;
  return MI_OK;
}

// Issue order and stream lanes: launches by cost, longest first, each to the
// least-loaded lane (LPT).
static void assign_lanes(mi_plan* P) {
  std::stable_sort(P->launches.begin(), P->launches.end(),
                   [](const Launch& x, const Launch& y) { return x.cost > y.cost; });
  std::vector<double> load(P->nlanes, 0.0);
  for (Launch& L : P->launches) {
    int best = 0;
    for (int k = 1; k < P->nlanes; ++k)
      if (load[k] < load[best]) best = k;
    L.lane = best;
    load[best] += L.cost;
  }
}

// Units: one per class launch, except that (with fusion on, for plans that
// write the unique array) the small register-kernel launches of each fusion
// group (similar VGPR use) become one fused launch, and so do the small
// LDS-wave-kernel launches of each wavefront count and the small
// global-table launches of each wavefront count; groups hold at most
// MI_FUSE_MAX launches, a group of one stays a plain launch.  Launch indices
// refer to the current order of P->launches.
static void build_units(mi_plan* P) {
  mi_ctx* ctx = P->ctx;
  P->units.clear();
  const int n = (int)P->launches.size();
  std::vector<char> used(n, 0);
  auto add_fused = [&](std::vector<int> idx, int kind, int nw, int acc) {
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int x, int y) { return P->launches[x].cost > P->launches[y].cost; });
    for (size_t g0 = 0; g0 < idx.size(); g0 += MI_FUSE_MAX) {
      const size_t g1 = std::min(idx.size(), g0 + MI_FUSE_MAX);
      if (g1 - g0 < 2) break;
      Unit U;
      U.kind = kind;
      U.nw = nw;
      U.acc = acc;
      int64_t blocks = 0;
      double cmax = 0.0, csum = 0.0;
      U.segs.nseg = (int32_t)(g1 - g0);
      for (size_t k = g0; k < g1; ++k) {
        const Launch& L = P->launches[idx[k]];
        const int s = (int)(k - g0);
        U.members.push_back(idx[k]);
        U.segs.bstart[s] = (int32_t)blocks;
        U.segs.cls[s] = L.cls;
        U.segs.task_off[s] = L.task_off;
        U.segs.ntasks[s] = L.ntasks;
        // kind 1: block -> task table; kind 3: the launch's slice of the HRR scratch
        U.segs.bt_off[s] = kind == 1 ? L.bt_off : (kind == 3 ? (int64_t)L.scratch_off : 0);
        U.segs.nq[s] = L.nq_shell;
        if (kind == 2) U.lds = std::max(U.lds, mi::wclass_lds_bytes(L.cls, nw));
        if (kind == 3) U.lds = std::max(U.lds, mi::wgclass_lds_bytes(L.cls, false));
        blocks += L.nblocks;
        cmax = std::max(cmax, L.cost);
        csum += L.cost;
        used[idx[k]] = 1;
      }
      if (blocks > INT32_MAX) {  // cannot happen for small launches; keep them apart
        for (int m : U.members) used[m] = 0;
        continue;
      }
      U.segs.bstart[U.segs.nseg] = (int32_t)blocks;
      U.nblocks = (int)blocks;
      // members run side by side: about the longest, plus a share of the rest
      U.cost = cmax + 0.25 * (csum - cmax);
      P->units.push_back(U);
    }
  };
  if (ctx->fuse && !P->bounds_only && P->order != MI_ORDER_JK) {
    // register kernels by fusion group (similar VGPR use); wave kernels by
    // wavefronts per block, in the instantiation of the group's largest
    // accumulator count (a few more VGPRs for the smaller classes)
    std::vector<std::vector<int>> reg(8);
    std::vector<std::vector<int>> wave(5), glob(9);
    std::vector<int> wacc(5, 2);
    for (int i = 0; i < n; ++i) {
      const Launch& L = P->launches[i];
      if (L.nq_shell >= MI_SMALL_LAUNCH) continue;
      const int g = L.var == 1 ? mi::register_group(L.cls) : -1;
      if (g >= 0 && g < 8) reg[g].push_back(i);
      if (L.var == 0 && L.nw >= 1 && L.nw <= 4) {
        wave[L.nw].push_back(i);
        wacc[L.nw] = std::max(wacc[L.nw], mi::wave_kernel_acc(L.cls));
      }
      // global-table launches by wavefronts per block
      if (L.var == 2 && L.nw >= 1 && L.nw <= 8) glob[L.nw].push_back(i);
    }
    for (int g = 0; g < 8; ++g) add_fused(reg[g], 1, 4, g);
    for (int nw = 1; nw <= 4; ++nw) add_fused(wave[nw], 2, nw, wacc[nw]);
    for (int nw = 1; nw <= 8; nw *= 2) add_fused(glob[nw], 3, nw, 0);
  }
  for (int i = 0; i < n; ++i)
    if (!used[i]) {
      Unit U;
      U.members.push_back(i);
      U.nblocks = P->launches[i].nblocks;
      U.nw = P->launches[i].nw;
      U.cost = P->launches[i].cost;
      P->units.push_back(U);
    }
}

// Units by cost, longest first, each to the least-loaded lane (LPT).
static void assign_unit_lanes(mi_plan* P) {
  std::stable_sort(P->units.begin(), P->units.end(), [](const Unit& x, const Unit& y) { return x.cost > y.cost; });
  std::vector<double> load(P->nlanes, 0.0);
  for (Unit& U : P->units) {
    int best = 0;
    for (int k = 1; k < P->nlanes; ++k)
      if (load[k] < load[best]) best = k;
    U.lane = best;
    load[best] += U.cost;
  }
}

int mi_plan_upload(mi_plan* P, hipStream_t alloc_stream) {
  int rc;
  const mi_plan::HostData& h = P->h;
  P->alloc_stream = alloc_stream;
  P->device_bytes = 0;
  if ((rc = upload(P, &P->d_shell_xyz, h.xyz.data(), h.xyz.size())) ||
      (rc = upload(P, &P->d_shell_prim_off, h.poff.data(), h.poff.size())) ||
      (rc = upload(P, &P->d_prim_exp, h.pexp.data(), h.pexp.size())) ||
      (rc = upload(P, &P->d_prim_cn, h.pcn.data(), h.pcn.size())) ||
      (rc = upload(P, &P->d_bf_scale, h.bfs.data(), h.bfs.size())) ||
      (rc = upload(P, &P->d_topo, h.topo.data(), h.topo.size())) ||
      (rc = upload<PairGeom>(P, &P->d_geom, nullptr, (size_t)P->npair)) ||
      (rc = upload<PrimPair>(P, &P->d_pp, nullptr, (size_t)P->npp)) ||
      (rc = upload(P, &P->d_pp_pair, h.pp_pair.data(), h.pp_pair.size())) ||
      (rc = upload(P, &P->d_pp_ab, h.pp_ab.data(), h.pp_ab.size())) ||
      (rc = upload(P, &P->d_tasks, h.tasks.data(), h.tasks.size())) ||
      (P->scratch_doubles && (rc = upload<double>(P, &P->d_scratch, nullptr, P->scratch_doubles))) ||
      (!h.bounds.empty() && (rc = upload(P, &P->d_bounds, h.bounds.data(), h.bounds.size()))) ||
      (!h.block_task.empty() && (rc = upload(P, &P->d_block_task, h.block_task.data(), h.block_task.size()))) ||
      (rc = upload<unsigned long long>(P, &P->d_run_ctr, nullptr, P->launches.size())))
    return rc;
  return MI_OK;
}

int mi_plan_set_mol_outputs(mi_plan* P, double* const* outs) {
  if (!P || !outs) return MI_EINVAL;
  if (P->d_tasks) return set_err(P->ctx, MI_EINVAL, "plan already uploaded");
  P->mol_out.assign(outs, outs + P->nmol);
  const double* base = outs[0];
  for (TaskDesc& t : P->h.tasks) {
    // the molecule of a task: its output offset is that molecule's mol_off
    const int m = (int)(std::upper_bound(P->mol_off.begin(), P->mol_off.end(), t.obase) - P->mol_off.begin()) - 1;
    if (m < 0 || m >= P->nmol || P->mol_off[m] != t.obase) return set_err(P->ctx, MI_EINVAL, "task outside the plan");
    t.obase = (int64_t)(outs[m] - base);
  }
  return MI_OK;
}

// Schwarz bounds of every shell pair, in the canonical pair order
// (molecule, pair class, list bucket, [first shell: J/K layout,]
// pair index) of a plan built with the same options: a bounds-only plan
// (diagonal quartets) runs on the context's bounds stream with stream-ordered
// memory and is read back.  Q = max |(ab|ab)|, V = max |(e0|e0)| (VRR outputs).
// A shell pair of a plan being built, before its topology is laid out.
struct PairKey {
  int mol, pc, bucket, pidx, A, B, K;
  int list;     // list split key: kbucket_fine (unique-array plans) or kbucket (J/K plans)
  double Q, V;
  int64_t ab0;  // first kept (ia, ib) code in ab_codes
  int ext;      // HRR amplification above amp_limit: quartets run in double-double
};
// Pairs handed from a build to its bounds-only build (same bases and options,
// so the same pairs in the same list order; the bounds-only build consumes them).
struct SharedPairs {
  std::vector<PairKey> keys;
  std::vector<int32_t> ab_codes;
};
static int plan_build(mi_ctx* ctx, const mi_basis* bases, int nmol, const mi_opts* opts, int flags,
                      SharedPairs* shared, mi_plan** out);

static int compute_pair_bounds(mi_ctx* ctx, const mi_basis* bases, int nmol, const mi_opts* opts,
                               SharedPairs* shared, std::vector<double>& Q, std::vector<double>& V) {
  mi_plan* B = nullptr;
  int rc = plan_build(ctx, bases, nmol, opts, MI_BUILD_BOUNDS_ONLY | MI_BUILD_DEFER_UPLOAD, shared, &B);
  if (rc) return rc;
  hipStream_t st = ctx->bstream;
  rc = mi_plan_upload(B, st);
  if (rc == MI_OK) rc = mi_plan_run(B, nullptr, 0, st, nullptr, nullptr);
  std::vector<PairGeom> g(B->npair);
  if (rc == MI_OK) {
    hipError_t e = hipMemcpyAsync(g.data(), B->d_geom, sizeof(PairGeom) * g.size(), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = set_err(ctx, MI_EDEVICE, "Schwarz bounds readback: %s", hipGetErrorString(e));
  }
  if (rc == MI_OK) {
    Q.resize(g.size());
    V.resize(g.size());
    for (size_t r = 0; r < g.size(); ++r) {
      Q[r] = g[r].Q;
      V[r] = g[r].V;
    }
  }
  mi_plan_destroy(B);
  return rc;
}

int mi_plan_build(mi_ctx* ctx, const mi_basis* bases, int nmol, const mi_opts* opts, int flags, mi_plan** out) {
  return plan_build(ctx, bases, nmol, opts, flags, nullptr, out);
}

static int plan_build(mi_ctx* ctx, const mi_basis* bases, int nmol, const mi_opts* opts, int flags,
                      SharedPairs* shared, mi_plan** out) {
  if (!ctx || !out || !bases || nmol <= 0) return set_err(ctx, MI_EINVAL, "bad plan arguments");
  *out = nullptr;
  int rc;
  for (int m = 0; m < nmol; ++m)
    if ((rc = validate(ctx, &bases[m]))) return rc;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const bool bounds_only = (flags & MI_BUILD_BOUNDS_ONLY) != 0;
  const double tau2 = (opts && opts->precision == MI_PREC_MIXED) ? opts->tau * opts->tau : 0.0;
  const double cut2 = (opts && opts->schwarz_cut > 0) ? opts->schwarz_cut * opts->schwarz_cut : 0.0;
  const int method = opts ? opts->method : MI_METHOD_HGP;
  if (method != MI_METHOD_HGP && method != MI_METHOD_RYS) return set_err(ctx, MI_EINVAL, "unknown method %d", method);
  const bool rys = method == MI_METHOD_RYS;
  mi_plan* P = new mi_plan();
  P->ctx = ctx;
  P->bounds_only = bounds_only;
  P->order = opts ? opts->order : MI_ORDER_CHRR;
  if (P->order != MI_ORDER_CHRR && P->order != MI_ORDER_IINDEX && P->order != MI_ORDER_JK) {
    delete P;
    return set_err(ctx, MI_EINVAL, "unknown order %d", P->order);
  }
  // Concatenated basis.
  std::vector<double> xyz, pexp, pcn, bfs;
  std::vector<int32_t> L, poff(1, 0), shmol;
  std::vector<int> mol_sh0(nmol + 1, 0), mol_bf0(nmol + 1, 0);
  P->mol_off.assign(nmol + 1, 0);
  for (int m = 0; m < nmol; ++m) {
    const mi_basis& b = bases[m];
    mol_sh0[m] = (int)L.size();
    mol_bf0[m] = (int)bfs.size();
    const int pbase = poff.back();
    for (int s2 = 0; s2 < b.nshell; ++s2) {
      for (int k = 0; k < 3; ++k) xyz.push_back(b.shell_xyz[3 * s2 + k]);
      L.push_back(b.shell_l[s2]);
      shmol.push_back(m);
      poff.push_back(pbase + b.shell_prim_off[s2 + 1]);
    }
    for (int p2 = 0; p2 < b.shell_prim_off[b.nshell]; ++p2) {
      pexp.push_back(b.prim_exp[p2]);
      pcn.push_back(b.prim_coef[p2] * b.prim_norm[p2]);
    }
    for (int f = 0; f < b.nbf; ++f) bfs.push_back(b.bf_scale[f]);
    P->mol_off[m + 1] = P->mol_off[m] + mi_n_unique(b.nbf);
  }
  mol_sh0[nmol] = (int)L.size();
  mol_bf0[nmol] = (int)bfs.size();
  const int ns = (int)L.size();
  P->nshell = ns;
  P->nbf = nmol == 1 ? bases[0].nbf : (int)bfs.size();
  P->nmol = nmol;
  P->n_unique = P->mol_off[nmol];
  std::vector<int> bfoff(ns + 1, 0), np(ns);
  for (int s2 = 0; s2 < ns; ++s2) {
    bfoff[s2 + 1] = bfoff[s2] + ncart(L[s2]);
    np[s2] = poff[s2 + 1] - poff[s2];
  }
  // Shell pairs (I >= J within a molecule), oriented by orient_pair (L_A >= L_B
  // except for the swapped pair classes).
  // Primitive pairs whose prefactor |K_ab| is below MI_PRIM_PAIR_CUT times the
  // largest of their shell pair are dropped (22% of them are exactly zero --
  // exp(-ab|AB|^2/zeta) underflows -- and the rest contribute below FP64
  // resolution); the kept (ia, ib) of a pair are listed in pp_ab.
  std::vector<PairKey> keys;
  std::vector<int32_t> ab_codes;
  std::vector<double> kk;
  if (shared) {  // the main build's pairs, already in list order
    keys = std::move(shared->keys);
    ab_codes = std::move(shared->ab_codes);
  }
  for (int m = 0; m < nmol && !shared; ++m)
    for (int I = mol_sh0[m]; I < mol_sh0[m + 1]; ++I)
      for (int J = mol_sh0[m]; J <= I; ++J) {
        int A, B;
        orient_pair(xyz.data(), L, poff, pexp, I, J, A, B);
        double r2 = 0.0;
        for (int k = 0; k < 3; ++k) r2 += (xyz[3 * A + k] - xyz[3 * B + k]) * (xyz[3 * A + k] - xyz[3 * B + k]);
        kk.assign((size_t)np[A] * np[B], 0.0);
        double kmax = 0.0;
        for (int ia = 0; ia < np[A]; ++ia)
          for (int ib = 0; ib < np[B]; ++ib) {
            const double a = pexp[poff[A] + ia], b = pexp[poff[B] + ib];
            const double v = std::fabs(pcn[poff[A] + ia] * pcn[poff[B] + ib] * std::exp(-a * b / (a + b) * r2) / (a + b));
            kk[(size_t)ia * np[B] + ib] = v;
            kmax = std::max(kmax, v);
          }
        const int64_t ab0 = (int64_t)ab_codes.size();
        for (int ia = 0; ia < np[A]; ++ia)
          for (int ib = 0; ib < np[B]; ++ib) {
            const double v = kk[(size_t)ia * np[B] + ib];
            if (v > MI_PRIM_PAIR_CUT * kmax || v == kmax) ab_codes.push_back(ia | (ib << 16));
          }
        const int K = (int)((int64_t)ab_codes.size() - ab0);
        const int i = I - mol_sh0[m], j = J - mol_sh0[m];
        // amplification of the kept primitive pairs in the chosen orientation;
        // above amp_limit (contracted shells mixing tight and diffuse
        // primitives several bohr apart, whichever way they are oriented) the
        // pair's quartets run in the double-double table kernel (hgp_wx_kernel):
        // 2.5% of the quartets of random contracted-f test bases, none of the
        // seeded cluster or 512-molecule set
        double amp = 1.0;
        if (L[B] > 0) {
          const double r = std::sqrt(r2);
          for (int64_t t = ab0; t < (int64_t)ab_codes.size(); ++t) {
            const double a = pexp[poff[A] + (ab_codes[t] & 0xFFFF)], b = pexp[poff[B] + (ab_codes[t] >> 16)];
            const double z = a + b, sw = 1.0 / std::sqrt(2.0 * z);
            const double ratio = (b / z * r + sw) / (a / z * r + sw);
            double pw = ratio;
            for (int e = 1; e < L[B]; ++e) pw *= ratio;
            amp = std::max(amp, pw);
          }
        }
        keys.push_back({m, pair_class(L[A], L[B]), kbucket(K), i * (i + 1) / 2 + j, A, B, K,
                        P->order == MI_ORDER_JK ? kbucket(K) : kbucket_fine(K), 0.0, 0.0, ab0,
                        amp > amp_limit(L[B]) ? 1 : 0});
      }
  // J/K layout of a plan without a cut: within a (class, bucket) list, pairs
  // of equal first shell A are contiguous, so along a task row (one bra
  // pair) the ket shell C changes rarely and the K_ac targets repeat too
  // (J_ab and K_bc repeat along any row).  A screening plan orders each list
  // by bound instead (its rectangles leave the screened quartets out);
  // splitting its lists by first shell as well was measured: about 6% less
  // J/K time on a 920-function cluster for 20x the plan-build time.
  const bool jk_layout = P->order == MI_ORDER_JK && cut2 == 0.0;
  // Order (molecule, class, ext, list bucket, pidx): the keys come in
  // molecule order with pidx increasing, so a stable counting sort over
  // (class, ext, list) within each molecule gives it in linear time.
  {
    auto list_id = [](const PairKey& k) { return (k.pc * 2 + k.ext) * 16 + k.list; };  // list < 16
    int ng = 0;
    for (const PairKey& k : keys) ng = std::max(ng, list_id(k) + 1);
    std::vector<PairKey> sorted(keys.size());
    std::vector<int> cnt;
    for (size_t s = 0, e; s < keys.size(); s = e) {
      for (e = s; e < keys.size() && keys[e].mol == keys[s].mol;) ++e;
      cnt.assign(ng + 1, 0);
      for (size_t r = s; r < e; ++r) ++cnt[list_id(keys[r]) + 1];
      for (int g = 0; g < ng; ++g) cnt[g + 1] += cnt[g];
      for (size_t r = s; r < e; ++r) sorted[s + cnt[list_id(keys[r])]++] = keys[r];
    }
    keys.swap(sorted);
  }
  // fn over each (molecule, class, ext, list) run of keys
  auto for_each_list = [&keys](auto fn) {
    for (size_t s = 0, e; s < keys.size(); s = e) {
      const PairKey& k = keys[s];
      for (e = s + 1; e < keys.size() && keys[e].mol == k.mol && keys[e].pc == k.pc && keys[e].ext == k.ext &&
                      keys[e].list == k.list;)
        ++e;
      fn(keys.begin() + s, keys.begin() + e);
    }
  };
  if (jk_layout)  // then by first shell, pidx order kept within equal A
    for_each_list([](std::vector<PairKey>::iterator b, std::vector<PairKey>::iterator e) {
      std::stable_sort(b, e, [](const PairKey& x, const PairKey& y) { return x.A < y.A; });
    });
  // Bounds (a device pass) when the plan screens or mixes precision: pairs
  // are then ordered by bound and screened quartets never enter a task.
  std::vector<double> kQ, kV;
  if (!bounds_only && !(flags & MI_BUILD_NO_BOUNDS) && (tau2 > 0.0 || cut2 > 0.0)) {
    SharedPairs sp{keys, ab_codes};
    if ((rc = compute_pair_bounds(ctx, bases, nmol, opts, &sp, kQ, kV))) {
      delete P;
      return rc;
    }
  }
  const bool has_bounds = !kQ.empty();
  if (has_bounds) {
    if (kQ.size() != keys.size()) {
      delete P;
      return set_err(ctx, MI_EDEVICE, "Schwarz bounds: %zu pairs, plan has %zu", kQ.size(), keys.size());
    }
    for (size_t r = 0; r < keys.size(); ++r) {
      keys[r].Q = kQ[r];
      keys[r].V = kV[r];
    }
    // within each (molecule, class, bucket) list: largest bound first
    if (!jk_layout)
      for_each_list([](std::vector<PairKey>::iterator b, std::vector<PairKey>::iterator e) {
        std::stable_sort(b, e, [](const PairKey& x, const PairKey& y) { return x.Q > y.Q; });
      });
  }
  const int npair = (int)keys.size();
  std::vector<PairTopo> topo(npair);
  std::vector<int32_t> pp_pair, pp_ab;
  int64_t ppoff = 0;
  for (int r = 0; r < npair; ++r) {
    const PairKey& k = keys[r];
    topo[r] = {k.A, k.B, bfoff[k.A], bfoff[k.B], (int32_t)ppoff, k.K, 0, k.pidx};
    for (int t = 0; t < k.K; ++t) {
      pp_pair.push_back(r);
      pp_ab.push_back(ab_codes[k.ab0 + t]);
    }
    ppoff += k.K;
  }
  if (ppoff > INT32_MAX) {
    delete P;
    return set_err(ctx, MI_ENOMEM, "too many primitive pairs (%lld)", (long long)ppoff);
  }
  P->npair = npair;
  P->npp = (int)ppoff;
  P->pair_mol.resize(npair);
  for (int r = 0; r < npair; ++r) P->pair_mol[r] = keys[r].mol;
  // Contiguous lists of equal (molecule, pc, ext, list bucket).
  struct List {
    int start, n, pc, Ksum, Kmax, mol;
    double qmax;  // bound of the list's first pair (its largest, once ordered by bound)
    int ext;
  };
  std::vector<List> lists;
  std::vector<int> mol_list0(nmol + 1, 0);
  for (int r = 0; r < npair;) {
    int e = r;
    int Ksum = 0, Kmax = 0;
    while (e < npair && keys[e].mol == keys[r].mol && keys[e].pc == keys[r].pc && keys[e].ext == keys[r].ext &&
           keys[e].list == keys[r].list) {
      Ksum += keys[e].K;
      Kmax = std::max(Kmax, keys[e].K);
      ++e;
    }
    lists.push_back({r, e - r, keys[r].pc, Ksum, Kmax, keys[r].mol, keys[r].Q, keys[r].ext});
    r = e;
  }
  for (size_t x = 0, m = 0; m <= (size_t)nmol; ++m) {
    while (x < lists.size() && lists[x].mol < (int)m) ++x;
    mol_list0[m] = (int)x;
  }
  // Tasks grouped by quartet class.  With Schwarz bounds and a cut, a list
  // pair becomes rectangles (bra rows x leading ket columns) that leave the
  // screened quartets out: in a list ordered by bound, the ket columns kept
  // for bra row i are a prefix of length n_i, non-increasing in i.  Rows are
  // grouped while n_i stays within 80% of the group's first (the few extra
  // quartets are screened in the kernels).
  std::vector<std::vector<TaskDesc>> per_class(mi::nclass());
  std::vector<std::vector<double>> kavgs(mi::nclass());
  std::vector<std::vector<char>> kext(mi::nclass());  // task runs in double-double (an ext list)
  std::vector<int64_t> cls_nq(mi::nclass(), 0);
  int64_t nq_shell = 0, nq_prim = 0;
  std::vector<int64_t> Kpre(npair + 1, 0);  // prefix sums of primitive pairs in record order
  for (int r = 0; r < npair; ++r) Kpre[r + 1] = Kpre[r] + keys[r].K;
  const bool screen_tasks = has_bounds && cut2 > 0.0;
  for (int m = 0; m < nmol && !bounds_only; ++m)
  for (int x = mol_list0[m]; x < mol_list0[m + 1]; ++x)
    for (int y = x; y < mol_list0[m + 1]; ++y) {
      const List* bra = &lists[x];
      const List* ket = &lists[y];
      // lists ordered by bound: nothing of the pair survives the cut
      if (screen_tasks && bra->qmax * ket->qmax < cut2) continue;
      if (ket->pc > bra->pc) std::swap(bra, ket);
      const double kavg = ((double)bra->Ksum / bra->n) * ((double)ket->Ksum / ket->n);
      const int bb = keys[bra->start].bucket, kb = keys[ket->start].bucket;
      // one task: bra records [b0, b0+nb) x ket records [k0, k0+nk), or the
      // triangle of [b0, b0+nb) with itself (same)
      auto emit = [&](int b0, int nb, int k0, int nk, int same) -> int {
        TaskDesc t{};
        t.braStart = b0;
        t.nBra = nb;
        t.ketStart = k0;
        t.nKet = nk;
        t.same = same;
        t.cls = quartet_class(bra->pc, ket->pc);
        t.nq = same ? (int64_t)nb * (nb + 1) / 2 : (int64_t)nb * nk;
        t.bf0 = mol_bf0[m];
        t.obase = P->mol_off[m];
        const int64_t nbm = mol_bf0[m + 1] - mol_bf0[m];
        t.M = nbm * (nbm + 1) / 2;
        t.G = bb;  // kernel chosen below, once the class's quartet count is known
        t.variant = kb;
        kavgs[t.cls].push_back(kavg);
        kext[t.cls].push_back((char)(bra->ext || ket->ext));
        const int64_t sk = Kpre[b0 + nb] - Kpre[b0];
        if (same) {
          // sum over p <= q of K_p K_q = ((sum K)^2 + sum K^2) / 2
          int64_t sk2 = 0;
          for (int r = b0; r < b0 + nb; ++r) sk2 += (int64_t)keys[r].K * keys[r].K;
          t.nprim = (sk * sk + sk2) / 2;
        } else {
          t.nprim = sk * (Kpre[k0 + nk] - Kpre[k0]);
        }
        nq_shell += t.nq;
        nq_prim += t.nprim;
        per_class[t.cls].push_back(t);
        cls_nq[t.cls] += t.nq;
        return MI_OK;
      };
      if (!screen_tasks) {
        rc = emit(bra->start, bra->n, ket->start, ket->n, x == y);
      } else {
        // n_i: ket columns with Q_i Q_j >= cut2 (kets ordered by Q, largest
        // first).  Called with i non-decreasing, so n_i only falls: a binary
        // search for the first row, then a pointer moved down.
        int ncur = -1;
        auto kept = [&](int i) -> int {
          const double qi = keys[bra->start + i].Q;
          if (!(qi > 0.0)) return 0;
          const double need = cut2 / qi;
          if (ncur < 0) {
            int lo = 0, hi = ket->n;  // first j with Q_j < need
            while (lo < hi) {
              const int mid = (lo + hi) >> 1;
              if (keys[ket->start + mid].Q >= need) lo = mid + 1; else hi = mid;
            }
            ncur = lo;
          }
          while (ncur > 0 && keys[ket->start + ncur - 1].Q < need) --ncur;
          return ncur;
        };
        int row = 0;
        rc = MI_OK;
        if (x == y) {
          // rows whose kept columns include all j <= i form a triangle
          while (row < bra->n && kept(row) > row) ++row;
          if (row > 0) rc = emit(bra->start, row, bra->start, row, 1);
        }
        while (rc == MI_OK && row < bra->n) {
          const int n0 = kept(row);
          if (n0 == 0) break;  // every later row keeps fewer
          int end = row + 1;
          while (end < bra->n && end - row < 256 && kept(end) * 5 >= n0 * 4) ++end;
          rc = emit(bra->start + row, end - row, ket->start, n0, 0);
          row = end;
        }
      }
      if (rc) {
        delete P;
        return rc;
      }
    }
  // kernel per task: the tuning regime follows the size of the class launch
  for (int c = 0; c < mi::nclass(); ++c)
    for (size_t k = 0; k < per_class[c].size(); ++k) {
      TaskDesc& t = per_class[c][k];
      const int bb = t.G, kb = t.variant;
      if (kext[c][k]) {  // a pair of the quartet cancels beyond FP64: the double-double kernel
        t.variant = 4;
        t.G = 1;
        continue;
      }
      if ((rc = choose_kernel(ctx, t, bb, kb, kavgs[c][k], cls_nq[c] < MI_SMALL_LAUNCH ? 0 : 1,
                              P->order == MI_ORDER_JK, rys))) {
        delete P;
        return rc;
      }
    }
  P->nq_shell = nq_shell;
  P->nq_prim = nq_prim;
  std::vector<TaskDesc> all_tasks;
  if ((rc = group_launches(ctx, P, per_class, all_tasks, P->launches))) {
    delete P;
    return rc;
  }
  // Schwarz pass: diagonal quartets (p|p) of every shell pair, needed for
  // FP32-VRR selection (MI_PREC_MIXED) and for screening (schwarz_cut > 0)
  // when the bounds were not computed with the plan (or for them).
  P->tau2 = tau2;
  P->cut2 = cut2;
  P->need_schwarz = bounds_only || ((tau2 > 0.0 || cut2 > 0.0) && !has_bounds);
  if (has_bounds) {
    P->h.bounds.resize(npair);
    for (int r = 0; r < npair; ++r) P->h.bounds[r] = make_double2(keys[r].Q, keys[r].V);
  }
  if (P->need_schwarz) {
    std::vector<std::vector<TaskDesc>> diag(mi::nclass());
    for (int m = 0; m < nmol; ++m)
      for (int x = mol_list0[m]; x < mol_list0[m + 1]; ++x) {
        const List& li = lists[x];
        TaskDesc t{};
        t.braStart = t.ketStart = li.start;
        t.nBra = t.nKet = li.n;
        t.same = 2;
        t.cls = quartet_class(li.pc, li.pc);
        t.nq = li.n;
        t.bf0 = mol_bf0[m];
        t.obase = P->mol_off[m];
        const int64_t nbm = mol_bf0[m + 1] - mol_bf0[m];
        t.M = nbm * (nbm + 1) / 2;
        const int bk = keys[li.start].bucket;
        if (li.ext) {
          t.variant = 4;
          t.G = 1;
        } else if ((rc = choose_kernel(ctx, t, bk, bk, ((double)li.Ksum / li.n) * ((double)li.Ksum / li.n), 0,
                                       false, rys))) {
          delete P;
          return rc;
        }
        int64_t sk2 = 0;
        for (int r = li.start; r < li.start + li.n; ++r) sk2 += (int64_t)keys[r].K * keys[r].K;
        t.nprim = sk2;
        diag[t.cls].push_back(t);
      }
    if ((rc = group_launches(ctx, P, diag, all_tasks, P->slaunches))) {
      delete P;
      return rc;
    }
  }
  // primitive-quartet counts per launch and a cost estimate (VRR+contraction
  // work per primitive quartet, HRR+scatter per shell quartet, plus a latency
  // floor per block wave)
  for (Launch& L : P->launches) {
    const WClassDesc& D = mi::wclass(L.cls);
    L.cost = (double)L.nq_prim * (D.nvrr + D.nacc) + (double)L.nq_shell * (D.nhrr + D.nout) +
             2.0e4 * (D.nvlev + D.nhlev + 1);
  }
  // Scratch of the global-table launches: the main launches may run
  // concurrently (stream lanes), so each gets its own slice; the Schwarz pass
  // runs before them on the caller's stream and reuses the same memory.
  {
    // slices start on 256-byte boundaries (the double-double kernel reads 16-byte pairs)
    auto take = [](size_t& o, size_t n) {
      const size_t at = (o + 31) & ~(size_t)31;
      o = at + n;
      return at;
    };
    size_t off = 0, soff = 0;
    auto need = [](const Launch& L) -> size_t {
      if (L.var == 2) return mi::wg_scratch_doubles(L.cls, L.nblocks);
      if (L.var == 4) return mi::wx_scratch_doubles(L.cls, L.nblocks);
      if (L.var == 5) return mi::ry_scratch_doubles(L.cls, L.nw, L.nblocks);
      return 0;
    };
    for (Launch& L : P->launches)
      if (L.var == 2 || L.var == 4 || L.var == 5) L.scratch_off = take(off, need(L));
    for (Launch& L : P->slaunches)
      if (L.var == 2 || L.var == 4 || L.var == 5) L.scratch_off = take(soff, need(L));
    P->scratch_doubles = std::max(off, soff);
  }
  P->nlanes = (flags & MI_BUILD_ONE_LANE) ? 1 : std::max(1, std::min(ctx->nstreams, (int)P->launches.size()));
  assign_lanes(P);
  build_units(P);
  assign_unit_lanes(P);
  // Host images of the device buffers; uploaded now, or by the caller of a
  // deferred build (the serving engine admits a plan by its device bytes).
  {
    mi_plan::HostData& h = P->h;
    h.xyz = std::move(xyz);
    h.poff = std::move(poff);
    h.pexp = std::move(pexp);
    h.pcn = std::move(pcn);
    h.bfs = std::move(bfs);
    h.topo = std::move(topo);
    h.pp_pair = std::move(pp_pair);
    h.pp_ab = std::move(pp_ab);
    h.tasks = std::move(all_tasks);
    auto nb = [](size_t n, size_t sz) { return sz * std::max<size_t>(n, 1); };
    P->device_bytes = nb(h.xyz.size(), 8) + nb(h.poff.size(), 4) + nb(h.pexp.size(), 8) + nb(h.pcn.size(), 8) +
                      nb(h.bfs.size(), 8) + nb(h.topo.size(), sizeof(PairTopo)) + nb(npair, sizeof(PairGeom)) +
                      nb(h.block_task.size(), 4) +
                      nb(P->npp, sizeof(PrimPair)) + nb(h.pp_pair.size(), 4) + nb(h.pp_ab.size(), 4) +
                      nb(h.tasks.size(), sizeof(TaskDesc)) +
                      (P->scratch_doubles ? 8 * P->scratch_doubles : 0) + 16 * h.bounds.size();
  }
  if (!(flags & MI_BUILD_DEFER_UPLOAD) && (rc = mi_plan_upload(P, nullptr))) {
    free_plan_buffers(P);
    delete P;
    return rc;
  }
  if (P->nlanes > 1) {
    hipError_t e = hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming);
    for (int k = 1; e == hipSuccess && k < P->nlanes; ++k)
      e = hipEventCreateWithFlags(&P->ev_join[k], hipEventDisableTiming);
    if (e != hipSuccess) {
      free_plan_buffers(P);
      delete P;
      return set_err(ctx, MI_EDEVICE, "hipEventCreate failed: %s", hipGetErrorString(e));
    }
  }
  *out = P;
  return MI_OK;
}

extern "C" {

int mi_plan_create(mi_ctx* ctx, const mi_basis* b, const mi_opts* opts, mi_plan** out) {
  return mi_plan_build(ctx, b, 1, opts, 0, out);
}

int mi_batch_plan_create(mi_ctx* ctx, const mi_basis* bases, int32_t nmol, const mi_opts* opts, mi_plan** out) {
  return mi_plan_build(ctx, bases, nmol, opts, 0, out);
}

int64_t mi_plan_out_len(const mi_plan* P) { return P ? P->n_unique : -1; }

int mi_plan_mol_offsets(const mi_plan* P, int64_t* offsets) {
  if (!P || !offsets) return MI_EINVAL;
  for (int m = 0; m <= P->nmol; ++m) offsets[m] = P->mol_off[m];
  return MI_OK;
}




}  // extern "C"

// A fused unit: segment tables index from the plan's task and block -> task
// arrays.
static int launch_fused(mi_ctx* ctx, const Unit& U, EriArgs a, const mi_plan* P, hipStream_t ls) {
  a.tasks = P->d_tasks;
  a.ntasks = 0;
  a.block_task = P->d_block_task;
  a.run_ctr = nullptr;
  if (U.kind == 1)
    HIP_TRY(ctx, mi::launch_fused_register(U.acc, a, U.segs, U.nblocks, ls));
  else if (U.kind == 2)
    HIP_TRY(ctx, mi::launch_fused_wave(U.acc, U.nw, a, U.segs, U.lds, U.nblocks, ctx->wtables, ls));
  else
    HIP_TRY(ctx, mi::launch_fused_wave_global(U.nw, a, U.segs, U.lds, U.nblocks, ctx->wtables, P->d_scratch, ls));
  return MI_OK;
}

static int launch_one(mi_ctx* ctx, const Launch& L, EriArgs a, const mi_plan* P, hipStream_t ls) {
  a.tasks = P->d_tasks + L.task_off;
  a.ntasks = L.ntasks;
  a.block_task = L.var == 1 ? P->d_block_task + L.bt_off : nullptr;
  a.run_ctr = (a.mode == MI_MODE_JK && P->d_run_ctr) ? P->d_run_ctr + (&L - P->launches.data()) : nullptr;
  if (L.var == 1)
    HIP_TRY(ctx, mi::launch_class_register(L.cls, a, L.nblocks, L.wpb, ls));
  else if (L.var == 0 || L.var == 3)
    HIP_TRY(ctx, mi::launch_class_wave(L.cls, L.nw, L.var == 3, a, L.nblocks, L.nq_shell, ctx->wtables, ls));
  else if (L.var == 4)
    HIP_TRY(ctx, mi::launch_class_extended(L.cls, a, L.nblocks, L.nq_shell, ctx->wtables,
                                           P->d_scratch + L.scratch_off, ls));
  else if (L.var == 5)
    HIP_TRY(ctx, mi::launch_class_rys(L.cls, L.nw, a, L.nblocks, L.nq_shell, ctx->wtables,
                                      P->d_scratch + L.scratch_off, ls));
  else if (L.var == 6)
    HIP_TRY(ctx, mi::launch_class_ryg(L.cls, L.nw, a, L.nblocks, L.nq_shell, ctx->wtables, ls));
  else
    HIP_TRY(ctx, mi::launch_class_wave_global(L.cls, L.nw, L.tl, a, L.nblocks, L.nq_shell, ctx->wtables,
                                              P->d_scratch + L.scratch_off, ls));
  return MI_OK;
}

// One execute on `stream`: shell-pair kernel, Schwarz pass when the plan
// needs bounds, then the class launches (forked over the plan's stream lanes
// unless ev != NULL, which times each launch on one stream).  jk != NULL
// accumulates Coulomb/exchange instead of writing the unique array.
int mi_plan_run(mi_plan* P, double* out, int64_t out_len, void* stream, hipEvent_t* ev, const JKArgs* jk,
                bool unit_events) {
  if (!P) return MI_EINVAL;
  mi_ctx* ctx = P->ctx;
  if (!jk && !P->bounds_only) {
    if (P->order == MI_ORDER_JK)
      return set_err(ctx, MI_EINVAL, "plan laid out for Coulomb/exchange builds (MI_ORDER_JK): no ERI output");
    if (!out) return set_err(ctx, MI_EINVAL, "output pointer is NULL");
    if (out_len != P->n_unique)
      return set_err(ctx, MI_ESIZE, "output length %lld, basis needs %lld", (long long)out_len,
                     (long long)P->n_unique);
  }
  hipStream_t s = (hipStream_t)stream;
  if (ev) HIP_TRY(ctx, hipEventRecord(ev[0], s));
  HIP_TRY(ctx, mi::launch_pair_prims(P->npp, P->d_pp_pair, P->d_pp_ab, P->d_topo, P->d_shell_xyz, P->d_shell_prim_off,
                                     P->d_prim_exp, P->d_prim_cn, P->d_pp, P->d_geom, P->d_bounds, s));
  EriArgs a{};
  a.topo = P->d_topo;
  a.geom = P->d_geom;
  a.pp = P->d_pp;
  a.shell_xyz = P->d_shell_xyz;
  a.bf_scale = P->d_bf_scale;
  a.boys = ctx->d_boys;
  a.out = out;
  a.order = P->order;
  a.nbf = P->nbf;
  if (P->need_schwarz) {
    EriArgs sa = a;
    sa.mode = MI_MODE_SCHWARZ;
    for (const Launch& L : P->slaunches) {
      int rc = launch_one(ctx, L, sa, P, s);
      if (rc) return rc;
    }
  }
  if (ev && jk) HIP_TRY(ctx, hipEventRecord(ev[1], s));
  a.tau2 = P->tau2;
  a.cut2 = P->cut2;
  if (jk) {
    a.mode = MI_MODE_JK;
    a.D = jk->D;
    a.Jacc = jk->J;
    a.Kacc = jk->K;
    const int64_t nn = (int64_t)P->nbf * P->nbf;
    HIP_TRY(ctx, mi::launch_zero(jk->J, nn, s));
    HIP_TRY(ctx, mi::launch_zero(jk->K, nn, s));
    // run counters of the wave-kernel launches (a kernel: graph-capture safe;
    // the all-zero bit pattern of a double is integer zero)
    if (P->d_run_ctr && !P->launches.empty())
      HIP_TRY(ctx, mi::launch_zero(reinterpret_cast<double*>(P->d_run_ctr), (int64_t)P->launches.size(), s));
  } else if (P->cut2 > 0.0 && !P->bounds_only) {
    // screened quartets are never written: their integrals are zero
    if (P->mol_out.empty()) {
      HIP_TRY(ctx, mi::launch_zero(out, P->n_unique, s));
    } else {
      for (int m = 0; m < P->nmol; ++m)
        HIP_TRY(ctx, mi::launch_zero(P->mol_out[m], P->mol_off[m + 1] - P->mol_off[m], s));
    }
  }
  // unique-array profiles: the zero fill is timed with the shell-pair kernel
  if (ev && !jk) HIP_TRY(ctx, hipEventRecord(ev[1], s));
  // Class launches are independent (disjoint output slots): fork them over the
  // plan's stream lanes and join back into the caller's stream.
  // Unique-array executes issue the plan's units (fused launches included);
  // J/K and per-launch profiles issue the class launches.
  const bool use_units = !jk && (!ev || unit_events) && !P->units.empty();
  const int nitem = (int)(use_units ? P->units.size() : P->launches.size());
  auto lane_of = [&](int i) { return use_units ? P->units[i].lane : P->launches[i].lane; };
  // Fork/join over the stream lanes.  Host submission costs a few
  // microseconds per call, so each lane waits for the fork just before its
  // first launch and records its join just after its last one (lanes without
  // launches do neither): the first kernels start as early as possible.
  const bool fork = !ev && P->nlanes > 1;
  int first[MI_MAX_STREAMS], last[MI_MAX_STREAMS];
  for (int k = 0; k < MI_MAX_STREAMS; ++k) first[k] = last[k] = -1;
  bool any_aux = false;
  if (fork)
    for (int i = 0; i < nitem; ++i) {
      const int k = lane_of(i);
      if (first[k] < 0) first[k] = i;
      last[k] = i;
      any_aux = any_aux || k > 0;
    }
  if (any_aux) HIP_TRY(ctx, hipEventRecord(P->ev_fork, s));
  for (int i = 0; i < nitem; ++i) {
    const int k = fork ? lane_of(i) : 0;
    hipStream_t ls = k > 0 ? ctx->aux[k - 1] : s;
    if (k > 0 && first[k] == i) HIP_TRY(ctx, hipStreamWaitEvent(ls, P->ev_fork, 0));
    int rc;
    if (use_units) {
      const Unit& U = P->units[i];
      rc = U.kind == 0 ? launch_one(ctx, P->launches[U.members[0]], a, P, ls) : launch_fused(ctx, U, a, P, ls);
    } else {
      rc = launch_one(ctx, P->launches[i], a, P, ls);
    }
    if (rc) return rc;
    if (ev) HIP_TRY(ctx, hipEventRecord(ev[i + 2], s));
    if (k > 0 && last[k] == i) HIP_TRY(ctx, hipEventRecord(P->ev_join[k], ls));
  }
  for (int k = 1; fork && k < P->nlanes; ++k)
    if (last[k] >= 0) HIP_TRY(ctx, hipStreamWaitEvent(s, P->ev_join[k], 0));
  if (jk) HIP_TRY(ctx, mi::launch_symmetrize(jk->J, jk->K, P->nbf, s));
  return MI_OK;
}

extern "C" {

int mi_plan_execute(mi_plan* P, double* out, int64_t out_len, void* stream) {
  return mi_plan_run(P, out, out_len, stream, nullptr, nullptr);
}

int mi_ctx_set_tuning(mi_ctx* ctx, const int32_t* rows, int32_t nrows) {
  if (!ctx || (nrows > 0 && !rows)) return MI_EINVAL;
  std::vector<TuneEntry> keep = ctx->tune;
  ctx->tune.assign((size_t)MI_NREGIME * mi::nclass() * MI_NBUCKET * MI_NBUCKET, TuneEntry());
  for (int r = 0; r < nrows; ++r) {
    const int32_t* q = rows + 6 * r;
    if (!tuning_row_ok(ctx, q[0], q[1], q[2], q[3], q[4], q[5])) {
      ctx->tune = keep;
      return set_err(ctx, MI_EINVAL, "tuning row %d (class %d buckets %d,%d variant %d G %d regime %d) unavailable", r,
                     q[0], q[1], q[2], q[3], q[4], q[5]);
    }
    set_tuning_row(ctx, q[0], q[1], q[2], q[3], q[4], q[5]);
  }
  return MI_OK;
}

int mi_ctx_set_streams(mi_ctx* ctx, int32_t nstreams) {
  if (!ctx || nstreams < 1) return MI_EINVAL;
  int avail = 1;
  while (avail < MI_MAX_STREAMS && ctx->aux[avail - 1]) ++avail;
  ctx->nstreams = std::min<int>(nstreams, avail);
  return MI_OK;
}

int mi_plan_graph_capture(mi_plan* P, double* out, int64_t out_len) {
  if (!P) return MI_EINVAL;
  mi_ctx* ctx = P->ctx;
  if (out_len != P->n_unique) return set_err(ctx, MI_ESIZE, "output length does not match the plan");
  hipStream_t cs = nullptr;
  HIP_TRY(ctx, hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  int rc = MI_OK;
  if (e == hipSuccess) {
    rc = mi_plan_run(P, out, out_len, cs, nullptr, nullptr);
    hipError_t e2 = hipStreamEndCapture(cs, &g);
    if (rc == MI_OK && e2 != hipSuccess) rc = set_err(ctx, MI_EDEVICE, "capture failed: %s", hipGetErrorString(e2));
  } else {
    rc = set_err(ctx, MI_EDEVICE, "hipStreamBeginCapture failed: %s", hipGetErrorString(e));
  }
  if (rc == MI_OK) {
    if (P->gexec) hipGraphExecDestroy(P->gexec);
    P->gexec = nullptr;
    e = hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = set_err(ctx, MI_EDEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  }
  if (g) hipGraphDestroy(g);
  hipStreamDestroy(cs);
  return rc;
}

int mi_plan_graph_launch(mi_plan* P, void* stream) {
  if (!P) return MI_EINVAL;
  if (!P->gexec) return set_err(P->ctx, MI_EINVAL, "no captured graph: call mi_plan_graph_capture first");
  HIP_TRY(P->ctx, hipGraphLaunch(P->gexec, (hipStream_t)stream));
  return MI_OK;
}

int mi_plan_jk_execute(mi_plan* P, const double* D, double* J, double* K, void* stream) {
  if (!P) return MI_EINVAL;
  if (P->nmol != 1) return set_err(P->ctx, MI_EINVAL, "Coulomb/exchange plans hold one molecule");
  if (!D || !J || !K) return set_err(P->ctx, MI_EINVAL, "NULL density or output");
  JKArgs jk{D, J, K};
  return mi_plan_run(P, nullptr, 0, stream, nullptr, &jk);
}

int mi_jk(mi_ctx* ctx, const mi_basis* b, const mi_opts* opts, const double* D, double* J, double* K) {
  if (!ctx) return MI_EINVAL;
  if (!b || !D || !J || !K) return set_err(ctx, MI_EINVAL, "NULL argument");
  mi_opts o{};
  if (opts) o = *opts;
  o.order = MI_ORDER_JK;
  mi_plan* P = nullptr;
  int rc = mi_plan_create(ctx, b, &o, &P);
  if (rc) return rc;
  const size_t nn = (size_t)b->nbf * b->nbf;
  double *dD = nullptr, *dJ = nullptr, *dK = nullptr;
  hipStream_t s = opts ? (hipStream_t)opts->stream : nullptr;
  hipError_t e = hipMalloc(&dD, sizeof(double) * nn);
  if (e == hipSuccess) e = hipMalloc(&dJ, sizeof(double) * nn);
  if (e == hipSuccess) e = hipMalloc(&dK, sizeof(double) * nn);
  if (e == hipSuccess) e = hipMemcpyAsync(dD, D, sizeof(double) * nn, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) rc = set_err(ctx, MI_ENOMEM, "device buffers: %s", hipGetErrorString(e));
  if (rc == MI_OK) rc = mi_plan_jk_execute(P, dD, dJ, dK, s);
  if (rc == MI_OK) {
    e = hipMemcpyAsync(J, dJ, sizeof(double) * nn, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(K, dK, sizeof(double) * nn, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = set_err(ctx, MI_EDEVICE, "copy back failed: %s", hipGetErrorString(e));
  }
  hipFree(dD);
  hipFree(dJ);
  hipFree(dK);
  mi_plan_destroy(P);
  return rc;
}

int mi_plan_pair_bounds(const mi_plan* P, double* Q, double* V, int32_t* mol) {
  if (!P || !Q || !V || !mol) return MI_EINVAL;
  if (!P->need_schwarz && !P->d_bounds)
    return set_err(P->ctx, MI_EINVAL, "plan has no Schwarz bounds (plain FP64 without screening)");
  if (!P->h.bounds.empty()) {  // computed when the plan was built (no device read)
    for (int r = 0; r < P->npair; ++r) {
      Q[r] = std::sqrt(P->h.bounds[r].x);
      V[r] = std::sqrt(P->h.bounds[r].y);
      mol[r] = P->pair_mol[r];
    }
    return MI_OK;
  }
  // filled by the Schwarz pass of the last execute, which may still run on any
  // stream (also a non-blocking one): wait for the device, then read
  std::vector<PairGeom> g(P->npair);
  HIP_TRY(P->ctx, hipSetDevice(P->ctx->device));
  HIP_TRY(P->ctx, hipDeviceSynchronize());
  HIP_TRY(P->ctx, hipMemcpy(g.data(), P->d_geom, sizeof(PairGeom) * g.size(), hipMemcpyDeviceToHost));
  for (int r = 0; r < P->npair; ++r) {
    Q[r] = std::sqrt(g[r].Q);
    V[r] = std::sqrt(g[r].V);
    mol[r] = P->pair_mol[r];
  }
  return MI_OK;
}

int mi_plan_nlaunch(const mi_plan* P) { return P ? (int)P->launches.size() : -1; }

int mi_plan_launch_info(const mi_plan* P, int32_t i, int32_t* cls4, int64_t* counts3, int32_t* kernel2) {
  if (!P || i < 0 || i >= (int)P->launches.size() || !cls4 || !counts3 || !kernel2) return MI_EINVAL;
  const Launch& L = P->launches[i];
  const WClassDesc& D = mi::wclass(L.cls);
  cls4[0] = D.la;
  cls4[1] = D.lb;
  cls4[2] = D.lc;
  cls4[3] = D.ld;
  counts3[0] = L.nq_shell;
  counts3[1] = L.nq_prim;
  counts3[2] = L.nblocks;
  kernel2[0] = L.var;
  kernel2[1] = L.nw;
  return MI_OK;
}

static int profile_run(mi_plan* P, double* out, int64_t out_len, void* stream, float* ms, const JKArgs* jk,
                       bool units = false) {
  if (!P || !ms) return MI_EINVAL;
  mi_ctx* ctx = P->ctx;
  const size_t nev = (units ? P->units.size() : P->launches.size()) + 2;
  std::vector<hipEvent_t> ev(nev, nullptr);
  for (auto& e : ev) HIP_TRY(ctx, hipEventCreate(&e));
  int rc = mi_plan_run(P, out, out_len, stream, ev.data(), jk, units);
  if (rc == MI_OK) {
    hipError_t e = hipEventSynchronize(ev[nev - 1]);
    if (e != hipSuccess) rc = set_err(ctx, MI_EDEVICE, "event sync failed: %s", hipGetErrorString(e));
    for (size_t i = 0; rc == MI_OK && i + 1 < nev; ++i) {
      e = hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]);
      if (e != hipSuccess) rc = set_err(ctx, MI_EDEVICE, "elapsed time failed: %s", hipGetErrorString(e));
    }
  }
  for (auto& e : ev) hipEventDestroy(e);
  return rc;
}

int mi_plan_profile(mi_plan* P, double* out, int64_t out_len, void* stream, float* ms) {
  return profile_run(P, out, out_len, stream, ms, nullptr);
}

int mi_plan_profile_units(mi_plan* P, double* out, int64_t out_len, void* stream, float* ms) {
  return profile_run(P, out, out_len, stream, ms, nullptr, true);
}

int mi_plan_nunit(const mi_plan* P) { return P ? (int)P->units.size() : -1; }

int mi_plan_unit_info(const mi_plan* P, int32_t u, int32_t* kind3, int32_t* members, int32_t cap) {
  if (!P || u < 0 || u >= (int)P->units.size() || !kind3 || (cap > 0 && !members)) return MI_EINVAL;
  const Unit& U = P->units[u];
  kind3[0] = U.kind;
  kind3[1] = U.nw;
  kind3[2] = (int32_t)U.members.size();
  for (int32_t k = 0; k < cap && k < (int32_t)U.members.size(); ++k) members[k] = U.members[k];
  return MI_OK;
}

int mi_ctx_set_fusion(mi_ctx* ctx, int32_t on) {
  if (!ctx) return MI_EINVAL;
  ctx->fuse = on != 0;
  return MI_OK;
}

int mi_plan_balance(mi_plan* P, double* out, int64_t out_len, void* stream) {
  if (!P) return MI_EINVAL;
  // Units are what a unique-array execute issues: time them and reassign
  // their lanes (class-launch lanes, used by J/K executes, stay as built).
  std::vector<float> ms(P->units.size() + 1);
  int rc = profile_run(P, out, out_len, stream, ms.data(), nullptr, true);
  if (rc) return rc;
  for (size_t i = 0; i < P->units.size(); ++i) P->units[i].cost = std::max(ms[i + 1], 1e-6f);
  assign_unit_lanes(P);
  if (P->gexec) {  // the captured graph has the old lanes
    hipGraphExecDestroy(P->gexec);
    P->gexec = nullptr;
  }
  return MI_OK;
}

int mi_plan_jk_profile(mi_plan* P, const double* D, double* J, double* K, void* stream, float* ms) {
  if (!P) return MI_EINVAL;
  if (P->nmol != 1) return set_err(P->ctx, MI_EINVAL, "Coulomb/exchange plans hold one molecule");
  if (!D || !J || !K) return set_err(P->ctx, MI_EINVAL, "NULL density or output");
  JKArgs jk{D, J, K};
  return profile_run(P, nullptr, 0, stream, ms, &jk);
}

void mi_plan_destroy(mi_plan* P) {
  if (!P) return;
  hipSetDevice(P->ctx->device);
  free_plan_buffers(P);
  delete P;
}

int mi_plan_stats(const mi_plan* P, int64_t* st, int32_t n) {
  if (!P || !st) return MI_EINVAL;
  const int64_t v[7] = {P->nbf, P->npair, P->npp, P->nq_shell, P->nq_prim, P->n_unique,
                        (int64_t)P->units.size() + 1};
  for (int i = 0; i < n && i < 7; ++i) st[i] = v[i];
  return MI_OK;
}

int mi_eri_unique(mi_ctx* ctx, const mi_basis* b, const mi_opts* opts, double* out, int64_t out_len) {
  if (!ctx) return MI_EINVAL;
  if (!out) return set_err(ctx, MI_EINVAL, "output pointer is NULL");
  int rc = validate(ctx, b);
  if (rc) return rc;
  if (out_len != mi_n_unique(b->nbf))
    return set_err(ctx, MI_ESIZE, "output length %lld, basis needs %lld", (long long)out_len,
                   (long long)mi_n_unique(b->nbf));
  mi_plan* P = nullptr;
  if ((rc = mi_plan_create(ctx, b, opts, &P))) return rc;
  double* d_out = nullptr;
  hipStream_t s = opts ? (hipStream_t)opts->stream : nullptr;
  hipError_t e = hipMalloc(&d_out, sizeof(double) * (size_t)out_len);
  if (e != hipSuccess) {
    mi_plan_destroy(P);
    return set_err(ctx, MI_ENOMEM, "hipMalloc of %lld doubles failed: %s", (long long)out_len, hipGetErrorString(e));
  }
  rc = mi_plan_execute(P, d_out, out_len, s);
  if (rc == MI_OK) {
    e = hipMemcpyAsync(out, d_out, sizeof(double) * (size_t)out_len, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = set_err(ctx, MI_EDEVICE, "copy back failed: %s", hipGetErrorString(e));
  }
  hipFree(d_out);
  mi_plan_destroy(P);
  return rc;
}

// mi_eri_unique without device allocation: the plan's device buffers (pair
// and primitive-pair records, tasks, block tables, HRR scratch of the
// global-table kernel) are carved from the static scratch area; the output is
// the caller's device buffer.  One lane (the caller's stream), no Schwarz pass
// at build time (mixed and screened plans run it inside the execute).  The
// host images of the buffers are freed on return, so the call synchronises
// its stream.  Calls that overlap on one device race on the shared area.
int mi_eri_unique_noalloc(mi_ctx* ctx, const mi_basis* b, const mi_opts* opts, double* out_device,
                          int64_t out_len) {
  if (!ctx) return MI_EINVAL;
  if (!out_device) return set_err(ctx, MI_EINVAL, "output pointer is NULL");
  int rc = validate(ctx, b);
  if (rc) return rc;
  if (out_len != mi_n_unique(b->nbf))
    return set_err(ctx, MI_ESIZE, "output length %lld, basis needs %lld", (long long)out_len,
                   (long long)mi_n_unique(b->nbf));
  if (opts && opts->order == MI_ORDER_JK) return set_err(ctx, MI_EINVAL, "order MI_ORDER_JK writes no unique array");
  if (!ctx->noalloc_scratch) return set_err(ctx, MI_EDEVICE, "static scratch area unavailable on device %d", ctx->device);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  mi_plan* P = nullptr;
  if ((rc = mi_plan_build(ctx, b, 1, opts, MI_BUILD_DEFER_UPLOAD | MI_BUILD_NO_BOUNDS | MI_BUILD_ONE_LANE, &P)))
    return rc;
  hipStream_t s = opts ? (hipStream_t)opts->stream : nullptr;
  P->arena = ctx->noalloc_scratch;
  P->arena_cap = MI_NOALLOC_SCRATCH_BYTES;
  P->arena_stream = s;
  rc = mi_plan_upload(P, nullptr);
  if (rc == MI_OK) rc = mi_plan_run(P, out_device, out_len, s, nullptr, nullptr);
  // the copies read P's host images: wait before they are freed (also after an error)
  const hipError_t e = hipStreamSynchronize(s);
  if (rc == MI_OK && e != hipSuccess) rc = set_err(ctx, MI_EDEVICE, "execute failed: %s", hipGetErrorString(e));
  mi_plan_destroy(P);
  return rc;
}

}  // extern "C"
