// mi_asan_driver.cpp -- the host side of libmi_hgp (plan builder, launch
// units, graphs, J/K, serving engine, allocation-free entry point, error
// paths) built with AddressSanitizer and UndefinedBehaviorSanitizer on the
// host only (device code unchanged) and driven through the C ABI on real
// inputs.  Test infrastructure: tests/test_gpu_host_sanitizers.py writes the
// bases and runs this program; any sanitizer report ends it with a nonzero
// status.  Built by tests/asan/Makefile.
//
// Input file (little-endian): int32 nmol, then per molecule int32 nshell,
// nprim, nbf, double xyz[3 nshell], int32 l[nshell], int32 off[nshell+1],
// double exp[nprim], coef[nprim], norm[nprim], bf_scale[nbf].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/mi_hgp.h"

namespace {

struct Mol {
  int32_t nshell = 0, nprim = 0, nbf = 0;
  std::vector<double> xyz, pexp, pcoef, pnorm, scale;
  std::vector<int32_t> l, off;
  mi_basis basis() const {
    return mi_basis{nshell, xyz.data(), l.data(), off.data(), pexp.data(), pcoef.data(), pnorm.data(), nbf,
                    scale.data()};
  }
};

int failures = 0;

#define CHECK(cond, ...)                      \
  do {                                        \
    if (!(cond)) {                            \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);      \
      std::fprintf(stderr, "\n");             \
      ++failures;                             \
    }                                         \
  } while (0)

#define HIPOK(x) CHECK((x) == hipSuccess, "%s", #x)

template <typename T>
bool rd(FILE* f, T* p, size_t n) {
  return std::fread(p, sizeof(T), n, f) == n;
}

std::vector<Mol> read_mols(const char* path) {
  std::vector<Mol> ms;
  FILE* f = std::fopen(path, "rb");
  if (!f) return ms;
  int32_t n = 0;
  if (!rd(f, &n, 1)) n = 0;
  for (int m = 0; m < n; ++m) {
    Mol x;
    if (!rd(f, &x.nshell, 1) || !rd(f, &x.nprim, 1) || !rd(f, &x.nbf, 1)) break;
    x.xyz.resize(3 * x.nshell);
    x.l.resize(x.nshell);
    x.off.resize(x.nshell + 1);
    x.pexp.resize(x.nprim);
    x.pcoef.resize(x.nprim);
    x.pnorm.resize(x.nprim);
    x.scale.resize(x.nbf);
    if (!rd(f, x.xyz.data(), x.xyz.size()) || !rd(f, x.l.data(), x.l.size()) || !rd(f, x.off.data(), x.off.size()) ||
        !rd(f, x.pexp.data(), x.pexp.size()) || !rd(f, x.pcoef.data(), x.pcoef.size()) ||
        !rd(f, x.pnorm.data(), x.pnorm.size()) || !rd(f, x.scale.data(), x.scale.size()))
      break;
    ms.push_back(std::move(x));
  }
  std::fclose(f);
  return ms;
}

bool close_enough(double a, double b) {
  const double d = std::fabs(a - b);
  return d <= 1e-12 || d <= 1e-10 * std::fabs(b);
}

mi_opts opts(int order, bool mixed) {
  mi_opts o{};
  o.order = order;
  o.precision = mixed ? MI_PREC_MIXED : MI_PREC_FP64;
  o.tau = mixed ? 1e-8 : 0.0;
  o.schwarz_cut = mixed ? 1e-12 : 0.0;
  o.method = MI_METHOD_HGP;
  return o;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s bases.bin\n", argv[0]);
    return 2;
  }
  std::vector<Mol> mols = read_mols(argv[1]);
  if (mols.size() < 4) {
    std::fprintf(stderr, "need at least 4 molecules in %s\n", argv[1]);
    return 2;
  }
  mi_ctx* ctx = nullptr;
  if (mi_create(0, &ctx) != MI_OK) {
    std::fprintf(stderr, "mi_create failed\n");
    return 2;
  }
  const int nb = (int)mols.size();
  std::vector<mi_basis> bases;
  for (const Mol& m : mols) bases.push_back(m.basis());

  // one-shot reference arrays of the first molecules (host arrays)
  const int nref = 4;
  std::vector<std::vector<double>> ref(nref);
  for (int m = 0; m < nref; ++m) {
    ref[m].assign(mi_n_unique(mols[m].nbf), 0.0);
    const mi_opts o = opts(MI_ORDER_CHRR, false);
    CHECK(mi_eri_unique(ctx, &bases[m], &o, ref[m].data(), (int64_t)ref[m].size()) == MI_OK, "eri_unique %d: %s", m,
          mi_last_error(ctx));
  }

  // batched plans, plain and mixed, eager and as a captured graph
  for (int mixed = 0; mixed < 2; ++mixed) {
    const mi_opts o = opts(MI_ORDER_CHRR, mixed);
    mi_plan* P = nullptr;
    CHECK(mi_batch_plan_create(ctx, bases.data(), nb, &o, &P) == MI_OK, "batch plan: %s", mi_last_error(ctx));
    if (!P) continue;
    const int64_t n = mi_plan_out_len(P);
    std::vector<int64_t> offs(nb + 1);
    CHECK(mi_plan_mol_offsets(P, offs.data()) == MI_OK, "offsets");
    double* d = nullptr;
    HIPOK(hipMalloc(&d, sizeof(double) * n));
    CHECK(mi_plan_execute(P, d, n, nullptr) == MI_OK, "execute: %s", mi_last_error(ctx));
    CHECK(mi_plan_graph_capture(P, d, n) == MI_OK, "capture: %s", mi_last_error(ctx));
    CHECK(mi_plan_graph_launch(P, nullptr) == MI_OK, "replay");
    HIPOK(hipDeviceSynchronize());
    std::vector<double> h(n);
    HIPOK(hipMemcpy(h.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int m = 0; m < nref; ++m) {
      int bad = 0;
      for (size_t k = 0; k < ref[m].size(); ++k) {
        const double tol = mixed ? 1e-10 : 0.0;
        const double x = h[offs[m] + k], y = ref[m][k];
        if (!(mixed ? std::fabs(x - y) <= tol : close_enough(x, y))) ++bad;
      }
      CHECK(bad == 0, "batch (mixed %d) molecule %d: %d slots off", mixed, m, bad);
    }
    // instrumentation entry points
    const int nl = mi_plan_nlaunch(P);
    std::vector<float> ms(nl + 2, 0.f);
    CHECK(mi_plan_profile(P, d, n, nullptr, ms.data()) == MI_OK, "profile");
    for (int i = 0; i < nl; ++i) {
      int32_t cls4[4], kern2[2];
      int64_t cnt3[3];
      CHECK(mi_plan_launch_info(P, i, cls4, cnt3, kern2) == MI_OK, "launch_info %d", i);
    }
    const int nu = mi_plan_nunit(P);
    std::vector<float> ums(nu + 1, 0.f);
    CHECK(mi_plan_profile_units(P, d, n, nullptr, ums.data()) == MI_OK, "profile_units");
    for (int u = 0; u < nu; ++u) {
      int32_t kind3[3], mem[64];
      CHECK(mi_plan_unit_info(P, u, kind3, mem, 64) == MI_OK, "unit_info %d", u);
    }
    CHECK(mi_plan_balance(P, d, n, nullptr) == MI_OK, "balance");
    int64_t st[8];
    CHECK(mi_plan_stats(P, st, 8) == MI_OK, "stats");
    if (mixed) {
      std::vector<double> Q(st[1]), V(st[1]);
      std::vector<int32_t> mol(st[1]);
      CHECK(mi_plan_pair_bounds(P, Q.data(), V.data(), mol.data()) == MI_OK, "pair bounds: %s", mi_last_error(ctx));
    }
    CHECK(mi_plan_execute(P, d, n - 1, nullptr) == MI_ESIZE, "wrong length accepted");
    mi_plan_destroy(P);
    HIPOK(hipFree(d));
  }

  // J/K: both plan layouts (J/K order without a cut groups pairs by first shell)
  {
    const Mol& m = mols[nb - 1];
    const int n = m.nbf;
    std::mt19937_64 rng(7);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<double> D((size_t)n * n), J0((size_t)n * n), K0((size_t)n * n), J1(J0), K1(K0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j <= i; ++j) D[(size_t)i * n + j] = D[(size_t)j * n + i] = nd(rng) / n;
    mi_opts o0 = opts(MI_ORDER_JK, false), o1 = opts(MI_ORDER_JK, true);
    const mi_basis b = m.basis();
    CHECK(mi_jk(ctx, &b, &o0, D.data(), J0.data(), K0.data()) == MI_OK, "jk fp64: %s", mi_last_error(ctx));
    CHECK(mi_jk(ctx, &b, &o1, D.data(), J1.data(), K1.data()) == MI_OK, "jk mixed: %s", mi_last_error(ctx));
    double dj = 0.0;
    for (size_t k = 0; k < J0.size(); ++k) dj = std::fmax(dj, std::fabs(J0[k] - J1[k]) + std::fabs(K0[k] - K1[k]));
    CHECK(dj <= 1e-9, "J/K mixed vs fp64 %.3e", dj);
  }

  // allocation-free entry point: two calls, the second smaller
  for (int m : {1, 0}) {
    const int64_t n = mi_n_unique(mols[m].nbf);
    double* d = nullptr;
    HIPOK(hipMalloc(&d, sizeof(double) * n));
    const mi_opts o = opts(MI_ORDER_CHRR, false);
    CHECK(mi_eri_unique_noalloc(ctx, &bases[m], &o, d, n) == MI_OK, "noalloc %d: %s", m, mi_last_error(ctx));
    std::vector<double> h(n);
    HIPOK(hipMemcpy(h.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost));
    int bad = 0;
    for (int64_t k = 0; k < n; ++k) bad += !close_enough(h[k], ref[m][k]);
    CHECK(bad == 0, "noalloc molecule %d: %d slots off", m, bad);
    HIPOK(hipFree(d));
  }

  // serving engine: ERI requests (merged into batches) and a J/K request
  {
    mi_server_opts so{4, 16, (int64_t)8 << 30};
    mi_server* S = nullptr;
    CHECK(mi_server_create(0, &so, &S) == MI_OK, "server create");
    if (S) {
      const int nreq = nb < 12 ? nb : 12;
      std::vector<double*> outs(nreq, nullptr);
      std::vector<int64_t> ids(nreq + 1, -1);
      const mi_opts o = opts(MI_ORDER_CHRR, false);
      for (int r = 0; r < nreq; ++r) {
        HIPOK(hipMalloc(&outs[r], sizeof(double) * mi_n_unique(mols[r].nbf)));
        CHECK(mi_server_submit(S, &bases[r], &o, MI_REQ_ERI, nullptr, outs[r], nullptr, nullptr, &ids[r]) == MI_OK,
              "submit %d: %s", r, mi_server_last_error(S));
        CHECK(mi_server_step(S) == MI_OK, "step");
      }
      const Mol& m = mols[nb - 1];
      const size_t nn = (size_t)m.nbf * m.nbf;
      double *dD = nullptr, *dJ = nullptr, *dK = nullptr;
      HIPOK(hipMalloc(&dD, 8 * nn));
      HIPOK(hipMalloc(&dJ, 8 * nn));
      HIPOK(hipMalloc(&dK, 8 * nn));
      HIPOK(hipMemset(dD, 0, 8 * nn));
      const mi_opts oj = opts(MI_ORDER_JK, true);
      const mi_basis bj = m.basis();
      CHECK(mi_server_submit(S, &bj, &oj, MI_REQ_JK, dD, nullptr, dJ, dK, &ids[nreq]) == MI_OK, "submit jk: %s",
            mi_server_last_error(S));
      CHECK(mi_server_drain(S) == MI_OK, "drain: %s", mi_server_last_error(S));
      for (int r = 0; r <= nreq; ++r) {
        int32_t state = -1;
        CHECK(mi_server_poll(S, ids[r], &state) == MI_OK && state == MI_REQ_DONE, "request %d state %d", r, state);
        CHECK(mi_server_release(S, ids[r]) == MI_OK, "release %d", r);
      }
      for (int r = 0; r < nref && r < nreq; ++r) {
        std::vector<double> h(ref[r].size());
        HIPOK(hipMemcpy(h.data(), outs[r], sizeof(double) * h.size(), hipMemcpyDeviceToHost));
        int bad = 0;
        for (size_t k = 0; k < h.size(); ++k) bad += !close_enough(h[k], ref[r][k]);
        CHECK(bad == 0, "served molecule %d: %d slots off", r, bad);
      }
      int64_t st[8];
      CHECK(mi_server_stats(S, st, 8) == MI_OK && st[3] == 0, "server failures");
      mi_server_destroy(S);
      for (double* p : outs) HIPOK(hipFree(p));
      HIPOK(hipFree(dD));
      HIPOK(hipFree(dJ));
      HIPOK(hipFree(dK));
    }
  }

  // error paths: no shells, a shell beyond the supported L, a short output
  {
    Mol bad = mols[0];
    mi_basis b = bad.basis();
    b.nshell = 0;
    const mi_opts o = opts(MI_ORDER_CHRR, false);
    std::vector<double> out(16);
    CHECK(mi_eri_unique(ctx, &b, &o, out.data(), 16) != MI_OK, "empty basis accepted");
    Mol hi = mols[0];
    hi.l[0] = mi_lmax() + 1;
    const mi_basis bh = hi.basis();
    CHECK(mi_eri_unique(ctx, &bh, &o, out.data(), 16) != MI_OK, "L beyond the tables accepted");
    const mi_basis b0 = mols[0].basis();
    CHECK(mi_eri_unique(ctx, &b0, &o, out.data(), 16) == MI_ESIZE, "short output accepted");
  }

  mi_destroy(ctx);
  // Leave without the runtime's exit-time teardown: the HSA runtime's static
  // destructors free memory after the sanitizer's device allocator has shut
  // down, which the sanitizer reports as a failed internal check (not a fault
  // of this library; every context and plan is destroyed above).
  if (failures) std::fprintf(stderr, "mi_asan_driver: %d failure(s)\n", failures);
  else std::printf("mi_asan_driver: ok (%d molecules)\n", nb);
  std::fflush(stdout);
  std::fflush(stderr);
  std::_Exit(failures ? 1 : 0);
}
