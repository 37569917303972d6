/*
 * mi_hgp.h -- C ABI of the MI355X-native Head-Gordon/Pople two-electron
 * integral engine (libmi_hgp.so).
 *
 * Drop-in boundary for MolecularIntegrals.jl's HGP path:
 *   all_twoe_ints_chrr(bfs::Basis)::Vector{Float64}     src/HGP.jl:22-34
 *   all_twoe_ints(bfs, ERI=coulomb)::Vector{Float64}    src/ERI.jl:104-112
 * Same basis in, same 8-fold-unique (ij|kl) array out, in either of the two
 * index orders those functions write (MI_ORDER_CHRR: enumeration of
 * iiterator, src/Basis.jl:203; MI_ORDER_IINDEX: iindex, src/Utils.jl:19).
 * Slots here are 0-based (the Julia slot minus one).
 *
 * Values are the package's tested definition, the normalised contracted
 * coulomb(bfs[i],bfs[j],bfs[k],bfs[l]) (src/ERI.jl:102, src/Basis.jl:86-98),
 * evaluated with HGP recurrences.  Plain calls (precision MI_PREC_FP64,
 * schwarz_cut 0) are bitwise deterministic.
 *
 * Plain pointers and sizes only; no exception crosses the ABI.  Every call
 * returns an MI_* status; mi_last_error() describes the last failure of a
 * context.  One context per host thread.
 */
#ifndef MI_HGP_H
#define MI_HGP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MI_OK 0
#define MI_EINVAL 1        /* malformed basis / options / sizes            */
#define MI_EUNSUPPORTED 2  /* angular momentum beyond the compiled tables  */
#define MI_EDEVICE 3       /* HIP runtime failure                          */
#define MI_ENOMEM 4        /* device allocation failed / budget exceeded   */
#define MI_ESIZE 5         /* output length does not match the basis       */

/* output orders */
#define MI_ORDER_IINDEX 0  /* all_twoe_ints:      slot = iindex(i,j,k,l) - 1   */
#define MI_ORDER_CHRR 1    /* all_twoe_ints_chrr: slot = enumeration index - 1 */
#define MI_ORDER_JK 2      /* no ERI output: a plan for mi_plan_jk_execute (without
                              a Schwarz cut its shell pairs are grouped by first
                              shell, so that more J/K contributions to one target
                              are summed before the atomics); mi_plan_execute on
                              it returns MI_EINVAL */

/* precision modes */
#define MI_PREC_FP64 0     /* everything in FP64                               */
#define MI_PREC_MIXED 1    /* FP32 VRR for quartets whose VRR-output bound < tau */

/* integral methods (mi_opts.method) */
#define MI_METHOD_HGP 0    /* Head-Gordon/Pople recurrences (all_twoe_ints_chrr, src/HGP.jl:22) */
#define MI_METHOD_RYS 1    /* Rys quadrature (all_twoe_ints_rys, src/Rys.jl:16): every
                              quartet class through the Rys kernel, in FP64 (precision
                              MI_PREC_MIXED screens and bounds as for HGP but keeps FP64) */

typedef struct mi_ctx mi_ctx;
typedef struct mi_plan mi_plan;

/*
 * Flattened Basis (MolecularIntegrals.Basis, src/Basis.jl:172-184):
 *  shells in bfs.shells order; functions in bfs.cgbfs order, i.e. shell by
 *  shell, each shell in shell_indices[L] order (src/Basis.jl:127-136).
 *  prim_coef[p]  raw Shell.coefs
 *  prim_norm[p]  PGBF norm of primitive p of the shell's first component
 *                (cgbf.pgbfs[p].norm, src/Basis.jl:29-33)
 *  bf_scale[i]   cgbfs[i].norm * cgbfs[i].pgbfs[1].norm / (first component's
 *                pgbfs[1].norm)
 *  With these, (ij|kl) = s_i s_j s_k s_l sum c'_p c'_q c'_r c'_s [pq|rs]
 *  where c'_p = prim_coef[p]*prim_norm[p], which equals coulomb() on CGBFs.
 */
typedef struct {
    int32_t nshell;
    const double *shell_xyz;        /* [3*nshell], bohr                  */
    const int32_t *shell_l;         /* [nshell], 0..MI_LMAX              */
    const int32_t *shell_prim_off;  /* [nshell+1]                        */
    const double *prim_exp;         /* [nprim]                           */
    const double *prim_coef;        /* [nprim]                           */
    const double *prim_norm;        /* [nprim]                           */
    int32_t nbf;                    /* sum over shells of (L+1)(L+2)/2   */
    const double *bf_scale;         /* [nbf]                             */
} mi_basis;

typedef struct {
    int32_t order;       /* MI_ORDER_*                                         */
    int32_t precision;   /* MI_PREC_*                                          */
    double tau;          /* MIXED: quartets with V_ab*V_cd < tau use FP32 VRR,
                            V = sqrt(max |(e0|e0)|) over a pair's contracted
                            VRR outputs (the V of mi_plan_pair_bounds): every
                            (e0|f0) of the quartet, i.e. every HRR input, is at
                            most V_ab V_cd.  Unlike schwarz_cut's Q, which bounds
                            the final integrals, V also bounds what the HRR
                            cancels, so FP32 rounding stays below ~tau * 1e-7. */
    double schwarz_cut;  /* 0: compute every quartet (plain); >0: skip below   */
    void *stream;        /* hipStream_t for *_device calls; NULL = default     */
    int32_t method;      /* MI_METHOD_* (0 = HGP)                               */
} mi_opts;

/* Context: device buffers, Boys grid and op tables for one GPU.
 * Replaces nothing in the reference (Julia keeps no device state). */
int mi_create(int device, mi_ctx **out);
void mi_destroy(mi_ctx *ctx);
const char *mi_last_error(const mi_ctx *ctx);
int32_t mi_lmax(void);                   /* largest supported shell L      */
/* The n-point Rys rule at T >= 0 as the Rys kernel evaluates it (host copy of
 * its tables): roots x[r] = t_r^2 in (0,1) and weights w[r] with
 * sum_r w_r x_r^k = F_k(T) for k < 2n.  Replaces Roots(n, X, roots, weights)
 * (src/Rys.jl:161-174, whose roots are u = t^2/(1-t^2)).  n = 1..mi_rys_nmax(). */
int mi_rys_rule(int32_t n, double T, double *x, double *w);
int32_t mi_rys_nmax(void);
int64_t mi_n_unique(int32_t nbf);        /* n(n+1)(n^2+n+2)/8, src/ERI.jl:106 */

/* Kernel selection (no counterpart in the reference).  rows: nrows x 6 int32
 * (quartet class, bra contraction bucket, ket contraction bucket, variant
 * [1 generated register kernel, 0 LDS wave kernel, 2 global-table kernel,
 * 3 LDS wave kernel reading its op tables from global memory, 5 Rys kernel
 * (1, 2 or 4 wavefronts), 6 Rys lane-group kernel (G = 16, 32 or 64 lanes
 * per shell quartet); 5 and 6 serve MI_METHOD_RYS plans, and HGP plans where
 * the table names them: the same integrals within rounding],
 * G = lanes per quartet for the register kernel or wavefronts per quartet /
 * block for the wave kernels, regime [0: class launches of fewer than 65536
 * shell quartets, 1: larger ones, 2: launches of MI_ORDER_JK plans (where a
 * class has no regime-2 entry they follow 0/1), -1: all]).  Replaces the whole table
 * (or nothing, on an invalid row); unlisted entries use the built-in
 * heuristic.  mi_create loads hgp_tuning.txt beside the library
 * (MI_HGP_TUNING overrides).  Affects plans created afterwards. */
int mi_ctx_set_tuning(mi_ctx *ctx, const int32_t *rows, int32_t nrows);
/* Number of streams class launches are spread over (1 = caller's stream only;
 * at most 16; default 4, HIP's default number of hardware queues -- more
 * streams overlap only if the process runs with GPU_MAX_HW_QUEUES raised to
 * match).  Affects plans created afterwards. */
int mi_ctx_set_streams(mi_ctx *ctx, int32_t nstreams);
/* Launch fusion (no counterpart in the reference, which has no device
 * launches).  on = 1 (the default): a plan that writes the unique array
 * issues its class launches of fewer than 65,536 shell quartets as fused
 * launches -- the generated register kernels in groups of similar register
 * use, the LDS wave kernel and the global-table kernel per wavefront count --
 * so latency-bound classes run side by side instead of one launch each.  on = 0: one launch
 * per class.  Results are bitwise the same; J/K executes and per-launch
 * profiles always use one launch per class.  Affects plans created
 * afterwards. */
int mi_ctx_set_fusion(mi_ctx *ctx, int32_t on);

/* One-shot drop-in: host basis in, host array out (caller-owned, out_len
 * must be mi_n_unique(nbf)).  Replaces all_twoe_ints_chrr (order CHRR,
 * src/HGP.jl:22) and all_twoe_ints (order IINDEX, src/ERI.jl:104); with
 * method MI_METHOD_RYS and order IINDEX, all_twoe_ints_rys (src/Rys.jl:16-25,
 * coulomb_rys :27-84, for every class up to (ff|ff), 7 roots; the reference's
 * Roots stops at 5). */
int mi_eri_unique(mi_ctx *ctx, const mi_basis *basis, const mi_opts *opts, double *out, int64_t out_len);

/* mi_eri_unique for callers that cannot allocate device memory (code that
 * must not call hipMalloc: tight serving loops, harnesses that forbid it).
 * No counterpart in the reference.  The output is the caller's DEVICE buffer
 * (out_len = mi_n_unique(nbf)); every other device buffer of the call comes
 * from a static 64 MiB scratch area that each device holds for the life of
 * the process, so the call makes no device allocation.  A basis whose plan
 * needs more returns MI_ENOMEM (about 60-300 functions fit, depending on
 * contraction and f shells).  Runs on opts->stream and synchronises it before
 * returning.  All calls on one device share the area: calls that overlap on
 * one device (other host threads) race.  Values are those of mi_eri_unique
 * (bit for bit in plain FP64). */
int mi_eri_unique_noalloc(mi_ctx *ctx, const mi_basis *basis, const mi_opts *opts, double *out_device,
                          int64_t out_len);

/* Plan / execute split: the plan holds the basis topology on the device
 * (shell pairs sorted by class, tasks); execute recomputes shell-pair data
 * and all ERIs into a caller-owned DEVICE buffer on opts->stream, without
 * host synchronisation.  The analogue of the reference's eri_fetcher
 * (src/Basis.jl:201-212) plus the threaded loop of src/HGP.jl:27-32. */
int mi_plan_create(mi_ctx *ctx, const mi_basis *basis, const mi_opts *opts, mi_plan **out);
int mi_plan_execute(mi_plan *plan, double *out_device, int64_t out_len, void *stream);
void mi_plan_destroy(mi_plan *plan);
/* Batched plan over nmol molecules (configuration "many molecules of uneven
 * size in one launch sequence"; no counterpart in the reference, which
 * handles one Basis per call).  Shell pairs and quartets stay inside each
 * molecule; the output is the molecules' unique arrays back to back, molecule
 * m starting at offsets[m] (mi_plan_mol_offsets, nmol+1 entries). */
int mi_batch_plan_create(mi_ctx *ctx, const mi_basis *bases, int32_t nmol, const mi_opts *opts, mi_plan **out);
int64_t mi_plan_out_len(const mi_plan *plan);
int mi_plan_mol_offsets(const mi_plan *plan, int64_t *offsets);

/* Direct Coulomb/exchange for one molecule, the ERI tensor never stored:
 *   J_ij = sum_kl (ij|kl) D_kl,  K_ij = sum_kl (ik|jl) D_kl  (D symmetric),
 * the index algebra of make2JmK (src/ERI.jl:114-128; G = 2J - K there).
 * D, J, K are nbf x nbf row-major.  opts->schwarz_cut > 0 skips quartets with
 * Q_ab Q_cd below it (Q = sqrt(max |(ab|ab)|)); MI_PREC_MIXED runs the VRR of
 * quartets with V_ab V_cd < tau (VRR-output bounds, as mi_opts.tau) in FP32.
 * Accumulation uses FP64 atomics, so
 * J and K agree across runs to rounding, not bit for bit.
 * mi_jk: host arrays.  mi_plan_jk_execute: device arrays, async on stream. */
int mi_jk(mi_ctx *ctx, const mi_basis *basis, const mi_opts *opts, const double *D, double *J, double *K);
int mi_plan_jk_execute(mi_plan *plan, const double *D_device, double *J_device, double *K_device, void *stream);

/* Per shell-pair bounds from the last execute of a plan that has a Schwarz
 * pass (MI_PREC_MIXED or schwarz_cut > 0): Q = sqrt(max |(ab|ab)|),
 * V = sqrt(max |(e0|e0)|) over the pair's VRR outputs, and the molecule of
 * each pair; npair entries in the plan's pair order (mi_plan_stats[1]).
 * Synchronous. */
int mi_plan_pair_bounds(const mi_plan *plan, double *Q, double *V, int32_t *mol);

/* Capture one execute (writing out_device) into a HIP graph owned by the
 * plan; mi_plan_graph_launch replays it on any stream.  The graph holds no
 * memset nodes. */
int mi_plan_graph_capture(mi_plan *plan, double *out_device, int64_t out_len);
int mi_plan_graph_launch(mi_plan *plan, void *stream);

/* Plan statistics: [0]=nbf [1]=shell pairs [2]=primitive pairs
 * [3]=shell quartets [4]=primitive quartets [5]=unique ERIs
 * [6]=kernel launches per execute. */
int mi_plan_stats(const mi_plan *plan, int64_t *stats, int32_t nstats);

/* Instrumentation (no counterpart in the reference, whose timing lives in
 * test/timing.jl).  A plan executes as launch -1 (shell-pair kernel) then
 * launches 0..n-1, one per quartet class present. */
int mi_plan_nlaunch(const mi_plan *plan);
/* cls4 = (la,lb,lc,ld); counts3 = (shell quartets, primitive quartets,
 * blocks); kernel2 = (variant, wavefronts per block): variant 1 = generated
 * register kernel (wavefronts 0), 0 = LDS wave kernel, 2 = global-table
 * kernel, 3 = LDS wave kernel with op tables in global memory, 4 =
 * double-double table kernel, 5 = Rys kernel, 6 = Rys lane-group kernel
 * (second entry: lanes per group).  Returns MI_OK
 * or MI_EINVAL on a bad index. */
int mi_plan_launch_info(const mi_plan *plan, int32_t i, int32_t *cls4, int64_t *counts3, int32_t *kernel2);
/* One execute with HIP events between launches on `stream`; synchronises and
 * writes nlaunch+1 durations in ms (ms[0] = shell-pair kernel, with the
 * Schwarz pass and the output zero fill of a screening plan). */
int mi_plan_profile(mi_plan *plan, double *out_device, int64_t out_len, void *stream, float *ms);
/* One profiled execute (output written as by mi_plan_execute), then the
 * launch units are reassigned to stream lanes by their measured durations
 * (longest first, each to the least-loaded lane) instead of the built-in cost
 * estimate.  Results do not change; a captured graph is dropped (capture
 * again).  For small, latency-bound plans executed many times. */
int mi_plan_balance(mi_plan *plan, double *out_device, int64_t out_len, void *stream);
/* Launch units of a unique-array execute (mi_ctx_set_fusion): unit u is one
 * class launch or a fused launch of several.  kind3 = (kind, wavefronts per
 * block, member count): kind 0 = one class launch, 1 = fused register
 * kernels, 2 = fused LDS wave kernels, 3 = fused global-table kernels;
 * members receives up to cap indices
 * (0..nlaunch-1, as mi_plan_launch_info) of the class launches it covers,
 * heaviest first. */
int mi_plan_nunit(const mi_plan *plan);
int mi_plan_unit_info(const mi_plan *plan, int32_t u, int32_t *kind3, int32_t *members, int32_t cap);
/* One execute issued as mi_plan_execute issues it, with HIP events between
 * the units on `stream`; synchronises and writes nunit+1 durations in ms
 * (ms[0] = shell-pair kernel, with the Schwarz pass and the output zero fill
 * of a screening plan). */
int mi_plan_profile_units(mi_plan *plan, double *out_device, int64_t out_len, void *stream, float *ms);
/* The same for one Coulomb/exchange execute (device arrays as in
 * mi_plan_jk_execute); ms[0] includes the Schwarz pass, ms[1] the zeroing
 * of J and K. */
int mi_plan_jk_profile(mi_plan *plan, const double *D_device, double *J_device, double *K_device, void *stream,
                       float *ms);

/* ---------------------------------------------------------------------------
 * Serving engine (no counterpart in the reference, whose drivers evaluate one
 * Basis per blocking call; callers: request servers that take ERI or J/K
 * requests for molecules of different sizes as they arrive).
 *
 * Requests are submitted with caller-owned DEVICE buffers (out_device of
 * mi_n_unique(nbf) doubles for MI_REQ_ERI; D, J, K of nbf*nbf for MI_REQ_JK)
 * and queued in arrival order; the basis is copied at submit.  mi_server_step
 * is the engine and never blocks: it retires finished batches (event query),
 * then admits the head of the queue while fewer than max_inflight batches are
 * on the device and the request's device bytes (plan, HRR scratch, output)
 * fit hbm_budget -- the head waits for memory otherwise.  ERI requests with
 * equal options at the head are merged (up to max_batch and the budget) into
 * one batched plan; a J/K request is a batch of its own.  Each batch runs on
 * its own stream, so requests admitted while others run overlap with them.
 * Device memory is allocated and freed stream-ordered on the batch stream.
 * A server and its requests belong to one host thread.
 * ------------------------------------------------------------------------- */
typedef struct mi_server mi_server;
typedef struct {
  int32_t max_inflight; /* batches on the device at once, 1..8 */
  int32_t max_batch;    /* ERI requests merged into one batch */
  int64_t hbm_budget;   /* device bytes the engine may hold for batches */
} mi_server_opts;
#define MI_REQ_ERI 0
#define MI_REQ_JK 1
/* request states (mi_server_poll) */
#define MI_REQ_QUEUED 0
#define MI_REQ_RUNNING 1
#define MI_REQ_DONE 2
#define MI_REQ_FAILED 3
int mi_server_create(int32_t device, const mi_server_opts *opts, mi_server **out);
void mi_server_destroy(mi_server *server);
const char *mi_server_last_error(const mi_server *server);
/* opts: order, precision, tau, schwarz_cut as for mi_eri_unique / mi_jk
 * (opts->stream is ignored).  *id receives the request id. */
int mi_server_submit(mi_server *server, const mi_basis *basis, const mi_opts *opts, int32_t kind,
                     const double *D_device, double *out_device, double *J_device, double *K_device, int64_t *id);
int mi_server_step(mi_server *server);
int mi_server_poll(mi_server *server, int64_t id, int32_t *state);
/* Forget a finished request (its id becomes unknown). */
int mi_server_release(mi_server *server, int64_t id);
/* Run the engine until the queue is empty and nothing is in flight
 * (blocking); MI_EDEVICE if any request failed. */
int mi_server_drain(mi_server *server);
/* st[0..7] = queued, running (requests on the device), done, failed,
 * batches launched, device bytes in use, peak device bytes, batches in
 * flight. */
int mi_server_stats(const mi_server *server, int64_t *stats, int32_t nstats);

#ifdef __cplusplus
}
#endif
#endif /* MI_HGP_H */
