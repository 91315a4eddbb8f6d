/* deghip.h -- C ABI of the MI355X (gfx950) EnsembleGPUKernel solver library (libdeghip.so).
 *
 * This is the boundary a host language binds (Julia `ccall`, Python `ctypes`). It replaces
 * the device half of DiffEqGPU's EnsembleGPUKernel path:
 *   vectorized_solve  (fixed step)   src/ensemblegpukernel/lowerlevel_solve.jl:53-131
 *   vectorized_asolve (adaptive)     src/ensemblegpukernel/lowerlevel_solve.jl:253-346
 *   ode_solve_kernel / ode_asolve_kernel   src/ensemblegpukernel/kernels.jl:1-152
 *   the KernelAbstractions ROCm backend    ext/AMDGPUExt.jl:1-17
 * The host side keeps output sizing and saveat normalisation (lowerlevel_solve.jl:64-109,
 * :271-324); this library receives the normalised arrays and fills them.
 *
 * Conventions: plain pointers and sizes only. Every array is structure-of-arrays with a
 * leading dimension `stride` (>= id_offset + n_traj):
 *   u0[c*stride + i], p[j*stride + i], ts[row*stride + i], us[(row*n + c)*stride + i].
 * Input and output arrays are device pointers owned by the caller. Functions return
 * DEG_OK or a negative DEG_E_* code; deg_last_error() describes the last failure of the
 * calling thread. A context is not thread-safe; the JIT cache is.
 */
#ifndef DEGHIP_H
#define DEGHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define DEG_OK 0
#define DEG_E_INVALID (-1)
#define DEG_E_HIP (-2)
#define DEG_E_COMPILE (-3)
#define DEG_E_UNSUPPORTED (-4)
#define DEG_E_NOMEM (-5)

/* methods: GPUTsit5, GPUVern7, GPURosenbrock23, GPUVern9, GPURodas4, GPURodas5P, GPUKvaerno3, GPUKvaerno5
 * (src/ensemblegpukernel/gpukernel_algorithms.jl:30, 57, 112, 83, 137, 162, 187, 212) */
#define DEG_TSIT5 0
#define DEG_VERN7 1
#define DEG_RB23 2
#define DEG_VERN9 3
#define DEG_RODAS4 4
#define DEG_RODAS5P 5
#define DEG_KVAERNO3 6
#define DEG_KVAERNO5 7
/* fixed-step SDE methods GPUEM, GPUSIEA (gpukernel_algorithms.jl:243, 266; kernels
 * perform_step/gpu_em_perform_step.jl:1-74, gpu_siea_perform_step.jl:49-139). Their kernels take
 * DEG_F_PARITY and the register cap only; the traced source also defines the diffusion (see
 * deg_kernel_compile) and the Wiener increments come from deg_draw_normal's generator. */
#define DEG_EM 8
#define DEG_SIEA 9

/* compile flags (one code object per (rhs, n, np, method, flags)) */
#define DEG_F_ADAPTIVE (1u << 0)   /* PI-controlled adaptive stepping (vectorized_asolve) */
#define DEG_F_PARITY (1u << 1)     /* no FP contraction: the reference's unfused Float32 sequence */
#define DEG_F_DETPOW (1u << 2)     /* controller powers by a deterministic polynomial pow */
#define DEG_F_KAHAN (1u << 3)      /* compensated accumulation of t */
#define DEG_F_PERSISTENT (1u << 4) /* persistent grid; lanes refill from a device work counter */
#define DEG_F_LEARNED (1u << 5)    /* MLP step-size proposal, accepted or rejected by the error estimate */
#define DEG_F_FASTPOW (1u << 6)    /* controller powers by hardware log2/exp2 */
#define DEG_F_STEPLOG (1u << 7)    /* write one log row per attempt (training data) */
#define DEG_F_SAVEAT_GLOBAL (1u << 8) /* saveat times read from global memory (needed above 4096 points) */
/* tail compaction: once the queue is empty, waves with few trajectories left hand their states to
 * another wave of the workgroup through LDS and leave (adaptive persistent kernels) */
#define DEG_F_COMPACT (1u << 12)
/* batched dense output (adaptive GPUTsit5, n <= 4): a step that crosses save points is stashed in
 * LDS and a wave interpolates for many stashed lanes at once; results are unchanged */
#define DEG_F_BATCH_SAVES (1u << 13)
/* stopping times: the step that reaches (within 100 eps) the next of deg_solve_args.tstops ends
 * exactly there (the `tstops` keyword, gpu_tsit5_perform_step.jl:18-27 and :158-164) */
#define DEG_F_TSTOPS (1u << 14)
/* the traced source defines callbacks (DEG_NDCB / DEG_NCCB and their functions, see
 * deg_kernel_compile), applied after every accepted step (integrator_utils.jl:274-325) */
#define DEG_F_CALLBACKS (1u << 15)
/* persistent adaptive ODE kernels: read each trajectory's inputs from deg_solve_args.inputs_ordered when it
 * is set (the code path is compiled in only with this flag) */
#define DEG_F_ORDERED_INPUTS (1u << 16)
/* register cap: compile for at least w (1..7) resident waves per SIMD (bits 9..11; 0: no cap) */
#define DEG_F_WAVES_SHIFT 9
#define DEG_F_WAVES(w) ((uint32_t)(w) << DEG_F_WAVES_SHIFT)

/* per-trajectory return codes written to retcode[] */
#define DEG_RET_SUCCESS 0
#define DEG_RET_DTMIN 1     /* the reference raises error("dt<dtmin") (gpu_tsit5_perform_step.jl:102) */
#define DEG_RET_MAXITERS 2  /* attempt cap reached (no reference counterpart: the reference loops on) */
#define DEG_RET_NONFINITE 3 /* final t or u not finite */
#define DEG_RET_TERMINATED 4 /* terminate! in a callback's affect (ReturnCode.Terminated) */
#define DEG_RET_FAILURE 5   /* an event time before the step start: the reference raises
                               error("Current interpolant only works between tprev and t") */

typedef struct deg_ctx deg_ctx;
typedef struct deg_kernel deg_kernel;

/* Arguments of one ensemble solve. Mirrors the keywords of vectorized_solve /
 * vectorized_asolve (lowerlevel_solve.jl:53-59, :253-260) after host normalisation. */
typedef struct deg_solve_args {
    uint32_t struct_size;    /* sizeof(deg_solve_args) */
    int32_t n;               /* state dimension; must equal the kernel's */
    int64_t n_traj;          /* trajectories in this call */
    int64_t stride;          /* leading dimension of the SoA arrays; 0 means n_traj */
    int64_t id_offset;       /* first trajectory (column) of this call */
    const float *u0;         /* [n][stride] */
    const float *p;          /* [np][stride]; may be NULL when np == 0 */
    const float *t0s;        /* [stride] per-trajectory tspan[1], or NULL for t0 */
    const float *tfs;        /* [stride] per-trajectory tspan[2], or NULL for tf */
    float t0, tf;            /* shared tspan; t0 is also the prefill value of unsaved ts rows */
    float dt;                /* fixed step, or initial step of the adaptive method */
    float abstol, reltol;    /* adaptive tolerances */
    const float *saveat;     /* [nsaveat] normalised saveat times, or NULL */
    int32_t nsaveat;
    int32_t nsave;           /* rows of ts/us (2 for endpoint-only output) */
    int32_t save_everystep;  /* fixed step: write every step (rows 1..nsave-1) */
    int32_t block_size;      /* 0: 256 */
    float *ts;               /* [nsave][stride] */
    float *us;               /* [nsave][n][stride] */
    int32_t *retcode;        /* [stride] or NULL */
    int32_t *naccept;        /* [stride] or NULL: accepted steps */
    int32_t *nreject;        /* [stride] or NULL: rejected attempts */
    int64_t max_attempts;    /* per-trajectory attempt cap; 0: 10^7 (ends lanes whose t stops advancing) */
    void *stream;            /* hipStream_t to launch on; NULL: the null (default) stream */
    int32_t grid_blocks;     /* persistent kernels: 0 fills the device */
    int32_t refill_batch;    /* persistent kernels: a wave refills once this many lanes are idle (1..64); 0: 4 */
    float *elapsed_ms;       /* optional: kernel time from HIP events around the launch (synchronises) */
    const void *mlp;         /* DEG_F_LEARNED: packed proposer weights (device), see deg_mlp_* */
    float *steplog;          /* DEG_F_STEPLOG: [steplog_cap][DEG_STEPLOG_COLS] rows (device) */
    uint32_t *steplog_count; /* DEG_F_STEPLOG: device row counter (zeroed by the caller) */
    int64_t steplog_cap;
    int32_t chunk;           /* persistent kernels: trajectory ids a wave claims per queue atomic; 0: 16 */
    int32_t mlp_every;       /* DEG_F_LEARNED: evaluate the proposer every k-th iteration (reusing its last
                                correction in between); 0 or 1: every attempt */
    const int32_t *order;    /* optional [n_traj] start order of trajectories id_offset + order[k] (e.g. from
                                deg_start_order or deg_order_by_span); a plain grid starts them in k order, a
                                persistent grid's 8 queues (one per XCD) take k = q, q + 8, q + 16, ... each,
                                so every XCD starts on its share of the first entries; NULL: ascending ids */
    const float *tstops;     /* DEG_F_TSTOPS: [ntstops] ascending stopping times shared by all trajectories
                                (device); the `tstops` keyword of vectorized_solve / vectorized_asolve */
    int32_t ntstops;
    int32_t reserved0;
    const uint64_t *seeds;   /* DEG_EM / DEG_SIEA: [stride] per-trajectory seeds (SDEProblem `seed`), or NULL */
    uint64_t seed;           /* DEG_EM / DEG_SIEA: the seed of every trajectory when seeds is NULL */
    int64_t traj_base;       /* DEG_EM / DEG_SIEA: ensemble index of column 0; the noise of column i is the
                                stream traj_base + i, so a chunked ensemble draws what one launch would */
    float *inputs_ordered;   /* optional, kernels compiled with DEG_F_ORDERED_INPUTS, with order: device array of (n + np) columns of
                                8 * ceil(n_traj / 8) floats, the trajectories' u0 then p values in queue-major
                                start-order position (order entry k at column index (k mod 8) * ceil(n_traj / 8)
                                + k / 8: the entries each of the 8 queues serves, contiguously).
                                deg_start_order / deg_start_order_model fill it when set, and deg_solve then
                                reads each trajectory's inputs from it in the order the queues hand them out
                                instead of gathering them from u0 and p (no counterpart in the reference).
                                NULL: inputs are read from u0 and p */
} deg_solve_args;

/* Context: one device (launches go to the stream given per call). */
int deg_ctx_create(int device, deg_ctx **out);
int deg_ctx_destroy(deg_ctx *ctx);
/* Number of compute units of the context's device. */
int deg_ctx_num_cus(deg_ctx *ctx);

/* Directory for compiled code objects (shared by all contexts; NULL resets to none). */
int deg_set_cache_dir(const char *dir);

/* Compile (or fetch from the cache) the solver for one traced right-hand side.
 * rhs_src defines, in HIP C:
 *   __device__ void deg_rhs(const float* u, const float* p, float t, float* du);
 * and for DEG_RB23, DEG_RODAS4, DEG_RODAS5P, DEG_KVAERNO3 and DEG_KVAERNO5 also deg_jac (row-major
 * J[r*n + c] = d du_r / d u_c) and deg_tgrad (analytic, forward-mode, or forward differences of deg_rhs
 * for autodiff = false, alg_utils.jl:17-27: the library only calls them). For those methods a constant mass matrix M (M u' = f,
 * ODEFunction(f; mass_matrix = M), nlsolve/type.jl:138 and the mass_matrix branches of the
 * Rosenbrock / Rodas steps) is given as
 *   #define DEG_MASS 1   #define DEG_MASS_ENTRIES <n*n float literals, row-major>
 * a singular M makes the rows with zero mass algebraic (a DAE); without DEG_MASS, M is the identity.
 * A problem with dual-number inputs (ForwardDiff duals in u0 / p, test/gpu_kernel_de/forward_diff.jl) is
 * an augmented state of n values followed by the P partials of each component, n (1 + P) entries in all,
 * with the parameters laid out the same way; the source then also defines
 *   #define DEG_DUAL_NV <n>   #define DEG_DUAL_P <P>
 * which selects the dual error norm (any ODE method; no callbacks, mass matrix, proposer or step logs). For
 * the implicit methods deg_jac then writes the n (1 + P) x n Jacobian over the value columns, J[r*n + c]
 * (rows n + i P + j: partial j of the Jacobian's row i), deg_tgrad all n (1 + P) entries, and the solver
 * factors W over the n values and solves each partial with that factor.
 * With DEG_F_CALLBACKS it also defines the callbacks (GPUDiscreteCallback / GPUContinuousCallback,
 * src/ensemblegpukernel/callbacks.jl:1-124), j counting from 0 in CallbackSet order:
 *   #define DEG_NDCB <discrete callbacks>   #define DEG_NCCB <continuous callbacks>
 *   bool  deg_dcb_cond(int j, const float* u, const float* p, float t);
 *   int   deg_dcb_affect(int j, float* u, float* p, float t);
 *   float deg_ccb_cond(int j, const float* u, const float* p, float t);
 *   int   deg_ccb_affect(int j, int neg, float* u, float* p, float t);   neg: affect_neg!
 *   int   deg_ccb_mode(int j);     bit 0 affect! present, bit 1 affect_neg! present,
 *                                  bits 2-3 rootfind (0 NoRootFind, 1 LeftRootFind, 2 RightRootFind)
 *   float deg_ccb_abstol(int j), deg_ccb_nudge(int j) (repeat_nudge), deg_ccb_dtrelax(int j);
 * an affect updates u (and p) in place and returns DEG_AFF_* bits.
 * For DEG_EM and DEG_SIEA, deg_rhs is the drift f and the source also defines the diffusion g
 * (SDEProblem(f, g, ...), the noise shape from noise_rate_prototype):
 *   #define DEG_NM <noise columns m>   #define DEG_DIAG_NOISE <1 diagonal, 0 general>
 *   void deg_noise(const float* u, const float* p, float t, float* g);
 * writing n entries (diagonal noise, multiplied elementwise with the n increments) or the n x m
 * matrix row-major g[r*m + c]. DEG_SIEA requires diagonal noise (lowerlevel_solve.jl:185-186).
 * Replaces the compile of ode_solve_kernel / ode_asolve_kernel for one problem type. */
#define DEG_AFF_TERMINATE 1u /* terminate!(integrator) */
#define DEG_AFF_KEEP_FSAL 2u /* u_modified!(integrator, false) */
int deg_kernel_compile(deg_ctx *ctx, const char *rhs_src, int n, int np, int method, uint32_t flags,
                       deg_kernel **out);
int deg_kernel_destroy(deg_kernel *k);
/* Resource figures of the compiled kernel: which = 0 VGPRs, 1 SGPRs, 2 scratch bytes/lane,
 * 3 max block size, 4 resident blocks per CU at block_size 256. */
int deg_kernel_attr(deg_kernel *k, int which, int *value);

/* Compile into the cache without a GPU (build-time warm-up). Writes the 16-hex-digit cache
 * key to key_out (>= 17 bytes) when non-NULL. */
int deg_jit_precompile(const char *rhs_src, int n, int np, int method, uint32_t flags, char *key_out);

/* The full HIP C++ source the JIT compiles for this configuration (for inspection).
 * Returns the length; copies at most cap-1 bytes plus a terminator into buf when non-NULL. */
int64_t deg_jit_source(const char *rhs_src, int n, int np, int method, uint32_t flags, char *buf, int64_t cap);

/* Launch one ensemble solve on args->stream (asynchronous unless elapsed_ms is set).
 * Replaces the kernel launch of vectorized_solve / vectorized_asolve
 * (lowerlevel_solve.jl:119-123, :339-343). A persistent kernel (DEG_F_PERSISTENT) keeps its work
 * counters in the kernel object, so launches of one kernel object must be ordered: one stream, or
 * streams synchronised between the launches (a second kernel object of the same configuration
 * has its own counters). */
int deg_solve(deg_kernel *k, const deg_solve_args *args);

/* Wait for all work on the context's device. */
int deg_sync(deg_ctx *ctx);

/* Last error message of the calling thread ("" when none). */
const char *deg_last_error(void);

/* Library version string. */
const char *deg_version(void);

/* sizeof(deg_solve_args) as compiled into the library, for bindings to check their layout. */
uint32_t deg_solve_args_size(void);

/* ---- parameter sweeps (no counterpart in the reference: its sweeps re-run solve() from the host) ---- */

/* Seeded uniform draws on the device: out[i] = lo + scale * U(seed, (id_offset + i) * rows + row)
 * for i in [0, n), U the SplitMix64 counter-based U[0,1) Float32 (diffeqgpu.jl_amd/rng.py draws
 * the same numbers on the host). Used to re-draw ensemble parameters for each sweep without a
 * host round trip; asynchronous on `stream`. Takes the place of the host-side `prob_func`
 * re-draws of an EnsembleProblem (src/solve.jl:187-202) for sweeps that are graph-captured. */
int deg_draw_uniform(float *out, int64_t n, int64_t id_offset, uint64_t seed, int32_t rows, int32_t row,
                     float lo, float scale, void *stream);

/* Standard normal draws of the SDE kernels' generator: out[c * n + i] = normal c (0 <= c < m) of
 * step `step` of noise stream id_offset + i under `seed`, for i in [0, n) -- the increment
 * sqrt(dt) * out of that step in a DEG_EM / DEG_SIEA solve (traj_base + column = id_offset + i).
 * Philox4x32-10 with key = seed and counter = (step, c / 4, id low, id high); Box-Muller on word
 * pairs with a deterministic ln and sincos. Asynchronous on `stream`. Takes the place of the
 * device `randn` of em_kernel / siea_kernel (gpu_em_perform_step.jl:48-53), exposed for tests
 * and for callers that want the same Wiener paths outside a solve. */
int deg_draw_normal(float *out, int64_t n, int64_t id_offset, uint64_t seed, uint32_t step, int32_t m,
                    void *stream);

/* Longest-first order of trajectories for the persistent queue: writes order[0..n) with a
 * permutation of 0..n-1 sorted by descending span tfs[i] - t0s[i] (NULL arrays: the shared t0/tf),
 * in 2047 buckets that split the range of the positive spans' float bit patterns evenly (about
 * 1/2000 octave for spans within a factor 2; no order within a bucket); zero, negative and NaN spans
 * come last. work: 4096 int32 of device memory. Asynchronous on `stream`, graph-capturable.
 * Passed as deg_solve_args.order it shortens the tail in which lanes of finished trajectories
 * idle (no counterpart in the reference, which launches one thread per trajectory). */
int deg_order_by_span(const float *t0s, const float *tfs, float t0, float tf, int64_t n, int32_t *order,
                      int32_t *work, void *stream);

/* Cost-ordered start (no counterpart in the reference, which starts trajectories in index order):
 * key[i] = |tf_i - t0_i| / h0_i for the n_traj trajectories of `a`, with h0 the standard initial step
 * estimate (Hairer, Norsett, Wanner; sc = abstol + |u0| reltol, h0 = 0.01 ||u0/sc|| / ||f(u0)/sc||)
 * computed by the kernel's own traced RHS, then order = the trajectories by descending key
 * (deg_order_by_span's bucketed counting sort; work: 4096 int32). Pass order as a->order to deg_solve:
 * trajectories likely to take the most steps start first, which shortens the tail of a persistent
 * grid. ODE kernels only; asynchronous on a->stream, graph-capturable; results do not depend on
 * the order. key: [n_traj] float, order: [n_traj] int32, device memory. */
int deg_start_order(deg_kernel *k, const deg_solve_args *a, float *key, int32_t *order, int32_t *work);

/* Cost-ordered start with a cost model (no counterpart in the reference): the key is the predicted
 * attempt count 2^(w . x) of a linear model on log2 features, x = [1, log2 key, log2(|u0_c| + 2^-20)
 * (c < n), log2(|p_j| + 2^-20) (j < np), log2 span] with key = span / h0 as in deg_start_order; model
 * holds the nmodel = 3 + n + np weights (device), fitted by the host on a pilot ensemble's attempt
 * counts (diffeqgpu.jl_amd/costmodel.py). Otherwise as deg_start_order. */
int deg_start_order_model(deg_kernel *k, const deg_solve_args *a, const float *model, int32_t nmodel, float *key,
                          int32_t *order, int32_t *work);

/* Allocation-free variant of deg_start_order (no counterpart in the reference; for callers that cannot
 * allocate device memory: code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The cost keys and the counting sort's counters live in a static device array of the library, one
 * per device, sized for at most DEG_NOALLOC_MAX_TRAJ trajectories (16 MiB + 16 KiB, held until the
 * process ends); larger n_traj is rejected with DEG_E_INVALID. Writes the same order as
 * deg_start_order (same keys, same buckets; order within a bucket is not fixed in either). Calls that
 * overlap on one device share that array and race: order them on one stream. */
#define DEG_NOALLOC_MAX_TRAJ 4194304
int deg_start_order_noalloc(deg_kernel *k, const deg_solve_args *a, int32_t *order);

typedef struct deg_graph deg_graph;
/* Capture every launch made on `stream` (non-NULL) by this thread between begin and end --
 * deg_draw_uniform, deg_solve, and any other stream work -- into an instantiated HIP graph.
 * Inside a capture deg_solve must not request elapsed_ms. The captured work contains no
 * memset nodes (the persistent queue is cleared by a kernel), so replays are exact. */
int deg_graph_begin(void *stream);
int deg_graph_end(void *stream, deg_graph **out);
/* Replay the captured work on `stream` (asynchronous). Replays on one stream are ordered; the
 * persistent kernels' work counters are per compiled kernel, so two replays that use the same
 * kernel must not run concurrently on different streams. */
int deg_graph_launch(deg_graph *g, void *stream);
int deg_graph_num_nodes(deg_graph *g, int64_t *count);
int deg_graph_destroy(deg_graph *g);

/* ---- learned step-size proposer (no counterpart in the reference, whose only controller is
 * the PI controller of integrator_utils.jl:1-11 and gpu_*_perform_step.jl; the PI controller
 * stays the fallback, and the embedded error estimate still accepts or rejects every step) ---- */

/* Step-log rows written by kernels compiled with DEG_F_STEPLOG: one row per attempt n that has
 * a successor, [x0..x5 at n, log2(dt_{n+1}/dt_n), log2 EEst_{n+1}, z_PI at n, accepted at n, 0, 0],
 * with x0 = log2 EEst_n, x1 = log2 EEst_{n-1}, x2 = log2 dt_n, x3 = log2 ||u||, x4 = log2 ||f(u)||,
 * x5 = log2(dt_n / dt_{n-1}) (RMS norms) and z_PI = log2 of the PI controller's step factor.
 * Rows past steplog_cap are dropped; the counter keeps counting. The learned solver (DEG_F_LEARNED)
 * steps with log2(dt_{n+1}/dt_n) = clamp(z_PI + MLP(x), log2 qmin, log2 qmax). */
#define DEG_STEPLOG_COLS 12
#define DEG_MLP_NPARAM 1313

/* sizeof the packed BF16 inference weights (the `mlp` argument of deg_solve_args). */
uint32_t deg_mlp_packed_size(void);
/* Columns of one partial-gradient row (DEG_MLP_NPARAM gradients, loss sum, row count). */
int32_t deg_mlp_partial_cols(void);

/* FP32 master parameters (DEG_MLP_NPARAM floats): W1[6][32], b1[32], W2[32][32], b2[32], w3[32], b3,
 * with h1 = relu(x' W1 + b1), h2 = relu(h1 W2 + b2), z = h2 . w3 + b3 and x' = (x - mean) * inv_std
 * (norm = [mean[6], inv_std[6]]). */

/* Partial gradients of the mean squared error over `nrows` step-log rows (optionally gathered
 * through idx) into partial[nblocks][deg_mlp_partial_cols()]; the target is the correction to the
 * PI factor that would have hit the error ratio rho: y = clamp(col6 + expo * (log2_rho - col7),
 * -ycap, ycap) - col8. */
int deg_mlp_grad(const float *rows, const int32_t *idx, int64_t nrows, const float *params, const float *norm,
                 float expo, float log2_rho, float ycap, float *partial, int32_t nblocks, void *stream);
/* One Adam step from the partial gradients (summed in block order); loss_out (optional, device)
 * receives the batch's mean squared error. step is 1-based. */
int deg_mlp_adam(float *params, float *m, float *v, const float *partial, int32_t nblocks, int32_t step, float lr,
                 float beta1, float beta2, float eps, float *loss_out, void *stream);
/* FP32 forward pass over n feature rows of leading dimension ld (>= 6). */
int deg_mlp_forward(const float *params, const float *norm, const float *x, int64_t ld, int64_t n, float *out,
                    void *stream);
/* Pack FP32 master parameters into the BF16 matrix-core layout the solver reads (device). */
int deg_mlp_pack(const float *params, const float *norm, float zlo, float zhi, void *packed, void *stream);
/* The solver's matrix-core evaluation over n feature rows (BF16 weights and activations, FP32
 * accumulation), for tests and measurement. */
int deg_mlp_infer(const void *packed, const float *x, int64_t ld, int64_t n, float *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* DEGHIP_H */
