/*
 * pelagos_relaybp.h -- C ABI of the MI355X Relay-BP-S decoder library
 * (libpelagos_relaybp.so, built from pelagos_amd/csrc/relaybp.hip).
 *
 * Plain C: raw pointers, sizes and an opaque graph handle; no torch types.
 * Device pointers are HIP device addresses on the graph's device; `stream`
 * is a hipStream_t passed as void* (NULL = default stream).
 *
 * Ownership: the caller owns every per-decode buffer (syndromes, priors,
 * gamma tables, outputs, workspace); the library owns only the graph layout
 * it builds in pel_graph_create.  All entry points return 0 on success and a
 * negative code on failure, with a message from pel_last_error() (per thread).
 *
 * The reference (quantum-supercomputing/pelagos) has no C ABI: its boundary is
 * the Python RelayBPDecoder launching Triton kernels.  Each entry point below
 * names the reference code it replaces.
 */
#ifndef PELAGOS_RELAYBP_H
#define PELAGOS_RELAYBP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PEL_ABI_VERSION 11

/* status codes */
#define PEL_OK 0
#define PEL_EINVAL (-1)    /* bad argument: maps to the reference's ValueError */
#define PEL_EDEVICE (-2)   /* HIP runtime error: maps to RuntimeError */
#define PEL_ENOMEM (-3)
#define PEL_EBUSY (-4)     /* a serving queue has no free entry (pel_serve_publish) */

typedef struct pel_graph pel_graph;

/* Learned memory-strength proposal: a per-variable MLP 8 -> 32 -> 32 -> 1
 * (relu, relu, then a logit z).  Each variable's output is a shift
 * (hi - lo) (sigmoid(z) - 1/2) of a learned leg's uniform draw u:
 * gamma = clamp(u + shift, lo, hi).  Parameters are one fp32 vector:
 * W1[32][8], b1[32], W2[32][32], b2[32], w3[32], b3. */
#define PEL_MLP_FEATURES 8
#define PEL_MLP_HIDDEN 32
#define PEL_MLP_PARAMS (32 * 8 + 32 + 32 * 32 + 32 + 32 + 1)

/* Relay-BP-S schedule parameters (RelayBPDecoder.__init__, decoder.py:271-307;
 * per-call overrides of decode(), decoder.py:2317-2361). */
typedef struct {
    int32_t pre_iter;        /* T0 (or decode(pre_iter_override=...)) */
    int32_t num_sets;        /* R: relay legs */
    int32_t set_max_iter;    /* Tr: iterations per relay leg */
    int32_t stop_nconv;      /* S: solutions to collect */
    int32_t plain;           /* algo == "plain": gamma 0 in the pre-iter leg */
    float gamma0;            /* pre-iter gamma for algo == "relay" */
    int32_t use_alpha;       /* normalized_min_sum_alpha is not None */
    int32_t use_beta;        /* offset_min_sum_beta is not None */
    float beta;              /* offset */
    const float *alpha_t;    /* device [max(T0,Tr)] alpha per in-leg iteration, or NULL */
    int32_t last_leg;        /* unconverged_correction == "last_leg" */
    int32_t skip_relay_legs; /* decode(skip_relay_legs=True) */
    int32_t leg_step;        /* 0: the whole schedule.  k >= 1 (plans made with PEL_PLAN_LEGWISE): only leg
                                k - 1 (0 the pre-iter leg, l the relay leg l - 1) for the shots still short
                                of stop_nconv solutions, resuming what the call with k - 1 on the same
                                buffers and workspace parked; every shot's outputs are as if the decode
                                ended after that leg.  Calls k = 1, 2, ... are the reference's
                                leg-synchronous schedule, whose fast_max_unconverged_pct exit checks the
                                found counts between legs (decoder.py:2890-2937, 3655-3835). */
} pel_params;

/* One batch of shots.  Shapes: B shots, C checks, V variables, R = num_sets. */
typedef struct {
    int32_t B;
    const uint8_t *syndromes;    /* [B,C] 0/1 */
    const float *prior_llr;      /* [V], or [B,V] when prior_batched (decode(prior_llr=...)) */
    const float *wj;             /* selection weights, same shape as prior_llr */
    int32_t prior_batched;
    const float *gamma_legs;     /* [R,V], or [R,B,V] when gamma_per_row; row l = relay leg l */
    int32_t gamma_per_row;
    int32_t gamma_by_position;   /* [R,V] rows already in the graph's variable order
                                    (pel_graph_permute_vars); 0 = caller's order */
    const float *init_gamma;     /* [B,V] pre-iter gamma (decode(init_gamma=...)) or NULL */
    const float *init_posterior; /* [B,V] (decode(init_posterior=...)) or NULL */
    uint8_t *errors;             /* out [B,V] */
    float *weights;              /* out [B]: best solution weight, +inf if none */
    int32_t *iterations;         /* out [B]: BP iterations run */
    int32_t *found_count;        /* out [B]: valid solutions found */
    float *last_posterior;       /* out [B,V] or NULL (decode(return_last_posterior=True)) */
    void *workspace;             /* device scratch of pel_plan.workspace_bytes */
    int64_t *shot_clock;         /* out [B,3] or NULL: device wall clock (100 MHz ticks) when the
                                    shot started and ended, and the workgroup that finished it */
    void *ev_begin;              /* hipEvent_t or NULL: recorded on the stream right before the */
    void *ev_end;                /* decode kernel and right after it (kernel-only timing) */
    /* Learned memory strengths (plans made with PEL_PLAN_LEARNED; no counterpart in the reference,
     * whose plug points for them are explicit_gammas / init_gamma, decoder.py:480-491, 2322). */
    const float *mlp;            /* network parameters [PEL_MLP_PARAMS], or NULL.  The network
                                    runs once per shot, on the state after the pre-iter leg when
                                    that state satisfies every check, and its per-variable shifts
                                    apply to those of relay legs 0 .. learned_legs-1 that start
                                    from a valid solution (the others keep their uniform rows) */
    int32_t learned_legs;
    float *feat_out;             /* out [B][learned_legs][V][8] or NULL: the network's features at
                                    the start of each of those legs (also without parameters) */
    uint8_t *hard_out;           /* out [B][learned_legs][V] or NULL: hard decision there; rows of
                                    legs a shot never reached are left untouched */
    float *gamma_out;            /* out [B][learned_legs][V] or NULL: the strengths each of those
                                    legs used (shifted or its uniform row) */
    float learned_lo, learned_hi; /* strength interval: a learned leg uses clamp(u + (hi - lo)
                                    (sigmoid(z) - 1/2), lo, hi) with u that leg's gamma_legs row */
    float learned_last_damp;     /* in [0, 1]: the last learned leg (learned_legs - 1) takes the shift
                                    times 1 - learned_last_damp (0: the full shift; 1: its uniform row) */
} pel_batch;

/* Launch plan for one (graph, B) pair: chosen by pel_plan_decode. */
typedef struct {
    int32_t block;           /* threads per workgroup (one shot per workgroup at a time) */
    int32_t grid;            /* persistent workgroups */
    int32_t placement;       /* bit0: variable state in LDS, bit1: check state in LDS,
                                bit2: sign words in LDS; a clear bit places it in workspace */
    int32_t lds_bytes;       /* dynamic LDS per workgroup */
    int64_t spill_stride;    /* workspace bytes per workgroup: state not in LDS, or (placements 7 and 15)
                                the copy of the running shot's first check pass; with learned
                                strengths also a shifted leg's strengths (4 V bytes) */
    int64_t workspace_bytes; /* total workspace the caller must provide */
    int32_t requeue_below;   /* shots whose first leg converged in fewer iterations are
                                re-queued behind all first legs (their posterior is parked in
                                the workspace); 0 = never.  Lowering it to 0 after planning is
                                allowed; raising it needs a plan made with it. */
    int32_t learned;         /* planned with PEL_PLAN_LEARNED */
    int32_t legwise;         /* planned with PEL_PLAN_LEGWISE */
} pel_plan;

/* pel_plan_decode flags */
#define PEL_PLAN_LEARNED 1   /* reserve LDS for learned memory strengths (fast placements only) */
#define PEL_PLAN_LEGWISE 2   /* workspace for leg-stepped decodes (pel_params.leg_step) */

/* ABI version (PEL_ABI_VERSION). */
int pel_abi_version(void);

/* Last error message of the calling thread ("" if none). */
const char *pel_last_error(void);

/* Build the device layout of a check matrix on the current HIP device.
 * chk_ptr [C+1], chk_var [E] are host arrays of H in CSR form; edge ids are
 * CSR positions.  Replaces CSRGraph._build_edge_representation
 * (src/pelagos/graph.py:101-211). */
int pel_graph_create(const int32_t *chk_ptr, const int32_t *chk_var, int32_t C, int32_t V,
                     int32_t E, pel_graph **out);

/* Free a graph (device layout included). */
int pel_graph_destroy(pel_graph *g);

/* Sizes of a graph: C, V, E and the per-shot state bytes (variables, checks,
 * sign words).  No counterpart in the reference beyond CSRGraph.get_stats
 * (graph.py:227-245). */
int pel_graph_info(const pel_graph *g, int32_t *C, int32_t *V, int32_t *E,
                   int64_t *var_state_bytes, int64_t *chk_state_bytes, int64_t *sign_bytes);

/* dst[r][v] = src[r][perm[v]] for r < rows: per-variable rows (e.g. a [R,V]
 * gamma table) into the graph's internal variable order, on `stream`.  No
 * counterpart in the reference; lets pel_decode read broadcast gamma rows
 * without a gather (pel_batch.gamma_by_position = 1). */
int pel_graph_permute_vars(const pel_graph *g, const float *src, float *dst, int64_t rows, void *stream);

/* Choose the launch plan for B shots.  block_override > 0 forces the block
 * size; placement_override >= 0 forces the placement bits (a state array that
 * does not fit the LDS budget is never placed there).  Replaces the
 * launch-configuration resolution of RelayBPDecoder._resolve_configs /
 * _get_tuned_meta (decoder.py:905-1222). */
int pel_plan_decode(const pel_graph *g, int32_t B, int32_t block_override,
                    int32_t placement_override, int32_t flags, pel_plan *out);

/* Decode a batch with one persistent launch: every shot runs the full
 * Relay-BP-S schedule (pre-iter leg, then relay legs until stop_nconv
 * solutions) on one workgroup at a time, and idle workgroups take the next
 * shot from a device queue: first every shot's pre-iter leg (heaviest
 * syndromes first), then the relay legs of the shots re-queued after a quick
 * first leg (pel_plan.requeue_below).  Results do not depend on the order.
 * Replaces RelayBPDecoder._decode_loop_single_gpu (decoder.py:3394-3873) and
 * the kernels it launches (kernels.py:446-1350). */
int pel_decode(const pel_graph *g, const pel_params *p, const pel_batch *b, const pel_plan *plan,
               void *stream);

/* Workspace bytes pel_decode_noalloc holds per GPU (32 MiB). */
#define PEL_NOALLOC_MAX_BYTES (32LL << 20)

/* pel_decode without a caller workspace and without any allocation: the same
 * launches, with the workspace in a library-owned device array of
 * PEL_NOALLOC_MAX_BYTES on each GPU (held until the process ends).  b->workspace
 * must be NULL; every other buffer is the caller's as in pel_decode.  A decode
 * whose workspace (the queue, order and re-queue arrays for B shots, plus
 * plan->spill_stride x plan->grid) exceeds PEL_NOALLOC_MAX_BYTES is refused with
 * PEL_EINVAL before anything is launched; a plan with requeue_below = 0 needs
 * no parked posteriors (about 7.4 MB for 4096 shots of the 936 x 8785 code).
 * The kernels reset every counter they rely on, so calls need no clearing in
 * between.  All calls on one GPU share its array: calls whose launches
 * overlap on one GPU (other streams, other threads) race and must be ordered
 * by the caller.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
int pel_decode_noalloc(const pel_graph *g, const pel_params *p, const pel_batch *b, const pel_plan *plan,
                       void *stream);

/* ---- several batches, of one code or several, in one persistent launch ----
 * A job is one batch of one graph: its own pel_params and pel_batch (buffers
 * and a workspace of pel_jobs_plan.workspace_bytes[i] bytes).  One launch
 * decodes up to PEL_MAX_JOBS jobs, taking them in order: a small code's shots
 * run one per wave (sixteen shots per workgroup at a time), a large code's one
 * per workgroup, and a workgroup moves to the next job as soon as the current
 * one has no shot left, so the slow last shots of one job overlap the next
 * jobs' work.  Each job's results equal a pel_decode of that batch (no
 * re-queue is used: the later jobs fill the tail).
 * No single counterpart in the reference, which decodes one batch of one code
 * per call; it replaces the difficult-case re-queue and the per-problem launch
 * configuration around the decode loop (decoder.py:1634-1683, 2890-2937,
 * 905-969) for callers with several batches in hand. */
#define PEL_MAX_JOBS 64

typedef struct {
    int32_t n_jobs;
    int32_t pad;
    int64_t table_bytes;                   /* device job table pel_decode_jobs needs */
    int32_t block;                         /* threads per workgroup */
    int32_t grid;                          /* persistent workgroups */
    int32_t lds_bytes;                     /* dynamic LDS per workgroup */
    int32_t group[PEL_MAX_JOBS];           /* threads per shot of job i: 64 (a shot per wave) or block */
    int32_t placement[PEL_MAX_JOBS];       /* as pel_plan.placement, for one shot group */
    int32_t learned[PEL_MAX_JOBS];         /* planned with PEL_PLAN_LEARNED */
    int32_t job_lds[PEL_MAX_JOBS];         /* LDS bytes of one shot group of job i */
    int64_t spill_stride[PEL_MAX_JOBS];    /* workspace bytes per shot group of job i */
    int64_t workspace_bytes[PEL_MAX_JOBS]; /* workspace job i's pel_batch must provide */
} pel_jobs_plan;

/* Plan a launch over n jobs: graphs[i] with shots[i] shots and plan flags
 * flags[i] (PEL_PLAN_LEARNED, or 0; flags may be NULL). */
int pel_plan_jobs(const pel_graph *const *graphs, const int32_t *shots, const int32_t *flags, int32_t n,
                  pel_jobs_plan *out);

/* Decode n jobs (graphs[i], params[i], batches[i]) with one launch on
 * `stream`.  Job i's batch must have the shot count and workspace its plan was
 * made for; `table` is device memory of plan->table_bytes for the jobs'
 * descriptors (written on the stream before the launch, so it may be reused
 * by the next call on the same stream); batches[0].ev_begin / ev_end, if set,
 * are recorded around the decode launch. */
int pel_decode_jobs(const pel_graph *const *graphs, const pel_params *params, const pel_batch *batches, int32_t n,
                    const pel_jobs_plan *plan, void *table, void *stream);

/* ---- serving: a persistent launch that takes shots as they are published --
 * The host publishes units of work into a queue in host memory that the
 * device reads: one shot of a shot-per-workgroup job, or up to
 * PEL_SERVE_WAVE_SHOTS shots (one request's) of a shot-per-wave job.  The
 * serving engine's workgroups claim units in publication order while the
 * launch runs, so shots published after the launch started join it at once,
 * and a slow shot holds one workgroup, not the shots published after it.  Each
 * unit names its own syndrome rows and result buffers (device memory); once
 * the engine has written a unit's results it sets the unit's done flag in host
 * memory.  A workgroup that finds no unit for idle_us microseconds, or finds
 * no unit left while the stop flag is set, leaves, so every launch ends by
 * itself; the caller
 * launches the engine again when it publishes work and the last launch has
 * ended (units published meanwhile wait in the queue).  Each shot's results
 * equal a pel_decode of that shot alone.
 * No single counterpart in the reference: it replaces the persistent worker
 * pool with ordered reassembly (process.py:114-254) and the per-batch decode
 * call (decoder.py:2317-2449) for callers whose shots arrive over time. */
#define PEL_SERVE_WAVE_SHOTS 64

typedef struct {
    int32_t job;              /* job index of the serving plan */
    int32_t count;            /* shots: 1 for a shot-per-workgroup job, 1..PEL_SERVE_WAVE_SHOTS for a shot-per-wave
                                 job (a unit that breaks this is marked done without being decoded) */
    int64_t syn_row_stride;   /* bytes from one shot's syndrome row to the next (C when contiguous) */
    const uint8_t *syndromes; /* device [count, C], 0/1 bytes, strided as given */
    uint8_t *errors;          /* device [count, V] */
    float *weights;           /* device [count] */
    int32_t *iterations;      /* device [count] */
    int32_t *found_count;     /* device [count] */
    int64_t syn_col_stride;   /* bytes from one check's syndrome bit to the next (1 when contiguous) */
} pel_serve_unit;

typedef struct pel_serve_queue pel_serve_queue;

/* A queue of `capacity` units: the ring, the published count and the stop flag
 * in host memory the device reads, the done flags in host memory the device
 * writes, the claim counter in device memory (on the current device). */
int pel_serve_queue_create(int32_t capacity, pel_serve_queue **out);
int pel_serve_queue_destroy(pel_serve_queue *q);

/* Host address of the done flags: flag[k % capacity] >= k + 1 once unit k
 * (counted from 0 in publication order) is done (an entry is reused only after
 * its unit is done, so a later unit's number there also means done).  Also the
 * capacity and the units published so far. */
int pel_serve_queue_info(const pel_serve_queue *q, const int64_t **done, int32_t *capacity, int64_t *published);

/* Publish n units; *first (may be NULL) receives the first one's number.
 * PEL_EBUSY, publishing nothing, if an entry they need still holds a unit
 * that is not done.  The units' buffers must stay valid and their syndromes
 * written (the producing work complete) until their done flags are set. */
int pel_serve_publish(pel_serve_queue *q, const pel_serve_unit *units, int32_t n, int64_t *first);

/* Units of [first, first + n) that are done (a negative status on bad arguments). */
int64_t pel_serve_done_count(const pel_serve_queue *q, int64_t first, int64_t n);

/* Set (1) or clear (0) the stop flag: with it set, the engine's workgroups
 * finish every unit published so far and leave without waiting for more. */
int pel_serve_set_stop(pel_serve_queue *q, int32_t stop);

/* Plan a serving engine over n jobs (graphs[i], flags[i] as in pel_plan_jobs).
 * Each job's pel_batch holds staging rows, not shots: B =
 * grid x (16 for a shot-per-wave job, else 1) syndrome and result rows the
 * engine stages units through, with the workspace of workspace_bytes[i]. */
int pel_plan_serve(const pel_graph *const *graphs, const int32_t *flags, int32_t n, pel_jobs_plan *out);

/* Launch the serving engine on `stream` over the jobs of a pel_plan_serve plan;
 * `table` as in pel_decode_jobs.  It returns at once; the launch ends when its
 * workgroups have found no unit for idle_us microseconds, or none left with
 * the stop flag set.  The queue, the jobs' buffers and the table must outlive
 * the launch. */
int pel_serve_launch(const pel_graph *const *graphs, const pel_params *params, const pel_batch *batches, int32_t n,
                     const pel_jobs_plan *plan, void *table, pel_serve_queue *q, int64_t idle_us, void *stream);

/* ---- float16 and float64 decodes ---------------------------------------------
 * RelayBPDecoder(data_type="float16" | "float64") (decoder.py:402-423; the kernels' dtype_enum
 * 0 / 2, e.g. kernels.py:1115-1126, 1313-1315, 680-715).  The batch's prior_llr, wj, gamma_legs
 * (caller's variable order, gamma_by_position = 0), init_gamma, init_posterior, weights and
 * last_posterior hold elements of that type; pel_params keeps its float32 scalars (gamma0, beta,
 * alpha_t), as the reference passes them.  The reference's own per-edge formulation: mu and nu of
 * every edge in a workspace slot per workgroup, one shot per workgroup at a time, the Relay-BP-S
 * schedule of pel_decode, leg-stepped decodes (pel_params.leg_step, plans made with
 * PEL_PLAN_LEGWISE) included; no learned strengths.  PEL_DTYPE_F32 is accepted too and gives
 * pel_decode's results. */
#define PEL_DTYPE_F16 0
#define PEL_DTYPE_F32 1
#define PEL_DTYPE_F64 2

/* Launch size for B shots of element type dtype: *grid workgroups and the workspace bytes
 * pel_decode_typed needs (a queue word, one state slot per workgroup and, with PEL_PLAN_LEGWISE in
 * flags, every shot's parked state for leg-stepped decodes). */
int pel_plan_typed(const pel_graph *g, int32_t B, int32_t dtype, int32_t flags, int32_t *grid,
                   int64_t *workspace_bytes);

/* Decode a batch whose floating-point buffers hold elements of type dtype, with grid workgroups
 * (1 .. the planned grid) and b->workspace of the planned size (planned with PEL_PLAN_LEGWISE for
 * p->leg_step > 0; calls k = 1, 2, ... must use the same buffers, workspace and grid).  Replaces
 * the dtype_enum 0 and 2 instantiations of the kernels RelayBPDecoder._decode_loop_single_gpu launches
 * (decoder.py:3394-3873, kernels.py:446-1350). */
int pel_decode_typed(const pel_graph *g, const pel_params *p, const pel_batch *b, int32_t dtype, int32_t grid,
                     void *stream);

/* ---- learned memory strengths (no counterpart in the reference) ----------
 * Training entry points for the MLP above.  x [n][8] features, y [n] soft
 * targets for sigmoid(z) in [0,1], w [n] sample weights.  All arrays are
 * device memory; work is ordered on `stream`. */

/* PEL_MLP_PARAMS. */
int pel_mlp_param_count(void);

/* Last error of the training entry points on the calling thread. */
const char *pel_mlp_last_error(void);

/* Workspace bytes pel_mlp_grad needs. */
int64_t pel_mlp_grad_workspace_bytes(void);

/* grad = d/dparams of mean_i w_i (softplus(z_i) - y_i z_i) and *loss its value
 * (loss may be NULL).  The sum runs in a fixed order (bit-reproducible). */
int pel_mlp_grad(const float *params, const float *x, const float *y, const float *w, int64_t n, float *grad,
                 float *loss, void *workspace, int64_t workspace_bytes, void *stream);

/* One Adam step (step counts from 1) on params with moment buffers m1, m2. */
int pel_mlp_adam(float *params, float *m1, float *m2, const float *grad, float lr, float beta1, float beta2,
                 float eps, int32_t step, void *stream);

/* z [n] = the network's logits for features x [n][8] (fp32). */
int pel_mlp_forward(const float *params, const float *x, int64_t n, float *z, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PELAGOS_RELAYBP_H */
