/*
 * relaybp_oracle.c -- CPU restatement of Relay-BP-S min-sum decoding.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP decoder in
 * pelagos_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls it.
 *
 * It restates, per shot and in plain fp32 C, the algorithm of the reference
 * (quantum-supercomputing/pelagos) throughput/flood-fill path:
 *
 *   schedule       src/pelagos/decoder.py:3394-3873  (_decode_loop_single_gpu)
 *   one iteration  src/pelagos/decoder.py:2610-2875  (_run_bp_iteration)
 *   C2V min-sum    src/pelagos/kernels.py:1051-1215  (c2v_min_sum_btile_kernel)
 *   V2C + gamma    src/pelagos/kernels.py:1219-1350  (v2c_and_gamma_btile_kernel)
 *   parity         src/pelagos/kernels.py:509-569    (parity_from_hard_compact_with_zero_kernel)
 *   select         src/pelagos/kernels.py:608-741    (check_and_select_kernel)
 *   freeze         src/pelagos/kernels.py:745-847    (freeze_finished_lanes_kernel)
 *   message init   src/pelagos/kernels.py:903-1000   (init_messages_with_priors_EB_kernel)
 *
 * It deliberately keeps the reference's explicit per-edge message arrays
 * (mu, nu) and its min1/min2/tie-count formulation of the check update, so it
 * is an independent formulation from the GPU kernel (which keeps a compressed
 * per-check state).  Each shot is decoded independently: the reference's lane
 * freezing makes every lane's result independent of the batch
 * (tests/test_decoder_fixes.py:184-268), so a per-shot loop is the same
 * algorithm.
 *
 * Arithmetic conventions shared with the HIP kernel (both built with
 * -ffp-contract=off):
 *   - lam_mixed = (1-g)*lam0 + g*post_prev, two roundings, no FMA
 *     (kernels.py:1309; the relay-bp Rust crate does not contract either);
 *   - the variable sum runs over a variable's edges in ascending edge id
 *     (= ascending check index), starting from 0.0f (kernels.py:1317-1325
 *     with BLOCK_SIZE fixed at 1, autotune.py:635);
 *   - the solution weight sum_j wj*[post_j <= 0] is accumulated in double and
 *     rounded once to float (exact, hence order-free, for these inputs).
 *
 * Parity status: pinned by tests/test_oracle_pins.py (closed forms, tree-code
 * exactness, brute-force ML, schedule invariants, relay-leg schedule closed
 * forms, planted-bug mutants 1-10).  The learned network's leg-0 features
 * (leg0_features) are pinned by tests/test_learned_gamma.py against the decode
 * state (closed forms, the pre-iter posterior, violated-check fractions, the
 * mean incoming message with a plain pre-iter leg) and against the pre-iter
 * leg's final messages (mu0, nu0) through the single-check update.
 *
 * Element type (ORC_ELEM, default 32): the same restatement for data_type
 * "float64" (64: every stored and computed value a double) and "float16" (16:
 * mu, nu, posteriors, priors, strengths and weights stored as IEEE half, check
 * minima and variable sums in fp32, the memory mix in half arithmetic), as the
 * reference's kernels do for dtype_enum 2 and 0 (kernels.py:1115-1126,
 * 1299-1350, 676-715).  Half values are held as their 16-bit patterns and
 * every half operation is computed exactly in double and rounded to nearest
 * even (to_half).  Parity status of these two: the conversions are pinned
 * against numpy's float16; the fp64 and fp16 arithmetic against the fp32
 * restatement on tree codes (same ML answers) and against the GPU; bit-level
 * agreement with the reference's Triton fp16 kernels is unpinned (they need a
 * GPU and Triton here).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Compile-time planted mistakes, used only by the mutation test that shows the
 * pinning checks reject a wrong oracle.  0 = the real oracle. */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

#ifndef ORC_ELEM
#define ORC_ELEM 32
#endif

#if ORC_ELEM == 16
/* IEEE half as its bit pattern. */
typedef uint16_t elem_t;
typedef float real_t;
static double from_half(uint16_t h)
{
    const int s = (h >> 15) & 1, e = (h >> 10) & 31, m = h & 1023;
    double v;
    if (e == 0) v = ldexp((double)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = ldexp((double)(m | 1024), e - 25);
    return s ? -v : v;
}
/* x rounded to the nearest half (ties to even); beyond the largest finite half
 * (65504) by half an ulp or more: infinity. */
static uint16_t to_half(double x)
{
    const uint16_t sgn = signbit(x) ? 0x8000u : 0u;
    const double a = fabs(x);
    if (isnan(x)) return sgn | 0x7e00u;
    if (a >= 65520.0) return sgn | 0x7c00u;
    if (a < ldexp(1.0, -14)) return sgn | (uint16_t)nearbyint(a * 16777216.0); /* units of 2^-24 */
    int e;
    frexp(a, &e);                                  /* a in [2^(e-1), 2^e) */
    double r = nearbyint(ldexp(a, 11 - e));        /* 1024 .. 2048 */
    if (r == 2048.0) { r = 1024.0; ++e; }
    const int E = e - 1 + 15;
    if (E >= 31) return sgn | 0x7c00u;
    return sgn | (uint16_t)(E << 10) | (uint16_t)((int)r - 1024);
}
#define LD(x) ((real_t)from_half(x))
#define ST(x) to_half((double)(x))
#define ELEM_ZERO ((elem_t)0)
#define RABS(x) fabsf(x)
#define RMAX(a, b) fmaxf(a, b)
#define NO_SOLUTION 60000.0                        /* kernels.py:713-714 */
/* (1 - g) lam0 + g post with half operands: each operation rounded to half */
static elem_t mix(elem_t g, elem_t l0, elem_t post)
{
    const double omg = from_half(to_half(1.0 - from_half(g)));
    const double t1 = from_half(to_half(omg * from_half(l0)));
    const double t2 = from_half(to_half(from_half(g) * from_half(post)));
    return to_half(t1 + t2);
}
static double weight_round(double acc) { return from_half(to_half((double)(float)acc)); }
#elif ORC_ELEM == 64
typedef double elem_t;
typedef double real_t;
#define LD(x) (x)
#define ST(x) ((double)(x))
#define ELEM_ZERO 0.0
#define RABS(x) fabs(x)
#define RMAX(a, b) fmax(a, b)
#define NO_SOLUTION 1e20
static elem_t mix(elem_t g, elem_t l0, elem_t post) { return (1.0 - g) * l0 + g * post; }
static double weight_round(double acc) { return acc; }
#else
typedef float elem_t;
typedef float real_t;
#define LD(x) (x)
#define ST(x) ((float)(x))
#define ELEM_ZERO 0.0f
#define RABS(x) fabsf(x)
#define RMAX(a, b) fmaxf(a, b)
#define NO_SOLUTION 1e20
static elem_t mix(elem_t g, elem_t l0, elem_t post) { return (1.0f - g) * l0 + g * post; }
static double weight_round(double acc) { return (double)(float)acc; }
#endif

typedef struct {
    int32_t pre_iter;        /* T0 */
    int32_t num_sets;        /* R  */
    int32_t set_max_iter;    /* Tr */
    int32_t stop_nconv;      /* S  */
    int32_t plain;           /* algo == "plain": gamma = 0 in the pre-iter leg (decoder.py:3489-3491) */
    float gamma0;            /* pre-iter gamma for algo == "relay" (decoder.py:3496-3500) */
    int32_t use_alpha;       /* normalized_min_sum_alpha is not None (decoder.py:1712) */
    int32_t use_beta;        /* offset_min_sum_beta is not None (decoder.py:1713) */
    float beta;
    const float *alpha_t;    /* [max(T0,Tr)] alpha for in-leg iteration t (decoder.py:2649-2660) */
    int32_t last_leg;        /* unconverged_correction == "last_leg" (decoder.py:2877-2888) */
    int32_t skip_relay_legs; /* decode(skip_relay_legs=True) (decoder.py:3625) */
} orc_params;

typedef struct {
    int32_t B;
    const uint8_t *syndromes;     /* [B,C] */
    const elem_t *prior_llr;      /* [V] or [B,V] when prior_batched */
    int32_t prior_batched;
    const elem_t *gamma_legs;     /* [R,V] or [R,B,V] when gamma_per_row */
    int32_t gamma_per_row;
    const elem_t *init_gamma;     /* [B,V] or NULL */
    const elem_t *init_posterior; /* [B,V] or NULL */
    uint8_t *errors;              /* [B,V] out */
    elem_t *weights;              /* [B]   out */
    int32_t *iterations;          /* [B]   out */
    int32_t *found_count;         /* [B]   out */
    elem_t *last_posterior;       /* [B,V] out or NULL */
    float *features0;             /* [B,V,8] out or NULL (float32 build): the learned network's
                                     features at the start of relay leg 0, for shots that reach it */
    elem_t *mu0, *nu0;            /* [B,E] out or NULL: the messages when the pre-iter leg ends */
} orc_batch;

typedef struct {
    int32_t C, V, E;
    const int32_t *row_ptr;   /* [C+1] CSR over checks; edge id = CSR position (graph.py:112-165) */
    const int32_t *row_var;   /* [E]   variable of each edge */
    int32_t *var_ptr;         /* [V+1] */
    int32_t *var_edges;       /* [E]   edge ids grouped by variable, ascending */
} orc_graph;

/* Variable-major edge lists in ascending edge id (a stable version of the
 * argsort at graph.py:144-154). */
static int build_var_lists(orc_graph *g)
{
    g->var_ptr = (int32_t *)calloc((size_t)g->V + 1, sizeof(int32_t));
    g->var_edges = (int32_t *)malloc(sizeof(int32_t) * (size_t)(g->E > 0 ? g->E : 1));
    int32_t *fill = (int32_t *)calloc((size_t)g->V + 1, sizeof(int32_t));
    if (!g->var_ptr || !g->var_edges || !fill) { free(fill); return -1; }
    for (int32_t e = 0; e < g->E; ++e) g->var_ptr[g->row_var[e] + 1]++;
    for (int32_t j = 0; j < g->V; ++j) g->var_ptr[j + 1] += g->var_ptr[j];
    for (int32_t e = 0; e < g->E; ++e) {
        int32_t j = g->row_var[e];
        g->var_edges[g->var_ptr[j] + fill[j]++] = e;
    }
    free(fill);
    return 0;
}

typedef struct {
    elem_t *mu, *nu;         /* [E] variable->check and check->variable messages */
    elem_t *post;            /* [V] posterior_ratios */
    uint8_t *hard;           /* [V] hard_dec */
    uint8_t *best;           /* [V] best_errors */
    uint8_t *first_hd;       /* [V] first_leg_hard_dec */
    const elem_t *gamma;     /* [V] current leg's memory strengths (or NULL with gamma_const) */
    elem_t gamma_const;
    uint8_t *flips;          /* [V] hard-decision flips during the leg (counted when count_flips) */
    int32_t count_flips;
    int32_t leg_t;           /* the in-leg iteration at which the last leg ended */
    elem_t *nu2;             /* [E] scratch: the messages of the check pass after a converged leg */
} orc_lane;

/* C2V, kernels.py:1128-1215: per check, the parity of negative incoming
 * messages XOR the syndrome bit, and the minimum |mu| over the other edges. */
static void check_update(const orc_graph *g, const orc_params *p, const uint8_t *syn,
                         float alpha, orc_lane *L)
{
    for (int32_t i = 0; i < g->C; ++i) {
        int32_t r0 = g->row_ptr[i], r1 = g->row_ptr[i + 1];
        if (r1 <= r0) continue;
        int deg_is_one = (r1 - r0) == 1;
        int neg = 0, cnt1 = 0;
        real_t min1 = (real_t)1e30, min2 = (real_t)1e30; /* the reference's start and padding value:
                                                     magnitudes are clipped at 1e30 (kernels.py:1120-1126, 1146) */
#if ORC_PLANT == 10
        min1 = min2 = INFINITY;                   /* planted: no clip */
#endif
        for (int32_t e = r0; e < r1; ++e) {
            real_t m = LD(L->mu[e]);
            real_t a = RABS(m);
            neg ^= (m < 0);
            if (a < min1) { min2 = min1; min1 = a; cnt1 = 1; }
            else if (a == min1) { cnt1++; min2 = a; }
            else if (a < min2) { min2 = a; }
        }
        int par = (neg & 1) ^ (syn[i] & 1);
#if ORC_PLANT == 1
        par = neg & 1;                            /* planted: syndrome sign dropped */
#endif
        real_t sign_prod = par == 0 ? (real_t)1 : (real_t)-1;
        for (int32_t e = r0; e < r1; ++e) {
            real_t m = LD(L->mu[e]);
            real_t a = RABS(m);
            real_t out = (a == min1 && cnt1 == 1) ? min2 : min1;
#if ORC_PLANT == 2
            out = min1;                           /* planted: own edge not excluded */
#endif
            if (deg_is_one) out = 0;
            if (p->use_alpha) out = (real_t)alpha * out;
            if (p->use_beta) out = RMAX(out - (real_t)p->beta, (real_t)0);
            real_t sgn = m >= 0 ? (real_t)1 : (real_t)-1;
            L->nu[e] = ST((sign_prod * sgn) * out);
        }
    }
}

/* V2C with memory mixing, kernels.py:1299-1350. */
static void variable_update(const orc_graph *g, const elem_t *lam0, orc_lane *L)
{
    for (int32_t j = 0; j < g->V; ++j) {
        elem_t gj = L->gamma ? L->gamma[j] : L->gamma_const;
        elem_t lm = mix(gj, lam0[j], L->post[j]);
#if ORC_PLANT == 3
        lm = gj * lam0[j] + (1.0f - gj) * L->post[j];   /* planted: memory weights swapped */
#endif
        real_t s = 0;
        for (int32_t q = g->var_ptr[j]; q < g->var_ptr[j + 1]; ++q) s += LD(L->nu[g->var_edges[q]]);
        real_t post = LD(lm) + s;
        if (L->count_flips && ((LD(L->post[j]) <= 0) != (post <= 0)) && L->flips[j] < 255) L->flips[j]++;
        L->post[j] = ST(post);
        L->hard[j] = post <= 0;
        for (int32_t q = g->var_ptr[j]; q < g->var_ptr[j + 1]; ++q) {
            int32_t e = g->var_edges[q];
#if ORC_PLANT == 4
            L->mu[e] = lm + s;                    /* planted: extrinsic subtraction dropped */
#else
            L->mu[e] = ST(LD(lm) + (s - LD(L->nu[e])));
#endif
        }
    }
}

/* parity_from_hard_*: every check XOR of hard decisions equals its syndrome bit. */
static int all_checks_ok(const orc_graph *g, const uint8_t *syn, const uint8_t *hard)
{
    for (int32_t i = 0; i < g->C; ++i) {
        int par = 0;
        for (int32_t e = g->row_ptr[i]; e < g->row_ptr[i + 1]; ++e) par ^= hard[g->row_var[e]];
        if (par != (syn[i] & 1)) return 0;
    }
    return 1;
}

/* init_messages_with_priors_EB_kernel: nu = 0, mu[e] = lam0[var(e)]. */
static void init_messages(const orc_graph *g, const elem_t *lam0, orc_lane *L)
{
    for (int32_t e = 0; e < g->E; ++e) { L->mu[e] = lam0[g->row_var[e]]; L->nu[e] = ELEM_ZERO; }
}

typedef struct {
    double best_w;           /* a value of the element type (or +inf) */
    int32_t found;
    int32_t iters;
} orc_shot;

/* One leg: up to T flood iterations; the lane freezes at its first valid
 * solution of the leg (freeze_finished_lanes_kernel with in_leg=True and
 * sought = found + 1).  Returns 1 if the lane converged in this leg. */
static int run_leg(const orc_graph *g, const orc_params *p, const uint8_t *syn,
                   const elem_t *lam0, const elem_t *wj, int32_t T, orc_lane *L, orc_shot *st)
{
    for (int32_t t = 0; t < T; ++t) {
        float alpha = p->use_alpha ? p->alpha_t[t] : 1.0f;
        check_update(g, p, syn, alpha, L);
        variable_update(g, lam0, L);
        st->iters++;
        L->leg_t = t;
        if (all_checks_ok(g, syn, L->hard)) {
            /* check_and_select_kernel: weight over bits (post <= 0) with wj. */
            double acc = 0.0;
            for (int32_t j = 0; j < g->V; ++j)
                if (LD(L->post[j]) <= 0) acc += (double)LD(wj[j]);
            double w = weight_round(acc);
#if ORC_PLANT == 7
            if (st->best_w >= NO_SOLUTION) {                  /* planted: the first solution is never replaced */
#elif ORC_PLANT == 9
            if (st->best_w >= NO_SOLUTION || w <= st->best_w) { /* planted: a tie replaces the best */
#else
            if (st->best_w >= NO_SOLUTION || w < st->best_w) {
#endif
                st->best_w = w;
                for (int32_t j = 0; j < g->V; ++j) L->best[j] = LD(L->post[j]) <= 0;
            }
            st->found++;
            /* freeze: hard_dec <- best_errors (kernels.py:806-817) */
            memcpy(L->hard, L->best, (size_t)g->V);
            return 1;
        }
    }
    return 0;
}

#if ORC_ELEM == 32
static float clamp4(float x) { return fminf(fmaxf(x, -4.0f), 4.0f); }

/* The learned network's features of every variable at the start of relay leg
 * 0, from the state the pre-iter leg ended in (no counterpart in the
 * reference; the decode kernel's feature pass, leg_features_t in
 * pelagos_amd/csrc/relaybp.hip, computes them from its compressed state):
 *   0 prior LLR / 8     1 posterior / 16     2 flips in the pre-iter leg / 8
 *   3 S / deg / 8 with S the sum of the incoming messages of the last variable update
 *   4 the smallest incoming message / 8 (0 for a variable in no check)
 *   5 deg / 8           6 fraction of the variable's checks violated     7 leg / 4 = 0
 * (1, 3, 4 clamped to [-4, 4]).  The incoming messages of feature 4 are the
 * ones a next variable update would read: after a leg that converged before
 * its last iteration, those of one more check pass (its alpha, from the final
 * mu); otherwise the last check pass's. */
static void leg0_features(const orc_graph *g, const orc_params *p, const uint8_t *syn, const elem_t *lam0,
                          int converged, orc_lane *L, float *f)
{
    const elem_t *nu4 = L->nu;
    if (converged && L->leg_t + 1 < p->pre_iter) {
        elem_t *keep = L->nu;
        L->nu = L->nu2;
        check_update(g, p, syn, p->use_alpha ? p->alpha_t[L->leg_t + 1] : 1.0f, L);
        L->nu = keep;
        nu4 = L->nu2;
    }
    for (int32_t j = 0; j < g->V; ++j) {
        const int32_t q0 = g->var_ptr[j], q1 = g->var_ptr[j + 1], deg = q1 - q0;
        float sum = 0.0f, mn = INFINITY;
        int32_t uns = 0;
        for (int32_t q = q0; q < q1; ++q) {
            const int32_t e = g->var_edges[q];
            sum += L->nu[e];
            mn = fminf(mn, nu4[e]);
            if (!converged) {
                /* the edge's check: e lies in [row_ptr[i], row_ptr[i+1]) */
                int32_t lo = 0, hi = g->C;
                while (hi - lo > 1) { const int32_t mid = (lo + hi) / 2; if (g->row_ptr[mid] <= e) lo = mid; else hi = mid; }
                int par = syn[lo] & 1;
                for (int32_t r = g->row_ptr[lo]; r < g->row_ptr[lo + 1]; ++r) par ^= L->post[g->row_var[r]] <= 0;
                uns += par;
            }
        }
        const float inv = deg > 0 ? 1.0f / (float)deg : 0.0f;
        float *o = f + 8 * (size_t)j;
        o[0] = lam0[j] * 0.125f;
        o[1] = clamp4(L->post[j] * 0.0625f);
        o[2] = (float)L->flips[j] * 0.125f;
        o[3] = clamp4(sum * inv * 0.125f);
        o[4] = deg > 0 ? clamp4(mn * 0.125f) : 0.0f;
        o[5] = (float)deg * 0.125f;
        o[6] = converged ? 0.0f : (float)uns * inv;
        o[7] = 0.0f;
    }
}
#endif

static void decode_shot(const orc_graph *g, const orc_params *p, const orc_batch *bt,
                        int32_t b, orc_lane *L, elem_t *wj_buf)
{
    const int32_t V = g->V;
    const uint8_t *syn = bt->syndromes + (size_t)b * g->C;
    const elem_t *lam0 = bt->prior_llr + (bt->prior_batched ? (size_t)b * V : 0);
    /* wj = lam0 with non-finite entries zeroed (decoder.py:736-753) */
    for (int32_t j = 0; j < V; ++j) wj_buf[j] = isfinite(LD(lam0[j])) ? lam0[j] : ELEM_ZERO;

    /* state reset (decoder.py:3430-3464; _reset_reused_tensors 1557-1592) */
    for (int32_t j = 0; j < V; ++j)
        L->post[j] = bt->init_posterior ? bt->init_posterior[(size_t)b * V + j] : lam0[j];
    memset(L->hard, 0, (size_t)V);
    memset(L->best, 0, (size_t)V);
    memset(L->first_hd, 0, (size_t)V);
    orc_shot st = {INFINITY, 0, 0};
    const int32_t S = p->stop_nconv;

    /* pre-iter leg (decoder.py:3470-3593) */
    init_messages(g, lam0, L);
    if (p->plain) { L->gamma = NULL; L->gamma_const = ELEM_ZERO; }
    else if (bt->init_gamma) { L->gamma = bt->init_gamma + (size_t)b * V; }
    else { L->gamma = NULL; L->gamma_const = ST(p->gamma0); }
    memset(L->flips, 0, (size_t)V);
    L->count_flips = bt->features0 != NULL;
    int pre_ok = 0;
    if (st.found < S) pre_ok = run_leg(g, p, syn, lam0, wj_buf, p->pre_iter, L, &st);
    L->count_flips = 0;
    if (bt->mu0) memcpy(bt->mu0 + (size_t)b * g->E, L->mu, sizeof(elem_t) * (size_t)g->E);
    if (bt->nu0) memcpy(bt->nu0 + (size_t)b * g->E, L->nu, sizeof(elem_t) * (size_t)g->E);
#if ORC_ELEM == 32
    if (bt->features0 && !p->skip_relay_legs && p->num_sets > 0 && st.found < S)
        leg0_features(g, p, syn, lam0, pre_ok, L, bt->features0 + (size_t)b * V * 8);
#else
    (void)pre_ok;
#endif
    memcpy(L->first_hd, L->hard, (size_t)V);            /* decoder.py:3597-3598 */

    if (!p->skip_relay_legs) {
        for (int32_t leg = 0; leg < p->num_sets; ++leg) { /* decoder.py:3672 */
            if (st.found >= S) break;                     /* lane stays frozen */
            int32_t row = leg;
#if ORC_PLANT == 8
            row = (leg + 1) % p->num_sets;                /* planted: leg l reads the next leg's row */
#endif
            L->gamma = bt->gamma_legs +
                       (bt->gamma_per_row ? ((size_t)row * bt->B + b) * V : (size_t)row * V);
#if ORC_PLANT != 5
            init_messages(g, lam0, L);                    /* decoder.py:3729-3739 */
#endif                                                    /* (planted 5: messages carried over) */
#if ORC_PLANT == 6
            for (int32_t j = 0; j < V; ++j) L->post[j] = lam0[j];  /* planted: posterior reset */
#endif
            run_leg(g, p, syn, lam0, wj_buf, p->set_max_iter, L, &st);
            if (leg == 0 && p->pre_iter == 0)             /* decoder.py:3809-3810 */
                memcpy(L->first_hd, L->hard, (size_t)V);
        }
    }

    /* output assembly (decoder.py:3841-3873) */
    uint8_t *out = bt->errors + (size_t)b * V;
    const uint8_t *src = isfinite(st.best_w) ? L->best : (p->last_leg ? L->hard : L->first_hd);
    memcpy(out, src, (size_t)V);
    bt->weights[b] = ST(st.best_w);
    bt->iterations[b] = st.iters;
    bt->found_count[b] = st.found;
    if (bt->last_posterior) memcpy(bt->last_posterior + (size_t)b * V, L->post, sizeof(elem_t) * V);
}

/* Decode a batch.  Returns 0 on success, -1 on allocation failure.
 * nthreads <= 0 uses the OpenMP default. */
int orc_decode(int32_t C, int32_t V, int32_t E, const int32_t *row_ptr, const int32_t *row_var,
               const orc_params *p, const orc_batch *bt, int32_t nthreads)
{
    orc_graph g = {C, V, E, row_ptr, row_var, NULL, NULL};
    if (build_var_lists(&g)) return -1;
    int rc = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel reduction(| : rc)
    {
        orc_lane L;
        size_t e1 = (size_t)(E > 0 ? E : 1);
        L.mu = (elem_t *)malloc(sizeof(elem_t) * e1);
        L.nu = (elem_t *)malloc(sizeof(elem_t) * e1);
        L.post = (elem_t *)malloc(sizeof(elem_t) * (size_t)V);
        L.hard = (uint8_t *)malloc((size_t)V);
        L.best = (uint8_t *)malloc((size_t)V);
        L.first_hd = (uint8_t *)malloc((size_t)V);
        L.flips = (uint8_t *)malloc((size_t)V);
        L.nu2 = (elem_t *)malloc(sizeof(elem_t) * e1);
        L.count_flips = 0;
        L.leg_t = 0;
        elem_t *wj = (elem_t *)malloc(sizeof(elem_t) * (size_t)V);
        L.gamma = NULL;
        L.gamma_const = ELEM_ZERO;
        if (!L.mu || !L.nu || !L.post || !L.hard || !L.best || !L.first_hd || !L.flips || !L.nu2 || !wj) {
            rc |= 1;
        } else {
#pragma omp for schedule(dynamic, 1)
            for (int32_t b = 0; b < bt->B; ++b) decode_shot(&g, p, bt, b, &L, wj);
        }
        free(L.mu); free(L.nu); free(L.post); free(L.hard); free(L.best); free(L.first_hd); free(L.flips); free(L.nu2);
        free(wj);
    }
    free(g.var_ptr);
    free(g.var_edges);
    return rc ? -1 : 0;
}

#if ORC_ELEM == 32
/* A single check update in isolation, for closed-form tests of C2V:
 * mu[deg] in, nu[deg] out (kernels.py:1128-1215). */
void orc_check_update_row(int32_t deg, const float *mu, uint8_t syn, int32_t use_alpha,
                          float alpha, int32_t use_beta, float beta, float *nu)
{
    int32_t row_ptr[2] = {0, deg};
    int32_t *row_var = (int32_t *)malloc(sizeof(int32_t) * (size_t)(deg > 0 ? deg : 1));
    for (int32_t k = 0; k < deg; ++k) row_var[k] = k;
    orc_graph g = {1, deg, deg, row_ptr, row_var, NULL, NULL};
    orc_params p;
    memset(&p, 0, sizeof p);
    p.use_alpha = use_alpha;
    p.use_beta = use_beta;
    p.beta = beta;
    orc_lane L;
    memset(&L, 0, sizeof L);
    L.mu = (float *)mu;
    L.nu = nu;
    check_update(&g, &p, &syn, alpha, &L);
    free(row_var);
}

#endif

#if ORC_ELEM == 16
/* The half conversions, for their pins. */
void orc_to_half(const double *x, uint16_t *h, int32_t n) { for (int32_t i = 0; i < n; ++i) h[i] = to_half(x[i]); }
void orc_from_half(const uint16_t *h, double *x, int32_t n) { for (int32_t i = 0; i < n; ++i) x[i] = from_half(h[i]); }
#endif

int orc_planted(void) { return ORC_PLANT; }
int orc_elem_bits(void) { return ORC_ELEM; }
