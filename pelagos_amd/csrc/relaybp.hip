// relaybp.hip -- MI355X (gfx950) Relay-BP-S min-sum decoder.
//
// Design (one persistent launch per decode):
//   * A workgroup owns one shot at a time and runs that shot's whole Relay-BP-S
//     schedule -- pre-iter leg, then relay legs until stop_nconv solutions --
//     with no host round trip.  When it finishes it takes the next shot from a
//     device work queue, so the heavy tail of slow shots is absorbed by
//     whichever workgroups are free (the reference's per-lane freezing,
//     kernels.py:745-847, becomes "this workgroup moves on").  The queue hands
//     out shots in decreasing syndrome weight, so the slow shots start early.
//   * Per-shot state is kept in compressed min-sum form instead of per-edge
//     messages (the reference's muT/nuT [E,B], decoder.py:1527-1529):
//       variables  LS[v] = {lam_mixed_v, S_v}   (8 B)  S_v = sum of incoming nu
//       checks     CS[c][r], c = 0..3 = {P1, -P1, P2, -P2}  (16 B) outgoing magnitudes
//                                     m1, m2 with the check's sign applied, and negated;
//                                     four planes, so lanes reading consecutive checks
//                                     hit distinct LDS banks
//       edge codes SB    2 bits per edge: sign of the variable->check message and
//                         "this edge holds the check's minimum"
//     Every message is a pure function of this state, bit for bit:
//       nu_e = CS[r][code_e]  (code = sign + 2 argmin; "-" flips the sign bit)
//       mu_e = lam_mixed_v + (S_v - nu_e)              (kernels.py:1339-1350)
//       post_v = lam_mixed_v + S_v                       (kernels.py:1328)
//     With the shot's prior LLRs this is 133 KB at 936 x 8785, which fits the
//     160 KB LDS of a CU, so the iteration loop touches HBM/L2 only for the graph
//     indices (shared by every shot).
//   * One iteration = check pass (a thread per check row, sequential over the
//     row: min1/min2/argmin/sign parity, plus the parity test of the previous
//     iteration's hard decision) and variable pass (a thread per variable,
//     sequential sum over its edges in ascending edge id), separated by
//     workgroup barriers.  No cross-lane reductions are needed.
//   * Rows and variables are renumbered by decreasing degree and stored as
//     64-wide ELL slices (one slice per wave), so the lanes of a wave do equal
//     work (row slices are then reordered to balance the SIMDs, and positions
//     within a slice chosen to spread LDS gathers over the banks) and element
//     k of every row of a slice is one coalesced 256-byte read.  The index entries are the LDS byte addresses of the state they
//     select, so a gather needs no address arithmetic.  Per-variable inputs
//     and outputs (priors, gammas, error bits) stay in the caller's order and
//     are mapped through the permutation.  Slots past a row's or a variable's
//     degree point at sentinel state (a variable with LLR +inf, a check whose
//     messages are +0), which leaves every result unchanged, so no lane tests
//     its own degree.
//
// Floating point follows the reference kernels exactly (fp32, no FMA
// contraction -- the build passes -ffp-contract=off and the memory mix uses
// explicit round-to-nearest intrinsics).
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "pelagos_relaybp.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define PEL_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess)                                                              \
            return fail(PEL_EDEVICE, std::string(#call) + ": " + hipGetErrorString(_e));   \
    } while (0)

constexpr int kSlice = 64;         // rows / variables per ELL slice (one wave)
constexpr int kChunk = 8;          // check-row edges per chunk (two chunks per edge-code word)
constexpr int kVChunk = 4;         // variable edges per chunk (column weights are small)
constexpr int kEdgesPerWord = 16;  // 2-bit edge codes
constexpr float kMinStart = 1e30f; // start of a check's running minima: the reference's clip (kernels.py:1120-1126)

// Queue counters at the start of the workspace: first-leg tickets handed out,
// re-queue tickets handed out, shots re-queued, shots that left their first leg.
enum { kCtrlFirst = 0, kCtrlTaken = 1, kCtrlPushed = 2, kCtrlLeft = 3, kCtrlWords = 64 };

// State arrays.  Placement bit k puts array k in LDS.  With LS, CS and SB all in
// LDS ("fast" placements 7 and 15) the layout is the canonical one the index
// entries are baked for: CS at byte 0, SB at sbb, LS at lsb, L0 after LS.
enum { kLS = 0, kCS = 1, kSB = 2, kL0 = 3, kNumArrays = 4 };

struct GraphDev {
    int32_t C, V, E;
    uint32_t sbb, lsb;          // canonical LDS byte offsets of SB and LS
    uint32_t cs_shift;          // log2 of the byte stride between the four CS planes
    // check side, by row position (slices of rows of similar degree)
    const int32_t *row_orig;    // [C]  original check index
    const int32_t *row_deg;     // [C]
    const int4 *rslice;         // [C/64] {first ELL element, first code word, width (whole chunks), 0}
    const uint32_t *row_ent;    // ELL (k, lane) at rslice.x + 64k + lane: lsb + 8 * (variable position);
                                // padding: the sentinel variable at position V
    // variable side, by variable position (variables sorted by decreasing degree)
    const int32_t *vperm;       // [V]  position -> caller's variable index
    const int32_t *vpos;        // [V]  caller's variable index -> position
    const int32_t *var_deg;     // [V]
    const int4 *vslice;         // [V/64] {first ELL element, max degree, 0, 0}
    const uint2 *var_ent;       // ELL {4 * row position, ((sbb + 4 * word) << 5) | bit}, ascending edge id;
                                // padding: the sentinel check at position C
};

struct KArgs {
    GraphDev g;
    int32_t B;
    int32_t pre_iter, num_sets, set_max_iter, stop_nconv;
    int32_t plain, use_alpha, use_beta, last_leg, skip_relay_legs;
    float gamma0, beta;
    const float *alpha_t;
    const uint8_t *syn;
    const float *prior;
    const float *wj;
    int32_t prior_batched;
    const float *wj_pos;     // wj and the priors in variable-position order (one copy for the
    const float *prior_pos;  // batch), or null when they are per shot
    const float *gamma_legs;
    int32_t gamma_per_row;
    int32_t gamma_by_position;  // gamma_legs rows are in variable-position order
    const float *init_gamma;
    const float *init_post;
    uint8_t *errors;
    float *weights;
    int32_t *iters;
    int32_t *found;
    float *last_post;
    int32_t *ctrl;         // queue counters (kCtrl*)
    const int32_t *order;  // [B] shot taken at each first-leg queue position
    int32_t requeue_below; // re-queue shots whose first leg converged in fewer iterations (0: never)
    int32_t *rq_slot;      // [B] re-queued shots in publication order (-1: not yet published)
    int4 *rq_state;        // [B][2] of a parked shot: {best weight bits, found, iterations, any iteration},
                           // {last leg converged, strength shifts computed, 0, 0}
    uint64_t *rq_bits;     // [B][ceil(V/64)] a parked shot's output bits (leg-stepped decodes)
    int32_t leg_step;      // -1: the whole schedule; k >= 0: only leg k (0 the pre-iter leg, k relay leg
                           // k-1) for shots short of stop_nconv solutions, every shot's outputs as if the
                           // decode ended after it, its state parked for the next leg
    float *rq_post;        // [B,V] posterior of a re-queued shot, by variable position
    int64_t *shot_clock;   // [B,3] start/end wall clock and workgroup, or null
    uint8_t *spill;
    int64_t spill_stride;
    uint32_t off[kNumArrays];  // byte offsets of the state arrays (LDS or spill slot)
    uint32_t scratch_off;      // LDS byte offset of the workgroup's reduction scratch
    uint32_t rtab, vtab;       // LDS byte addresses of the row and variable slice descriptor tables
    uint32_t rowinfo;          // LDS byte address of the shot's row flags (u8 by row position: bit 0 the
                               // syndrome, bit 1 degree > 0, bit 2 degree == 1)
    uint32_t best_lds;         // LDS byte address of the shot's output bits (one per variable position,
                               // 64 positions per 8-byte word): its best solution so far, or the
                               // hard decision the unconverged correction falls back to
    // learned memory strengths (plans made with PEL_PLAN_LEARNED)
    const float *mlp;          // network parameters, or null (then features are only recorded)
    int32_t learned_legs;      // relay legs 0 .. learned_legs-1 get per-variable strengths
    float *feat_out;           // [B][learned_legs][V][8] features at each such leg's start, or null
    uint8_t *hard_out;         // [B][learned_legs][V] hard decision there, or null
    float *gamma_out;          // [B][learned_legs][V] strengths used, or null
    uint32_t flips_lds, unsat_lds, gamma_lds;  // LDS addresses: flip counts u8 [V+1], violated checks
                                               // u8 [C+1], learned shifts i8 [V]
    uint32_t gamma_slot;       // byte offset in the group's workspace slot of a shifted leg's
                               // strengths, fp32 [V] by position
    float learned_lo, learned_hi;              // strength interval of the relay legs
    float last_scale;                          // the shift's factor on the last learned leg
    int8_t *rq_gamma;          // [B][V] relay-leg-0 strength shifts of re-queued shots (by position)
    const uint32_t *mlp_frag;  // [64][kMlpStride] the network as each lane's MFMA operands
};

template <int PLACE>
struct IsFast {
    static constexpr bool value = (PLACE & 7) == 7;
};

// LDS addresses as plain integers: the kernel has no static LDS, so the
// dynamic region starts at LDS address 0 and a baked byte offset is the
// address itself (checked on the host when a fast placement is planned).
template <typename T>
__device__ __forceinline__ __attribute__((address_space(3))) T *lds_at(uint32_t addr) {
    return reinterpret_cast<__attribute__((address_space(3))) T *>(addr);
}

// Device memory through global-address-space pointers.  A load through a
// generic pointer compiles to a flat load, which counts against the LDS
// counter as well as the memory counter, so every wait for an LDS result would
// also wait for the index loads in flight; global loads let the two overlap.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gmem(const T *p) {
    return (const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem_w(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}
typedef float vec2f __attribute__((ext_vector_type(2)));
typedef float vec4f __attribute__((ext_vector_type(4)));

// One shot's working memory.  In the fast placements every access is an LDS
// access at a baked address; otherwise it goes through per-array pointers.
template <int PLACE>
struct ShotMem {
    uint32_t lb;            // LDS base of the shot's group: 0 when a workgroup runs the shot, the wave's
                            // region when each wave of a workgroup runs its own shot
    uint8_t *ls, *cs, *sb;  // array bases (LDS or workspace slot)
    vec4f *first;           // fast placements: the workgroup's copy of LDS [0, first_bytes) (CS and SB)
    int32_t first_words;    // after a shot's first check pass, in 16-byte words
    const float *l0g;   // the caller's prior array (caller's order) when kL0 is clear
    uint32_t lsb, sbb, l0b;  // canonical LDS addresses of LS, SB and L0
    uint32_t cs_shift;       // CS plane stride = 1 << cs_shift bytes

    static constexpr bool kFast = IsFast<PLACE>::value;

    // LDS at the group-relative byte address addr.
    template <typename T>
    __device__ __forceinline__ __attribute__((address_space(3))) T *lds(uint32_t addr) const {
        return lds_at<T>(addr + lb);
    }
    // An array the placement keeps in LDS (its bit set) or in the workspace slot.
    template <int BIT, typename T>
    __device__ __forceinline__ T ld(const uint8_t *p) const {
        if (PLACE & (1 << BIT)) return *(const __attribute__((address_space(3))) T *)p;
        return *gmem(reinterpret_cast<const T *>(p));
    }
    template <int BIT, typename T>
    __device__ __forceinline__ void st(uint8_t *p, T x) const {
        if (PLACE & (1 << BIT)) *(__attribute__((address_space(3))) T *)p = x;
        else *gmem_w(reinterpret_cast<T *>(p)) = x;
    }

    __device__ __forceinline__ float2 ls_ent(uint32_t ent) const {  // ent = lsb + 8 * position
        if (kFast) {
            const vec2f r = *lds<vec2f>(ent);
            return make_float2(r.x, r.y);
        }
        const vec2f r = ld<kLS, vec2f>(ls + (ent - lsb));
        return make_float2(r.x, r.y);
    }
    // Prior of the variable whose LS entry is ent (placement bit kL0, which
    // only the fast placement 15 has): LS and L0 are parallel arrays.
    __device__ __forceinline__ float l0_ent(uint32_t ent) const {
        return *lds<float>((ent >> 1) + (l0b - (lsb >> 1)));
    }
    __device__ __forceinline__ void ls_set(int v, float2 x) const {
        if (kFast) *lds<vec2f>(lsb + 8u * (uint32_t)v) = vec2f{x.x, x.y};
        else st<kLS, vec2f>(ls + 8u * (uint32_t)v, vec2f{x.x, x.y});
    }
    __device__ __forceinline__ float l0_pos(int v) const { return *lds<float>(l0b + 4u * (uint32_t)v); }
    __device__ __forceinline__ void l0_set(int v, float x) const { *lds<float>(l0b + 4u * (uint32_t)v) = x; }
    __device__ __forceinline__ float2 ls_pos(int v) const {
        if (kFast) {
            const vec2f r = *lds<vec2f>(lsb + 8u * (uint32_t)v);
            return make_float2(r.x, r.y);
        }
        const vec2f r = ld<kLS, vec2f>(ls + 8u * (uint32_t)v);
        return make_float2(r.x, r.y);
    }
    __device__ __forceinline__ void cs_set(int r, float4 x) const {
        const uint32_t o = 4u * (uint32_t)r, pl = 1u << cs_shift;
        if (kFast) {
            *lds<float>(o) = x.x;
            *lds<float>(o + pl) = x.y;
            *lds<float>(o + 2u * pl) = x.z;
            *lds<float>(o + 3u * pl) = x.w;
        } else {
            st<kCS, float>(cs + o, x.x);
            st<kCS, float>(cs + o + pl, x.y);
            st<kCS, float>(cs + o + 2u * pl, x.z);
            st<kCS, float>(cs + o + 3u * pl, x.w);
        }
    }
    __device__ __forceinline__ float cs_msg(uint32_t e) const {  // e = 4 * row position + (code << cs_shift)
        if (kFast) return *lds<float>(e);
        return ld<kCS, float>(cs + e);
    }
    __device__ __forceinline__ uint32_t sb_get(uint32_t w) const {
        if (kFast) return *lds<uint32_t>(sbb + 4u * w);
        return ld<kSB, uint32_t>(sb + 4u * w);
    }
    __device__ __forceinline__ void sb_set(uint32_t w, uint32_t x) const {
        if (kFast) *lds<uint32_t>(sbb + 4u * w) = x;
        else st<kSB, uint32_t>(sb + 4u * w, x);
    }
    __device__ __forceinline__ uint32_t sb_ent(uint32_t e) const {  // e = ((sbb + 4 word) << 5) | bit
        if (kFast) return *lds<uint32_t>(e >> 5);
        return ld<kSB, uint32_t>(sb + ((e >> 5) - sbb));
    }
};

// Memory strength of the variable at position v for one leg, as the variable
// pass reads it: a row of V values (in position order or the caller's order) or
// a constant.
struct GammaRow {
    const float *row;
    const int32_t *perm;  // non-null: row is in the caller's order
    float konst;
    __device__ __forceinline__ float operator()(int v) const {
        if (!row) return konst;
        return perm ? gmem(row)[gmem(perm)[v]] : gmem(row)[v];
    }
};

// A leg's strengths before learned shifts are applied: the row, and optionally
// per-variable shifts in LDS (a shifted leg writes clamp(u + shift, lo, hi) to
// a row of its own before its iterations).
struct GammaSrc {
    const float *row;
    const int32_t *perm;  // non-null: row is in the caller's order
    float konst;
    uint32_t lds = 0;     // non-zero: learned per-variable shifts (int8 steps of (hi - lo) / 254, by
    float lo = 0.f, hi = 0.f;  // position) at this LDS address, added to the row's strength and
                               // clamped to [lo, hi]
    float scale = 1.f;         // factor on the shifts (the last learned leg's damping)
    __device__ __forceinline__ float base(int v) const {
        if (!row) return konst;
        return perm ? gmem(row)[gmem(perm)[v]] : gmem(row)[v];
    }
    __device__ __forceinline__ float operator()(int v) const {
        const float u = base(v);
        if (!lds) return u;
        const float d = (float)*lds_at<int8_t>(lds + (uint32_t)v) * ((hi - lo) * (scale / 254.f));
        return fminf(fmaxf(u + d, lo), hi);
    }
};

struct ShotState {
    float best_w;
    int32_t found;
    int32_t iters;
    int32_t any_iter;
    int32_t last_ok;  // the last leg ended at a valid solution (its final state satisfies every check)
    int32_t shifted;  // the network gave this shot strength shifts for its learned legs
    int32_t first_ok; // the workgroup's copy of this shot's first check pass is written
};

// Values every thread of the workgroup holds identically (LDS broadcasts,
// workgroup reductions) are passed through readfirstlane so the compiler
// treats the branches they control as uniform.
__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ double uniform_f64(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(u & 0xffffffffu));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ int4 uniform_i4(int4 v) {
    return make_int4(uniform_i32(v.x), uniform_i32(v.y), uniform_i32(v.z), uniform_i32(v.w));
}

// An opaque copy of the thread index.  Branches on a particular thread (the
// workgroup leader, lane 0 of a wave) test this value so that the optimizer
// cannot carry the outcome of one such test to the next one: jump-threading
// "thread 0 stores results" into "thread 0 fetches the next shot" gives the
// persistent loop a second back edge that other lanes take without lane 0,
// which then never fetches a new shot.
__device__ __forceinline__ unsigned opaque_tid() {
    unsigned v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "v"((unsigned)threadIdx.x));
    return v;
}

// Element u of one ELL chunk: the chunk base is uniform (scalar registers),
// the lane's byte offset is a loop-invariant 32-bit register and u's offset a
// compile-time immediate, so the load needs no per-element address arithmetic.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t load_ell(const uint32_t *__restrict__ chunk, uint32_t lane_off, int u) {
    return *gmem(reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(chunk) + (size_t)lane_off +
                                                    (size_t)(u * kSlice * 4)));
}
__device__ __forceinline__ uint2 load_ell(const uint2 *__restrict__ chunk, uint32_t lane_off, int u) {
    const u32x2 x = *gmem(reinterpret_cast<const u32x2 *>(reinterpret_cast<const char *>(chunk) + (size_t)lane_off +
                                                          (size_t)(u * kSlice * 8)));
    return make_uint2(x.x, x.y);
}

// A shot runs on a group of BLOCK threads: the whole workgroup (BLOCK is the
// workgroup size) or, with BLOCK == 64, one wave of a workgroup whose waves run
// a shot each (the shot-per-wave mode of small codes).  Shot-level code takes
// its thread index, barriers and leader from these.
template <int BLOCK>
__device__ __forceinline__ int gtid() {
    return (int)(threadIdx.x & (unsigned)(BLOCK - 1));
}
// Group barrier.  A one-wave group needs no s_barrier: the workgroup-scope
// fences still order its LDS and global accesses across its lanes (the waits
// __syncthreads has), without stopping the workgroup's other waves.
template <int BLOCK>
__device__ __forceinline__ void gsync() {
    if (BLOCK == 64) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        __syncthreads();
    }
}
template <int BLOCK>
__device__ __forceinline__ bool gleader() {
    return (opaque_tid() & (unsigned)(BLOCK - 1)) == 0u;
}

// Sum over the group (double, the same butterfly order in every wave, then
// the waves in order, from +0).
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double *s_red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (BLOCK == 64) return uniform_f64(0.0 + v);
    const int wid = threadIdx.x >> 6;
    if ((opaque_tid() & 63u) == 0) s_red[wid] = v;
    __syncthreads();
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) tot += s_red[w];
    __syncthreads();
    return uniform_f64(tot);
}

// OR of a predicate over the workgroup with one barrier: a ballot per wave, one
// flag per wave in LDS.  Each wave rewrites its own flag on every call; two
// calls are always separated by another barrier, so no flag is overwritten
// while a slower wave still reads the previous round.
template <int BLOCK>
__device__ __forceinline__ int block_or(int pred, int *s_flag) {
    const bool any = __ballot(pred) != 0ull;
    if (BLOCK == 64) return any ? 1 : 0;
    if ((opaque_tid() & 63u) == 0) s_flag[threadIdx.x >> 6] = any ? 1 : 0;
    __syncthreads();
    int r = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) r |= s_flag[w];
    return uniform_i32(r);
}

// Slice descriptors, read from the group's LDS copy (one broadcast read).
template <int PLACE>
__device__ __forceinline__ int4 row_slice(const KArgs &a, const ShotMem<PLACE> &m, int pos) {
    const vec4f r = *m.template lds<vec4f>(a.rtab + 16u * (uint32_t)(pos >> 6));
    return uniform_i4(make_int4(__float_as_int(r.x), __float_as_int(r.y), __float_as_int(r.z), __float_as_int(r.w)));
}
template <int PLACE>
__device__ __forceinline__ int2 var_slice(const KArgs &a, const ShotMem<PLACE> &m, int v) {
    const vec2f r = *m.template lds<vec2f>(a.vtab + 8u * (uint32_t)(v >> 6));
    return make_int2(uniform_i32(__float_as_int(r.x)), uniform_i32(__float_as_int(r.y)));
}

// The shot's row flags, read by every check pass from LDS: written once per
// shot (syndrome bit, degree > 0, degree == 1 by row position) so that a check
// pass does not start with dependent global loads of the syndrome and the row
// degree, whose latency every wave would wait out.
template <int BLOCK, int PLACE>
__device__ __forceinline__ void row_info_init(const KArgs &a, const ShotMem<PLACE> &m, const uint8_t *__restrict__ syn) {
    for (int pos = gtid<BLOCK>(); pos < a.g.C; pos += BLOCK) {
        const int deg = a.g.row_deg[pos];
        *m.template lds<uint8_t>(a.rowinfo + (uint32_t)pos) =
            (uint8_t)((gmem(syn)[gmem(a.g.row_orig)[pos]] & 1u) | (deg > 0 ? 2u : 0u) | (deg == 1 ? 4u : 0u));
    }
}

// Edge codes: 16 slots of a row share one 32-bit word, slot k's 2-bit code at
// bits 31-2k (sign of the incoming message mu_e) and 30-2k (the edge holds the
// row's minimum), so CS plane code*stride holds, for row r:
//   0: +m1 * par   1: +m2 * par   2: -m1 * par   3: -m2 * par.
//
// One chunk of kChunk edge slots of a row: gathers each slot's variable state
// (and, after the first iteration, its previous check message, selected by the
// slot's code in wcur, whose bits 15-2u.. are slot u's), then updates the
// running minimum, second minimum, argmin and hard-decision parity.  Each slot
// shifts `bits` left by two and brings its sign in at bit 1 (v_alignbit takes
// mu's bits 31 and 30; bit 30 is cleared afterwards).  After the first
// iteration mu = lam_mixed + (S - nu) is never -0 (S starts at +0 and sums
// from +0, so S - nu and with it mu can be -0 only if lam_mixed and S were),
// so its sign bit is exactly mu < 0 for the finite states that finite priors
// give.  The first iteration's mu is the prior itself, whose sign is taken by
// comparison.
template <bool FIRST, int PLACE>
__device__ __forceinline__ uint32_t check_chunk(const KArgs &a, const ShotMem<PLACE> &m, const uint32_t (&e)[kChunk],
                                                int k0, uint32_t wcur, uint32_t qrow, uint32_t bits, float &m1,
                                                float &m2, uint32_t &arg_lo, uint32_t &arg_base, uint64_t &hpar) {
    float x[kChunk], y[kChunk], nuo[kChunk];
    bool any_lt = false;
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
        if (FIRST) {
            if (PLACE & (1 << kL0)) {
                x[u] = m.l0_ent(e[u]);
            } else {
                const int v = (int)((e[u] - m.lsb) >> 3);
                x[u] = v < a.g.V ? gmem(m.l0g)[gmem(a.g.vperm)[v]] : INFINITY;
            }
            y[u] = 0.f;
            nuo[u] = 0.f;
        } else {
            const float2 ls = m.ls_ent(e[u]);
            x[u] = ls.x;
            y[u] = ls.y;
            nuo[u] = m.cs_msg((__builtin_amdgcn_ubfe(wcur, 14u - 2u * (uint32_t)u, 2u) << m.cs_shift) | qrow);
        }
    }
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
        float mu = x[u];
        if (FIRST) {
            bits = (bits << 2) | (mu < 0.f ? 2u : 0u);
        } else {
            hpar ^= __ballot(x[u] + y[u] <= 0.f);
            mu = x[u] + (y[u] - nuo[u]);
            bits = __builtin_amdgcn_alignbit(bits, __float_as_uint(mu), 30u);
        }
        const float am = fabsf(mu);
        const bool lt = am < m1;
        m2 = __builtin_amdgcn_fmed3f(m1, am, m2);
        arg_lo = lt ? (uint32_t)u : arg_lo;  // an inline constant: no copy of k0 + u per slot
        any_lt |= lt;
        m1 = lt ? am : m1;
    }
    arg_base = any_lt ? (uint32_t)k0 : arg_base;  // the chunk of the last new minimum
    return bits;
}

// Check pass for row position pos (kernels.py:1128-1215).  FIRST: messages were
// just reset (mu_e = lam0_v, kernels.py:981-1000), so nothing is reconstructed
// and no parity is tested.  Otherwise reconstructs mu_e = lam_mixed + (S - nu_e)
// from the previous iteration's state and also returns 1 if the previous
// iteration's hard decision violates the check (kernels.py:509-569 fused into
// the check pass).
//
// The wave walks its slice in chunks of kChunk edges up to the slice's widest
// row, two chunks (one 16-slot code word) per step.  The index loads of each
// chunk are issued one chunk ahead, so their L2 latency overlaps the previous
// chunk's gathers and arithmetic.  Slots past the row's end hold the sentinel
// variable (LLR +inf), whose message is +inf: positive, never below the
// running minimum and never a hard-decision 1, so it changes nothing.
//
// The running reduction keeps the reference's tie semantics (kernels.py:
// 1160-1178): the argmin edge is the first to attain the minimum and m2 equals
// m1 when two edges share it; med3(m1, a, m2) is the new second minimum for
// every order of a against m1 <= m2.  Both start at 1e30 as the reference's
// do (kernels.py:1120-1126, padding masked to 1e30 at 1146), so an outgoing
// magnitude is min(1e30, min over the other edges): messages stay clipped
// at 1e30 when posteriors grow past it under strong memory.
template <bool FIRST, int PLACE>
__device__ __forceinline__ int check_row(const KArgs &a, const ShotMem<PLACE> &m, int pos, float alpha,
                                         const uint8_t *__restrict__ syn) {
    const int4 rs = row_slice(a, m, pos);  // {ent, word, width, 0}
    const int lane = pos & 63;
    const uint32_t info = *m.template lds<uint8_t>(a.rowinfo + (uint32_t)pos);
    const uint32_t s_i = info & 1u;
    const uint32_t wbase = (uint32_t)(rs.y + lane);
    const uint32_t lane_off = 4u * (uint32_t)lane;
    const uint32_t qrow = 4u * (uint32_t)pos;  // this row's messages CS[0..3][pos]
    const uint32_t *__restrict__ ent = a.g.row_ent + rs.x;
    float m1 = kMinStart, m2 = kMinStart;
    uint32_t arg_lo = 0, arg_base = 0;  // argmin = arg_base + arg_lo
    uint64_t hpar = 0;  // lane mask: parity of hard decisions so far
    uint32_t neg = 0;
    uint32_t ea[kChunk], eb[kChunk];
    if (rs.z > 0) {
#pragma unroll
        for (int u = 0; u < kChunk; ++u) ea[u] = load_ell(ent, lane_off, u);
    }
    for (int k0 = 0; k0 < rs.z; k0 += 2 * kChunk) {
        const uint32_t word = wbase + (uint32_t)(k0 >> 4) * kSlice;
        const uint32_t wold = FIRST ? 0u : m.sb_get(word);
        const bool two = k0 + kChunk < rs.z;
        if (two) {
            const uint32_t *__restrict__ c = ent + (int64_t)(k0 + kChunk) * kSlice;
#pragma unroll
            for (int u = 0; u < kChunk; ++u) eb[u] = load_ell(c, lane_off, u);
        }
        uint32_t bits = check_chunk<FIRST, PLACE>(a, m, ea, k0, wold >> 16, qrow, 0u, m1, m2, arg_lo, arg_base, hpar);
        if (two) {
            if (k0 + 2 * kChunk < rs.z) {
                const uint32_t *__restrict__ c = ent + (int64_t)(k0 + 2 * kChunk) * kSlice;
#pragma unroll
                for (int u = 0; u < kChunk; ++u) ea[u] = load_ell(c, lane_off, u);
            }
            bits = check_chunk<FIRST, PLACE>(a, m, eb, k0 + kChunk, wold, qrow, bits, m1, m2, arg_lo, arg_base, hpar);
        } else {
            bits <<= 16;  // a word's first chunk holds its high half
        }
        bits &= 0xAAAAAAAAu;  // the signs; the minimum flags start clear
        m.sb_set(word, bits);
        neg ^= __popc(bits);
    }
    if (info & 2u) {
        // Mark the edge holding the minimum (after the sign words above).
        const uint32_t argk = arg_base + arg_lo;
        const uint32_t aw = wbase + (argk >> 4) * kSlice;
        m.sb_set(aw, m.sb_get(aw) | (1u << (30u - 2u * (argk & 15u))));
        float o1 = m1, o2 = m2;
        if (info & 4u) o1 = o2 = 0.f;  // degree-1 checks send zero (kernels.py:1192-1193)
        if (a.use_alpha) {
            o1 = alpha * o1;
            o2 = alpha * o2;
        }
        if (a.use_beta) {
            o1 = fmaxf(o1 - a.beta, 0.f);
            o2 = fmaxf(o2 - a.beta, 0.f);
        }
        const uint32_t par = ((neg ^ s_i) & 1u) << 31;
        const uint32_t p1 = __float_as_uint(o1) ^ par, p2 = __float_as_uint(o2) ^ par;
        m.cs_set(pos, make_float4(__uint_as_float(p1), __uint_as_float(p2), __uint_as_float(p1 ^ 0x80000000u),
                                      __uint_as_float(p2 ^ 0x80000000u)));
    }
    const bool hp = (hpar >> (threadIdx.x & 63)) & 1ull;
    return FIRST ? 0 : (int)(hp ^ (s_i != 0u));
}

// Parity of the current hard decision on row position pos (after the last
// iteration of a leg); padding slots read the sentinel variable (+inf, a 0).
template <int PLACE>
__device__ __forceinline__ int parity_row(const KArgs &a, const ShotMem<PLACE> &m, int pos,
                                          const uint8_t *__restrict__ syn) {
    const int4 rs = row_slice(a, m, pos);
    bool hpar = (*m.template lds<uint8_t>(a.rowinfo + (uint32_t)pos) & 1u) != 0u;
    const uint32_t lane_off = 4u * (uint32_t)(pos & 63);
    for (int k0 = 0; k0 < rs.z; k0 += kChunk) {
        const uint32_t *__restrict__ chunk = a.g.row_ent + rs.x + (int64_t)k0 * kSlice;
        uint32_t e[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) e[u] = load_ell(chunk, lane_off, u);
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
            const float2 ls = m.ls_ent(e[u]);
            hpar ^= (ls.x + ls.y <= 0.f);
        }
    }
    return (int)hpar;
}

// W consecutive incoming messages of one variable, added to s in slot order
// (ascending edge id).  Entry {4 * row position, ((sbb + 4 word) << 5) | bit}:
// the 2-bit code is read at the bit offset held in the entry's low 5 bits
// (v_bfe takes its offset modulo 32) and selects CS[code][row].  Padding slots
// select the sentinel check's +0, which leaves s unchanged (s is never -0).
template <int W, int PLACE>
__device__ __forceinline__ void var_messages(const ShotMem<PLACE> &m, const uint2 *__restrict__ chunk,
                                             uint32_t lane_off, float &s) {
    uint2 e[W];
#pragma unroll
    for (int u = 0; u < W; ++u) e[u] = load_ell(chunk, lane_off, u);
    float nu[W];
#pragma unroll
    for (int u = 0; u < W; ++u) {
        const uint32_t w = m.sb_ent(e[u].y);
        nu[u] = m.cs_msg(e[u].x + (__builtin_amdgcn_ubfe(w, e[u].y, 2u) << m.cs_shift));
    }
#pragma unroll
    for (int u = 0; u < W; ++u) s += nu[u];
}

// Variable pass for variable position v (kernels.py:1299-1350).  The incoming
// messages are added one at a time in ascending edge id.
template <int PLACE, bool FLIPS = false>
__device__ __forceinline__ void var_update(const KArgs &a, const ShotMem<PLACE> &m, int v, const GammaRow &gs) {
    const int2 vs = var_slice(a, m, v);  // {ent, width}
    const float g = gs(v);
    const float l0 = (PLACE & (1 << kL0)) ? m.l0_pos(v) : gmem(m.l0g)[gmem(a.g.vperm)[v]];
    const float2 ls = m.ls_pos(v);
    const float post_prev = ls.x + ls.y;
    const float lm = __fadd_rn(__fmul_rn(1.0f - g, l0), __fmul_rn(g, post_prev));
    const uint32_t lane_off = 8u * (uint32_t)(v & 63);
    const uint2 *__restrict__ base = a.g.var_ent + vs.x;
    float s = 0.f;
    // The slice's width (uniform): up to 8 messages are gathered in one round
    // (index loads, then code words, then messages), wider slices in chunks.
    switch (vs.y) {
        case 1: var_messages<1, PLACE>(m, base, lane_off, s); break;
        case 2: var_messages<2, PLACE>(m, base, lane_off, s); break;
        case 3: var_messages<3, PLACE>(m, base, lane_off, s); break;
        case 4: var_messages<4, PLACE>(m, base, lane_off, s); break;
        case 5: var_messages<5, PLACE>(m, base, lane_off, s); break;
        case 6: var_messages<6, PLACE>(m, base, lane_off, s); break;
        case 7: var_messages<7, PLACE>(m, base, lane_off, s); break;
        case 8: var_messages<8, PLACE>(m, base, lane_off, s); break;
        default: {
            int q = 0;
            for (; q + kVChunk <= vs.y; q += kVChunk)
                var_messages<kVChunk, PLACE>(m, base + (int64_t)q * kSlice, lane_off, s);
            const int rest = vs.y - q;
            if (rest == 3) var_messages<3, PLACE>(m, base + (int64_t)q * kSlice, lane_off, s);
            else if (rest == 2) var_messages<2, PLACE>(m, base + (int64_t)q * kSlice, lane_off, s);
            else if (rest == 1) var_messages<1, PLACE>(m, base + (int64_t)q * kSlice, lane_off, s);
        }
    }
    m.ls_set(v, make_float2(lm, s));
    if (FLIPS && ((post_prev <= 0.f) != (lm + s <= 0.f))) {
        // hard-decision flips during the leg (a feature of the learned strengths)
        __attribute__((address_space(3))) uint8_t *c = m.template lds<uint8_t>(a.flips_lds + (uint32_t)v);
        *c = (uint8_t)min((int)*c + 1, 255);
    }
}

// ---- learned memory strengths --------------------------------------------------------------
// At the start of relay leg r < learned_legs every variable gets its own memory
// strength from the network declared in pelagos_relaybp.h, evaluated on
// features of the state the previous leg ended in.  No counterpart in the
// reference (its plug points are explicit_gammas / init_gamma,
// decoder.py:480-491, 2322).
//
// Network evaluation runs on the matrix cores (bf16 operands, fp32
// accumulation), transposed so that each layer's output is already the next
// layer's B operand (H1^T = W1 X^T, H2^T = W2 H1^T, z = w3 H2^T with w3 on every
// row of its A operand, 32 variables per product; see row32 below): no data moves
// between layers.

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));


constexpr int kMW1 = 0, kMB1 = 32 * 8, kMW2 = kMB1 + 32, kMB2 = kMW2 + 32 * 32, kMW3 = kMB2 + 32, kMB3 = kMW3 + 32;
static_assert(kMB3 + 1 == PEL_MLP_PARAMS, "parameter layout");

typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// The network on the matrix cores, 32 variables per product (v_mfma_f32_32x32x16_bf16).
// A 32x32 product leaves lane l the column l % 32 (a variable) and the rows
// row32(l, e) = 8 (e / 4) + 4 (l / 32) + e % 4 of its accumulators e = 0..15 (the
// hidden units); its A operand gives lane l the row l % 32 and K slots 8 (l / 32)
// .. +7, its B operand the same K slots of column l % 32.  So the hidden units a
// lane holds after one layer are the K slots it feeds into the next one, in any
// order the weights of that layer are laid out for: no data moves between layers.
__host__ __device__ constexpr int row32(int half, int e) { return 8 * (e / 4) + 4 * half + (e % 4); }

// This lane's slices of the network.
struct MlpRegs {
    bf16x8_t w1;     // layer 1, row h = lane % 32: K slots 0-6 the weights of features 0-6 and
                     // slot 7 the bias b1[h], as the bf16 nearest each (lanes 0-31) and the bf16
                     // nearest its remainder (lanes 32-63)
    bf16x8_t w2[2];  // layer 2, row h: the weights of hidden units row32(lane / 32, 8 i + e)
    f32x16_t b2;     // layer-2 accumulator start: b2 of rows row32(lane / 32, e)
    bf16x8_t w3[2];  // layer 3 as a 32x32 A operand whose rows all equal w3: the weights of
                     // hidden units row32(lane / 32, 8 i + e)
    float b3;
};

// The network as each lane's MFMA operands, laid out [lane][kMlpStride words]
// so that a lane loads its registers with ten 16-byte loads from one base
// address (immediate offsets): words 0-3 w1, 4-11 w2, 12-27 b2, 28-35 w3, 36 b3,
// 37-39 zero.
constexpr int kMlpWords = 37, kMlpStride = 40;

// Two fp32 values rounded to bf16 (nearest even), low first: one
// v_cvt_pk_bf16_f32 (two scalar conversions compile to one each plus a v_perm).
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(vec2f{a, b}, bf16x2_t));
}

// max(x, 0) of both bf16 halves of a word (v_pk_max_i16): the sign bit makes
// every negative bf16 (and -0) a negative 16-bit integer.
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t w) {
    typedef short s16x2_t __attribute__((ext_vector_type(2)));
    const s16x2_t v = __builtin_bit_cast(s16x2_t, w);
    const s16x2_t z = {0, 0};
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, z));
}

// The bf16 nearest x (as a float) and the bf16 nearest the remainder: together
// they carry x to about 16 significant bits.
__device__ __forceinline__ void split_bf16(float x, float &hi, float &lo) {
    hi = (float)(__bf16)x;
    lo = (float)(__bf16)(x - hi);
}

// One wave: lane l writes its operand words (weights rounded to bf16 once).
// The leg-index feature (7) is 0 wherever the network runs (the start of relay
// leg 0), so its weights never contribute; its K slot carries the bias instead,
// against a constant 1 in the features.
__global__ void __launch_bounds__(64) mlp_frag_kernel(const float *__restrict__ p, uint32_t *__restrict__ frag) {
    const int l = threadIdx.x, h = l & 31, half = l >> 5;
    auto put = [&](int w, uint32_t x) { frag[l * kMlpStride + w] = x; };
    float w1[8];
    for (int k = 0; k < 8; ++k) {
        float hi, lo;
        split_bf16(k < 7 ? p[kMW1 + h * 8 + k] : p[kMB1 + h], hi, lo);
        w1[k] = half ? lo : hi;
    }
    for (int k = 0; k < 4; ++k) put(k, pack_bf16x2(w1[2 * k], w1[2 * k + 1]));
    for (int i = 0; i < 2; ++i)
        for (int k = 0; k < 4; ++k)
            put(4 + 4 * i + k, pack_bf16x2(p[kMW2 + h * 32 + row32(half, 8 * i + 2 * k)],
                                           p[kMW2 + h * 32 + row32(half, 8 * i + 2 * k + 1)]));
    for (int e = 0; e < 16; ++e) put(12 + e, __float_as_uint(p[kMB2 + row32(half, e)]));
    for (int k = 0; k < 8; ++k)
        put(28 + k, pack_bf16x2(p[kMW3 + row32(half, 2 * k)], p[kMW3 + row32(half, 2 * k + 1)]));
    put(36, __float_as_uint(p[kMB3]));
    for (int w = kMlpWords; w < kMlpStride; ++w) put(w, 0u);
}

__device__ __forceinline__ void load_mlp(const uint32_t *__restrict__ frag, MlpRegs &r) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 *__restrict__ base = reinterpret_cast<const u32x4 *>(frag + (threadIdx.x & 63) * kMlpStride);
    // global loads (a flat load would also hold up the LDS waits in between)
    uint32_t w[kMlpStride];
#pragma unroll
    for (int i = 0; i < kMlpStride / 4; ++i) {
        const u32x4 q = *gmem(base + i);
        w[4 * i] = q.x;
        w[4 * i + 1] = q.y;
        w[4 * i + 2] = q.z;
        w[4 * i + 3] = q.w;
    }
    r.w1 = __builtin_bit_cast(bf16x8_t, u32x4{w[0], w[1], w[2], w[3]});
#pragma unroll
    for (int i = 0; i < 2; ++i)
        r.w2[i] = __builtin_bit_cast(bf16x8_t, u32x4{w[4 + 4 * i], w[5 + 4 * i], w[6 + 4 * i], w[7 + 4 * i]});
#pragma unroll
    for (int e = 0; e < 16; ++e) r.b2[e] = __uint_as_float(w[12 + e]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
        r.w3[i] = __builtin_bit_cast(bf16x8_t, u32x4{w[28 + 4 * i], w[29 + 4 * i], w[30 + 4 * i], w[31 + 4 * i]});
    r.b3 = __uint_as_float(w[36]);
}

// w3 . h2 of column l % 32 of a 32-variable group (without b3) from the layer-1
// input x (B operand): layer 1 (bias through K slot 7), bf16 ReLU, layer 2 from
// b2, bf16 ReLU, and layer 3 as two more products whose A rows all equal w3, so
// that every accumulator of lane l holds the column's sum.
__device__ __forceinline__ float mlp_group_logit(const MlpRegs &r, bf16x8_t x) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const f32x16_t zero = {};
    const f32x16_t d1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(r.w1, x, zero, 0, 0, 0);
    uint32_t h1[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) h1[k] = relu_bf16x2(pack_bf16x2(d1[2 * k], d1[2 * k + 1]));
    f32x16_t d2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
        r.w2[0], __builtin_bit_cast(bf16x8_t, u32x4{h1[0], h1[1], h1[2], h1[3]}), r.b2, 0, 0, 0);
    d2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
        r.w2[1], __builtin_bit_cast(bf16x8_t, u32x4{h1[4], h1[5], h1[6], h1[7]}), d2, 0, 0, 0);
    uint32_t h2[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) h2[k] = relu_bf16x2(pack_bf16x2(d2[2 * k], d2[2 * k + 1]));
    f32x16_t d3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
        r.w3[0], __builtin_bit_cast(bf16x8_t, u32x4{h2[0], h2[1], h2[2], h2[3]}), zero, 0, 0, 0);
    d3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
        r.w3[1], __builtin_bit_cast(bf16x8_t, u32x4{h2[4], h2[5], h2[6], h2[7]}), d3, 0, 0, 0);
    return d3[0];
}

// Logit of the variable whose features this lane packed (bf16 pairs: features
// 0-1, 2-3, 4-5, 6 and the constant 1 of the bias slot), for all 64 lanes of
// the wave at once (all must be active): columns of group A are the wave's
// variables 0-31, of group B 32-63.  Lane l of a product needs the features of
// column l % 32 (both halves carry them, against the hi and the lo weights):
// one permlane32 swap of a word with itself gives group A's word (lanes 32-63
// take lanes 0-31's) and group B's (lanes 0-31 take lanes 32-63's).  Lane l's
// own column is column l % 32 of group A (l < 32) or of group B.
__device__ __forceinline__ float mlp_logit(const MlpRegs &r, const uint32_t (&pk)[4]) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    uint32_t xa[4], xb[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const auto t = __builtin_amdgcn_permlane32_swap(pk[w], pk[w], false, false);
        xa[w] = t[0];
        xb[w] = t[1];
    }
    const float za = mlp_group_logit(r, __builtin_bit_cast(bf16x8_t, u32x4{xa[0], xa[1], xa[2], xa[3]}));
    const float zb = mlp_group_logit(r, __builtin_bit_cast(bf16x8_t, u32x4{xb[0], xb[1], xb[2], xb[3]}));
    return ((threadIdx.x & 32u) ? zb : za) + r.b3;
}

__device__ __forceinline__ float clamp4(float x) { return fminf(fmaxf(x, -4.f), 4.f); }

// 1/deg correctly rounded, for deg >= 1: the reciprocal instruction (within 1
// ulp) and one Newton step in fused multiply-adds give it for every integer in
// [1, 65535] (an exhaustive check over every float within 1 ulp of 1/deg); wider
// columns take the division.
__device__ __forceinline__ float inv_degree(int deg) {
    float d = (float)deg;
    if (deg > 65535) {
        asm volatile("" : "+v"(d));  // keeps the division in this branch
        return 1.f / d;
    }
    const float r = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(r, __builtin_fmaf(-d, r, 1.f), r);
}

// Features of every variable at the start of relay leg r (the state the
// previous leg ended in), scaled to O(1):
//   0 prior LLR / 8            1 posterior / 16 (clamped to +-4)
//   2 hard-decision flips during the previous leg / 8
//   3 mean incoming check message / 8      4 smallest incoming message / 8
//   5 column weight / 8        6 fraction of the variable's checks violated
//   7 r / 4
// Every lane computes the eight features of one variable (a wave takes a
// 64-variable slice).  Then, with parameters, the network gives every variable
// its strength shift for leg r: the slice is evaluated as two 32-variable
// matrix-core groups (mlp_logit), and lane l's own logit comes back to lane l.
// The shift is
// (hi - lo) (sigmoid(z) - 1/2), kept in LDS as the nearest multiple of
// (hi - lo) / 254 (int8), and the leg uses gamma = clamp(u + shift, lo, hi)
// with u the leg's own random draw, whose disorder lets successive legs
// explore differently; a network that outputs z = 0 reproduces the uniform
// legs exactly.  known_valid: the state satisfies every check (the previous
// leg converged), so feature 6 is 0 without a parity pass.  apply: evaluate
// the network and shift the leg (else the leg keeps its uniform row).  fresh: no
// check pass has run in this shot (pre_iter 0), so the check state holds no
// messages of it and feature 4 is 0 (every message is still 0).  The flip counts
// are cleared for the next leg.  Optionally records features and hard
// decisions in the caller's order.
// Slices per wave whose features are gathered before their network evaluation.
// More than one overlaps their loads, but spills the learned kernel's registers
// (it sits at the 128-VGPR limit of 16 waves per CU), so it stays 1.
constexpr int kFeatBatch = 1;

// What the features of variable position vc are computed from (see
// leg_features_t): its state, degree, prior and the smallest of its incoming
// messages (and the count of its violated checks).
template <int PLACE, bool KNOWN_VALID>
struct VarFeatures {
    float2 ls;
    float l0, mn;
    int deg, uns;
    int2 vs;
    // Columns k0 .. k0 + W - 1 of the slice (W <= its width; WHOLE: k0 = 0 and
    // W is the width): the W index loads are issued together (uniform base, the
    // lane's offset, immediates), then the code words, then the messages; slots
    // past this variable's degree are not counted.
    template <int W, bool WHOLE = false>
    __device__ __forceinline__ void gather(const KArgs &a, const ShotMem<PLACE> &m, int vc, int k0) {
        const uint2 *__restrict__ base = a.g.var_ent + vs.x + (int64_t)k0 * kSlice;
        const uint32_t lane_off = 8u * (uint32_t)(vc & 63);
        uint2 e[W];
#pragma unroll
        for (int u = 0; u < W; ++u) e[u] = load_ell(base, lane_off, u);
        float msg[W];
        int un[W];
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const uint32_t w = m.sb_ent(e[u].y);
            msg[u] = m.cs_msg(e[u].x + (__builtin_amdgcn_ubfe(w, e[u].y, 2u) << m.cs_shift));
            un[u] = KNOWN_VALID ? 0 : (int)*m.template lds<uint8_t>(a.unsat_lds + (e[u].x >> 2));
        }
        if (WHOLE && __ballot(deg != W) == 0) {
            // every variable of the slice has its width (variables sit by decreasing
            // degree, so this is every slice but a few): every slot counts
#pragma unroll
            for (int u = 0; u < W; ++u) {
                mn = fminf(mn, msg[u]);
                uns += un[u];
            }
        } else {
            // (the opaque copy keeps the slot tests in this branch)
            int dg = deg;
            asm volatile("" : "+v"(dg));
#pragma unroll
            for (int u = 0; u < W; ++u) {  // selects, not branches: no load is sunk into a branch
                const bool in = k0 + u < dg;
                mn = fminf(mn, in ? msg[u] : INFINITY);
                uns += in ? un[u] : 0;
            }
        }
    }
    __device__ __forceinline__ void start(const KArgs &a, const ShotMem<PLACE> &m, int vc) {
        ls = m.ls_pos(vc);
        deg = a.g.var_deg[vc];
        l0 = (PLACE & (1 << kL0)) ? m.l0_pos(vc) : gmem(m.l0g)[gmem(a.g.vperm)[vc]];
        vs = var_slice(a, m, vc);
        mn = INFINITY;
        uns = 0;
        switch (vs.y) {
            case 0: break;
            case 1: gather<1, true>(a, m, vc, 0); break;
            case 2: gather<2, true>(a, m, vc, 0); break;
            case 3: gather<3, true>(a, m, vc, 0); break;
            case 4: gather<4, true>(a, m, vc, 0); break;
            case 5: gather<5, true>(a, m, vc, 0); break;
            case 6: gather<6, true>(a, m, vc, 0); break;
            case 7: gather<7, true>(a, m, vc, 0); break;
            case 8: gather<8, true>(a, m, vc, 0); break;
            default: {
                int k0 = 0;
                for (; k0 + 4 <= vs.y; k0 += 4) gather<4>(a, m, vc, k0);
                for (; k0 < vs.y; ++k0) gather<1>(a, m, vc, k0);
            }
        }
    }
};

template <int BLOCK, int PLACE, bool KNOWN_VALID>
__device__ __forceinline__ void leg_features_t(const KArgs &a, int b, const ShotMem<PLACE> &m, int r, const uint8_t *__restrict__ syn,
                               bool apply, bool fresh) {
    constexpr bool known_valid = KNOWN_VALID;
    const int C = a.g.C, V = a.g.V, L = a.learned_legs;
    if (!known_valid) {
        for (int i = gtid<BLOCK>(); i <= C; i += BLOCK)
            *m.template lds<uint8_t>(a.unsat_lds + (uint32_t)i) = (uint8_t)(i < C ? parity_row<PLACE>(a, m, i, syn) : 0);
        gsync<BLOCK>();
    }
    const int lane = gtid<BLOCK>() & 63;
    const bool net = a.mlp && apply;
    // A wave takes kFeatBatch of its slices at a time (v0 is uniform, so every
    // lane of a wave runs the same rounds; the MFMAs need all 64 lanes), and
    // the network's operands stay in registers over the wave's slices.  Lanes
    // past V, and batch slots past the last slice, compute a copy of variable
    // V - 1 (in the last slice) and store nothing.
    MlpRegs mr;
    if (net) load_mlp(a.mlp_frag, mr);
    for (int vb = gtid<BLOCK>() & ~63; vb < V; vb += kFeatBatch * BLOCK) {
        VarFeatures<PLACE, KNOWN_VALID> vf[kFeatBatch];
#pragma unroll
        for (int j = 0; j < kFeatBatch; ++j) {
            vf[j].start(a, m, min(vb + j * BLOCK + lane, V - 1));
            if (fresh) vf[j].mn = 0.f;  // no check pass has run: every message is still 0
        }
        uint32_t pk[kFeatBatch][4];
#pragma unroll
        for (int j = 0; j < kFeatBatch; ++j) {
            const VarFeatures<PLACE, KNOWN_VALID> &x = vf[j];
            const int v = vb + j * BLOCK + lane;
            const int vc = min(v, V - 1);
            float f[8];
            const float inv_deg = x.deg > 0 ? inv_degree(x.deg) : 0.f;
            const float post = x.ls.x + x.ls.y;
            f[0] = x.l0 * 0.125f;
            f[1] = clamp4(post * 0.0625f);
            f[2] = (float)*m.template lds<uint8_t>(a.flips_lds + (uint32_t)vc) * 0.125f;
            f[3] = clamp4(x.ls.y * inv_deg * 0.125f);
            f[4] = x.deg > 0 ? clamp4(x.mn * 0.125f) : 0.f;
            f[5] = (float)x.deg * 0.125f;
            f[6] = known_valid ? 0.f : (float)x.uns * inv_deg;
            f[7] = (float)r * 0.25f;
            if (v < V && (a.feat_out || a.hard_out)) {
                const size_t row = ((size_t)b * L + r) * V + a.g.vperm[v];
                if (a.feat_out) {
                    *reinterpret_cast<float4 *>(a.feat_out + row * 8) = make_float4(f[0], f[1], f[2], f[3]);
                    *reinterpret_cast<float4 *>(a.feat_out + row * 8 + 4) = make_float4(f[4], f[5], f[6], f[7]);
                }
                if (a.hard_out) a.hard_out[row] = post <= 0.f ? 1 : 0;
            }
            pk[j][0] = pack_bf16x2(f[0], f[1]);
            pk[j][1] = pack_bf16x2(f[2], f[3]);
            pk[j][2] = pack_bf16x2(f[4], f[5]);
            pk[j][3] = pack_bf16x2(f[6], 1.f);  // feature 7 (0 at leg 0) gives its slot to the bias
        }
        if (net) {
#pragma unroll
            for (int j = 0; j < kFeatBatch; ++j) {
                const int v = vb + j * BLOCK + lane;
                if (uniform_i32(vb + j * BLOCK) >= V) break;
                const float z = mlp_logit(mr, pk[j]);
                // shift of the leg's own draw in steps of (hi - lo) / 254: 254 (sigmoid(z) - 1/2)
                // rounded to the nearest integer (ties to even), in [-127, 127]
                // (hardware exp2 and reciprocal: within a few ulp of the correctly rounded
                // sigmoid, which moves a shift by one step only at a rounding tie)
                const float sig = __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-z * 1.4426950408889634f));
                const float steps = __builtin_rintf(254.f * (sig - 0.5f));
                if (v < V) *m.template lds<int8_t>(a.gamma_lds + (uint32_t)v) = (int8_t)(int)steps;
            }
        }
#pragma unroll
        for (int j = 0; j < kFeatBatch; ++j) {
            const int v = vb + j * BLOCK + lane;
            if (v < V) *m.template lds<uint8_t>(a.flips_lds + (uint32_t)v) = 0;
        }
    }
    gsync<BLOCK>();
}

template <int BLOCK, int PLACE>
__device__ __forceinline__ void leg_features(const KArgs &a, int b, const ShotMem<PLACE> &m, int r, const uint8_t *__restrict__ syn,
                             bool known_valid, bool apply, bool fresh = false) {
    if (known_valid) leg_features_t<BLOCK, PLACE, true>(a, b, m, r, syn, apply, false);
    else leg_features_t<BLOCK, PLACE, false>(a, b, m, r, syn, apply, fresh);
}

// The strengths a learned leg uses (shifted or its uniform row), in the
// caller's order, for gamma_out.
template <int BLOCK>
__device__ __forceinline__ void record_strengths(const KArgs &a, int b, int r, const GammaSrc &gs) {
    const int V = a.g.V, L = a.learned_legs;
    for (int v = gtid<BLOCK>(); v < V; v += BLOCK) a.gamma_out[((size_t)b * L + r) * V + a.g.vperm[v]] = gs(v);
}

constexpr int kBatchU = 4;  // positions per thread handled together by the per-shot loops

// The shot's output bits from its current state: bit v = (post_v <= 0), or 0
// everywhere when no iteration has run yet (the reference's hard_dec starts at
// zero).  A wave packs the 64 positions of one output word with a ballot.
// Readers of the words are separated from this by a workgroup barrier.
template <int BLOCK, int PLACE>
__device__ __forceinline__ void best_from_state(const KArgs &a, const ShotMem<PLACE> &m, bool any_iter) {
    const int V = a.g.V;
    for (int vb = (int)(gtid<BLOCK>() & ~63u); vb < V; vb += BLOCK) {
        const float2 ls = m.ls_pos(min(vb + (int)(gtid<BLOCK>() & 63u), V));  // position V: the sentinel, +inf
        const uint64_t bits = __ballot(any_iter && ls.x + ls.y <= 0.f);
        if ((opaque_tid() & 63u) == 0) *m.template lds<uint64_t>(a.best_lds + 8u * (uint32_t)(vb >> 6)) = bits;
    }
}

// The shot's error row, in the caller's order, from its output bits: the only
// write of the row, made once by the workgroup that completes the shot.
template <int BLOCK, int PLACE>
__device__ __forceinline__ void write_errors(const KArgs &a, const ShotMem<PLACE> &m, int b) {
    const int V = a.g.V;
    uint8_t *err = a.errors + (size_t)b * V;
    gsync<BLOCK>();  // every wave's output words are in LDS
    for (int j0 = gtid<BLOCK>(); j0 < V; j0 += kBatchU * BLOCK) {
        int v[kBatchU];
        uint32_t w[kBatchU];
#pragma unroll
        for (int u = 0; u < kBatchU; ++u) v[u] = a.g.vpos[min(j0 + u * BLOCK, V - 1)];
#pragma unroll
        for (int u = 0; u < kBatchU; ++u) w[u] = *m.template lds<uint32_t>(a.best_lds + 4u * (uint32_t)(v[u] >> 5));
#pragma unroll
        for (int u = 0; u < kBatchU; ++u)
            if (j0 + u * BLOCK < V) err[j0 + u * BLOCK] = (uint8_t)((w[u] >> (v[u] & 31)) & 1u);
    }
}

// f(j, state) for every caller-order index j this thread owns, with the state
// of its position; kBatchU at a time, so that their index loads and state
// reads are in flight together.
template <int BLOCK, int PLACE, typename F>
__device__ __forceinline__ void for_each_caller_index(const KArgs &a, const ShotMem<PLACE> &m, F f) {
    const int V = a.g.V;
    for (int j0 = gtid<BLOCK>(); j0 < V; j0 += kBatchU * BLOCK) {
        int v[kBatchU];
        float2 ls[kBatchU];
#pragma unroll
        for (int u = 0; u < kBatchU; ++u) v[u] = a.g.vpos[min(j0 + u * BLOCK, V - 1)];
#pragma unroll
        for (int u = 0; u < kBatchU; ++u) ls[u] = m.ls_pos(v[u]);
#pragma unroll
        for (int u = 0; u < kBatchU; ++u)
            if (j0 + u * BLOCK < V) f(j0 + u * BLOCK, ls[u]);
    }
}

// The shot's first check pass back from the workgroup's copy into LDS [0,
// first_words * 16) (fast placements), with loads that write LDS directly
// (global_load_lds_dwordx4: no registers, the data lands in LDS when the load
// completes).  Thread t reloads words t, t + BLOCK, ..., the ones it saved; a
// wave's 64 words are consecutive, as the instruction writes them (lane l at
// the wave's base + 16 l).  The barrier after which any wave reads them waits
// for them (a workgroup barrier waits for the wave's loads to LDS).
template <int BLOCK, int PLACE>
__device__ __forceinline__ void restore_first(const ShotMem<PLACE> &m) {
    const int lane = gtid<BLOCK>() & 63;
    for (int i = gtid<BLOCK>(); i - lane < m.first_words; i += BLOCK) {
        if (i < m.first_words)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(m.first + i),
                                             (__attribute__((address_space(3))) void *)lds_at<uint8_t>(
                                                 m.lb + 16u * (uint32_t)(i - lane)),
                                             16, 0, 0);
    }
}

// check_and_select_kernel (kernels.py:657-741) for a shot whose hard decision
// satisfies every check: weight = sum of wj over bits with post <= 0 (double
// accumulation, rounded once), strict improvement or first solution wins.  The
// winner goes to the shot's output bits (best_errors, kernels.py:722-734).
template <int BLOCK, int PLACE>
__device__ __forceinline__ void select_solution(const KArgs &a, const ShotMem<PLACE> &m, const float *__restrict__ wj,
                                ShotState &st, double *s_red) {
    const int V = a.g.V;
    if (st.best_w < 1e20f) {
        // A best solution exists (its bits are the output bits).  A leg that found
        // the same one again (the common case) cannot replace it: equal weight is
        // no strict improvement.  Comparing the bits (a ballot per 64 positions) is
        // cheaper than the weight sum, which needs the weights from L2 and a double
        // reduction.  The reduction's scratch holds the comparison's flags here; a
        // barrier separates them from the weight sum that may follow.
        int diff = 0;
        for (int vb = (int)(gtid<BLOCK>() & ~63u); vb < V; vb += BLOCK) {
            const float2 ls = m.ls_pos(min(vb + (int)(gtid<BLOCK>() & 63u), V));  // position V: +inf, a 0
            const uint64_t bits = __ballot(ls.x + ls.y <= 0.f);
            diff |= bits != *m.template lds<uint64_t>(a.best_lds + 8u * (uint32_t)(vb >> 6)) ? 1 : 0;
        }
        if (!block_or<BLOCK>(diff, reinterpret_cast<int *>(s_red))) {
            st.found += 1;
            return;
        }
        gsync<BLOCK>();
    }
    double acc = 0.0;
    // Position order: contiguous state reads.  The double sum of n fp32
    // weights is exact, whatever the order, while the set weights span less
    // than 2^(29 - log2 n) in magnitude (2^15 at n = V = 8785).
    if (a.wj_pos) {
        // weights in position order too: kBatchU contiguous loads per thread in
        // flight together, and a select rather than a branch
        for (int v0 = gtid<BLOCK>(); v0 < V; v0 += kBatchU * BLOCK) {
            float2 ls[kBatchU];
            float w[kBatchU];
#pragma unroll
            for (int u = 0; u < kBatchU; ++u) {
                ls[u] = m.ls_pos(min(v0 + u * BLOCK, V));  // position V: the sentinel (+inf, never set)
                w[u] = gmem(a.wj_pos)[min(v0 + u * BLOCK, V - 1)];
            }
#pragma unroll
            for (int u = 0; u < kBatchU; ++u) acc += ls[u].x + ls[u].y <= 0.f ? (double)w[u] : 0.0;
        }
    } else {
        // per-shot weights in the caller's order: gathered only for the few set bits
        for (int v = gtid<BLOCK>(); v < V; v += BLOCK) {
            const float2 ls = m.ls_pos(v);
            if (ls.x + ls.y <= 0.f) acc += (double)gmem(wj)[gmem(a.g.vperm)[v]];
        }
    }
    const float w = (float)block_sum<BLOCK>(acc, s_red);
    if (st.best_w >= 1e20f || w < st.best_w) {
        st.best_w = w;
        best_from_state<BLOCK, PLACE>(a, m, true);
    }
    st.found += 1;
}

// One leg of up to T iterations.  The shot leaves the leg at its first valid
// solution (freeze_finished_lanes_kernel with sought = found + 1).  FLIPS:
// count each variable's hard-decision flips (a feature of the learned
// strengths; only legs whose end state feeds the network count them).
template <int BLOCK, int PLACE, bool FLIPS = false>
__device__ __forceinline__ bool run_leg(const KArgs &a, const ShotMem<PLACE> &m, int T, const GammaRow &gs,
                        const float *__restrict__ wj, const uint8_t *__restrict__ syn,
                        ShotState &st, double *s_red, int *s_flag) {
    const int C = a.g.C, V = a.g.V;
    for (int t = 0; t < T; ++t) {
        const float alpha = a.use_alpha ? gmem(a.alpha_t)[t] : 1.0f;
        int bad = 0;
        // A leg's first check pass starts from messages reset to the priors
        // (kernels.py:981-1000), so its output (CS and SB) depends only on the
        // shot's priors, syndrome and alpha_t[0]: every leg of a shot computes
        // the same words.  In the fast placements (CS and SB at LDS [0, lsb))
        // the first one is copied to the workgroup's slot and later legs copy it
        // back.  Each thread restores exactly the words it saved.
        // A one-wave shot (a small code) recomputes the pass: a few rows per lane cost
        // less than the copy's round trip through L2.
        constexpr bool kCopyFirst = ShotMem<PLACE>::kFast && BLOCK > 64;
        const bool restore = t == 0 && kCopyFirst && st.first_ok;
        if (restore) {
            restore_first<BLOCK, PLACE>(m);
            // the reloaded words are in LDS before this wave goes on (a one-wave
            // group's barrier does not wait for loads; a workgroup barrier would)
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        } else if (t == 0) {
            for (int i = gtid<BLOCK>(); i < C; i += BLOCK) check_row<true, PLACE>(a, m, i, alpha, syn);
        } else {
            for (int i = gtid<BLOCK>(); i < C; i += BLOCK) bad |= check_row<false, PLACE>(a, m, i, alpha, syn);
        }
        bad = block_or<BLOCK>(bad, s_flag);
        if (t == 0 && kCopyFirst && !st.first_ok) {
            // read before the barrier that ends this iteration; the next check
            // pass rewrites CS and SB only after it
            for (int i = gtid<BLOCK>(); i < m.first_words; i += BLOCK) gmem_w(m.first)[i] = *m.template lds<vec4f>(16u * (uint32_t)i);
            st.first_ok = 1;
        }
        if (t > 0 && !bad) {  // iteration t-1 converged
            select_solution<BLOCK, PLACE>(a, m, wj, st, s_red);
            return true;
        }
        for (int v = gtid<BLOCK>(); v < V; v += BLOCK) var_update<PLACE, FLIPS>(a, m, v, gs);
        gsync<BLOCK>();
        st.iters += 1;
        st.any_iter = 1;
    }
    if (T > 0) {
        int bad = 0;
        for (int i = gtid<BLOCK>(); i < C; i += BLOCK) bad |= parity_row<PLACE>(a, m, i, syn);
        bad = block_or<BLOCK>(bad, s_flag);
        if (!bad) {
            select_solution<BLOCK, PLACE>(a, m, wj, st, s_red);
            return true;
        }
    }
    return false;
}

// Per-shot slices of the batch inputs.
struct ShotIO {
    const uint8_t *syn;
    const float *l0;
    const float *wj;
};

__device__ __forceinline__ ShotIO shot_io(const KArgs &a, int b) {
    const size_t pofs = a.prior_batched ? (size_t)b * a.g.V : 0;
    return ShotIO{a.syn + (size_t)b * a.g.C, a.prior + pofs, a.wj + pofs};
}

// Stage 1 of a shot: state reset (decoder.py:3430-3464) and the pre-iter leg
// (decoder.py:3470-3598).  The variable state starts as lam_mixed = post
// (prior or init_posterior), S = 0, so that lam_mixed + S == post exactly.
// Returns true when relay legs remain to run.
template <int BLOCK, int PLACE, bool LEARN>
__device__ __forceinline__ bool first_leg(const KArgs &a, int b, ShotMem<PLACE> &m, ShotState &st, double *s_red,
                          int *s_flag, bool l0_ready = false) {
    const int V = a.g.V;
    const ShotIO io = shot_io(a, b);
    if ((PLACE & (1 << kL0)) && l0_ready) {
        // The priors in LDS are this shot's already (shared by the batch, and the
        // group's previous shot had the same graph and priors): the state starts
        // from them by position, with no global loads.
        for (int v = gtid<BLOCK>(); v < V; v += BLOCK) {
            m.ls_set(v, make_float2(m.l0_pos(v), 0.f));
            if (LEARN) *m.template lds<uint8_t>(a.flips_lds + (uint32_t)v) = 0;
        }
    } else
    for (int j0 = gtid<BLOCK>(); j0 < V; j0 += kBatchU * BLOCK) {  // kBatchU per thread at a time
        int v[kBatchU];
        float lj[kBatchU], p0[kBatchU];
#pragma unroll
        for (int u = 0; u < kBatchU; ++u) {
            const int j = min(j0 + u * BLOCK, V - 1);
            v[u] = a.g.vpos[j];
            lj[u] = io.l0[j];
            p0[u] = a.init_post ? a.init_post[(size_t)b * V + j] : lj[u];
        }
#pragma unroll
        for (int u = 0; u < kBatchU; ++u) {
            if (j0 + u * BLOCK >= V) break;
            m.ls_set(v[u], make_float2(p0[u], 0.f));
            if (PLACE & (1 << kL0)) m.l0_set(v[u], lj[u]);
            if (LEARN) *m.template lds<uint8_t>(a.flips_lds + (uint32_t)v[u]) = 0;
        }
    }
    row_info_init<BLOCK, PLACE>(a, m, io.syn);
    m.l0g = io.l0;
    gsync<BLOCK>();

    st = ShotState{INFINITY, 0, 0, 0, 0, 0, 0};
    GammaRow g0{nullptr, nullptr, a.plain ? 0.f : a.gamma0};
    if (!a.plain && a.init_gamma) {
        g0.row = a.init_gamma + (size_t)b * V;
        g0.perm = a.g.vperm;
    }
    if (st.found < a.stop_nconv) {
        // the network reads the flips of this leg
        if (LEARN && a.learned_legs > 0)
            st.last_ok = run_leg<BLOCK, PLACE, true>(a, m, a.pre_iter, g0, io.wj, io.syn, st, s_red, s_flag);
        else
            st.last_ok = run_leg<BLOCK, PLACE, false>(a, m, a.pre_iter, g0, io.wj, io.syn, st, s_red, s_flag);
    }
    // first_leg_hard_dec (decoder.py:3597-3598): the output bits while no solution exists.
    if (!a.last_leg && st.found == 0) best_from_state<BLOCK, PLACE>(a, m, st.any_iter != 0);
    return !a.skip_relay_legs && a.num_sets > 0 && st.found < a.stop_nconv;
}

// The memory strengths of relay leg `leg` for shot b: row leg of the table.
__device__ __forceinline__ GammaSrc leg_gamma(const KArgs &a, int b, int leg) {
    GammaSrc gl{nullptr, nullptr, 0.f};
    if (a.gamma_per_row) {
        gl.row = a.gamma_legs + ((size_t)leg * a.B + b) * a.g.V;
        gl.perm = a.g.vperm;
    } else {
        gl.row = a.gamma_legs + (size_t)leg * a.g.V;
        gl.perm = a.gamma_by_position ? nullptr : a.g.vperm;
    }
    return gl;
}

// A shifted leg's strengths clamp(u + shift, lo, hi), by position, written once
// to the group's workspace slot: the leg's iterations then read one float per
// variable, as a uniform leg reads its row, instead of the row, the shift and
// the clamp in every variable pass.
template <int BLOCK, int PLACE>
__device__ __forceinline__ GammaRow leg_strengths(const KArgs &a, const ShotMem<PLACE> &m, const GammaSrc &gl) {
    if (!gl.lds) return GammaRow{gl.row, gl.perm, gl.konst};
    float *g = reinterpret_cast<float *>(reinterpret_cast<uint8_t *>(m.first) + a.gamma_slot);
    for (int v = gtid<BLOCK>(); v < a.g.V; v += BLOCK) *gmem_w(g + v) = gl(v);
    gsync<BLOCK>();
    return GammaRow{g, nullptr, 0.f};
}

// Stage 2: relay legs until stop_nconv solutions (decoder.py:3655-3835), then
// the unconverged correction and the optional last posterior.
template <int BLOCK, int PLACE, bool LEARN>
__device__ __forceinline__ void relay_legs(const KArgs &a, int b, ShotMem<PLACE> &m, ShotState &st, bool run, double *s_red,
                           int *s_flag, bool leg0_ready = false, int leg_begin = 0, int leg_end = 1 << 30) {
    const int V = a.g.V;
    const ShotIO io = shot_io(a, b);
    if (run) {
        for (int leg = leg_begin; leg < min(a.num_sets, leg_end) && st.found < a.stop_nconv; ++leg) {
            GammaSrc gl = leg_gamma(a, b, leg);
            bool count = false;
            if (LEARN && leg < a.learned_legs) {
                // The network runs once per shot, on the state after the pre-iter leg
                // (the start of relay leg 0), and only when that state is a valid
                // solution; its per-variable shifts then apply to every learned leg
                // that starts from a valid solution (the previous leg converged).  A
                // leg that starts from a state violating some check runs on its
                // uniform row, the exploration such a state needs.  Traces record the
                // features at the start of every learned leg.  A re-queued shot (its
                // first leg converged) had the network run before it was parked.
                const bool trace = a.feat_out || a.hard_out;
                if (leg == 0) {
                    st.shifted = leg0_ready ? st.shifted : (a.mlp && st.last_ok);
                    if (!leg0_ready && (st.shifted || trace))
                        leg_features<BLOCK, PLACE>(a, b, m, 0, io.syn, st.last_ok != 0, st.shifted != 0, st.any_iter == 0);
                } else if (trace) {
                    leg_features<BLOCK, PLACE>(a, b, m, leg, io.syn, st.last_ok != 0, false, st.any_iter == 0);
                }
                if (st.shifted && st.last_ok) {
                    gl.lds = a.gamma_lds + m.lb;
                    gl.lo = a.learned_lo;
                    gl.hi = a.learned_hi;
                    gl.scale = leg == a.learned_legs - 1 ? a.last_scale : 1.f;
                }
                if (a.gamma_out) record_strengths<BLOCK>(a, b, leg, gl);
                count = trace && leg + 1 < a.learned_legs;
            }
            const GammaRow gr = leg_strengths<BLOCK, PLACE>(a, m, gl);
            if (count)
                st.last_ok = run_leg<BLOCK, PLACE, true>(a, m, a.set_max_iter, gr, io.wj, io.syn, st, s_red, s_flag);
            else
                st.last_ok = run_leg<BLOCK, PLACE, false>(a, m, a.set_max_iter, gr, io.wj, io.syn, st, s_red, s_flag);
            if (leg == 0 && a.pre_iter == 0 && !a.last_leg && st.found == 0)
                best_from_state<BLOCK, PLACE>(a, m, st.any_iter != 0);  // decoder.py:3809-3810
        }
    }
    // last_leg correction: the live hard decision for shots with no solution.
    if (a.last_leg && st.found == 0) best_from_state<BLOCK, PLACE>(a, m, st.any_iter != 0);
    write_errors<BLOCK, PLACE>(a, m, b);
    if (a.last_post) {
        float *lp = a.last_post + (size_t)b * V;
        for_each_caller_index<BLOCK, PLACE>(a, m, [&](int j, float2 ls) { lp[j] = ls.x + ls.y; });
    }
}

// Re-queue a shot between its first leg and its relay legs.  At a leg boundary
// the messages are about to be reset (kernels.py:935-975), so the posterior and
// the leg bookkeeping are the whole state: the posterior goes to rq_post, the
// scalars to rq_state, and the shot is published in the next rq_slot.  The
// parking workgroup writes nothing else of the shot (its error row and
// counters are written only by the workgroup that completes it; the first
// solution is the sign of the parked posterior).  The resuming workgroup may
// run on another XCD, so the stores must be visible device-wide before the
// leader publishes: each wave waits until its own stores have reached the L2,
// the barrier orders that before the leader's agent-scope release fence, and
// that one fence (one L2 write-back) then covers the whole workgroup.  A fence
// in every thread would write back and invalidate the XCD's L2 once per wave.
template <int BLOCK, int PLACE, bool LEARN>
__device__ __forceinline__ void requeue_shot(const KArgs &a, int b, const ShotMem<PLACE> &m, const ShotState &st) {
    const int V = a.g.V;
    float *post = a.rq_post + (size_t)b * V;
    for (int v = gtid<BLOCK>(); v < V; v += BLOCK) {
        const float2 ls = m.ls_pos(v);
        post[v] = ls.x + ls.y;
        if (LEARN && a.mlp) a.rq_gamma[(size_t)b * V + v] = *m.template lds<int8_t>(a.gamma_lds + (uint32_t)v);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    gsync<BLOCK>();
    if (gleader<BLOCK>()) {
        __threadfence();
        a.rq_state[2 * b] = make_int4(__float_as_int(st.best_w), st.found, st.iters, st.any_iter);
        a.rq_state[2 * b + 1] = make_int4(st.last_ok, st.shifted, 0, 0);
        const int slot = atomicAdd(&a.ctrl[kCtrlPushed], 1);
        __hip_atomic_store(&a.rq_slot[slot], b, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&a.ctrl[kCtrlLeft], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Resume a parked shot: posterior back as lam_mixed = post, S = 0, and its
// output bits: from the posterior's signs for a re-queued shot (the first
// leg's solution, its only one so far), else the parked bits.  A re-queued
// shot's parked state was acquired by the leader (wait_requeued) before the
// workgroup barrier that precedes every read of it here; a leg-stepped
// decode's was written by an earlier launch.
template <int BLOCK, int PLACE, bool LEARN>
__device__ __forceinline__ void resume_shot(const KArgs &a, int b, ShotMem<PLACE> &m, bool bits_from_post = true) {
    const int V = a.g.V;
    const ShotIO io = shot_io(a, b);
    const float *post = a.rq_post + (size_t)b * V;
    for (int v0 = gtid<BLOCK>(); v0 < V; v0 += kBatchU * BLOCK) {  // kBatchU per thread at a time
        float p[kBatchU], l[kBatchU];
        int8_t gq[kBatchU];
#pragma unroll
        for (int u = 0; u < kBatchU; ++u) {
            const int v = min(v0 + u * BLOCK, V - 1);
            p[u] = post[v];
            if (PLACE & (1 << kL0)) l[u] = a.prior_pos ? a.prior_pos[v] : io.l0[a.g.vperm[v]];
            if (LEARN && a.mlp) gq[u] = a.rq_gamma[(size_t)b * V + v];
        }
#pragma unroll
        for (int u = 0; u < kBatchU; ++u) {
            const int v = v0 + u * BLOCK;
            if (v >= V) break;
            m.ls_set(v, make_float2(p[u], 0.f));
            if (PLACE & (1 << kL0)) m.l0_set(v, l[u]);
            if (LEARN) {
                *m.template lds<uint8_t>(a.flips_lds + (uint32_t)v) = 0;  // leg 0's flips start from zero
                if (a.mlp) *m.template lds<int8_t>(a.gamma_lds + (uint32_t)v) = gq[u];
            }
        }
    }
    row_info_init<BLOCK, PLACE>(a, m, io.syn);
    m.l0g = io.l0;
    gsync<BLOCK>();
    if (bits_from_post) {
        best_from_state<BLOCK, PLACE>(a, m, true);
    } else {
        const int nw = (V + 63) >> 6;
        for (int w = gtid<BLOCK>(); w < nw; w += BLOCK)
            *m.template lds<uint64_t>(a.best_lds + 8u * (uint32_t)w) = a.rq_bits[(size_t)b * nw + w];
    }
}

// Leg-stepped decodes: the whole state of a shot after the leg a launch ran
// (posterior, strength shifts, output bits, scalars), for the next launch.
template <int BLOCK, int PLACE, bool LEARN>
__device__ __forceinline__ void park_shot(const KArgs &a, int b, const ShotMem<PLACE> &m, const ShotState &st) {
    const int V = a.g.V;
    float *post = a.rq_post + (size_t)b * V;
    gsync<BLOCK>();  // every wave's output words are in LDS
    for (int v = gtid<BLOCK>(); v < V; v += BLOCK) {
        const float2 ls = m.ls_pos(v);
        post[v] = ls.x + ls.y;
        if (LEARN && a.mlp) a.rq_gamma[(size_t)b * V + v] = *m.template lds<int8_t>(a.gamma_lds + (uint32_t)v);
    }
    const int nw = (V + 63) >> 6;
    for (int w = gtid<BLOCK>(); w < nw; w += BLOCK)
        a.rq_bits[(size_t)b * nw + w] = *m.template lds<uint64_t>(a.best_lds + 8u * (uint32_t)w);
    if (gleader<BLOCK>()) {
        a.rq_state[2 * b] = make_int4(__float_as_int(st.best_w), st.found, st.iters, st.any_iter);
        a.rq_state[2 * b + 1] = make_int4(st.last_ok, st.shifted, 0, 0);
    }
}

// One launch of a leg-stepped decode for shot b: leg a.leg_step only, for a
// shot still short of stop_nconv solutions (the outputs of the others are
// final), then the outputs as if the decode ended here and the parked state.
// The learned network runs, as in a whole decode, on the state after the
// pre-iter leg; its shifts are parked with the rest.
template <int BLOCK, int PLACE, bool LEARN>
__device__ __forceinline__ void legwise_shot(const KArgs &a, int b, ShotMem<PLACE> &m, ShotState &st, double *s_red,
                                             int *s_flag) {
    if (a.leg_step == 0) {
        const bool more = first_leg<BLOCK, PLACE, LEARN>(a, b, m, st, s_red, s_flag);
        if (LEARN && more && a.learned_legs > 0) {
            st.shifted = a.mlp != nullptr && st.last_ok;
            if (st.shifted) leg_features<BLOCK, PLACE>(a, b, m, 0, a.syn + (size_t)b * a.g.C, true, true);
        }
        relay_legs<BLOCK, PLACE, LEARN>(a, b, m, st, false, s_red, s_flag);
        park_shot<BLOCK, PLACE, LEARN>(a, b, m, st);
        return;
    }
    const int4 s0 = uniform_i4(a.rq_state[2 * b]);
    const int4 s1 = uniform_i4(a.rq_state[2 * b + 1]);
    st = ShotState{__int_as_float(s0.x), s0.y, s0.z, s0.w, s1.x, s1.y, 0};
    if (st.found >= a.stop_nconv) return;  // done in an earlier launch: its outputs are written
    resume_shot<BLOCK, PLACE, LEARN>(a, b, m, false);
    relay_legs<BLOCK, PLACE, LEARN>(a, b, m, st, true, s_red, s_flag, true, a.leg_step - 1, a.leg_step);
    park_shot<BLOCK, PLACE, LEARN>(a, b, m, st);
}

// Leader only: the shot behind re-queue ticket t2, or B when no shot will ever
// be published there.  Every shot leaves stage 1 (re-queued or not) within one
// leg of bounded length, and each workgroup holding one is running, so once
// all B have left (kCtrlLeft == B) the number published is final and the wait
// ends.
__device__ __forceinline__ int wait_requeued(const KArgs &a, int t2) {
    if (t2 >= a.B) return a.B;
    for (;;) {
        const int s = __hip_atomic_load(&a.rq_slot[t2], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (s >= 0) return s;
        const int left = __hip_atomic_load(&a.ctrl[kCtrlLeft], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (left >= a.B) {
            const int pushed = __hip_atomic_load(&a.ctrl[kCtrlPushed], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            if (t2 >= pushed) return a.B;
        }
        __builtin_amdgcn_s_sleep(4);
    }
}

// A shot group's working memory: LDS at the group's base lb, the rest in the
// group's workspace slot.
template <int PLACE>
__device__ __forceinline__ ShotMem<PLACE> shot_mem(const KArgs &a, uint8_t *smem, uint32_t lb, uint8_t *slot) {
    ShotMem<PLACE> m;
    m.lb = lb;
    m.ls = ((PLACE & (1 << kLS)) ? smem + lb : slot) + a.off[kLS];
    m.cs = ((PLACE & (1 << kCS)) ? smem + lb : slot) + a.off[kCS];
    m.sb = ((PLACE & (1 << kSB)) ? smem + lb : slot) + a.off[kSB];
    m.l0g = nullptr;
    m.lsb = a.g.lsb;
    m.sbb = a.g.sbb;
    m.l0b = a.off[kL0];
    m.cs_shift = a.g.cs_shift;
    m.first = reinterpret_cast<vec4f *>(slot);
    m.first_words = (int32_t)(a.g.lsb / 16u);  // CS at 0, SB at sbb, LS at lsb
    return m;
}

// The group's copy of the slice descriptor tables, and the sentinels that
// padding slots point at (no pass ever writes them).  Every later read is
// separated from these writes by the group barrier at the end of a shot's
// state reset.
template <int BLOCK, int PLACE>
__device__ __forceinline__ void group_setup(const KArgs &a, const ShotMem<PLACE> &m) {
    const int nrs = (a.g.C + kSlice - 1) / kSlice, nvs = (a.g.V + kSlice - 1) / kSlice;
    for (int i = gtid<BLOCK>(); i < nrs; i += BLOCK) {
        const int4 r = a.g.rslice[i];
        *m.template lds<vec4f>(a.rtab + 16u * (uint32_t)i) =
            vec4f{__int_as_float(r.x), __int_as_float(r.y), __int_as_float(r.z), __int_as_float(r.w)};
    }
    for (int i = gtid<BLOCK>(); i < nvs; i += BLOCK) {
        const int4 r = a.g.vslice[i];
        *m.template lds<vec2f>(a.vtab + 8u * (uint32_t)i) = vec2f{__int_as_float(r.x), __int_as_float(r.y)};
    }
    if (gleader<BLOCK>()) {
        m.ls_set(a.g.V, make_float2(INFINITY, 0.f));
        if (PLACE & (1 << kL0)) m.l0_set(a.g.V, INFINITY);
        m.cs_set(a.g.C, make_float4(0.f, 0.f, 0.f, 0.f));
    }
}

template <int BLOCK, int PLACE, bool LEARN>
__global__ void __launch_bounds__(BLOCK) relay_decode_kernel(KArgs a) {
    // All LDS is dynamic (state arrays, then the reduction scratch), so the
    // baked state addresses are absolute LDS addresses.
    extern __shared__ __align__(16) uint8_t smem[];
    double *s_red = reinterpret_cast<double *>(smem + a.scratch_off);
    int *s_flag = reinterpret_cast<int *>(smem + a.scratch_off + 8u * (BLOCK / 64));
    int *s_next = s_flag + BLOCK / 64;
    uint8_t *slot = a.spill + (int64_t)blockIdx.x * a.spill_stride;
    ShotMem<PLACE> m = shot_mem<PLACE>(a, smem, 0, slot);  // the workgroup runs one shot at a time
    group_setup<BLOCK, PLACE>(a, m);
    // The leader writes the finished shot's scalars and takes the next ticket in
    // one branch: first-leg tickets in queue order, then re-queued shots.
    int prev = -1;
    ShotState st{0.f, 0, 0, 0, 0, 0, 0};
    bool l0_loaded = false;  // the priors in LDS are this decode's (a first leg or a resume wrote them)
    for (;;) {
        if (opaque_tid() == 0) {
            const int64_t now = a.shot_clock ? (int64_t)wall_clock64() : 0;
            if (prev >= 0) {
                a.weights[prev] = st.best_w;
                a.iters[prev] = st.iters;
                a.found[prev] = st.found;
                if (a.shot_clock) {
                    a.shot_clock[3 * prev + 1] = now;
                    a.shot_clock[3 * prev + 2] = blockIdx.x;
                }
            }
            int nb = a.B, stage = 1;
            const int t = atomicAdd(&a.ctrl[kCtrlFirst], 1);
            if (t < a.B) {
                nb = a.order[t];
                if (a.shot_clock) a.shot_clock[3 * nb] = now;
            } else if (a.requeue_below > 0) {
                stage = 2;
                nb = wait_requeued(a, atomicAdd(&a.ctrl[kCtrlTaken], 1));
            }
            s_next[0] = nb;
            s_next[1] = stage;
        }
        __syncthreads();
        const int b = uniform_i32(s_next[0]);
        const int stage = uniform_i32(s_next[1]);
        __syncthreads();
        if (b >= a.B) break;
        if (a.leg_step >= 0) {
            legwise_shot<BLOCK, PLACE, LEARN>(a, b, m, st, s_red, s_flag);
            prev = b;
            continue;
        }
        bool more = true;
        if (stage == 1) {
            more = first_leg<BLOCK, PLACE, LEARN>(a, b, m, st, s_red, s_flag,
                                                  l0_loaded && !a.prior_batched && !a.init_post);
            l0_loaded = true;
            if (a.requeue_below > 0) {
                if (more && st.found > 0 && st.iters < a.requeue_below) {
                    // the first leg converged: relay leg 0 starts from a valid solution
                    if (LEARN && a.learned_legs > 0 && (a.mlp || a.feat_out || a.hard_out))
                        leg_features<BLOCK, PLACE>(a, b, m, 0, a.syn + (size_t)b * a.g.C, true, a.mlp != nullptr);
                    requeue_shot<BLOCK, PLACE, LEARN>(a, b, m, st);
                    prev = -1;
                    continue;
                }
                if (opaque_tid() == 0)
                    __hip_atomic_fetch_add(&a.ctrl[kCtrlLeft], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else {
            resume_shot<BLOCK, PLACE, LEARN>(a, b, m);
            l0_loaded = true;
            const int4 rs = uniform_i4(a.rq_state[2 * b]);
            // parked shots are those whose first leg converged; the network ran before
            // they were parked
            st = ShotState{__int_as_float(rs.x), rs.y, rs.z, rs.w, 1, a.mlp != nullptr, 0};
        }
        relay_legs<BLOCK, PLACE, LEARN>(a, b, m, st, more, s_red, s_flag, stage == 2);
        prev = b;
    }
}

// ---- one persistent launch over several jobs ----------------------------------------------------
// A job is one batch of one code (its own graph, parameters, buffers and
// workspace).  The launch's workgroups (kEngineBlock threads, one per CU)
// take the jobs in order.  A large code's shots run on the whole workgroup, one
// first-leg ticket at a time, in the job's syndrome-weight order (as in
// relay_decode_kernel without re-queue); a small code's shots run one per wave,
// each wave taking the job's next shot until none is left.  A workgroup moves
// on when the current job has no shot left, so one job's last, slow shots
// overlap the next jobs' work instead of leaving CUs idle.
constexpr int kEngineBlock = 1024;
constexpr int kEngineWaves = kEngineBlock / 64;
constexpr int kMaxJobs = PEL_MAX_JOBS;

// A job's shot routine: its placement, plus these flags.
enum { kKindWave = 16, kKindLearn = 32 };

// One job's descriptor in the launch's job table (device memory, written by
// jobs_setup_kernel before the engine runs and only read by it).
struct JobDesc {
    KArgs a;
    int32_t kind;     // placement | kKindWave (a shot per wave) | kKindLearn
    uint32_t stride;  // LDS bytes per wave of a shot-per-wave job
};

struct EngineArgs {
    int32_t n_jobs;
    uint32_t ctl_lds;  // LDS byte address of the workgroup's two control words
};

// Descriptors for the job table travel as kernel arguments, kSetupJobs per
// setup launch.
constexpr int kSetupJobs = 7;
struct SetupArgs {
    JobDesc *dst;
    int32_t count;
    int32_t pad;
    JobDesc src[kSetupJobs];
};
static_assert(sizeof(SetupArgs) <= 4096, "kernel arguments are at most 4 KiB");
static_assert(sizeof(JobDesc) % 4 == 0, "descriptors are copied as words");

// Copies its descriptors into the job table word by word, reading them in
// place from the kernel-argument segment.
__global__ void __launch_bounds__(256) jobs_setup_kernel(SetupArgs) {
    const SetupArgs &sa = *(const SetupArgs *)__builtin_amdgcn_kernarg_segment_ptr();
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&sa.src[0]);
    uint32_t *dst = reinterpret_cast<uint32_t *>(sa.dst);
    const int words = sa.count * (int)(sizeof(JobDesc) / 4);
    for (int i = threadIdx.x; i < words; i += 256) dst[i] = src[i];
}

// One shot of a job on a group (the workgroup, or BLOCK = 64: one wave at LDS
// base lb): its whole schedule, then the leader writes its scalars.
template <int BLOCK, int PLACE, bool LEARN>
__device__ __forceinline__ void engine_shot(const KArgs &a, int b, uint8_t *smem, uint32_t lb, uint8_t *slot,
                                            bool reuse = false) {
    ShotMem<PLACE> m = shot_mem<PLACE>(a, smem, lb, slot);
    double *s_red = reinterpret_cast<double *>(smem + lb + a.scratch_off);
    int *s_flag = reinterpret_cast<int *>(smem + lb + a.scratch_off + 8u * (BLOCK / 64));
    // reuse: the group's previous shot had this graph, layout and prior array, so
    // the slice tables, the sentinels and (when the batch shares its priors) the
    // priors in LDS are this shot's already
    if (!reuse) group_setup<BLOCK, PLACE>(a, m);
    ShotState st{0.f, 0, 0, 0, 0, 0, 0};
    const bool more = first_leg<BLOCK, PLACE, LEARN>(a, b, m, st, s_red, s_flag,
                                                     reuse && !a.prior_batched && !a.init_post);
    relay_legs<BLOCK, PLACE, LEARN>(a, b, m, st, more, s_red, s_flag);
    if (gleader<BLOCK>()) {
        a.weights[b] = st.best_w;
        a.iters[b] = st.iters;
        a.found[b] = st.found;
    }
}

// Shot-per-wave jobs from job j on: every wave of the workgroup takes the job's
// next shot (lane 0 takes the ticket) until the job has none left, then goes
// on by itself to the next job while the jobs keep job j's kind and per-wave
// LDS stride (its region of LDS is laid out the same), skipping empty ones.  A
// wave thus never waits for the workgroup's slower waves between such jobs
// (a launch of many batches of a small code); the workgroup meets again only
// at a job of another kind, or at the end.
template <int PLACE>
__device__ __forceinline__ void engine_wave_jobs(const JobDesc *__restrict__ jobs, int j, int n_jobs, uint8_t *smem) {
    const uint32_t wave = threadIdx.x >> 6;
    const int kind = uniform_i32(jobs[j].kind);
    const uint32_t stride = (uint32_t)uniform_i32((int)jobs[j].stride);
    // what the wave's region of LDS holds from its last shot (as in engine_kernel)
    const void *last_graph = nullptr, *last_prior = nullptr;
    for (int jj = j; jj < n_jobs; ++jj) {
        if (uniform_i32(jobs[jj].kind) != kind || (uint32_t)uniform_i32((int)jobs[jj].stride) != stride) break;
        const KArgs &a = jobs[jj].a;
        if (a.B <= 0) continue;  // an empty job has no workspace
        uint8_t *slot = a.spill + ((int64_t)blockIdx.x * kEngineWaves + wave) * a.spill_stride;
        for (;;) {
            int t = 0;
            if ((opaque_tid() & 63u) == 0) t = atomicAdd(&a.ctrl[kCtrlFirst], 1);
            t = uniform_i32(t);  // lane 0's ticket (the first active lane)
            if (t >= a.B) break;
            const bool reuse = a.g.rslice == last_graph && a.prior == last_prior;
            last_graph = a.g.rslice;
            last_prior = a.prior;
            engine_shot<64, PLACE, false>(a, a.order[t], smem, wave * stride, slot, reuse);
        }
    }
}

// Job kinds as bits of an engine instantiation's KINDS mask: the launch uses
// the smallest instantiation whose routines cover its jobs, so a launch of
// plain jobs does not carry the learned routines' register pressure.
constexpr int kind_bit(int kind) {
    return kind == 15 ? 1 : kind == 7 ? 2 : kind == 3 ? 4 : kind == 1 ? 8 : kind == 0 ? 16
         : kind == (15 | kKindLearn) ? 32 : kind == (7 | kKindLearn) ? 64 : kind == (15 | kKindWave) ? 128
         : kind == (7 | kKindWave) ? 256 : 0;
}
constexpr int kKindsPlainWg = 1 | 2 | 4 | 8 | 16;
constexpr int kKindsWave = 128 | 256;
constexpr int kKindsLearn = 32 | 64;
constexpr int kKindsAll = kKindsPlainWg | kKindsWave | kKindsLearn;
// the instantiations, smallest first

template <int KINDS>
__global__ void __launch_bounds__(kEngineBlock) engine_kernel(EngineArgs e, const JobDesc *__restrict__ jobs) {
    extern __shared__ __align__(16) uint8_t smem[];
    int *ctl = reinterpret_cast<int *>(smem + e.ctl_lds);
    int j = 0;
    // what the workgroup's LDS holds from its last whole-workgroup shot
    const void *last_graph = nullptr, *last_prior = nullptr;
    int last_kind = -1;
    for (;;) {
        // The leader finds the first job from j on with a shot left, and takes a
        // whole-workgroup job's ticket (a shot-per-wave job's waves take theirs).
        if (opaque_tid() == 0) {
            int jj = j, t = 0;
            for (; jj < e.n_jobs; ++jj) {
                const KArgs &a = jobs[jj].a;
                if (a.B <= 0) continue;  // an empty job has no workspace
                if (jobs[jj].kind & kKindWave) {
                    if (__hip_atomic_load(&a.ctrl[kCtrlFirst], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < a.B) break;
                } else {
                    t = atomicAdd(&a.ctrl[kCtrlFirst], 1);
                    if (t < a.B) break;
                }
            }
            ctl[0] = jj;
            ctl[1] = t;
        }
        __syncthreads();
        const int jj = uniform_i32(ctl[0]);
        const int t = uniform_i32(ctl[1]);
        __syncthreads();
        if (jj >= e.n_jobs) break;
        j = jj;
        const KArgs &a = jobs[j].a;
        uint8_t *slot = a.spill + (int64_t)blockIdx.x * a.spill_stride;
        const int kind = uniform_i32(jobs[j].kind);
        const int b = kind & kKindWave ? 0 : a.order[t];
        const bool reuse = kind == last_kind && a.g.rslice == last_graph && a.prior == last_prior;
        last_kind = kind;  // a shot-per-wave job lays out LDS per wave: never reused
        last_graph = kind & kKindWave ? nullptr : a.g.rslice;
        last_prior = a.prior;
#define PEL_ENGINE_CASE(K, CALL) \
    else if ((KINDS & kind_bit(K)) && kind == (K)) CALL;
        if (false) {
        }
        PEL_ENGINE_CASE(15, (engine_shot<kEngineBlock, 15, false>(a, b, smem, 0, slot, reuse)))
        PEL_ENGINE_CASE(7, (engine_shot<kEngineBlock, 7, false>(a, b, smem, 0, slot, reuse)))
        PEL_ENGINE_CASE(3, (engine_shot<kEngineBlock, 3, false>(a, b, smem, 0, slot, reuse)))
        PEL_ENGINE_CASE(1, (engine_shot<kEngineBlock, 1, false>(a, b, smem, 0, slot, reuse)))
        PEL_ENGINE_CASE(0, (engine_shot<kEngineBlock, 0, false>(a, b, smem, 0, slot, reuse)))
        PEL_ENGINE_CASE(15 | kKindLearn, (engine_shot<kEngineBlock, 15, true>(a, b, smem, 0, slot, reuse)))
        PEL_ENGINE_CASE(7 | kKindLearn, (engine_shot<kEngineBlock, 7, true>(a, b, smem, 0, slot, reuse)))
        PEL_ENGINE_CASE(15 | kKindWave, (engine_wave_jobs<15>(jobs, j, e.n_jobs, smem)))
        PEL_ENGINE_CASE(7 | kKindWave, (engine_wave_jobs<7>(jobs, j, e.n_jobs, smem)))
#undef PEL_ENGINE_CASE
    }
}

typedef void (*EngineFn)(EngineArgs, const JobDesc *);

// The smallest engine instantiation whose routines cover the job kinds in mask.
EngineFn pick_engine(int mask) {
    if ((mask & ~1) == 0) return engine_kernel<1>;
    if ((mask & ~32) == 0) return engine_kernel<32>;
    if ((mask & ~(1 | 128)) == 0) return engine_kernel<1 | 128>;
    if ((mask & ~(1 | 8 | 32 | 128)) == 0) return engine_kernel<1 | 8 | 32 | 128>;  // the three-code mix
    if ((mask & ~(kKindsPlainWg | kKindsWave)) == 0) return engine_kernel<kKindsPlainWg | kKindsWave>;
    return engine_kernel<kKindsAll>;
}

// ---- float16 and float64 decodes: the per-edge formulation ------------------------------------
// data_type "float16" / "float64" (decoder.py:402-423, the kernels' dtype_enum 0 / 2).  The fast
// kernels above keep a compressed fp32 state whose LDS addresses are baked into the graph layout;
// these other element types run the reference's own formulation instead: every edge's mu and nu
// in the workgroup's workspace slot, a check pass over rows and a variable pass over variables,
// one workgroup per shot taken from a queue, the schedule of first_leg / relay_legs.  Element
// type S is what the reference stores (mu, nu, posterior, priors, strengths, weights); the
// arithmetic follows its kernels for that dtype:
//   float64  everything in double (kernels.py:1117-1126 min1/min2, 1313-1325 sums);
//   float16  check minima and variable sums in fp32 (kernels.py:1118-1126, 1315-1325), the memory
//            mix (1 - g) lam0 + g post in half arithmetic (fp16 operands, kernels.py:1309), every
//            stored message and posterior rounded to half; the hard decision from the fp32
//            posterior, the solution bits from the stored one (1328-1337 vs 690-692), and
//            60000 as "no solution yet" (713-714).
// float32 through this path gives the fast kernels' results bit for bit (a test checks it).

constexpr int kTypedBlock = 256;

struct TypedArgs {
    int32_t C, V, E, B;
    const int32_t *row_ptr, *row_var, *var_ptr, *var_edges;  // CSR (edge id = position); per
                                                              // variable, edge ids ascending
    int32_t pre_iter, num_sets, set_max_iter, stop_nconv;
    int32_t plain, use_alpha, use_beta, last_leg, skip_relay_legs;
    float gamma0, beta;
    const float *alpha_t;
    const uint8_t *syn;
    const void *prior, *wj;  // S [V] or [B,V]
    int32_t prior_batched;
    const void *gamma_legs;  // S [R,V] or [R,B,V] (caller's variable order)
    int32_t gamma_per_row;
    const void *init_gamma, *init_post;  // S [B,V] or null
    uint8_t *errors;
    void *weights;                       // S [B]
    int32_t *iters, *found;
    void *last_post;                     // S [B,V] or null
    int32_t *ctrl;                       // [0] queue ticket counter (cleared by typed_clear_kernel)
    uint8_t *slots;
    int64_t slot_stride;
    int32_t leg_step;                    // 0: the whole schedule; k >= 1: only leg k - 1 (0 the pre-iter
                                         // leg), for shots short of stop_nconv solutions (as pel_params)
    uint8_t *park;                       // leg-stepped decodes: [B] shot states between launches
    int64_t park_stride;
};

// A parked shot (leg-stepped decodes): best weight, found and iteration counts, then its
// posterior (V elements) and its hard, best and first-leg bits (V bytes each).
constexpr int kTypedParkHead = 16;

template <typename S>
struct TypedMath {                 // float64: double arithmetic throughout
    typedef double C;
    static constexpr double kNoSolution = 1e20;
    // the double sum of the weights as the element type: float32 rounds it once (as the fast
    // kernels do), float16 rounds that float once more, float64 keeps it
    __device__ static double weight(double acc) { return acc; }
    __device__ static S mix(S g, S l0, S post) { return __dadd_rn(__dmul_rn(1.0 - g, l0), __dmul_rn(g, post)); }
};
template <>
struct TypedMath<float> {
    typedef float C;
    static constexpr double kNoSolution = 1e20;
    __device__ static double weight(double acc) { return (double)(float)acc; }
    __device__ static float mix(float g, float l0, float post) {
        return __fadd_rn(__fmul_rn(1.0f - g, l0), __fmul_rn(g, post));
    }
};
template <>
struct TypedMath<_Float16> {
    typedef float C;
    static constexpr double kNoSolution = 60000.0;
    __device__ static double weight(double acc) { return (double)(_Float16)(float)acc; }
    // fp16 operands give fp16 operations, each rounded (the file is built without contraction)
    __device__ static _Float16 mix(_Float16 g, _Float16 l0, _Float16 post) {
        const _Float16 one = (_Float16)1.0f;
        return (_Float16)((one - g) * l0) + (_Float16)(g * post);
    }
};

// Check pass (c2v_min_sum_btile_kernel, kernels.py:1128-1215) over the group's rows.
template <typename S>
__device__ __forceinline__ void typed_check_pass(const TypedArgs &a, const uint8_t *syn, float alpha, const S *mu,
                                                 S *nu) {
    typedef typename TypedMath<S>::C C;
    for (int i = threadIdx.x; i < a.C; i += kTypedBlock) {
        const int r0 = a.row_ptr[i], r1 = a.row_ptr[i + 1];
        if (r1 <= r0) continue;
        int neg = 0, cnt1 = 0;
        C min1 = (C)1e30, min2 = (C)1e30;  // kernels.py:1120-1126: magnitudes clipped at 1e30
        for (int e = r0; e < r1; ++e) {
            const C m = (C)mu[e];
            const C am = m < (C)0 ? -m : m;
            neg ^= (m < (C)0);
            if (am < min1) {
                min2 = min1;
                min1 = am;
                cnt1 = 1;
            } else if (am == min1) {
                ++cnt1;
                min2 = am;
            } else if (am < min2) {
                min2 = am;
            }
        }
        const C sign_prod = ((neg & 1) ^ (syn[i] & 1)) ? (C)-1 : (C)1;
        const bool deg_one = r1 - r0 == 1;
        for (int e = r0; e < r1; ++e) {
            const C m = (C)mu[e];
            const C am = m < (C)0 ? -m : m;
            C out = (am == min1 && cnt1 == 1) ? min2 : min1;
            if (deg_one) out = (C)0;
            if (a.use_alpha) out = (C)alpha * out;
            if (a.use_beta) out = fmax(out - (C)a.beta, (C)0);
            const C sgn = m >= (C)0 ? (C)1 : (C)-1;
            nu[e] = (S)((sign_prod * sgn) * out);
        }
    }
}

// Variable pass (v2c_and_gamma_btile_kernel, kernels.py:1299-1350): memory mix, the incoming
// messages summed in ascending edge id from 0, posterior, hard decision, outgoing messages.
template <typename S>
__device__ __forceinline__ void typed_var_pass(const TypedArgs &a, const S *l0, const S *grow, S gconst, S *post,
                                               uint8_t *hard, S *mu, const S *nu) {
    typedef typename TypedMath<S>::C C;
    for (int j = threadIdx.x; j < a.V; j += kTypedBlock) {
        const S g = grow ? grow[j] : gconst;
        const S lm = TypedMath<S>::mix(g, l0[j], post[j]);
        C s = (C)0;
        for (int q = a.var_ptr[j]; q < a.var_ptr[j + 1]; ++q) s += (C)nu[a.var_edges[q]];
        const C pc = (C)lm + s;
        post[j] = (S)pc;
        hard[j] = pc <= (C)0;
        for (int q = a.var_ptr[j]; q < a.var_ptr[j + 1]; ++q) {
            const int e = a.var_edges[q];
            mu[e] = (S)((C)lm + (s - (C)nu[e]));
        }
    }
}

struct TypedShot {
    double best_w;
    int found, iters;
};

// One leg of up to T iterations; the shot leaves it at its first valid solution
// (check_and_select_kernel, freeze_finished_lanes_kernel).  Returns true if it converged.
template <typename S>
__device__ bool typed_leg(const TypedArgs &a, const uint8_t *syn, const S *l0, const S *wj, const S *grow, S gconst,
                          int T, S *mu, S *nu, S *post, uint8_t *hard, uint8_t *best, TypedShot &st, int *s_flag,
                          double *s_red) {
    typedef typename TypedMath<S>::C C;
    for (int t = 0; t < T; ++t) {
        typed_check_pass<S>(a, syn, a.use_alpha ? a.alpha_t[t] : 1.0f, mu, nu);
        __syncthreads();
        typed_var_pass<S>(a, l0, grow, gconst, post, hard, mu, nu);
        if (threadIdx.x == 0) *s_flag = 0;
        __syncthreads();
        st.iters += 1;
        int bad = 0;
        for (int i = threadIdx.x; i < a.C; i += kTypedBlock) {
            int par = 0;
            for (int e = a.row_ptr[i]; e < a.row_ptr[i + 1]; ++e) par ^= hard[a.row_var[e]];
            bad |= par != (syn[i] & 1);
        }
        if (bad) atomicOr(s_flag, 1);
        __syncthreads();
        const int any_bad = *s_flag;
        __syncthreads();
        if (any_bad) continue;
        // the weight of the stored posterior's bits, accumulated in double (exact for these inputs)
        double acc = 0.0;
        for (int j = threadIdx.x; j < a.V; j += kTypedBlock)
            if ((C)post[j] <= (C)0) acc += (double)wj[j];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
        __syncthreads();
        double tot = 0.0;
        for (int w = 0; w < kTypedBlock / 64; ++w) tot += s_red[w];
        const double w = TypedMath<S>::weight(tot);
        const bool better = st.best_w >= TypedMath<S>::kNoSolution || w < st.best_w;
        if (better) st.best_w = w;
        for (int j = threadIdx.x; j < a.V; j += kTypedBlock) {
            if (better) best[j] = (C)post[j] <= (C)0;
            hard[j] = better ? (uint8_t)((C)post[j] <= (C)0) : best[j];  // freeze: hard <- best_errors
        }
        st.found += 1;
        __syncthreads();
        return true;
    }
    return false;
}

template <typename S>
__global__ void __launch_bounds__(kTypedBlock) typed_decode_kernel(TypedArgs a) {
    __shared__ int s_shot, s_flag;
    __shared__ double s_red[kTypedBlock / 64];
    const int V = a.V, E = a.E;
    uint8_t *slot = a.slots + (int64_t)blockIdx.x * a.slot_stride;
    S *mu = reinterpret_cast<S *>(slot);
    S *nu = mu + E;
    S *post = nu + E;
    uint8_t *hard = reinterpret_cast<uint8_t *>(post + V);
    uint8_t *best = hard + V;
    uint8_t *fhd = best + V;
    for (;;) {
        if (threadIdx.x == 0) s_shot = atomicAdd(&a.ctrl[0], 1);
        __syncthreads();
        const int b = s_shot;
        __syncthreads();
        if (b >= a.B) break;
        const uint8_t *syn = a.syn + (size_t)b * a.C;
        const size_t pofs = a.prior_batched ? (size_t)b * V : 0;
        const S *l0 = static_cast<const S *>(a.prior) + pofs;
        const S *wj = static_cast<const S *>(a.wj) + pofs;
        // state reset (decoder.py:3430-3464)
        for (int j = threadIdx.x; j < V; j += kTypedBlock) {
            post[j] = a.init_post ? static_cast<const S *>(a.init_post)[(size_t)b * V + j] : l0[j];
            hard[j] = best[j] = fhd[j] = 0;
        }
        TypedShot st{(double)INFINITY, 0, 0};
        uint8_t *pk = a.leg_step > 0 ? a.park + (int64_t)b * a.park_stride : nullptr;
        S *ppost = pk ? reinterpret_cast<S *>(pk + kTypedParkHead) : nullptr;
        uint8_t *pbits = pk ? reinterpret_cast<uint8_t *>(ppost + V) : nullptr;
        int leg_begin = 0, leg_end = a.num_sets;
        if (a.leg_step <= 1) {
            // the pre-iter leg (decoder.py:3470-3598): gamma 0 (plain), init_gamma, or gamma0
            const S *grow = (!a.plain && a.init_gamma) ? static_cast<const S *>(a.init_gamma) + (size_t)b * V : nullptr;
            const S gconst = a.plain ? (S)0.0f : (S)a.gamma0;
            for (int e = threadIdx.x; e < E; e += kTypedBlock) {  // init_messages_with_priors_EB_kernel
                mu[e] = l0[a.row_var[e]];
                nu[e] = (S)0.0f;
            }
            __syncthreads();
            if (st.found < a.stop_nconv)
                typed_leg<S>(a, syn, l0, wj, grow, gconst, a.pre_iter, mu, nu, post, hard, best, st, &s_flag, s_red);
            for (int j = threadIdx.x; j < V; j += kTypedBlock) fhd[j] = hard[j];  // decoder.py:3597-3598
            __syncthreads();
            if (a.leg_step == 1) leg_end = 0;
        } else {
            // relay leg leg_step - 2 of a leg-stepped decode, from the state the last launch parked
            st.best_w = *reinterpret_cast<const double *>(pk);
            st.found = *reinterpret_cast<const int32_t *>(pk + 8);
            st.iters = *reinterpret_cast<const int32_t *>(pk + 12);
            __syncthreads();  // every thread has read the scalars before any could change them
            if (st.found >= a.stop_nconv) continue;  // finished in an earlier launch: outputs final
            for (int j = threadIdx.x; j < V; j += kTypedBlock) {
                post[j] = ppost[j];
                hard[j] = pbits[j];
                best[j] = pbits[V + j];
                fhd[j] = pbits[2 * V + j];
            }
            __syncthreads();
            leg_begin = a.leg_step - 2;
            leg_end = leg_begin + 1;
        }
        if (!a.skip_relay_legs) {
            for (int leg = leg_begin; leg < min(a.num_sets, leg_end) && st.found < a.stop_nconv; ++leg) {
                const S *row = static_cast<const S *>(a.gamma_legs) +
                               (a.gamma_per_row ? ((size_t)leg * a.B + b) * V : (size_t)leg * V);
                for (int e = threadIdx.x; e < E; e += kTypedBlock) {  // message reset (decoder.py:3729-3739)
                    mu[e] = l0[a.row_var[e]];
                    nu[e] = (S)0.0f;
                }
                __syncthreads();
                typed_leg<S>(a, syn, l0, wj, row, (S)0.0f, a.set_max_iter, mu, nu, post, hard, best, st, &s_flag,
                             s_red);
                if (leg == 0 && a.pre_iter == 0) {  // decoder.py:3809-3810
                    for (int j = threadIdx.x; j < V; j += kTypedBlock) fhd[j] = hard[j];
                    __syncthreads();
                }
            }
        }
        // output assembly (decoder.py:3841-3873); a leg-stepped launch leaves them as if the decode
        // ended after its leg, and parks the state for the next one
        const uint8_t *src = __builtin_isfinite(st.best_w) ? best : (a.last_leg ? hard : fhd);
        for (int j = threadIdx.x; j < V; j += kTypedBlock) {
            a.errors[(size_t)b * V + j] = src[j];
            if (a.last_post) static_cast<S *>(a.last_post)[(size_t)b * V + j] = post[j];
            if (pk) {
                ppost[j] = post[j];
                pbits[j] = hard[j];
                pbits[V + j] = best[j];
                pbits[2 * V + j] = fhd[j];
            }
        }
        if (threadIdx.x == 0) {
            static_cast<S *>(a.weights)[b] = (S)st.best_w;
            a.iters[b] = st.iters;
            a.found[b] = st.found;
            if (pk) {
                *reinterpret_cast<double *>(pk) = st.best_w;
                *reinterpret_cast<int32_t *>(pk + 8) = st.found;
                *reinterpret_cast<int32_t *>(pk + 12) = st.iters;
            }
        }
        __syncthreads();
    }
}

__global__ void typed_clear_kernel(int32_t *ctrl) {
    if (threadIdx.x == 0) ctrl[0] = 0;
}

// ---- serving: a persistent launch that takes units as they are published -----------------------
// The host writes units into a ring in host memory, then raises the published
// count (and fences, so the stores leave the CPU).  A workgroup's leader claims
// the next unit with a compare-and-swap on the claim word (device memory),
// below a published count it read from memory, and copies the unit into LDS.
// The unit's shots are staged through the job's rows of this workgroup (a
// shot-per-workgroup job) or of each wave (a shot-per-wave job, whose waves
// take the unit's shots one at a time), decoded by the job's routine, and
// their results copied to the unit's buffers.  Then every wave waits for its
// stores to complete, the workgroup synchronises, and the leader sets the
// unit's done flag.
//
// The claim word (device memory) holds the units claimed so far, the number of
// workgroups holding a unit, and a closed bit.  A waiting leader backs off
// (bounded sleeps); once it has waited idle_ticks, no workgroup holds a unit
// and no unit is left, it sets the closed bit in the same compare-and-swap
// that claims compete on, and every workgroup leaves at its next look.  So a
// launch either has all its workgroups taking units or is ending, never a few
// stragglers holding slow shots while new units wait for them: a launch stays
// open while any shot runs.  A leader that sees the stop flag with no unit
// left leaves at once.  A unit in hand is always finished.
struct ServeArgs {
    const int64_t *published;     // host memory: units published so far
    const int32_t *stop;          // host memory: non-zero makes waiting workgroups leave
    const pel_serve_unit *units;  // host memory: the ring, capacity entries
    int64_t *done;                // host memory: done[k % capacity] = k + 1 once unit k is complete
    uint64_t *claim;              // device memory: the claim word
    int64_t idle_ticks;           // wall-clock ticks (100 MHz) a workgroup waits before it may close
    int32_t capacity;
    int32_t n_jobs;
    uint32_t ctl_lds;             // LDS byte address of the workgroup's control block
};
constexpr int kServeBusyShift = 40;
constexpr uint64_t kServeCount = (1ull << kServeBusyShift) - 1;  // units claimed
constexpr uint64_t kServeBusy = 1ull << kServeBusyShift;         // one workgroup holding a unit
constexpr uint64_t kServeClosed = 1ull << 63;

// Reopens the claim word for a new launch (the last one ended closed, with no
// unit held).
__global__ void __launch_bounds__(64) serve_open_kernel(uint64_t *claim) {
    if (threadIdx.x == 0) __hip_atomic_fetch_and(claim, ~kServeClosed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The workgroup's control block in LDS, 8-byte words: the unit number (-1:
// leave), the unit (eight words), the leader's last read of the published
// count and the time it last finished a unit, and the waves' shot counter.
// Everything a workgroup needs after a shot is read back from here, so no
// register holds it across the shot.
constexpr int kServeCtlBytes = 128;
enum { kCtlUnit = 0, kCtlWords = 1, kCtlSeen = 9, kCtlIdle = 10, kCtlNext = 11 };
static_assert(sizeof(pel_serve_unit) == 64, "a unit is eight 8-byte words");
static_assert(8 * (kCtlNext + 1) <= kServeCtlBytes, "control block");

template <typename T>
__device__ __forceinline__ T *uniform_ptr(int64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return reinterpret_cast<T *>((uint64_t)lo | ((uint64_t)hi << 32));
}

__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
    return ((int64_t)uniform_i32((int)(v >> 32)) << 32) | (uint32_t)uniform_i32((int)v);
}

// The unit in the control block, as uniform values.
__device__ __forceinline__ pel_serve_unit ctl_unit(const int64_t *ctl) {
    pel_serve_unit u;
    u.job = uniform_i32((int)ctl[kCtlWords]);
    u.count = uniform_i32((int)(ctl[kCtlWords] >> 32));
    u.syn_row_stride = uniform_i64(ctl[kCtlWords + 1]);
    u.syndromes = uniform_ptr<const uint8_t>(ctl[kCtlWords + 2]);
    u.errors = uniform_ptr<uint8_t>(ctl[kCtlWords + 3]);
    u.weights = uniform_ptr<float>(ctl[kCtlWords + 4]);
    u.iterations = uniform_ptr<int32_t>(ctl[kCtlWords + 5]);
    u.found_count = uniform_ptr<int32_t>(ctl[kCtlWords + 6]);
    u.syn_col_stride = uniform_i64(ctl[kCtlWords + 7]);
    return u;
}

// n bytes from src to dst by the threads of a BLOCK-thread group.
template <int BLOCK>
__device__ __forceinline__ void copy_bytes(uint8_t *__restrict__ dst, const uint8_t *__restrict__ src, int n) {
    for (int i = gtid<BLOCK>(); i < n; i += BLOCK) dst[i] = src[i];
}

// Memory the engine shares with the host and with other kernels (the units'
// syndromes and results) is read and written with system-scope accesses, which
// go to memory instead of the caches: a unit's syndromes may have been written
// after this launch read the same addresses, and its results must be in memory
// before its done flag.  No acquire fence (an L2 invalidate of the whole XCD,
// which would also evict the graph indices every other shot on the XCD reads
// each iteration) is needed for either.
template <typename T>
__device__ __forceinline__ T sys_load(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <typename T>
__device__ __forceinline__ void sys_store(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Shot i's syndrome row of the unit (strided as the unit says) into the
// staging row dst of C bytes.
template <int BLOCK>
__device__ __forceinline__ void serve_stage(uint8_t *__restrict__ dst, const pel_serve_unit &u, int i, int C) {
    const uint8_t *src = u.syndromes + (int64_t)i * u.syn_row_stride;
    for (int c = gtid<BLOCK>(); c < C; c += BLOCK) dst[c] = sys_load(src + (int64_t)c * u.syn_col_stride);
}

// Shot i of the unit from staging row b: its error row and scalars.
template <int BLOCK>
__device__ __forceinline__ void serve_results(const KArgs &a, int b, const pel_serve_unit &u, int i) {
    const int V = a.g.V;
    const uint8_t *src = a.errors + (size_t)b * V;
    uint8_t *dst = u.errors + (size_t)i * V;
    for (int j = gtid<BLOCK>(); j < V; j += BLOCK) sys_store(dst + j, src[j]);
    if (gleader<BLOCK>()) {
        sys_store(reinterpret_cast<uint32_t *>(u.weights) + i, __float_as_uint(a.weights[b]));
        sys_store(u.iterations + i, a.iters[b]);
        sys_store(u.found_count + i, a.found[b]);
    }
}

// The unit's one shot on the whole workgroup through its staging row blockIdx.x.
template <int PLACE, bool LEARN>
__device__ __forceinline__ void serve_group_unit(const KArgs &a, uint8_t *smem, const int64_t *ctl) {
    const int b = (int)blockIdx.x;
    const size_t C = (size_t)a.g.C;
    serve_stage<kEngineBlock>(const_cast<uint8_t *>(a.syn) + (size_t)b * C, ctl_unit(ctl), 0, (int)C);
    __syncthreads();
    engine_shot<kEngineBlock, PLACE, LEARN>(a, b, smem, 0, a.spill + (int64_t)b * a.spill_stride);
    __syncthreads();  // the shot's error row and scalars, for every thread's copy
    serve_results<kEngineBlock>(a, b, ctl_unit(ctl), 0);
}

// The unit's shots on the waves of the workgroup, each wave taking the next
// one (counter in LDS) through its own staging row.
template <int PLACE>
__device__ __forceinline__ void serve_wave_unit(const KArgs &a, uint32_t stride, uint8_t *smem, int64_t *ctl) {
    const uint32_t wave = threadIdx.x >> 6;
    const int b = (int)(blockIdx.x * kEngineWaves + wave);
    int *next = reinterpret_cast<int *>(ctl + kCtlNext);
    for (;;) {
        int i = 0;
        if ((opaque_tid() & 63u) == 0) i = atomicAdd(next, 1);
        i = uniform_i32(i);
        const pel_serve_unit u = ctl_unit(ctl);
        if (i >= u.count) break;
        const size_t C = (size_t)a.g.C;
        serve_stage<64>(const_cast<uint8_t *>(a.syn) + (size_t)b * C, u, i, (int)C);
        gsync<64>();
        engine_shot<64, PLACE, false>(a, b, smem, wave * stride, a.spill + (int64_t)b * a.spill_stride);
        gsync<64>();
        serve_results<64>(a, b, ctl_unit(ctl), i);
    }
}

// The leader's wait for the next unit: its number, or -1 to leave.
__device__ __forceinline__ int64_t serve_claim(const ServeArgs &s, int64_t *ctl) {
    int64_t seen = ctl[kCtlSeen];
    const int64_t idle_from = ctl[kCtlIdle];
    int64_t k = -1;
    int nap = 1;
    for (;;) {
        uint64_t w = __hip_atomic_load(s.claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (w & kServeClosed) break;
        const int64_t c = (int64_t)(w & kServeCount);
        if (c >= seen) seen = sys_load(s.published);
        if (c < seen) {
            if (__hip_atomic_compare_exchange_strong(s.claim, &w, w + 1 + kServeBusy, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                k = c;
                break;
            }
            continue;
        }
        if (__hip_atomic_load(s.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0) break;
        if ((int64_t)wall_clock64() - idle_from > s.idle_ticks && (w & ~kServeClosed) >> kServeBusyShift == 0) {
            if (__hip_atomic_compare_exchange_strong(s.claim, &w, w | kServeClosed, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                break;
            continue;
        }
        for (int z = 0; z < nap; ++z) __builtin_amdgcn_s_sleep(127);  // about 3.4 us each
        nap = nap < 8 ? 2 * nap : 8;
    }
    ctl[kCtlSeen] = seen;
    ctl[kCtlUnit] = k;
    if (k >= 0) {
        const int64_t *src = reinterpret_cast<const int64_t *>(s.units + k % s.capacity);
        for (int w = 0; w < 8; ++w) ctl[kCtlWords + w] = sys_load(src + w);
        ctl[kCtlNext] = 0;
    }
    return k;
}

template <int KINDS>
__global__ void __launch_bounds__(kEngineBlock) serve_kernel(ServeArgs s, const JobDesc *__restrict__ jobs) {
    extern __shared__ __align__(16) uint8_t smem[];
    int64_t *ctl = reinterpret_cast<int64_t *>(smem + s.ctl_lds);
    if (opaque_tid() == 0) {
        ctl[kCtlSeen] = 0;
        ctl[kCtlIdle] = wall_clock64();
    }
    for (;;) {
        if (opaque_tid() == 0) (void)serve_claim(s, ctl);
        __syncthreads();
        if (uniform_i64(ctl[kCtlUnit]) < 0) break;
        const pel_serve_unit u = ctl_unit(ctl);
        // a malformed unit (the host checks them) is marked done without work
        const bool ok = u.job >= 0 && u.job < s.n_jobs && u.count >= 1 && u.count <= PEL_SERVE_WAVE_SHOTS &&
                        u.syndromes && u.errors && u.weights && u.iterations && u.found_count;
        const int kind = ok ? uniform_i32(jobs[u.job].kind) : -1;
        if (ok && ((kind & kKindWave) || u.count == 1)) {
            const JobDesc &jd = jobs[u.job];
            const KArgs &a = jd.a;
#define PEL_SERVE_CASE(K, CALL) \
    else if ((KINDS & kind_bit(K)) && kind == (K)) CALL;
            if (false) {
            }
            PEL_SERVE_CASE(15, (serve_group_unit<15, false>(a, smem, ctl)))
            PEL_SERVE_CASE(7, (serve_group_unit<7, false>(a, smem, ctl)))
            PEL_SERVE_CASE(3, (serve_group_unit<3, false>(a, smem, ctl)))
            PEL_SERVE_CASE(1, (serve_group_unit<1, false>(a, smem, ctl)))
            PEL_SERVE_CASE(0, (serve_group_unit<0, false>(a, smem, ctl)))
            PEL_SERVE_CASE(15 | kKindLearn, (serve_group_unit<15, true>(a, smem, ctl)))
            PEL_SERVE_CASE(7 | kKindLearn, (serve_group_unit<7, true>(a, smem, ctl)))
            PEL_SERVE_CASE(15 | kKindWave, (serve_wave_unit<15>(a, jd.stride, smem, ctl)))
            PEL_SERVE_CASE(7 | kKindWave, (serve_wave_unit<7>(a, jd.stride, smem, ctl)))
#undef PEL_SERVE_CASE
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's result stores are in memory
        __syncthreads();
        if (opaque_tid() == 0) {
            const int64_t k = ctl[kCtlUnit];
            sys_store(s.done + k % s.capacity, k + 1);
            __hip_atomic_fetch_sub(s.claim, kServeBusy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ctl[kCtlIdle] = wall_clock64();
        }
    }
}

typedef void (*ServeFn)(ServeArgs, const JobDesc *);

ServeFn pick_serve(int mask) {
    if ((mask & ~(1 | 8 | 32 | 128)) == 0) return serve_kernel<1 | 8 | 32 | 128>;  // the three-code mix
    return serve_kernel<kKindsAll>;
}

// Work-queue order of the first legs: shots by decreasing syndrome weight
// (heavier syndromes tend to need more iterations, so they start first).
// syndrome_key_kernel: one wave per shot computes the bucket key (weights of
// kKeyBuckets - 1 and more share bucket 0) and clears the shot's re-queue slot.
// order_shots_kernel: one workgroup counting-sorts the keys and resets the
// queue counters.
constexpr int kKeyBuckets = 4096;

// Shots whose first leg converged in fewer iterations than this are re-queued
// behind every first leg, so that shots whose first leg is slow (the ones that
// go on to need many relay legs) start as early as possible.  On the
// 936 x 8785 circuit-shaped workload 99 % of first legs converge within 11
// iterations.
constexpr int kDefaultRequeueBelow = 16;
constexpr int kOrderBlock = 1024;

__global__ void __launch_bounds__(kOrderBlock) syndrome_key_kernel(const uint8_t *syn, int32_t B, int32_t C,
                                                                   int32_t *key, int32_t *rq_slot) {
    const int b = blockIdx.x * (kOrderBlock / 64) + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b >= B) return;
    const uint8_t *s = syn + (size_t)b * C;
    int w = 0;
    for (int i = lane; i < C; i += 64) w += s[i] & 1;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
    if (lane == 0) {
        key[b] = kKeyBuckets - 1 - min(w, kKeyBuckets - 1);
        rq_slot[b] = -1;
    }
}

__global__ void __launch_bounds__(kOrderBlock) order_shots_kernel(const int32_t *key, int32_t B, int32_t *order,
                                                                  int32_t *ctrl) {
    __shared__ int hist[kKeyBuckets];
    __shared__ int wsum[kOrderBlock / 64];
    constexpr int kPer = kKeyBuckets / kOrderBlock;
    const int tid = threadIdx.x;
    for (int k = tid; k < kKeyBuckets; k += kOrderBlock) hist[k] = 0;
    __syncthreads();
    for (int b = tid; b < B; b += kOrderBlock) atomicAdd(&hist[key[b]], 1);
    __syncthreads();
    // Exclusive scan: thread tid owns buckets kPer*tid .. kPer*tid + kPer - 1.
    int c[kPer], loc = 0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        c[u] = hist[kPer * tid + u];
        loc += c[u];
    }
    int inc = loc;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(inc, off, 64);
        if ((tid & 63) >= off) inc += t;
    }
    if ((tid & 63) == 63) wsum[tid >> 6] = inc;
    __syncthreads();
    int run = inc - loc;
    for (int w = 0; w < (tid >> 6); ++w) run += wsum[w];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        hist[kPer * tid + u] = run;
        run += c[u];
    }
    __syncthreads();
    for (int b = tid; b < B; b += kOrderBlock) order[atomicAdd(&hist[key[b]], 1)] = b;
    if (tid < kCtrlWords) ctrl[tid] = 0;
}

// dst[r][v] = src[r][vperm[v]] for rows r: per-variable data into position order.
__global__ void permute_vars_kernel(const float *src, float *dst, const int32_t *vperm, int32_t V, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / V;
        const int32_t v = (int32_t)(i - r * V);
        dst[i] = src[r * V + vperm[v]];
    }
}

typedef void (*KernelFn)(KArgs);

}  // namespace

// The workspace of pel_decode_noalloc: one copy of this array on every GPU the
// library runs on, held until the process ends.  Host code never uses the
// array's name as an address (in host code it is only a stand-in for the
// symbol): scratch_on_device looks up each device's own address once, with
// that device current, and every launch takes that address.
static __device__ __align__(256) uint8_t scratch[PEL_NOALLOC_MAX_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxDevices];

// The address of `scratch` on device dev (the current device).
int scratch_on_device(int dev, uint8_t **out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(PEL_EINVAL, "pel_decode_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(PEL_EDEVICE, std::string("pel_decode_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<uint8_t *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return PEL_OK;
}

// Placements: 15 and 7 keep every state array but the priors (7) in LDS in the
// canonical layout; 3 and 1 keep the variable state (and 3 also the check
// state) in LDS; 0 keeps nothing there.
constexpr int kPlacements[] = {0, 1, 3, 7, 15};

template <int BLOCK>
KernelFn pick_place(int place, bool learn) {
    if (learn) return place == 15 ? relay_decode_kernel<BLOCK, 15, true> : relay_decode_kernel<BLOCK, 7, true>;
    switch (place) {
        case 15: return relay_decode_kernel<BLOCK, 15, false>;
        case 7: return relay_decode_kernel<BLOCK, 7, false>;
        case 3: return relay_decode_kernel<BLOCK, 3, false>;
        case 1: return relay_decode_kernel<BLOCK, 1, false>;
        default: return relay_decode_kernel<BLOCK, 0, false>;
    }
}

// learn: the kernel with learned memory strengths (fast placements only).
KernelFn pick_kernel(int block, int place, bool learn) {
    switch (block) {
        case 64: return pick_place<64>(place, learn);
        case 128: return pick_place<128>(place, learn);
        case 256: return pick_place<256>(place, learn);
        case 512: return pick_place<512>(place, learn);
        default: return pick_place<1024>(place, learn);
    }
}

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

struct StateSizes {
    int64_t sz[kNumArrays];  // LS, CS, SB, L0 bytes per shot
    int64_t nrs, nvs;        // row and variable slices (their descriptors are copied to LDS)
    int64_t C, V;
};

// LDS addresses of the learned-strength arrays (after the slice tables).
struct LearnLds {
    uint32_t flips, unsat, gamma;
    uint32_t slot;  // workspace-slot byte offset of a shifted leg's strengths (fp32 [V])
};

}  // namespace

struct pel_graph {
    int device;
    int32_t C, V, E, sb_words, max_row_deg, cs_shift;
    void *dev_mem;  // one allocation holding every array of d
    GraphDev d;
    // plain CSR for the float16 / float64 decodes: rows (edge id = position) and, per
    // variable, its edge ids ascending
    const int32_t *row_ptr, *row_var, *var_ptr, *var_edges;
    StateSizes state() const {
        // each array has one sentinel entry past its last variable or check
        return {{align_up(8LL * (V + 1), 16), 4LL << cs_shift, align_up(4LL * std::max(sb_words, 1), 16),
                 align_up(4LL * (V + 1), 16)},
                (C + kSlice - 1) / kSlice,
                (V + kSlice - 1) / kSlice,
                C,
                V};
    }
};

namespace {

// Workspace layout: [0,256) queue counters | order [B] | key [B] | rq_slot [B]
// | rq_state [B][2] | rq_post [B,V] (only when shots are parked: re-queue or
// leg-stepped decodes) | rq_gamma [B,V] int8 (parked with learned strengths)
// | rq_bits [B][ceil(V/64)] u64 (leg-stepped) | mlp_frag (learned plans) |
// spill slots
struct WsLayout {
    int64_t order, key, rq_slot, rq_state, rq_post, rq_gamma, rq_bits, mlp_frag, by_pos, spill, total_fixed;
};

WsLayout ws_layout(const pel_graph *g, int64_t B, bool requeue, bool learn, bool legwise = false) {
    const int64_t nb = std::max<int64_t>(B, 1);
    const bool park = requeue || legwise;
    WsLayout w;
    w.order = 256;
    w.key = align_up(w.order + 4 * nb, 256);
    w.rq_slot = align_up(w.key + 4 * nb, 256);
    w.rq_state = align_up(w.rq_slot + 4 * nb, 256);
    w.rq_post = align_up(w.rq_state + 32 * nb, 256);
    w.rq_gamma = align_up(w.rq_post + (park ? 4 * nb * g->V : 0), 256);
    w.rq_bits = align_up(w.rq_gamma + (park && learn ? nb * g->V : 0), 256);
    w.mlp_frag = align_up(w.rq_bits + (legwise ? 8 * nb * ((g->V + 63) / 64) : 0), 256);
    w.by_pos = align_up(w.mlp_frag + (learn ? 4LL * kMlpStride * 64 : 0), 256);  // wj, priors by position
    w.spill = align_up(w.by_pos + 2 * align_up(4LL * g->V, 256), 256);
    w.total_fixed = w.spill;
    return w;
}

// Byte offsets of the state arrays for a placement and of the workgroup's
// reduction scratch (always in LDS, after the arrays); LDS and spill totals.
void place_arrays(const StateSizes &st, int place, int block, bool learn, uint32_t off[kNumArrays],
                  uint32_t *scratch, int64_t *lds, int64_t *spill, LearnLds *ll = nullptr) {
    *lds = 0;
    *spill = 0;
    if ((place & 7) == 7) {  // canonical LDS layout: CS | SB | LS | L0
        off[kCS] = 0;
        off[kSB] = (uint32_t)st.sz[kCS];
        off[kLS] = (uint32_t)(st.sz[kCS] + st.sz[kSB]);
        off[kL0] = (uint32_t)(st.sz[kCS] + st.sz[kSB] + st.sz[kLS]);
        *lds = st.sz[kCS] + st.sz[kSB] + st.sz[kLS] + ((place & (1 << kL0)) ? st.sz[kL0] : 0);
        if (!(place & (1 << kL0))) off[kL0] = 0;
        *spill = st.sz[kCS] + st.sz[kSB];  // the slot holds the copy of a shot's first check pass
    } else {
        for (int k = 0; k < kNumArrays; ++k) {
            if (place & (1 << k)) {
                off[k] = (uint32_t)*lds;
                *lds += st.sz[k];
            } else if (k == kL0) {
                off[k] = 0;  // read from the caller's prior, no copy
            } else {
                off[k] = (uint32_t)*spill;
                *spill += st.sz[k];
            }
        }
    }
    // scratch: s_red doubles, s_flag ints, s_next[2]; then the slice descriptor
    // tables: rows {ent, word, width, 0}, variables {ent, width}; the row flags
    // (u8 per row); the output bits (8 bytes per 64 variable positions)
    *scratch = (uint32_t)*lds;
    *lds += align_up(12LL * (block / 64) + 8, 16) + 16LL * st.nrs + align_up(8LL * st.nvs, 16) + align_up(st.C, 16) +
            align_up(8LL * st.nvs, 16);
    LearnLds l{0, 0, 0, 0};
    if (learn) {  // flip counts u8 [V+1], violated checks u8 [C+1], strength shifts i8 [V]
        l.flips = (uint32_t)*lds;
        *lds += align_up(st.V + 1, 16);
        l.unsat = (uint32_t)*lds;
        *lds += align_up(st.C + 1, 16);
        l.gamma = (uint32_t)*lds;
        *lds += align_up(st.V, 16);
        // the workspace slot: a shifted leg's strengths
        *spill = align_up(*spill, 256);
        l.slot = (uint32_t)*spill;
        *spill += 4 * st.V;
    }
    if (ll) *ll = l;
}

// LDS byte addresses of the slice descriptor tables after the scratch.
uint32_t row_table_addr(uint32_t scratch, int block) {
    return scratch + (uint32_t)align_up(12LL * (block / 64) + 8, 16);
}

// Order of the row slices across the CU's SIMDs.  The check pass gives slice s
// to wave s mod (waves per workgroup), and a workgroup's waves sit on the four
// SIMDs in turn, so slice s runs on SIMD s % 4 for every block size that is a
// multiple of 256.  With rows sorted by degree the widest slices would meet on
// SIMD 0.  The full slices (rorig holds the rows by decreasing degree) are
// reassigned: widest first to the least-loaded SIMD with a free position, then
// pairwise swaps while one lowers the larger of the two SIMDs' chunk totals.
// Rows keep their degree order within a slice; a partial last slice stays last.
void balance_row_slices(const int32_t *chk_ptr, int32_t C, std::vector<int32_t> &rorig) {
    const int32_t nfull = C / kSlice;
    if (nfull < 8) return;
    auto chunks_at = [&](int32_t p) {
        const int32_t d = chk_ptr[rorig[p] + 1] - chk_ptr[rorig[p]];
        return (d + kChunk - 1) / kChunk;
    };
    std::vector<int32_t> chunks(nfull);
    for (int32_t s = 0; s < nfull; ++s) chunks[s] = chunks_at(s * kSlice);
    int64_t load[4] = {0, 0, 0, 0};
    if (C % kSlice) load[nfull % 4] += chunks_at(nfull * kSlice);
    std::vector<int32_t> cls[4];
    size_t cap[4];
    for (int r = 0; r < 4; ++r) cap[r] = r < nfull ? (size_t)((nfull - 1 - r) / 4 + 1) : 0;
    for (int32_t s = 0; s < nfull; ++s) {
        int best = -1;
        for (int r = 0; r < 4; ++r)
            if (cls[r].size() < cap[r] && (best < 0 || load[r] < load[best])) best = r;
        cls[best].push_back(s);
        load[best] += chunks[s];
    }
    auto improve = [&]() {
        for (int x = 0; x < 4; ++x)
            for (int y = 0; y < 4; ++y) {
                if (x == y) continue;
                for (size_t i = 0; i < cls[x].size(); ++i)
                    for (size_t j = 0; j < cls[y].size(); ++j) {
                        const int32_t a = cls[x][i], b = cls[y][j];
                        const int64_t nx = load[x] - chunks[a] + chunks[b], ny = load[y] - chunks[b] + chunks[a];
                        if (std::max(nx, ny) < std::max(load[x], load[y])) {
                            cls[x][i] = b;
                            cls[y][j] = a;
                            load[x] = nx;
                            load[y] = ny;
                            return true;
                        }
                    }
            }
        return false;
    };
    for (int guard = 0; guard < 10000 && improve(); ++guard) {
    }
    std::vector<int32_t> out(rorig);
    for (int32_t p = 0; p < nfull; ++p) {
        const int32_t src = cls[p % 4][(size_t)(p / 4)];
        std::copy(rorig.begin() + (int64_t)src * kSlice, rorig.begin() + (int64_t)(src + 1) * kSlice,
                  out.begin() + (int64_t)p * kSlice);
    }
    rorig.swap(out);
}

// Positions of the variables within each degree class.  At slot k of a row
// slice, a check pass reads the 8-byte state of the k-th variable of each of
// its 64 rows.  An LDS read of 8-byte words serves 16 lanes per pass, and two
// lanes of a pass collide when their variables' positions are equal mod 16 (the
// state array is 8-byte entries over 32 four-byte banks) unless they read the
// same entry.  Variables of equal degree are interchangeable (same slice
// width), so each in turn takes the residue mod 16 that the fewest variables
// it is read together with already hold, among the residues its class still
// has positions for; the class's positions that are equal to r mod 16 then go
// to its variables of residue r in their previous order.
void bank_spread_vars(const int32_t *chk_ptr, const int32_t *chk_var, int32_t C, int32_t V,
                      const std::vector<int32_t> &rorig, const std::vector<int32_t> &vdeg0,
                      std::vector<int32_t> &vperm) {
    std::vector<std::vector<int32_t>> members;  // read group -> variables
    std::vector<std::vector<int32_t>> vgroups(V);
    const int32_t nrs = (C + kSlice - 1) / kSlice;
    for (int32_t sl = 0; sl < nrs; ++sl) {
        int32_t width = 0;
        for (int32_t p = sl * kSlice; p < std::min(C, (sl + 1) * kSlice); ++p)
            width = std::max(width, chk_ptr[rorig[p] + 1] - chk_ptr[rorig[p]]);
        for (int32_t g = 0; g < kSlice / 16; ++g)
            for (int32_t k = 0; k < width; ++k) {
                std::vector<int32_t> mem;
                for (int32_t l = 16 * g; l < 16 * (g + 1); ++l) {
                    const int32_t p = sl * kSlice + l;
                    if (p >= C) break;
                    const int32_t row = rorig[p];
                    if (k < chk_ptr[row + 1] - chk_ptr[row]) mem.push_back(chk_var[chk_ptr[row] + k]);
                }
                if (mem.size() < 2) continue;
                const int32_t id = (int32_t)members.size();
                for (int32_t v : mem) vgroups[v].push_back(id);
                members.push_back(std::move(mem));
            }
    }
    std::vector<int8_t> res(V, -1);
    for (int32_t a = 0; a < V;) {
        int32_t b = a;
        while (b < V && vdeg0[vperm[b]] == vdeg0[vperm[a]]) ++b;
        int32_t cap[16] = {0};
        for (int32_t p = a; p < b; ++p) cap[p & 15]++;
        for (int32_t i = a; i < b; ++i) {
            const int32_t v = vperm[i];
            int32_t cnt[16] = {0};
            for (int32_t g : vgroups[v])
                for (int32_t u : members[g])
                    if (u != v && res[u] >= 0) cnt[res[u]]++;
            int32_t best = -1;
            for (int32_t r = 0; r < 16; ++r)
                if (cap[r] > 0 && (best < 0 || cnt[r] < cnt[best])) best = r;
            res[v] = (int8_t)best;
            cap[best]--;
        }
        std::vector<int32_t> next[16];
        for (int32_t i = a; i < b; ++i) next[res[vperm[i]]].push_back(vperm[i]);
        int32_t taken[16] = {0};
        std::vector<int32_t> out(b - a);
        for (int32_t p = a; p < b; ++p) out[p - a] = next[p & 15][(size_t)taken[p & 15]++];
        std::copy(out.begin(), out.end(), vperm.begin() + a);
        a = b;
    }
}

// Lanes of the rows within each 16-lane group.  The variable pass gathers, for
// the 64 variables of a slice at one slot, each one's check message and code
// word; both sit at an LDS bank fixed by the row's lane (lane mod 32 over 32
// banks, lane mod 64 over 64), so two variables of one gather whose rows share
// a bank collide unless they read the same word.  A row may take any lane of
// its 16-lane group without changing a check-pass gather group (the rows of
// one 16-lane group), so the variable positions above stay as good.  Each row
// in turn, widest first, takes the free lane whose banks the fewest rows it is
// gathered with already hold; then pairs of a group swap lanes while a swap
// lowers that count.
void bank_spread_rows(const int32_t *chk_ptr, const int32_t *chk_var, int32_t C, int32_t V,
                      const std::vector<int32_t> &vposv, std::vector<int32_t> &rorig) {
    std::vector<std::vector<int32_t>> vrows(V);  // by variable position: its rows in slot order
    for (int32_t i = 0; i < C; ++i)
        for (int32_t e = chk_ptr[i]; e < chk_ptr[i + 1]; ++e) vrows[vposv[chk_var[e]]].push_back(i);
    std::vector<std::vector<std::pair<int32_t, int32_t>>> occ(C);  // row -> (row gathered with it, times)
    {
        std::vector<std::vector<int32_t>> cnt(C);
        std::vector<int32_t> rows;
        for (int32_t v0 = 0; v0 < V; v0 += kSlice) {
            const int32_t v1 = std::min(V, v0 + kSlice);
            size_t w = 0;
            for (int32_t v = v0; v < v1; ++v) w = std::max(w, vrows[v].size());
            for (size_t u = 0; u < w; ++u) {
                rows.clear();
                for (int32_t v = v0; v < v1; ++v)
                    if (u < vrows[v].size()) rows.push_back(vrows[v][u]);
                for (int32_t a : rows)
                    for (int32_t b : rows)
                        if (a != b) cnt[a].push_back(b);
            }
        }
        for (int32_t r = 0; r < C; ++r) {
            std::sort(cnt[r].begin(), cnt[r].end());
            for (size_t i = 0; i < cnt[r].size();) {
                size_t j = i;
                while (j < cnt[r].size() && cnt[r][j] == cnt[r][i]) ++j;
                occ[r].emplace_back(cnt[r][i], (int32_t)(j - i));
                i = j;
            }
        }
    }
    std::vector<int32_t> pos(C, -1);
    auto cost = [&](int32_t r, int32_t p) {
        int64_t c = 0;
        for (const auto &o : occ[r]) {
            const int32_t q = pos[o.first];
            if (q < 0) continue;
            if ((q & 31) == (p & 31)) c += o.second;
            if ((q & 63) == (p & 63)) c += o.second;
        }
        return c;
    };
    auto degree = [&](int32_t r) { return chk_ptr[r + 1] - chk_ptr[r]; };
    for (int32_t g0 = 0; g0 < C; g0 += 16) {
        const int32_t n = std::min(C - g0, 16);
        std::vector<int32_t> rows(rorig.begin() + g0, rorig.begin() + g0 + n);
        std::stable_sort(rows.begin(), rows.end(), [&](int32_t x, int32_t y) { return degree(x) > degree(y); });
        std::vector<bool> used(n, false);
        for (int32_t r : rows) {
            int32_t best = -1;
            int64_t bc = 0;
            for (int32_t l = 0; l < n; ++l) {
                if (used[l]) continue;
                const int64_t c = cost(r, g0 + l);
                if (best < 0 || c < bc) best = l, bc = c;
            }
            used[best] = true;
            pos[r] = g0 + best;
            rorig[g0 + best] = r;
        }
    }
    for (int sweep = 0; sweep < 4; ++sweep) {
        bool moved = false;
        for (int32_t g0 = 0; g0 < C; g0 += 16) {
            const int32_t n = std::min(C - g0, 16);
            for (int32_t i = 0; i < n; ++i)
                for (int32_t j = i + 1; j < n; ++j) {
                    const int32_t a = rorig[g0 + i], b = rorig[g0 + j];
                    const int64_t before = cost(a, g0 + i) + cost(b, g0 + j);
                    pos[a] = g0 + j;
                    pos[b] = g0 + i;
                    const int64_t after = cost(a, g0 + j) + cost(b, g0 + i);
                    if (after < before) {
                        rorig[g0 + i] = b;
                        rorig[g0 + j] = a;
                        moved = true;
                    } else {
                        pos[a] = g0 + i;
                        pos[b] = g0 + j;
                    }
                }
        }
        if (!moved) break;
    }
}

}  // namespace

extern "C" {

int pel_abi_version(void) { return PEL_ABI_VERSION; }

const char *pel_last_error(void) { return g_last_error.c_str(); }

int pel_graph_create(const int32_t *chk_ptr, const int32_t *chk_var, int32_t C, int32_t V, int32_t E,
                     pel_graph **out) {
    if (!out) return fail(PEL_EINVAL, "pel_graph_create: out is NULL");
    *out = nullptr;
    if (!chk_ptr || (!chk_var && E > 0)) return fail(PEL_EINVAL, "pel_graph_create: NULL CSR array");
    if (C <= 0 || V <= 0 || E < 0) return fail(PEL_EINVAL, "pel_graph_create: empty matrix");
    if (chk_ptr[0] != 0 || chk_ptr[C] != E)
        return fail(PEL_EINVAL, "pel_graph_create: chk_ptr must start at 0 and end at E");
    int32_t maxdeg = 0;
    for (int32_t i = 0; i < C; ++i) {
        const int32_t d = chk_ptr[i + 1] - chk_ptr[i];
        if (d < 0) return fail(PEL_EINVAL, "pel_graph_create: chk_ptr not non-decreasing");
        maxdeg = std::max(maxdeg, d);
    }
    for (int32_t e = 0; e < E; ++e)
        if (chk_var[e] < 0 || chk_var[e] >= V)
            return fail(PEL_EINVAL, "pel_graph_create: column index out of range");

    // Row positions: checks sorted by decreasing degree (stable).
    std::vector<int32_t> rorig(C), rpos(C), rdeg(C);
    std::iota(rorig.begin(), rorig.end(), 0);
    std::stable_sort(rorig.begin(), rorig.end(), [&](int32_t x, int32_t y) {
        return (chk_ptr[x + 1] - chk_ptr[x]) > (chk_ptr[y + 1] - chk_ptr[y]);
    });
    balance_row_slices(chk_ptr, C, rorig);
    for (int32_t p = 0; p < C; ++p) {
        rpos[rorig[p]] = p;
        rdeg[p] = chk_ptr[rorig[p] + 1] - chk_ptr[rorig[p]];
    }
    // Variable positions: variables sorted by decreasing degree (stable).
    std::vector<int32_t> vdeg0(V, 0);
    for (int32_t e = 0; e < E; ++e) vdeg0[chk_var[e]]++;
    std::vector<int32_t> vperm(V), vposv(V), vdeg(V);
    std::iota(vperm.begin(), vperm.end(), 0);
    std::stable_sort(vperm.begin(), vperm.end(), [&](int32_t x, int32_t y) { return vdeg0[x] > vdeg0[y]; });
    bank_spread_vars(chk_ptr, chk_var, C, V, rorig, vdeg0, vperm);
    for (int32_t v = 0; v < V; ++v) {
        vposv[vperm[v]] = v;
        vdeg[v] = vdeg0[vperm[v]];
    }
    bank_spread_rows(chk_ptr, chk_var, C, V, vposv, rorig);
    for (int32_t p = 0; p < C; ++p) {
        rpos[rorig[p]] = p;
        rdeg[p] = chk_ptr[rorig[p] + 1] - chk_ptr[rorig[p]];
    }
    // Row slices: width padded to whole chunks; edge-code words of 16 edges.
    const int32_t nrs = (C + kSlice - 1) / kSlice;
    std::vector<int4> rs(nrs);
    int64_t nidx = 0, nword = 0;
    for (int32_t s = 0; s < nrs; ++s) {
        int32_t w = 0;
        for (int32_t p = s * kSlice; p < std::min(C, (s + 1) * kSlice); ++p) w = std::max(w, rdeg[p]);
        const int32_t wpad = (w + kChunk - 1) / kChunk * kChunk;
        rs[s] = make_int4((int32_t)nidx, (int32_t)nword, wpad, 0);
        nidx += (int64_t)wpad * kSlice;
        nword += (int64_t)((wpad + kEdgesPerWord - 1) / kEdgesPerWord) * kSlice;
    }
    // Canonical LDS layout (must match pel_graph::state and place_arrays).
    int32_t cs_shift = 4;
    while ((1LL << cs_shift) < 4LL * (C + 1)) ++cs_shift;  // plane stride: a power of two >= 4 (C + 1)
    const int64_t cs_bytes = 4LL << cs_shift;
    const int64_t sb_bytes = align_up(4LL * std::max<int64_t>(nword, 1), 16);
    const int64_t lsb = cs_bytes + sb_bytes;
    if (nidx > INT32_MAX / 2 || lsb + 8LL * (V + 1) > (1LL << 26))
        return fail(PEL_EINVAL, "pel_graph_create: graph too large for 32-bit state addressing");
    // padding -> the sentinel variable at position V
    std::vector<uint32_t> row_ent(std::max<int64_t>(nidx, 1), (uint32_t)(lsb + 8LL * V));
    for (int32_t p = 0; p < C; ++p) {
        const int32_t s = p / kSlice, lane = p % kSlice, i = rorig[p];
        for (int32_t k = 0; k < rdeg[p]; ++k)
            row_ent[rs[s].x + (int64_t)k * kSlice + lane] = (uint32_t)(lsb + 8LL * vposv[chk_var[chk_ptr[i] + k]]);
    }
    // Variable slices (position order); entries in ascending edge id (= ascending check).
    const int32_t nvs = (V + kSlice - 1) / kSlice;
    std::vector<int4> vs(nvs);
    int64_t nve = 0;
    for (int32_t s = 0; s < nvs; ++s) {
        const int32_t w = vdeg[s * kSlice];  // widest variable comes first
        vs[s] = make_int4((int32_t)nve, w, 0, 0);
        nve += (int64_t)w * kSlice;
    }
    if (nve > INT32_MAX / 2) return fail(PEL_EINVAL, "pel_graph_create: graph too large for 32-bit indexing");
    // padding -> the sentinel check at position C (all four messages +0), code word 0
    const uint32_t pad_sb = (uint32_t)(cs_bytes << 5);
    std::vector<uint2> var_ent(std::max<int64_t>(nve, 1), make_uint2(4u * (uint32_t)C, pad_sb));
    std::vector<int32_t> fill(V, 0);
    for (int32_t i = 0; i < C; ++i) {
        const int32_t p = rpos[i], s = p / kSlice, lane = p % kSlice;
        for (int32_t e = chk_ptr[i]; e < chk_ptr[i + 1]; ++e) {
            const int32_t v = vposv[chk_var[e]], k = e - chk_ptr[i];
            const int64_t word = (int64_t)rs[s].y + (k / kEdgesPerWord) * kSlice + lane;
            const uint32_t sbaddr =
                (uint32_t)((cs_bytes + 4 * word) << 5) | (uint32_t)(2 * (kEdgesPerWord - 1 - k % kEdgesPerWord));
            var_ent[vs[v / kSlice].x + (int64_t)fill[v]++ * kSlice + (v % kSlice)] = make_uint2(4u * (uint32_t)p, sbaddr);
        }
    }
    // Plain CSR for the float16 / float64 decodes.
    std::vector<int32_t> cvp(V + 1, 0), cve(std::max<int32_t>(E, 1), 0);
    for (int32_t e = 0; e < E; ++e) cvp[chk_var[e] + 1]++;
    for (int32_t j = 0; j < V; ++j) cvp[j + 1] += cvp[j];
    {
        std::vector<int32_t> cf(cvp.begin(), cvp.end() - 1);
        for (int32_t e = 0; e < E; ++e) cve[cf[chk_var[e]]++] = e;  // ascending edge id per variable
    }
    // One device allocation for every array.
    int64_t off = 0;
    auto take = [&](int64_t bytes) {
        const int64_t o = off;
        off = align_up(off + std::max<int64_t>(bytes, 4), 256);
        return o;
    };
    const int64_t o_ro = take(4LL * C), o_rd = take(4LL * C), o_rs = take(16LL * nrs),
                  o_re = take(4LL * (int64_t)row_ent.size()), o_vp = take(4LL * V), o_vq = take(4LL * V),
                  o_vd = take(4LL * V), o_vs = take(16LL * nvs), o_ve = take(8LL * (int64_t)var_ent.size()),
                  o_crp = take(4LL * (C + 1)), o_crv = take(4LL * std::max<int32_t>(E, 1)), o_cvp = take(4LL * (V + 1)),
                  o_cve = take(4LL * std::max<int32_t>(E, 1));
    const int64_t total = off;
    std::vector<uint8_t> host((size_t)total, 0);
    memcpy(host.data() + o_ro, rorig.data(), 4ULL * C);
    memcpy(host.data() + o_rd, rdeg.data(), 4ULL * C);
    memcpy(host.data() + o_rs, rs.data(), 16ULL * nrs);
    memcpy(host.data() + o_re, row_ent.data(), 4ULL * row_ent.size());
    memcpy(host.data() + o_vp, vperm.data(), 4ULL * V);
    memcpy(host.data() + o_vq, vposv.data(), 4ULL * V);
    memcpy(host.data() + o_vd, vdeg.data(), 4ULL * V);
    memcpy(host.data() + o_vs, vs.data(), 16ULL * nvs);
    memcpy(host.data() + o_ve, var_ent.data(), 8ULL * var_ent.size());
    memcpy(host.data() + o_crp, chk_ptr, 4ULL * (C + 1));
    if (E > 0) memcpy(host.data() + o_crv, chk_var, 4ULL * E);
    memcpy(host.data() + o_cvp, cvp.data(), 4ULL * (V + 1));
    memcpy(host.data() + o_cve, cve.data(), 4ULL * cve.size());

    pel_graph *g = new pel_graph();
    hipError_t e = hipGetDevice(&g->device);
    if (e == hipSuccess) e = hipMalloc(&g->dev_mem, (size_t)total);
    if (e != hipSuccess) {
        delete g;
        return fail(PEL_ENOMEM, std::string("pel_graph_create: ") + hipGetErrorString(e));
    }
    e = hipMemcpy(g->dev_mem, host.data(), (size_t)total, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(g->dev_mem);
        delete g;
        return fail(PEL_EDEVICE, std::string("pel_graph_create: hipMemcpy: ") + hipGetErrorString(e));
    }
    const uint8_t *base = static_cast<const uint8_t *>(g->dev_mem);
    g->C = C;
    g->V = V;
    g->E = E;
    g->sb_words = (int32_t)nword;
    g->max_row_deg = maxdeg;
    g->cs_shift = cs_shift;
    g->d.cs_shift = (uint32_t)cs_shift;
    g->d.C = C;
    g->d.V = V;
    g->d.E = E;
    g->d.sbb = (uint32_t)cs_bytes;
    g->d.lsb = (uint32_t)lsb;
    g->d.row_orig = reinterpret_cast<const int32_t *>(base + o_ro);
    g->d.row_deg = reinterpret_cast<const int32_t *>(base + o_rd);
    g->d.rslice = reinterpret_cast<const int4 *>(base + o_rs);
    g->d.row_ent = reinterpret_cast<const uint32_t *>(base + o_re);
    g->d.vperm = reinterpret_cast<const int32_t *>(base + o_vp);
    g->d.vpos = reinterpret_cast<const int32_t *>(base + o_vq);
    g->d.var_deg = reinterpret_cast<const int32_t *>(base + o_vd);
    g->d.vslice = reinterpret_cast<const int4 *>(base + o_vs);
    g->d.var_ent = reinterpret_cast<const uint2 *>(base + o_ve);
    g->row_ptr = reinterpret_cast<const int32_t *>(base + o_crp);
    g->row_var = reinterpret_cast<const int32_t *>(base + o_crv);
    g->var_ptr = reinterpret_cast<const int32_t *>(base + o_cvp);
    g->var_edges = reinterpret_cast<const int32_t *>(base + o_cve);
    *out = g;
    return PEL_OK;
}

int pel_graph_destroy(pel_graph *g) {
    if (!g) return PEL_OK;
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != g->device) (void)hipSetDevice(g->device);
    hipError_t e = hipFree(g->dev_mem);
    if (cur != g->device) (void)hipSetDevice(cur);
    delete g;
    if (e != hipSuccess) return fail(PEL_EDEVICE, std::string("pel_graph_destroy: ") + hipGetErrorString(e));
    return PEL_OK;
}

int pel_graph_info(const pel_graph *g, int32_t *C, int32_t *V, int32_t *E, int64_t *var_state_bytes,
                   int64_t *chk_state_bytes, int64_t *sign_bytes) {
    if (!g) return fail(PEL_EINVAL, "pel_graph_info: NULL graph");
    const StateSizes s = g->state();
    if (C) *C = g->C;
    if (V) *V = g->V;
    if (E) *E = g->E;
    if (var_state_bytes) *var_state_bytes = s.sz[kLS] + s.sz[kL0];
    if (chk_state_bytes) *chk_state_bytes = s.sz[kCS];
    if (sign_bytes) *sign_bytes = s.sz[kSB];
    return PEL_OK;
}

int pel_graph_permute_vars(const pel_graph *g, const float *src, float *dst, int64_t rows, void *stream) {
    if (!g || (!src && rows > 0) || (!dst && rows > 0) || rows < 0)
        return fail(PEL_EINVAL, "pel_graph_permute_vars: bad argument");
    if (rows == 0) return PEL_OK;
    const int64_t n = rows * (int64_t)g->V;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(permute_vars_kernel, dim3((unsigned)blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       src, dst, g->d.vperm, g->V, n);
    PEL_HIP(hipGetLastError());
    return PEL_OK;
}

int pel_plan_decode(const pel_graph *g, int32_t B, int32_t block_override, int32_t placement_override,
                    int32_t flags, pel_plan *out) {
    if (!g || !out) return fail(PEL_EINVAL, "pel_plan_decode: NULL argument");
    if (B < 0) return fail(PEL_EINVAL, "pel_plan_decode: negative batch");
    int dev = 0;
    PEL_HIP(hipGetDevice(&dev));
    if (dev != g->device) return fail(PEL_EINVAL, "pel_plan_decode: current device differs from the graph's");
    int max_lds = 0, n_cu = 0;
    PEL_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    PEL_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    int block = block_override;
    if (block <= 0) {
        const int want = std::max(g->C, g->V / 8);
        block = 64;
        while (block < want && block < 1024) block *= 2;
    }
    if (block != 64 && block != 128 && block != 256 && block != 512 && block != 1024)
        return fail(PEL_EINVAL, "pel_plan_decode: block must be 64, 128, 256, 512 or 1024");
    if (flags & ~(PEL_PLAN_LEARNED | PEL_PLAN_LEGWISE)) return fail(PEL_EINVAL, "pel_plan_decode: unknown flags");
    const bool learn = (flags & PEL_PLAN_LEARNED) != 0;
    const bool legwise = (flags & PEL_PLAN_LEGWISE) != 0;
    const StateSizes st = g->state();
    const int64_t budget = (int64_t)max_lds;
    int place = -1;
    for (int p : kPlacements) {
        if (learn && (p & 7) != 7) continue;  // learned strengths run in the fast placements
        uint32_t o[kNumArrays], sc = 0;
        int64_t l = 0, s = 0;
        place_arrays(st, p, block, learn, o, &sc, &l, &s);
        if (l <= budget) place = p;
    }
    if (place < 0)
        return fail(PEL_EINVAL, "pel_plan_decode: learned strengths need the graph's state in LDS, which does not fit");
    if (placement_override >= 0) {
        bool known = false;
        for (int p : kPlacements) known |= (p == placement_override);
        if (!known) return fail(PEL_EINVAL, "pel_plan_decode: placement must be 0, 1, 3, 7 or 15");
        if (placement_override > place || (learn && (placement_override & 7) != 7))
            return fail(PEL_EINVAL, "pel_plan_decode: requested placement exceeds the LDS budget");
        place = placement_override;
    }
    uint32_t offs[kNumArrays], scr_off = 0;
    int64_t lds = 0, spill = 0;
    place_arrays(st, place, block, learn, offs, &scr_off, &lds, &spill);
    KernelFn fn = pick_kernel(block, place, learn);
    {
        // LDS is addressed absolutely (the dynamic region must start at address 0),
        // and fast placements use the layout the graph's entries were baked for.
        hipFuncAttributes fa;
        PEL_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(fn)));
        if (fa.sharedSizeBytes != 0 ||
            ((place & 7) == 7 && (offs[kLS] != g->d.lsb || offs[kSB] != g->d.sbb || offs[kCS] != 0)))
            return fail(PEL_EDEVICE, "pel_plan_decode: LDS layout does not match the graph's baked addresses");
    }
    if (lds > 65536) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
        (void)hipGetLastError();
    }
    int per_cu = 0;
    PEL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(fn), block,
                                                         (size_t)lds));
    if (per_cu <= 0) return fail(PEL_EDEVICE, "pel_plan_decode: kernel cannot be resident (LDS/registers)");
    int64_t grid = (int64_t)per_cu * n_cu;
    grid = std::min<int64_t>(grid, std::max<int32_t>(B, 1));
    out->requeue_below = kDefaultRequeueBelow;
    out->learned = learn ? 1 : 0;
    out->legwise = legwise ? 1 : 0;
    const WsLayout w = ws_layout(g, B, out->requeue_below > 0, learn, legwise);
    out->block = block;
    out->grid = (int32_t)grid;
    out->placement = place;
    out->lds_bytes = (int32_t)lds;
    out->spill_stride = align_up(spill, 256);
    out->workspace_bytes = w.total_fixed + out->spill_stride * grid;
    return PEL_OK;
}

}  // extern "C"

namespace {

// Checks of one batch that do not depend on the launch layout.
int check_batch(const std::string &w_, const pel_graph *g, const pel_params *p, const pel_batch *b, bool learned_plan,
                bool noalloc) {
    if (!g || !p || !b) return fail(PEL_EINVAL, w_ + ": NULL argument");
    if (b->B < 0) return fail(PEL_EINVAL, w_ + ": negative batch");
    if (p->pre_iter < 0 || p->num_sets < 0 || p->set_max_iter < 0)
        return fail(PEL_EINVAL, w_ + ": iteration counts must be >= 0");
    if (p->stop_nconv < 1) return fail(PEL_EINVAL, w_ + ": stop_nconv must be >= 1");
    if (b->B == 0) return PEL_OK;
    if (!b->syndromes || !b->prior_llr || !b->wj || !b->errors || !b->weights || !b->iterations ||
        !b->found_count || (!noalloc && !b->workspace))
        return fail(PEL_EINVAL, w_ + ": NULL buffer");
    if (noalloc && b->workspace)
        return fail(PEL_EINVAL, w_ + ": takes no workspace (it uses the library's own); pass NULL");
    if (p->num_sets > 0 && !p->skip_relay_legs && !b->gamma_legs)
        return fail(PEL_EINVAL, w_ + ": gamma_legs is NULL with num_sets > 0");
    if (b->gamma_per_row && b->gamma_by_position)
        return fail(PEL_EINVAL, w_ + ": per-shot gamma rows are taken in the caller's order");
    if (p->use_alpha && !p->alpha_t) return fail(PEL_EINVAL, w_ + ": alpha_t is NULL with use_alpha");
    if (b->learned_legs < 0) return fail(PEL_EINVAL, w_ + ": learned_legs must be >= 0");
    if ((b->learned_legs > 0 || b->mlp || b->feat_out || b->hard_out || b->gamma_out) && !learned_plan)
        return fail(PEL_EINVAL, w_ + ": learned strengths need a plan made with PEL_PLAN_LEARNED");
    if ((b->mlp || b->feat_out || b->hard_out || b->gamma_out) && b->learned_legs == 0)
        return fail(PEL_EINVAL, w_ + ": learned strengths or records with learned_legs == 0");
    if (b->gamma_out && !b->mlp) return fail(PEL_EINVAL, w_ + ": gamma_out needs network parameters");
    if (b->mlp && !(b->learned_lo <= b->learned_hi))
        return fail(PEL_EINVAL, w_ + ": learned_lo must not exceed learned_hi");
    if (!(b->learned_last_damp >= 0.f && b->learned_last_damp <= 1.f))
        return fail(PEL_EINVAL, w_ + ": learned_last_damp must lie in [0, 1]");
    if (b->mlp && (p->skip_relay_legs || p->num_sets < b->learned_legs))
        return fail(PEL_EINVAL, w_ + ": learned legs need that many relay legs");
    int dev = 0;
    PEL_HIP(hipGetDevice(&dev));
    if (dev != g->device) return fail(PEL_EINVAL, w_ + ": current device differs from the graph's");
    return PEL_OK;
}

// A shot group's layout: the arrays' offsets, the reduction scratch and the
// learned-strength arrays in LDS, the LDS and workspace-slot bytes.
struct GroupLayout {
    uint32_t offs[kNumArrays], scr_off;
    int64_t lds, spill;
    LearnLds ll;
};

GroupLayout group_layout(const pel_graph *g, int place, int group, bool learn) {
    GroupLayout L;
    place_arrays(g->state(), place, group, learn, L.offs, &L.scr_off, &L.lds, &L.spill, &L.ll);
    return L;
}

// The kernel arguments of one batch with workspace ws laid out as w.
void fill_args(KArgs &a, const pel_graph *g, const pel_params *p, const pel_batch *b, uint8_t *ws,
               const WsLayout &w, bool learn, int32_t requeue_below, int group, const GroupLayout &L,
               int64_t spill_stride, bool legwise = false) {
    const StateSizes st = g->state();
    memset(&a, 0, sizeof a);
    a.g = g->d;
    a.B = b->B;
    a.pre_iter = p->pre_iter;
    a.num_sets = p->num_sets;
    a.set_max_iter = p->set_max_iter;
    a.stop_nconv = p->stop_nconv;
    a.plain = p->plain;
    a.use_alpha = p->use_alpha;
    a.use_beta = p->use_beta;
    a.last_leg = p->last_leg;
    a.skip_relay_legs = p->skip_relay_legs;
    a.gamma0 = p->gamma0;
    a.beta = p->beta;
    a.alpha_t = p->alpha_t;
    a.syn = b->syndromes;
    a.prior = b->prior_llr;
    a.wj = b->wj;
    a.prior_batched = b->prior_batched;
    if (!b->prior_batched) {
        a.wj_pos = reinterpret_cast<const float *>(ws + w.by_pos);
        a.prior_pos = reinterpret_cast<const float *>(ws + w.by_pos + align_up(4LL * g->V, 256));
    }
    a.gamma_legs = b->gamma_legs;
    a.gamma_per_row = b->gamma_per_row;
    a.gamma_by_position = b->gamma_by_position;
    a.init_gamma = b->init_gamma;
    a.init_post = b->init_posterior;
    a.errors = b->errors;
    a.weights = b->weights;
    a.iters = b->iterations;
    a.found = b->found_count;
    a.last_post = b->last_posterior;
    a.shot_clock = b->shot_clock;
    a.ctrl = reinterpret_cast<int32_t *>(ws);
    a.order = reinterpret_cast<int32_t *>(ws + w.order);
    a.leg_step = legwise ? p->leg_step - 1 : -1;
    a.requeue_below = a.leg_step >= 0 ? 0 : requeue_below;
    const bool park = a.requeue_below > 0 || a.leg_step >= 0;
    a.rq_slot = reinterpret_cast<int32_t *>(ws + w.rq_slot);
    a.rq_state = reinterpret_cast<int4 *>(ws + w.rq_state);
    a.rq_post = park ? reinterpret_cast<float *>(ws + w.rq_post) : nullptr;
    a.rq_gamma = park && learn ? reinterpret_cast<int8_t *>(ws + w.rq_gamma) : nullptr;
    a.rq_bits = a.leg_step >= 0 ? reinterpret_cast<uint64_t *>(ws + w.rq_bits) : nullptr;
    a.mlp_frag = learn ? reinterpret_cast<const uint32_t *>(ws + w.mlp_frag) : nullptr;
    a.mlp = b->mlp;
    a.learned_legs = learn ? b->learned_legs : 0;
    a.feat_out = b->feat_out;
    a.hard_out = b->hard_out;
    a.gamma_out = b->gamma_out;
    a.flips_lds = L.ll.flips;
    a.unsat_lds = L.ll.unsat;
    a.gamma_lds = L.ll.gamma;
    a.gamma_slot = L.ll.slot;
    a.learned_lo = b->learned_lo;
    a.learned_hi = b->learned_hi;
    a.last_scale = 1.f - b->learned_last_damp;
    a.spill = ws + w.spill;
    a.spill_stride = spill_stride;
    for (int k = 0; k < kNumArrays; ++k) a.off[k] = L.offs[k];
    a.scratch_off = L.scr_off;
    a.rtab = row_table_addr(L.scr_off, group);
    a.vtab = a.rtab + (uint32_t)(16LL * st.nrs);
    a.rowinfo = a.vtab + (uint32_t)align_up(8LL * st.nvs, 16);
    a.best_lds = a.rowinfo + (uint32_t)align_up(st.C, 16);
}

// The per-decode launches before the decode kernel: syndrome weights and the
// first-leg queue order (which also reset the queue counters), the shared
// weights and priors in position order, and the network's operand layout.
// (order = false: the serving engine, which takes no shot from the queue.)
int launch_prep(const pel_graph *g, const pel_batch *b, const KArgs &a, uint8_t *ws, const WsLayout &w,
                hipStream_t s, bool order = true) {
    if (order) {
        int32_t *key = reinterpret_cast<int32_t *>(ws + w.key);
        const int key_blocks = (int)((b->B + kOrderBlock / 64 - 1) / (kOrderBlock / 64));
        hipLaunchKernelGGL(syndrome_key_kernel, dim3(key_blocks), dim3(kOrderBlock), 0, s, b->syndromes, b->B, g->C,
                           key, a.rq_slot);
        PEL_HIP(hipGetLastError());
        hipLaunchKernelGGL(order_shots_kernel, dim3(1), dim3(kOrderBlock), 0, s, key, b->B,
                           const_cast<int32_t *>(a.order), a.ctrl);
        PEL_HIP(hipGetLastError());
    }
    if (!b->prior_batched) {  // the shared weights and priors in position order, once per decode
        const unsigned pblocks = (unsigned)std::min<int64_t>((g->V + 255) / 256, 1024);
        hipLaunchKernelGGL(permute_vars_kernel, dim3(pblocks), dim3(256), 0, s, b->wj, const_cast<float *>(a.wj_pos),
                           g->d.vperm, g->V, (int64_t)g->V);
        PEL_HIP(hipGetLastError());
        hipLaunchKernelGGL(permute_vars_kernel, dim3(pblocks), dim3(256), 0, s, b->prior_llr,
                           const_cast<float *>(a.prior_pos), g->d.vperm, g->V, (int64_t)g->V);
        PEL_HIP(hipGetLastError());
    }
    if (b->mlp) {  // the network as per-lane MFMA operands, once per decode
        hipLaunchKernelGGL(mlp_frag_kernel, dim3(1), dim3(64), 0, s, b->mlp,
                           reinterpret_cast<uint32_t *>(ws + w.mlp_frag));
        PEL_HIP(hipGetLastError());
    }
    return PEL_OK;
}

// pel_decode and pel_decode_noalloc: the same launches, with the workspace from
// the batch (pel_decode) or the library's per-GPU scratch array (noalloc).
int decode_impl(const char *who, const pel_graph *g, const pel_params *p, const pel_batch *b, const pel_plan *plan,
                bool noalloc, void *stream) {
    const std::string w_ = who;
    if (!plan) return fail(PEL_EINVAL, w_ + ": NULL argument");
    int rc = check_batch(w_, g, p, b, plan->learned != 0, noalloc);
    if (rc != PEL_OK || b->B == 0) return rc;
    bool known = false;
    for (int pl : kPlacements) known |= (pl == plan->placement);
    if (plan->block != 64 && plan->block != 128 && plan->block != 256 && plan->block != 512 && plan->block != 1024)
        return fail(PEL_EINVAL, w_ + ": plan block must be 64, 128, 256, 512 or 1024");
    const bool learn = plan->learned != 0;
    const GroupLayout L = group_layout(g, plan->placement, plan->block, learn);
    if (plan->requeue_below < 0) return fail(PEL_EINVAL, w_ + ": requeue_below must be >= 0");
    if (p->leg_step < 0 || p->leg_step > p->num_sets + 1)
        return fail(PEL_EINVAL, w_ + ": leg_step must lie in 0 .. num_sets + 1");
    if (p->leg_step > 0 && !plan->legwise)
        return fail(PEL_EINVAL, w_ + ": leg-stepped decodes need a plan made with PEL_PLAN_LEGWISE");
    if (p->leg_step > 0 && (b->feat_out || b->hard_out || b->shot_clock || p->skip_relay_legs))
        return fail(PEL_EINVAL, w_ + ": leg-stepped decodes record no traces or clocks and run the relay legs");
    const WsLayout w = ws_layout(g, b->B, plan->requeue_below > 0, learn, plan->legwise != 0);
    const int64_t need = w.total_fixed + plan->spill_stride * plan->grid;
    if (!known || (learn && (plan->placement & 7) != 7) || L.lds != plan->lds_bytes ||
        align_up(L.spill, 256) != plan->spill_stride || plan->grid <= 0 || (!noalloc && plan->workspace_bytes < need))
        return fail(PEL_EINVAL, w_ + ": plan does not match this graph and batch size");
    uint8_t *ws = static_cast<uint8_t *>(b->workspace);
    if (noalloc) {
        if (need > (int64_t)PEL_NOALLOC_MAX_BYTES)
            return fail(PEL_EINVAL, w_ + ": this decode needs " + std::to_string(need) +
                                        " bytes of workspace, more than the " + std::to_string(PEL_NOALLOC_MAX_BYTES) +
                                        " the library holds per GPU (use fewer shots, a plan with requeue_below = 0, "
                                        "or pel_decode with a workspace)");
        int dev = 0;
        PEL_HIP(hipGetDevice(&dev));
        rc = scratch_on_device(dev, &ws);
        if (rc != PEL_OK) return rc;
    }
    KArgs a;
    fill_args(a, g, p, b, ws, w, learn, plan->requeue_below, plan->block, L, plan->spill_stride, p->leg_step > 0);
    hipStream_t s = static_cast<hipStream_t>(stream);
    rc = launch_prep(g, b, a, ws, w, s);
    if (rc != PEL_OK) return rc;
    if (b->ev_begin) PEL_HIP(hipEventRecord(static_cast<hipEvent_t>(b->ev_begin), s));
    KernelFn fn = pick_kernel(plan->block, plan->placement, learn);
    hipLaunchKernelGGL(fn, dim3(plan->grid), dim3(plan->block), (size_t)plan->lds_bytes, s, a);
    PEL_HIP(hipGetLastError());
    if (b->ev_end) PEL_HIP(hipEventRecord(static_cast<hipEvent_t>(b->ev_end), s));
    return PEL_OK;
}

// The role of one job in an engine launch: a shot per wave (a small code whose
// single-shot plan is one wave, in a fast placement that fits sixteen times in
// the LDS budget), else a shot per workgroup in the largest placement that
// fits (a fast one for learned strengths).  Returns the LDS bytes the job needs
// (per workgroup), or -1 when it does not fit.
int64_t engine_role(const pel_graph *g, bool learn, int64_t budget, int *group, int *place, GroupLayout *L) {
    const int want = std::max(g->C, g->V / 8);
    if (!learn && want <= 64) {
        for (int p : {15, 7}) {
            const GroupLayout l = group_layout(g, p, 64, false);
            if (kEngineWaves * align_up(l.lds, 16) <= budget) {
                *group = 64;
                *place = p;
                *L = l;
                return kEngineWaves * align_up(l.lds, 16);
            }
        }
    }
    int best = -1;
    GroupLayout lb{};
    for (int p : kPlacements) {
        if (learn && (p & 7) != 7) continue;
        const GroupLayout l = group_layout(g, p, kEngineBlock, learn);
        if (l.lds <= budget) {
            best = p;
            lb = l;
        }
    }
    if (best < 0) return -1;
    *group = kEngineBlock;
    *place = best;
    *L = lb;
    return lb.lds;
}

// Plans an engine launch over n jobs: pel_decode_jobs over shots[i] shots
// (serve = false), or the serving engine (serve = true: every workgroup that
// can be resident, and each job's batch holds its staging rows).
int plan_engine(const std::string &w_, const pel_graph *const *graphs, const int32_t *shots, const int32_t *flags,
                int32_t n, pel_jobs_plan *out, bool serve) {
    if (!graphs || (!serve && !shots) || !out) return fail(PEL_EINVAL, w_ + ": NULL argument");
    if (n < 1 || n > kMaxJobs) return fail(PEL_EINVAL, w_ + ": between 1 and PEL_MAX_JOBS jobs");
    int dev = 0, max_lds = 0, n_cu = 0;
    PEL_HIP(hipGetDevice(&dev));
    PEL_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    PEL_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    memset(out, 0, sizeof *out);
    const int64_t ctl = serve ? kServeCtlBytes : 16;
    const int64_t budget = (int64_t)max_lds - ctl;  // the control words sit after every job's LDS
    int64_t lds = 0, units = 0;
    for (int32_t i = 0; i < n; ++i) {
        const pel_graph *g = graphs[i];
        if (!g) return fail(PEL_EINVAL, w_ + ": NULL graph");
        if (g->device != dev) return fail(PEL_EINVAL, w_ + ": a graph is on another device");
        if (!serve && shots[i] < 0) return fail(PEL_EINVAL, w_ + ": negative batch");
        const int32_t fl = flags ? flags[i] : 0;
        if (fl & ~PEL_PLAN_LEARNED) return fail(PEL_EINVAL, w_ + ": unknown flags");
        const bool learn = (fl & PEL_PLAN_LEARNED) != 0;
        int group = 0, place = -1;
        GroupLayout L;
        const int64_t need = engine_role(g, learn, budget, &group, &place, &L);
        if (need < 0) return fail(PEL_EINVAL, w_ + ": a job's state does not fit the LDS budget");
        out->group[i] = group;
        out->placement[i] = place;
        out->learned[i] = learn ? 1 : 0;
        out->job_lds[i] = (int32_t)L.lds;
        out->spill_stride[i] = align_up(L.spill, 256);
        lds = std::max(lds, need);
        if (!serve) units += group == 64 ? (shots[i] + kEngineWaves - 1) / kEngineWaves : shots[i];
    }
    lds = align_up(lds, 16) + ctl;
    int mask = 0;
    for (int32_t i = 0; i < n; ++i)
        mask |= kind_bit(out->placement[i] | (out->group[i] == 64 ? kKindWave : 0) | (out->learned[i] ? kKindLearn : 0));
    const void *fn = serve ? reinterpret_cast<const void *>(pick_serve(mask))
                           : reinterpret_cast<const void *>(pick_engine(mask));
    hipFuncAttributes fa;
    PEL_HIP(hipFuncGetAttributes(&fa, fn));
    if (fa.sharedSizeBytes != 0) return fail(PEL_EDEVICE, w_ + ": the engine kernel has static LDS");
    if (lds > 65536) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
    }
    int per_cu = 0;
    PEL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kEngineBlock, (size_t)lds));
    if (per_cu <= 0) return fail(PEL_EDEVICE, w_ + ": the engine kernel cannot be resident");
    const int64_t resident = (int64_t)per_cu * n_cu;
    const int64_t grid = serve ? resident : std::max<int64_t>(1, std::min<int64_t>(resident, units));
    out->n_jobs = n;
    out->table_bytes = (int64_t)sizeof(JobDesc) * n;
    out->block = kEngineBlock;
    out->grid = (int32_t)grid;
    out->lds_bytes = (int32_t)lds;
    for (int32_t i = 0; i < n; ++i) {
        const bool learn = out->learned[i] != 0;
        const int64_t slots = grid * (out->group[i] == 64 ? kEngineWaves : 1);
        const int32_t B = serve ? (int32_t)slots : shots[i];
        out->workspace_bytes[i] = ws_layout(graphs[i], B, false, learn).total_fixed + out->spill_stride[i] * slots;
    }
    return PEL_OK;
}

// Whether a batch has per-shot inputs or records, which a serving job (whose
// rows are staging rows, not shots) cannot take.
bool p_has_per_shot(const pel_params *p, const pel_batch *b) {
    return b->prior_batched || b->gamma_per_row || b->init_gamma || b->init_posterior || b->last_posterior ||
           b->shot_clock || b->feat_out || b->hard_out || b->gamma_out || p->leg_step != 0;
}

// The job table's descriptors of n jobs under an engine plan (serve: the
// serving engine, whose batches are staging rows); *mask the job kinds.
int job_descs(const std::string &w_, const pel_graph *const *graphs, const pel_params *params,
              const pel_batch *batches, int32_t n, const pel_jobs_plan *plan, bool serve, std::vector<JobDesc> &desc,
              int *mask) {
    if (!graphs || !params || !batches || !plan) return fail(PEL_EINVAL, w_ + ": NULL argument");
    if (n < 1 || n > kMaxJobs || plan->n_jobs != n || plan->block != kEngineBlock || plan->grid <= 0 ||
        plan->table_bytes < (int64_t)sizeof(JobDesc) * n)
        return fail(PEL_EINVAL, w_ + ": plan does not match these jobs");
    desc.assign((size_t)n, JobDesc{});
    memset(desc.data(), 0, sizeof(JobDesc) * desc.size());
    int64_t lds_need = 0;
    *mask = 0;
    for (int32_t i = 0; i < n; ++i) {
        const pel_graph *g = graphs[i];
        const pel_batch *b = &batches[i];
        const bool learn = plan->learned[i] != 0;
        int rc = check_batch(w_, g, &params[i], b, learn, false);
        if (rc != PEL_OK) return rc;
        const int group = plan->group[i], place = plan->placement[i];
        bool known = false;
        for (int pl : kPlacements) known |= (pl == place);
        if (!known || (group != 64 && group != kEngineBlock) || (group == 64 && (learn || (place & 7) != 7)) ||
            (learn && (place & 7) != 7))
            return fail(PEL_EINVAL, w_ + ": plan does not match these jobs");
        const GroupLayout L = group_layout(g, place, group, learn);
        const WsLayout w = ws_layout(g, b->B, false, learn);
        const int64_t slots = (int64_t)plan->grid * (group == 64 ? kEngineWaves : 1);
        if (L.lds != plan->job_lds[i] || align_up(L.spill, 256) != plan->spill_stride[i] ||
            plan->workspace_bytes[i] < w.total_fixed + plan->spill_stride[i] * slots)
            return fail(PEL_EINVAL, w_ + ": plan does not match these jobs");
        if (serve && (b->B != slots || p_has_per_shot(&params[i], b)))
            return fail(PEL_EINVAL, w_ + ": a serving job's batch is grid x (16 or 1) staging rows with shared "
                                         "priors and gamma rows, no records and no leg steps");
        lds_need = std::max(lds_need, group == 64 ? kEngineWaves * align_up(L.lds, 16) : L.lds);
        fill_args(desc[i].a, g, &params[i], b, static_cast<uint8_t *>(b->workspace), w, learn, 0, group, L,
                  plan->spill_stride[i]);
        desc[i].a.shot_clock = nullptr;
        desc[i].kind = place | (group == 64 ? kKindWave : 0) | (learn ? kKindLearn : 0);
        desc[i].stride = (uint32_t)align_up(L.lds, 16);
        *mask |= kind_bit(desc[i].kind);
    }
    if (lds_need + (serve ? kServeCtlBytes : 16) > plan->lds_bytes)
        return fail(PEL_EINVAL, w_ + ": plan does not match these jobs");
    return PEL_OK;
}

// The job table, kSetupJobs descriptors per setup launch on stream s.
int write_job_table(const std::vector<JobDesc> &desc, JobDesc *tab, hipStream_t s) {
    const int32_t n = (int32_t)desc.size();
    for (int32_t i0 = 0; i0 < n; i0 += kSetupJobs) {
        SetupArgs sa;
        memset(&sa, 0, sizeof sa);
        sa.dst = tab + i0;
        sa.count = std::min<int32_t>(kSetupJobs, n - i0);
        for (int32_t k = 0; k < sa.count; ++k) sa.src[k] = desc[(size_t)(i0 + k)];
        hipLaunchKernelGGL(jobs_setup_kernel, dim3(1), dim3(256), 0, s, sa);
        PEL_HIP(hipGetLastError());
    }
    return PEL_OK;
}

int decode_jobs_impl(const pel_graph *const *graphs, const pel_params *params, const pel_batch *batches, int32_t n,
                     const pel_jobs_plan *plan, void *table, void *stream) {
    const std::string w_ = "pel_decode_jobs";
    if (!table) return fail(PEL_EINVAL, w_ + ": NULL argument");
    std::vector<JobDesc> desc;
    int mask = 0;
    int rc = job_descs(w_, graphs, params, batches, n, plan, false, desc, &mask);
    if (rc != PEL_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (int32_t i = 0; i < n; ++i) {
        if (batches[i].B == 0) continue;
        const WsLayout w = ws_layout(graphs[i], batches[i].B, false, plan->learned[i] != 0);
        rc = launch_prep(graphs[i], &batches[i], desc[i].a, static_cast<uint8_t *>(batches[i].workspace), w, s);
        if (rc != PEL_OK) return rc;
    }
    JobDesc *tab = static_cast<JobDesc *>(table);
    rc = write_job_table(desc, tab, s);
    if (rc != PEL_OK) return rc;
    EngineArgs e;
    memset(&e, 0, sizeof e);
    e.n_jobs = n;
    e.ctl_lds = (uint32_t)(plan->lds_bytes - 16);
    if (batches[0].ev_begin) PEL_HIP(hipEventRecord(static_cast<hipEvent_t>(batches[0].ev_begin), s));
    hipLaunchKernelGGL(pick_engine(mask), dim3(plan->grid), dim3(kEngineBlock), (size_t)plan->lds_bytes, s, e,
                       static_cast<const JobDesc *>(tab));
    PEL_HIP(hipGetLastError());
    if (batches[0].ev_end) PEL_HIP(hipEventRecord(static_cast<hipEvent_t>(batches[0].ev_end), s));
    return PEL_OK;
}

}  // namespace

// The serving queue: one block of host memory the device reads and writes
// (the published count and stop flag, the unit ring, the done flags) and the
// claim counter in device memory.
struct pel_serve_queue {
    int device = 0;
    int32_t capacity = 0;
    uint8_t *host = nullptr;     // host address of the block
    uint8_t *mapped = nullptr;   // its device address
    uint64_t *claim = nullptr;   // device memory: the claim word
    int64_t next = 0;            // units published
    std::mutex mu;
    static constexpr int64_t kUnitsOff = 64;
    int64_t *published() const { return reinterpret_cast<int64_t *>(host); }
    int32_t *stop() const { return reinterpret_cast<int32_t *>(host + 8); }
    pel_serve_unit *units() const { return reinterpret_cast<pel_serve_unit *>(host + kUnitsOff); }
    int64_t *done() const { return reinterpret_cast<int64_t *>(host + kUnitsOff + 64LL * capacity); }
    int64_t bytes() const { return kUnitsOff + 64LL * capacity + 8LL * capacity; }
};

namespace {

int serve_launch_impl(const pel_graph *const *graphs, const pel_params *params, const pel_batch *batches, int32_t n,
                      const pel_jobs_plan *plan, void *table, pel_serve_queue *q, int64_t idle_us, void *stream) {
    const std::string w_ = "pel_serve_launch";
    if (!table || !q) return fail(PEL_EINVAL, w_ + ": NULL argument");
    if (idle_us < 0 || idle_us > 60LL * 1000 * 1000) return fail(PEL_EINVAL, w_ + ": idle_us must lie in 0 .. 6e7");
    std::vector<JobDesc> desc;
    int mask = 0;
    int rc = job_descs(w_, graphs, params, batches, n, plan, true, desc, &mask);
    if (rc != PEL_OK) return rc;
    if (graphs[0]->device != q->device) return fail(PEL_EINVAL, w_ + ": the queue is on another device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (int32_t i = 0; i < n; ++i) {
        const WsLayout w = ws_layout(graphs[i], batches[i].B, false, plan->learned[i] != 0);
        rc = launch_prep(graphs[i], &batches[i], desc[i].a, static_cast<uint8_t *>(batches[i].workspace), w, s, false);
        if (rc != PEL_OK) return rc;
    }
    JobDesc *tab = static_cast<JobDesc *>(table);
    rc = write_job_table(desc, tab, s);
    if (rc != PEL_OK) return rc;
    hipLaunchKernelGGL(serve_open_kernel, dim3(1), dim3(64), 0, s, q->claim);
    PEL_HIP(hipGetLastError());
    ServeArgs sa;
    memset(&sa, 0, sizeof sa);
    sa.published = reinterpret_cast<const int64_t *>(q->mapped);
    sa.stop = reinterpret_cast<const int32_t *>(q->mapped + 8);
    sa.units = reinterpret_cast<const pel_serve_unit *>(q->mapped + pel_serve_queue::kUnitsOff);
    sa.done = reinterpret_cast<int64_t *>(q->mapped + pel_serve_queue::kUnitsOff + 64LL * q->capacity);
    sa.claim = q->claim;
    sa.idle_ticks = idle_us * 100;  // the wall clock runs at 100 MHz
    sa.capacity = q->capacity;
    sa.n_jobs = n;
    sa.ctl_lds = (uint32_t)(plan->lds_bytes - kServeCtlBytes);
    hipLaunchKernelGGL(pick_serve(mask), dim3(plan->grid), dim3(kEngineBlock), (size_t)plan->lds_bytes, s, sa,
                       static_cast<const JobDesc *>(tab));
    PEL_HIP(hipGetLastError());
    return PEL_OK;
}

}  // namespace

extern "C" {

int pel_decode(const pel_graph *g, const pel_params *p, const pel_batch *b, const pel_plan *plan, void *stream) {
    return decode_impl("pel_decode", g, p, b, plan, false, stream);
}

int pel_decode_noalloc(const pel_graph *g, const pel_params *p, const pel_batch *b, const pel_plan *plan,
                       void *stream) {
    return decode_impl("pel_decode_noalloc", g, p, b, plan, true, stream);
}

int pel_plan_jobs(const pel_graph *const *graphs, const int32_t *shots, const int32_t *flags, int32_t n,
                  pel_jobs_plan *out) {
    return plan_engine("pel_plan_jobs", graphs, shots, flags, n, out, false);
}

int pel_decode_jobs(const pel_graph *const *graphs, const pel_params *params, const pel_batch *batches, int32_t n,
                    const pel_jobs_plan *plan, void *table, void *stream) {
    return decode_jobs_impl(graphs, params, batches, n, plan, table, stream);
}

int pel_plan_serve(const pel_graph *const *graphs, const int32_t *flags, int32_t n, pel_jobs_plan *out) {
    return plan_engine("pel_plan_serve", graphs, nullptr, flags, n, out, true);
}

int pel_serve_launch(const pel_graph *const *graphs, const pel_params *params, const pel_batch *batches, int32_t n,
                     const pel_jobs_plan *plan, void *table, pel_serve_queue *q, int64_t idle_us, void *stream) {
    return serve_launch_impl(graphs, params, batches, n, plan, table, q, idle_us, stream);
}

int pel_serve_queue_create(int32_t capacity, pel_serve_queue **out) {
    if (!out) return fail(PEL_EINVAL, "pel_serve_queue_create: NULL argument");
    *out = nullptr;
    if (capacity < 1 || capacity > (1 << 24)) return fail(PEL_EINVAL, "pel_serve_queue_create: capacity must lie in 1 .. 2^24");
    pel_serve_queue *q = new pel_serve_queue();
    q->capacity = capacity;
    PEL_HIP(hipGetDevice(&q->device));
    void *h = nullptr;
    hipError_t e = hipHostMalloc(&h, (size_t)q->bytes(), hipHostMallocCoherent | hipHostMallocMapped);
    if (e != hipSuccess) {
        delete q;
        return fail(PEL_ENOMEM, std::string("pel_serve_queue_create: ") + hipGetErrorString(e));
    }
    q->host = static_cast<uint8_t *>(h);
    memset(q->host, 0, (size_t)q->bytes());
    void *d = nullptr;
    e = hipHostGetDevicePointer(&d, h, 0);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&q->claim), 8);
    if (e == hipSuccess) e = hipMemset(q->claim, 0, 8);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        if (q->claim) (void)hipFree(q->claim);
        (void)hipHostFree(h);
        delete q;
        return fail(PEL_EDEVICE, std::string("pel_serve_queue_create: ") + hipGetErrorString(e));
    }
    q->mapped = static_cast<uint8_t *>(d);
    *out = q;
    return PEL_OK;
}

int pel_serve_queue_destroy(pel_serve_queue *q) {
    if (!q) return PEL_OK;
    hipError_t e1 = hipFree(q->claim);
    hipError_t e2 = hipHostFree(q->host);
    delete q;
    if (e1 != hipSuccess || e2 != hipSuccess)
        return fail(PEL_EDEVICE, std::string("pel_serve_queue_destroy: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    return PEL_OK;
}

int pel_serve_queue_info(const pel_serve_queue *q, const int64_t **done, int32_t *capacity, int64_t *published) {
    if (!q) return fail(PEL_EINVAL, "pel_serve_queue_info: NULL queue");
    if (done) *done = q->done();
    if (capacity) *capacity = q->capacity;
    if (published) *published = __atomic_load_n(q->published(), __ATOMIC_ACQUIRE);
    return PEL_OK;
}

int pel_serve_publish(pel_serve_queue *q, const pel_serve_unit *units, int32_t n, int64_t *first) {
    if (!q || (!units && n > 0)) return fail(PEL_EINVAL, "pel_serve_publish: NULL argument");
    if (n < 0) return fail(PEL_EINVAL, "pel_serve_publish: negative count");
    for (int32_t i = 0; i < n; ++i) {
        const pel_serve_unit &u = units[i];
        if (u.job < 0 || u.count < 1 || u.count > PEL_SERVE_WAVE_SHOTS || !u.syndromes || !u.errors || !u.weights ||
            !u.iterations || !u.found_count)
            return fail(PEL_EINVAL, "pel_serve_publish: a unit has a bad job, count or NULL buffer");
    }
    std::lock_guard<std::mutex> lock(q->mu);
    const int64_t k0 = q->next;
    if (n > q->capacity) return fail(PEL_EBUSY, "pel_serve_publish: more units than the queue holds");
    for (int32_t i = 0; i < n; ++i) {  // entry k % capacity is free once unit k - capacity is done
        const int64_t k = k0 + i;
        if (k >= q->capacity && __atomic_load_n(q->done() + k % q->capacity, __ATOMIC_ACQUIRE) < k - q->capacity + 1)
            return fail(PEL_EBUSY, "pel_serve_publish: the queue is full");
    }
    for (int32_t i = 0; i < n; ++i) q->units()[(k0 + i) % q->capacity] = units[i];
    __atomic_store_n(q->published(), k0 + n, __ATOMIC_RELEASE);
    std::atomic_thread_fence(std::memory_order_seq_cst);  // drains the store buffers: the device sees it now
    q->next = k0 + n;
    if (first) *first = k0;
    return PEL_OK;
}

int pel_serve_set_stop(pel_serve_queue *q, int32_t stop) {
    if (!q) return fail(PEL_EINVAL, "pel_serve_set_stop: NULL queue");
    __atomic_store_n(q->stop(), stop ? 1 : 0, __ATOMIC_RELEASE);
    std::atomic_thread_fence(std::memory_order_seq_cst);
    return PEL_OK;
}

int64_t pel_serve_done_count(const pel_serve_queue *q, int64_t first, int64_t n) {
    if (!q || first < 0 || n < 0) return fail(PEL_EINVAL, "pel_serve_done_count: bad argument");
    int64_t done = 0;
    for (int64_t k = first; k < first + n; ++k)  // a later unit in the entry implies this one was done
        done += __atomic_load_n(q->done() + k % q->capacity, __ATOMIC_ACQUIRE) >= k + 1;
    return done;
}

}  // extern "C"

// ---- float16 / float64 decodes -------------------------------------------------------------------
namespace {

int64_t typed_elem_bytes(int32_t dtype) { return dtype == PEL_DTYPE_F16 ? 2 : dtype == PEL_DTYPE_F32 ? 4 : 8; }

int64_t typed_slot_bytes(const pel_graph *g, int32_t dtype) {
    return align_up((2LL * g->E + g->V) * typed_elem_bytes(dtype) + 3LL * g->V, 256);
}

int64_t typed_park_bytes(const pel_graph *g, int32_t dtype) {
    return align_up(kTypedParkHead + (int64_t)g->V * typed_elem_bytes(dtype) + 3LL * g->V, 16);
}

}  // namespace

extern "C" {

int pel_plan_typed(const pel_graph *g, int32_t B, int32_t dtype, int32_t flags, int32_t *grid,
                   int64_t *workspace_bytes) {
    if (!g || !grid || !workspace_bytes) return fail(PEL_EINVAL, "pel_plan_typed: NULL argument");
    if (dtype != PEL_DTYPE_F16 && dtype != PEL_DTYPE_F32 && dtype != PEL_DTYPE_F64)
        return fail(PEL_EINVAL, "pel_plan_typed: dtype must be PEL_DTYPE_F16, PEL_DTYPE_F32 or PEL_DTYPE_F64");
    if (B < 0) return fail(PEL_EINVAL, "pel_plan_typed: negative batch size");
    if (flags & ~PEL_PLAN_LEGWISE) return fail(PEL_EINVAL, "pel_plan_typed: flags other than PEL_PLAN_LEGWISE");
    int dev = 0, n_cu = 0;
    PEL_HIP(hipGetDevice(&dev));
    PEL_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    *grid = std::max(1, std::min<int32_t>(std::max(B, 1), 2 * n_cu));
    *workspace_bytes = 256 + (int64_t)*grid * typed_slot_bytes(g, dtype) +
                       ((flags & PEL_PLAN_LEGWISE) ? (int64_t)B * typed_park_bytes(g, dtype) : 0);
    return PEL_OK;
}

int pel_decode_typed(const pel_graph *g, const pel_params *p, const pel_batch *b, int32_t dtype, int32_t grid,
                     void *stream) {
    const std::string w_ = "pel_decode_typed";
    if (!g || !p || !b) return fail(PEL_EINVAL, w_ + ": NULL argument");
    int32_t want_grid = 0;
    int64_t need = 0;
    int rc = pel_plan_typed(g, b->B, dtype, p->leg_step > 0 ? PEL_PLAN_LEGWISE : 0, &want_grid, &need);
    if (rc != PEL_OK) return rc;
    if (b->B == 0) return PEL_OK;
    if (grid < 1 || grid > want_grid) return fail(PEL_EINVAL, w_ + ": grid must lie in 1 .. the planned grid");
    if (!b->syndromes || !b->prior_llr || !b->wj || !b->errors || !b->weights || !b->iterations || !b->found_count ||
        !b->workspace || (p->num_sets > 0 && !p->skip_relay_legs && !b->gamma_legs) || (p->use_alpha && !p->alpha_t))
        return fail(PEL_EINVAL, w_ + ": NULL buffer");
    if (p->pre_iter < 0 || p->num_sets < 0 || p->set_max_iter < 0 || p->stop_nconv < 1)
        return fail(PEL_EINVAL, w_ + ": bad schedule parameters");
    if (b->mlp || b->feat_out || b->hard_out || b->gamma_out || b->learned_legs)
        return fail(PEL_EINVAL, w_ + ": learned strengths run in float32 only (pel_decode)");
    if (p->leg_step < 0 || p->leg_step > p->num_sets + 1)
        return fail(PEL_EINVAL, w_ + ": leg_step must lie in 0 .. num_sets + 1");
    if (p->leg_step > 0 && p->skip_relay_legs)
        return fail(PEL_EINVAL, w_ + ": leg-stepped decodes run the relay legs");
    if (b->gamma_by_position) return fail(PEL_EINVAL, w_ + ": gamma rows must be in the caller's variable order");
    TypedArgs a{};
    a.C = g->C;
    a.V = g->V;
    a.E = g->E;
    a.B = b->B;
    a.row_ptr = g->row_ptr;
    a.row_var = g->row_var;
    a.var_ptr = g->var_ptr;
    a.var_edges = g->var_edges;
    a.pre_iter = p->pre_iter;
    a.num_sets = p->num_sets;
    a.set_max_iter = p->set_max_iter;
    a.stop_nconv = p->stop_nconv;
    a.plain = p->plain;
    a.use_alpha = p->use_alpha;
    a.use_beta = p->use_beta;
    a.last_leg = p->last_leg;
    a.skip_relay_legs = p->skip_relay_legs;
    a.gamma0 = p->gamma0;
    a.beta = p->beta;
    a.alpha_t = p->alpha_t;
    a.syn = b->syndromes;
    a.prior = b->prior_llr;
    a.wj = b->wj;
    a.prior_batched = b->prior_batched;
    a.gamma_legs = b->gamma_legs;
    a.gamma_per_row = b->gamma_per_row;
    a.init_gamma = b->init_gamma;
    a.init_post = b->init_posterior;
    a.errors = b->errors;
    a.weights = b->weights;
    a.iters = b->iterations;
    a.found = b->found_count;
    a.last_post = b->last_posterior;
    uint8_t *ws = static_cast<uint8_t *>(b->workspace);
    a.ctrl = reinterpret_cast<int32_t *>(ws);
    a.slots = ws + 256;
    a.slot_stride = typed_slot_bytes(g, dtype);
    a.leg_step = p->leg_step;
    a.park = p->leg_step > 0 ? a.slots + (int64_t)grid * a.slot_stride : nullptr;
    a.park_stride = typed_park_bytes(g, dtype);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(typed_clear_kernel, dim3(1), dim3(64), 0, s, a.ctrl);
    PEL_HIP(hipGetLastError());
    if (b->ev_begin) PEL_HIP(hipEventRecord(static_cast<hipEvent_t>(b->ev_begin), s));
    if (dtype == PEL_DTYPE_F16)
        hipLaunchKernelGGL(typed_decode_kernel<_Float16>, dim3(grid), dim3(kTypedBlock), 0, s, a);
    else if (dtype == PEL_DTYPE_F32)
        hipLaunchKernelGGL(typed_decode_kernel<float>, dim3(grid), dim3(kTypedBlock), 0, s, a);
    else
        hipLaunchKernelGGL(typed_decode_kernel<double>, dim3(grid), dim3(kTypedBlock), 0, s, a);
    PEL_HIP(hipGetLastError());
    if (b->ev_end) PEL_HIP(hipEventRecord(static_cast<hipEvent_t>(b->ev_end), s));
    return PEL_OK;
}

}  // extern "C"
