// SPDX-License-Identifier: Apache-2.0
// The device's private state and the helpers that the one-world step
// (device_hip.hip) and the batched step of many worlds (device_batch.hip)
// share: the record arrays of a world, its staging of host writes, buffer
// growth, and the per-thread bodies of the constraint passes.
#ifndef NPH_DEVICE_IMPL_H
#define NPH_DEVICE_IMPL_H

#include <hip/hip_runtime.h>

#include "nph_device.h"
#include "nph_joint.h"
#include "nph_learned.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NPH_BLOCK 256U
#define NPH_SCAN_BLOCK 1024U
// The narrow phase keeps its clip buffers in private memory: one wave a block.
#define NPH_COLLIDE_BLOCK 64U

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        const hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                                    \
            fprintf(stderr, "nepheline_amd: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                           \
            return e_ == hipErrorOutOfMemory ? NPH_RESULT_OUT_OF_MEMORY : NPH_RESULT_DEVICE_LOST;  \
        }                                                                                          \
    } while (0)

// A hull or mesh block and its device copy.
typedef struct HullEntry {
    const void *host;
    const void *device;
} HullEntry;

// Record writes from the host (body and shape creation, velocity changes)
// wait here and reach HBM together, in one copy and one scatter, before the
// next device read or stage.  A slot written twice keeps its last record.
typedef struct Staged {
    u32 rec_words;  // record size in 4-byte words
    u32 *recs;      // count * rec_words
    u32 *slots;     // count
    u32 count, cap;
    u32 *where;     // per slot: staged index + 1, or 0
    u32 cap_where;
    u32 *dev;       // device copy: slots then records
    u64 dev_words;
} Staged;

struct NphDevice {
    hipStream_t stream;
    int shared_stream; // a batch's stream: the batch destroys it
    hipEvent_t ev[4];
    f64 solve_ms;
    f64 collide_ms;
    f64 passes_ms;
    // Records.
    NphBodyRec *bodies;
    NphShapeRec *shapes;
    NphContactRec *contacts;
    u32 cap_bodies, cap_shapes, cap_contacts;
    // The contact records as the step's narrow phase found them, put back
    // when the step is abandoned before its changes are applied.
    NphContactRec *contacts_saved;
    u32 cap_contacts_saved, contacts_saved_count;
    // Per step.
    u32 *solver_of_body; // cap_bodies
    u32 *awake_bodies;   // cap_bodies
    NphSolverBody *sbodies; // cap_bodies + 1
    NphSolverParams *params; // cap_bodies
    u32 *list;  // awake contacts / graph contacts (cap_list)
    u32 cap_list;
    u32 *codes; // change codes, per list entry
    // The narrow phase's team lists (32-vertex and large-hull pairs) and
    // their counts.
    struct NphSortedEntry *team_small;
    struct NphSortedEntry *team_big;
    struct NphSortedEntry *team_mesh;
    u32 cap_team;
    u32 *team_counts;
    NphContactInit *inits;
    u32 cap_inits;
    NphGroup *groups;
    u32 *lane_contact, *lane_body_a, *lane_body_b, *lane_flags;
    f32 *lanes;
    u32 cap_groups;
    f32 *points;
    u32 cap_point_rows;
    NphGrown *grown; // cap_shapes + 1 (count first)
    u32 *grown_blocks; // per NPH_BLOCK awake bodies: grown shapes (cap_bodies / NPH_BLOCK + 1)
    // Host staging (pinned).
    u32 *h_u32;
    u32 cap_h_u32;
    u32 *h_list;  // the awake list being collided
    u64 cap_h_list;
    u32 *h_codes; // change codes read back
    u64 cap_h_codes;
    NphContactInit *h_inits;
    u64 cap_h_inits;
    u32 collide_first; // entries [0, collide_first) queued by collide_begin
    // Per-launch timing of the colour passes (off unless asked for).
    int time_kernels;
    hipEvent_t *kev;
    u32 kev_pairs;
    f64 kernel_ms;
    u64 kernel_launches;
    hipEvent_t cev[4];
    NphGrown *h_grown;
    u32 cap_h_grown;
    HullEntry *hulls;
    u32 hull_count, hull_cap;
    Staged staged_bodies;
    Staged staged_shapes;
    // Joints by slot: host-written configuration, device-written state.
    NphJointConf *joint_conf;
    NphJointState *joint_state;
    u32 cap_joints;
    Staged staged_jconf;
    Staged staged_jstate;
    u32 *step_joints; // the graph's joints for a step
    u32 cap_step_joints;
    // Between nph_dev_solve and nph_dev_solve_finish.
    NphStepData pending_sd;
    u32 pending_timed;
    // Continuous collision: fast bodies (count first) and their packed rows.
    u32 *fast_list;
    NphFastBody *fast_pack;
    u32 cap_fast_pack;
    u32 *h_fast_count;
    u64 cap_h_fast_count;
    NphFastBody *h_fast_pack;
    u64 cap_h_fast_pack;
    // The learned warm start (off unless a network is set), residual
    // measurement and the contact log.
    int learn_on;
    int measure;
    u32 learn_mode;
    int gate_on;
    f32 shift[16], scale[16], out_scale[4];
    f32 *mlp_params;     // NPH_AMD_MLP_PARAMS
    u16 *lx;             // lanes * 4 network rows of 16 BF16
    f32 *lraw;           // lanes * 4 rows of 16 raw features
    f32 *ly;             // lanes * 4 rows of 4 outputs
    f32 *lres;           // per lane residual
    u32 cap_learn_lanes;
    u32 *gate;           // per contact record: 1 = plain warm start for good
    u32 *wstats;         // NPH_WSTATS words
    u32 *h_wstats;
    f32 *log_x, *log_t;  // log rows: 16 features, 4 targets
    u32 *log_count;
    u32 log_cap;
    u32 mesh_shapes;     // mesh shapes made: the mesh pairs' kernel runs
    u32 big_hull_shapes; // shapes of hulls beyond 32 vertices: the large-hull kernels run
    u64 learn_hash;      // FNV-1a of the network's parameters and settings
    // Pair finding (pairs_hip.hip): the trees' copies, the shapes' filter
    // data and moved stamps, the filter set.
    NphTreeNode *pf_nodes[3];
    u32 pf_node_cap[3];
    NphPairShape *pf_shapes;
    u32 *pf_stamp;
    u32 pf_shape_cap;
    u64 *pf_set;
    u32 pf_set_cap;
    u32 pf_search; // searches so far: the moved stamps' value
};

// Buffer growth: a new allocation of new_bytes, the old contents copied
// (keep) in stream order, the old one freed.
NphResult devi_grow_buffer(void **p, u64 old_bytes, u64 new_bytes, hipStream_t s, int keep);
u32 devi_grow_cap(u32 cap, u32 need);
// A pinned host buffer of at least `bytes`.
NphResult devi_reserve_pinned(NphDevice *d, void **p, u64 *cap, u64 bytes);
// A world's statistics words of the learned warm start, allocated once.
NphResult devi_reserve_stats(NphDevice *d);
// A world's device on `stream` (a stream of its own when NULL).
NphResult devi_create(NphDevice **out, hipStream_t stream);
// The host's staged record writes of a world, to the device (stream-ordered).
NphResult devi_flush_staged(NphDevice *d);

static inline u32 devi_blocks(u32 n, u32 block) { return (n + block - 1U) / block; }

// A world's narrow-phase records, for the batched narrow phase.
typedef struct NphCollideWorld {
    NphContactRec *contacts;
    const NphShapeRec *shapes;
    const NphBodyRec *bodies;
    u32 *gate;
} NphCollideWorld;

// One contact of the batched narrow phase's list, sorted by pair type.
typedef struct NphSortedEntry {
    u32 world;
    u32 record;
    u32 index; // the entry's place in the half's list (its change code's)
    u32 pad;
} NphSortedEntry;

// One pair of the standalone collision queries.
typedef struct NphQueryPair {
    NphGeom a;
    NphGeom b;
    NphPhysicsPose pose_a;
    NphPhysicsPose pose_b;
    f32 speculative;
    u32 big; // a hull beyond 32 vertices: the large-hull kernel takes it
    NphCache cache;
} NphQueryPair;


enum { PASS_WARM = 0, PASS_SOLVE = 1, PASS_RELAX = 2, PASS_RESTITUTION = 3 };

// One lane of a group in a constraint pass.
__device__ static inline void run_lane(const NphStepData *d, const NphStepParams *sp, u32 pass,
                                       u32 g, u32 l)
{
    switch (pass) {
    case PASS_WARM:
        lane_warm_start(d, g, l);
        break;
    case PASS_SOLVE:
        lane_solve(d, g, l, sp->inv_h, sp->push_max, 0.0F, 1);
        break;
    case PASS_RELAX:
        lane_solve(d, g, l, sp->inv_h, sp->push_max, 0.0F, 0);
        break;
    default:
        lane_restitution(d, g, l);
        break;
    }
}

// A joint's pass: warm start, or the soft (solve) or rigid (relax) rows.
// The joint's state is held in registers for the pass.
__device__ static inline void run_joint(const NphStepData *d, const NphStepParams *sp, u32 pass,
                                        u32 slot)
{
    const NphJointConf cf = d->joint_conf[slot];
    NphJointState st = d->joint_state[slot];
    if (pass == PASS_WARM) {
        joint_warm_start(d->sbodies, &cf, &st);
    } else {
        NphJointStep step;
        step.h = sp->h;
        step.inv_h = sp->inv_h;
        step.soft.bias = sp->joint_bias;
        step.soft.mass = sp->joint_mass;
        step.soft.impulse = sp->joint_impulse;
        joint_solve(d->sbodies, &step, &cf, &st, pass == PASS_SOLVE);
    }
    NphJointState *out = &d->joint_state[slot];
    out->linear_impulse = st.linear_impulse;
    out->angular_impulse = st.angular_impulse;
    out->friction_impulse = st.friction_impulse;
    out->lower_impulse = st.lower_impulse;
    out->upper_impulse = st.upper_impulse;
    out->swing_impulse = st.swing_impulse;
    out->motor_impulse = st.motor_impulse;
    out->spring_impulse = st.spring_impulse;
}

// Thread t of one colour's launch: joints [jfirst, jfirst + jcount) take the
// first threads (rounded up to whole waves), groups [first, first + count)
// the rest, a lane a thread.
__device__ static inline void color_pass_thread(const NphStepData *d, const NphStepParams *sp,
                                                u32 pass, u32 first, u32 count, u32 jfirst,
                                                u32 jcount, u32 t)
{
    const u32 jpad = (jcount + 63U) & ~63U;
    if (t < jpad) {
        if (t < jcount) {
            run_joint(d, sp, pass, d->step_joints[jfirst + t]);
        }
        return;
    }
    const u32 u = t - jpad;
    if (u < count * NPH_LANES) {
        run_lane(d, sp, pass, first + u / NPH_LANES, u % NPH_LANES);
    }
}

// The overflow of one world: its joints, then its groups (one contact
// each), in order, on one thread (solver.c:771-779).
__device__ static inline void overflow_pass(const NphStepData *d, const NphStepParams *sp, u32 pass,
                                            u32 first, u32 count, u32 jfirst, u32 jcount)
{
    for (u32 j = jfirst; j < jfirst + jcount; ++j) {
        run_joint(d, sp, pass, d->step_joints[j]);
    }
    for (u32 g = first; g < first + count; ++g) {
        run_lane(d, sp, pass, g, 0U);
    }
}

// A contact record the host made, cleared and linked (contact.c:111-127).
__device__ static inline void init_contact_record(NphContactRec *contacts, u32 *gate,
                                                  const NphContactInit *in)
{
    NphContactRec *c = &contacts[in->record];
    gate[in->record] = 0U;
    memset(c, 0, sizeof *c);
    c->shape_a = in->shape_a;
    c->shape_b = in->shape_b;
    c->body_a = in->body_a;
    c->body_b = in->body_b;
    c->normal = v3(0.0F, 1.0F, 0.0F);
    c->child = in->child;
}

// The shapes of awake body i whose fat boxes grew in finalize.
__device__ static inline u32 grown_of_body(const NphStepData *d, u32 i)
{
    u32 mine = 0U;
    const NphBodyRec *b = &d->bodies[d->awake_bodies[i]];
    if (b->enlarged != 0U) {
        for (u32 s = b->shape_head; s != NPH_NULL32; s = d->shapes[s].next) {
            mine += (d->shapes[s].flags & NPH_SHAPE_ENLARGED) != 0U ? 1U : 0U;
        }
    }
    return mine;
}

// Body i's grown shapes written from out[pos], their flags cleared.
__device__ static inline void grown_write(const NphStepData *d, u32 i, NphGrown *out)
{
    NphBodyRec *b = &d->bodies[d->awake_bodies[i]];
    u32 k = 0U;
    for (u32 s = b->shape_head; s != NPH_NULL32; s = d->shapes[s].next) {
        NphShapeRec *sh = &d->shapes[s];
        if ((sh->flags & NPH_SHAPE_ENLARGED) != 0U) {
            NphGrown *g = &out[k];
            g->shape = s;
            g->pad = 0U;
            g->fat_aabb = sh->fat_aabb;
            sh->flags &= ~NPH_SHAPE_ENLARGED;
            k += 1U;
        }
    }
    b->enlarged = 0U;
}

// ---------------------------------------------------------------------------
// The learned warm start, a lane at a time (device_hip.hip launches it for one
// world, device_batch.hip for many).
// ---------------------------------------------------------------------------
// Statistics words: maxima as the bits of non-negative floats (ordered as
// u32), sums as floats, counts; the gate's threshold persists across steps.
enum {
    WS_MAX_PLAIN = 0,
    WS_MAX_LEARNED,
    WS_COUNT_PLAIN,
    WS_COUNT_LEARNED,
    WS_GATED,
    WS_MAX_PEN,
    WS_SUM_PLAIN,
    WS_SUM_LEARNED,
    WS_CLEARED, // the words above are cleared each step
    WS_THRESHOLD = WS_CLEARED,
    NPH_WSTATS = 16
};

typedef struct NphLearnArgs {
    u16 *x;          // network rows (NULL: no inference)
    f32 *raw;        // raw feature rows (NULL: no log)
    const f32 *y;    // network outputs
    f32 *res;        // per lane residual
    u32 *gate;       // per contact record
    u32 *stats;
    f32 *log_x;
    f32 *log_t;
    u32 *log_count;
    u32 log_cap;
    u32 mode;
    u32 gate_on;
    u32 group_count;
    f32 inv_h;
    f32 shift[16];
    f32 scale[16];
    f32 out_scale[4];
} NphLearnArgs;

__device__ static inline u16 bf16_bits(f32 f)
{
    u32 u = __float_as_uint(f); // round to nearest even (finite values)
    u += 0x7FFFU + ((u >> 16U) & 1U);
    return (u16)(u >> 16U);
}

__device__ static inline int lane_active(const NphStepData *d, const NphLearnArgs *a, u32 L,
                                         const NphContactRec **c)
{
    const u32 g = L / NPH_LANES;
    if (g >= a->group_count || L % NPH_LANES >= d->groups[g].lane_count) {
        return 0;
    }
    *c = &d->contacts[d->lane_contact[L]];
    return 1;
}

// Each lane's four network rows (zero where there is no point) and, when
// logging, its points' raw features and carried-nothing flags.  L < the
// world's groups * 8.
__device__ static inline void learned_features_at(const NphStepData *d, const NphLearnArgs *a,
                                                  u32 L)
{
    const NphContactRec *c = NULL;
    const u32 points = lane_active(d, a, L, &c) ? c->point_count : 0U;
    const u32 g = L / NPH_LANES;
    const u32 l = L % NPH_LANES;
    for (u32 k = 0U; k < NPH_LEARNED_ROWS_PER_LANE; ++k) {
        const u64 row = (u64)L * NPH_LEARNED_ROWS_PER_LANE + k;
        f32 f[NPH_F_COUNT];
        f32 fresh = 0.0F;
        f32 carried = 0.0F;
        if (k < points) {
            learned_features(d, g, l, k, f);
            const u64 R = (u64)(d->groups[g].first_point + k) * NPH_LANES + l;
            carried = PF(d, NPH_P_NIMP, R);
            fresh = (carried == 0.0F && PF(d, NPH_P_TI1, R) == 0.0F &&
                     PF(d, NPH_P_TI2, R) == 0.0F)
                        ? 1.0F
                        : 0.0F;
        } else {
            for (u32 i = 0U; i < NPH_F_COUNT; ++i) {
                f[i] = 0.0F;
            }
        }
        if (a->x != NULL) {
            u16 *xr = a->x + row * NPH_LEARNED_IN;
            for (u32 i = 0U; i < NPH_LEARNED_IN; ++i) {
                xr[i] = (i < NPH_F_COUNT && k < points)
                            ? bf16_bits((f[i] - a->shift[i]) * a->scale[i])
                            : (u16)0U;
            }
        }
        if (a->raw != NULL) {
            f32 *fr = a->raw + row * NPH_LEARNED_IN;
            for (u32 i = 0U; i < NPH_F_COUNT; ++i) {
                fr[i] = f[i];
            }
            fr[14] = carried;
            fr[15] = fresh;
        }
    }
}

// The network's impulses into the chosen points of a lane the gate leaves
// to it; such lanes are flagged.
__device__ static inline void learned_apply_at(const NphStepData *d, const NphLearnArgs *a, u32 L)
{
    const NphContactRec *c = NULL;
    if (!lane_active(d, a, L, &c)) {
        return;
    }
    const u32 record = d->lane_contact[L];
    if (a->gate_on != 0U && a->gate[record] != 0U) {
        return;
    }
    // With every point taken by the network, one contact record in
    // NPH_LEARNED_CONTROL stays plain: the sample the gate measures the
    // plain solver's residual on.
    if (a->mode == NPH_AMD_WARM_ALL_POINTS && record % NPH_LEARNED_CONTROL == 0U) {
        return;
    }
    const u32 g = L / NPH_LANES;
    const u32 l = L % NPH_LANES;
    const f32 friction = LF(d, NPH_L_FRICTION, L);
    u32 any = 0U;
    for (u32 k = 0U; k < c->point_count; ++k) {
        const u64 R = (u64)(d->groups[g].first_point + k) * NPH_LANES + l;
        const int fresh = PF(d, NPH_P_NIMP, R) == 0.0F && PF(d, NPH_P_TI1, R) == 0.0F &&
                          PF(d, NPH_P_TI2, R) == 0.0F;
        if (a->mode == NPH_AMD_WARM_ALL_POINTS || fresh) {
            const f32 *y = a->y + ((u64)L * NPH_LEARNED_ROWS_PER_LANE + k) * NPH_LEARNED_OUT;
            f32 n, t1, t2;
            learned_impulses(y, a->out_scale, PF(d, NPH_P_NMASS, R), friction, &n, &t1, &t2);
            PF(d, NPH_P_NIMP, R) = n;
            PF(d, NPH_P_TI1, R) = t1;
            PF(d, NPH_P_TI2, R) = t2;
            any = 1U;
        }
    }
    if (any != 0U) {
        d->lane_flags[L] |= NPH_LANE_LEARNED;
    }
}

// Every lane's residual and penetration after the step, into the
// statistics by how the lane started: reduced across the wave first, so
// each wave adds to the shared words once.
__device__ static inline f32 wave_max(f32 v)
{
    for (u32 o = 32U; o > 0U; o >>= 1U) {
        v = nph_maxf(v, __shfl_xor(v, (int)o, 64));
    }
    return v;
}

__device__ static inline f32 wave_sum(f32 v)
{
    for (u32 o = 32U; o > 0U; o >>= 1U) {
        v += __shfl_xor(v, (int)o, 64);
    }
    return v;
}

// Every lane of the wave calls this (the statistics are reduced across the
// wave, then each wave adds to the shared words once); `in_range` marks the
// lanes that hold a lane of the world.
__device__ static inline void learned_residual_at(const NphStepData *d, const NphLearnArgs *a,
                                                  u32 L, int in_range)
{
    const NphContactRec *c = NULL;
    f32 res = 0.0F;
    f32 pen = 0.0F;
    int active = 0;
    int learned = 0;
    if (in_range && lane_active(d, a, L, &c)) {
        learned_residual(d, a->inv_h, L / NPH_LANES, L % NPH_LANES, c, &res, &pen);
        a->res[L] = res;
        active = 1;
        learned = (d->lane_flags[L] & NPH_LANE_LEARNED) != 0U;
    }
    const f32 rp = active && !learned ? res : 0.0F;
    const f32 rl = active && learned ? res : 0.0F;
    const f32 max_p = wave_max(rp);
    const f32 max_l = wave_max(rl);
    const f32 sum_p = wave_sum(rp);
    const f32 sum_l = wave_sum(rl);
    const f32 max_pen = wave_max(active ? nph_maxf(pen, 0.0F) : 0.0F);
    const u64 bp = __ballot(active && !learned);
    const u64 bl = __ballot(active && learned);
    if ((threadIdx.x & 63U) == 0U && (bp | bl) != 0U) {
        atomicMax(&a->stats[WS_MAX_PLAIN], __float_as_uint(max_p));
        atomicMax(&a->stats[WS_MAX_LEARNED], __float_as_uint(max_l));
        atomicAdd(&a->stats[WS_COUNT_PLAIN], (u32)__popcll(bp));
        atomicAdd(&a->stats[WS_COUNT_LEARNED], (u32)__popcll(bl));
        atomicAdd((f32 *)&a->stats[WS_SUM_PLAIN], sum_p);
        atomicAdd((f32 *)&a->stats[WS_SUM_LEARNED], sum_l);
        atomicMax(&a->stats[WS_MAX_PEN], __float_as_uint(max_pen));
    }
}

// The gate's threshold: the step's largest plain residual, or the last
// one while no contact started plainly.
__device__ static inline void learned_threshold(u32 *stats)
{
    if (stats[WS_COUNT_PLAIN] > 0U) {
        stats[WS_THRESHOLD] = stats[WS_MAX_PLAIN];
    }
}

__device__ static inline void learned_gate_at(const NphStepData *d, const NphLearnArgs *a, u32 L)
{
    const NphContactRec *c = NULL;
    if (!lane_active(d, a, L, &c) || (d->lane_flags[L] & NPH_LANE_LEARNED) == 0U) {
        return;
    }
    if (a->res[L] > __uint_as_float(a->stats[WS_THRESHOLD])) {
        a->gate[d->lane_contact[L]] = 1U;
        atomicAdd(&a->stats[WS_GATED], 1U);
    }
}

#endif // NPH_DEVICE_IMPL_H
