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
    // Per step.
    u32 *solver_of_body; // cap_bodies
    u32 *awake_bodies;   // cap_bodies
    NphSolverBody *sbodies; // cap_bodies + 1
    NphSolverParams *params; // cap_bodies
    u32 *list;  // awake contacts / graph contacts (cap_list)
    u32 cap_list;
    u32 *codes; // change codes, per list entry
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
};

// Buffer growth: a new allocation of new_bytes, the old contents copied
// (keep) in stream order, the old one freed.
NphResult devi_grow_buffer(void **p, u64 old_bytes, u64 new_bytes, hipStream_t s, int keep);
u32 devi_grow_cap(u32 cap, u32 need);
// A pinned host buffer of at least `bytes`.
NphResult devi_reserve_pinned(NphDevice *d, void **p, u64 *cap, u64 bytes);
// A world's device on `stream` (a stream of its own when NULL).
NphResult devi_create(NphDevice **out, hipStream_t stream);
// The host's staged record writes of a world, to the device (stream-ordered).
NphResult devi_flush_staged(NphDevice *d);

static inline u32 devi_blocks(u32 n, u32 block) { return (n + block - 1U) / block; }

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

#endif // NPH_DEVICE_IMPL_H
