// SPDX-License-Identifier: Apache-2.0
// The boundary between the host bookkeeping (C17) and the device that runs
// the step's data-parallel stages.  One implementation ships: HIP on
// gfx950 (device_hip.hip).  All records live in device memory; the host
// sends what the serial bookkeeping decided (new contacts, the constraint
// graph's colour lists) and reads back what it needs to decide the next
// serial step (changed contacts, grown proxies).
#ifndef NPH_DEVICE_H
#define NPH_DEVICE_H

#include "nph_step.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct NphDevice NphDevice;

#define NPH_GRAPH_COLORS 24U // world_internal.h:359
#define NPH_GRAPH_DYNAMIC_COLORS 20U
#define NPH_GRAPH_OVERFLOW NPH_GRAPH_COLORS

// A contact record the host made this step (contact.c:111-127 on the device).
typedef struct NphContactInit {
    u32 record;
    u32 shape_a;
    u32 shape_b;
    u32 body_a;
    u32 body_b;
    u32 child; // a mesh pair's next record, or NPH_NULL32
} NphContactInit;

// A shape whose fat box grew in finalize, in the awake list's order and each
// body's shape-list order (solver.c:868-885).
typedef struct NphGrown {
    u32 shape;
    u32 pad;
    NphAABB fat_aabb;
} NphGrown;

typedef struct NphSolveInput {
    const u32 *awake_bodies; // body slots, solver order
    u32 awake_count;
    u32 body_slots; // slots in use (the solver map's size)
    // The graph's contacts: colours 0..23, then the overflow, each list in
    // its own order (the reference's colour arrays).
    const u32 *graph_contacts;
    u32 color_count[NPH_GRAPH_COLORS + 1U];
    // The graph's joints (slots) in the same colour order.
    const u32 *graph_joints;
    u32 joint_color_count[NPH_GRAPH_COLORS + 1U];
    NphStepParams params;
} NphSolveInput;

// ---------------------------------------------------------------------------
// Pair finding (contact.c:703-854): the host keeps the three AABB trees as
// the reference builds them; the device holds a copy of each, brought up to
// date with the nodes the host wrote since the last search, and runs every
// moved proxy's query over them, count -> scan -> fill, so the new pairs
// come out in the move buffer's order and each proxy's in its trees' order
// (pairs_hip.hip).
// ---------------------------------------------------------------------------
// The trees, in the order a dynamic proxy's query visits them.
#define NPH_TREE_STATIC 0U
#define NPH_TREE_KINEMATIC 1U
#define NPH_TREE_DYNAMIC 2U

typedef struct NphTreeNode {
    NphAABB aabb;
    u64 category_bits;
    u32 child1; // a leaf's user data, low word
    u32 child2; // a leaf's user data, high word
    u32 parent;
    u16 height; // 0: a leaf
    u16 flags;
} NphTreeNode; // 48 B (aabb_tree.h:52-67)

// What the pair filters read of a shape (contact.c:703-746, shape.c:21-27).
typedef struct NphPairShape {
    u64 category;
    u64 mask;
    i32 group;
    u32 body;
    u32 flags; // NPH_PAIR_*
    u32 pad;
} NphPairShape; // 32 B

enum {
    NPH_PAIR_UNTOUCHABLE = 1U, // a sensor or a ghost
    NPH_PAIR_DYNAMIC = 2U,     // its body is dynamic
    NPH_PAIR_JOINTED = 4U,     // its body has joints
};

// The filter set's keys: shape pairs that have a contact, and body pairs
// that a joint keeps from colliding.
#define NPH_PAIR_BODY_KEY (1ULL << 63U)
NPH_HD u64 nph_pair_key(u32 a, u32 b)
{
    return (a < b ? ((u64)a << 32U) | b : ((u64)b << 32U) | a) + 1ULL; // contact.c:12-15
}

// One world's part of a search: what the host changed since the world's
// last search, and its move buffer.
typedef struct NphPairInput {
    NphDevice *dev;
    const NphTreeNode *tree_nodes[3];
    const u32 *tree_dirty[3]; // nodes written since the last search
    u32 tree_dirty_count[3];
    u32 tree_root[3];
    u32 tree_capacity[3];
    u32 tree_height[3];
    const u32 *shape_slots; // shapes whose filter data changed, and that data
    const NphPairShape *shape_info;
    u32 shape_count;
    u32 shape_capacity; // shape slots the device must hold
    const u64 *set_keys; // filter-set keys to bring up to date, and whether each is in
    const u8 *set_in;
    u32 set_count;
    u32 set_capacity; // a power of two
    u32 set_reset;    // the set starts over (empty, then set_keys)
    const u32 *moved_keys; // proxy << 2 | tree, the move buffer's order
    const u32 *moved_shapes;
    u32 moved_count;
} NphPairInput;

typedef struct NphPairSearch NphPairSearch;
NphResult nph_pairs_create(NphPairSearch **out);
// The last search's host packing time (ms), its upload's device time (ms)
// and the upload's bytes.
void nph_pairs_last_detail(const NphPairSearch *p, f64 out[3]);
void nph_pairs_destroy(NphPairSearch *p);
// The new pairs of n worlds in one search: world k's are pairs
// [(*starts)[k], (*starts)[k + 1]) of *pairs, each (shape a, shape b), in
// order; valid until the next call.
NphResult nph_pairs_find(NphPairSearch *p, const NphPairInput *in, u32 n, const u32 **pairs,
                         const u32 **starts);
// Device milliseconds of the last search (its stream's events).
f64 nph_pairs_last_ms(const NphPairSearch *p);

NphResult nph_dev_create(NphDevice **out);
void nph_dev_destroy(NphDevice *d);
const char *nph_dev_name(void);
// Grows the record arrays (contents kept) to at least these many slots.
NphResult nph_dev_reserve(NphDevice *d, u32 bodies, u32 shapes, u32 contacts);
// A device copy of a hull or mesh block of `bytes`, made once per block and
// kept for the device's life; *out is its device address.
NphResult nph_dev_block(NphDevice *d, const void *block, u32 bytes, const void **out);
// A mesh shape exists: the narrow phase runs its mesh-pair kernel too.
void nph_dev_note_mesh_shape(NphDevice *d);
// A shape of a hull beyond 32 vertices was made: the large-hull narrow
// phase runs from now on.
void nph_dev_note_big_hull_shape(NphDevice *d);
NphResult nph_dev_write_body(NphDevice *d, u32 slot, const NphBodyRec *b);
// Joint records by slot: the configuration (when conf is not NULL) and, with
// reset, accumulated impulses from zero; staged as the body writes are.
NphResult nph_dev_reserve_joints(NphDevice *d, u32 joints);
NphResult nph_dev_write_joint(NphDevice *d, u32 slot, const NphJointConf *conf, int reset);
NphResult nph_dev_read_joint(NphDevice *d, u32 slot, NphJointState *out);
NphResult nph_dev_write_shape(NphDevice *d, u32 slot, const NphShapeRec *s);
NphResult nph_dev_read_bodies(NphDevice *d, NphBodyRec *out, u32 count);
NphResult nph_dev_read_shapes(NphDevice *d, NphShapeRec *out, u32 count);
NphResult nph_dev_read_contacts(NphDevice *d, NphContactRec *out, u32 count);
// Queued behind earlier work; the list may be reused on return.
NphResult nph_dev_init_contacts(NphDevice *d, const NphContactInit *list, u32 count);
// The narrow phase in two parts, so the host can look for new pairs while
// the device works: collide_begin queues list entries [0, count) (the awake
// contacts when the step starts) and returns at once; collide_end queues
// entries [first, count) where first is begin's count (the contacts made
// since, after init_contacts), waits for both, and points *codes at the change
// code of every entry (0: none), valid until the next call.
NphResult nph_dev_collide_begin(NphDevice *d, const u32 *awake, u32 count, f32 step_dt);
// The first `count` contact records saved before a step changes them, and
// put back if the step is abandoned before its changes reach the host.
NphResult nph_dev_contacts_checkpoint(NphDevice *d, u32 count);
NphResult nph_dev_contacts_restore(NphDevice *d);
NphResult nph_dev_collide_end(NphDevice *d, const u32 *awake, u32 count, f32 step_dt,
                              const u32 **codes);
// A fast body (continuous collision) after finalize: its record (pose at
// the step's end, boxes not yet refreshed unless a bullet), its solver rows.
typedef struct NphFastBody {
    NphBodyRec body;
    NphSolverBody sb;
    NphSolverParams params;
    u32 index; // awake index
    u32 slot;  // body slot
    u32 pad[2];
} NphFastBody;

// The solver stages and finalize.  With params.continuous, the fast bodies
// come back through *fast (valid until the next call) for the host to sweep
// and rewrite (nph_dev_write_body/shape) before nph_dev_solve_finish.
NphResult nph_dev_solve(NphDevice *d, const NphSolveInput *in, const NphFastBody **fast,
                        u32 *fast_count);
// The grown proxies into *grown (the device's buffer, valid until the next
// call), *grown_count of them, after the host's writes reach the device.
NphResult nph_dev_solve_finish(NphDevice *d, const NphGrown **grown, u32 *grown_count);
// Seconds of device time the last solve took (event-timed), and the last
// collide.
f64 nph_dev_last_solve_ms(const NphDevice *d);
f64 nph_dev_last_collide_ms(const NphDevice *d);
// The constraint passes alone (warm start, solve, relax over every substep,
// then restitution): the stream's events around them.
f64 nph_dev_last_passes_ms(const NphDevice *d);
// Per-launch event timing of the colour-pass kernel, off by default (it adds
// two event records a launch): enabling clears the totals; out[0] = device
// milliseconds summed over the timed launches, out[1] = their count.
void nph_dev_time_kernels(NphDevice *d, int enable);
void nph_dev_kernel_times(const NphDevice *d, f64 out[2]);

// The learned warm start: the network (NULL: off), residual measurement,
// its statistics of the last solve, and the contact log.
NphResult nph_dev_set_warm_start(NphDevice *d, const NphAmdWarmStart *ws);
void nph_dev_measure_residual(NphDevice *d, int enable);
NphResult nph_dev_warm_stats(NphDevice *d, NphAmdWarmStats *out);
NphResult nph_dev_log_contacts(NphDevice *d, u32 capacity);
NphResult nph_dev_contact_log(NphDevice *d, u32 *rows, const f32 **features,
                              const f32 **targets);

// A learned warm start, residual measurement or contact log is on.
int nph_dev_learned_active(const NphDevice *d);
// The contact log is on.
int nph_dev_logging(const NphDevice *d);

// ---------------------------------------------------------------------------
// Many worlds stepped together: each world keeps its own records (its
// NphDevice), their work goes on one stream, and every stage of the step is
// one launch over all the worlds' items (device_batch.hip).
// ---------------------------------------------------------------------------
typedef struct NphDeviceBatch NphDeviceBatch;

// Launch choices of the batched step (the autotuner picks them).
typedef struct NphBatchTuning {
    u32 pass_block;  // threads a block of the colour passes: 64, 128 or 256
    u32 body_block;  // threads a block of the per-body and per-lane stages
    u32 use_graph;   // the substep loop launched as a captured HIP graph
    u32 pad;
} NphBatchTuning;

NphResult nph_devb_create(NphDeviceBatch **out);
void nph_devb_destroy(NphDeviceBatch *b);
// A world's device whose work goes on the batch's stream.
NphResult nph_devb_world(NphDeviceBatch *b, NphDevice **out);
// A world's device (made by any batch of the same owner) moved to b's
// stream; every earlier operation of the world must have completed.
void nph_devb_adopt(NphDeviceBatch *b, NphDevice *d);
void nph_devb_set_tuning(NphDeviceBatch *b, const NphBatchTuning *t);
NphBatchTuning nph_devb_tuning(const NphDeviceBatch *b);

// One world's part of a narrow-phase launch: the records made since the
// last launch (inits), then entries [first, count) of its awake contacts.
typedef struct NphCollidePart {
    NphDevice *dev;
    const NphContactInit *inits;
    u32 init_count;
    const u32 *list;
    u32 first;
    u32 count;
} NphCollidePart;

// Half 0 queues the contacts alive when the step starts and returns at once
// (the lists are copied before it returns); half 1 queues the rest, waits,
// and points codes[0] / codes[1] at the change codes of each half, the
// parts' entries one after another in part order (valid until the next
// call).
NphResult nph_devb_collide(NphDeviceBatch *b, const NphCollidePart *parts, u32 n, f32 step_dt,
                           u32 half, const u32 *codes[2]);
// The solver stages and finalize for every world (all worlds share one
// NphStepParams: in[0]'s); the grown proxies of world k are
// (*grown)[(*starts)[k] .. (*starts)[k + 1]), in that world's awake order.
// Continuous collision and the contact log are not batched; worlds with a
// learned warm start must share one network.
NphResult nph_devb_solve(NphDeviceBatch *b, NphDevice *const *devs, const NphSolveInput *in,
                         u32 n, const NphGrown **grown, const u32 **starts);
// b's later work waits for the first narrow-phase half `first` queued
// last (so that half runs with the whole device before b's).
NphResult nph_devb_follow(NphDeviceBatch *b, const NphDeviceBatch *first);
// The same in two calls: begin queues the solve and returns (the inputs may
// change at once); end waits for it and returns the grown proxies.
NphResult nph_devb_solve_begin(NphDeviceBatch *b, NphDevice *const *devs, const NphSolveInput *in,
                               u32 n);
NphResult nph_devb_solve_end(NphDeviceBatch *b, const NphGrown **grown, const u32 **starts);
// Device milliseconds of the last batched solve: out[0] the whole solve,
// out[1] the constraint passes, out[2] the narrow phase (both halves); and
// the host's milliseconds blocked on the device: out[3] for the narrow
// phase's change codes, out[4] for the solve's grown proxies.
void nph_devb_times(const NphDeviceBatch *b, f64 out[5]);
// The learned warm start's statistics summed over the measured worlds and
// steps since the last reset: out = residual sums (plain, learned), contact
// counts (plain, learned), largest residuals (plain, learned), largest
// penetration, gated contacts; *steps = steps measured.
NphResult nph_devb_warm_stats(NphDeviceBatch *b, f64 out[8], u32 *steps, int reset);
// Per-launch timing of the batched colour passes (not in graph mode):
// out[0] = milliseconds summed, out[1] = launches.
void nph_devb_time_kernels(NphDeviceBatch *b, int enable);
// out: colour-pass milliseconds and launches, convex narrow-phase kernel
// milliseconds, launches and pairs, since timing was enabled.
void nph_devb_kernel_times(const NphDeviceBatch *b, f64 out[5]);

#ifdef __cplusplus
}
#endif

#endif // NPH_DEVICE_H
