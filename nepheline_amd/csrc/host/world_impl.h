// SPDX-License-Identifier: Apache-2.0
// The host side of the world: the serial bookkeeping the reference keeps on
// the CPU (engine/physics/src/world_internal.h), restated around a device
// that holds every numeric record.
#ifndef NPH_WORLD_IMPL_H
#define NPH_WORLD_IMPL_H

#include "../nph_device.h"
#include "../nph_joint.h"

#include <stdlib.h>
#include <string.h>

// ---------------------------------------------------------------------------
// Growable arrays over the world's allocator
// ---------------------------------------------------------------------------
typedef struct NphVecBuf {
    void *data;
    u32 len;
    u32 cap;
} NphVecBuf;

// The allocator sees the caller's file and line.
NphResult vbuf_reserve_at(NphAllocator *a, NphVecBuf *v, u32 elem, u32 cap, NphCallsite site);
NphResult vbuf_push_at(NphAllocator *a, NphVecBuf *v, u32 elem, const void *item, NphCallsite site);
#define NPH_SITE ((NphCallsite){__FILE__, (u32)__LINE__})
#define vbuf_reserve(a, v, elem, cap) vbuf_reserve_at((a), (v), (elem), (cap), NPH_SITE)
#define vbuf_push(a, v, elem, item) vbuf_push_at((a), (v), (elem), (item), NPH_SITE)
void vbuf_free(NphAllocator *a, NphVecBuf *v, u32 elem);
#define VB(v, T) ((T *)(v).data)

// ---------------------------------------------------------------------------
// The dynamic AABB tree (aabb_tree.c): proxy ids, insertion, growth, the
// partial SAH rebuild and the query order fix which pairs a step creates
// and in what order.
// ---------------------------------------------------------------------------
#define NPH_TREE_NULL NPH_NULL32

typedef struct NphTree {
    NphAllocator *alloc;
    NphTreeNode *nodes; // node_capacity
    u32 root;
    u32 node_count;
    u32 node_capacity;
    u32 proxy_count;
    u32 *free_ids; // free node ids, the next one taken last
    u32 free_count;
    u32 *dirty;    // nodes whose box, categories, children or height changed
    u32 dirty_count;
    u8 *dirty_mark;
    u32 rebuild_capacity;
    u32 *leaf_indices;
    NphAABB *leaf_boxes;
    NphVec3 *leaf_centers;
    u32 *bins_scratch;
} NphTree;

typedef int (*NphTreeQueryFn)(u32 proxy, u64 user_data, void *context);

NphResult tree_init(NphTree *t, NphAllocator *alloc, u32 initial_capacity);
void tree_free(NphTree *t);
NphResult tree_create_proxy(NphTree *t, NphAABB aabb, u64 category, u64 user_data, u32 *out);
void tree_enlarge_proxy(NphTree *t, u32 proxy, NphAABB aabb);
void tree_query(const NphTree *t, NphAABB aabb, u64 mask, NphTreeQueryFn fn, void *ctx);
NphResult tree_rebuild(NphTree *t, int full);
u64 tree_user_data(const NphTree *t, u32 proxy);
void tree_destroy_proxy(NphTree *t, u32 proxy);
void tree_move_proxy(NphTree *t, u32 proxy, NphAABB aabb);
void tree_set_category_bits(NphTree *t, u32 proxy, u64 category_bits);
// The write list taken (the device's copy is up to date): cleared.
void tree_dirty_clear(NphTree *t);
u32 tree_height(const NphTree *t);

// ---------------------------------------------------------------------------
// Records
// ---------------------------------------------------------------------------
enum {
    HB_AWAKE = 1U << 0U,
    HB_ENABLE_SLEEP = 1U << 1U,
    HB_GYROSCOPIC = 1U << 2U,
    HB_BULLET = 1U << 5U,
    HB_MASS_SET = 1U << 6U, // mass properties set by the caller
    HB_VISITED = 1U << 8U,  // an island split's mark
};

typedef struct HBody {
    u32 handle; // 0: free slot
    u32 type;
    u32 flags;
    u32 shape_head;
    u32 shape_count;
    u32 contact_head; // contact << 1 | side
    u32 contact_count;
    u32 joint_head; // joint << 1 | side
    u32 joint_count;
    u32 awake_slot;
    u32 island; // its island (dynamic bodies), else NPH_NULL32
    u32 island_prev;
    u32 island_next;
    u32 move_event; // this step's move event, when it has one
    u64 user_data;
} HBody;

enum {
    HS_CONTACT_EVENTS = 1U << 0U,
    HS_HIT_EVENTS = 1U << 1U,
    HS_IN_MOVE_BUFFER = 1U << 3U,
    HS_SENSOR = 1U << 4U,
    HS_GHOST = 1U << 5U,
    HS_PAIR_DIRTY = 1U << 6U, // its pair-filter data waits for the device
};

typedef struct HShape {
    u32 handle;
    u32 body;
    u32 next;
    u32 flags;
    NphPhysicsGeometry geometry; // as given (host hull address)
    NphPhysicsPose pose;
    f32 density;
    f32 friction;
    f32 restitution;
    u32 combine;
    NphPhysicsFilter filter;
    u32 tree;
    u32 proxy;
    NphAABB aabb;
    NphAABB fat_aabb;
    u64 user_data;
} HShape;

enum {
    HC_TOUCHING = 1U << 0U,
    HC_EVENTS = 1U << 1U,
    HC_HIT_EVENTS = 1U << 2U,
    HC_ISLAND = 1U << 3U,  // links its bodies' islands
    HC_VISITED = 1U << 8U, // an island split's mark
};

typedef struct HContact {
    u32 shape_a; // NPH_NULL32: free record
    u32 shape_b;
    u32 body_a;
    u32 body_b;
    u32 flags;
    u32 prev[2];
    u32 next[2];
    u32 color;
    u32 color_slot;
    u32 awake_slot;
    u32 parent; // a mesh pair's child: its pair's record (else NPH_NULL32)
    u32 child;  // a mesh pair's next record (else NPH_NULL32)
    u32 island; // the island it links, when HC_ISLAND
    u32 island_prev;
    u32 island_next;
} HContact;

// joint.c / world_internal.h:280-330, less the step's numbers (the device's).
typedef struct HJoint {
    u32 handle; // 0: free slot
    u32 type;
    u32 body_a; // slots
    u32 body_b;
    u32 flags; // NPH_JOINT_COLLIDE_CONNECTED
    u32 prev[2];
    u32 next[2];
    u32 color;
    u32 color_slot;
    u32 island; // the island it links, when HJ_ISLAND
    u32 island_prev;
    u32 island_next;
    NphPhysicsPose frame_a;
    NphPhysicsPose frame_b;
    NphJointSettings s;
    u64 user_data;
} HJoint;

enum {
    HJ_ISLAND = 1U << 4U,  // links its bodies' islands (HJoint.flags)
    HJ_VISITED = 1U << 5U, // an island split's mark
};

// A set of bodies joined by touching contacts and joints, which sleeps and
// wakes as one (island.c): its bodies', contacts' and joints' lists, the
// union-find parent of a merge not yet carried out, its awake list slot.
typedef struct HIsland {
    u32 live;
    u32 awake;
    u32 body_head, body_tail, body_count;
    u32 contact_head, contact_tail, contact_count;
    u32 joint_head, joint_tail, joint_count;
    u32 parent;
    u32 awake_slot;
    u32 removed; // constraints removed since the last split: it may have come apart
    f32 sleep_time;
} HIsland;

typedef struct NphSlots {
    NphVecBuf generation; // u32
    NphVecBuf free;       // u32
} NphSlots;

// (shape a << 32 | shape b) -> contact; open addressing, linear probing.
typedef struct NphPairMap {
    u64 *keys; // 0: empty; ~0: removed
    u32 *values;
    u32 cap; // power of two
    u32 count;
    u32 used; // count + removed
} NphPairMap;


struct NphPhysicsWorld {
    NphAllocator *alloc;
    NphPhysicsWorldDesc desc;
    NphDevice *dev;
    f32 hit_threshold;
    int stepping;

    NphSlots body_slots;
    NphVecBuf bodies;    // HBody
    NphVecBuf body_recs; // NphBodyRec, the host copy (fresh unless bodies_stale)
    int bodies_stale;
    NphSlots shape_slots;
    NphVecBuf shapes;     // HShape
    NphVecBuf contacts;   // HContact
    NphVecBuf free_contacts; // u32
    NphVecBuf awake_bodies;   // u32 body slots
    NphVecBuf awake_contacts; // u32
    NphPairMap pairs;
    NphVecBuf inits; // NphContactInit of records made since the last collide
    NphVecBuf sweep_shapes; // NphShapeRec: a fast body's shapes beyond 64

    NphTree trees[3];
    NphVecBuf move_buffer; // u32 proxy keys

    NphVecBuf color_bits[NPH_GRAPH_COLORS]; // u64 words over body slots
    NphVecBuf color_contacts[NPH_GRAPH_COLORS + 1U]; // u32; [24] = overflow
    NphVecBuf graph; // u32, the colour lists concatenated for the device

    NphSlots joint_slots;
    NphVecBuf joints;                              // HJoint
    NphVecBuf color_joints[NPH_GRAPH_COLORS + 1U]; // u32 joint slots; [24] = overflow
    NphVecBuf step_joints;                         // u32, the colours' joints concatenated
    u32 joint_count;                               // live joints

    NphVecBuf changed; // u32
    NphVecBuf codes;   // u32

    NphVecBuf islands;        // HIsland
    NphVecBuf free_islands;   // u32
    NphVecBuf awake_islands;  // u32
    NphVecBuf island_work;    // u32: merge pairs, split seeds
    NphVecBuf island_stack;   // u32: a split's depth-first stack
    NphVecBuf sleep_list;     // u32: islands that fall asleep this step

    NphVecBuf begin_events; // NphPhysicsContactBeginEvent
    NphVecBuf hit_events;   // NphPhysicsContactHitEvent
    u32 hit_event_shapes;   // live shapes created with hit events; none means no contact reports hits
    u32 kinematic_bodies;   // live kinematic bodies (a body keeps its type for life)
    NphVecBuf end_events;
    NphVecBuf move_events; // NphPhysicsBodyMoveEvent, built at the first sync after the step
    int move_events_ready;
    u32 move_pending; // the awake bodies the step's move events cover
    NphVecBuf contact_view; // NphContactRec, readback scratch
    NphVecBuf shape_view;   // NphShapeRec as the device holds them, for queries
    int shape_view_ok;      // shape_view is current

    // Pair finding on the device: the search, what changed since the last
    // one (shapes' filter data, filter-set keys), and the set's size.
    NphPairSearch *pair_search;
    NphVecBuf pf_dirty_shapes; // u32 shape slots
    NphVecBuf pf_touched;      // u64 set keys
    NphVecBuf pf_keys;         // u64: a set that starts over
    NphVecBuf pf_key_in;       // u8
    NphVecBuf pf_shape_info;   // NphPairShape
    NphVecBuf pf_moved_shapes; // u32
    u32 pf_set_cap;            // 0: none yet
    u32 pf_set_used;           // slots taken (removed ones too), at most
    int pf_lost;               // a change went unrecorded: everything goes again
    f64 pairs_device_ms;
    f64 rebuild_ms; // the last step's tree rebuilds

    NphPhysicsStats stats;
    f32 step_dt;
    f32 last_substep; // the last step's substep length (the joint reaction's)
    u32 solved_contacts; // the last step's graph contacts
    u32 solved_joints;     // the last step's joints in colours (not the overflow)
    u32 solved_joint_ends; // their dynamic bodies
};

u32 phys_handle(u32 slot, u32 generation);
u32 phys_slot(u32 handle);
NphResult slots_alloc(NphAllocator *a, NphSlots *s, u32 *slot, u32 *generation);
void slots_release(NphAllocator *a, NphSlots *s, u32 slot);

HBody *world_body(const NphPhysicsWorld *w, NphPhysicsBody h);
HShape *world_shape(const NphPhysicsWorld *w, NphPhysicsShape h);
NphResult world_sync_bodies(NphPhysicsWorld *w); // device -> host copy when stale
// The device's contact records into contact_view (NphContactRec by record).
NphResult world_read_contacts(NphPhysicsWorld *w);
NphResult world_upload_body(NphPhysicsWorld *w, u32 slot);
NphResult world_wake_body(NphPhysicsWorld *w, u32 slot);

// island.c restated: islands, waking and sleep.
NphResult islands_add_body(NphPhysicsWorld *w, u32 body);
void islands_remove_body(NphPhysicsWorld *w, u32 body);
void islands_link_contact(NphPhysicsWorld *w, u32 contact);
void islands_unlink_contact(NphPhysicsWorld *w, u32 contact);
void islands_link_joint(NphPhysicsWorld *w, u32 joint);
void islands_unlink_joint(NphPhysicsWorld *w, u32 joint);
u32 islands_root(NphPhysicsWorld *w, u32 island);
NphResult islands_merge_pending(NphPhysicsWorld *w);
NphResult islands_sleep_and_split(NphPhysicsWorld *w);
NphResult islands_sleep_now(NphPhysicsWorld *w, u32 island);
// A body into (awake) or out of the awake list, then its contacts' and
// joints' awake and graph membership.
NphResult world_set_awake_list(NphPhysicsWorld *w, u32 body, int awake);
void islands_free(NphPhysicsWorld *w);
// contact.c / joint.c: a body's contacts' and joints' membership brought in
// line with their bodies.
NphResult contacts_sync_body(NphPhysicsWorld *w, u32 body);
// cast.c: convex proxies, GJK with poses and a simplex cache, casts and
// overlaps against one geometry.
typedef struct NphProxy {
    const NphVec3 *points;
    u32 count;
    f32 radius;
} NphProxy;
typedef struct NphSimplexCache {
    u32 count;
    u8 index_a[4];
    u8 index_b[4];
} NphSimplexCache;
typedef struct NphProxyDistance {
    NphDistance d;
    u32 simplex_count; // 4: the cores overlap
    u32 iterations;    // of the GJK loop
} NphProxyDistance;
NphProxy proxy_of(const NphPhysicsGeometry *g, NphVec3 storage[8]);
void proxy_gjk(const NphProxy *a, NphPhysicsPose pa, const NphProxy *b, NphPhysicsPose pb,
               NphSimplexCache *cache, NphGjk *g);
void proxy_distance(const NphProxy *a, NphPhysicsPose pa, const NphProxy *b, NphPhysicsPose pb,
                    int rounded, NphSimplexCache *cache, NphProxyDistance *out);
f32 proxy_gap(const NphProxy *m, NphPhysicsPose pose_m, const NphProxy *t, NphPhysicsPose pose_t,
              NphVec3 n);
NphVec3 proxy_contact_axis(const NphProxy *m, NphPhysicsPose pose_m, const NphProxy *t,
                           NphPhysicsPose pose_t, const NphSimplexCache *features, NphVec3 hint);
int raycast_local(const NphPhysicsGeometry *g, NphVec3 origin, NphVec3 delta,
                  NphPhysicsCastHit *out);
int cast_convex(const NphPhysicsGeometry *a, NphPhysicsPose pose_a, NphVec3 delta, f32 margin,
                const NphPhysicsGeometry *b, NphPhysicsPose pose_b, f32 limit,
                NphPhysicsCastHit *out);
int overlap_convex(const NphPhysicsGeometry *a, NphPhysicsPose pose_a, const NphPhysicsGeometry *b,
                   NphPhysicsPose pose_b);
int overlap_mesh_local(const NphPhysicsGeometry *a, NphQuat rotation, NphDVec3 position,
                       const NphPhysicsMesh *mesh);

// contact.c: the device pair search's change records and its host steps.
void pairs_mark_shape(NphPhysicsWorld *w, u32 shape);
void pairs_mark_body(NphPhysicsWorld *w, u32 body);
void pairs_touch_key(NphPhysicsWorld *w, u64 key);
NphResult pairs_prepare(NphPhysicsWorld *w, NphPairInput *in);
void pairs_sent(NphPhysicsWorld *w);
NphResult pairs_finish(NphPhysicsWorld *w, const u32 *pairs, u32 count);
NphResult joints_sync(NphPhysicsWorld *w, u32 joint);

// contact.c restated
// New pairs a moved proxy's counting pass keeps, so that the filling pass
// queries again only the proxies with more.
#define NPH_PAIR_CACHE 8U
NphResult contacts_update_pairs(NphPhysicsWorld *w);
// The step in phases, for the batch that steps many worlds together.
NphResult world_create(const NphPhysicsWorldDesc *desc, NphDeviceBatch *batch,
                       NphPhysicsWorld **out);
void world_step_begin(NphPhysicsWorld *w, f32 dt);
// The step's end: island merges, then sleep and splits when sleep is on.
NphResult world_step_end(NphPhysicsWorld *w);
// The solver's input (the awake list, the graph's colour lists, the step's
// numbers); in->awake_count = 0 when nothing is awake.
NphResult world_solve_input(NphPhysicsWorld *w, f32 dt, NphSolveInput *in);
// After the device's solve: grown proxies, pairs' bookkeeping, hit events.
NphResult world_solve_post(NphPhysicsWorld *w, const NphGrown *grown, u32 grown_count);
f64 world_now_ms(void);
NphResult contacts_reserve_changes(NphPhysicsWorld *w);
NphResult contacts_reserve_new(NphPhysicsWorld *w, u32 pairs);
NphResult contacts_collide_begin(NphPhysicsWorld *w);
NphResult contacts_collide_end(NphPhysicsWorld *w);
// The narrow phase's change codes of the awake list, entries [0, na) from
// codes_a and [na, na + nb) from codes_b, applied in list order.
NphResult contacts_apply_codes(NphPhysicsWorld *w, const u32 *codes_a, u32 na, const u32 *codes_b,
                               u32 nb);
NphResult contacts_post_solve(NphPhysicsWorld *w, const NphGrown *grown, u32 count);
NphResult move_buffer_push(NphPhysicsWorld *w, u32 shape);
u32 graph_pick_color(NphPhysicsWorld *w, u32 body_a, u32 body_b);
void graph_set_bodies(NphPhysicsWorld *w, u32 color, u32 body_a, u32 body_b, int set);
void contacts_destroy(NphPhysicsWorld *w, u32 contact, int events);
HContact *contacts_at(NphPhysicsWorld *w, u32 contact);

NphResult world_shape_record(NphPhysicsWorld *w, u32 s, NphShapeRec *out); // as the device holds it
NphGeom world_host_geom(const HShape *s); // the geometry with the host's hull address

// continuous.c restated: the fast bodies' sweeps, between finalize and the
// grown proxies.
NphResult continuous_sweep(NphPhysicsWorld *w, const NphFastBody *fast, u32 count);

// joint.c restated
int joints_allow_contact(const NphPhysicsWorld *w, u32 body_a, u32 body_b);
void joints_free(NphPhysicsWorld *w);
void joints_destroy(NphPhysicsWorld *w, u32 joint, int wake);
void world_update_mass(NphPhysicsWorld *w, u32 slot);
NphBodyRec *world_rec(NphPhysicsWorld *w, u32 slot);
NphResult world_refresh_shape(NphPhysicsWorld *w, u32 shape); // host record -> device

NphResult pairmap_init(NphAllocator *a, NphPairMap *m);
void pairmap_free(NphAllocator *a, NphPairMap *m);

#endif // NPH_WORLD_IMPL_H
