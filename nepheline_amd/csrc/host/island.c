// SPDX-License-Identifier: Apache-2.0
// Islands, waking and sleep, as the reference keeps them (island.c:139-735),
// on the host: every decision below fixes an order (the awake body list's
// swap-removals, which bodies leave the solver, which contacts leave the
// graph) that the device's step then follows.
//
// An island is a persistent set of dynamic bodies joined by touching
// contacts and joints, with its bodies, contacts and joints in intrusive
// lists. A contact that begins touching between two dynamic bodies links
// their islands through a union-find parent; the unions are carried out at
// the step's end, the smaller island's lists appended to the larger's. A
// contact or joint that goes away only counts a possible split; the island
// is split into its connected parts when it is about to sleep. An awake
// island falls asleep when the stillest-moving of its bodies has been still
// for the world's time to sleep (its bodies' sleep clocks run in the
// device's finalize); a sleeping island wakes whole when anything touches or
// changes one of its bodies.
#include "world_impl.h"

static HIsland *isl(NphPhysicsWorld *w, u32 i) { return &VB(w->islands, HIsland)[i]; }
static HBody *bod(NphPhysicsWorld *w, u32 i) { return &VB(w->bodies, HBody)[i]; }
static HContact *con(NphPhysicsWorld *w, u32 i) { return &VB(w->contacts, HContact)[i]; }
static HJoint *jnt(NphPhysicsWorld *w, u32 i) { return &VB(w->joints, HJoint)[i]; }
static int is_awake(const HBody *b) { return (b->flags & HB_AWAKE) != 0U; }

// ---------------------------------------------------------------------------
// The awake lists
// ---------------------------------------------------------------------------
static void awake_island_drop(NphPhysicsWorld *w, u32 island)
{
    HIsland *is = isl(w, island);
    if (is->awake_slot == NPH_NULL32) {
        return;
    }
    u32 *list = VB(w->awake_islands, u32);
    const u32 last = list[w->awake_islands.len - 1U];
    list[is->awake_slot] = last;
    isl(w, last)->awake_slot = is->awake_slot;
    w->awake_islands.len -= 1U;
    is->awake_slot = NPH_NULL32;
}

static NphResult awake_island_put(NphPhysicsWorld *w, u32 island)
{
    if (isl(w, island)->awake_slot != NPH_NULL32) {
        return NPH_RESULT_OK;
    }
    NphResult r = vbuf_push(w->alloc, &w->awake_islands, 4U, &island);
    if (r == NPH_RESULT_OK) {
        isl(w, island)->awake_slot = w->awake_islands.len - 1U;
    }
    return r;
}

// A body's record edited on the host and sent to the device again (the host
// copy brought up to date first).
static NphBodyRec *edit_rec(NphPhysicsWorld *w, u32 body)
{
    return world_sync_bodies(w) == NPH_RESULT_OK ? world_rec(w, body) : NULL;
}

// A body into the awake list (its sleep clock restarted) or out of it by
// swap, at rest: what motion is left is dropped, so it wakes as it fell
// asleep. Its contacts and joints are not touched here.
static NphResult body_list(NphPhysicsWorld *w, u32 body, int awake)
{
    NphBodyRec *rec = edit_rec(w, body);
    if (rec == NULL) {
        return NPH_RESULT_DEVICE_LOST;
    }
    HBody *b = bod(w, body);
    if (awake) {
        rec->sleep_time = 0.0F;
        NphResult r = world_upload_body(w, body);
        if (r != NPH_RESULT_OK || b->awake_slot != NPH_NULL32) {
            return r;
        }
        r = vbuf_push(w->alloc, &w->awake_bodies, 4U, &body);
        if (r == NPH_RESULT_OK) {
            b = bod(w, body);
            b->awake_slot = w->awake_bodies.len - 1U;
            b->flags |= HB_AWAKE;
        }
        return r;
    }
    if (b->awake_slot == NPH_NULL32) {
        return NPH_RESULT_OK;
    }
    u32 *list = VB(w->awake_bodies, u32);
    const u32 last = list[w->awake_bodies.len - 1U];
    list[b->awake_slot] = last;
    bod(w, last)->awake_slot = b->awake_slot;
    w->awake_bodies.len -= 1U;
    b->awake_slot = NPH_NULL32;
    b->flags &= ~(u32)HB_AWAKE;
    rec->v = v3_zero();
    rec->w = v3_zero();
    rec->force = v3_zero();
    rec->torque = v3_zero();
    return world_upload_body(w, body);
}

NphResult world_set_awake_list(NphPhysicsWorld *w, u32 body, int awake)
{
    NphResult r = body_list(w, body, awake);
    return r == NPH_RESULT_OK ? contacts_sync_body(w, body) : r;
}

// ---------------------------------------------------------------------------
// Island records and their lists
// ---------------------------------------------------------------------------
static NphResult island_make(NphPhysicsWorld *w, int awake, u32 *out)
{
    u32 index;
    if (w->free_islands.len > 0U) {
        w->free_islands.len -= 1U;
        index = VB(w->free_islands, u32)[w->free_islands.len];
    } else {
        NphResult r = vbuf_reserve(w->alloc, &w->islands, sizeof(HIsland), w->islands.len + 1U);
        if (r == NPH_RESULT_OK) {
            // Room to free every island without allocating.
            r = vbuf_reserve(w->alloc, &w->free_islands, 4U, w->islands.len + 1U);
        }
        if (r != NPH_RESULT_OK) {
            return r;
        }
        index = w->islands.len;
        w->islands.len += 1U;
    }
    HIsland *is = isl(w, index);
    memset(is, 0, sizeof *is);
    is->live = 1U;
    is->body_head = is->body_tail = NPH_NULL32;
    is->contact_head = is->contact_tail = NPH_NULL32;
    is->joint_head = is->joint_tail = NPH_NULL32;
    is->parent = NPH_NULL32;
    is->awake_slot = NPH_NULL32;
    if (awake) {
        is->awake = 1U;
        NphResult r = awake_island_put(w, index);
        if (r != NPH_RESULT_OK) {
            is->live = 0U;
            (void)vbuf_push(w->alloc, &w->free_islands, 4U, &index); // reserved
            return r;
        }
    }
    *out = index;
    return NPH_RESULT_OK;
}

static void island_release(NphPhysicsWorld *w, u32 island)
{
    awake_island_drop(w, island);
    isl(w, island)->live = 0U;
    (void)vbuf_push(w->alloc, &w->free_islands, 4U, &island); // reserved when made
}

// The three intrusive lists share one shape: head and tail in the island,
// prev and next in the member.
typedef struct Links {
    u32 *island;
    u32 *prev;
    u32 *next;
} Links;

static Links body_links(NphPhysicsWorld *w, u32 i)
{
    HBody *b = bod(w, i);
    Links l = {&b->island, &b->island_prev, &b->island_next};
    return l;
}

static Links contact_links(NphPhysicsWorld *w, u32 i)
{
    HContact *c = con(w, i);
    Links l = {&c->island, &c->island_prev, &c->island_next};
    return l;
}

static Links joint_links(NphPhysicsWorld *w, u32 i)
{
    HJoint *j = jnt(w, i);
    Links l = {&j->island, &j->island_prev, &j->island_next};
    return l;
}

enum { MEMBER_BODY = 0, MEMBER_CONTACT = 1, MEMBER_JOINT = 2 };

static Links member_links(NphPhysicsWorld *w, int kind, u32 i)
{
    return kind == MEMBER_BODY ? body_links(w, i)
                               : (kind == MEMBER_CONTACT ? contact_links(w, i) : joint_links(w, i));
}

static void island_ends(HIsland *is, int kind, u32 **head, u32 **tail, u32 **count)
{
    if (kind == MEMBER_BODY) {
        *head = &is->body_head;
        *tail = &is->body_tail;
        *count = &is->body_count;
    } else if (kind == MEMBER_CONTACT) {
        *head = &is->contact_head;
        *tail = &is->contact_tail;
        *count = &is->contact_count;
    } else {
        *head = &is->joint_head;
        *tail = &is->joint_tail;
        *count = &is->joint_count;
    }
}

static void island_push(NphPhysicsWorld *w, u32 island, int kind, u32 member)
{
    u32 *head, *tail, *count;
    island_ends(isl(w, island), kind, &head, &tail, &count);
    const Links l = member_links(w, kind, member);
    *l.island = island;
    *l.prev = *tail;
    *l.next = NPH_NULL32;
    if (*tail != NPH_NULL32) {
        *member_links(w, kind, *tail).next = member;
    } else {
        *head = member;
    }
    *tail = member;
    *count += 1U;
}

static void island_cut(NphPhysicsWorld *w, int kind, u32 member)
{
    const Links l = member_links(w, kind, member);
    u32 *head, *tail, *count;
    island_ends(isl(w, *l.island), kind, &head, &tail, &count);
    if (*l.prev != NPH_NULL32) {
        *member_links(w, kind, *l.prev).next = *l.next;
    } else {
        *head = *l.next;
    }
    if (*l.next != NPH_NULL32) {
        *member_links(w, kind, *l.next).prev = *l.prev;
    } else {
        *tail = *l.prev;
    }
    *count -= 1U;
    *l.island = NPH_NULL32;
    *l.prev = NPH_NULL32;
    *l.next = NPH_NULL32;
}

// ---------------------------------------------------------------------------
// Membership
// ---------------------------------------------------------------------------

// A new dynamic body: an island of its own, awake (a body made asleep then
// leaves the awake list alone, its island still counted awake).
NphResult islands_add_body(NphPhysicsWorld *w, u32 body)
{
    u32 island = NPH_NULL32;
    NphResult r = island_make(w, 1, &island);
    if (r == NPH_RESULT_OK) {
        island_push(w, island, MEMBER_BODY, body);
    }
    return r;
}

void islands_remove_body(NphPhysicsWorld *w, u32 body)
{
    const u32 island = bod(w, body)->island;
    if (island == NPH_NULL32) {
        return;
    }
    island_cut(w, MEMBER_BODY, body);
    HIsland *is = isl(w, island);
    if (is->body_count == 0U) {
        island_release(w, island);
    } else {
        is->removed += 1U; // its contacts went first: what is left may fall apart
    }
}

u32 islands_root(NphPhysicsWorld *w, u32 island)
{
    u32 root = island;
    while (isl(w, root)->parent != NPH_NULL32) {
        root = isl(w, root)->parent;
    }
    for (u32 i = island; isl(w, i)->parent != NPH_NULL32;) { // path compression
        const u32 up = isl(w, i)->parent;
        isl(w, i)->parent = root;
        i = up;
    }
    return root;
}

// The root the two bodies' islands share once linked: the one with more
// bodies (ties: the lower index) becomes the other's parent.
static u32 island_join(NphPhysicsWorld *w, u32 body_a, u32 body_b)
{
    const u32 ra = islands_root(w, bod(w, body_a)->island);
    const u32 rb = islands_root(w, bod(w, body_b)->island);
    if (ra == rb) {
        return ra;
    }
    const u32 na = isl(w, ra)->body_count;
    const u32 nb = isl(w, rb)->body_count;
    const int keep_a = na > nb || (na == nb && ra < rb);
    isl(w, keep_a ? rb : ra)->parent = keep_a ? ra : rb;
    return keep_a ? ra : rb;
}

void islands_link_contact(NphPhysicsWorld *w, u32 contact)
{
    const HContact *c = con(w, contact);
    island_push(w, island_join(w, c->body_a, c->body_b), MEMBER_CONTACT, contact);
    con(w, contact)->flags |= HC_ISLAND;
}

void islands_unlink_contact(NphPhysicsWorld *w, u32 contact)
{
    isl(w, con(w, contact)->island)->removed += 1U;
    island_cut(w, MEMBER_CONTACT, contact);
    con(w, contact)->flags &= ~(u32)HC_ISLAND;
}

void islands_link_joint(NphPhysicsWorld *w, u32 joint)
{
    const HJoint *j = jnt(w, joint);
    island_push(w, island_join(w, j->body_a, j->body_b), MEMBER_JOINT, joint);
    jnt(w, joint)->flags |= HJ_ISLAND;
}

void islands_unlink_joint(NphPhysicsWorld *w, u32 joint)
{
    isl(w, jnt(w, joint)->island)->removed += 1U;
    island_cut(w, MEMBER_JOINT, joint);
    jnt(w, joint)->flags &= ~(u32)HJ_ISLAND;
}

// ---------------------------------------------------------------------------
// Merges (island.c:427-492)
// ---------------------------------------------------------------------------

// Island `from`'s members appended to `into`'s lists in list order; `from`
// freed.
static void island_absorb(NphPhysicsWorld *w, u32 into, u32 from)
{
    for (int kind = MEMBER_BODY; kind <= MEMBER_JOINT; ++kind) {
        u32 *head, *tail, *count;
        island_ends(isl(w, from), kind, &head, &tail, &count);
        for (u32 m = *head; m != NPH_NULL32;) {
            const u32 next = *member_links(w, kind, m).next;
            island_push(w, into, kind, m);
            m = next;
        }
        island_ends(isl(w, from), kind, &head, &tail, &count);
        *head = *tail = NPH_NULL32;
        *count = 0U;
    }
    isl(w, into)->removed += isl(w, from)->removed;
    island_release(w, from);
}

NphResult islands_merge_pending(NphPhysicsWorld *w)
{
    // Every child and its final root first (freeing an island must not cut
    // a later child's path), in the awake list's order; then the merges.
    w->island_work.len = 0U;
    for (u32 i = 0U; i < w->awake_islands.len; ++i) {
        const u32 island = VB(w->awake_islands, u32)[i];
        if (isl(w, island)->parent == NPH_NULL32) {
            continue;
        }
        const u32 pair[2] = {island, islands_root(w, island)};
        NphResult r = vbuf_push(w->alloc, &w->island_work, 8U, pair);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    const u32 merges = w->island_work.len;
    for (u32 i = 0U; i < merges; ++i) {
        const u32 child = VB(w->island_work, u32)[2U * i];
        const u32 root = VB(w->island_work, u32)[2U * i + 1U];
        isl(w, child)->parent = NPH_NULL32;
        island_absorb(w, root, child);
    }
    return NPH_RESULT_OK;
}

// ---------------------------------------------------------------------------
// Waking and sleep (island.c:494-735)
// ---------------------------------------------------------------------------
static NphResult island_wake(NphPhysicsWorld *w, u32 island)
{
    HIsland *is = isl(w, island);
    if (is->awake != 0U) {
        // Awake already: the call still restarts its bodies' sleep clocks.
        NphResult r = NPH_RESULT_OK;
        for (u32 b = is->body_head; b != NPH_NULL32 && r == NPH_RESULT_OK; b = bod(w, b)->island_next) {
            NphBodyRec *rec = edit_rec(w, b);
            if (rec == NULL) {
                return NPH_RESULT_DEVICE_LOST;
            }
            rec->sleep_time = 0.0F;
            r = world_upload_body(w, b);
        }
        isl(w, island)->sleep_time = 0.0F;
        return r;
    }
    is->awake = 1U;
    is->sleep_time = 0.0F;
    NphResult r = awake_island_put(w, island);
    for (u32 b = isl(w, island)->body_head; b != NPH_NULL32 && r == NPH_RESULT_OK;
         b = bod(w, b)->island_next) {
        r = body_list(w, b, 1);
    }
    for (u32 b = isl(w, island)->body_head; b != NPH_NULL32 && r == NPH_RESULT_OK;
         b = bod(w, b)->island_next) {
        r = contacts_sync_body(w, b);
    }
    return r;
}

NphResult world_wake_body(NphPhysicsWorld *w, u32 slot)
{
    const HBody *b = bod(w, slot);
    if (b->type == NPH_PHYSICS_BODY_STATIC) {
        return NPH_RESULT_OK;
    }
    if (b->type == NPH_PHYSICS_BODY_DYNAMIC) {
        return island_wake(w, islands_root(w, b->island));
    }
    if (is_awake(b)) { // kinematic, awake: its clock restarts
        NphBodyRec *rec = edit_rec(w, slot);
        if (rec == NULL) {
            return NPH_RESULT_DEVICE_LOST;
        }
        rec->sleep_time = 0.0F;
        return world_upload_body(w, slot);
    }
    return world_set_awake_list(w, slot, 1);
}

static NphResult island_sleep(NphPhysicsWorld *w, u32 island)
{
    NphResult r = NPH_RESULT_OK;
    for (u32 b = isl(w, island)->body_head; b != NPH_NULL32 && r == NPH_RESULT_OK;
         b = bod(w, b)->island_next) {
        // Its move event of this step says so, at rest.
        const HBody *hb = bod(w, b);
        if (hb->move_event < w->move_events.len) {
            NphPhysicsBodyMoveEvent *e =
                &VB(w->move_events, NphPhysicsBodyMoveEvent)[hb->move_event];
            if (e->body.value == hb->handle) {
                e->fell_asleep = 1U;
                e->linear_velocity = v3_zero();
                e->angular_velocity = v3_zero();
            }
        }
        r = body_list(w, b, 0);
    }
    for (u32 b = isl(w, island)->body_head; b != NPH_NULL32 && r == NPH_RESULT_OK;
         b = bod(w, b)->island_next) {
        r = contacts_sync_body(w, b);
    }
    if (r == NPH_RESULT_OK) {
        isl(w, island)->awake = 0U;
        awake_island_drop(w, island);
    }
    return r;
}

NphResult islands_sleep_now(NphPhysicsWorld *w, u32 island)
{
    return isl(w, island)->awake != 0U ? island_sleep(w, island) : NPH_RESULT_OK;
}

// The neighbours of body x joined to it by an island-linked contact or joint
// not yet visited: marked, pushed, and the link appended to `part`.
static void split_visit(NphPhysicsWorld *w, u32 x, u32 part)
{
    for (u32 key = bod(w, x)->contact_head; key != NPH_NULL32;) {
        HContact *c = con(w, key >> 1U);
        const u32 contact = key >> 1U;
        const u32 other = (key & 1U) == 0U ? c->body_b : c->body_a;
        key = c->next[key & 1U];
        if ((c->flags & HC_ISLAND) == 0U || (c->flags & HC_VISITED) != 0U) {
            continue;
        }
        c->flags |= HC_VISITED;
        island_push(w, part, MEMBER_CONTACT, contact);
        if ((bod(w, other)->flags & HB_VISITED) == 0U) {
            bod(w, other)->flags |= HB_VISITED;
            (void)vbuf_push(w->alloc, &w->island_stack, 4U, &other); // reserved
        }
    }
    for (u32 key = bod(w, x)->joint_head; key != NPH_NULL32;) {
        HJoint *j = jnt(w, key >> 1U);
        const u32 joint = key >> 1U;
        const u32 other = (key & 1U) == 0U ? j->body_b : j->body_a;
        key = j->next[key & 1U];
        if ((j->flags & HJ_ISLAND) == 0U || (j->flags & HJ_VISITED) != 0U) {
            continue;
        }
        j->flags |= HJ_VISITED;
        island_push(w, part, MEMBER_JOINT, joint);
        if ((bod(w, other)->flags & HB_VISITED) == 0U) {
            bod(w, other)->flags |= HB_VISITED;
            (void)vbuf_push(w->alloc, &w->island_stack, 4U, &other); // reserved
        }
    }
}

// The island's connected parts (bodies joined by linked contacts and
// joints), depth first from its bodies in list order; the first part keeps
// the island's record.
static NphResult island_split(NphPhysicsWorld *w, u32 island)
{
    w->island_work.len = 0U;
    for (u32 b = isl(w, island)->body_head; b != NPH_NULL32; b = bod(w, b)->island_next) {
        NphResult r = vbuf_push(w->alloc, &w->island_work, 4U, &b);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    const u32 n = w->island_work.len;
    NphResult r = vbuf_reserve(w->alloc, &w->island_stack, 4U, n > 0U ? n : 1U);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    // Detach the contacts and joints; the parts rebuild every list.
    HIsland *is = isl(w, island);
    for (u32 c = is->contact_head; c != NPH_NULL32;) {
        const u32 next = con(w, c)->island_next;
        con(w, c)->island = NPH_NULL32;
        c = next;
    }
    for (u32 j = is->joint_head; j != NPH_NULL32;) {
        const u32 next = jnt(w, j)->island_next;
        jnt(w, j)->island = NPH_NULL32;
        j = next;
    }
    is->body_head = is->body_tail = NPH_NULL32;
    is->contact_head = is->contact_tail = NPH_NULL32;
    is->joint_head = is->joint_tail = NPH_NULL32;
    is->body_count = is->contact_count = is->joint_count = 0U;
    for (u32 i = 0U; i < n && r == NPH_RESULT_OK; ++i) {
        const u32 seed = VB(w->island_work, u32)[i];
        if ((bod(w, seed)->flags & HB_VISITED) != 0U) {
            continue;
        }
        u32 part = island;
        if (isl(w, island)->body_count != 0U) {
            r = island_make(w, 1, &part);
            if (r != NPH_RESULT_OK) {
                break;
            }
        }
        w->island_stack.len = 0U;
        bod(w, seed)->flags |= HB_VISITED;
        (void)vbuf_push(w->alloc, &w->island_stack, 4U, &seed); // reserved
        while (w->island_stack.len > 0U) {
            w->island_stack.len -= 1U;
            const u32 x = VB(w->island_stack, u32)[w->island_stack.len];
            island_push(w, part, MEMBER_BODY, x);
            split_visit(w, x, part);
        }
    }
    // Marks off, over every body that was in the island and its links.
    for (u32 i = 0U; i < n; ++i) {
        const u32 b = VB(w->island_work, u32)[i];
        bod(w, b)->flags &= ~(u32)HB_VISITED;
        for (u32 key = bod(w, b)->contact_head; key != NPH_NULL32;) {
            HContact *c = con(w, key >> 1U);
            c->flags &= ~(u32)HC_VISITED;
            key = c->next[key & 1U];
        }
        for (u32 key = bod(w, b)->joint_head; key != NPH_NULL32;) {
            HJoint *j = jnt(w, key >> 1U);
            j->flags &= ~(u32)HJ_VISITED;
            key = j->next[key & 1U];
        }
    }
    isl(w, island)->removed = 0U;
    return r;
}

NphResult islands_sleep_and_split(NphPhysicsWorld *w)
{
    NphResult r = world_sync_bodies(w); // the sleep clocks finalize ran
    if (r != NPH_RESULT_OK) {
        return r;
    }
    const f32 time_to_sleep = w->desc.time_to_sleep;
    // An island's clock is its stillest-moving body's. One that lost
    // constraints may have come apart: the one whose sleepiest body is the
    // sleepiest of all such islands is split (one a step), and sleeps in
    // parts later; the others that are still long enough fall asleep, in
    // the awake list's order.
    w->sleep_list.len = 0U;
    u32 split = NPH_NULL32;
    f32 split_time = 0.0F;
    for (u32 i = 0U; i < w->awake_islands.len; ++i) {
        const u32 island = VB(w->awake_islands, u32)[i];
        HIsland *is = isl(w, island);
        f32 least = 1e30F;
        f32 most = 0.0F;
        for (u32 b = is->body_head; b != NPH_NULL32; b = bod(w, b)->island_next) {
            const f32 t = (bod(w, b)->flags & HB_ENABLE_SLEEP) != 0U ? world_rec(w, b)->sleep_time
                                                                       : 0.0F;
            least = nph_minf(least, t);
            most = nph_maxf(most, t);
        }
        is->sleep_time = least;
        if (is->removed > 0U) {
            if (most >= time_to_sleep && most > split_time) {
                split = island;
                split_time = most;
            }
        } else if (least >= time_to_sleep) {
            r = vbuf_push(w->alloc, &w->sleep_list, 4U, &island);
            if (r != NPH_RESULT_OK) {
                return r;
            }
        }
    }
    for (u32 i = 0U; i < w->sleep_list.len && r == NPH_RESULT_OK; ++i) {
        const u32 island = VB(w->sleep_list, u32)[i];
        if (isl(w, island)->live != 0U && isl(w, island)->awake != 0U) {
            r = island_sleep(w, island);
        }
    }
    if (r == NPH_RESULT_OK && split != NPH_NULL32) {
        r = island_split(w, split);
    }
    // Kinematic bodies sleep on their own clocks.
    for (u32 i = 0U; i < w->awake_bodies.len && r == NPH_RESULT_OK;) {
        const u32 b = VB(w->awake_bodies, u32)[i];
        const HBody *hb = bod(w, b);
        if (hb->type == NPH_PHYSICS_BODY_KINEMATIC && (hb->flags & HB_ENABLE_SLEEP) != 0U &&
            world_rec(w, b)->sleep_time >= time_to_sleep) {
            r = world_set_awake_list(w, b, 0);
            continue; // the slot holds another body now
        }
        i += 1U;
    }
    return r;
}

void islands_free(NphPhysicsWorld *w)
{
    vbuf_free(w->alloc, &w->islands, sizeof(HIsland));
    vbuf_free(w->alloc, &w->free_islands, 4U);
    vbuf_free(w->alloc, &w->awake_islands, 4U);
    vbuf_free(w->alloc, &w->island_work, 4U);
    vbuf_free(w->alloc, &w->island_stack, 4U);
    vbuf_free(w->alloc, &w->sleep_list, 4U);
}
