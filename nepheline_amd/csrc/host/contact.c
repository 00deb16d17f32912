// SPDX-License-Identifier: Apache-2.0
// Contacts on the host: the pool and its free list, the pair map, the
// broadphase's new pairs, the serial update of what began and stopped
// touching, and the constraint graph's colours.  Restated from
// engine/physics/src/contact.c (:12-263, :591-854) and island.c (:34-135);
// the per-contact narrow phase runs on the device (nph_dev_collide).
#include "world_impl.h"

#include <omp.h>

// ---------------------------------------------------------------------------
// The pair map
// ---------------------------------------------------------------------------
#define MAP_EMPTY 0ULL
#define MAP_GONE (~0ULL)

// contact.c:12-15; +1 keeps key 0 free for empty buckets.
static u64 pair_key(u32 a, u32 b) { return nph_pair_key(a, b); }

static u64 mix64(u64 x)
{
    x ^= x >> 33U;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33U;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33U;
    return x;
}

static void *halloc(NphAllocator *a, u64 size)
{
    NphCallsite cs = {__FILE__, (u32)__LINE__};
    return a->vtable->alloc(a, size, 16U, cs);
}

static void hfree(NphAllocator *a, void *p, u64 size)
{
    NphCallsite cs = {__FILE__, (u32)__LINE__};
    if (p != NULL) {
        a->vtable->free(a, p, size, 16U, cs);
    }
}

NphResult pairmap_init(NphAllocator *a, NphPairMap *m)
{
    memset(m, 0, sizeof *m);
    m->cap = 1024U;
    m->keys = (u64 *)halloc(a, (u64)m->cap * 8U);
    m->values = (u32 *)halloc(a, (u64)m->cap * 4U);
    if (m->keys == NULL || m->values == NULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    memset(m->keys, 0, (size_t)m->cap * 8U);
    return NPH_RESULT_OK;
}

void pairmap_free(NphAllocator *a, NphPairMap *m)
{
    hfree(a, m->keys, (u64)m->cap * 8U);
    hfree(a, m->values, (u64)m->cap * 4U);
    memset(m, 0, sizeof *m);
}

static u32 map_find(const NphPairMap *m, u64 key)
{
    u32 i = (u32)mix64(key) & (m->cap - 1U);
    for (;;) {
        const u64 k = m->keys[i];
        if (k == key) {
            return i;
        }
        if (k == MAP_EMPTY) {
            return NPH_NULL32;
        }
        i = (i + 1U) & (m->cap - 1U);
    }
}

static NphResult map_insert(NphAllocator *a, NphPairMap *m, u64 key, u32 value);

// The map rebuilt at capacity `cap` (removed keys dropped).
static NphResult map_rebuild(NphAllocator *a, NphPairMap *m, u32 cap)
{
    NphPairMap n;
    memset(&n, 0, sizeof n);
    n.cap = cap;
    n.keys = (u64 *)halloc(a, (u64)n.cap * 8U);
    n.values = (u32 *)halloc(a, (u64)n.cap * 4U);
    if (n.keys == NULL || n.values == NULL) {
        hfree(a, n.keys, (u64)n.cap * 8U);
        hfree(a, n.values, (u64)n.cap * 4U);
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    memset(n.keys, 0, (size_t)n.cap * 8U);
    for (u32 i = 0U; i < m->cap; ++i) {
        if (m->keys[i] != MAP_EMPTY && m->keys[i] != MAP_GONE) {
            map_insert(a, &n, m->keys[i], m->values[i]);
        }
    }
    pairmap_free(a, m);
    *m = n;
    return NPH_RESULT_OK;
}

static NphResult map_grow(NphAllocator *a, NphPairMap *m)
{
    return map_rebuild(a, m, m->count * 4U > m->cap ? m->cap * 2U : m->cap);
}

// Room for `extra` more keys without growing.
static NphResult map_reserve(NphAllocator *a, NphPairMap *m, u32 extra)
{
    if ((u64)(m->used + extra) * 2U <= m->cap) {
        return NPH_RESULT_OK;
    }
    u32 cap = m->cap;
    while ((u64)(m->count + extra) * 2U > cap) {
        cap *= 2U;
    }
    return map_rebuild(a, m, cap);
}

static NphResult map_insert(NphAllocator *a, NphPairMap *m, u64 key, u32 value)
{
    if ((m->used + 1U) * 2U > m->cap) {
        NphResult r = map_grow(a, m);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    u32 i = (u32)mix64(key) & (m->cap - 1U);
    while (m->keys[i] != MAP_EMPTY && m->keys[i] != MAP_GONE) {
        i = (i + 1U) & (m->cap - 1U);
    }
    if (m->keys[i] == MAP_EMPTY) {
        m->used += 1U;
    }
    m->keys[i] = key;
    m->values[i] = value;
    m->count += 1U;
    return NPH_RESULT_OK;
}

static void map_remove(NphPairMap *m, u64 key)
{
    const u32 i = map_find(m, key);
    if (i != NPH_NULL32) {
        m->keys[i] = MAP_GONE;
        m->count -= 1U;
    }
}

// ---------------------------------------------------------------------------
// Lifetime (contact.c:27-237)
// ---------------------------------------------------------------------------
static HContact *contact_at(NphPhysicsWorld *w, u32 i) { return &VB(w->contacts, HContact)[i]; }
static HBody *body_at(NphPhysicsWorld *w, u32 i) { return &VB(w->bodies, HBody)[i]; }
static HShape *shape_at(NphPhysicsWorld *w, u32 i) { return &VB(w->shapes, HShape)[i]; }
static int body_awake(const HBody *b) { return (b->flags & HB_AWAKE) != 0U; }

static void list_insert(NphPhysicsWorld *w, u32 contact, u32 side, u32 body)
{
    HBody *b = body_at(w, body);
    HContact *c = contact_at(w, contact);
    const u32 key = (contact << 1U) | side;
    c->prev[side] = NPH_NULL32;
    c->next[side] = b->contact_head;
    if (b->contact_head != NPH_NULL32) {
        contact_at(w, b->contact_head >> 1U)->prev[b->contact_head & 1U] = key;
    }
    b->contact_head = key;
    b->contact_count += 1U;
}

static void list_remove(NphPhysicsWorld *w, u32 contact, u32 side, u32 body)
{
    HBody *b = body_at(w, body);
    HContact *c = contact_at(w, contact);
    const u32 prev = c->prev[side];
    const u32 next = c->next[side];
    if (prev != NPH_NULL32) {
        contact_at(w, prev >> 1U)->next[prev & 1U] = next;
    } else {
        b->contact_head = next;
    }
    if (next != NPH_NULL32) {
        contact_at(w, next >> 1U)->prev[next & 1U] = prev;
    }
    b->contact_count -= 1U;
}

static NphResult awake_list_add(NphPhysicsWorld *w, u32 contact)
{
    HContact *c = contact_at(w, contact);
    if (c->awake_slot != NPH_NULL32) {
        return NPH_RESULT_OK;
    }
    NphResult r = vbuf_push(w->alloc, &w->awake_contacts, 4U, &contact);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    c->awake_slot = w->awake_contacts.len - 1U;
    return NPH_RESULT_OK;
}

static void awake_list_remove(NphPhysicsWorld *w, u32 contact)
{
    HContact *c = contact_at(w, contact);
    const u32 slot = c->awake_slot;
    if (slot == NPH_NULL32) {
        return;
    }
    u32 *list = VB(w->awake_contacts, u32);
    const u32 last = list[w->awake_contacts.len - 1U];
    list[slot] = last;
    contact_at(w, last)->awake_slot = slot;
    w->awake_contacts.len -= 1U;
    c->awake_slot = NPH_NULL32;
}

// ---------------------------------------------------------------------------
// The constraint graph (island.c:34-104)
// ---------------------------------------------------------------------------
static int bit_test(const NphVecBuf *bits, u32 i)
{
    return (VB(*bits, u64)[i >> 6U] >> (i & 63U)) & 1U;
}

static void bit_set(NphVecBuf *bits, u32 i, int on)
{
    u64 *wds = VB(*bits, u64);
    if (on) {
        wds[i >> 6U] |= 1ULL << (i & 63U);
    } else {
        wds[i >> 6U] &= ~(1ULL << (i & 63U));
    }
}

u32 graph_pick_color(NphPhysicsWorld *w, u32 body_a, u32 body_b)
{
    const int dyn_a = body_at(w, body_a)->type == NPH_PHYSICS_BODY_DYNAMIC;
    const int dyn_b = body_at(w, body_b)->type == NPH_PHYSICS_BODY_DYNAMIC;
    if (dyn_a && dyn_b) {
        for (u32 k = 0U; k < NPH_GRAPH_DYNAMIC_COLORS; ++k) {
            if (!bit_test(&w->color_bits[k], body_a) && !bit_test(&w->color_bits[k], body_b)) {
                return k;
            }
        }
        return NPH_GRAPH_OVERFLOW;
    }
    const u32 body = dyn_a ? body_a : body_b;
    for (u32 k = 1U; k < NPH_GRAPH_COLORS; ++k) {
        if (!bit_test(&w->color_bits[k], body)) {
            return k;
        }
    }
    return NPH_GRAPH_OVERFLOW;
}

void graph_set_bodies(NphPhysicsWorld *w, u32 color, u32 body_a, u32 body_b, int set)
{
    if (color == NPH_GRAPH_OVERFLOW) {
        return;
    }
    const u32 bodies[2] = {body_a, body_b};
    for (u32 i = 0U; i < 2U; ++i) {
        if (body_at(w, bodies[i])->type == NPH_PHYSICS_BODY_DYNAMIC) {
            bit_set(&w->color_bits[color], bodies[i], set);
        }
    }
}

static NphResult graph_add(NphPhysicsWorld *w, u32 contact)
{
    HContact *c = contact_at(w, contact);
    const u32 color = graph_pick_color(w, c->body_a, c->body_b);
    NphVecBuf *list = &w->color_contacts[color];
    NphResult r = vbuf_push(w->alloc, list, 4U, &contact);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    graph_set_bodies(w, color, c->body_a, c->body_b, 1);
    c->color = color;
    c->color_slot = list->len - 1U;
    return NPH_RESULT_OK;
}

static void graph_remove(NphPhysicsWorld *w, u32 contact)
{
    HContact *c = contact_at(w, contact);
    const u32 color = c->color;
    NphVecBuf *list = &w->color_contacts[color];
    u32 *d = VB(*list, u32);
    const u32 slot = c->color_slot;
    const u32 last = d[list->len - 1U];
    d[slot] = last;
    contact_at(w, last)->color_slot = slot;
    list->len -= 1U;
    graph_set_bodies(w, color, c->body_a, c->body_b, 0);
    c->color = NPH_NULL32;
    c->color_slot = NPH_NULL32;
}

// contact.c:239-263: the pair's awake membership, then the graph membership
// of its record and (a mesh pair's) children, in chain order.
static NphResult contact_sync(NphPhysicsWorld *w, u32 contact)
{
    const HContact *c = contact_at(w, contact);
    const HBody *a = body_at(w, c->body_a);
    const HBody *b = body_at(w, c->body_b);
    const int awake = body_awake(a) || body_awake(b);
    NphResult r = NPH_RESULT_OK;
    if (awake) {
        r = awake_list_add(w, contact);
    } else {
        awake_list_remove(w, contact);
    }
    if (r != NPH_RESULT_OK) {
        return r;
    }
    const int solvable = awake && (body_awake(a) || a->type != NPH_PHYSICS_BODY_DYNAMIC) &&
                         (body_awake(b) || b->type != NPH_PHYSICS_BODY_DYNAMIC);
    for (u32 k = contact; k != NPH_NULL32; k = contact_at(w, k)->child) {
        const HContact *rk = contact_at(w, k);
        const int in_graph = solvable && (rk->flags & HC_TOUCHING) != 0U;
        if (in_graph && rk->color == NPH_NULL32) {
            r = graph_add(w, k);
            if (r != NPH_RESULT_OK) {
                return r;
            }
        } else if (!in_graph && rk->color != NPH_NULL32) {
            graph_remove(w, k);
        }
    }
    return NPH_RESULT_OK;
}

static int is_surface(const HShape *s) { return s->geometry.type == NPH_PHYSICS_SHAPE_MESH; }

static NphResult contact_alloc(NphPhysicsWorld *w, u32 *out)
{
    if (w->free_contacts.len > 0U) {
        *out = VB(w->free_contacts, u32)[w->free_contacts.len - 1U];
        w->free_contacts.len -= 1U;
        return NPH_RESULT_OK;
    }
    NphResult r = vbuf_reserve(w->alloc, &w->contacts, sizeof(HContact), w->contacts.len + 1U);
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(w->alloc, &w->free_contacts, 4U, w->contacts.len + 1U);
    }
    if (r == NPH_RESULT_OK) {
        r = nph_dev_reserve(w->dev, 0U, 0U, w->contacts.len + 1U);
    }
    if (r != NPH_RESULT_OK) {
        return r;
    }
    w->contacts.len += 1U;
    *out = w->contacts.len - 1U;
    return NPH_RESULT_OK;
}

// contact.c:129-179: one record, or a mesh pair's four (the pair's, then
// its children, each from the free list in turn).
static NphResult contact_create(NphPhysicsWorld *w, u32 shape_a, u32 shape_b)
{
    const HShape *sa = shape_at(w, shape_a);
    const HShape *sb = shape_at(w, shape_b);
    const u32 records =
        is_surface(sa) || is_surface(sb) ? NPH_PHYSICS_MESH_MAX_MANIFOLDS : 1U;
    u32 rec[NPH_PHYSICS_MESH_MAX_MANIFOLDS] = {NPH_NULL32, NPH_NULL32, NPH_NULL32, NPH_NULL32};
    NphResult r = NPH_RESULT_OK;
    u32 made = 0U;
    while (made < records && r == NPH_RESULT_OK) {
        r = contact_alloc(w, &rec[made]);
        made += r == NPH_RESULT_OK ? 1U : 0U;
    }
    const u32 index = rec[0];
    if (r == NPH_RESULT_OK) {
        r = map_insert(w->alloc, &w->pairs, pair_key(shape_a, shape_b), index);
    }
    if (r == NPH_RESULT_OK) {
        pairs_touch_key(w, pair_key(shape_a, shape_b));
    }
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(w->alloc, &w->awake_contacts, 4U, w->contacts.len);
    }
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(w->alloc, &w->inits, sizeof(NphContactInit), w->inits.len + records);
    }
    if (r != NPH_RESULT_OK) {
        for (u32 k = 0U; k < made; ++k) {
            (void)vbuf_push(w->alloc, &w->free_contacts, 4U, &rec[k]); // reserved
        }
        return r;
    }
    // The shapes may have moved in memory (contact_alloc grows nothing of
    // theirs), so read them again.
    sa = shape_at(w, shape_a);
    sb = shape_at(w, shape_b);
    for (u32 k = 0U; k < records; ++k) {
        HContact *c = contact_at(w, rec[k]);
        memset(c, 0, sizeof *c);
        c->shape_a = shape_a;
        c->shape_b = shape_b;
        c->body_a = sa->body;
        c->body_b = sb->body;
        c->color = NPH_NULL32;
        c->color_slot = NPH_NULL32;
        c->awake_slot = NPH_NULL32;
        c->island = NPH_NULL32;
        c->island_prev = NPH_NULL32;
        c->island_next = NPH_NULL32;
        c->parent = k == 0U ? NPH_NULL32 : index;
        c->child = k + 1U < records ? rec[k + 1U] : NPH_NULL32;
        const NphContactInit init = {rec[k], shape_a, shape_b, c->body_a, c->body_b, c->child};
        (void)vbuf_push(w->alloc, &w->inits, sizeof init, &init); // reserved
    }
    HContact *c = contact_at(w, index);
    if (((sa->flags | sb->flags) & HS_CONTACT_EVENTS) != 0U) {
        c->flags |= HC_EVENTS;
    }
    if (((sa->flags | sb->flags) & HS_HIT_EVENTS) != 0U) {
        c->flags |= HC_HIT_EVENTS;
    }
    list_insert(w, index, 0U, c->body_a);
    list_insert(w, index, 1U, c->body_b);
    if (body_awake(body_at(w, c->body_a)) || body_awake(body_at(w, c->body_b))) {
        r = awake_list_add(w, index);
    }
    return r;
}

// contact.c:181-237: a mesh pair's children go first, each out of the
// graph and back to the free list, then the pair's own record.
static void contact_destroy(NphPhysicsWorld *w, u32 contact, int events)
{
    HContact *c = contact_at(w, contact);
    for (u32 k = c->child; k != NPH_NULL32;) {
        HContact *child = contact_at(w, k);
        const u32 next = child->child;
        if (child->color != NPH_NULL32) {
            graph_remove(w, k);
        }
        child->shape_a = NPH_NULL32;
        child->shape_b = NPH_NULL32;
        child->flags = 0U;
        child->parent = NPH_NULL32;
        child->child = NPH_NULL32;
        (void)vbuf_push(w->alloc, &w->free_contacts, 4U, &k); // reserved when made
        k = next;
    }
    c->child = NPH_NULL32;
    if ((c->flags & HC_TOUCHING) != 0U && events && (c->flags & HC_EVENTS) != 0U) {
        NphPhysicsContactEndEvent e;
        e.shape_a.value = shape_at(w, c->shape_a)->handle;
        e.shape_b.value = shape_at(w, c->shape_b)->handle;
        (void)vbuf_push(w->alloc, &w->end_events, sizeof e, &e);
    }
    if (c->color != NPH_NULL32) {
        graph_remove(w, contact);
    }
    if ((c->flags & HC_ISLAND) != 0U) {
        islands_unlink_contact(w, contact);
    }
    awake_list_remove(w, contact);
    list_remove(w, contact, 0U, c->body_a);
    list_remove(w, contact, 1U, c->body_b);
    map_remove(&w->pairs, pair_key(c->shape_a, c->shape_b));
    pairs_touch_key(w, pair_key(c->shape_a, c->shape_b));
    c->shape_a = NPH_NULL32;
    c->shape_b = NPH_NULL32;
    c->flags = 0U;
    (void)vbuf_push(w->alloc, &w->free_contacts, 4U, &contact); // reserved when made
}

void contacts_destroy(NphPhysicsWorld *w, u32 contact, int events)
{
    contact_destroy(w, contact, events);
}

// contact.c:265-281: the body's contacts, then its joints.
NphResult contacts_sync_body(NphPhysicsWorld *w, u32 body)
{
    for (u32 key = body_at(w, body)->contact_head; key != NPH_NULL32;) {
        const u32 c = key >> 1U;
        const u32 next = contact_at(w, c)->next[key & 1U];
        NphResult r = contact_sync(w, c);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        key = next;
    }
    for (u32 key = body_at(w, body)->joint_head; key != NPH_NULL32;) {
        const u32 j = key >> 1U;
        const u32 next = VB(w->joints, HJoint)[j].next[key & 1U];
        NphResult r = joints_sync(w, j);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        key = next;
    }
    return NPH_RESULT_OK;
}

HContact *contacts_at(NphPhysicsWorld *w, u32 contact) { return contact_at(w, contact); }

// contact.c:591-655
static NphResult apply_change(NphPhysicsWorld *w, u32 contact, u32 change)
{
    HContact *c = contact_at(w, contact);
    if ((change & NPH_CHANGE_DISJOINT) != 0U) {
        contact_destroy(w, contact, 1);
        return NPH_RESULT_OK;
    }
    // A mesh pair's children touch as their manifolds say; their graph
    // membership follows in contact_sync.
    {
        u32 i = 1U;
        for (u32 k = c->child; k != NPH_NULL32; k = contact_at(w, k)->child, ++i) {
            const u32 bits = (change >> (NPH_CHANGE_CHILD_SHIFT + 2U * (i - 1U))) & 3U;
            HContact *child = contact_at(w, k);
            if ((bits & NPH_CHANGE_STARTED) != 0U) {
                child->flags |= HC_TOUCHING;
            } else if ((bits & NPH_CHANGE_STOPPED) != 0U) {
                child->flags &= ~(u32)HC_TOUCHING;
            }
        }
    }
    if ((change & NPH_CHANGE_STARTED) != 0U) {
        c->flags |= HC_TOUCHING;
        if ((c->flags & HC_EVENTS) != 0U) {
            NphPhysicsContactBeginEvent e;
            e.shape_a.value = shape_at(w, c->shape_a)->handle;
            e.shape_b.value = shape_at(w, c->shape_b)->handle;
            NphResult r = vbuf_push(w->alloc, &w->begin_events, sizeof e, &e);
            if (r != NPH_RESULT_OK) {
                return r;
            }
        }
        // A sleeping dynamic body touched by an awake one wakes with its
        // island; the two islands then link.
        const u32 ends[2] = {c->body_a, c->body_b};
        for (u32 k = 0U; k < 2U; ++k) {
            const HBody *b = body_at(w, ends[k]);
            if (b->type == NPH_PHYSICS_BODY_DYNAMIC && !body_awake(b)) {
                NphResult r = world_wake_body(w, ends[k]);
                if (r != NPH_RESULT_OK) {
                    return r;
                }
            }
        }
        if (body_at(w, c->body_a)->type == NPH_PHYSICS_BODY_DYNAMIC &&
            body_at(w, c->body_b)->type == NPH_PHYSICS_BODY_DYNAMIC) {
            islands_link_contact(w, contact);
        }
        return contact_sync(w, contact);
    }
    if ((change & NPH_CHANGE_STOPPED) != 0U) {
        c->flags &= ~(u32)HC_TOUCHING;
        if ((c->flags & HC_EVENTS) != 0U) {
            NphPhysicsContactEndEvent e;
            e.shape_a.value = shape_at(w, c->shape_a)->handle;
            e.shape_b.value = shape_at(w, c->shape_b)->handle;
            NphResult r = vbuf_push(w->alloc, &w->end_events, sizeof e, &e);
            if (r != NPH_RESULT_OK) {
                return r;
            }
        }
        if ((c->flags & HC_ISLAND) != 0U) {
            islands_unlink_contact(w, contact);
        }
        return contact_sync(w, contact);
    }
    if ((change & NPH_CHANGE_CHILDREN) != 0U) {
        return contact_sync(w, contact);
    }
    return NPH_RESULT_OK;
}

// The narrow phase of the contacts that exist when the step starts, queued
// on the device before the host looks for new pairs: a contact's manifold
// depends only on its own pair's state, and new contacts join the end of the
// awake list, so the two parts together are the reference's one pass over
// the list (contact.c:657-690).
NphResult contacts_collide_begin(NphPhysicsWorld *w)
{
    NphResult r = nph_dev_init_contacts(w->dev, VB(w->inits, NphContactInit), w->inits.len);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    w->inits.len = 0U;
    // The records as they stand, put back if the step is abandoned before
    // the changes this narrow phase finds reach the host.
    r = nph_dev_contacts_checkpoint(w->dev, w->contacts.len);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    return nph_dev_collide_begin(w->dev, VB(w->awake_contacts, u32), w->awake_contacts.len,
                                 w->step_dt);
}

// Room for the changes of every awake contact, taken before the narrow
// phase's second half so that a lack of memory abandons the step before
// any change is applied: the change lists, an event each, the awake list
// and each colour's list grown by every contact at most, and the solver's
// lists.
NphResult contacts_reserve_changes(NphPhysicsWorld *w)
{
    const u32 n = w->awake_contacts.len > 0U ? w->awake_contacts.len : 1U;
    NphResult r = vbuf_reserve(w->alloc, &w->changed, 4U, n);
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(w->alloc, &w->codes, 4U, n);
    }
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(w->alloc, &w->begin_events, sizeof(NphPhysicsContactBeginEvent),
                         w->begin_events.len + n);
    }
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(w->alloc, &w->end_events, sizeof(NphPhysicsContactEndEvent),
                         w->end_events.len + n);
    }
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(w->alloc, &w->awake_contacts, 4U, w->contacts.len);
    }
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS && r == NPH_RESULT_OK; ++c) {
        r = vbuf_reserve(w->alloc, &w->color_contacts[c], 4U, w->color_contacts[c].len + n);
    }
    // The solver's lists: every contact and joint at most.
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(w->alloc, &w->graph, 4U, w->contacts.len > 0U ? w->contacts.len : 1U);
    }
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(w->alloc, &w->step_joints, 4U, w->joints.len > 0U ? w->joints.len : 1U);
    }
    return r;
}

// Room for `pairs` new contacts.
NphResult contacts_reserve_new(NphPhysicsWorld *w, u32 pairs)
{
    const u32 records = NPH_PHYSICS_MESH_MAX_MANIFOLDS * pairs;
    const u32 most = w->contacts.len + records;
    NphResult r = vbuf_reserve(w->alloc, &w->contacts, sizeof(HContact), most);
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(w->alloc, &w->free_contacts, 4U, most);
    }
    if (r == NPH_RESULT_OK) {
        r = nph_dev_reserve(w->dev, 0U, 0U, most);
    }
    if (r == NPH_RESULT_OK) {
        r = map_reserve(w->alloc, &w->pairs, pairs);
    }
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(w->alloc, &w->awake_contacts, 4U, most);
    }
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(w->alloc, &w->inits, sizeof(NphContactInit), w->inits.len + records);
    }
    return r;
}

NphResult contacts_collide_end(NphPhysicsWorld *w)
{
    NphResult r = nph_dev_init_contacts(w->dev, VB(w->inits, NphContactInit), w->inits.len);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    w->inits.len = 0U;
    const u32 n = w->awake_contacts.len;
    const u32 *codes = NULL;
    r = nph_dev_collide_end(w->dev, VB(w->awake_contacts, u32), n, w->step_dt, &codes);
    if (r != NPH_RESULT_OK || n == 0U) {
        return r;
    }
    return contacts_apply_codes(w, codes, n, NULL, 0U);
}

NphResult contacts_apply_codes(NphPhysicsWorld *w, const u32 *codes_a, u32 na, const u32 *codes_b,
                               u32 nb)
{
    const u32 n = na + nb;
    if (n == 0U) {
        return NPH_RESULT_OK;
    }
    NphResult r;
    if ((r = vbuf_reserve(w->alloc, &w->changed, 4U, n)) != NPH_RESULT_OK ||
        (r = vbuf_reserve(w->alloc, &w->codes, 4U, n)) != NPH_RESULT_OK) {
        return r;
    }
    // The changed records in list order, taken before any change edits the
    // list (contact.c:677-687).
    const u32 *list = VB(w->awake_contacts, u32);
    u32 *changed = VB(w->changed, u32);
    u32 *changed_codes = VB(w->codes, u32);
    u32 m = 0U;
    for (u32 i = 0U; i < n; ++i) {
        const u32 code = i < na ? codes_a[i] : codes_b[i - na];
        if (code != 0U) {
            changed[m] = list[i];
            changed_codes[m] = code;
            m += 1U;
        }
    }
    for (u32 i = 0U; i < m; ++i) {
        r = apply_change(w, changed[i], changed_codes[i]);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    return NPH_RESULT_OK;
}

// ---------------------------------------------------------------------------
// New pairs (contact.c:691-854): the queries run on the device
// (pairs_hip.hip); the host keeps the trees, tells the device what changed,
// and makes the contacts of the pairs that come back, in their order.
// ---------------------------------------------------------------------------
void pairs_mark_shape(NphPhysicsWorld *w, u32 shape)
{
    HShape *s = shape_at(w, shape);
    if ((s->flags & HS_PAIR_DIRTY) != 0U) {
        return;
    }
    if (vbuf_push(w->alloc, &w->pf_dirty_shapes, 4U, &shape) == NPH_RESULT_OK) {
        s->flags |= HS_PAIR_DIRTY;
    } else {
        w->pf_lost = 1; // every shape goes again
    }
}

void pairs_mark_body(NphPhysicsWorld *w, u32 body)
{
    for (u32 s = body_at(w, body)->shape_head; s != NPH_NULL32; s = shape_at(w, s)->next) {
        pairs_mark_shape(w, s);
    }
}

void pairs_touch_key(NphPhysicsWorld *w, u64 key)
{
    if (vbuf_push(w->alloc, &w->pf_touched, 8U, &key) != NPH_RESULT_OK) {
        w->pf_lost = 1; // the set starts over
    }
}

// contact.c:703-746's shape data, as the device's filters read it.
static NphPairShape pair_shape_of(const NphPhysicsWorld *w, u32 slot)
{
    NphPairShape p;
    memset(&p, 0, sizeof p);
    const HShape *s = &VB(w->shapes, HShape)[slot];
    if (s->handle == 0U) {
        return p;
    }
    const HBody *b = &VB(w->bodies, HBody)[s->body];
    p.category = s->filter.category;
    p.mask = s->filter.mask;
    p.group = s->filter.group;
    p.body = s->body;
    p.flags = ((s->flags & (HS_SENSOR | HS_GHOST)) != 0U ? NPH_PAIR_UNTOUCHABLE : 0U) |
              (b->type == NPH_PHYSICS_BODY_DYNAMIC ? NPH_PAIR_DYNAMIC : 0U) |
              (b->joint_count != 0U ? NPH_PAIR_JOINTED : 0U);
    return p;
}

// Whether the set holds the key now: a shape pair with a contact, or a
// body pair a joint keeps apart (joint.c:302-318).
static int key_in_set(const NphPhysicsWorld *w, u64 key)
{
    if ((key & NPH_PAIR_BODY_KEY) == 0U) {
        return map_find(&w->pairs, key) != NPH_NULL32;
    }
    const u64 k = (key & ~NPH_PAIR_BODY_KEY) - 1ULL;
    const u32 a = (u32)(k >> 32U);
    const u32 b = (u32)k;
    if (a >= w->bodies.len || b >= w->bodies.len) {
        return 0;
    }
    const HBody *ba = &VB(w->bodies, HBody)[a];
    const HBody *bb = &VB(w->bodies, HBody)[b];
    return ba->joint_count != 0U && bb->joint_count != 0U && !joints_allow_contact(w, a, b);
}

static int u64_order(const void *x, const void *y)
{
    const u64 a = *(const u64 *)x;
    const u64 b = *(const u64 *)y;
    return a < b ? -1 : (a > b ? 1 : 0);
}

// The keys the set holds now, for a set that starts over.
static NphResult all_set_keys(NphPhysicsWorld *w)
{
    w->pf_keys.len = 0U;
    for (u32 i = 0U; i < w->pairs.cap; ++i) {
        const u64 k = w->pairs.keys[i];
        if (k != MAP_EMPTY && k != MAP_GONE) {
            NphResult r = vbuf_push(w->alloc, &w->pf_keys, 8U, &k);
            if (r != NPH_RESULT_OK) {
                return r;
            }
        }
    }
    const u32 shape_keys = w->pf_keys.len;
    for (u32 j = 0U; j < w->joints.len; ++j) {
        const HJoint *jt = &VB(w->joints, HJoint)[j];
        if (jt->handle == 0U || (jt->flags & NPH_JOINT_COLLIDE_CONNECTED) != 0U) {
            continue;
        }
        const u64 k = nph_pair_key(jt->body_a, jt->body_b) | NPH_PAIR_BODY_KEY;
        NphResult r = vbuf_push(w->alloc, &w->pf_keys, 8U, &k);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    // Two joints may keep one pair apart: each key once.
    u64 *keys = VB(w->pf_keys, u64);
    const u32 n = w->pf_keys.len - shape_keys;
    if (n > 1U) {
        qsort(keys + shape_keys, n, 8U, u64_order);
        u32 m = 1U;
        for (u32 i = 1U; i < n; ++i) {
            if (keys[shape_keys + i] != keys[shape_keys + m - 1U]) {
                keys[shape_keys + m++] = keys[shape_keys + i];
            }
        }
        w->pf_keys.len = shape_keys + m;
    }
    return NPH_RESULT_OK;
}

static u32 pow2_at_least(u32 n)
{
    u32 c = 1024U;
    while (c < n) {
        c *= 2U;
    }
    return c;
}

// The trees' rebuild (contact.c:811-818), and the stats of the move buffer.
static NphResult pairs_begin(NphPhysicsWorld *w)
{
    const f64 t0 = world_now_ms();
    for (u32 t = NPH_TREE_KINEMATIC; t <= NPH_TREE_DYNAMIC; ++t) {
        NphResult r = tree_rebuild(&w->trees[t], 0);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    w->rebuild_ms = world_now_ms() - t0;
    w->stats.moved_proxies = w->move_buffer.len;
    w->stats.new_pairs = 0U;
    return NPH_RESULT_OK;
}

NphResult pairs_prepare(NphPhysicsWorld *w, NphPairInput *in)
{
    NphResult r = pairs_begin(w);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    memset(in, 0, sizeof *in);
    in->dev = w->dev;
    for (u32 t = 0U; t < 3U; ++t) {
        const NphTree *tr = &w->trees[t];
        in->tree_nodes[t] = tr->nodes;
        in->tree_dirty[t] = tr->dirty;
        in->tree_dirty_count[t] = tr->dirty_count;
        in->tree_root[t] = tr->root;
        in->tree_capacity[t] = tr->node_capacity;
        in->tree_height[t] = tree_height(tr);
    }
    // Shapes whose filter data changed (every live one after a lost mark).
    if (w->pf_lost) {
        for (u32 s = 0U; s < w->shapes.len; ++s) {
            if (shape_at(w, s)->handle != 0U) {
                pairs_mark_shape(w, s);
            }
        }
    }
    const u32 ns = w->pf_dirty_shapes.len;
    if ((r = vbuf_reserve(w->alloc, &w->pf_shape_info, sizeof(NphPairShape), ns + 1U)) !=
        NPH_RESULT_OK) {
        return r;
    }
    for (u32 i = 0U; i < ns; ++i) {
        const u32 slot = VB(w->pf_dirty_shapes, u32)[i];
        VB(w->pf_shape_info, NphPairShape)[i] = pair_shape_of(w, slot);
    }
    in->shape_slots = VB(w->pf_dirty_shapes, u32);
    in->shape_info = VB(w->pf_shape_info, NphPairShape);
    in->shape_count = ns;
    in->shape_capacity = w->shapes.len;
    // The set's keys that changed, each once, and whether each is in; or,
    // when the set would pass half full, every key it holds, in a new set.
    u64 *touched = VB(w->pf_touched, u64);
    u32 nt = w->pf_touched.len;
    if (nt > 1U) {
        qsort(touched, nt, 8U, u64_order);
        u32 m = 1U;
        for (u32 i = 1U; i < nt; ++i) {
            if (touched[i] != touched[m - 1U]) {
                touched[m++] = touched[i];
            }
        }
        nt = m;
    }
    u32 adds = 0U;
    for (u32 i = 0U; i < nt; ++i) {
        adds += key_in_set(w, touched[i]) ? 1U : 0U;
    }
    if (w->pf_lost || w->pf_set_cap == 0U || 2ULL * ((u64)w->pf_set_used + adds) > w->pf_set_cap) {
        if ((r = all_set_keys(w)) != NPH_RESULT_OK) {
            return r;
        }
        w->pf_set_cap = pow2_at_least(4U * w->pf_keys.len);
        w->pf_set_used = w->pf_keys.len;
        if ((r = vbuf_reserve(w->alloc, &w->pf_key_in, 1U, w->pf_keys.len + 1U)) != NPH_RESULT_OK) {
            return r;
        }
        memset(w->pf_key_in.data, 1, (size_t)w->pf_keys.len + 1U);
        in->set_keys = VB(w->pf_keys, u64);
        in->set_count = w->pf_keys.len;
        in->set_reset = 1U;
    } else {
        if ((r = vbuf_reserve(w->alloc, &w->pf_key_in, 1U, nt + 1U)) != NPH_RESULT_OK) {
            return r;
        }
        for (u32 i = 0U; i < nt; ++i) {
            VB(w->pf_key_in, u8)[i] = key_in_set(w, touched[i]) ? 1U : 0U;
        }
        w->pf_set_used += adds;
        in->set_keys = touched;
        in->set_count = nt;
    }
    in->set_in = VB(w->pf_key_in, u8);
    in->set_capacity = w->pf_set_cap;
    // The move buffer and its shapes.
    const u32 n = w->move_buffer.len;
    if ((r = vbuf_reserve(w->alloc, &w->pf_moved_shapes, 4U, n + 1U)) != NPH_RESULT_OK) {
        return r;
    }
    for (u32 i = 0U; i < n; ++i) {
        const u32 key = VB(w->move_buffer, u32)[i];
        VB(w->pf_moved_shapes, u32)[i] = (u32)tree_user_data(&w->trees[key & 3U], key >> 2U);
    }
    in->moved_keys = VB(w->move_buffer, u32);
    in->moved_shapes = VB(w->pf_moved_shapes, u32);
    in->moved_count = n;
    return NPH_RESULT_OK;
}

// The device holds what the input sent: the change lists start over.
void pairs_sent(NphPhysicsWorld *w)
{
    for (u32 t = 0U; t < 3U; ++t) {
        tree_dirty_clear(&w->trees[t]);
    }
    for (u32 i = 0U; i < w->pf_dirty_shapes.len; ++i) {
        shape_at(w, VB(w->pf_dirty_shapes, u32)[i])->flags &= ~(u32)HS_PAIR_DIRTY;
    }
    w->pf_dirty_shapes.len = 0U;
    w->pf_touched.len = 0U;
    w->pf_lost = 0;
}

// contact.c:844-853: the move buffer emptied, then a contact for each new
// pair in order.
NphResult pairs_finish(NphPhysicsWorld *w, const u32 *pairs, u32 count)
{
    // Room for every pair's contact (a mesh pair takes four records) before
    // the move buffer is let go: a lack of memory then abandons the step
    // with the search still to do.
    NphResult room = contacts_reserve_new(w, count);
    if (room != NPH_RESULT_OK) {
        return room;
    }
    for (u32 i = 0U; i < w->move_buffer.len; ++i) {
        shape_at(w, VB(w->pf_moved_shapes, u32)[i])->flags &= ~(u32)HS_IN_MOVE_BUFFER;
    }
    w->move_buffer.len = 0U;
    for (u32 i = 0U; i < count; ++i) {
        const u32 a = pairs[2U * i];
        const u32 b = pairs[2U * i + 1U];
        if (map_find(&w->pairs, pair_key(a, b)) == NPH_NULL32) {
            NphResult r = contact_create(w, a, b);
            if (r != NPH_RESULT_OK) {
                return r;
            }
            w->stats.new_pairs += 1U;
        }
    }
    return NPH_RESULT_OK;
}

NphResult contacts_update_pairs(NphPhysicsWorld *w)
{
    if (w->move_buffer.len == 0U) {
        return pairs_begin(w);
    }
    NphPairInput in;
    NphResult r = pairs_prepare(w, &in);
    if (r == NPH_RESULT_OK && w->pair_search == NULL) {
        r = nph_pairs_create(&w->pair_search);
    }
    const u32 *pairs = NULL;
    const u32 *starts = NULL;
    if (r == NPH_RESULT_OK) {
        r = nph_pairs_find(w->pair_search, &in, 1U, &pairs, &starts);
    }
    if (r != NPH_RESULT_OK) {
        return r;
    }
    pairs_sent(w);
    w->pairs_device_ms = nph_pairs_last_ms(w->pair_search);
    return pairs_finish(w, pairs + 2U * starts[0], starts[1] - starts[0]);
}

NphResult move_buffer_push(NphPhysicsWorld *w, u32 shape)
{
    HShape *s = shape_at(w, shape);
    if ((s->flags & HS_IN_MOVE_BUFFER) != 0U) {
        return NPH_RESULT_OK;
    }
    const u32 key = (s->proxy << 2U) | s->tree;
    NphResult r = vbuf_push(w->alloc, &w->move_buffer, 4U, &key);
    if (r == NPH_RESULT_OK) {
        s->flags |= HS_IN_MOVE_BUFFER;
    }
    return r;
}

// solver.c:868-885: grown proxies into the trees and the move buffer, in the
// awake list's order (the device reports them in that order).
NphResult contacts_post_solve(NphPhysicsWorld *w, const NphGrown *grown, u32 count)
{
    for (u32 i = 0U; i < count; ++i) {
        HShape *s = shape_at(w, grown[i].shape);
        s->fat_aabb = grown[i].fat_aabb;
        tree_enlarge_proxy(&w->trees[s->tree], s->proxy, s->fat_aabb);
        NphResult r = move_buffer_push(w, grown[i].shape);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    return NPH_RESULT_OK;
}
