// SPDX-License-Identifier: Apache-2.0
// Pair finding on the device (engine/physics/src/contact.c:703-854): every
// moved proxy's query over the three AABB trees, filtered as the reference
// filters, in one launch for one world or for a batch of worlds.
//
// The trees are the host's (host/tree.c builds them node for node as the
// reference does, and their shape fixes the order a query meets its
// leaves); each world's device copy takes the nodes the host wrote since
// the world's last search.  A query is the reference's depth-first walk,
// the first child first, one thread per moved proxy: first counted, then,
// at the places an exclusive scan of the counts gives, filled.  So the
// pairs come out in the move buffer's order and each proxy's in its trees'
// order (static, kinematic, dynamic for a dynamic proxy; the dynamic tree
// alone otherwise), with no sort.
//
// The filters (contact.c:703-746) read each shape's NphPairShape (body,
// category, mask, group, sensor/ghost, body dynamic, body jointed), a
// per-shape stamp set for the proxies moved in this search (the lower key
// of two moved proxies reports their pair), and a hash set of keys: shape
// pairs that already have a contact, and body pairs a joint keeps apart.
// The set is open addressing with linear probing; the host keeps it at
// most half full (counting removed slots) and starts it over, larger, when
// it would pass that.
#include "device_impl.h"

#include <time.h>

static f64 pf_now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (f64)ts.tv_sec * 1e3 + (f64)ts.tv_nsec * 1e-6;
}

#define PF_BLOCK 256U
#define PF_QBLOCK 64U // a query's walk is long and branchy: small blocks spread it over the CUs
#define PF_SMALL 64U  // moved proxies a one-block search takes
#define PF_SCAN 1024U
#define PF_STACK 64U        // depth-first stack of a tree of height up to 62
#define PF_STACK_DEEP 1024U // aabb_tree.h:50, for taller trees
#define PF_EMPTY 0ULL
#define PF_GONE (~0ULL)

// A world's device state for one search.
struct PfWorld {
    NphTreeNode *nodes[3];
    u32 root[3];
    u32 search; // this search's stamp
    NphPairShape *shapes;
    u32 *stamp;
    u64 *set;
    u32 set_mask;
    u32 pad;
};

// Upload records.
struct PfNode {
    u32 world;
    u32 tree;
    u32 index;
    u32 pad;
    NphTreeNode node;
};
struct PfShape {
    u32 world;
    u32 slot;
    u32 pad[2];
    NphPairShape info;
};
struct PfKey {
    u32 world;
    u32 in;
    u64 key;
};
struct PfMoved {
    u32 world;
    u32 key;
    u32 shape;
    u32 pad;
};

static_assert(sizeof(PfNode) == 64U && sizeof(PfShape) == 48U && sizeof(PfKey) == 16U &&
                  sizeof(PfMoved) == 16U,
              "upload record sizes");

struct NphPairSearch {
    hipStream_t stream;
    hipEvent_t ev[3]; // the search's start and end, the upload's start
    f64 last_ms;
    f64 last_gather_ms; // host: the records packed for the upload
    f64 last_upload_ms; // device: the upload's copy
    u64 last_upload_bytes;
    u8 *h_up; // pinned upload
    u64 cap_h_up;
    u8 *d_up;
    u64 cap_d_up;
    u32 *d_counts; // per moved proxy
    u32 *d_offsets;
    u32 *d_chunks; // per PF_SCAN chunk, then the total
    u32 cap_moved;
    u32 *d_pairs; // 2 words a pair
    u32 cap_pairs;
    u32 *d_starts; // per world + 1
    u32 cap_starts;
    u32 *h_starts; // pinned
    u64 cap_h_starts; // bytes
    u32 *h_pairs; // pinned
    u64 cap_h_pairs; // bytes
    u64 *at; // per world: its first node, shape, key and moved record
    u32 cap_at;
};

// ---------------------------------------------------------------------------
// The filter set
// ---------------------------------------------------------------------------
__device__ static inline u32 pf_slot(u64 key, u32 mask)
{
    u64 x = key; // a 64-bit finalizer (Stafford's mix 13)
    x ^= x >> 30U;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27U;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 31U;
    return (u32)x & mask;
}

__device__ static inline int pf_set_has(const PfWorld &w, u64 key)
{
    u32 i = pf_slot(key, w.set_mask);
    for (u32 n = 0U; n <= w.set_mask; ++n) {
        const u64 k = w.set[i];
        if (k == key) {
            return 1;
        }
        if (k == PF_EMPTY) {
            return 0;
        }
        i = (i + 1U) & w.set_mask;
    }
    return 0;
}

// Keys are distinct within a search: a removal needs no atomics; an
// insertion takes the first empty slot of its probe sequence that no other
// insertion took first.
__device__ static inline void pf_set_remove(const PfWorld &w, u64 key)
{
    u32 i = pf_slot(key, w.set_mask);
    for (u32 n = 0U; n <= w.set_mask; ++n) {
        const u64 k = w.set[i];
        if (k == key) {
            w.set[i] = PF_GONE;
            return;
        }
        if (k == PF_EMPTY) {
            return;
        }
        i = (i + 1U) & w.set_mask;
    }
}

__device__ static inline void pf_set_insert(const PfWorld &w, u64 key)
{
    u32 i = pf_slot(key, w.set_mask);
    for (u32 n = 0U; n <= w.set_mask; ++n) {
        const u64 k = w.set[i];
        if (k == key) {
            return;
        }
        if (k == PF_EMPTY) {
            const u64 prev = atomicCAS((unsigned long long *)&w.set[i], (unsigned long long)PF_EMPTY,
                                       (unsigned long long)key);
            if (prev == PF_EMPTY || prev == key) {
                return;
            }
        }
        i = (i + 1U) & w.set_mask;
    }
}

// ---------------------------------------------------------------------------
// The query
// ---------------------------------------------------------------------------

// contact.c:703-746: whether the moved shape `m` and the leaf's shape make
// a new pair.
__device__ static inline int pf_accept(const PfWorld &w, const PfMoved &m, const NphPairShape &me,
                                       u32 other, u32 other_key)
{
    if (other == m.shape) {
        return 0;
    }
    if (w.stamp[other] == w.search && other_key < m.key) {
        return 0; // both moved: the lower key reports the pair
    }
    const NphPairShape o = w.shapes[other];
    if (o.body == me.body || ((me.flags | o.flags) & NPH_PAIR_UNTOUCHABLE) != 0U ||
        ((me.flags | o.flags) & NPH_PAIR_DYNAMIC) == 0U) {
        return 0;
    }
    if (me.group == o.group && me.group != 0) {
        if (me.group < 0) {
            return 0;
        }
    } else if ((me.category & o.mask) == 0U || (o.category & me.mask) == 0U) {
        return 0;
    }
    if ((me.flags & o.flags & NPH_PAIR_JOINTED) != 0U &&
        pf_set_has(w, nph_pair_key(me.body, o.body) | NPH_PAIR_BODY_KEY)) {
        return 0;
    }
    return !pf_set_has(w, nph_pair_key(m.shape, other));
}

// The moved proxy's new pairs counted, and written from `out` when FILL.
template <u32 STACK, bool FILL>
__device__ static u32 pf_query(const PfWorld &w, const PfMoved &m, u32 *out)
{
    const u32 tree = m.key & 3U;
    const NphAABB box = w.nodes[tree][m.key >> 2U].aabb;
    const NphPairShape me = w.shapes[m.shape];
    const u64 mask = me.group > 0 ? ~0ULL : me.mask;
    u32 count = 0U;
    u32 stack[STACK];
    for (u32 q = tree == NPH_TREE_DYNAMIC ? NPH_TREE_STATIC : NPH_TREE_DYNAMIC; q < 3U; ++q) {
        if (w.root[q] == NPH_NULL32) {
            continue;
        }
        const NphTreeNode *nodes = w.nodes[q];
        u32 top = 0U;
        stack[top++] = w.root[q];
        while (top > 0U) {
            const u32 i = stack[--top];
            const NphTreeNode *n = &nodes[i];
            if ((n->category_bits & mask) == 0U || !aabb_overlaps(n->aabb, box)) {
                continue;
            }
            if (n->height == 0U) {
                if (pf_accept(w, m, me, n->child1, (i << 2U) | q)) {
                    if (FILL) {
                        out[2U * count] = m.shape;
                        out[2U * count + 1U] = n->child1;
                    }
                    count += 1U;
                }
            } else if (top + 2U <= STACK) {
                stack[top++] = n->child2;
                stack[top++] = n->child1;
            }
        }
    }
    return count;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
struct PfArgs {
    const PfWorld *worlds;
    const PfNode *nodes;
    const PfShape *shapes;
    const PfKey *keys;
    const PfMoved *moved;
    const u32 *first; // per world + 1: its first moved proxy
    u32 node_count, shape_count, key_count, moved_count, world_count;
    u32 *counts;
    u32 *offsets;
    u32 *chunks;
    u32 *pairs;
    u32 pair_cap;
    u32 *starts;
};

// Node and shape writes, set removals, this search's stamps.
__global__ void __launch_bounds__(PF_BLOCK) k_pf_apply(PfArgs a)
{
    u32 t = blockIdx.x * PF_BLOCK + threadIdx.x;
    if (t < a.node_count) {
        const PfNode &u = a.nodes[t];
        a.worlds[u.world].nodes[u.tree][u.index] = u.node;
        return;
    }
    t -= a.node_count;
    if (t < a.shape_count) {
        const PfShape &u = a.shapes[t];
        a.worlds[u.world].shapes[u.slot] = u.info;
        return;
    }
    t -= a.shape_count;
    if (t < a.key_count) {
        const PfKey &u = a.keys[t];
        if (u.in == 0U) {
            pf_set_remove(a.worlds[u.world], u.key);
        }
        return;
    }
    t -= a.key_count;
    if (t < a.moved_count) {
        const PfMoved &m = a.moved[t];
        const PfWorld &w = a.worlds[m.world];
        w.stamp[m.shape] = w.search;
    }
}

__global__ void __launch_bounds__(PF_BLOCK) k_pf_insert(PfArgs a)
{
    const u32 t = blockIdx.x * PF_BLOCK + threadIdx.x;
    if (t < a.key_count && a.keys[t].in != 0U) {
        pf_set_insert(a.worlds[a.keys[t].world], a.keys[t].key);
    }
}

template <u32 STACK>
__global__ void __launch_bounds__(PF_QBLOCK) k_pf_count(PfArgs a)
{
    const u32 t = blockIdx.x * PF_QBLOCK + threadIdx.x;
    if (t < a.moved_count) {
        const PfMoved m = a.moved[t];
        a.counts[t] = pf_query<STACK, false>(a.worlds[m.world], m, NULL);
    }
}

// Exclusive scan over a block of PF_SCAN threads (one value each).
__device__ static u32 pf_block_scan(u32 v, u32 *total)
{
    __shared__ u32 wave_sum[PF_SCAN / 64U];
    const u32 lane = threadIdx.x & 63U;
    const u32 wave = threadIdx.x >> 6U;
    u32 x = v;
    for (u32 off = 1U; off < 64U; off <<= 1U) {
        const u32 y = (u32)__shfl_up((int)x, off, 64);
        x += lane >= off ? y : 0U;
    }
    if (lane == 63U) {
        wave_sum[wave] = x;
    }
    __syncthreads();
    if (wave == 0U) {
        u32 s = lane < PF_SCAN / 64U ? wave_sum[lane] : 0U;
        for (u32 off = 1U; off < 64U; off <<= 1U) {
            const u32 y = (u32)__shfl_up((int)s, off, 64);
            s += lane >= off ? y : 0U;
        }
        if (lane < PF_SCAN / 64U) {
            wave_sum[lane] = s; // inclusive over waves
        }
    }
    __syncthreads();
    const u32 before = wave == 0U ? 0U : wave_sum[wave - 1U];
    *total = wave_sum[PF_SCAN / 64U - 1U];
    return before + x - v;
}

// Counts -> offsets within each chunk of PF_SCAN, and the chunks' sums.
__global__ void __launch_bounds__(PF_SCAN) k_pf_scan_chunk(PfArgs a)
{
    const u32 t = blockIdx.x * PF_SCAN + threadIdx.x;
    const u32 v = t < a.moved_count ? a.counts[t] : 0U;
    u32 total = 0U;
    const u32 off = pf_block_scan(v, &total);
    if (t < a.moved_count) {
        a.offsets[t] = off;
    }
    if (threadIdx.x == 0U) {
        a.chunks[blockIdx.x] = total;
    }
}

// The chunks' sums -> their bases (one block, any number of chunks); the
// total after them.
__global__ void __launch_bounds__(PF_SCAN) k_pf_scan_bases(PfArgs a, u32 chunk_count)
{
    u32 carry = 0U;
    for (u32 base = 0U; base < chunk_count; base += PF_SCAN) {
        const u32 c = base + threadIdx.x;
        const u32 v = c < chunk_count ? a.chunks[c] : 0U;
        u32 total = 0U;
        const u32 off = pf_block_scan(v, &total);
        __syncthreads(); // every read of this round's sums before the writes
        if (c < chunk_count) {
            a.chunks[c] = carry + off;
        }
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0U) {
        a.chunks[chunk_count] = carry;
    }
}

// Global offsets, then the pairs where they fit the buffer.
template <u32 STACK>
__global__ void __launch_bounds__(PF_QBLOCK) k_pf_fill(PfArgs a)
{
    const u32 t = blockIdx.x * PF_QBLOCK + threadIdx.x;
    if (t < a.moved_count) {
        const u32 off = a.chunks[t / PF_SCAN] + a.offsets[t];
        a.offsets[t] = off;
        if (off + a.counts[t] <= a.pair_cap) {
            const PfMoved m = a.moved[t];
            (void)pf_query<STACK, true>(a.worlds[m.world], m, a.pairs + 2ULL * off);
        }
    }
}

// The pairs again, every one (after the buffer grew).
template <u32 STACK>
__global__ void __launch_bounds__(PF_QBLOCK) k_pf_refill(PfArgs a)
{
    const u32 t = blockIdx.x * PF_QBLOCK + threadIdx.x;
    if (t < a.moved_count) {
        const PfMoved m = a.moved[t];
        (void)pf_query<STACK, true>(a.worlds[m.world], m, a.pairs + 2ULL * a.offsets[t]);
    }
}

__global__ void __launch_bounds__(PF_BLOCK) k_pf_starts(PfArgs a, u32 chunk_count)
{
    const u32 k = blockIdx.x * PF_BLOCK + threadIdx.x;
    if (k <= a.world_count) {
        const u32 f = a.first[k];
        a.starts[k] = f < a.moved_count ? a.offsets[f] : a.chunks[chunk_count];
    }
}

// Up to PF_SCAN moved proxies in one block: count, scan, fill, starts.
template <u32 STACK>
__global__ void __launch_bounds__(PF_SCAN) k_pf_small(PfArgs a)
{
    const u32 t = threadIdx.x;
    PfMoved m;
    u32 c = 0U;
    if (t < a.moved_count) {
        m = a.moved[t];
        c = pf_query<STACK, false>(a.worlds[m.world], m, NULL);
        a.counts[t] = c;
    }
    u32 total = 0U;
    const u32 off = pf_block_scan(c, &total);
    if (t < a.moved_count) {
        a.offsets[t] = off;
        if (off + c <= a.pair_cap) {
            (void)pf_query<STACK, true>(a.worlds[m.world], m, a.pairs + 2ULL * off);
        }
    }
    if (t == 0U) {
        a.chunks[0] = 0U;
        a.chunks[1] = total;
    }
    __threadfence_block();
    __syncthreads();
    for (u32 k = t; k <= a.world_count; k += PF_SCAN) {
        const u32 f = a.first[k];
        a.starts[k] = f < a.moved_count ? a.offsets[f] : total;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
NphResult nph_pairs_create(NphPairSearch **out)
{
    NphPairSearch *p = (NphPairSearch *)calloc(1, sizeof *p);
    if (p == NULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&p->ev[0]) != hipSuccess || hipEventCreate(&p->ev[1]) != hipSuccess ||
        hipEventCreate(&p->ev[2]) != hipSuccess) {
        free(p);
        return NPH_RESULT_DEVICE_LOST;
    }
    *out = p;
    return NPH_RESULT_OK;
}

void nph_pairs_destroy(NphPairSearch *p)
{
    if (p == NULL) {
        return;
    }
    (void)hipStreamSynchronize(p->stream);
    void *bufs[] = {p->d_up, p->d_counts, p->d_offsets, p->d_chunks, p->d_pairs, p->d_starts};
    for (u32 i = 0U; i < sizeof bufs / sizeof bufs[0]; ++i) {
        if (bufs[i] != NULL) {
            (void)hipFree(bufs[i]);
        }
    }
    free(p->at);
    void *pinned[] = {p->h_up, p->h_starts, p->h_pairs};
    for (u32 i = 0U; i < sizeof pinned / sizeof pinned[0]; ++i) {
        if (pinned[i] != NULL) {
            (void)hipHostFree(pinned[i]);
        }
    }
    for (u32 i = 0U; i < 3U; ++i) {
        (void)hipEventDestroy(p->ev[i]);
    }
    (void)hipStreamDestroy(p->stream);
    free(p);
}

f64 nph_pairs_last_ms(const NphPairSearch *p) { return p->last_ms; }

void nph_pairs_last_detail(const NphPairSearch *p, f64 out[3])
{
    out[0] = p->last_gather_ms;
    out[1] = p->last_upload_ms;
    out[2] = (f64)p->last_upload_bytes;
}

static NphResult pf_pinned(void **p, u64 *cap, u64 bytes)
{
    if (bytes <= *cap) {
        return NPH_RESULT_OK;
    }
    u64 c = *cap < 4096U ? 4096U : *cap;
    while (c < bytes) {
        c *= 2U;
    }
    if (*p != NULL) {
        HIP_TRY(hipHostFree(*p));
        *p = NULL;
        *cap = 0U;
    }
    HIP_TRY(hipHostMalloc(p, c, hipHostMallocDefault));
    *cap = c;
    return NPH_RESULT_OK;
}

static NphResult pf_device(void **p, u64 *cap, u64 bytes, hipStream_t s)
{
    if (bytes <= *cap) {
        return NPH_RESULT_OK;
    }
    u64 c = *cap < 4096U ? 4096U : *cap;
    while (c < bytes) {
        c *= 2U;
    }
    const NphResult r = devi_grow_buffer(p, 0U, c, s, 0);
    if (r == NPH_RESULT_OK) {
        *cap = c;
    }
    return r;
}

// The world's device arrays grown to the host's capacities; a set that
// starts over starts empty.
static NphResult pf_reserve_world(NphDevice *d, const NphPairInput *in, hipStream_t s)
{
    for (u32 t = 0U; t < 3U; ++t) {
        if (in->tree_capacity[t] > d->pf_node_cap[t]) {
            NphResult r = devi_grow_buffer((void **)&d->pf_nodes[t],
                                           (u64)d->pf_node_cap[t] * sizeof(NphTreeNode),
                                           (u64)in->tree_capacity[t] * sizeof(NphTreeNode), s, 1);
            if (r != NPH_RESULT_OK) {
                return r;
            }
            d->pf_node_cap[t] = in->tree_capacity[t];
        }
    }
    if (in->shape_capacity > d->pf_shape_cap) {
        const u32 old = d->pf_shape_cap;
        const u32 cap = devi_grow_cap(old, in->shape_capacity);
        NphResult r = devi_grow_buffer((void **)&d->pf_shapes, (u64)old * sizeof(NphPairShape),
                                       (u64)cap * sizeof(NphPairShape), s, 1);
        if (r == NPH_RESULT_OK) {
            r = devi_grow_buffer((void **)&d->pf_stamp, (u64)old * 4U, (u64)cap * 4U, s, 1);
        }
        if (r != NPH_RESULT_OK) {
            return r;
        }
        HIP_TRY(hipMemsetAsync(d->pf_stamp + old, 0, (u64)(cap - old) * 4U, s));
        d->pf_shape_cap = cap;
    }
    if (in->set_capacity != d->pf_set_cap) {
        NphResult r = devi_grow_buffer((void **)&d->pf_set, 0U, (u64)in->set_capacity * 8U, s, 0);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        d->pf_set_cap = in->set_capacity;
    }
    if (in->set_reset != 0U || in->set_capacity != d->pf_set_cap) {
        HIP_TRY(hipMemsetAsync(d->pf_set, 0, (u64)in->set_capacity * 8U, s));
    }
    return NPH_RESULT_OK;
}

static u64 pf_align(u64 x) { return (x + 63U) & ~63ULL; }

NphResult nph_pairs_find(NphPairSearch *p, const NphPairInput *in, u32 n, const u32 **pairs,
                         const u32 **starts)
{
    const hipStream_t s = p->stream;
    if (n > p->cap_at) {
        const u32 c = devi_grow_cap(p->cap_at, n);
        u64 *at = (u64 *)realloc(p->at, (size_t)c * 4U * 8U);
        if (at == NULL) {
            return NPH_RESULT_OUT_OF_MEMORY;
        }
        p->at = at;
        p->cap_at = c;
    }
    u64 nn = 0U, ns = 0U, nk = 0U, nm = 0U;
    u32 deep = 0U;
    for (u32 k = 0U; k < n; ++k) {
        p->at[4U * k] = nn;
        p->at[4U * k + 1U] = ns;
        p->at[4U * k + 2U] = nk;
        p->at[4U * k + 3U] = nm;
        if (in[k].set_capacity == 0U || (in[k].set_capacity & (in[k].set_capacity - 1U)) != 0U) {
            return NPH_RESULT_INVALID_ARGUMENT;
        }
        NphResult r = pf_reserve_world(in[k].dev, &in[k], s);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        for (u32 t = 0U; t < 3U; ++t) {
            nn += in[k].tree_dirty_count[t];
            deep |= in[k].tree_height[t] + 2U > PF_STACK ? 1U : 0U;
        }
        ns += in[k].shape_count;
        nk += in[k].set_count;
        nm += in[k].moved_count;
    }
    if (nm + nn + ns + nk > 0xFFFFFFFFULL - PF_BLOCK) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    // The upload, packed.
    const u64 o_nodes = 0U;
    const u64 o_shapes = pf_align(o_nodes + nn * sizeof(PfNode));
    const u64 o_keys = pf_align(o_shapes + ns * sizeof(PfShape));
    const u64 o_moved = pf_align(o_keys + nk * sizeof(PfKey));
    const u64 o_worlds = pf_align(o_moved + nm * sizeof(PfMoved));
    const u64 o_first = pf_align(o_worlds + (u64)n * sizeof(PfWorld));
    const u64 bytes = pf_align(o_first + ((u64)n + 1U) * 4U);
    NphResult r = pf_pinned((void **)&p->h_up, &p->cap_h_up, bytes);
    if (r == NPH_RESULT_OK) {
        r = pf_device((void **)&p->d_up, &p->cap_d_up, bytes, s);
    }
    if (r != NPH_RESULT_OK) {
        return r;
    }
    PfNode *un = (PfNode *)(p->h_up + o_nodes);
    PfShape *us = (PfShape *)(p->h_up + o_shapes);
    PfKey *uk = (PfKey *)(p->h_up + o_keys);
    PfMoved *um = (PfMoved *)(p->h_up + o_moved);
    PfWorld *uw = (PfWorld *)(p->h_up + o_worlds);
    u32 *uf = (u32 *)(p->h_up + o_first);
    // Each world's records in parallel (the node gather is the bulk of it).
    const f64 g0 = pf_now_ms();
#pragma omp parallel for schedule(dynamic, 4)
    for (u32 k = 0U; k < n; ++k) {
        u64 at_n = p->at[4U * k], at_s = p->at[4U * k + 1U];
        u64 at_k = p->at[4U * k + 2U], at_m = p->at[4U * k + 3U];
        const NphPairInput *w = &in[k];
        NphDevice *d = w->dev;
        d->pf_search += 1U;
        PfWorld *pw = &uw[k];
        memset(pw, 0, sizeof *pw);
        for (u32 t = 0U; t < 3U; ++t) {
            pw->nodes[t] = d->pf_nodes[t];
            pw->root[t] = w->tree_root[t];
            for (u32 i = 0U; i < w->tree_dirty_count[t]; ++i) {
                PfNode *u = &un[at_n++];
                u->world = k;
                u->tree = t;
                u->index = w->tree_dirty[t][i];
                u->pad = 0U;
                u->node = w->tree_nodes[t][u->index];
            }
        }
        pw->search = d->pf_search;
        pw->shapes = d->pf_shapes;
        pw->stamp = d->pf_stamp;
        pw->set = d->pf_set;
        pw->set_mask = w->set_capacity - 1U;
        for (u32 i = 0U; i < w->shape_count; ++i) {
            PfShape *u = &us[at_s++];
            memset(u, 0, sizeof *u);
            u->world = k;
            u->slot = w->shape_slots[i];
            u->info = w->shape_info[i];
        }
        for (u32 i = 0U; i < w->set_count; ++i) {
            PfKey *u = &uk[at_k++];
            u->world = k;
            u->in = w->set_in[i] != 0U ? 1U : 0U;
            u->key = w->set_keys[i];
        }
        uf[k] = (u32)at_m;
        for (u32 i = 0U; i < w->moved_count; ++i) {
            PfMoved *u = &um[at_m++];
            u->world = k;
            u->key = w->moved_keys[i];
            u->shape = w->moved_shapes[i];
            u->pad = 0U;
        }
    }
    uf[n] = (u32)nm;
    p->last_gather_ms = pf_now_ms() - g0;
    p->last_upload_bytes = bytes;
    // Scratch for the counts and the pairs.
    const u32 moved = (u32)nm;
    const u32 chunk_count = (moved + PF_SCAN - 1U) / PF_SCAN;
    if (moved > p->cap_moved) {
        const u32 c = devi_grow_cap(p->cap_moved, moved);
        r = devi_grow_buffer((void **)&p->d_counts, 0U, (u64)c * 4U, s, 0);
        if (r == NPH_RESULT_OK) {
            r = devi_grow_buffer((void **)&p->d_offsets, 0U, (u64)c * 4U, s, 0);
        }
        if (r == NPH_RESULT_OK) {
            r = devi_grow_buffer((void **)&p->d_chunks, 0U, ((u64)c / PF_SCAN + 2U) * 4U, s, 0);
        }
        if (r != NPH_RESULT_OK) {
            return r;
        }
        p->cap_moved = c;
    }
    if (p->d_chunks == NULL) {
        r = devi_grow_buffer((void **)&p->d_chunks, 0U, 64U * 4U, s, 0);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    if (n + 1U > p->cap_starts) {
        const u32 c = devi_grow_cap(p->cap_starts, n + 1U);
        r = devi_grow_buffer((void **)&p->d_starts, 0U, (u64)c * 4U, s, 0);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        p->cap_starts = c;
    }
    r = pf_pinned((void **)&p->h_starts, &p->cap_h_starts, ((u64)n + 1U) * 4U);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    if (p->d_pairs == NULL) {
        r = devi_grow_buffer((void **)&p->d_pairs, 0U, 2ULL * 4096U * 4U, s, 0);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        p->cap_pairs = 4096U;
    }
    HIP_TRY(hipEventRecord(p->ev[2], s));
    HIP_TRY(hipMemcpyAsync(p->d_up, p->h_up, bytes, hipMemcpyHostToDevice, s));
    PfArgs a;
    a.worlds = (const PfWorld *)(p->d_up + o_worlds);
    a.nodes = (const PfNode *)(p->d_up + o_nodes);
    a.shapes = (const PfShape *)(p->d_up + o_shapes);
    a.keys = (const PfKey *)(p->d_up + o_keys);
    a.moved = (const PfMoved *)(p->d_up + o_moved);
    a.first = (const u32 *)(p->d_up + o_first);
    a.node_count = (u32)nn;
    a.shape_count = (u32)ns;
    a.key_count = (u32)nk;
    a.moved_count = moved;
    a.world_count = n;
    a.counts = p->d_counts;
    a.offsets = p->d_offsets;
    a.chunks = p->d_chunks;
    a.pairs = p->d_pairs;
    a.pair_cap = p->cap_pairs;
    a.starts = p->d_starts;
    HIP_TRY(hipEventRecord(p->ev[0], s));
    const u64 apply = nn + ns + nk + nm;
    if (apply > 0U) {
        hipLaunchKernelGGL(k_pf_apply, dim3(devi_blocks((u32)apply, PF_BLOCK)), dim3(PF_BLOCK), 0, s, a);
        HIP_TRY(hipGetLastError());
    }
    if (nk > 0U) {
        hipLaunchKernelGGL(k_pf_insert, dim3(devi_blocks((u32)nk, PF_BLOCK)), dim3(PF_BLOCK), 0, s, a);
        HIP_TRY(hipGetLastError());
    }
    const u32 blocks = devi_blocks(moved > 0U ? moved : 1U, PF_QBLOCK);
    if (moved <= PF_SMALL) {
        if (deep) {
            hipLaunchKernelGGL(k_pf_small<PF_STACK_DEEP>, dim3(1), dim3(PF_SCAN), 0, s, a);
        } else {
            hipLaunchKernelGGL(k_pf_small<PF_STACK>, dim3(1), dim3(PF_SCAN), 0, s, a);
        }
        HIP_TRY(hipGetLastError());
    } else {
        if (deep) {
            hipLaunchKernelGGL(k_pf_count<PF_STACK_DEEP>, dim3(blocks), dim3(PF_QBLOCK), 0, s, a);
        } else {
            hipLaunchKernelGGL(k_pf_count<PF_STACK>, dim3(blocks), dim3(PF_QBLOCK), 0, s, a);
        }
        hipLaunchKernelGGL(k_pf_scan_chunk, dim3(chunk_count), dim3(PF_SCAN), 0, s, a);
        hipLaunchKernelGGL(k_pf_scan_bases, dim3(1), dim3(PF_SCAN), 0, s, a, chunk_count);
        if (deep) {
            hipLaunchKernelGGL(k_pf_fill<PF_STACK_DEEP>, dim3(blocks), dim3(PF_QBLOCK), 0, s, a);
        } else {
            hipLaunchKernelGGL(k_pf_fill<PF_STACK>, dim3(blocks), dim3(PF_QBLOCK), 0, s, a);
        }
        hipLaunchKernelGGL(k_pf_starts, dim3(devi_blocks(n + 1U, PF_BLOCK)), dim3(PF_BLOCK), 0, s, a,
                           chunk_count);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(p->h_starts, p->d_starts, ((u64)n + 1U) * 4U, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const u32 total = p->h_starts[n];
    if (total > p->cap_pairs) {
        const u32 c = devi_grow_cap(p->cap_pairs, total);
        r = devi_grow_buffer((void **)&p->d_pairs, 0U, 2ULL * c * 4U, s, 0);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        p->cap_pairs = c;
        a.pairs = p->d_pairs;
        a.pair_cap = c;
        if (deep) {
            hipLaunchKernelGGL(k_pf_refill<PF_STACK_DEEP>, dim3(blocks), dim3(PF_QBLOCK), 0, s, a);
        } else {
            hipLaunchKernelGGL(k_pf_refill<PF_STACK>, dim3(blocks), dim3(PF_QBLOCK), 0, s, a);
        }
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(p->ev[1], s));
    if (total > 0U) {
        r = pf_pinned((void **)&p->h_pairs, &p->cap_h_pairs, 2ULL * total * 4U);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        HIP_TRY(hipMemcpyAsync(p->h_pairs, p->d_pairs, 2ULL * total * 4U, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0.0F;
    if (hipEventElapsedTime(&ms, p->ev[0], p->ev[1]) == hipSuccess) {
        p->last_ms = (f64)ms;
    }
    if (hipEventElapsedTime(&ms, p->ev[2], p->ev[0]) == hipSuccess) {
        p->last_upload_ms = (f64)ms;
    }
    *pairs = p->h_pairs;
    *starts = p->h_starts;
    return NPH_RESULT_OK;
}
