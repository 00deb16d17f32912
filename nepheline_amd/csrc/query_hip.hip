// SPDX-License-Identifier: Apache-2.0
// The reference's standalone narrow-phase queries (collide.h:97-121) on the
// device: nph_physics_collide and nph_physics_manifold_reduce run the same
// device code the world step's narrow phase runs (nph_collide.h), one
// thread a pair, and nph_amd_collide_many runs many pairs in one launch.
//
// The queries own a device context of their own (stream, buffers, device
// copies of the hull blocks they have seen), made on first use and guarded
// by a mutex: the reference's functions are free functions, callable from
// any thread.  There is no CPU fallback: without a HIP device the calls
// report no contact and say why on stderr.
//
// nph_amd_collide_many_static is the same query for callers that cannot
// allocate device memory: it stages the pairs, results and hull blocks of a
// call in one fixed device array (`scratch`, below) and calls no allocator.

#include "device_impl.h"
#include "device_narrow.h"

#include <mutex>

// One surface query on one team: a convex shape against one triangle
// (mode 0) or against a mesh (mode 1), the world step's mesh narrow phase
// (nph_surface.h). The shape's place and turn are in the surface's frame.
struct NphSurfaceQuery {
    NphGeom a;
    NphQuat rotation;
    NphDVec3 position;          // mesh: the shape in the mesh's frame (64 bits)
    NphVec3 tri[3];             // triangle: its corners relative to the shape
    u32 active;
    u32 two_sided;
    const NphPhysicsMesh *mesh; // device copy
    f32 speculative;
    u32 mode;
};

namespace {

// A hull block the queries have seen, and its device copy.
struct QueryBlock {
    const void *host;
    const void *device;
};

struct QueryContext {
    std::mutex lock;
    int ready = 0;
    int failed = 0;
    hipStream_t stream = NULL;
    NphQueryPair *pairs = NULL;
    NphManifold *manifolds = NULL;
    NphCache *caches = NULL;
    u32 cap = 0U;
    NphMPoint *points = NULL;
    u32 cap_points = 0U;
    u32 *count = NULL;
    QueryBlock *blocks = NULL;
    u32 block_count = 0U, block_cap = 0U;
    NphSurfaceQuery *surface_in = NULL; // one surface query's input and output
    NphMeshManifolds *surface_out = NULL;
};

QueryContext g_query;

NphResult query_ready(QueryContext *q)
{
    if (q->ready) {
        return NPH_RESULT_OK;
    }
    if (q->failed) {
        return NPH_RESULT_UNSUPPORTED;
    }
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1 ||
        hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&q->count, 4U) != hipSuccess) {
        fprintf(stderr, "nepheline_amd: no HIP device: the collision queries run only on an "
                        "MI355X\n");
        q->failed = 1;
        return NPH_RESULT_UNSUPPORTED;
    }
    q->ready = 1;
    return NPH_RESULT_OK;
}

// The device copy of a hull or mesh block of `bytes` bytes (made once per
// block).
NphResult query_block(QueryContext *q, const void *block, u32 bytes, const void **out)
{
    for (u32 i = 0U; i < q->block_count; ++i) {
        if (q->blocks[i].host == block) {
            *out = q->blocks[i].device;
            return NPH_RESULT_OK;
        }
    }
    if (q->block_count == q->block_cap) {
        const u32 c = q->block_cap == 0U ? 16U : 2U * q->block_cap;
        QueryBlock *n = (QueryBlock *)realloc(q->blocks, (size_t)c * sizeof(QueryBlock));
        if (n == NULL) {
            return NPH_RESULT_OUT_OF_MEMORY;
        }
        q->blocks = n;
        q->block_cap = c;
    }
    void *p = NULL;
    HIP_TRY(hipMalloc(&p, bytes));
    HIP_TRY(hipMemcpy(p, block, bytes, hipMemcpyHostToDevice));
    q->blocks[q->block_count].host = block;
    q->blocks[q->block_count].device = p;
    q->block_count += 1U;
    *out = p;
    return NPH_RESULT_OK;
}

NphResult query_geom(QueryContext *q, const NphPhysicsGeometry *g, NphGeom *out)
{
    const void *block = NULL;
    if (g->type == NPH_PHYSICS_SHAPE_HULL) {
        const NphResult r = query_block(q, g->hull, g->hull->byte_size, &block);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    *out = geom_from(g, block);
    return NPH_RESULT_OK;
}

bool big_hull(const NphPhysicsGeometry *g)
{
    return g->type == NPH_PHYSICS_SHAPE_HULL && g->hull->vertex_count > NPH_HULL_MAX_VERTICES;
}

int convex(NphPhysicsShapeType t)
{
    return t == NPH_PHYSICS_SHAPE_SPHERE || t == NPH_PHYSICS_SHAPE_CAPSULE ||
           t == NPH_PHYSICS_SHAPE_BOX || t == NPH_PHYSICS_SHAPE_HULL;
}

// The fixed device array of nph_amd_collide_many_static: a call's pairs,
// manifolds, caches and hull blocks, laid out from its start. Each device
// holds its own copy until the process ends. In host code `scratch` names a
// host-side stand-in, never the device memory: kernels and copies take the
// address hipGetSymbolAddress gives for the current device.
static __device__ __align__(256) unsigned char scratch[NPH_AMD_STATIC_SCRATCH_BYTES];

// The device address of `scratch`, looked up once per device (with that
// device current) on the first call made on it.
struct StaticScratch {
    std::mutex lock;
    void *ptr[NPH_AMD_MAX_DEVICES];
};
StaticScratch g_static;

NphResult static_scratch(void **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= (int)NPH_AMD_MAX_DEVICES) {
        fprintf(stderr, "nepheline_amd: no HIP device: the collision queries run only on an "
                        "MI355X\n");
        return NPH_RESULT_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> guard(g_static.lock);
    if (g_static.ptr[dev] == NULL) {
        void *p = NULL;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == NULL) {
            fprintf(stderr, "nepheline_amd: the static scratch of device %d has no address: %s\n",
                    dev, hipGetErrorString(e));
            return NPH_RESULT_DEVICE_LOST;
        }
        g_static.ptr[dev] = p;
    }
    *out = g_static.ptr[dev];
    return NPH_RESULT_OK;
}

static inline u64 align256(u64 x) { return (x + 255U) & ~(u64)255U; }

// A hull block of one static call: its place in the scratch array.
struct StaticBlock {
    const NphPhysicsHull *host;
    u64 offset;
};

// The distinct hull blocks of a static call (at most NPH_AMD_STATIC_MAX_HULLS),
// each placed after `base` bytes of the scratch array; false when one is
// refused or there are too many.
NphResult static_blocks(const NphAmdCollideQuery *queries, u32 n, u64 base, StaticBlock *blocks,
                        u32 *count, u64 *end)
{
    u32 k = 0U;
    u64 at = base;
    for (u32 i = 0U; i < n; ++i) {
        const NphPhysicsGeometry *g[2] = {&queries[i].a, &queries[i].b};
        for (u32 s = 0U; s < 2U; ++s) {
            if (!convex(g[s]->type)) {
                return NPH_RESULT_INVALID_ARGUMENT;
            }
            if (g[s]->type != NPH_PHYSICS_SHAPE_HULL) {
                continue;
            }
            const NphPhysicsHull *h = g[s]->hull;
            u32 j = 0U;
            while (j < k && blocks[j].host != h) {
                ++j;
            }
            if (j < k) {
                continue;
            }
            if (k == NPH_AMD_STATIC_MAX_HULLS) {
                fprintf(stderr, "nepheline_amd: nph_amd_collide_many_static takes at most %u "
                                "distinct hull blocks a call\n",
                        NPH_AMD_STATIC_MAX_HULLS);
                return NPH_RESULT_INVALID_ARGUMENT;
            }
            blocks[k].host = h;
            blocks[k].offset = at;
            at = align256(at + h->byte_size);
            ++k;
        }
    }
    *count = k;
    *end = at;
    return NPH_RESULT_OK;
}

} // namespace

// A team a pair (device_narrow.h); hulls beyond 32 vertices in the
// large-hull build.
__global__ void __launch_bounds__(NPH_TEAM_BLOCK) k_query_teams(const NphQueryPair *pairs, u32 n,
                                                                NphManifold *out, NphCache *caches)
{
    query_teams<NPH_HULL_MAX_VERTICES, NPH_CLIP_CAPACITY, NPH_TEAM_BLOCK>(pairs, n, 0, out, caches);
}

__global__ void __launch_bounds__(NPH_BIG_TEAM_BLOCK) k_query_big_teams(const NphQueryPair *pairs, u32 n,
                                                                        NphManifold *out,
                                                                        NphCache *caches)
{
    query_teams<NPH_BIG_HULL_VERTICES, NPH_BIG_CLIP_CAPACITY, NPH_BIG_TEAM_BLOCK>(pairs, n, 1, out,
                                                                                  caches);
}

// Launches the pair queries of `pairs` on `s`.
static void query_launch(hipStream_t s, const NphQueryPair *pairs, u32 n, bool any_big,
                         NphManifold *out, NphCache *caches)
{
    hipLaunchKernelGGL(k_query_teams, dim3(devi_team_grid(n, NPH_TEAM_BLOCK)), dim3(NPH_TEAM_BLOCK),
                       0, s, pairs, n, out, caches);
    if (any_big) {
        hipLaunchKernelGGL(k_query_big_teams, dim3(devi_team_grid(n, NPH_BIG_TEAM_BLOCK)),
                           dim3(NPH_BIG_TEAM_BLOCK), 0, s, pairs, n, out, caches);
    }
}

template <u32 CAP>
__global__ void __launch_bounds__(NPH_SURFACE_TEAM) k_query_surface(const NphSurfaceQuery *in,
                                                                    NphMeshManifolds *out)
{
    __shared__ NphSurfaceLds<NPH_SURFACE_TEAM, CAP> S;
    const NphTeam<NPH_SURFACE_TEAM> tm;
    const NphSurfaceQuery x = *in;
    if (x.mode == 1U) {
        surface_collide(tm, &S, &x.a, x.rotation, x.position, x.mesh, x.speculative);
        if (tm.lane == 0U) {
            *out = S.out;
        }
        return;
    }
    const NphSurfaceShape sh = surface_shape(tm, &S, &x.a, x.rotation);
    NphFacet f;
    NphTriContact c;
    tri_clear(&c);
    if (facet_make(&f, x.tri, x.active, (int)x.two_sided, sh.center)) {
        facet_contact(tm, &S, sh, f, nph_maxf(x.speculative, 0.0F), &c);
    }
    if (tm.lane == 0U) {
        c.point_count = manifold_reduce(c.points, c.point_count, c.normal);
        NphManifold *m = &out->manifolds[0];
        manifold_clear(m);
        if (c.point_count > 0U) {
            m->normal = c.normal;
        }
        for (u32 i = 0U; i < c.point_count; ++i) {
            manifold_push(m, c.points[i].point, c.points[i].separation, c.points[i].id);
        }
        out->count = 1U;
    }
}

__global__ void __launch_bounds__(64) k_query_reduce(NphMPoint *points, u32 count, NphVec3 normal,
                                                     u32 *kept)
{
    if (blockIdx.x == 0U && threadIdx.x == 0U) {
        *kept = manifold_reduce(points, count, normal);
    }
}

extern "C" {

bool nph_physics_collide_supported(NphPhysicsShapeType a, NphPhysicsShapeType b)
{
    return convex(a) && convex(b);
}

NphResult nph_amd_collide_many(const NphAmdCollideQuery *queries, u32 n, NphPhysicsManifold *out,
                               NphPhysicsCollideCache *caches)
{
    if (n == 0U) {
        return NPH_RESULT_OK;
    }
    if (queries == NULL || out == NULL) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    QueryContext *q = &g_query;
    std::lock_guard<std::mutex> guard(q->lock);
    NphResult r = query_ready(q);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    if (n > q->cap) {
        u32 c = q->cap < 64U ? 64U : q->cap;
        while (c < n) {
            c *= 2U;
        }
        HIP_TRY(hipStreamSynchronize(q->stream));
        void *old[3] = {q->pairs, q->manifolds, q->caches};
        for (u32 i = 0U; i < 3U; ++i) {
            if (old[i] != NULL) {
                HIP_TRY(hipFree(old[i]));
            }
        }
        q->pairs = NULL;
        q->manifolds = NULL;
        q->caches = NULL;
        q->cap = 0U;
        HIP_TRY(hipMalloc((void **)&q->pairs, (u64)c * sizeof(NphQueryPair)));
        HIP_TRY(hipMalloc((void **)&q->manifolds, (u64)c * sizeof(NphManifold)));
        HIP_TRY(hipMalloc((void **)&q->caches, (u64)c * sizeof(NphCache)));
        q->cap = c;
    }
    NphQueryPair *h = (NphQueryPair *)malloc((size_t)n * sizeof(NphQueryPair));
    if (h == NULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    bool any_big = false;
    for (u32 i = 0U; i < n && r == NPH_RESULT_OK; ++i) {
        const NphAmdCollideQuery *x = &queries[i];
        memset(&h[i], 0, sizeof h[i]);
        if (!convex(x->a.type) || !convex(x->b.type)) {
            r = NPH_RESULT_INVALID_ARGUMENT;
            break;
        }
        r = query_geom(q, &x->a, &h[i].a);
        if (r == NPH_RESULT_OK) {
            r = query_geom(q, &x->b, &h[i].b);
        }
        h[i].pose_a = x->pose_a;
        h[i].pose_b = x->pose_b;
        h[i].speculative = x->speculative_distance;
        h[i].big = big_hull(&x->a) || big_hull(&x->b) ? 1U : 0U;
        any_big |= h[i].big != 0U;
        if (caches != NULL) {
            memcpy(&h[i].cache, &caches[i], sizeof h[i].cache);
        }
    }
    if (r == NPH_RESULT_OK) {
        const hipError_t e1 = hipMemcpyAsync(q->pairs, h, (u64)n * sizeof(NphQueryPair),
                                             hipMemcpyHostToDevice, q->stream);
        if (e1 == hipSuccess) {
            query_launch(q->stream, q->pairs, n, any_big, q->manifolds, q->caches);
        }
        const hipError_t e2 = e1 == hipSuccess ? hipGetLastError() : e1;
        const hipError_t e3 =
            e2 == hipSuccess ? hipMemcpyAsync(out, q->manifolds, (u64)n * sizeof(NphManifold),
                                              hipMemcpyDeviceToHost, q->stream)
                             : e2;
        const hipError_t e4 = (e3 == hipSuccess && caches != NULL)
                                  ? hipMemcpyAsync(caches, q->caches, (u64)n * sizeof(NphCache),
                                                   hipMemcpyDeviceToHost, q->stream)
                                  : e3;
        const hipError_t e5 = e4 == hipSuccess ? hipStreamSynchronize(q->stream) : e4;
        if (e5 != hipSuccess) {
            fprintf(stderr, "nepheline_amd: collision query failed: %s\n", hipGetErrorString(e5));
            r = NPH_RESULT_DEVICE_LOST;
        }
    }
    free(h);
    return r;
}

void nph_physics_collide(const NphPhysicsGeometry *a, NphPhysicsPose pose_a,
                         const NphPhysicsGeometry *b, NphPhysicsPose pose_b,
                         f32 speculative_distance, NphPhysicsCollideCache *cache,
                         NphPhysicsManifold *out)
{
    NphAmdCollideQuery qy;
    qy.a = *a;
    qy.pose_a = pose_a;
    qy.b = *b;
    qy.pose_b = pose_b;
    qy.speculative_distance = speculative_distance;
    qy.pad = 0U;
    NphPhysicsCollideCache local;
    memset(&local, 0, sizeof local);
    NphPhysicsCollideCache *c = cache != NULL ? cache : &local;
    if (nph_amd_collide_many(&qy, 1U, out, c) != NPH_RESULT_OK) {
        out->normal = v3(0.0F, 1.0F, 0.0F);
        out->point_count = 0U;
    }
}

// Runs one surface query and copies its manifolds back; false without a
// device or on a refused input (the caller then reports no contact).
static bool surface_query(NphSurfaceQuery *x, const NphPhysicsGeometry *a, NphMeshManifolds *out)
{
    QueryContext *q = &g_query;
    std::lock_guard<std::mutex> guard(q->lock);
    if (query_ready(q) != NPH_RESULT_OK || !convex(a->type) || query_geom(q, a, &x->a) !=
                                                                     NPH_RESULT_OK) {
        return false;
    }
    if (x->mode == 1U) {
        const void *mesh = NULL;
        if (query_block(q, x->mesh, x->mesh->byte_size, &mesh) != NPH_RESULT_OK) {
            return false;
        }
        x->mesh = (const NphPhysicsMesh *)mesh;
    }
    if (q->surface_in == NULL &&
        (hipMalloc((void **)&q->surface_in, sizeof(NphSurfaceQuery)) != hipSuccess ||
         hipMalloc((void **)&q->surface_out, sizeof(NphMeshManifolds)) != hipSuccess)) {
        q->surface_in = NULL;
        return false;
    }
    hipError_t e = hipMemcpyAsync(q->surface_in, x, sizeof *x, hipMemcpyHostToDevice, q->stream);
    if (e == hipSuccess) {
        if (big_hull(a)) {
            hipLaunchKernelGGL(k_query_surface<NPH_BIG_CLIP_CAPACITY>, dim3(1),
                               dim3(NPH_SURFACE_TEAM), 0, q->stream, q->surface_in,
                               q->surface_out);
        } else {
            hipLaunchKernelGGL(k_query_surface<NPH_CLIP_CAPACITY>, dim3(1), dim3(NPH_SURFACE_TEAM),
                               0, q->stream, q->surface_in, q->surface_out);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        e = hipMemcpyAsync(out, q->surface_out, sizeof *out, hipMemcpyDeviceToHost, q->stream);
    }
    if (e == hipSuccess) {
        e = hipStreamSynchronize(q->stream);
    }
    if (e != hipSuccess) {
        fprintf(stderr, "nepheline_amd: surface query failed: %s\n", hipGetErrorString(e));
        return false;
    }
    return true;
}

void nph_physics_collide_triangle(const NphPhysicsGeometry *a, NphPhysicsPose pose_a,
                                  const NphPhysicsTriangle *t, f32 speculative_distance,
                                  NphPhysicsManifold *out)
{
    // collide_mesh.c:631-657: the triangle moved next to the shape, which
    // stays turned; points moved back by the shape's place.
    NphSurfaceQuery x;
    memset(&x, 0, sizeof x);
    x.rotation = pose_a.rotation;
    for (u32 i = 0U; i < 3U; ++i) {
        x.tri[i] = v3_sub(t->vertices[i], pose_a.position);
    }
    x.active = t->active_edges & NPH_PHYSICS_MESH_EDGES;
    x.two_sided = t->two_sided != 0 ? 1U : 0U;
    x.speculative = speculative_distance;
    x.mode = 0U;
    NphMeshManifolds mm;
    manifold_clear((NphManifold *)out);
    if (!surface_query(&x, a, &mm)) {
        return;
    }
    const NphManifold *m = &mm.manifolds[0];
    out->normal = m->normal;
    out->point_count = m->point_count;
    for (u32 i = 0U; i < m->point_count; ++i) {
        out->points[i].point = v3_add(m->points[i].point, pose_a.position);
        out->points[i].separation = m->points[i].separation;
        out->points[i].id = m->points[i].id;
    }
}

void nph_physics_collide_mesh(const NphPhysicsGeometry *a, NphPhysicsPose pose_a,
                              const NphPhysicsMesh *mesh, NphPhysicsPose pose_mesh,
                              f32 speculative_distance, NphPhysicsMeshManifolds *out)
{
    // collide_mesh.c:1033-1079: the shape placed in the mesh's frame in 64
    // bits, the manifolds turned back and moved by the shape's place.
    NphSurfaceQuery x;
    memset(&x, 0, sizeof x);
    x.rotation = q_mul(q_conj(pose_mesh.rotation), pose_a.rotation);
    x.position = drotate_inverse(pose_mesh.rotation,
                                 dv3((f64)pose_a.position.x - (f64)pose_mesh.position.x,
                                     (f64)pose_a.position.y - (f64)pose_mesh.position.y,
                                     (f64)pose_a.position.z - (f64)pose_mesh.position.z));
    x.mesh = mesh;
    x.speculative = speculative_distance;
    x.mode = 1U;
    NphMeshManifolds mm;
    out->count = 0U;
    if (mesh == NULL || !surface_query(&x, a, &mm)) {
        return;
    }
    out->count = mm.count;
    for (u32 i = 0U; i < mm.count; ++i) {
        const NphManifold *m = &mm.manifolds[i];
        NphPhysicsManifold *o = &out->manifolds[i];
        o->normal = q_rotate(pose_mesh.rotation, m->normal);
        o->point_count = m->point_count;
        for (u32 j = 0U; j < m->point_count; ++j) {
            o->points[j].point = v3_add(pose_a.position, q_rotate(pose_mesh.rotation,
                                                                  m->points[j].point));
            o->points[j].separation = m->points[j].separation;
            o->points[j].id = m->points[j].id;
        }
    }
}

NphResult nph_amd_collide_many_static(const NphAmdCollideQuery *queries, u32 n,
                                      NphPhysicsManifold *out, NphPhysicsCollideCache *caches)
{
    if (n == 0U) {
        return NPH_RESULT_OK;
    }
    if (queries == NULL || out == NULL) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    // The call's layout in the scratch array: pairs, manifolds, caches, then
    // the hull blocks. Everything is checked before the device is touched.
    const u64 at_manifolds = align256((u64)n * sizeof(NphQueryPair));
    const u64 at_caches = align256(at_manifolds + (u64)n * sizeof(NphManifold));
    const u64 at_blocks = align256(at_caches + (u64)n * sizeof(NphCache));
    StaticBlock blocks[NPH_AMD_STATIC_MAX_HULLS];
    u32 block_count = 0U;
    u64 end = 0U;
    NphResult r = static_blocks(queries, n, at_blocks, blocks, &block_count, &end);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    if (end > NPH_AMD_STATIC_SCRATCH_BYTES) {
        fprintf(stderr, "nepheline_amd: nph_amd_collide_many_static: %u pairs and %u hull blocks "
                        "need %llu bytes, over the %u-byte static scratch\n",
                n, block_count, (unsigned long long)end, NPH_AMD_STATIC_SCRATCH_BYTES);
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    void *base = NULL;
    r = static_scratch(&base);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    unsigned char *dev = (unsigned char *)base;
    NphQueryPair *pairs = (NphQueryPair *)dev;
    NphManifold *manifolds = (NphManifold *)(dev + at_manifolds);
    NphCache *dcaches = (NphCache *)(dev + at_caches);
    for (u32 j = 0U; j < block_count; ++j) {
        HIP_TRY(hipMemcpy(dev + blocks[j].offset, blocks[j].host, blocks[j].host->byte_size,
                          hipMemcpyHostToDevice));
    }
    // The pairs in device form, a chunk at a time through a buffer on the
    // stack (hipMemcpy returns once the chunk is on the device).
    enum { CHUNK = 64 };
    NphQueryPair h[CHUNK];
    bool any_big = false;
    for (u32 first = 0U; first < n; first += CHUNK) {
        const u32 m = n - first < (u32)CHUNK ? n - first : (u32)CHUNK;
        for (u32 i = 0U; i < m; ++i) {
            const NphAmdCollideQuery *x = &queries[first + i];
            memset(&h[i], 0, sizeof h[i]);
            const NphPhysicsGeometry *g[2] = {&x->a, &x->b};
            NphGeom *o[2] = {&h[i].a, &h[i].b};
            for (u32 s = 0U; s < 2U; ++s) {
                const void *address = NULL;
                if (g[s]->type == NPH_PHYSICS_SHAPE_HULL) {
                    u32 j = 0U;
                    while (blocks[j].host != g[s]->hull) {
                        ++j; // present: static_blocks placed every hull
                    }
                    address = dev + blocks[j].offset;
                }
                *o[s] = geom_from(g[s], address);
            }
            h[i].pose_a = x->pose_a;
            h[i].pose_b = x->pose_b;
            h[i].speculative = x->speculative_distance;
            h[i].big = big_hull(&x->a) || big_hull(&x->b) ? 1U : 0U;
            any_big |= h[i].big != 0U;
            if (caches != NULL) {
                memcpy(&h[i].cache, &caches[first + i], sizeof h[i].cache);
            }
        }
        HIP_TRY(hipMemcpy(pairs + first, h, (u64)m * sizeof(NphQueryPair),
                          hipMemcpyHostToDevice));
    }
    query_launch(0, pairs, n, any_big, manifolds, dcaches);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, manifolds, (u64)n * sizeof(NphManifold), hipMemcpyDeviceToHost));
    if (caches != NULL) {
        HIP_TRY(hipMemcpy(caches, dcaches, (u64)n * sizeof(NphCache), hipMemcpyDeviceToHost));
    }
    return NPH_RESULT_OK;
}

u32 nph_physics_manifold_reduce(NphPhysicsManifoldPoint *points, u32 count, NphVec3 normal)
{
    if (count <= NPH_PHYSICS_MANIFOLD_MAX_POINTS || points == NULL) {
        return count;
    }
    QueryContext *q = &g_query;
    std::lock_guard<std::mutex> guard(q->lock);
    if (query_ready(q) != NPH_RESULT_OK) {
        return count;
    }
    if (count > q->cap_points) {
        u32 c = q->cap_points < 256U ? 256U : q->cap_points;
        while (c < count) {
            c *= 2U;
        }
        if (hipStreamSynchronize(q->stream) != hipSuccess) {
            return count;
        }
        if (q->points != NULL) {
            (void)hipFree(q->points);
            q->points = NULL;
            q->cap_points = 0U;
        }
        if (hipMalloc((void **)&q->points, (u64)c * sizeof(NphMPoint)) != hipSuccess) {
            return count;
        }
        q->cap_points = c;
    }
    u32 kept = count;
    if (hipMemcpyAsync(q->points, points, (u64)count * sizeof(NphMPoint), hipMemcpyHostToDevice,
                       q->stream) != hipSuccess) {
        return count;
    }
    hipLaunchKernelGGL(k_query_reduce, dim3(1), dim3(64), 0, q->stream, q->points, count, normal,
                       q->count);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&kept, q->count, 4U, hipMemcpyDeviceToHost, q->stream) != hipSuccess ||
        hipStreamSynchronize(q->stream) != hipSuccess) {
        fprintf(stderr, "nepheline_amd: manifold reduction failed\n");
        return count;
    }
    if (hipMemcpy(points, q->points, (u64)kept * sizeof(NphMPoint), hipMemcpyDeviceToHost) !=
        hipSuccess) {
        return count;
    }
    return kept;
}

} // extern "C"
