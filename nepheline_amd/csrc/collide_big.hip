// SPDX-License-Identifier: Apache-2.0
// The narrow phase for convex pairs with a hull of more than 32 vertices:
// the same collide code as the step's kernels (nph_collide.h), built here
// with B's vertex array sized for the reference's largest hull (256
// vertices, hull.h:45), B's face normals recomputed rather than kept, and
// a clip capacity of 132: hull faces have at most 64 vertices (the host's
// limit), so a clipped face holds at most 64 + 64 points and neither this
// capacity nor the reference's 264 (collide_hull.c:26) ever truncates. The
// step's convex kernels skip these pairs (NPH_SHAPE_BIG_HULL on a shape,
// `big` on a query) and these kernels take only them, so each pair runs in
// exactly one. The frame stays near the mesh kernel's: a 15 KB one (B's
// normals kept, capacity 264) could not get its scratch on batch streams
// (HSA_STATUS_ERROR_OUT_OF_RESOURCES). Mesh pairs need none of this:
// against a triangle the per-shape arrays hold the triangle, and a clipped
// face fits the 32-vertex build's 68 points.
#define NPH_HULL_MAX_VERTICES 256U
#define NPH_CLIP_CAPACITY     132U
#define NPH_CACHE_B_NORMALS   0

#include "device_impl.h"

__global__ void __launch_bounds__(NPH_COLLIDE_BLOCK) k_collide_big(NphContactRec *contacts, const NphShapeRec *shapes,
                              const NphBodyRec *bodies, const u32 *list, u32 *codes, u32 n, f32 step_dt)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || contacts[list[i]].child != NPH_NULL32 ||
        !contact_has_big_hull(&contacts[list[i]], shapes)) {
        return;
    }
    contact_update(contacts, list[i], shapes, bodies, step_dt);
    codes[i] = contacts[list[i]].change;
}

__global__ void __launch_bounds__(NPH_COLLIDE_BLOCK) k_b_collide_big(const NphSortedEntry *__restrict__ sorted, u32 n,
                                                      const NphCollideWorld *__restrict__ cw,
                                                      u32 *codes, f32 step_dt)
{
    const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) {
        return;
    }
    const NphSortedEntry e = sorted[j];
    const NphCollideWorld c = cw[e.world];
    if (c.contacts[e.record].child != NPH_NULL32 ||
        !contact_has_big_hull(&c.contacts[e.record], c.shapes)) {
        return;
    }
    contact_update(c.contacts, e.record, c.shapes, c.bodies, step_dt);
    codes[e.index] = c.contacts[e.record].change;
}

__global__ void __launch_bounds__(64) k_query_collide_big(const NphQueryPair *pairs, u32 n,
                                                          NphManifold *out, NphCache *caches)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || pairs[i].big == 0U) {
        return;
    }
    const NphQueryPair p = pairs[i];
    NphCache cache = p.cache;
    NphManifold m;
    nph_collide(&p.a, p.pose_a, &p.b, p.pose_b, p.speculative, &cache, &m);
    out[i] = m;
    caches[i] = cache;
}

void devi_collide_big(hipStream_t s, NphContactRec *contacts, const NphShapeRec *shapes,
                      const NphBodyRec *bodies, const u32 *list, u32 *codes, u32 n, f32 step_dt)
{
    hipLaunchKernelGGL(k_collide_big, dim3(devi_blocks(n, NPH_COLLIDE_BLOCK)),
                       dim3(NPH_COLLIDE_BLOCK), 0, s, contacts, shapes, bodies, list, codes, n,
                       step_dt);
}

void devi_b_collide_big(hipStream_t s, const NphSortedEntry *sorted, u32 n,
                        const NphCollideWorld *cw, u32 *codes, f32 step_dt)
{
    hipLaunchKernelGGL(k_b_collide_big, dim3(devi_blocks(n, NPH_COLLIDE_BLOCK)),
                       dim3(NPH_COLLIDE_BLOCK), 0, s, sorted, n, cw, codes, step_dt);
}

void devi_query_collide_big(hipStream_t s, const NphQueryPair *pairs, u32 n, NphManifold *out,
                            NphCache *caches)
{
    hipLaunchKernelGGL(k_query_collide_big, dim3(devi_blocks(n, 64U)), dim3(64), 0, s, pairs, n,
                       out, caches);
}
