// SPDX-License-Identifier: Apache-2.0
// The narrow phase of the convex pairs in two passes, shared by the
// one-world step (device_hip.hip), the batched step (device_batch.hip) and
// the pair queries (query_hip.hip).
//
//  1. The lane pass, one contact a lane: the fat-box test (contact.c:523),
//     the closed forms of spheres, capsules and boxes with spheres, and for
//     every other pair (boxes and hulls) an entry appended to a team list;
//     the wave's entries get their places from one ballot and one atomic
//     add a wave (stream compaction). Hulls of more than 32 vertices go to a
//     list of their own.
//  2. The team pass, one pair a team of NPH_TEAM lanes (nph_polytope.h),
//     over the list the lane pass filled: a persistent grid whose teams
//     stride over the entries the list holds on the device, so the host
//     launches it without reading the count back.
// A pair's manifold depends only on its own record, so the order the lists
// take does not matter to the results; each change code goes to the
// contact's place in the narrow-phase list.
#ifndef NPH_DEVICE_NARROW_H
#define NPH_DEVICE_NARROW_H

#include "device_impl.h"
#include "nph_polytope.h"

#define NPH_TEAM 16U
#define NPH_TEAM_BLOCK 128U    // eight teams a block (hulls of at most 32 vertices)
#define NPH_BIG_TEAM_BLOCK 64U // four teams a block (hulls of up to 256 vertices)
#define NPH_TEAM_GRID_MAX 2048U
#define NPH_BIG_HULL_VERTICES 256U
#define NPH_BIG_CLIP_CAPACITY (NPH_BIG_HULL_VERTICES + 8U) // the reference's

// Mesh pairs: a team of NPH_SURFACE_TEAM lanes a pair, four teams a block;
// two with the large hulls' clip capacity (worlds that hold such hulls).
#define NPH_MESH_TEAM_BLOCK 64U
#define NPH_BIG_MESH_TEAM_BLOCK 32U

// The team lists of one narrow-phase launch: entries for 32-vertex teams,
// entries for large-hull teams, mesh pairs, and their three counts.
typedef struct NphTeamLists {
    NphSortedEntry *small;
    NphSortedEntry *big;
    NphSortedEntry *mesh;
    u32 *counts;
} NphTeamLists;

// Blocks of a persistent team launch for at most `n` entries.
static inline u32 devi_team_grid(u32 n, u32 block)
{
    const u32 teams = block / NPH_TEAM;
    const u32 b = (n + teams - 1U) / teams;
    return b == 0U ? 1U : (b < NPH_TEAM_GRID_MAX ? b : NPH_TEAM_GRID_MAX);
}

// One world's records, for the passes' world accessor.
struct NphOneWorld {
    NphCollideWorld c;
    __device__ inline NphCollideWorld operator()(u32) const { return c; }
};

// Many worlds' records (the batch's table).
struct NphManyWorlds {
    const NphCollideWorld *cw;
    __device__ inline NphCollideWorld operator()(u32 w) const { return cw[w]; }
};

__device__ inline void narrow_counts_clear(u32 *counts)
{
    if (threadIdx.x < 3U) {
        counts[threadIdx.x] = 0U;
    }
}

// Appends the entries of the wave's lanes that pass `p`, in lane order, at
// places reserved by one atomic add. Every lane of the wave takes part.
__device__ inline void wave_append(bool p, const NphSortedEntry &e, NphSortedEntry *list, u32 *count)
{
    const u64 m = __ballot(p ? 1 : 0);
    if (m == 0ULL) {
        return;
    }
    const u32 lane = __lane_id();
    const u32 leader = (u32)__builtin_ctzll(m);
    u32 base = 0U;
    if (lane == leader) {
        base = atomicAdd(count, (u32)__popcll(m));
    }
    base = (u32)__shfl((int)base, (int)leader);
    if (p) {
        list[base + (u32)__popcll(m & ((1ULL << lane) - 1ULL))] = e;
    }
}

// The lane pass for entry `e` (valid false: a lane past the list's end,
// which still takes part in the compaction).
template <class WorldOf>
__device__ inline void narrow_lane(bool valid, const NphSortedEntry &e, const WorldOf &world,
                                   u32 *codes, f32 step_dt, const NphTeamLists &lists)
{
    int route = 0; // 1: a 32-vertex team, 2: a large-hull team, 3: a mesh team
    if (valid) {
        const NphCollideWorld c = world(e.world);
        NphContactRec *rec = &c.contacts[e.record];
        if (rec->child != NPH_NULL32) {
            // A mesh pair (a chain of records): its fat-box test here
            // (contact.c:523-526), the rest on a mesh team.
            rec->change = 0U;
            if (!aabb_overlaps(c.shapes[rec->shape_a].fat_aabb, c.shapes[rec->shape_b].fat_aabb)) {
                rec->change = NPH_CHANGE_DISJOINT;
                codes[e.index] = NPH_CHANGE_DISJOINT;
            } else {
                route = 3;
            }
        } else {
            NphPairJob job;
            if (!contact_job(rec, c.shapes, c.bodies, step_dt, &job)) {
                rec->change = NPH_CHANGE_DISJOINT;
                codes[e.index] = NPH_CHANGE_DISJOINT;
            } else {
                NphManifold m;
                int done = 1;
                if (pair_wants_team(job.ga.type, job.gb.type)) {
                    // A cached axis that still separates ends the pair here,
                    // as it would end it on the team.
                    done = cached_axis_separates(&job.ga, job.pa, &job.gb, job.pb,
                                                 job.speculative, &job.cache);
                    if (done) {
                        manifold_clear(&m);
                        if (job.ga.type < job.gb.type) {
                            m.normal = v3_neg(m.normal);
                        }
                    } else {
                        route = contact_has_big_hull(rec, c.shapes) ? 2 : 1;
                    }
                } else if (pair_has_closed_form(job.ga.type, job.gb.type)) {
                    collide_closed_form(&job.ga, job.pa, &job.gb, job.pb, job.speculative, &m);
                } else {
                    NphLaneScratch sc;
                    lane_collide(&job.ga, job.pa, &job.gb, job.pb, job.speculative, &job.cache,
                                 &sc, &m);
                }
                if (done) {
                    contact_store(rec, c.shapes, &job, &m);
                    codes[e.index] = rec->change;
                }
            }
        }
    }
    wave_append(route == 1, e, lists.small, &lists.counts[0]);
    wave_append(route == 2, e, lists.big, &lists.counts[1]);
    wave_append(route == 3, e, lists.mesh, &lists.counts[2]);
}

// The team pass over a team list of `*count` entries (BLOCK threads a
// block, teams of NPH_TEAM lanes, hulls of at most V vertices).
template <u32 V, u32 CAP, u32 BLOCK, class WorldOf>
__device__ inline void narrow_teams(const NphSortedEntry *list, const u32 *count,
                                    const WorldOf &world, u32 *codes, f32 step_dt)
{
    constexpr u32 TEAMS = BLOCK / NPH_TEAM;
    __shared__ NphPolyScratch<V, CAP> scratch[TEAMS];
    const NphTeam<NPH_TEAM> tm;
    const u32 slot = threadIdx.x / NPH_TEAM;
    const u32 n = *count;
    for (u32 t = blockIdx.x * TEAMS + slot; t < n; t += gridDim.x * TEAMS) {
        const NphSortedEntry e = list[t];
        const NphCollideWorld c = world(e.world);
        NphContactRec *rec = &c.contacts[e.record];
        NphPairJob job;
        (void)contact_job(rec, c.shapes, c.bodies, step_dt, &job); // the lane pass kept it
        NphManifold m;
        team_collide(tm, &scratch[slot], &job.ga, job.pa, &job.gb, job.pb, job.speculative,
                     &job.cache, &m);
        if (tm.lane == 0U) {
            contact_store(rec, c.shapes, &job, &m);
            codes[e.index] = rec->change;
        }
        tm.sync();
    }
}

// The mesh pairs of a mesh list of `*count` entries, a team a pair
// (nph_surface.h), persistent like the convex teams; lane 0 carries the
// manifolds into the pair's chain of records.
template <u32 CAP, u32 BLOCK, class WorldOf>
__device__ inline void narrow_mesh_teams(const NphSortedEntry *list, const u32 *count,
                                         const WorldOf &world, u32 *codes, f32 step_dt)
{
    constexpr u32 TEAMS = BLOCK / NPH_SURFACE_TEAM;
    __shared__ NphSurfaceLds<NPH_SURFACE_TEAM, CAP> lds[TEAMS];
    const NphTeam<NPH_SURFACE_TEAM> tm;
    const u32 slot = threadIdx.x / NPH_SURFACE_TEAM;
    const u32 n = *count;
    for (u32 t = blockIdx.x * TEAMS + slot; t < n; t += gridDim.x * TEAMS) {
        const NphSortedEntry e = list[t];
        const NphCollideWorld c = world(e.world);
        NphMeshFrame f;
        mesh_frame(&c.contacts[e.record], c.shapes, c.bodies, step_dt, &f);
        surface_collide(tm, &lds[slot], f.other, f.rotation, f.position, f.mesh, f.speculative);
        if (tm.lane == 0U) {
            mesh_store(c.contacts, e.record, c.shapes, &f, &lds[slot].out);
            codes[e.index] = c.contacts[e.record].change;
        }
        tm.sync();
    }
}

// A pair query on a team (no records): pairs flagged `big` in the large-hull
// build only, the others in the 32-vertex build.
template <u32 V, u32 CAP, u32 BLOCK>
__device__ inline void query_teams(const NphQueryPair *pairs, u32 n, int big, NphManifold *out,
                                   NphCache *caches)
{
    constexpr u32 TEAMS = BLOCK / NPH_TEAM;
    __shared__ NphPolyScratch<V, CAP> scratch[TEAMS];
    const NphTeam<NPH_TEAM> tm;
    const u32 slot = threadIdx.x / NPH_TEAM;
    for (u32 i = blockIdx.x * TEAMS + slot; i < n; i += gridDim.x * TEAMS) {
        if ((pairs[i].big != 0U) != (big != 0)) {
            continue;
        }
        const NphQueryPair p = pairs[i];
        NphCache cache = p.cache;
        NphManifold m;
        team_collide(tm, &scratch[slot], &p.a, p.pose_a, &p.b, p.pose_b, p.speculative, &cache,
                     &m);
        if (tm.lane == 0U) {
            out[i] = m;
            caches[i] = cache;
        }
        tm.sync();
    }
}

#endif // NPH_DEVICE_NARROW_H
