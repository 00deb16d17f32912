// SPDX-License-Identifier: Apache-2.0
// TEST INFRASTRUCTURE ONLY.  A plain-C restatement of the reference's
// soft-step solver arithmetic and state hash, used by tests/ (and nothing
// else) as the checker of the HIP world step.  It never runs in the product.
//
// Pinned: every function here is compared bit for bit against the
// reference's own code compiled from its sources (oracle/_ref, through
// ref_shim.c) on seeded random inputs, and the hash against the reference's
// regression vectors (engine/core/tests/test_hash.c:21-33).  See
// tests/test_oracle.py.
//
// Layouts are the reference's, so the same bytes go to both:
//   OrSolverBody  = PhysSolverBody   (world_internal.h:387-400, 64 B)
//   OrGroup       = PhysContactLanes (world_internal.h:430-458)
//   OrPoints      = PhysPointLanes   (world_internal.h:460-479)
//   OrPrepareIn   = PhysPrepareLanes (world_internal.h:482-489)
#ifndef NPH_ORACLE_H
#define NPH_ORACLE_H

#include <stdint.h>

#define OR_LANES 8
#define OR_MAX_POINTS 4

typedef struct OrSolverBody {
    float v[4];
    float w[4];
    float dp[4];
    float dq[4];
} OrSolverBody;

typedef struct OrGroup {
    uint32_t body_a[OR_LANES];
    uint32_t body_b[OR_LANES];
    uint32_t contact[OR_LANES];
    uint32_t write_mask_a;
    uint32_t write_mask_b;
    uint32_t lane_count;
    uint32_t point_count;
    uint32_t first_point;
    uint32_t scale_hits;
    uint32_t pad[2];
    float inv_mass_a[OR_LANES];
    float inv_mass_b[OR_LANES];
    float normal[3][OR_LANES];
    float tangent1[3][OR_LANES];
    float tangent2[3][OR_LANES];
    float friction[OR_LANES];
    float restitution[OR_LANES];
    float bias_rate[OR_LANES];
    float mass_scale[OR_LANES];
    float impulse_scale[OR_LANES];
    float impact_lanes[OR_LANES];
    float impacted[OR_LANES];
} OrGroup;

typedef struct OrPoints {
    float anchor_a[3][OR_LANES];
    float anchor_b[3][OR_LANES];
    float adjusted_separation[OR_LANES];
    float normal_mass[OR_LANES];
    float normal_impulse[OR_LANES];
    float max_normal_impulse[OR_LANES];
    float relative_velocity[OR_LANES];
    float substep_velocity[OR_LANES];
    float tangent_mass[3][OR_LANES];
    float tangent_impulse[2][OR_LANES];
    float ang_a_n[3][OR_LANES];
    float ang_b_n[3][OR_LANES];
    float ang_a_t1[3][OR_LANES];
    float ang_b_t1[3][OR_LANES];
    float ang_a_t2[3][OR_LANES];
    float ang_b_t2[3][OR_LANES];
} OrPoints;

typedef struct OrPrepareIn {
    float inv_inertia_a[9][OR_LANES];
    float inv_inertia_b[9][OR_LANES];
    float tangent_impulse[OR_MAX_POINTS][3][OR_LANES];
    float active[OR_MAX_POINTS][OR_LANES];
} OrPrepareIn;

#ifdef __cplusplus
extern "C" {
#endif

// solver_impl.h:276-385
void oracle_prepare(const OrSolverBody *bodies, OrGroup *g, OrPoints *points,
                    const OrPrepareIn *in);
// solver_impl.h:181-258 (with the TGS hit scaling when g->scale_hits)
void oracle_warm_start(OrSolverBody *bodies, OrGroup *g, OrPoints *points);
// solver_impl.h:484-602
void oracle_solve(OrSolverBody *bodies, OrGroup *g, OrPoints *points, float inv_h, float push_max,
                  float slop, int use_bias);
// solver_impl.h:611-668
void oracle_restitution(OrSolverBody *bodies, OrGroup *g, OrPoints *points, float threshold);
// solver.c:48-65: out = {bias_rate, mass_scale, impulse_scale}
void oracle_make_soft(float hertz, float zeta, float h, float out[3]);
// engine/core/src/hash.c:73-120 (wyhash construction)
uint64_t oracle_hash64(const void *data, uint64_t len, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif // NPH_ORACLE_H
