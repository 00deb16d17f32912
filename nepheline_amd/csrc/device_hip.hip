// SPDX-License-Identifier: Apache-2.0
// The world step's device stages on an MI355X (gfx950), in HIP.
//
// Records (bodies, shapes, contacts) are arrays in HBM indexed by the host's
// slots.  The step's per-item work is the __host__ __device__ code of
// nph_step.h: one thread per awake contact for the narrow phase, one per body
// for integration and finalize, one per lane for the constraint passes.  A
// colour's lanes touch disjoint dynamic bodies, so a colour pass is one
// launch; passes follow the reference's order (the overflow's groups first,
// one after another on one thread, then colours 0..23).
//
// Compiled with -ffp-contract=off, correctly rounded f32 division and
// square root, and denormals kept, so each operation rounds as the
// reference's SSE code does.

#include <hip/hip_runtime.h>

#include "nph_device.h"

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

typedef struct HullEntry {
    const NphPhysicsHull *host;
    const NphPhysicsHull *device;
} HullEntry;

struct NphDevice {
    hipStream_t stream;
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
    u32 *compact; // changed records then codes, 2 * cap_list + 1
    NphContactInit *inits;
    u32 cap_inits;
    NphGroup *groups;
    u32 *lane_contact, *lane_body_a, *lane_body_b, *lane_flags;
    f32 *lanes;
    u32 cap_groups;
    f32 *points;
    u32 cap_point_rows;
    NphGrown *grown; // cap_shapes + 1 (count first)
    // Host staging (pinned).
    u32 *h_u32;
    u32 cap_h_u32;
    NphGrown *h_grown;
    u32 cap_h_grown;
    HullEntry *hulls;
    u32 hull_count, hull_cap;
};

const char *nph_dev_name(void) { return "hip:gfx950"; }

static NphResult grow_buffer(void **p, u64 old_bytes, u64 new_bytes, hipStream_t s, int keep)
{
    void *n = NULL;
    HIP_TRY(hipMalloc(&n, new_bytes));
    if (keep && *p != NULL && old_bytes > 0U) {
        HIP_TRY(hipMemcpyAsync(n, *p, old_bytes, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    if (*p != NULL) {
        HIP_TRY(hipFree(*p));
    }
    *p = n;
    return NPH_RESULT_OK;
}

static u32 grow_cap(u32 cap, u32 need)
{
    u32 c = cap < 64U ? 64U : cap;
    while (c < need) {
        c = c * 2U;
    }
    return c;
}

NphResult nph_dev_create(NphDevice **out)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
        fprintf(stderr, "nepheline_amd: no HIP device: the world step runs only on an MI355X\n");
        return NPH_RESULT_UNSUPPORTED;
    }
    NphDevice *d = (NphDevice *)calloc(1, sizeof *d);
    if (d == NULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
        free(d);
        return NPH_RESULT_DEVICE_LOST;
    }
    for (u32 i = 0U; i < 4U; ++i) {
        if (hipEventCreate(&d->ev[i]) != hipSuccess) {
            free(d);
            return NPH_RESULT_DEVICE_LOST;
        }
    }
    *out = d;
    return NPH_RESULT_OK;
}

void nph_dev_destroy(NphDevice *d)
{
    if (d == NULL) {
        return;
    }
    (void)hipStreamSynchronize(d->stream);
    void *bufs[] = {d->bodies, d->shapes, d->contacts, d->solver_of_body, d->awake_bodies,
                    d->sbodies, d->params, d->list, d->codes, d->compact, d->inits, d->groups,
                    d->lane_contact, d->lane_body_a, d->lane_body_b, d->lane_flags, d->lanes,
                    d->points, d->grown};
    for (u32 i = 0U; i < sizeof bufs / sizeof bufs[0]; ++i) {
        if (bufs[i] != NULL) {
            (void)hipFree(bufs[i]);
        }
    }
    for (u32 i = 0U; i < d->hull_count; ++i) {
        (void)hipFree((void *)d->hulls[i].device);
    }
    free(d->hulls);
    if (d->h_u32 != NULL) {
        (void)hipHostFree(d->h_u32);
    }
    if (d->h_grown != NULL) {
        (void)hipHostFree(d->h_grown);
    }
    for (u32 i = 0U; i < 4U; ++i) {
        (void)hipEventDestroy(d->ev[i]);
    }
    (void)hipStreamDestroy(d->stream);
    free(d);
}

f64 nph_dev_last_solve_ms(const NphDevice *d) { return d->solve_ms; }
f64 nph_dev_last_collide_ms(const NphDevice *d) { return d->collide_ms; }
f64 nph_dev_last_passes_ms(const NphDevice *d) { return d->passes_ms; }

NphResult nph_dev_reserve(NphDevice *d, u32 bodies, u32 shapes, u32 contacts)
{
    if (bodies > d->cap_bodies) {
        const u32 c = grow_cap(d->cap_bodies, bodies);
        NphResult r;
        if ((r = grow_buffer((void **)&d->bodies, (u64)d->cap_bodies * sizeof(NphBodyRec),
                             (u64)c * sizeof(NphBodyRec), d->stream, 1)) != NPH_RESULT_OK ||
            (r = grow_buffer((void **)&d->solver_of_body, 0, (u64)c * 4U, d->stream, 0)) !=
                NPH_RESULT_OK ||
            (r = grow_buffer((void **)&d->awake_bodies, 0, (u64)c * 4U, d->stream, 0)) !=
                NPH_RESULT_OK ||
            (r = grow_buffer((void **)&d->sbodies, 0, (u64)(c + 1U) * sizeof(NphSolverBody),
                             d->stream, 0)) != NPH_RESULT_OK ||
            (r = grow_buffer((void **)&d->params, 0, (u64)c * sizeof(NphSolverParams), d->stream,
                             0)) != NPH_RESULT_OK) {
            return r;
        }
        d->cap_bodies = c;
    }
    if (shapes > d->cap_shapes) {
        const u32 c = grow_cap(d->cap_shapes, shapes);
        NphResult r;
        if ((r = grow_buffer((void **)&d->shapes, (u64)d->cap_shapes * sizeof(NphShapeRec),
                             (u64)c * sizeof(NphShapeRec), d->stream, 1)) != NPH_RESULT_OK ||
            (r = grow_buffer((void **)&d->grown, 0, (u64)(c + 1U) * sizeof(NphGrown), d->stream,
                             0)) != NPH_RESULT_OK) {
            return r;
        }
        d->cap_shapes = c;
    }
    if (contacts > d->cap_contacts) {
        const u32 c = grow_cap(d->cap_contacts, contacts);
        NphResult r = grow_buffer((void **)&d->contacts, (u64)d->cap_contacts * sizeof(NphContactRec),
                                  (u64)c * sizeof(NphContactRec), d->stream, 1);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        d->cap_contacts = c;
    }
    return NPH_RESULT_OK;
}

static NphResult reserve_list(NphDevice *d, u32 n)
{
    if (n <= d->cap_list) {
        return NPH_RESULT_OK;
    }
    const u32 c = grow_cap(d->cap_list, n);
    NphResult r;
    if ((r = grow_buffer((void **)&d->list, 0, (u64)c * 4U, d->stream, 0)) != NPH_RESULT_OK ||
        (r = grow_buffer((void **)&d->codes, 0, (u64)c * 4U, d->stream, 0)) != NPH_RESULT_OK ||
        (r = grow_buffer((void **)&d->compact, 0, (u64)(2U * c + 1U) * 4U, d->stream, 0)) !=
            NPH_RESULT_OK) {
        return r;
    }
    d->cap_list = c;
    return NPH_RESULT_OK;
}

static NphResult reserve_host_u32(NphDevice *d, u32 n)
{
    if (n <= d->cap_h_u32) {
        return NPH_RESULT_OK;
    }
    const u32 c = grow_cap(d->cap_h_u32, n);
    if (d->h_u32 != NULL) {
        HIP_TRY(hipHostFree(d->h_u32));
        d->h_u32 = NULL;
    }
    HIP_TRY(hipHostMalloc((void **)&d->h_u32, (u64)c * 4U, hipHostMallocDefault));
    d->cap_h_u32 = c;
    return NPH_RESULT_OK;
}

static NphResult reserve_groups(NphDevice *d, u32 groups)
{
    if (groups <= d->cap_groups) {
        return NPH_RESULT_OK;
    }
    const u32 c = grow_cap(d->cap_groups, groups);
    const u64 lanes = (u64)c * NPH_LANES;
    const u64 rows = (u64)c * NPH_PHYSICS_MANIFOLD_MAX_POINTS;
    NphResult r;
    if ((r = grow_buffer((void **)&d->groups, 0, (u64)c * sizeof(NphGroup), d->stream, 0)) !=
            NPH_RESULT_OK ||
        (r = grow_buffer((void **)&d->lane_contact, 0, lanes * 4U, d->stream, 0)) != NPH_RESULT_OK ||
        (r = grow_buffer((void **)&d->lane_body_a, 0, lanes * 4U, d->stream, 0)) != NPH_RESULT_OK ||
        (r = grow_buffer((void **)&d->lane_body_b, 0, lanes * 4U, d->stream, 0)) != NPH_RESULT_OK ||
        (r = grow_buffer((void **)&d->lane_flags, 0, lanes * 4U, d->stream, 0)) != NPH_RESULT_OK ||
        (r = grow_buffer((void **)&d->lanes, 0, lanes * NPH_L_FIELDS * 4U, d->stream, 0)) !=
            NPH_RESULT_OK ||
        (r = grow_buffer((void **)&d->points, 0, rows * NPH_LANES * NPH_P_FIELDS * 4U, d->stream,
                         0)) != NPH_RESULT_OK) {
        return r;
    }
    d->cap_groups = c;
    d->cap_point_rows = (u32)rows;
    return NPH_RESULT_OK;
}

NphResult nph_dev_hull(NphDevice *d, const NphPhysicsHull *hull, const NphPhysicsHull **out)
{
    for (u32 i = 0U; i < d->hull_count; ++i) {
        if (d->hulls[i].host == hull) {
            *out = d->hulls[i].device;
            return NPH_RESULT_OK;
        }
    }
    if (d->hull_count == d->hull_cap) {
        const u32 c = d->hull_cap == 0U ? 16U : 2U * d->hull_cap;
        HullEntry *n = (HullEntry *)realloc(d->hulls, (size_t)c * sizeof(HullEntry));
        if (n == NULL) {
            return NPH_RESULT_OUT_OF_MEMORY;
        }
        d->hulls = n;
        d->hull_cap = c;
    }
    void *p = NULL;
    HIP_TRY(hipMalloc(&p, hull->byte_size));
    HIP_TRY(hipMemcpy(p, hull, hull->byte_size, hipMemcpyHostToDevice));
    d->hulls[d->hull_count].host = hull;
    d->hulls[d->hull_count].device = (const NphPhysicsHull *)p;
    d->hull_count += 1U;
    *out = (const NphPhysicsHull *)p;
    return NPH_RESULT_OK;
}

NphResult nph_dev_write_body(NphDevice *d, u32 slot, const NphBodyRec *b)
{
    HIP_TRY(hipMemcpyAsync(&d->bodies[slot], b, sizeof *b, hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return NPH_RESULT_OK;
}

NphResult nph_dev_write_shape(NphDevice *d, u32 slot, const NphShapeRec *s)
{
    HIP_TRY(hipMemcpyAsync(&d->shapes[slot], s, sizeof *s, hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return NPH_RESULT_OK;
}

NphResult nph_dev_read_bodies(NphDevice *d, NphBodyRec *out, u32 count)
{
    if (count > 0U) {
        HIP_TRY(hipMemcpyAsync(out, d->bodies, (u64)count * sizeof *out, hipMemcpyDeviceToHost,
                               d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
    }
    return NPH_RESULT_OK;
}

NphResult nph_dev_read_shapes(NphDevice *d, NphShapeRec *out, u32 count)
{
    if (count > 0U) {
        HIP_TRY(hipMemcpyAsync(out, d->shapes, (u64)count * sizeof *out, hipMemcpyDeviceToHost,
                               d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
    }
    return NPH_RESULT_OK;
}

NphResult nph_dev_read_contacts(NphDevice *d, NphContactRec *out, u32 count)
{
    if (count > 0U) {
        HIP_TRY(hipMemcpyAsync(out, d->contacts, (u64)count * sizeof *out, hipMemcpyDeviceToHost,
                               d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
    }
    return NPH_RESULT_OK;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(NPH_BLOCK) k_init_contacts(NphContactRec *contacts, const NphContactInit *list, u32 n)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const NphContactInit in = list[i];
    NphContactRec *c = &contacts[in.record];
    memset(c, 0, sizeof *c);
    c->shape_a = in.shape_a;
    c->shape_b = in.shape_b;
    c->body_a = in.body_a;
    c->body_b = in.body_b;
    c->normal = v3(0.0F, 1.0F, 0.0F);
}

__global__ void __launch_bounds__(NPH_COLLIDE_BLOCK) k_collide(NphContactRec *contacts, const NphShapeRec *shapes,
                          const NphBodyRec *bodies, const u32 *list, u32 *codes, u32 n, f32 step_dt)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    NphContactRec *c = &contacts[list[i]];
    contact_update(c, shapes, bodies, step_dt);
    codes[i] = c->change;
}

// Exclusive scan of one value per thread over a block of NPH_SCAN_BLOCK.
__device__ static u32 block_exclusive_scan(u32 v, u32 *total)
{
    __shared__ u32 s[NPH_SCAN_BLOCK];
    const u32 t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (u32 off = 1U; off < NPH_SCAN_BLOCK; off <<= 1U) {
        const u32 x = t >= off ? s[t - off] : 0U;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    const u32 incl = s[t];
    *total = s[NPH_SCAN_BLOCK - 1U];
    __syncthreads();
    return incl - v;
}

// Order-keeping compaction of the changed records: out[0] = count, then the
// records, then their codes.
__global__ void __launch_bounds__(NPH_SCAN_BLOCK) k_compact_changes(const u32 *list, const u32 *codes, u32 n, u32 *out)
{
    u32 base = 0U;
    for (u32 start = 0U; start < n; start += NPH_SCAN_BLOCK) {
        const u32 i = start + threadIdx.x;
        const u32 code = i < n ? codes[i] : 0U;
        u32 total = 0U;
        const u32 pos = block_exclusive_scan(code != 0U ? 1U : 0U, &total);
        if (code != 0U) {
            out[1U + base + pos] = list[i];
            out[1U + n + base + pos] = code;
        }
        base += total;
    }
    if (threadIdx.x == 0U) {
        out[0] = base;
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_solver_map_fill(u32 *map, u32 slots, u32 dummy)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < slots) {
        map[i] = dummy;
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_solver_map_set(u32 *map, const u32 *awake, u32 n)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        map[awake[i]] = i;
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_prepare_bodies(NphStepData d, NphStepParams sp)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.body_count) {
        const u32 slot = d.awake_bodies[i];
        prepare_body(&sp, &d.bodies[slot], slot, &d.sbodies[i], &d.params[i]);
    } else if (i == d.body_count) {
        solver_body_dummy(&d.sbodies[i]);
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_group_points(NphStepData d, u32 group_count)
{
    const u32 g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= group_count) {
        return;
    }
    NphGroup *grp = &d.groups[g];
    u32 p = 0U;
    for (u32 l = 0U; l < grp->lane_count; ++l) {
        p = nph_maxu(p, d.contacts[d.lane_contact[g * NPH_LANES + l]].point_count);
    }
    grp->point_count = p;
}

__global__ void __launch_bounds__(NPH_SCAN_BLOCK) k_group_scan(NphStepData d, u32 group_count)
{
    u32 base = 0U;
    for (u32 start = 0U; start < group_count; start += NPH_SCAN_BLOCK) {
        const u32 g = start + threadIdx.x;
        const u32 p = g < group_count ? d.groups[g].point_count : 0U;
        u32 total = 0U;
        const u32 pos = block_exclusive_scan(p, &total);
        if (g < group_count) {
            d.groups[g].first_point = base + pos;
        }
        base += total;
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_lane_prepare(NphStepData d, NphStepParams sp, u32 group_count)
{
    const u32 L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L < group_count * NPH_LANES) {
        lane_prepare(&d, &sp, L / NPH_LANES, L % NPH_LANES);
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_group_impact(NphStepData d, u32 group_count)
{
    const u32 g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < group_count) {
        group_finish_prepare(&d, g);
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_integrate_velocities(NphStepData d, NphStepParams sp)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.body_count) {
        integrate_velocity(&sp, &d.params[i], &d.sbodies[i]);
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_integrate_positions(NphStepData d, NphStepParams sp)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.body_count) {
        integrate_position(&sp, &d.params[i], &d.sbodies[i]);
    }
}

enum { PASS_WARM = 0, PASS_SOLVE = 1, PASS_RELAX = 2, PASS_RESTITUTION = 3 };

__device__ static void run_lane(const NphStepData *d, const NphStepParams *sp, u32 pass, u32 g,
                                u32 l)
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

// One colour's groups [first, first + count): a thread per lane.
__global__ void __launch_bounds__(NPH_BLOCK) k_color_pass(NphStepData d, NphStepParams sp, u32 pass, u32 first, u32 count)
{
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count * NPH_LANES) {
        run_lane(&d, &sp, pass, first + t / NPH_LANES, t % NPH_LANES);
    }
}

// The overflow: its groups (one contact each) in order, on one thread.
__global__ void __launch_bounds__(64) k_overflow_pass(NphStepData d, NphStepParams sp, u32 pass, u32 first, u32 count)
{
    if (blockIdx.x == 0U && threadIdx.x == 0U) {
        for (u32 g = first; g < first + count; ++g) {
            run_lane(&d, &sp, pass, g, 0U);
        }
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_restitution_masks(NphStepData d, f32 threshold, u32 group_count)
{
    const u32 g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < group_count) {
        group_restitution_masks(&d, threshold, g);
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_store(NphStepData d, u32 group_count)
{
    const u32 L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L < group_count * NPH_LANES) {
        lane_store(&d, L / NPH_LANES, L % NPH_LANES);
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_finalize(NphStepData d, NphStepParams sp)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.body_count) {
        finalize_body(&sp, &d.params[i], &d.sbodies[i], &d.bodies[d.awake_bodies[i]], d.shapes);
    }
}

// The grown shapes in the awake list's order, each body's shapes in its
// list's order; their flags cleared.  out[0].shape = count.
__global__ void __launch_bounds__(NPH_SCAN_BLOCK) k_collect_grown(NphStepData d, NphGrown *out)
{
    u32 base = 0U;
    for (u32 start = 0U; start < d.body_count; start += NPH_SCAN_BLOCK) {
        const u32 i = start + threadIdx.x;
        u32 mine = 0U;
        NphBodyRec *b = 0;
        if (i < d.body_count) {
            b = &d.bodies[d.awake_bodies[i]];
            if (b->enlarged != 0U) {
                for (u32 s = b->shape_head; s != NPH_NULL32; s = d.shapes[s].next) {
                    mine += (d.shapes[s].flags & NPH_SHAPE_ENLARGED) != 0U ? 1U : 0U;
                }
            }
        }
        u32 total = 0U;
        const u32 pos = block_exclusive_scan(mine, &total);
        if (mine != 0U) {
            u32 k = 0U;
            for (u32 s = b->shape_head; s != NPH_NULL32; s = d.shapes[s].next) {
                NphShapeRec *sh = &d.shapes[s];
                if ((sh->flags & NPH_SHAPE_ENLARGED) != 0U) {
                    NphGrown *g = &out[1U + base + pos + k];
                    g->shape = s;
                    g->pad = 0U;
                    g->fat_aabb = sh->fat_aabb;
                    sh->flags &= ~NPH_SHAPE_ENLARGED;
                    k += 1U;
                }
            }
            b->enlarged = 0U;
        }
        base += total;
    }
    if (threadIdx.x == 0U) {
        out[0].shape = base;
    }
}

static u32 blocks_for(u32 n) { return (n + NPH_BLOCK - 1U) / NPH_BLOCK; }

NphResult nph_dev_init_contacts(NphDevice *d, const NphContactInit *list, u32 count)
{
    if (count == 0U) {
        return NPH_RESULT_OK;
    }
    if (count > d->cap_inits) {
        const u32 c = grow_cap(d->cap_inits, count);
        NphResult r = grow_buffer((void **)&d->inits, 0, (u64)c * sizeof(NphContactInit), d->stream, 0);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        d->cap_inits = c;
    }
    HIP_TRY(hipMemcpyAsync(d->inits, list, (u64)count * sizeof *list, hipMemcpyHostToDevice,
                           d->stream));
    hipLaunchKernelGGL(k_init_contacts, dim3(blocks_for(count)), dim3(NPH_BLOCK), 0, d->stream,
                       d->contacts, d->inits, count);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(d->stream));
    return NPH_RESULT_OK;
}

NphResult nph_dev_collide(NphDevice *d, const u32 *awake, u32 count, f32 step_dt, u32 *changed,
                          u32 *codes, u32 *changed_count)
{
    *changed_count = 0U;
    d->collide_ms = 0.0;
    if (count == 0U) {
        return NPH_RESULT_OK;
    }
    NphResult r = reserve_list(d, count);
    if (r == NPH_RESULT_OK) {
        r = reserve_host_u32(d, 2U * count + 1U);
    }
    if (r != NPH_RESULT_OK) {
        return r;
    }
    HIP_TRY(hipMemcpyAsync(d->list, awake, (u64)count * 4U, hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipEventRecord(d->ev[0], d->stream));
    hipLaunchKernelGGL(k_collide, dim3((count + NPH_COLLIDE_BLOCK - 1U) / NPH_COLLIDE_BLOCK),
                       dim3(NPH_COLLIDE_BLOCK), 0, d->stream, d->contacts, d->shapes, d->bodies,
                       d->list, d->codes, count, step_dt);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_compact_changes, dim3(1), dim3(NPH_SCAN_BLOCK), 0, d->stream, d->list,
                       d->codes, count, d->compact);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(d->ev[1], d->stream));
    HIP_TRY(hipMemcpyAsync(d->h_u32, d->compact, 4U, hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    const u32 n = d->h_u32[0];
    if (n > 0U) {
        HIP_TRY(hipMemcpyAsync(changed, d->compact + 1U, (u64)n * 4U, hipMemcpyDeviceToHost,
                               d->stream));
        HIP_TRY(hipMemcpyAsync(codes, d->compact + 1U + count, (u64)n * 4U, hipMemcpyDeviceToHost,
                               d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
    }
    float ms = 0.0F;
    (void)hipEventElapsedTime(&ms, d->ev[0], d->ev[1]);
    d->collide_ms = ms;
    *changed_count = n;
    return NPH_RESULT_OK;
}

NphResult nph_dev_solve(NphDevice *d, const NphSolveInput *in, const NphGrown **grown,
                        u32 *grown_count)
{
    *grown = NULL;
    *grown_count = 0U;
    d->solve_ms = 0.0;
    const u32 n = in->awake_count;
    if (n == 0U) {
        return NPH_RESULT_OK;
    }
    // Groups: eight contacts of a colour each, one of the overflow each
    // (solver.c:810-843); colours 0..23 then the overflow.
    u32 first[NPH_GRAPH_COLORS + 1U];
    u32 count[NPH_GRAPH_COLORS + 1U];
    u32 groups = 0U;
    u32 graph = 0U;
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
        const u32 per = c == NPH_GRAPH_OVERFLOW ? 1U : NPH_LANES;
        first[c] = groups;
        count[c] = (in->color_count[c] + per - 1U) / per;
        groups += count[c];
        graph += in->color_count[c];
    }
    NphResult r = reserve_groups(d, groups > 0U ? groups : 1U);
    if (r == NPH_RESULT_OK) {
        r = reserve_host_u32(d, groups * (NPH_LANES + 1U) + 1U);
    }
    if (r == NPH_RESULT_OK) {
        r = reserve_list(d, graph > 0U ? graph : 1U);
    }
    if (r != NPH_RESULT_OK) {
        return r;
    }
    // The group table: lane contacts and lane counts, laid out on the host.
    u32 *lane_contact = d->h_u32;
    NphGroup *gtab = (NphGroup *)malloc((size_t)(groups > 0U ? groups : 1U) * sizeof(NphGroup));
    if (gtab == NULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    {
        u32 g = 0U;
        u32 at = 0U;
        for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
            const u32 per = c == NPH_GRAPH_OVERFLOW ? 1U : NPH_LANES;
            const u32 m = in->color_count[c];
            for (u32 i = 0U; i < m; i += per) {
                const u32 lanes = nph_minu(per, m - i);
                memset(&gtab[g], 0, sizeof gtab[g]);
                gtab[g].lane_count = lanes;
                for (u32 l = 0U; l < NPH_LANES; ++l) {
                    lane_contact[g * NPH_LANES + l] =
                        l < lanes ? in->graph_contacts[at + i + l] : NPH_NULL32;
                }
                g += 1U;
            }
            at += m;
        }
    }
    if (groups > 0U) {
        HIP_TRY(hipMemcpyAsync(d->lane_contact, lane_contact, (u64)groups * NPH_LANES * 4U,
                               hipMemcpyHostToDevice, d->stream));
        HIP_TRY(hipMemcpyAsync(d->groups, gtab, (u64)groups * sizeof(NphGroup),
                               hipMemcpyHostToDevice, d->stream));
    }
    HIP_TRY(hipMemcpyAsync(d->awake_bodies, in->awake_bodies, (u64)n * 4U, hipMemcpyHostToDevice,
                           d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream)); // the host buffers are reused below
    free(gtab);

    NphStepData sd;
    memset(&sd, 0, sizeof sd);
    sd.bodies = d->bodies;
    sd.shapes = d->shapes;
    sd.contacts = d->contacts;
    sd.solver_of_body = d->solver_of_body;
    sd.awake_bodies = d->awake_bodies;
    sd.sbodies = d->sbodies;
    sd.params = d->params;
    sd.groups = d->groups;
    sd.lane_contact = d->lane_contact;
    sd.lane_body_a = d->lane_body_a;
    sd.lane_body_b = d->lane_body_b;
    sd.lane_flags = d->lane_flags;
    sd.lanes = d->lanes;
    sd.points = d->points;
    sd.lane_stride = d->cap_groups * NPH_LANES;
    sd.point_stride = d->cap_point_rows * NPH_LANES;
    sd.body_count = n;
    const NphStepParams sp = in->params;
    hipStream_t s = d->stream;

    HIP_TRY(hipEventRecord(d->ev[0], s));
    hipLaunchKernelGGL(k_solver_map_fill, dim3(blocks_for(in->body_slots)), dim3(NPH_BLOCK), 0, s,
                       d->solver_of_body, in->body_slots, n);
    hipLaunchKernelGGL(k_solver_map_set, dim3(blocks_for(n)), dim3(NPH_BLOCK), 0, s,
                       d->solver_of_body, d->awake_bodies, n);
    hipLaunchKernelGGL(k_prepare_bodies, dim3(blocks_for(n + 1U)), dim3(NPH_BLOCK), 0, s, sd, sp);
    if (groups > 0U) {
        hipLaunchKernelGGL(k_group_points, dim3(blocks_for(groups)), dim3(NPH_BLOCK), 0, s, sd,
                           groups);
        hipLaunchKernelGGL(k_group_scan, dim3(1), dim3(NPH_SCAN_BLOCK), 0, s, sd, groups);
        hipLaunchKernelGGL(k_lane_prepare, dim3(blocks_for(groups * NPH_LANES)), dim3(NPH_BLOCK),
                           0, s, sd, sp, groups);
        hipLaunchKernelGGL(k_group_impact, dim3(blocks_for(groups)), dim3(NPH_BLOCK), 0, s, sd,
                           groups);
    }
    HIP_TRY(hipGetLastError());
    // One constraint pass: the overflow first, then colours 0..23 (solver.c:774-789).
#define PASS(kind)                                                                                 \
    do {                                                                                           \
        if (count[NPH_GRAPH_OVERFLOW] > 0U) {                                                      \
            hipLaunchKernelGGL(k_overflow_pass, dim3(1), dim3(64), 0, s, sd, sp, (u32)(kind),      \
                               first[NPH_GRAPH_OVERFLOW], count[NPH_GRAPH_OVERFLOW]);               \
        }                                                                                          \
        for (u32 c = 0U; c < NPH_GRAPH_COLORS; ++c) {                                              \
            if (count[c] > 0U) {                                                                   \
                hipLaunchKernelGGL(k_color_pass, dim3(blocks_for(count[c] * NPH_LANES)),           \
                                   dim3(NPH_BLOCK), 0, s, sd, sp, (u32)(kind), first[c], count[c]); \
            }                                                                                      \
        }                                                                                          \
    } while (0)
    HIP_TRY(hipEventRecord(d->ev[2], s));
    for (u32 step = 0U; step < sp.substeps; ++step) {
        hipLaunchKernelGGL(k_integrate_velocities, dim3(blocks_for(n)), dim3(NPH_BLOCK), 0, s, sd,
                           sp);
        PASS(PASS_WARM);
        PASS(PASS_SOLVE);
        hipLaunchKernelGGL(k_integrate_positions, dim3(blocks_for(n)), dim3(NPH_BLOCK), 0, s, sd,
                           sp);
        PASS(PASS_RELAX);
    }
    if (groups > 0U) {
        hipLaunchKernelGGL(k_restitution_masks, dim3(blocks_for(groups)), dim3(NPH_BLOCK), 0, s,
                           sd, sp.restitution_threshold, groups);
        PASS(PASS_RESTITUTION);
    }
    HIP_TRY(hipEventRecord(d->ev[3], s));
    if (groups > 0U) {
        hipLaunchKernelGGL(k_store, dim3(blocks_for(groups * NPH_LANES)), dim3(NPH_BLOCK), 0, s, sd,
                           groups);
    }
#undef PASS
    hipLaunchKernelGGL(k_finalize, dim3(blocks_for(n)), dim3(NPH_BLOCK), 0, s, sd, sp);
    hipLaunchKernelGGL(k_collect_grown, dim3(1), dim3(NPH_SCAN_BLOCK), 0, s, sd, d->grown);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(d->ev[1], s));
    // The count, then the records.
    if (d->cap_h_grown < 1U) {
        HIP_TRY(hipHostMalloc((void **)&d->h_grown, 64U * sizeof(NphGrown), hipHostMallocDefault));
        d->cap_h_grown = 64U;
    }
    HIP_TRY(hipMemcpyAsync(d->h_grown, d->grown, sizeof(NphGrown), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const u32 m = d->h_grown[0].shape;
    if (m + 1U > d->cap_h_grown) {
        const u32 c = grow_cap(d->cap_h_grown, m + 1U);
        HIP_TRY(hipHostFree(d->h_grown));
        d->h_grown = NULL;
        HIP_TRY(hipHostMalloc((void **)&d->h_grown, (u64)c * sizeof(NphGrown), hipHostMallocDefault));
        d->cap_h_grown = c;
    }
    if (m > 0U) {
        HIP_TRY(hipMemcpyAsync(d->h_grown, d->grown, (u64)(m + 1U) * sizeof(NphGrown),
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    float ms = 0.0F;
    (void)hipEventElapsedTime(&ms, d->ev[0], d->ev[1]);
    d->solve_ms = ms;
    ms = 0.0F;
    (void)hipEventElapsedTime(&ms, d->ev[2], d->ev[3]);
    d->passes_ms = ms;
    *grown = d->h_grown + 1;
    *grown_count = m;
    return NPH_RESULT_OK;
}
