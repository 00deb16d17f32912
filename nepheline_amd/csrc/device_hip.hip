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



#include "device_impl.h"
#include "device_narrow.h"

#include <hip/hip_ext.h>

const char *nph_dev_name(void) { return "hip:gfx950"; }

NphResult devi_grow_buffer(void **p, u64 old_bytes, u64 new_bytes, hipStream_t s, int keep)
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

u32 devi_grow_cap(u32 cap, u32 need)
{
    u32 c = cap < 64U ? 64U : cap;
    while (c < need) {
        c = c * 2U;
    }
    return c;
}

NphResult nph_dev_create(NphDevice **out) { return devi_create(out, NULL); }

NphResult devi_create(NphDevice **out, hipStream_t stream)
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
    if (stream != NULL) {
        d->stream = stream;
        d->shared_stream = 1;
    } else if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
        free(d);
        return NPH_RESULT_DEVICE_LOST;
    }
    for (u32 i = 0U; i < 4U; ++i) {
        if (hipEventCreate(&d->ev[i]) != hipSuccess) {
            free(d);
            return NPH_RESULT_DEVICE_LOST;
        }
    }
    for (u32 i = 0U; i < 4U; ++i) {
        if (hipEventCreate(&d->cev[i]) != hipSuccess) {
            free(d);
            return NPH_RESULT_DEVICE_LOST;
        }
    }
    d->staged_bodies.rec_words = (u32)(sizeof(NphBodyRec) / 4U);
    d->staged_shapes.rec_words = (u32)(sizeof(NphShapeRec) / 4U);
    d->staged_jconf.rec_words = (u32)(sizeof(NphJointConf) / 4U);
    d->staged_jstate.rec_words = (u32)(sizeof(NphJointState) / 4U);
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
                    d->sbodies, d->params, d->list, d->codes, d->inits, d->groups,
                    d->lane_contact, d->lane_body_a, d->lane_body_b, d->lane_flags, d->lanes,
                    d->points, d->grown, d->grown_blocks, d->joint_conf, d->joint_state, d->step_joints,
                    d->fast_list, d->fast_pack, d->gate, d->mlp_params, d->lx, d->lraw, d->ly,
                    d->lres, d->wstats, d->log_x, d->log_t, d->log_count, d->team_small,
                    d->team_big, d->team_mesh, d->team_counts, d->pf_nodes[0], d->pf_nodes[1],
                    d->pf_nodes[2], d->pf_shapes, d->pf_stamp, d->pf_set, d->contacts_saved};
    for (u32 i = 0U; i < sizeof bufs / sizeof bufs[0]; ++i) {
        if (bufs[i] != NULL) {
            (void)hipFree(bufs[i]);
        }
    }
    for (u32 i = 0U; i < d->hull_count; ++i) {
        (void)hipFree((void *)d->hulls[i].device);
    }
    free(d->hulls);
    Staged *st[4] = {&d->staged_bodies, &d->staged_shapes, &d->staged_jconf, &d->staged_jstate};
    for (u32 i = 0U; i < 4U; ++i) {
        free(st[i]->recs);
        free(st[i]->slots);
        free(st[i]->where);
        if (st[i]->dev != NULL) {
            (void)hipFree(st[i]->dev);
        }
    }
    if (d->h_u32 != NULL) {
        (void)hipHostFree(d->h_u32);
    }
    void *pinned[] = {d->h_list, d->h_codes, d->h_inits, d->h_fast_count, d->h_fast_pack,
                      d->h_wstats};
    for (u32 i = 0U; i < 6U; ++i) {
        if (pinned[i] != NULL) {
            (void)hipHostFree(pinned[i]);
        }
    }
    if (d->h_grown != NULL) {
        (void)hipHostFree(d->h_grown);
    }
    for (u32 i = 0U; i < 4U; ++i) {
        (void)hipEventDestroy(d->ev[i]);
        (void)hipEventDestroy(d->cev[i]);
    }
    for (u32 i = 0U; i < 2U * d->kev_pairs; ++i) {
        (void)hipEventDestroy(d->kev[i]);
    }
    free(d->kev);
    if (!d->shared_stream) {
        (void)hipStreamDestroy(d->stream);
    }
    free(d);
}

f64 nph_dev_last_solve_ms(const NphDevice *d) { return d->solve_ms; }
f64 nph_dev_last_collide_ms(const NphDevice *d) { return d->collide_ms; }
f64 nph_dev_last_passes_ms(const NphDevice *d) { return d->passes_ms; }

void nph_dev_time_kernels(NphDevice *d, int enable)
{
    d->time_kernels = enable;
    d->kernel_ms = 0.0;
    d->kernel_launches = 0U;
}

void nph_dev_kernel_times(const NphDevice *d, f64 out[2])
{
    out[0] = d->kernel_ms;
    out[1] = (f64)d->kernel_launches;
}

NphResult nph_dev_reserve(NphDevice *d, u32 bodies, u32 shapes, u32 contacts)
{
    if (bodies > d->cap_bodies) {
        const u32 c = devi_grow_cap(d->cap_bodies, bodies);
        NphResult r;
        if ((r = devi_grow_buffer((void **)&d->bodies, (u64)d->cap_bodies * sizeof(NphBodyRec),
                             (u64)c * sizeof(NphBodyRec), d->stream, 1)) != NPH_RESULT_OK ||
            (r = devi_grow_buffer((void **)&d->solver_of_body, 0, (u64)c * 4U, d->stream, 0)) !=
                NPH_RESULT_OK ||
            (r = devi_grow_buffer((void **)&d->awake_bodies, 0, (u64)c * 4U, d->stream, 0)) !=
                NPH_RESULT_OK ||
            (r = devi_grow_buffer((void **)&d->sbodies, 0, (u64)(c + 1U) * sizeof(NphSolverBody),
                             d->stream, 0)) != NPH_RESULT_OK ||
            (r = devi_grow_buffer((void **)&d->params, 0, (u64)c * sizeof(NphSolverParams), d->stream,
                             0)) != NPH_RESULT_OK ||
            (r = devi_grow_buffer((void **)&d->grown_blocks, 0, (u64)(c / NPH_BLOCK + 1U) * 4U,
                             d->stream, 0)) != NPH_RESULT_OK ||
            (r = devi_grow_buffer((void **)&d->fast_list, 0, (u64)(c + 1U) * 4U, d->stream, 0)) !=
                NPH_RESULT_OK) {
            return r;
        }
        d->cap_bodies = c;
    }
    if (shapes > d->cap_shapes) {
        const u32 c = devi_grow_cap(d->cap_shapes, shapes);
        NphResult r;
        if ((r = devi_grow_buffer((void **)&d->shapes, (u64)d->cap_shapes * sizeof(NphShapeRec),
                             (u64)c * sizeof(NphShapeRec), d->stream, 1)) != NPH_RESULT_OK ||
            (r = devi_grow_buffer((void **)&d->grown, 0, (u64)(c + 1U) * sizeof(NphGrown), d->stream,
                             0)) != NPH_RESULT_OK) {
            return r;
        }
        d->cap_shapes = c;
    }
    if (contacts > d->cap_contacts) {
        const u32 c = devi_grow_cap(d->cap_contacts, contacts);
        NphResult r = devi_grow_buffer((void **)&d->contacts, (u64)d->cap_contacts * sizeof(NphContactRec),
                                  (u64)c * sizeof(NphContactRec), d->stream, 1);
        if (r == NPH_RESULT_OK) {
            r = devi_grow_buffer((void **)&d->gate, (u64)d->cap_contacts * 4U, (u64)c * 4U, d->stream, 1);
        }
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
    const u32 c = devi_grow_cap(d->cap_list, n);
    NphResult r;
    if ((r = devi_grow_buffer((void **)&d->list, 0, (u64)c * 4U, d->stream, 0)) != NPH_RESULT_OK ||
        (r = devi_grow_buffer((void **)&d->codes, 0, (u64)c * 4U, d->stream, 0)) != NPH_RESULT_OK) {
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
    const u32 c = devi_grow_cap(d->cap_h_u32, n);
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
    const u32 c = devi_grow_cap(d->cap_groups, groups);
    const u64 lanes = (u64)c * NPH_LANES;
    const u64 rows = (u64)c * NPH_PHYSICS_MANIFOLD_MAX_POINTS;
    NphResult r;
    if ((r = devi_grow_buffer((void **)&d->groups, 0, (u64)c * sizeof(NphGroup), d->stream, 0)) !=
            NPH_RESULT_OK ||
        (r = devi_grow_buffer((void **)&d->lane_contact, 0, lanes * 4U, d->stream, 0)) != NPH_RESULT_OK ||
        (r = devi_grow_buffer((void **)&d->lane_body_a, 0, lanes * 4U, d->stream, 0)) != NPH_RESULT_OK ||
        (r = devi_grow_buffer((void **)&d->lane_body_b, 0, lanes * 4U, d->stream, 0)) != NPH_RESULT_OK ||
        (r = devi_grow_buffer((void **)&d->lane_flags, 0, lanes * 4U, d->stream, 0)) != NPH_RESULT_OK ||
        (r = devi_grow_buffer((void **)&d->lanes, 0, lanes * NPH_L_FIELDS * 4U, d->stream, 0)) !=
            NPH_RESULT_OK ||
        (r = devi_grow_buffer((void **)&d->points, 0, rows * NPH_LANES * NPH_P_FIELDS * 4U, d->stream,
                         0)) != NPH_RESULT_OK) {
        return r;
    }
    d->cap_groups = c;
    d->cap_point_rows = (u32)rows;
    return NPH_RESULT_OK;
}

NphResult nph_dev_block(NphDevice *d, const void *block, u32 bytes, const void **out)
{
    for (u32 i = 0U; i < d->hull_count; ++i) {
        if (d->hulls[i].host == block) {
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
    HIP_TRY(hipMalloc(&p, bytes));
    HIP_TRY(hipMemcpy(p, block, bytes, hipMemcpyHostToDevice));
    d->hulls[d->hull_count].host = block;
    d->hulls[d->hull_count].device = p;
    d->hull_count += 1U;
    *out = p;
    return NPH_RESULT_OK;
}

static_assert(sizeof(NphBodyRec) % 4U == 0U && sizeof(NphShapeRec) % 4U == 0U &&
                  sizeof(NphJointConf) % 4U == 0U && sizeof(NphJointState) % 4U == 0U,
              "records are copied as 4-byte words");

static NphResult stage_put(Staged *st, u32 slot, const void *rec)
{
    if (slot >= st->cap_where) {
        const u32 c = devi_grow_cap(st->cap_where, slot + 1U);
        u32 *n = (u32 *)realloc(st->where, (size_t)c * 4U);
        if (n == NULL) {
            return NPH_RESULT_OUT_OF_MEMORY;
        }
        memset(n + st->cap_where, 0, (size_t)(c - st->cap_where) * 4U);
        st->where = n;
        st->cap_where = c;
    }
    u32 i = st->where[slot];
    if (i == 0U) {
        if (st->count == st->cap) {
            const u32 c = devi_grow_cap(st->cap, st->count + 1U);
            u32 *r = (u32 *)realloc(st->recs, (size_t)c * st->rec_words * 4U);
            if (r == NULL) {
                return NPH_RESULT_OUT_OF_MEMORY;
            }
            st->recs = r;
            u32 *q = (u32 *)realloc(st->slots, (size_t)c * 4U);
            if (q == NULL) {
                return NPH_RESULT_OUT_OF_MEMORY;
            }
            st->slots = q;
            st->cap = c;
        }
        i = ++st->count;
        st->where[slot] = i;
        st->slots[i - 1U] = slot;
    }
    memcpy(st->recs + (u64)(i - 1U) * st->rec_words, rec, (size_t)st->rec_words * 4U);
    return NPH_RESULT_OK;
}

// Record i of the staged list to its slot, one word a thread.
__global__ void __launch_bounds__(NPH_BLOCK) k_scatter_records(u32 *dst, const u32 *staged,
                                                               u32 count, u32 words)
{
    const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (u64)count * words) {
        const u32 i = (u32)(t / words);
        const u32 w = (u32)(t % words);
        dst[(u64)staged[i] * words + w] = staged[count + (u64)i * words + w];
    }
}

static NphResult stage_flush(NphDevice *d, Staged *st, void *dst)
{
    if (st->count == 0U) {
        return NPH_RESULT_OK;
    }
    const u64 words = (u64)st->count * (1U + st->rec_words);
    if (words > st->dev_words) {
        u64 c = st->dev_words < 4096U ? 4096U : st->dev_words;
        while (c < words) {
            c *= 2U;
        }
        NphResult r = devi_grow_buffer((void **)&st->dev, 0, c * 4U, d->stream, 0);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        st->dev_words = c;
    }
    HIP_TRY(hipMemcpyAsync(st->dev, st->slots, (u64)st->count * 4U, hipMemcpyHostToDevice,
                           d->stream));
    HIP_TRY(hipMemcpyAsync(st->dev + st->count, st->recs, (u64)st->count * st->rec_words * 4U,
                           hipMemcpyHostToDevice, d->stream));
    const u64 threads = (u64)st->count * st->rec_words;
    hipLaunchKernelGGL(k_scatter_records, dim3((u32)((threads + NPH_BLOCK - 1U) / NPH_BLOCK)),
                       dim3(NPH_BLOCK), 0, d->stream, (u32 *)dst, st->dev, st->count,
                       st->rec_words);
    HIP_TRY(hipGetLastError());
    // The host lists are reused at once.
    HIP_TRY(hipStreamSynchronize(d->stream));
    for (u32 i = 0U; i < st->count; ++i) {
        st->where[st->slots[i]] = 0U;
    }
    st->count = 0U;
    return NPH_RESULT_OK;
}

// Every device read and stage starts here.
NphResult devi_flush_staged(NphDevice *d)
{
    NphResult r = stage_flush(d, &d->staged_bodies, d->bodies);
    if (r == NPH_RESULT_OK) {
        r = stage_flush(d, &d->staged_shapes, d->shapes);
    }
    if (r == NPH_RESULT_OK) {
        r = stage_flush(d, &d->staged_jconf, d->joint_conf);
    }
    if (r == NPH_RESULT_OK) {
        r = stage_flush(d, &d->staged_jstate, d->joint_state);
    }
    return r;
}

NphResult nph_dev_reserve_joints(NphDevice *d, u32 joints)
{
    if (joints <= d->cap_joints) {
        return NPH_RESULT_OK;
    }
    const u32 c = devi_grow_cap(d->cap_joints, joints);
    NphResult r;
    if ((r = devi_grow_buffer((void **)&d->joint_conf, (u64)d->cap_joints * sizeof(NphJointConf),
                         (u64)c * sizeof(NphJointConf), d->stream, 1)) != NPH_RESULT_OK ||
        (r = devi_grow_buffer((void **)&d->joint_state, (u64)d->cap_joints * sizeof(NphJointState),
                         (u64)c * sizeof(NphJointState), d->stream, 1)) != NPH_RESULT_OK) {
        return r;
    }
    d->cap_joints = c;
    return NPH_RESULT_OK;
}

NphResult nph_dev_write_joint(NphDevice *d, u32 slot, const NphJointConf *conf, int reset)
{
    if (slot >= d->cap_joints) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    NphResult r = NPH_RESULT_OK;
    if (conf != NULL) {
        r = stage_put(&d->staged_jconf, slot, conf);
    }
    if (r == NPH_RESULT_OK && reset) {
        NphJointState zero;
        memset(&zero, 0, sizeof zero);
        r = stage_put(&d->staged_jstate, slot, &zero);
    }
    return r;
}

NphResult nph_dev_read_joint(NphDevice *d, u32 slot, NphJointState *out)
{
    if (slot >= d->cap_joints) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    NphResult r = devi_flush_staged(d);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    HIP_TRY(hipMemcpyAsync(out, &d->joint_state[slot], sizeof *out, hipMemcpyDeviceToHost,
                           d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return NPH_RESULT_OK;
}

NphResult nph_dev_write_body(NphDevice *d, u32 slot, const NphBodyRec *b)
{
    if (slot >= d->cap_bodies) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    return stage_put(&d->staged_bodies, slot, b);
}

NphResult nph_dev_write_shape(NphDevice *d, u32 slot, const NphShapeRec *s)
{
    if (slot >= d->cap_shapes) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    return stage_put(&d->staged_shapes, slot, s);
}

NphResult nph_dev_read_bodies(NphDevice *d, NphBodyRec *out, u32 count)
{
    {
        const NphResult r = devi_flush_staged(d);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    if (count > 0U) {
        HIP_TRY(hipMemcpyAsync(out, d->bodies, (u64)count * sizeof *out, hipMemcpyDeviceToHost,
                               d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
    }
    return NPH_RESULT_OK;
}

NphResult nph_dev_read_shapes(NphDevice *d, NphShapeRec *out, u32 count)
{
    {
        const NphResult r = devi_flush_staged(d);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    if (count > 0U) {
        HIP_TRY(hipMemcpyAsync(out, d->shapes, (u64)count * sizeof *out, hipMemcpyDeviceToHost,
                               d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
    }
    return NPH_RESULT_OK;
}

NphResult nph_dev_read_contacts(NphDevice *d, NphContactRec *out, u32 count)
{
    {
        const NphResult r = devi_flush_staged(d);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    if (count > 0U) {
        HIP_TRY(hipMemcpyAsync(out, d->contacts, (u64)count * sizeof *out, hipMemcpyDeviceToHost,
                               d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
    }
    return NPH_RESULT_OK;
}

// The first `count` contact records copied aside (stream-ordered), for
// nph_dev_contacts_restore.
NphResult nph_dev_contacts_checkpoint(NphDevice *d, u32 count)
{
    const NphResult r = devi_flush_staged(d);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    d->contacts_saved_count = 0U;
    if (count == 0U) {
        return NPH_RESULT_OK;
    }
    if (count > d->cap_contacts_saved) {
        const u32 cap = devi_grow_cap(d->cap_contacts_saved, count);
        const NphResult g = devi_grow_buffer((void **)&d->contacts_saved, 0U,
                                             (u64)cap * sizeof(NphContactRec), d->stream, 0);
        if (g != NPH_RESULT_OK) {
            return g;
        }
        d->cap_contacts_saved = cap;
    }
    HIP_TRY(hipMemcpyAsync(d->contacts_saved, d->contacts, (u64)count * sizeof(NphContactRec),
                           hipMemcpyDeviceToDevice, d->stream));
    d->contacts_saved_count = count;
    return NPH_RESULT_OK;
}

// The records the last checkpoint saved put back, after the device's work
// so far (the narrow phase queued since) has run.
NphResult nph_dev_contacts_restore(NphDevice *d)
{
    if (d->contacts_saved_count > 0U) {
        HIP_TRY(hipMemcpyAsync(d->contacts, d->contacts_saved,
                               (u64)d->contacts_saved_count * sizeof(NphContactRec),
                               hipMemcpyDeviceToDevice, d->stream));
    }
    HIP_TRY(hipStreamSynchronize(d->stream));
    return NPH_RESULT_OK;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(NPH_BLOCK) k_init_contacts(NphContactRec *contacts, u32 *gate, const NphContactInit *list, u32 n)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const NphContactInit in = list[i];
    init_contact_record(contacts, gate, &in);
}

// The convex narrow phase (device_narrow.h): lanes, then teams.
__global__ void __launch_bounds__(64) k_narrow_clear(u32 *counts) { narrow_counts_clear(counts); }

__global__ void __launch_bounds__(NPH_BLOCK) k_narrow_lanes(NphOneWorld w, const u32 *list, u32 *codes, u32 n,
                                                            f32 step_dt, NphTeamLists lists)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    NphSortedEntry e;
    e.world = 0U;
    e.record = i < n ? list[i] : 0U;
    e.index = i;
    e.pad = 0U;
    narrow_lane(i < n, e, w, codes, step_dt, lists);
}

__global__ void __launch_bounds__(NPH_TEAM_BLOCK) k_narrow_teams(NphOneWorld w, const NphSortedEntry *list,
                                                                 const u32 *count, u32 *codes,
                                                                 f32 step_dt)
{
    narrow_teams<NPH_HULL_MAX_VERTICES, NPH_CLIP_CAPACITY, NPH_TEAM_BLOCK>(list, count, w, codes,
                                                                            step_dt);
}

__global__ void __launch_bounds__(NPH_BIG_TEAM_BLOCK) k_narrow_big_teams(NphOneWorld w, const NphSortedEntry *list,
                                                                         const u32 *count, u32 *codes,
                                                                         f32 step_dt)
{
    narrow_teams<NPH_BIG_HULL_VERTICES, NPH_BIG_CLIP_CAPACITY, NPH_BIG_TEAM_BLOCK>(
        list, count, w, codes, step_dt);
}

// The mesh pairs the lane pass listed, a team a pair.
__global__ void __launch_bounds__(NPH_MESH_TEAM_BLOCK) k_narrow_mesh_teams(NphOneWorld w, const NphSortedEntry *list,
                                                                           const u32 *count, u32 *codes,
                                                                           f32 step_dt)
{
    narrow_mesh_teams<NPH_CLIP_CAPACITY, NPH_MESH_TEAM_BLOCK>(list, count, w, codes, step_dt);
}

__global__ void __launch_bounds__(NPH_BIG_MESH_TEAM_BLOCK) k_narrow_big_mesh_teams(NphOneWorld w, const NphSortedEntry *list,
                                                                                   const u32 *count, u32 *codes,
                                                                                   f32 step_dt)
{
    narrow_mesh_teams<NPH_BIG_CLIP_CAPACITY, NPH_BIG_MESH_TEAM_BLOCK>(list, count, w, codes,
                                                                      step_dt);
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

// The step's joints prepared (solver.c:1036, joint_solver.c:726-824).
__global__ void __launch_bounds__(NPH_BLOCK) k_joint_prepare(NphStepData d, NphStepParams sp)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.joint_count) {
        const u32 slot = d.step_joints[i];
        const NphJointConf *cf = &d.joint_conf[slot];
        joint_prepare(cf, &d.joint_state[slot], &d.bodies[cf->body_a], &d.bodies[cf->body_b],
                      d.solver_of_body[cf->body_a], d.solver_of_body[cf->body_b], d.params, sp.h);
    }
}

// One colour's joints [jfirst, jfirst + jcount) and groups [first, first +
// count): the joints take the first threads (rounded up to whole waves),
// the groups' lanes the rest. A colour's constraints share no dynamic body,
// so their order inside the launch is free.
__global__ void __launch_bounds__(NPH_BLOCK) k_color_pass(NphStepData d, NphStepParams sp, u32 pass, u32 first, u32 count, u32 jfirst, u32 jcount)
{
    color_pass_thread(&d, &sp, pass, first, count, jfirst, jcount,
                      blockIdx.x * blockDim.x + threadIdx.x);
}

// The overflow: its joints, then its groups (one contact each), in order, on
// one thread (solver.c:771-779).
__global__ void __launch_bounds__(64) k_overflow_pass(NphStepData d, NphStepParams sp, u32 pass, u32 first, u32 count, u32 jfirst, u32 jcount)
{
    if (blockIdx.x == 0U && threadIdx.x == 0U) {
        overflow_pass(&d, &sp, pass, first, count, jfirst, jcount);
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
// list's order; their flags cleared.  Two launches: a count per block of
// NPH_BLOCK awake bodies, then each block places its bodies' shapes after the
// counts of the blocks before it.  out[0].shape = count.
__device__ static u32 block_sum(u32 v)
{
    __shared__ u32 part[NPH_BLOCK / 64U];
    const u32 lane = threadIdx.x & 63U;
    const u32 wave = threadIdx.x >> 6U;
    for (u32 off = 32U; off > 0U; off >>= 1U) {
        v += __shfl_down(v, off, 64);
    }
    if (lane == 0U) {
        part[wave] = v;
    }
    __syncthreads();
    u32 total = 0U;
    for (u32 w = 0U; w < NPH_BLOCK / 64U; ++w) {
        total += part[w];
    }
    __syncthreads();
    return total;
}

__global__ void __launch_bounds__(NPH_BLOCK) k_grown_count(NphStepData d, u32 *block_counts)
{
    const u32 i = blockIdx.x * NPH_BLOCK + threadIdx.x;
    const u32 total = block_sum(i < d.body_count ? grown_of_body(&d, i) : 0U);
    if (threadIdx.x == 0U) {
        block_counts[blockIdx.x] = total;
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_grown_scatter(NphStepData d,
                                                             const u32 *block_counts, NphGrown *out)
{
    __shared__ u32 scan[NPH_BLOCK];
    u32 before = 0U;
    for (u32 j = threadIdx.x; j < blockIdx.x; j += NPH_BLOCK) {
        before += block_counts[j];
    }
    const u32 base = block_sum(before);
    const u32 i = blockIdx.x * NPH_BLOCK + threadIdx.x;
    const u32 mine = i < d.body_count ? grown_of_body(&d, i) : 0U;
    scan[threadIdx.x] = mine;
    __syncthreads();
    for (u32 off = 1U; off < NPH_BLOCK; off <<= 1U) {
        const u32 x = threadIdx.x >= off ? scan[threadIdx.x - off] : 0U;
        __syncthreads();
        scan[threadIdx.x] += x;
        __syncthreads();
    }
    const u32 pos = base + scan[threadIdx.x] - mine;
    if (mine != 0U) {
        grown_write(&d, i, &out[1U + pos]);
    }
    if (blockIdx.x == gridDim.x - 1U && threadIdx.x == NPH_BLOCK - 1U) {
        out[0].shape = base + scan[NPH_BLOCK - 1U];
    }
}

// The fast bodies (continuous collision) in the awake list's order: counts
// per block, then placement; out[0] = count, then awake indices.
__global__ void __launch_bounds__(NPH_BLOCK) k_fast_count(NphStepData d, u32 *block_counts)
{
    const u32 i = blockIdx.x * NPH_BLOCK + threadIdx.x;
    const u32 fast = i < d.body_count &&
                     (d.bodies[d.awake_bodies[i]].flags & NPH_BODY_FAST) != 0U ? 1U : 0U;
    const u32 total = block_sum(fast);
    if (threadIdx.x == 0U) {
        block_counts[blockIdx.x] = total;
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_fast_scatter(NphStepData d,
                                                            const u32 *block_counts, u32 *out)
{
    __shared__ u32 scan[NPH_BLOCK];
    u32 before = 0U;
    for (u32 j = threadIdx.x; j < blockIdx.x; j += NPH_BLOCK) {
        before += block_counts[j];
    }
    const u32 base = block_sum(before);
    const u32 i = blockIdx.x * NPH_BLOCK + threadIdx.x;
    const u32 mine = i < d.body_count &&
                     (d.bodies[d.awake_bodies[i]].flags & NPH_BODY_FAST) != 0U ? 1U : 0U;
    scan[threadIdx.x] = mine;
    __syncthreads();
    for (u32 off = 1U; off < NPH_BLOCK; off <<= 1U) {
        const u32 x = threadIdx.x >= off ? scan[threadIdx.x - off] : 0U;
        __syncthreads();
        scan[threadIdx.x] += x;
        __syncthreads();
    }
    if (mine != 0U) {
        out[1U + base + scan[threadIdx.x] - 1U] = i;
    }
    if (blockIdx.x == gridDim.x - 1U && threadIdx.x == NPH_BLOCK - 1U) {
        out[0] = base + scan[NPH_BLOCK - 1U];
    }
}

// Each fast body's record and solver rows, for the host; its flag cleared
// (the sweep is done once).
__global__ void __launch_bounds__(NPH_BLOCK) k_fast_pack(NphStepData d, const u32 *list, u32 count,
                                                         NphFastBody *out)
{
    const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) {
        const u32 i = list[1U + k];
        const u32 slot = d.awake_bodies[i];
        NphBodyRec *b = &d.bodies[slot];
        b->flags &= ~(u32)NPH_BODY_FAST;
        out[k].body = *b;
        out[k].sb = d.sbodies[i];
        out[k].params = d.params[i];
        out[k].index = i;
        out[k].slot = slot;
    }
}

// ---------------------------------------------------------------------------
// The learned warm start
// ---------------------------------------------------------------------------

// The learned warm start's kernels: each a thread a lane of the one world.
__global__ void __launch_bounds__(NPH_BLOCK) k_learned_features(NphStepData d, NphLearnArgs a)
{
    const u32 L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L < a.group_count * NPH_LANES) {
        learned_features_at(&d, &a, L);
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_learned_apply(NphStepData d, NphLearnArgs a)
{
    const u32 L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L < a.group_count * NPH_LANES) {
        learned_apply_at(&d, &a, L);
    }
}

__global__ void __launch_bounds__(64) k_learned_clear(u32 *stats)
{
    if (threadIdx.x < WS_CLEARED) {
        stats[threadIdx.x] = 0U;
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_learned_residual(NphStepData d, NphLearnArgs a)
{
    const u32 L = blockIdx.x * blockDim.x + threadIdx.x;
    learned_residual_at(&d, &a, L, L < a.group_count * NPH_LANES);
}

__global__ void __launch_bounds__(64) k_learned_threshold(u32 *stats)
{
    if (threadIdx.x == 0U) {
        learned_threshold(stats);
    }
}

__global__ void __launch_bounds__(NPH_BLOCK) k_learned_gate(NphStepData d, NphLearnArgs a)
{
    const u32 L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L < a.group_count * NPH_LANES) {
        learned_gate_at(&d, &a, L);
    }
}

// Appends every point's raw features and converged impulses to the log
// (one counter update a wave).
__global__ void __launch_bounds__(NPH_BLOCK) k_learned_log(NphStepData d, NphLearnArgs a)
{
    const u32 L = blockIdx.x * blockDim.x + threadIdx.x;
    const NphContactRec *c = NULL;
    const u32 points =
        (L < a.group_count * NPH_LANES && lane_active(&d, &a, L, &c)) ? c->point_count : 0U;
    // Inclusive scan of the points over the wave.
    const u32 lane = threadIdx.x & 63U;
    u32 incl = points;
    for (u32 o = 1U; o < 64U; o <<= 1U) {
        const u32 v = (u32)__shfl_up((int)incl, o, 64);
        if (lane >= o) {
            incl += v;
        }
    }
    const u32 total = (u32)__shfl((int)incl, 63, 64);
    u32 base = 0U;
    if (lane == 63U && total > 0U) {
        base = atomicAdd(a.log_count, total);
    }
    base = (u32)__shfl((int)base, 63, 64);
    if (points == 0U) {
        return;
    }
    const u32 g = L / NPH_LANES;
    const u32 l = L % NPH_LANES;
    for (u32 k = 0U; k < points; ++k) {
        const u32 at = base + incl - points + k;
        if (at >= a.log_cap) {
            return;
        }
        const u64 R = (u64)(d.groups[g].first_point + k) * NPH_LANES + l;
        const f32 *fr = a.raw + ((u64)L * NPH_LEARNED_ROWS_PER_LANE + k) * NPH_LEARNED_IN;
        f32 *xo = a.log_x + (u64)at * NPH_LEARNED_IN;
        for (u32 i = 0U; i < NPH_LEARNED_IN; ++i) {
            xo[i] = fr[i];
        }
        f32 *to = a.log_t + (u64)at * NPH_LEARNED_OUT;
        to[0] = PF(&d, NPH_P_NIMP, R);
        to[1] = PF(&d, NPH_P_TI1, R);
        to[2] = PF(&d, NPH_P_TI2, R);
        to[3] = fr[15];
    }
}

static u32 blocks_for(u32 n) { return (n + NPH_BLOCK - 1U) / NPH_BLOCK; }

static NphResult learned_before(NphDevice *d, const NphStepData *sd, u32 groups, f32 inv_h);
static NphResult learned_after(NphDevice *d, const NphStepData *sd, u32 groups, f32 inv_h);

// A pinned host buffer of at least `bytes`; the stream is drained before an
// old one is freed (a queued copy may still read it).
NphResult devi_reserve_pinned(NphDevice *d, void **p, u64 *cap, u64 bytes)
{
    if (bytes <= *cap) {
        return NPH_RESULT_OK;
    }
    u64 c = *cap < 4096U ? 4096U : *cap;
    while (c < bytes) {
        c *= 2U;
    }
    if (*p != NULL) {
        HIP_TRY(hipStreamSynchronize(d->stream));
        HIP_TRY(hipHostFree(*p));
        *p = NULL;
        *cap = 0U;
    }
    HIP_TRY(hipHostMalloc(p, c, hipHostMallocDefault));
    *cap = c;
    return NPH_RESULT_OK;
}

// Queued, not waited for: the records are staged in pinned memory, so the
// caller's list may change at once.
NphResult nph_dev_init_contacts(NphDevice *d, const NphContactInit *list, u32 count)
{
    {
        const NphResult fr = devi_flush_staged(d);
        if (fr != NPH_RESULT_OK) {
            return fr;
        }
    }
    if (count == 0U) {
        return NPH_RESULT_OK;
    }
    if (count > d->cap_inits) {
        const u32 c = devi_grow_cap(d->cap_inits, count);
        NphResult r = devi_grow_buffer((void **)&d->inits, 0, (u64)c * sizeof(NphContactInit), d->stream, 0);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        d->cap_inits = c;
    }
    // The previous step's copy out of the staging buffer is done: every step
    // ends with the stream drained.
    NphResult r = devi_reserve_pinned(d, (void **)&d->h_inits, &d->cap_h_inits,
                                 (u64)count * sizeof *list);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    memcpy(d->h_inits, list, (size_t)count * sizeof *list);
    HIP_TRY(hipMemcpyAsync(d->inits, d->h_inits, (u64)count * sizeof *list, hipMemcpyHostToDevice,
                           d->stream));
    hipLaunchKernelGGL(k_init_contacts, dim3(blocks_for(count)), dim3(NPH_BLOCK), 0, d->stream,
                       d->contacts, d->gate, d->inits, count);
    HIP_TRY(hipGetLastError());
    return NPH_RESULT_OK;
}

// The narrow phase over list entries [first, count): the entries staged in
// pinned memory, copied up, one thread each.
static NphResult collide_range(NphDevice *d, const u32 *awake, u32 first, u32 count, f32 step_dt,
                               hipEvent_t e0, hipEvent_t e1)
{
    if (count > d->cap_list) {
        // Keep what an earlier range wrote (stream-ordered copies).
        const u32 c = devi_grow_cap(d->cap_list, count);
        NphResult r;
        if ((r = devi_grow_buffer((void **)&d->list, (u64)d->cap_list * 4U, (u64)c * 4U, d->stream,
                             1)) != NPH_RESULT_OK ||
            (r = devi_grow_buffer((void **)&d->codes, (u64)d->cap_list * 4U, (u64)c * 4U, d->stream,
                             1)) != NPH_RESULT_OK) {
            return r;
        }
        d->cap_list = c;
    }
    // Entries [0, first) went up with an earlier range: growing may drop them.
    {
        NphResult r = devi_reserve_pinned(d, (void **)&d->h_list, &d->cap_h_list, (u64)count * 4U);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    const u32 n = count - first;
    memcpy(d->h_list + first, awake + first, (size_t)n * 4U);
    HIP_TRY(hipMemcpyAsync(d->list + first, d->h_list + first, (u64)n * 4U, hipMemcpyHostToDevice,
                           d->stream));
    if (d->team_counts == NULL) {
        HIP_TRY(hipMalloc((void **)&d->team_counts, 64U));
    }
    if (n > d->cap_team) {
        const u32 c = devi_grow_cap(d->cap_team, n);
        NphResult r;
        if ((r = devi_grow_buffer((void **)&d->team_small, 0, (u64)c * sizeof(NphSortedEntry),
                                  d->stream, 0)) != NPH_RESULT_OK ||
            (r = devi_grow_buffer((void **)&d->team_big, 0, (u64)c * sizeof(NphSortedEntry),
                                  d->stream, 0)) != NPH_RESULT_OK ||
            (r = devi_grow_buffer((void **)&d->team_mesh, 0, (u64)c * sizeof(NphSortedEntry),
                                  d->stream, 0)) != NPH_RESULT_OK) {
            return r;
        }
        d->cap_team = c;
    }
    HIP_TRY(hipEventRecord(e0, d->stream));
    NphOneWorld w;
    w.c.contacts = d->contacts;
    w.c.shapes = d->shapes;
    w.c.bodies = d->bodies;
    w.c.gate = d->gate;
    NphTeamLists lists;
    lists.small = d->team_small;
    lists.big = d->team_big;
    lists.mesh = d->team_mesh;
    lists.counts = d->team_counts;
    hipLaunchKernelGGL(k_narrow_clear, dim3(1), dim3(64), 0, d->stream, d->team_counts);
    hipLaunchKernelGGL(k_narrow_lanes, dim3(devi_blocks(n, NPH_BLOCK)), dim3(NPH_BLOCK), 0,
                       d->stream, w, d->list + first, d->codes + first, n, step_dt, lists);
    hipLaunchKernelGGL(k_narrow_teams, dim3(devi_team_grid(n, NPH_TEAM_BLOCK)),
                       dim3(NPH_TEAM_BLOCK), 0, d->stream, w, d->team_small, d->team_counts,
                       d->codes + first, step_dt);
    if (d->big_hull_shapes > 0U) {
        hipLaunchKernelGGL(k_narrow_big_teams, dim3(devi_team_grid(n, NPH_BIG_TEAM_BLOCK)),
                           dim3(NPH_BIG_TEAM_BLOCK), 0, d->stream, w, d->team_big,
                           d->team_counts + 1, d->codes + first, step_dt);
    }
    if (d->mesh_shapes > 0U && d->big_hull_shapes > 0U) {
        hipLaunchKernelGGL(k_narrow_big_mesh_teams, dim3(devi_team_grid(n, NPH_BIG_MESH_TEAM_BLOCK)),
                           dim3(NPH_BIG_MESH_TEAM_BLOCK), 0, d->stream, w, d->team_mesh,
                           d->team_counts + 2, d->codes + first, step_dt);
    } else if (d->mesh_shapes > 0U) {
        hipLaunchKernelGGL(k_narrow_mesh_teams, dim3(devi_team_grid(n, NPH_MESH_TEAM_BLOCK)),
                           dim3(NPH_MESH_TEAM_BLOCK), 0, d->stream, w, d->team_mesh,
                           d->team_counts + 2, d->codes + first, step_dt);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, d->stream));
    return NPH_RESULT_OK;
}

void nph_dev_note_mesh_shape(NphDevice *d) { d->mesh_shapes += 1U; }

void nph_dev_note_big_hull_shape(NphDevice *d) { d->big_hull_shapes += 1U; }

NphResult nph_dev_collide_begin(NphDevice *d, const u32 *awake, u32 count, f32 step_dt)
{
    {
        const NphResult fr = devi_flush_staged(d);
        if (fr != NPH_RESULT_OK) {
            return fr;
        }
    }
    d->collide_ms = 0.0;
    d->collide_first = count;
    if (count == 0U) {
        return NPH_RESULT_OK;
    }
    return collide_range(d, awake, 0U, count, step_dt, d->cev[0], d->cev[1]);
}

NphResult nph_dev_collide_end(NphDevice *d, const u32 *awake, u32 count, f32 step_dt,
                              const u32 **codes)
{
    *codes = NULL;
    const u32 first = d->collide_first;
    if (count < first) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    if (count > first) {
        NphResult r = collide_range(d, awake, first, count, step_dt, d->cev[2], d->cev[3]);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    if (count > 0U) {
        NphResult r = devi_reserve_pinned(d, (void **)&d->h_codes, &d->cap_h_codes, (u64)count * 4U);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        HIP_TRY(hipMemcpyAsync(d->h_codes, d->codes, (u64)count * 4U, hipMemcpyDeviceToHost,
                               d->stream));
    }
    HIP_TRY(hipStreamSynchronize(d->stream));
    f64 total = 0.0;
    float ms = 0.0F;
    if (first > 0U && hipEventElapsedTime(&ms, d->cev[0], d->cev[1]) == hipSuccess) {
        total += ms;
    }
    if (count > first && hipEventElapsedTime(&ms, d->cev[2], d->cev[3]) == hipSuccess) {
        total += ms;
    }
    d->collide_ms = total;
    *codes = d->h_codes;
    return NPH_RESULT_OK;
}

NphResult nph_dev_solve(NphDevice *d, const NphSolveInput *in, const NphFastBody **fast,
                        u32 *fast_count)
{
    *fast = NULL;
    *fast_count = 0U;
    d->pending_sd.body_count = 0U;
    {
        const NphResult fr = devi_flush_staged(d);
        if (fr != NPH_RESULT_OK) {
            return fr;
        }
    }
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
    u32 jfirst[NPH_GRAPH_COLORS + 1U];
    u32 jcount[NPH_GRAPH_COLORS + 1U];
    u32 joints = 0U;
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
        jfirst[c] = joints;
        jcount[c] = in->joint_color_count[c];
        joints += jcount[c];
    }
    if (joints > d->cap_step_joints) {
        const u32 c = devi_grow_cap(d->cap_step_joints, joints);
        NphResult gr = devi_grow_buffer((void **)&d->step_joints, 0, (u64)c * 4U, d->stream, 0);
        if (gr != NPH_RESULT_OK) {
            return gr;
        }
        d->cap_step_joints = c;
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
    if (joints > 0U) {
        HIP_TRY(hipMemcpyAsync(d->step_joints, in->graph_joints, (u64)joints * 4U,
                               hipMemcpyHostToDevice, d->stream));
    }
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
    sd.joint_conf = d->joint_conf;
    sd.joint_state = d->joint_state;
    sd.step_joints = d->step_joints;
    sd.joint_count = joints;
    const NphStepParams sp = in->params;
    hipStream_t s = d->stream;

    HIP_TRY(hipEventRecord(d->ev[0], s));
    hipLaunchKernelGGL(k_solver_map_fill, dim3(blocks_for(in->body_slots)), dim3(NPH_BLOCK), 0, s,
                       d->solver_of_body, in->body_slots, n);
    hipLaunchKernelGGL(k_solver_map_set, dim3(blocks_for(n)), dim3(NPH_BLOCK), 0, s,
                       d->solver_of_body, d->awake_bodies, n);
    hipLaunchKernelGGL(k_prepare_bodies, dim3(blocks_for(n + 1U)), dim3(NPH_BLOCK), 0, s, sd, sp);
    if (joints > 0U) {
        hipLaunchKernelGGL(k_joint_prepare, dim3(blocks_for(joints)), dim3(NPH_BLOCK), 0, s, sd,
                           sp);
    }
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
    r = learned_before(d, &sd, groups, sp.inv_h);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    // One constraint pass: the overflow first, then colours 0..23 (solver.c:774-789).
    u32 timed = 0U; // colour-pass launches bracketed by events this step
#define PASS(kind)                                                                                 \
    do {                                                                                           \
        const int joints_ = (kind) != PASS_RESTITUTION;                                            \
        const u32 ojc_ = joints_ ? jcount[NPH_GRAPH_OVERFLOW] : 0U;                                \
        if (count[NPH_GRAPH_OVERFLOW] > 0U || ojc_ > 0U) {                                         \
            hipLaunchKernelGGL(k_overflow_pass, dim3(1), dim3(64), 0, s, sd, sp, (u32)(kind),      \
                               first[NPH_GRAPH_OVERFLOW], count[NPH_GRAPH_OVERFLOW],                \
                               jfirst[NPH_GRAPH_OVERFLOW], ojc_);                                   \
        }                                                                                          \
        for (u32 c = 0U; c < NPH_GRAPH_COLORS; ++c) {                                              \
            const u32 jc_ = joints_ ? jcount[c] : 0U;                                              \
            if (count[c] > 0U || jc_ > 0U) {                                                       \
                const int ev_ = d->time_kernels && timed < d->kev_pairs;                           \
                const dim3 grid_(blocks_for(((jc_ + 63U) & ~63U) + count[c] * NPH_LANES));         \
                if (ev_) {                                                                         \
                    /* the dispatch packet itself stamps the kernel's start and end */             \
                    hipExtLaunchKernelGGL(k_color_pass, grid_, dim3(NPH_BLOCK), 0, s,              \
                                          d->kev[2U * timed], d->kev[2U * timed + 1U], 0, sd, sp,  \
                                          (u32)(kind), first[c], count[c], jfirst[c], jc_);        \
                    timed += 1U;                                                                   \
                } else {                                                                           \
                    hipLaunchKernelGGL(k_color_pass, grid_, dim3(NPH_BLOCK), 0, s, sd, sp,         \
                                       (u32)(kind), first[c], count[c], jfirst[c], jc_);           \
                }                                                                                  \
            }                                                                                      \
        }                                                                                          \
    } while (0)
    if (d->time_kernels) {
        // Enough event pairs for every colour of every pass of this step.
        const u32 need = (3U * sp.substeps + 1U) * NPH_GRAPH_COLORS;
        while (d->kev_pairs < need) {
            hipEvent_t *n = (hipEvent_t *)realloc(d->kev, (size_t)(2U * need) * sizeof *n);
            if (n == NULL) {
                return NPH_RESULT_OUT_OF_MEMORY;
            }
            d->kev = n;
            for (u32 i = 2U * d->kev_pairs; i < 2U * need; ++i) {
                HIP_TRY(hipEventCreate(&d->kev[i]));
            }
            d->kev_pairs = need;
        }
    }
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
    r = learned_after(d, &sd, groups, sp.inv_h);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    hipLaunchKernelGGL(k_finalize, dim3(blocks_for(n)), dim3(NPH_BLOCK), 0, s, sd, sp);
    HIP_TRY(hipGetLastError());
    d->pending_sd = sd;
    d->pending_timed = timed;
    if (sp.continuous == 0U) {
        return NPH_RESULT_OK;
    }
    // The fast bodies in the awake list's order, packed with their solver
    // rows for the host's sweeps.
    hipLaunchKernelGGL(k_fast_count, dim3(blocks_for(n)), dim3(NPH_BLOCK), 0, s, sd,
                       d->grown_blocks);
    hipLaunchKernelGGL(k_fast_scatter, dim3(blocks_for(n)), dim3(NPH_BLOCK), 0, s, sd,
                       d->grown_blocks, d->fast_list);
    HIP_TRY(hipGetLastError());
    NphResult pr = devi_reserve_pinned(d, (void **)&d->h_fast_count, &d->cap_h_fast_count, 64U);
    if (pr != NPH_RESULT_OK) {
        return pr;
    }
    HIP_TRY(hipMemcpyAsync(d->h_fast_count, d->fast_list, 4U, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const u32 f = d->h_fast_count[0];
    if (f == 0U) {
        return NPH_RESULT_OK;
    }
    if (f > d->cap_fast_pack) {
        const u32 c = devi_grow_cap(d->cap_fast_pack, f);
        pr = devi_grow_buffer((void **)&d->fast_pack, 0, (u64)c * sizeof(NphFastBody), s, 0);
        if (pr != NPH_RESULT_OK) {
            return pr;
        }
        d->cap_fast_pack = c;
    }
    pr = devi_reserve_pinned(d, (void **)&d->h_fast_pack, &d->cap_h_fast_pack,
                        (u64)f * sizeof(NphFastBody));
    if (pr != NPH_RESULT_OK) {
        return pr;
    }
    hipLaunchKernelGGL(k_fast_pack, dim3(blocks_for(f)), dim3(NPH_BLOCK), 0, s, sd, d->fast_list,
                       f, d->fast_pack);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(d->h_fast_pack, d->fast_pack, (u64)f * sizeof(NphFastBody),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *fast = d->h_fast_pack;
    *fast_count = f;
    return NPH_RESULT_OK;
}

NphResult nph_dev_solve_finish(NphDevice *d, const NphGrown **grown, u32 *grown_count)
{
    *grown = NULL;
    *grown_count = 0U;
    {
        const NphResult fr = devi_flush_staged(d); // the host's sweeps
        if (fr != NPH_RESULT_OK) {
            return fr;
        }
    }
    const NphStepData sd = d->pending_sd;
    const u32 n = sd.body_count;
    if (n == 0U) {
        return NPH_RESULT_OK;
    }
    hipStream_t s = d->stream;
    hipLaunchKernelGGL(k_grown_count, dim3(blocks_for(n)), dim3(NPH_BLOCK), 0, s, sd,
                       d->grown_blocks);
    hipLaunchKernelGGL(k_grown_scatter, dim3(blocks_for(n)), dim3(NPH_BLOCK), 0, s, sd,
                       d->grown_blocks, d->grown);
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
        const u32 c = devi_grow_cap(d->cap_h_grown, m + 1U);
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
    for (u32 i = 0U; i < d->pending_timed; ++i) {
        float kms = 0.0F;
        HIP_TRY(hipEventElapsedTime(&kms, d->kev[2U * i], d->kev[2U * i + 1U]));
        d->kernel_ms += kms;
        d->kernel_launches += 1U;
    }
    d->pending_timed = 0U;
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

// ---------------------------------------------------------------------------
// The learned warm start: configuration, statistics, the contact log
// ---------------------------------------------------------------------------

NphResult devi_reserve_stats(NphDevice *d)
{
    if (d->wstats != NULL) {
        return NPH_RESULT_OK;
    }
    HIP_TRY(hipMalloc((void **)&d->wstats, NPH_WSTATS * 4U));
    HIP_TRY(hipMemsetAsync(d->wstats, 0, NPH_WSTATS * 4U, d->stream));
    const u32 inf = 0x7F800000U; // no threshold until some contact starts plainly
    HIP_TRY(hipMemcpyAsync(d->wstats + WS_THRESHOLD, &inf, 4U, hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipHostMalloc((void **)&d->h_wstats, NPH_WSTATS * 4U, hipHostMallocDefault));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return NPH_RESULT_OK;
}

NphResult nph_dev_set_warm_start(NphDevice *d, const NphAmdWarmStart *ws)
{
    if (ws == NULL) {
        d->learn_on = 0;
        return NPH_RESULT_OK;
    }
    if (ws->params == NULL || ws->mode > NPH_AMD_WARM_ALL_POINTS) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    if (d->mlp_params == NULL) {
        HIP_TRY(hipMalloc((void **)&d->mlp_params, NPH_AMD_MLP_PARAMS * 4U));
    }
    NphResult r = devi_reserve_stats(d);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    HIP_TRY(hipMemcpyAsync(d->mlp_params, ws->params, NPH_AMD_MLP_PARAMS * 4U,
                           hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    memcpy(d->shift, ws->feature_shift, sizeof d->shift);
    memcpy(d->scale, ws->feature_scale, sizeof d->scale);
    memcpy(d->out_scale, ws->output_scale, sizeof d->out_scale);
    d->learn_mode = ws->mode;
    d->gate_on = ws->gate ? 1 : 0;
    d->learn_on = 1;
    // Worlds stepped in one batch share one network: this names it.
    u64 h = 0xcbf29ce484222325ULL;
    const u8 *parts[5] = {(const u8 *)ws->params, (const u8 *)d->shift, (const u8 *)d->scale,
                          (const u8 *)d->out_scale, (const u8 *)&d->learn_mode};
    const u64 sizes[5] = {NPH_AMD_MLP_PARAMS * 4U, sizeof d->shift, sizeof d->scale,
                          sizeof d->out_scale, sizeof d->learn_mode};
    for (u32 k = 0U; k < 5U; ++k) {
        for (u64 i = 0U; i < sizes[k]; ++i) {
            h = (h ^ parts[k][i]) * 0x100000001b3ULL;
        }
    }
    d->learn_hash = h ^ (u64)d->gate_on;
    return NPH_RESULT_OK;
}

void nph_dev_measure_residual(NphDevice *d, int enable) { d->measure = enable; }

int nph_dev_learned_active(const NphDevice *d)
{
    return d->learn_on || d->measure || d->log_cap > 0U;
}

int nph_dev_logging(const NphDevice *d) { return d->log_cap > 0U; }

NphResult nph_dev_warm_stats(NphDevice *d, NphAmdWarmStats *out)
{
    memset(out, 0, sizeof *out);
    if (d->wstats == NULL) {
        return NPH_RESULT_OK;
    }
    HIP_TRY(hipMemcpyAsync(d->h_wstats, d->wstats, NPH_WSTATS * 4U, hipMemcpyDeviceToHost,
                           d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    const u32 *w = d->h_wstats;
    f32 f;
    memcpy(&f, &w[WS_MAX_PLAIN], 4U);
    out->residual_max_plain = f;
    memcpy(&f, &w[WS_MAX_LEARNED], 4U);
    out->residual_max_learned = f;
    memcpy(&f, &w[WS_SUM_PLAIN], 4U);
    out->residual_sum_plain = f;
    memcpy(&f, &w[WS_SUM_LEARNED], 4U);
    out->residual_sum_learned = f;
    memcpy(&f, &w[WS_MAX_PEN], 4U);
    out->penetration_max = f;
    out->contacts_plain = w[WS_COUNT_PLAIN];
    out->contacts_learned = w[WS_COUNT_LEARNED];
    out->gated = w[WS_GATED];
    return NPH_RESULT_OK;
}

NphResult nph_dev_log_contacts(NphDevice *d, u32 capacity)
{
    HIP_TRY(hipStreamSynchronize(d->stream));
    void *old[3] = {d->log_x, d->log_t, d->log_count};
    for (u32 i = 0U; i < 3U; ++i) {
        if (old[i] != NULL) {
            HIP_TRY(hipFree(old[i]));
        }
    }
    d->log_x = NULL;
    d->log_t = NULL;
    d->log_count = NULL;
    d->log_cap = 0U;
    if (capacity == 0U) {
        return NPH_RESULT_OK;
    }
    HIP_TRY(hipMalloc((void **)&d->log_x, (u64)capacity * NPH_LEARNED_IN * 4U));
    HIP_TRY(hipMalloc((void **)&d->log_t, (u64)capacity * NPH_LEARNED_OUT * 4U));
    HIP_TRY(hipMalloc((void **)&d->log_count, 4U));
    HIP_TRY(hipMemsetAsync(d->log_count, 0, 4U, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    d->log_cap = capacity;
    return NPH_RESULT_OK;
}

NphResult nph_dev_contact_log(NphDevice *d, u32 *rows, const f32 **features, const f32 **targets)
{
    *rows = 0U;
    *features = d->log_x;
    *targets = d->log_t;
    if (d->log_cap == 0U) {
        return NPH_RESULT_OK;
    }
    u32 n = 0U;
    HIP_TRY(hipMemcpyAsync(&n, d->log_count, 4U, hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    *rows = n < d->log_cap ? n : d->log_cap;
    return NPH_RESULT_OK;
}

// The per-lane buffers of the learned kernels for `lanes` lanes.
static NphResult reserve_learn(NphDevice *d, u32 lanes)
{
    NphResult r = devi_reserve_stats(d);
    if (r != NPH_RESULT_OK || lanes <= d->cap_learn_lanes) {
        return r;
    }
    const u32 c = devi_grow_cap(d->cap_learn_lanes, lanes);
    const u64 rows = (u64)c * NPH_LEARNED_ROWS_PER_LANE;
    if ((r = devi_grow_buffer((void **)&d->lx, 0, rows * NPH_LEARNED_IN * 2U, d->stream, 0)) !=
            NPH_RESULT_OK ||
        (r = devi_grow_buffer((void **)&d->lraw, 0, rows * NPH_LEARNED_IN * 4U, d->stream, 0)) !=
            NPH_RESULT_OK ||
        (r = devi_grow_buffer((void **)&d->ly, 0, rows * NPH_LEARNED_OUT * 4U, d->stream, 0)) !=
            NPH_RESULT_OK ||
        (r = devi_grow_buffer((void **)&d->lres, 0, (u64)c * 4U, d->stream, 0)) != NPH_RESULT_OK) {
        return r;
    }
    d->cap_learn_lanes = c;
    return NPH_RESULT_OK;
}

static NphLearnArgs learn_args(const NphDevice *d, u32 groups, f32 inv_h)
{
    NphLearnArgs a;
    memset(&a, 0, sizeof a);
    a.x = d->learn_on ? d->lx : NULL;
    a.raw = d->log_cap > 0U ? d->lraw : NULL;
    a.y = d->ly;
    a.res = d->lres;
    a.gate = d->gate;
    a.stats = d->wstats;
    a.log_x = d->log_x;
    a.log_t = d->log_t;
    a.log_count = d->log_count;
    a.log_cap = d->log_cap;
    a.mode = d->learn_mode;
    a.gate_on = d->gate_on ? 1U : 0U;
    a.group_count = groups;
    a.inv_h = inv_h;
    memcpy(a.shift, d->shift, sizeof a.shift);
    memcpy(a.scale, d->scale, sizeof a.scale);
    memcpy(a.out_scale, d->out_scale, sizeof a.out_scale);
    return a;
}

// After the lanes are prepared: features, the network and its impulses.
static NphResult learned_before(NphDevice *d, const NphStepData *sd, u32 groups, f32 inv_h)
{
    if (groups == 0U || (!d->learn_on && d->log_cap == 0U && !d->measure)) {
        return NPH_RESULT_OK;
    }
    const u32 lanes = groups * NPH_LANES;
    NphResult r = reserve_learn(d, lanes);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    const NphLearnArgs a = learn_args(d, groups, inv_h);
    hipStream_t s = d->stream;
    if (a.x != NULL || a.raw != NULL) {
        hipLaunchKernelGGL(k_learned_features, dim3(blocks_for(lanes)), dim3(NPH_BLOCK), 0, s, *sd,
                           a);
        HIP_TRY(hipGetLastError());
    }
    if (d->learn_on) {
        r = nph_amd_mlp_forward(d->lx, lanes * NPH_LEARNED_ROWS_PER_LANE, d->mlp_params, d->ly,
                                (void *)s);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        hipLaunchKernelGGL(k_learned_apply, dim3(blocks_for(lanes)), dim3(NPH_BLOCK), 0, s, *sd,
                           a);
        HIP_TRY(hipGetLastError());
    }
    return NPH_RESULT_OK;
}

// After the store: residuals, the gate, the log.
static NphResult learned_after(NphDevice *d, const NphStepData *sd, u32 groups, f32 inv_h)
{
    if (!d->learn_on && d->log_cap == 0U && !d->measure) {
        return NPH_RESULT_OK;
    }
    hipStream_t s = d->stream;
    if (d->learn_on || d->measure) {
        // The statistics are this step's, also when nothing touches.
        const NphResult r = devi_reserve_stats(d);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        hipLaunchKernelGGL(k_learned_clear, dim3(1), dim3(64), 0, s, d->wstats);
    }
    if (groups == 0U) {
        HIP_TRY(hipGetLastError());
        return NPH_RESULT_OK;
    }
    const u32 lanes = groups * NPH_LANES;
    const NphLearnArgs a = learn_args(d, groups, inv_h);
    if (d->learn_on || d->measure) {
        hipLaunchKernelGGL(k_learned_residual, dim3(blocks_for(lanes)), dim3(NPH_BLOCK), 0, s, *sd,
                           a);
        if (d->learn_on && d->gate_on) {
            hipLaunchKernelGGL(k_learned_threshold, dim3(1), dim3(64), 0, s, d->wstats);
            hipLaunchKernelGGL(k_learned_gate, dim3(blocks_for(lanes)), dim3(NPH_BLOCK), 0, s, *sd,
                               a);
        }
    }
    if (d->log_cap > 0U) {
        hipLaunchKernelGGL(k_learned_log, dim3(blocks_for(lanes)), dim3(NPH_BLOCK), 0, s, *sd, a);
    }
    HIP_TRY(hipGetLastError());
    return NPH_RESULT_OK;
}
