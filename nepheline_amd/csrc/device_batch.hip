// SPDX-License-Identifier: Apache-2.0
// Many independent worlds stepped together on one MI355X (gfx950).
//
// Each world keeps its own records (bodies, shapes, contacts, joints: its
// NphDevice), and the host keeps each world's bookkeeping.  What is merged is
// the launches: every stage of the step is one launch over the items of all
// the worlds, so a batch of 1 024 worlds of 16 to 20 000 bodies costs the
// launches of one world, and each colour pass fills the chip with the lanes
// of every world instead of the few hundred of one.
//
// Work is laid out in waves.  A wave (64 threads) serves one world: a wave
// table lists, per wave, the world and the world-local index range it
// covers.  The wave index is wave-uniform, so the table entry and the
// world's descriptor (NphStepData: the world's record arrays and its slices
// of the batch's per-step arrays) load through the scalar unit, and the
// per-thread code is the one-world step's (device_impl.h), operation for
// operation: a world's results are bit for bit the ones it gets alone.
//
// Per-step arrays (solver bodies, lanes, points, groups) are the batch's,
// each world a contiguous slice; lane and point fields stay structure of
// arrays over the whole batch (one stride), so a world's slice is an offset
// of the base pointer.  A world's groups never mix with another's: each
// colour's contacts are grouped eight at a time per world, as the reference
// groups them (solver.c:810-843).

#include "device_impl.h"
#include "device_narrow.h"

#include <hip/hip_ext.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#define NPH_WAVE 64U
#define NPH_COLLIDE_BLOCK 64U
#define NPH_SCAN 1024U
#define NPH_SLOTS (NPH_GRAPH_COLORS + 1U)

typedef struct NphWave {
    u32 world;
    u32 base; // first index this wave covers
    u32 end;  // the world's end of the range
    u32 pad;
} NphWave;

// One world's share of a colour: groups [first, first + count), joints
// [jfirst, jfirst + jcount), both world-local.
typedef struct NphRange {
    u32 first, count, jfirst, jcount;
} NphRange;

// A world's part in the learned warm start: its gate and statistics, its
// lanes' place in the batch's network rows.
typedef struct NphLearnWorld {
    u32 *gate;
    u32 *stats;
    u32 lane_base; // first lane of the world in the batch
    u32 groups;
    u32 learn;     // the network starts its points
    u32 measure;   // residuals measured (learn implies it)
    u32 pad[2];
} NphLearnWorld;

enum { ACC_SUM_PLAIN = 0, ACC_SUM_LEARNED, ACC_COUNT_PLAIN, ACC_COUNT_LEARNED, ACC_MAX_PLAIN,
       ACC_MAX_LEARNED, ACC_MAX_PEN, ACC_GATED, ACC_WORDS };

// A pinned host buffer and its device twin, filled on the host and sent up
// with one copy.
typedef struct Stage {
    u8 *host;
    u8 *dev;
    u64 cap;
    u64 used;
} Stage;

typedef struct Half {
    Stage stage;
    u32 *codes; // device: change codes per list entry
    u32 cap_codes;
    u32 list_count;
} Half;

enum {
    T_SLOTS = 0,  // body slots (the solver map)
    T_BODIES,     // awake bodies + the dummy
    T_GROUPS,
    T_LANES,
    T_JOINTS,
    T_COLOR0,     // + colour: that colour's joints (whole waves), then lanes
    T_COUNT = T_COLOR0 + NPH_GRAPH_COLORS
};

// Host threads that fill the stages world by world (each world's slice of
// every staged array is known from prefix sums first, so the worlds are
// independent): a fixed pool, woken per call, worlds taken one at a time.
class FillPool {
  public:
    FillPool()
    {
        unsigned n = std::thread::hardware_concurrency();
        n = n == 0U ? 1U : (n > 16U ? 16U : n);
        for (unsigned t = 1U; t < n; ++t) {
            workers_.emplace_back([this] { loop(); });
        }
    }
    ~FillPool()
    {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
            ++epoch_;
        }
        cv_.notify_all();
        for (auto &w : workers_) {
            w.join();
        }
    }
    // body(k) for every k in [0, n), on the pool and the calling thread.
    void run(u32 n, const std::function<void(u32)> &body)
    {
        if (n < 8U || workers_.empty()) {
            for (u32 k = 0U; k < n; ++k) {
                body(k);
            }
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            body_ = &body;
            n_ = n;
            next_.store(0U);
            busy_ = (u32)workers_.size();
            ++epoch_;
        }
        cv_.notify_all();
        drain();
        std::unique_lock<std::mutex> g(m_);
        done_.wait(g, [this] { return busy_ == 0U; });
        body_ = nullptr;
    }

  private:
    void drain()
    {
        for (u32 k = next_.fetch_add(1U); k < n_; k = next_.fetch_add(1U)) {
            (*body_)(k);
        }
    }
    void loop()
    {
        u64 seen = 0U;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return epoch_ != seen; });
                seen = epoch_;
                if (stop_) {
                    return;
                }
            }
            drain();
            {
                std::lock_guard<std::mutex> g(m_);
                busy_ -= 1U;
            }
            done_.notify_one();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(u32)> *body_ = nullptr;
    u32 n_ = 0U;
    std::atomic<u32> next_{0U};
    u32 busy_ = 0U;
    u64 epoch_ = 0U;
    bool stop_ = false;
};

struct NphDeviceBatch {
    hipStream_t stream;
    hipEvent_t ev[8]; // solve [0, 1], passes [2, 3], narrow phase halves [4, 5] and [6, 7]
    NphBatchTuning tune;
    Half half[2];
    u32 *h_codes; // both halves' codes, read back
    u64 cap_h_codes;
    // The narrow phase's list sorted by pair type (shared by the halves,
    // which run one after the other), and the key counters.
    NphSortedEntry *sorted;
    u32 cap_sorted;
    // The narrow phase's team lists (32-vertex and large-hull pairs) and
    // their counts.
    NphSortedEntry *team_small;
    NphSortedEntry *team_big;
    NphSortedEntry *team_mesh;
    u32 cap_team_small, cap_team_big, cap_team_mesh;
    u32 *team_counts;
    u32 *keys; // counts [0, 64), cursors [64, 128)
    Stage solve;
    // Per-step device arrays of the solve, the worlds' slices.
    NphSolverBody *sbodies; // awake bodies + one dummy a world
    NphSolverParams *params;
    u32 cap_sbodies, cap_params;
    u32 *lane_body_a, *lane_body_b, *lane_flags;
    f32 *lanes;
    u32 cap_lanes; // lanes (groups * 8): the lane fields' stride
    f32 *points;
    u32 cap_point_lanes; // point rows * 8: the point fields' stride
    u32 *gcount, *gpos, *bsum; // grown shapes per awake body, their places
    u32 cap_gbodies, cap_bsum;
    NphGrown *grown;
    u32 cap_grown;
    u32 *wstart; // per world: first grown record; [n] = total
    u32 cap_wstart;
    NphGrown *h_grown;
    u64 cap_h_grown;
    u32 *h_wstart;
    u64 cap_h_wstart;
    // The learned warm start over the batch: network rows, outputs and
    // residuals of every world's lanes (its slice at its first lane), and
    // the residual statistics summed over worlds and steps (doubles:
    // residual sums, contact counts, maxima, gated contacts).
    u16 *lx;
    f32 *ly;
    f32 *lres;
    u32 cap_learn_lanes;
    f64 *acc;
    f64 *h_acc;
    u32 acc_steps;
    // The substep loop as a graph, instantiated once and updated.
    hipGraphExec_t exec;
    FillPool *pool;
    std::vector<u64> offs; // per-world prefix sums of the staged arrays
    // Timing.
    f64 solve_ms, passes_ms, collide_ms;
    f64 wait_ms[2]; // host blocked on the device: narrow phase's codes, solve's grown proxies
    u32 pending_n;  // worlds of the solve begun and not yet ended
    u32 pending_timed;
    int time_kernels;
    hipEvent_t *kev;
    u32 kev_pairs;
    f64 kernel_ms;
    u64 kernel_launches;
    // The convex narrow-phase kernel's own duration per half (dispatch-
    // stamped events), its launches and the pairs they took.
    hipEvent_t cev[4];
    int cev_made;
    u32 cev_pairs[2];
    f64 collide_kernel_ms;
    u64 collide_kernel_launches;
    u64 collide_kernel_pairs;
};

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// The wave's entry of a table: the wave index is read from the first lane,
// so everything derived from it is uniform.
#define WAVE_ITEM(tab, nwaves)                                                                     \
    const u32 wv_ = (u32)__builtin_amdgcn_readfirstlane(                                           \
        (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));                                      \
    if (wv_ >= (nwaves)) {                                                                         \
        return;                                                                                    \
    }                                                                                              \
    const NphWave e = (tab)[wv_];                                                                  \
    const u32 i = e.base + (threadIdx.x & (NPH_WAVE - 1U));                                        \
    if (i >= e.end) {                                                                              \
        return;                                                                                    \
    }

typedef const NphWave *__restrict__ Tab;
typedef const NphStepData *__restrict__ Worlds;

__global__ void __launch_bounds__(256) k_b_init(Tab tab, u32 nw,
                                                const NphCollideWorld *__restrict__ cw,
                                                const NphContactInit *__restrict__ inits)
{
    WAVE_ITEM(tab, nw)
    const NphCollideWorld c = cw[e.world];
    const NphContactInit in = inits[i];
    init_contact_record(c.contacts, c.gate, &in);
}

// The narrow phase runs over the list sorted by the pair's shape types, so
// a wave runs one collide routine instead of every routine its mixed pairs
// would call: a counting sort over NPH_PAIR_KEYS keys (per-block counts in
// LDS, one global atomic a key a block).  The order inside a key is the
// order the blocks reserve their ranges in; each contact's manifold depends
// only on its own pair, and its change code goes back to its list index.
#define NPH_PAIR_KEYS 64U

__device__ static inline u32 pair_key_of(const NphCollideWorld &c, u32 rec)
{
    const NphContactRec *cr = &c.contacts[rec];
    const u32 ta = nph_minu(c.shapes[cr->shape_a].geom.type, 7U);
    const u32 tb = nph_minu(c.shapes[cr->shape_b].geom.type, 7U);
    return ta < tb ? ta * 8U + tb : tb * 8U + ta;
}

__global__ void __launch_bounds__(64) k_b_keys_clear(u32 *count)
{
    if (threadIdx.x < NPH_PAIR_KEYS) {
        count[threadIdx.x] = 0U;
    }
}

// Entry i of the list (batch-wide index) in this thread, or none.
__device__ static inline int list_item(Tab tab, u32 nw, u32 *world, u32 *i)
{
    const u32 wv = (u32)__builtin_amdgcn_readfirstlane(
        (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    if (wv >= nw) {
        return 0;
    }
    const NphWave e = tab[wv];
    *world = e.world;
    *i = e.base + (threadIdx.x & (NPH_WAVE - 1U));
    return *i < e.end;
}

__global__ void __launch_bounds__(256) k_b_key_count(Tab tab, u32 nw,
                                                     const NphCollideWorld *__restrict__ cw,
                                                     const u32 *__restrict__ list, u32 *count)
{
    __shared__ u32 c[NPH_PAIR_KEYS];
    if (threadIdx.x < NPH_PAIR_KEYS) {
        c[threadIdx.x] = 0U;
    }
    __syncthreads();
    u32 w = 0U, i = 0U;
    if (list_item(tab, nw, &w, &i)) {
        atomicAdd(&c[pair_key_of(cw[w], list[i])], 1U);
    }
    __syncthreads();
    if (threadIdx.x < NPH_PAIR_KEYS && c[threadIdx.x] != 0U) {
        atomicAdd(&count[threadIdx.x], c[threadIdx.x]);
    }
}

// Each key's first place: the counts of the keys before it.
__global__ void __launch_bounds__(64) k_b_key_scan(const u32 *count, u32 *cursor)
{
    if (threadIdx.x == 0U) {
        u32 base = 0U;
        for (u32 k = 0U; k < NPH_PAIR_KEYS; ++k) {
            cursor[k] = base;
            base += count[k];
        }
    }
}

__global__ void __launch_bounds__(256) k_b_key_scatter(Tab tab, u32 nw,
                                                       const NphCollideWorld *__restrict__ cw,
                                                       const u32 *__restrict__ list, u32 *cursor,
                                                       NphSortedEntry *out)
{
    __shared__ u32 c[NPH_PAIR_KEYS];
    __shared__ u32 base[NPH_PAIR_KEYS];
    if (threadIdx.x < NPH_PAIR_KEYS) {
        c[threadIdx.x] = 0U;
    }
    __syncthreads();
    u32 w = 0U, i = 0U, key = 0U, rank = 0U;
    const int mine = list_item(tab, nw, &w, &i);
    if (mine) {
        key = pair_key_of(cw[w], list[i]);
        rank = atomicAdd(&c[key], 1U);
    }
    __syncthreads();
    if (threadIdx.x < NPH_PAIR_KEYS && c[threadIdx.x] != 0U) {
        base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], c[threadIdx.x]);
    }
    __syncthreads();
    if (mine) {
        NphSortedEntry e;
        e.world = w;
        e.record = list[i];
        e.index = i;
        e.pad = 0U;
        out[base[key] + rank] = e;
    }
}

// The narrow phase over the sorted list (device_narrow.h): lanes, then
// teams (convex pairs, large hulls, mesh pairs); a wave of the lane pass
// covers pairs of one type, and so do most teams of a wave of the team
// pass, whose list keeps the waves' order.
__global__ void __launch_bounds__(64) k_b_narrow_clear(u32 *counts) { narrow_counts_clear(counts); }

__global__ void __launch_bounds__(NPH_BLOCK) k_b_narrow_lanes(const NphSortedEntry *__restrict__ sorted, u32 n,
                                                              NphManyWorlds w, u32 *codes,
                                                              f32 step_dt, NphTeamLists lists)
{
    const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
    NphSortedEntry e;
    if (j < n) {
        e = sorted[j];
    } else {
        e.world = e.record = e.index = e.pad = 0U;
    }
    narrow_lane(j < n, e, w, codes, step_dt, lists);
}

__global__ void __launch_bounds__(NPH_TEAM_BLOCK) k_b_narrow_teams(const NphSortedEntry *list, const u32 *count,
                                                                   NphManyWorlds w, u32 *codes,
                                                                   f32 step_dt)
{
    narrow_teams<NPH_HULL_MAX_VERTICES, NPH_CLIP_CAPACITY, NPH_TEAM_BLOCK>(list, count, w, codes,
                                                                            step_dt);
}

__global__ void __launch_bounds__(NPH_BIG_TEAM_BLOCK) k_b_narrow_big_teams(const NphSortedEntry *list,
                                                                           const u32 *count,
                                                                           NphManyWorlds w, u32 *codes,
                                                                           f32 step_dt)
{
    narrow_teams<NPH_BIG_HULL_VERTICES, NPH_BIG_CLIP_CAPACITY, NPH_BIG_TEAM_BLOCK>(
        list, count, w, codes, step_dt);
}

__global__ void __launch_bounds__(NPH_MESH_TEAM_BLOCK) k_b_narrow_mesh_teams(const NphSortedEntry *list,
                                                                             const u32 *count,
                                                                             NphManyWorlds w, u32 *codes,
                                                                             f32 step_dt)
{
    narrow_mesh_teams<NPH_CLIP_CAPACITY, NPH_MESH_TEAM_BLOCK>(list, count, w, codes, step_dt);
}

__global__ void __launch_bounds__(NPH_BIG_MESH_TEAM_BLOCK) k_b_narrow_big_mesh_teams(const NphSortedEntry *list,
                                                                                     const u32 *count,
                                                                                     NphManyWorlds w, u32 *codes,
                                                                                     f32 step_dt)
{
    narrow_mesh_teams<NPH_BIG_CLIP_CAPACITY, NPH_BIG_MESH_TEAM_BLOCK>(list, count, w, codes,
                                                                      step_dt);
}

__global__ void __launch_bounds__(256) k_b_map_fill(Tab tab, u32 nw, Worlds ws)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    d.solver_of_body[i] = d.body_count;
}

__global__ void __launch_bounds__(256) k_b_map_set(Tab tab, u32 nw, Worlds ws)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    if (i < d.body_count) {
        d.solver_of_body[d.awake_bodies[i]] = i;
    }
}

__global__ void __launch_bounds__(256) k_b_prepare_bodies(Tab tab, u32 nw, Worlds ws,
                                                          NphStepParams sp)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    if (i < d.body_count) {
        const u32 slot = d.awake_bodies[i];
        prepare_body(&sp, &d.bodies[slot], slot, &d.sbodies[i], &d.params[i]);
    } else {
        solver_body_dummy(&d.sbodies[i]);
    }
}

__global__ void __launch_bounds__(256) k_b_joint_prepare(Tab tab, u32 nw, Worlds ws,
                                                         NphStepParams sp)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    const u32 slot = d.step_joints[i];
    const NphJointConf *cf = &d.joint_conf[slot];
    joint_prepare(cf, &d.joint_state[slot], &d.bodies[cf->body_a], &d.bodies[cf->body_b],
                  d.solver_of_body[cf->body_a], d.solver_of_body[cf->body_b], d.params, sp.h);
}

__global__ void __launch_bounds__(256) k_b_group_points(Tab tab, u32 nw, Worlds ws)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    NphGroup *grp = &d.groups[i];
    u32 p = 0U;
    for (u32 l = 0U; l < grp->lane_count; ++l) {
        p = nph_maxu(p, d.contacts[d.lane_contact[i * NPH_LANES + l]].point_count);
    }
    grp->point_count = p;
}

// Exclusive scan of one value a thread over a block of NPH_SCAN threads.
__device__ static u32 scan_block(u32 v, u32 *total)
{
    __shared__ u32 s[NPH_SCAN];
    const u32 t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (u32 off = 1U; off < NPH_SCAN; off <<= 1U) {
        const u32 x = t >= off ? s[t - off] : 0U;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    const u32 incl = s[t];
    *total = s[NPH_SCAN - 1U];
    __syncthreads();
    return incl - v;
}

// Each world's groups' first point rows, a block a world.
__global__ void __launch_bounds__(NPH_SCAN) k_b_group_scan(Worlds ws, const u32 *wgroups)
{
    const NphStepData d = ws[blockIdx.x];
    const u32 n = wgroups[blockIdx.x];
    u32 base = 0U;
    for (u32 start = 0U; start < n; start += NPH_SCAN) {
        const u32 g = start + threadIdx.x;
        const u32 p = g < n ? d.groups[g].point_count : 0U;
        u32 total = 0U;
        const u32 pos = scan_block(p, &total);
        if (g < n) {
            d.groups[g].first_point = base + pos;
        }
        base += total;
    }
}

__global__ void __launch_bounds__(256) k_b_lane_prepare(Tab tab, u32 nw, Worlds ws,
                                                        NphStepParams sp)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    lane_prepare(&d, &sp, i / NPH_LANES, i % NPH_LANES);
}

__global__ void __launch_bounds__(256) k_b_group_impact(Tab tab, u32 nw, Worlds ws)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    group_finish_prepare(&d, i);
}

__global__ void __launch_bounds__(256) k_b_integrate_velocities(Tab tab, u32 nw, Worlds ws,
                                                                NphStepParams sp)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    if (i < d.body_count) {
        integrate_velocity(&sp, &d.params[i], &d.sbodies[i]);
    }
}

__global__ void __launch_bounds__(256) k_b_integrate_positions(Tab tab, u32 nw, Worlds ws,
                                                               NphStepParams sp)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    if (i < d.body_count) {
        integrate_position(&sp, &d.params[i], &d.sbodies[i]);
    }
}

// Colour c of every world: a world's threads are laid out as in its own
// launch (joints rounded up to whole waves, then lanes); restitution runs
// no joint rows.
__global__ void __launch_bounds__(256) k_b_color_pass(Tab tab, u32 nw, Worlds ws,
                                                      const NphRange *__restrict__ ranges,
                                                      u32 color, NphStepParams sp, u32 pass)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    const NphRange r = ranges[e.world * NPH_SLOTS + color];
    if (pass == PASS_RESTITUTION && i < ((r.jcount + 63U) & ~63U)) {
        return;
    }
    color_pass_thread(&d, &sp, pass, r.first, r.count, r.jfirst, r.jcount, i);
}

// The overflow of every world that has one, a thread a world.
__global__ void __launch_bounds__(64) k_b_overflow(const u32 *__restrict__ which, u32 n, Worlds ws,
                                                   const NphRange *__restrict__ ranges,
                                                   NphStepParams sp, u32 pass)
{
    const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) {
        return;
    }
    const u32 w = which[k];
    const NphStepData d = ws[w];
    const NphRange r = ranges[w * NPH_SLOTS + NPH_GRAPH_OVERFLOW];
    overflow_pass(&d, &sp, pass, r.first, r.count, r.jfirst,
                  pass == PASS_RESTITUTION ? 0U : r.jcount);
}

__global__ void __launch_bounds__(256) k_b_restitution_masks(Tab tab, u32 nw, Worlds ws,
                                                             f32 threshold)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    group_restitution_masks(&d, threshold, i);
}

__global__ void __launch_bounds__(256) k_b_store(Tab tab, u32 nw, Worlds ws)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    lane_store(&d, i / NPH_LANES, i % NPH_LANES);
}

__global__ void __launch_bounds__(256) k_b_finalize(Tab tab, u32 nw, Worlds ws, NphStepParams sp)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    if (i < d.body_count) {
        finalize_body(&sp, &d.params[i], &d.sbodies[i], &d.bodies[d.awake_bodies[i]], d.shapes);
    }
}

// The learned warm start: the batch's arguments narrowed to world w.
__device__ static inline NphLearnArgs world_learn(const NphLearnArgs &a, const NphLearnWorld &lw)
{
    NphLearnArgs r = a;
    r.x = a.x + (u64)lw.lane_base * NPH_LEARNED_ROWS_PER_LANE * NPH_LEARNED_IN;
    r.y = a.y + (u64)lw.lane_base * NPH_LEARNED_ROWS_PER_LANE * NPH_LEARNED_OUT;
    r.res = a.res + lw.lane_base;
    r.gate = lw.gate;
    r.stats = lw.stats;
    r.group_count = lw.groups;
    return r;
}

typedef const NphLearnWorld *__restrict__ LearnWorlds;

__global__ void __launch_bounds__(256) k_b_learned_features(Tab tab, u32 nw, Worlds ws,
                                                            LearnWorlds lws, NphLearnArgs a)
{
    WAVE_ITEM(tab, nw)
    const NphLearnWorld lw = lws[e.world];
    if (lw.learn == 0U) {
        return;
    }
    const NphStepData d = ws[e.world];
    const NphLearnArgs aw = world_learn(a, lw);
    learned_features_at(&d, &aw, i);
}

__global__ void __launch_bounds__(256) k_b_learned_apply(Tab tab, u32 nw, Worlds ws,
                                                         LearnWorlds lws, NphLearnArgs a)
{
    WAVE_ITEM(tab, nw)
    const NphLearnWorld lw = lws[e.world];
    if (lw.learn == 0U) {
        return;
    }
    const NphStepData d = ws[e.world];
    const NphLearnArgs aw = world_learn(a, lw);
    learned_apply_at(&d, &aw, i);
}

// A thread a world: the statistics words cleared, or the gate's threshold.
__global__ void __launch_bounds__(64) k_b_learned_clear(LearnWorlds lws, u32 n)
{
    const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && lws[k].measure != 0U) {
        for (u32 i = 0U; i < WS_CLEARED; ++i) {
            lws[k].stats[i] = 0U;
        }
    }
}

__global__ void __launch_bounds__(64) k_b_learned_threshold(LearnWorlds lws, u32 n)
{
    const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && lws[k].learn != 0U) {
        learned_threshold(lws[k].stats);
    }
}

// Every lane of a wave reaches the wave's reductions, in range or not.
__global__ void __launch_bounds__(256) k_b_learned_residual(Tab tab, u32 nw, Worlds ws,
                                                            LearnWorlds lws, NphLearnArgs a)
{
    const u32 wv = (u32)__builtin_amdgcn_readfirstlane(
        (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    if (wv >= nw) {
        return;
    }
    const NphWave e = tab[wv];
    const NphLearnWorld lw = lws[e.world];
    if (lw.measure == 0U) {
        return;
    }
    const u32 i = e.base + (threadIdx.x & (NPH_WAVE - 1U));
    const NphStepData d = ws[e.world];
    const NphLearnArgs aw = world_learn(a, lw);
    learned_residual_at(&d, &aw, i, i < e.end);
}

__global__ void __launch_bounds__(256) k_b_learned_gate(Tab tab, u32 nw, Worlds ws,
                                                        LearnWorlds lws, NphLearnArgs a)
{
    WAVE_ITEM(tab, nw)
    const NphLearnWorld lw = lws[e.world];
    if (lw.learn == 0U) {
        return;
    }
    const NphStepData d = ws[e.world];
    const NphLearnArgs aw = world_learn(a, lw);
    learned_gate_at(&d, &aw, i);
}

// The step's statistics of every measured world added to the batch's
// running totals (one block; double sums, in world order per thread, then
// over the block's threads in a fixed order).
__global__ void __launch_bounds__(256) k_b_learned_accumulate(LearnWorlds lws, u32 n, f64 *acc)
{
    __shared__ f64 part[ACC_WORDS][256];
    f64 v[ACC_WORDS];
    for (u32 j = 0U; j < ACC_WORDS; ++j) {
        v[j] = 0.0;
    }
    for (u32 k = threadIdx.x; k < n; k += blockDim.x) {
        if (lws[k].measure == 0U) {
            continue;
        }
        const u32 *st = lws[k].stats;
        v[ACC_SUM_PLAIN] += (f64)__uint_as_float(st[WS_SUM_PLAIN]);
        v[ACC_SUM_LEARNED] += (f64)__uint_as_float(st[WS_SUM_LEARNED]);
        v[ACC_COUNT_PLAIN] += (f64)st[WS_COUNT_PLAIN];
        v[ACC_COUNT_LEARNED] += (f64)st[WS_COUNT_LEARNED];
        v[ACC_MAX_PLAIN] = fmax(v[ACC_MAX_PLAIN], (f64)__uint_as_float(st[WS_MAX_PLAIN]));
        v[ACC_MAX_LEARNED] = fmax(v[ACC_MAX_LEARNED], (f64)__uint_as_float(st[WS_MAX_LEARNED]));
        v[ACC_MAX_PEN] = fmax(v[ACC_MAX_PEN], (f64)__uint_as_float(st[WS_MAX_PEN]));
        v[ACC_GATED] += (f64)st[WS_GATED];
    }
    for (u32 j = 0U; j < ACC_WORDS; ++j) {
        part[j][threadIdx.x] = v[j];
    }
    __syncthreads();
    if (threadIdx.x == 0U) {
        for (u32 t = 1U; t < blockDim.x; ++t) {
            for (u32 j = 0U; j < ACC_WORDS; ++j) {
                if (j == ACC_MAX_PLAIN || j == ACC_MAX_LEARNED || j == ACC_MAX_PEN) {
                    v[j] = fmax(v[j], part[j][t]);
                } else {
                    v[j] += part[j][t];
                }
            }
        }
        for (u32 j = 0U; j < ACC_WORDS; ++j) {
            if (j == ACC_MAX_PLAIN || j == ACC_MAX_LEARNED || j == ACC_MAX_PEN) {
                acc[j] = fmax(acc[j], v[j]);
            } else {
                acc[j] += v[j];
            }
        }
    }
}

// Grown shapes: a count per awake body of the batch (world k's bodies at
// aw_off[k]), an exclusive scan over all of them, then each body writes its
// shapes at its place: every world's grown records in its awake order, the
// worlds one after another.
__global__ void __launch_bounds__(256) k_b_grown_count(Tab tab, u32 nw, Worlds ws,
                                                       const u32 *__restrict__ aw_off, u32 *gcount)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    if (i < d.body_count) {
        gcount[aw_off[e.world] + i] = grown_of_body(&d, i);
    }
}

__global__ void __launch_bounds__(256) k_b_grown_scatter(Tab tab, u32 nw, Worlds ws,
                                                         const u32 *__restrict__ aw_off,
                                                         const u32 *__restrict__ gcount,
                                                         const u32 *__restrict__ gpos,
                                                         NphGrown *out)
{
    WAVE_ITEM(tab, nw)
    const NphStepData d = ws[e.world];
    if (i < d.body_count) {
        const u32 a = aw_off[e.world] + i;
        if (gcount[a] != 0U) {
            grown_write(&d, i, &out[gpos[a]]);
        }
    }
}

__global__ void __launch_bounds__(NPH_SCAN) k_scan_blocks(const u32 *in, u32 *out, u32 n,
                                                          u32 *bsum)
{
    const u32 i = blockIdx.x * NPH_SCAN + threadIdx.x;
    u32 total = 0U;
    const u32 pos = scan_block(i < n ? in[i] : 0U, &total);
    if (i < n) {
        out[i] = pos;
    }
    if (threadIdx.x == 0U) {
        bsum[blockIdx.x] = total;
    }
}

// The block sums scanned in place (one block, in chunks); out[n] = total.
__global__ void __launch_bounds__(NPH_SCAN) k_scan_sums(u32 *bsum, u32 nb, u32 *out, u32 n)
{
    u32 base = 0U;
    for (u32 start = 0U; start < nb; start += NPH_SCAN) {
        const u32 b = start + threadIdx.x;
        const u32 v = b < nb ? bsum[b] : 0U;
        u32 total = 0U;
        const u32 pos = scan_block(v, &total);
        if (b < nb) {
            bsum[b] = base + pos;
        }
        base += total;
    }
    if (threadIdx.x == 0U) {
        out[n] = base;
    }
}

__global__ void __launch_bounds__(NPH_SCAN) k_scan_add(u32 *out, u32 n, const u32 *bsum)
{
    const u32 i = blockIdx.x * NPH_SCAN + threadIdx.x;
    if (i < n) {
        out[i] += bsum[blockIdx.x];
    }
}

__global__ void __launch_bounds__(256) k_b_world_starts(const u32 *aw_off, const u32 *gpos, u32 n,
                                                        u32 *wstart)
{
    const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k <= n) {
        wstart[k] = gpos[aw_off[k]];
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static NphResult stage_reserve(Stage *st, u64 bytes, hipStream_t s)
{
    if (bytes <= st->cap) {
        return NPH_RESULT_OK;
    }
    u64 c = st->cap < 65536U ? 65536U : st->cap;
    while (c < bytes) {
        c *= 2U;
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (st->host != NULL) {
        HIP_TRY(hipHostFree(st->host));
        st->host = NULL;
    }
    if (st->dev != NULL) {
        HIP_TRY(hipFree(st->dev));
        st->dev = NULL;
    }
    st->cap = 0U;
    HIP_TRY(hipHostMalloc((void **)&st->host, c, hipHostMallocDefault));
    HIP_TRY(hipMalloc((void **)&st->dev, c));
    st->cap = c;
    return NPH_RESULT_OK;
}

// Byte offset of the next array of `bytes` in a stage being laid out.
static u64 stage_take(u64 *at, u64 bytes)
{
    const u64 o = (*at + 255U) & ~(u64)255U;
    *at = o + bytes;
    return o;
}

static NphResult pinned_reserve(hipStream_t s, void **p, u64 *cap, u64 bytes)
{
    if (bytes <= *cap) {
        return NPH_RESULT_OK;
    }
    u64 c = *cap < 4096U ? 4096U : *cap;
    while (c < bytes) {
        c *= 2U;
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (*p != NULL) {
        HIP_TRY(hipHostFree(*p));
        *p = NULL;
    }
    *cap = 0U;
    HIP_TRY(hipHostMalloc(p, c, hipHostMallocDefault));
    *cap = c;
    return NPH_RESULT_OK;
}

// Device array growth (contents dropped: per-step data).
template <class T> static NphResult dev_reserve(T **p, u32 *cap, u64 need, hipStream_t s)
{
    if (need <= *cap) {
        return NPH_RESULT_OK;
    }
    if (need > 0xFFFFFFFFULL / 2U) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    const u32 c = devi_grow_cap(*cap, (u32)need);
    NphResult r = devi_grow_buffer((void **)p, 0, (u64)c * sizeof(T), s, 0);
    if (r == NPH_RESULT_OK) {
        *cap = c;
    }
    return r;
}

// Waves covering [0, count) of world w, from tab[at]; their indices shifted
// by `shift` (the world's start in a concatenated array, or 0).
static u32 add_waves(NphWave *tab, u32 at, u32 w, u32 count, u32 shift = 0U)
{
    for (u32 base = 0U; base < count; base += NPH_WAVE) {
        tab[at].world = w;
        tab[at].base = base + shift;
        tab[at].end = count + shift;
        tab[at].pad = 0U;
        at += 1U;
    }
    return at;
}

static u32 waves_of(u32 count) { return (count + NPH_WAVE - 1U) / NPH_WAVE; }

NphResult nph_devb_create(NphDeviceBatch **out)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
        fprintf(stderr, "nepheline_amd: no HIP device: the world step runs only on an MI355X\n");
        return NPH_RESULT_UNSUPPORTED;
    }
    NphDeviceBatch *b = new (std::nothrow) NphDeviceBatch();
    if (b == NULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    b->pool = new (std::nothrow) FillPool();
    if (b->pool == NULL) {
        delete b;
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b->pool;
        delete b;
        return NPH_RESULT_DEVICE_LOST;
    }
    for (u32 i = 0U; i < 8U; ++i) {
        if (hipEventCreate(&b->ev[i]) != hipSuccess) {
            nph_devb_destroy(b);
            return NPH_RESULT_DEVICE_LOST;
        }
    }
    b->tune.pass_block = 256U;
    b->tune.body_block = 256U;
    b->tune.use_graph = 0U;
    if (hipMalloc((void **)&b->keys, 2U * NPH_PAIR_KEYS * 4U) != hipSuccess ||
        hipMalloc((void **)&b->acc, ACC_WORDS * sizeof(f64)) != hipSuccess ||
        hipHostMalloc((void **)&b->h_acc, ACC_WORDS * sizeof(f64), hipHostMallocDefault) !=
            hipSuccess ||
        hipMemsetAsync(b->acc, 0, ACC_WORDS * sizeof(f64), b->stream) != hipSuccess) {
        nph_devb_destroy(b);
        return NPH_RESULT_DEVICE_LOST;
    }
    *out = b;
    return NPH_RESULT_OK;
}

void nph_devb_destroy(NphDeviceBatch *b)
{
    if (b == NULL) {
        return;
    }
    (void)hipStreamSynchronize(b->stream);
    if (b->exec != NULL) {
        (void)hipGraphExecDestroy(b->exec);
    }
    Stage *st[3] = {&b->half[0].stage, &b->half[1].stage, &b->solve};
    for (u32 i = 0U; i < 3U; ++i) {
        if (st[i]->host != NULL) {
            (void)hipHostFree(st[i]->host);
        }
        if (st[i]->dev != NULL) {
            (void)hipFree(st[i]->dev);
        }
    }
    void *bufs[] = {b->half[0].codes, b->half[1].codes, b->sbodies, b->params, b->lane_body_a,
                    b->lane_body_b, b->lane_flags, b->lanes, b->points, b->gcount, b->gpos,
                    b->bsum, b->grown, b->wstart};
    for (u32 i = 0U; i < sizeof bufs / sizeof bufs[0]; ++i) {
        if (bufs[i] != NULL) {
            (void)hipFree(bufs[i]);
        }
    }
    void *lbufs[] = {b->lx,     b->ly,         b->lres,      b->acc,
                     b->sorted, b->keys,       b->team_small, b->team_big,
                     b->team_mesh, b->team_counts};
    for (u32 i = 0U; i < sizeof lbufs / sizeof lbufs[0]; ++i) {
        if (lbufs[i] != NULL) {
            (void)hipFree(lbufs[i]);
        }
    }
    void *pinned[] = {b->h_codes, b->h_grown, b->h_wstart, b->h_acc};
    for (u32 i = 0U; i < 4U; ++i) {
        if (pinned[i] != NULL) {
            (void)hipHostFree(pinned[i]);
        }
    }
    for (u32 i = 0U; i < 8U; ++i) {
        (void)hipEventDestroy(b->ev[i]);
    }
    for (u32 i = 0U; i < 2U * b->kev_pairs; ++i) {
        (void)hipEventDestroy(b->kev[i]);
    }
    if (b->cev_made) {
        for (u32 i = 0U; i < 4U; ++i) {
            (void)hipEventDestroy(b->cev[i]);
        }
    }
    free(b->kev);
    (void)hipStreamDestroy(b->stream);
    delete b->pool;
    delete b;
}

NphResult nph_devb_world(NphDeviceBatch *b, NphDevice **out) { return devi_create(out, b->stream); }

void nph_devb_adopt(NphDeviceBatch *b, NphDevice *d) { d->stream = b->stream; }

void nph_devb_set_tuning(NphDeviceBatch *b, const NphBatchTuning *t)
{
    b->tune = *t;
    if (b->tune.pass_block != 64U && b->tune.pass_block != 128U && b->tune.pass_block != 256U) {
        b->tune.pass_block = 256U;
    }
    if (b->tune.body_block != 64U && b->tune.body_block != 128U && b->tune.body_block != 256U) {
        b->tune.body_block = 256U;
    }
}

NphBatchTuning nph_devb_tuning(const NphDeviceBatch *b) { return b->tune; }

void nph_devb_times(const NphDeviceBatch *b, f64 out[5])
{
    out[0] = b->solve_ms;
    out[1] = b->passes_ms;
    out[2] = b->collide_ms;
    out[3] = b->wait_ms[0];
    out[4] = b->wait_ms[1];
}

void nph_devb_time_kernels(NphDeviceBatch *b, int enable)
{
    b->time_kernels = enable;
    b->kernel_ms = 0.0;
    b->kernel_launches = 0U;
    b->collide_kernel_ms = 0.0;
    b->collide_kernel_launches = 0U;
    b->collide_kernel_pairs = 0U;
    b->cev_pairs[0] = b->cev_pairs[1] = 0U;
}

void nph_devb_kernel_times(const NphDeviceBatch *b, f64 out[5])
{
    out[0] = b->kernel_ms;
    out[1] = (f64)b->kernel_launches;
    out[2] = b->collide_kernel_ms;
    out[3] = (f64)b->collide_kernel_launches;
    out[4] = (f64)b->collide_kernel_pairs;
}

NphResult nph_devb_follow(NphDeviceBatch *b, const NphDeviceBatch *first)
{
    HIP_TRY(hipStreamWaitEvent(b->stream, first->ev[5], 0));
    return NPH_RESULT_OK;
}

NphResult nph_devb_warm_stats(NphDeviceBatch *b, f64 out[8], u32 *steps, int reset)
{
    HIP_TRY(hipMemcpyAsync(b->h_acc, b->acc, ACC_WORDS * sizeof(f64), hipMemcpyDeviceToHost,
                           b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    memcpy(out, b->h_acc, ACC_WORDS * sizeof(f64));
    *steps = b->acc_steps;
    if (reset) {
        HIP_TRY(hipMemsetAsync(b->acc, 0, ACC_WORDS * sizeof(f64), b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        b->acc_steps = 0U;
    }
    return NPH_RESULT_OK;
}

NphResult nph_devb_collide(NphDeviceBatch *b, const NphCollidePart *parts, u32 n, f32 step_dt,
                           u32 half, const u32 *codes[2])
{
    if (half > 1U) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    Half *h = &b->half[half];
    hipStream_t s = b->stream;
    // Every world's staged record writes first, in stream order.
    for (u32 k = 0U; k < n; ++k) {
        const NphResult r = devi_flush_staged(parts[k].dev);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    u64 inits = 0U, list = 0U, iw = 0U, lw = 0U;
    int mesh = 0;
    int big = 0;
    for (u32 k = 0U; k < n; ++k) {
        const NphCollidePart *p = &parts[k];
        if (p->count < p->first) {
            return NPH_RESULT_INVALID_ARGUMENT;
        }
        inits += p->init_count;
        list += p->count - p->first;
        iw += waves_of(p->init_count);
        lw += waves_of(p->count - p->first);
        mesh |= p->dev->mesh_shapes > 0U;
        big |= p->dev->big_hull_shapes > 0U;
    }
    if (list > 0xFFFFFFF0ULL || inits > 0xFFFFFFF0ULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    u64 at = 0U;
    const u64 o_cw = stage_take(&at, (u64)n * sizeof(NphCollideWorld));
    const u64 o_init = stage_take(&at, inits * sizeof(NphContactInit));
    const u64 o_list = stage_take(&at, list * 4U);
    const u64 o_iw = stage_take(&at, iw * sizeof(NphWave));
    const u64 o_lw = stage_take(&at, lw * sizeof(NphWave));
    NphResult r = stage_reserve(&h->stage, at, s);
    if (r == NPH_RESULT_OK) {
        r = dev_reserve(&h->codes, &h->cap_codes, list > 0U ? list : 1U, s);
    }
    if (r == NPH_RESULT_OK) {
        r = dev_reserve(&b->sorted, &b->cap_sorted, list > 0U ? list : 1U, s);
    }
    if (r != NPH_RESULT_OK) {
        return r;
    }
    u8 *hs = h->stage.host;
    NphCollideWorld *cw = (NphCollideWorld *)(hs + o_cw);
    NphContactInit *hi = (NphContactInit *)(hs + o_init);
    u32 *hl = (u32 *)(hs + o_list);
    NphWave *tiw = (NphWave *)(hs + o_iw);
    NphWave *tlw = (NphWave *)(hs + o_lw);
    // Each world's place in the concatenated arrays and tables, then the
    // worlds filled in parallel.
    b->offs.resize(4U * (size_t)n);
    u64 *off = b->offs.data();
    {
        u64 ai = 0U, al = 0U, ti = 0U, tl = 0U;
        for (u32 k = 0U; k < n; ++k) {
            const NphCollidePart *p = &parts[k];
            off[4U * k] = ai;
            off[4U * k + 1U] = al;
            off[4U * k + 2U] = ti;
            off[4U * k + 3U] = tl;
            ai += p->init_count;
            al += p->count - p->first;
            ti += waves_of(p->init_count);
            tl += waves_of(p->count - p->first);
        }
    }
    b->pool->run(n, [&](u32 k) {
        const NphCollidePart *p = &parts[k];
        const u64 ai = off[4U * k], al = off[4U * k + 1U];
        cw[k].contacts = p->dev->contacts;
        cw[k].shapes = p->dev->shapes;
        cw[k].bodies = p->dev->bodies;
        cw[k].gate = p->dev->gate;
        if (p->init_count > 0U) {
            memcpy(hi + ai, p->inits, (size_t)p->init_count * sizeof *hi);
        }
        const u32 m = p->count - p->first;
        if (m > 0U) {
            memcpy(hl + al, p->list + p->first, (size_t)m * 4U);
        }
        // Table entries carry batch-wide indices.
        (void)add_waves(tiw, (u32)off[4U * k + 2U], k, p->init_count, (u32)ai);
        (void)add_waves(tlw, (u32)off[4U * k + 3U], k, m, (u32)al);
    });
    h->list_count = (u32)list;
    u8 *ds = h->stage.dev;
    if (at > 0U) {
        HIP_TRY(hipMemcpyAsync(ds, hs, at, hipMemcpyHostToDevice, s));
    }
    HIP_TRY(hipEventRecord(b->ev[4U + 2U * half], s));
    if (iw > 0U) {
        hipLaunchKernelGGL(k_b_init, dim3(devi_blocks((u32)iw * NPH_WAVE, 256U)), dim3(256), 0, s,
                           (const NphWave *)(ds + o_iw), (u32)iw,
                           (const NphCollideWorld *)(ds + o_cw),
                           (const NphContactInit *)(ds + o_init));
    }
    if (lw > 0U) {
        const NphWave *tab = (const NphWave *)(ds + o_lw);
        const NphCollideWorld *dcw = (const NphCollideWorld *)(ds + o_cw);
        const u32 *dl = (const u32 *)(ds + o_list);
        const u32 m = (u32)list;
        const dim3 grid(devi_blocks((u32)lw * NPH_WAVE, 256U));
        hipLaunchKernelGGL(k_b_keys_clear, dim3(1), dim3(64), 0, s, b->keys);
        hipLaunchKernelGGL(k_b_key_count, grid, dim3(256), 0, s, tab, (u32)lw, dcw, dl, b->keys);
        hipLaunchKernelGGL(k_b_key_scan, dim3(1), dim3(64), 0, s, b->keys,
                           b->keys + NPH_PAIR_KEYS);
        hipLaunchKernelGGL(k_b_key_scatter, grid, dim3(256), 0, s, tab, (u32)lw, dcw, dl,
                           b->keys + NPH_PAIR_KEYS, b->sorted);
        if (b->time_kernels && !b->cev_made) {
            for (u32 e = 0U; e < 4U; ++e) {
                HIP_TRY(hipEventCreate(&b->cev[e]));
            }
            b->cev_made = 1;
        }
        if (b->team_counts == NULL) {
            HIP_TRY(hipMalloc((void **)&b->team_counts, 64U));
        }
        r = dev_reserve(&b->team_small, &b->cap_team_small, m, s);
        if (r == NPH_RESULT_OK) {
            r = dev_reserve(&b->team_big, &b->cap_team_big, m, s);
        }
        if (r == NPH_RESULT_OK) {
            r = dev_reserve(&b->team_mesh, &b->cap_team_mesh, m, s);
        }
        if (r != NPH_RESULT_OK) {
            return r;
        }
        NphManyWorlds w;
        w.cw = dcw;
        NphTeamLists lists;
        lists.small = b->team_small;
        lists.big = b->team_big;
        lists.mesh = b->team_mesh;
        lists.counts = b->team_counts;
        hipLaunchKernelGGL(k_b_narrow_clear, dim3(1), dim3(64), 0, s, b->team_counts);
        const dim3 lane_grid(devi_blocks(m, NPH_BLOCK));
        const dim3 team_grid(devi_team_grid(m, NPH_TEAM_BLOCK));
        if (b->time_kernels) {
            // The dispatch packets stamp the lane pass's start and the team
            // pass's end.
            hipExtLaunchKernelGGL(k_b_narrow_lanes, lane_grid, dim3(NPH_BLOCK), 0, s,
                                  b->cev[2U * half], NULL, 0, b->sorted, m, w, h->codes, step_dt,
                                  lists);
            hipExtLaunchKernelGGL(k_b_narrow_teams, team_grid, dim3(NPH_TEAM_BLOCK), 0, s, NULL,
                                  b->cev[2U * half + 1U], 0, b->team_small, b->team_counts, w,
                                  h->codes, step_dt);
            b->cev_pairs[half] = m;
        } else {
            hipLaunchKernelGGL(k_b_narrow_lanes, lane_grid, dim3(NPH_BLOCK), 0, s, b->sorted, m, w,
                               h->codes, step_dt, lists);
            hipLaunchKernelGGL(k_b_narrow_teams, team_grid, dim3(NPH_TEAM_BLOCK), 0, s,
                               b->team_small, b->team_counts, w, h->codes, step_dt);
        }
        if (big) {
            hipLaunchKernelGGL(k_b_narrow_big_teams, dim3(devi_team_grid(m, NPH_BIG_TEAM_BLOCK)),
                               dim3(NPH_BIG_TEAM_BLOCK), 0, s, b->team_big, b->team_counts + 1, w,
                               h->codes, step_dt);
        }
        if (mesh && big) {
            hipLaunchKernelGGL(k_b_narrow_big_mesh_teams,
                               dim3(devi_team_grid(m, NPH_BIG_MESH_TEAM_BLOCK)),
                               dim3(NPH_BIG_MESH_TEAM_BLOCK), 0, s, b->team_mesh,
                               b->team_counts + 2, w, h->codes, step_dt);
        } else if (mesh) {
            hipLaunchKernelGGL(k_b_narrow_mesh_teams, dim3(devi_team_grid(m, NPH_MESH_TEAM_BLOCK)),
                               dim3(NPH_MESH_TEAM_BLOCK), 0, s, b->team_mesh, b->team_counts + 2,
                               w, h->codes, step_dt);
        }
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(b->ev[5U + 2U * half], s));
    if (half == 0U) {
        return NPH_RESULT_OK;
    }
    // Both halves' codes back.
    const u64 l0 = b->half[0].list_count;
    const u64 l1 = h->list_count;
    r = pinned_reserve(s, (void **)&b->h_codes, &b->cap_h_codes, (l0 + l1 + 1U) * 4U);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    if (l0 > 0U) {
        HIP_TRY(hipMemcpyAsync(b->h_codes, b->half[0].codes, l0 * 4U, hipMemcpyDeviceToHost, s));
    }
    if (l1 > 0U) {
        HIP_TRY(hipMemcpyAsync(b->h_codes + l0, h->codes, l1 * 4U, hipMemcpyDeviceToHost, s));
    }
    const auto w0 = std::chrono::steady_clock::now();
    HIP_TRY(hipStreamSynchronize(s));
    b->wait_ms[0] =
        std::chrono::duration<f64, std::milli>(std::chrono::steady_clock::now() - w0).count();
    float ms = 0.0F;
    b->collide_ms = 0.0;
    if (hipEventElapsedTime(&ms, b->ev[4], b->ev[5]) == hipSuccess) {
        b->collide_ms += ms;
    }
    if (hipEventElapsedTime(&ms, b->ev[6], b->ev[7]) == hipSuccess) {
        b->collide_ms += ms;
    }
    for (u32 k = 0U; k < 2U; ++k) {
        if (b->cev_pairs[k] > 0U) {
            float kms = 0.0F;
            HIP_TRY(hipEventElapsedTime(&kms, b->cev[2U * k], b->cev[2U * k + 1U]));
            b->collide_kernel_ms += kms;
            b->collide_kernel_launches += 1U;
            b->collide_kernel_pairs += b->cev_pairs[k];
            b->cev_pairs[k] = 0U;
        }
    }
    codes[0] = b->h_codes;
    codes[1] = b->h_codes + l0;
    b->half[0].list_count = 0U;
    return NPH_RESULT_OK;
}

// The launches of one constraint pass over the batch: every world's
// overflow first, then colours 0..23 (solver.c:774-789).
typedef struct PassPlan {
    const NphWave *tab[T_COUNT];
    u32 nw[T_COUNT];
    const NphStepData *ws;
    const NphRange *ranges;
    const u32 *over;
    u32 n_over;
} PassPlan;

static NphResult launch_pass(NphDeviceBatch *b, const PassPlan *pl, const NphStepParams *sp,
                             u32 pass, int timed, u32 *timed_count)
{
    hipStream_t s = b->stream;
    if (pl->n_over > 0U) {
        hipLaunchKernelGGL(k_b_overflow, dim3(devi_blocks(pl->n_over, 64U)), dim3(64), 0, s,
                           pl->over, pl->n_over, pl->ws, pl->ranges, *sp, pass);
    }
    const u32 pb = b->tune.pass_block;
    for (u32 c = 0U; c < NPH_GRAPH_COLORS; ++c) {
        const u32 nw = pl->nw[T_COLOR0 + c];
        if (nw == 0U) {
            continue;
        }
        const int ev = timed && *timed_count < b->kev_pairs;
        const dim3 grid(devi_blocks(nw * NPH_WAVE, pb));
        if (ev) {
            // The dispatch packet itself stamps the kernel's start and end.
            hipExtLaunchKernelGGL(k_b_color_pass, grid, dim3(pb), 0, s,
                                  b->kev[2U * *timed_count], b->kev[2U * *timed_count + 1U], 0,
                                  pl->tab[T_COLOR0 + c], nw, pl->ws, pl->ranges, c, *sp, pass);
            *timed_count += 1U;
        } else {
            hipLaunchKernelGGL(k_b_color_pass, grid, dim3(pb), 0, s, pl->tab[T_COLOR0 + c], nw,
                               pl->ws, pl->ranges, c, *sp, pass);
        }
    }
    return NPH_RESULT_OK;
}

static NphResult launch_substeps(NphDeviceBatch *b, const PassPlan *pl, const NphStepParams *sp,
                                 int timed, u32 *timed_count)
{
    hipStream_t s = b->stream;
    const u32 bb = b->tune.body_block;
    const u32 nb = pl->nw[T_BODIES];
    NphResult r = NPH_RESULT_OK;
    for (u32 step = 0U; step < sp->substeps && r == NPH_RESULT_OK; ++step) {
        hipLaunchKernelGGL(k_b_integrate_velocities, dim3(devi_blocks(nb * NPH_WAVE, bb)),
                           dim3(bb), 0, s, pl->tab[T_BODIES], nb, pl->ws, *sp);
        r = launch_pass(b, pl, sp, PASS_WARM, timed, timed_count);
        if (r == NPH_RESULT_OK) {
            r = launch_pass(b, pl, sp, PASS_SOLVE, timed, timed_count);
        }
        hipLaunchKernelGGL(k_b_integrate_positions, dim3(devi_blocks(nb * NPH_WAVE, bb)),
                           dim3(bb), 0, s, pl->tab[T_BODIES], nb, pl->ws, *sp);
        if (r == NPH_RESULT_OK) {
            r = launch_pass(b, pl, sp, PASS_RELAX, timed, timed_count);
        }
    }
    if (r == NPH_RESULT_OK && pl->nw[T_GROUPS] > 0U) {
        hipLaunchKernelGGL(k_b_restitution_masks, dim3(devi_blocks(pl->nw[T_GROUPS] * NPH_WAVE, bb)),
                           dim3(bb), 0, s, pl->tab[T_GROUPS], pl->nw[T_GROUPS], pl->ws,
                           sp->restitution_threshold);
        r = launch_pass(b, pl, sp, PASS_RESTITUTION, timed, timed_count);
    }
    if (r == NPH_RESULT_OK) {
        HIP_TRY(hipGetLastError());
    }
    return r;
}

// The substep loop captured as a graph and launched; the executable graph
// is kept and updated in place while the launch sequence keeps its shape.
static NphResult launch_substeps_graph(NphDeviceBatch *b, const PassPlan *pl,
                                       const NphStepParams *sp)
{
    hipStream_t s = b->stream;
    hipGraph_t graph = NULL;
    u32 none = 0U;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const NphResult r = launch_substeps(b, pl, sp, 0, &none);
    const hipError_t ce = hipStreamEndCapture(s, &graph);
    if (r != NPH_RESULT_OK) {
        if (graph != NULL) {
            (void)hipGraphDestroy(graph);
        }
        return r;
    }
    HIP_TRY(ce);
    int updated = 0;
    if (b->exec != NULL) {
        hipGraphNode_t bad = NULL;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(b->exec, graph, &bad, &res) == hipSuccess &&
            res == hipGraphExecUpdateSuccess) {
            updated = 1;
        } else {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(b->exec);
            b->exec = NULL;
        }
    }
    if (!updated) {
        const hipError_t ie = hipGraphInstantiate(&b->exec, graph, NULL, NULL, 0);
        if (ie != hipSuccess) {
            b->exec = NULL;
            (void)hipGraphDestroy(graph);
            HIP_TRY(ie);
        }
    }
    const hipError_t le = hipGraphLaunch(b->exec, s);
    (void)hipGraphDestroy(graph);
    HIP_TRY(le);
    return NPH_RESULT_OK;
}

NphResult nph_devb_solve(NphDeviceBatch *b, NphDevice *const *devs, const NphSolveInput *in,
                         u32 n, const NphGrown **grown, const u32 **starts)
{
    const NphResult r = nph_devb_solve_begin(b, devs, in, n);
    return r == NPH_RESULT_OK ? nph_devb_solve_end(b, grown, starts) : r;
}

NphResult nph_devb_solve_begin(NphDeviceBatch *b, NphDevice *const *devs, const NphSolveInput *in,
                               u32 n)
{
    hipStream_t s = b->stream;
    b->pending_n = 0U;
    b->pending_timed = 0U;
    for (u32 k = 0U; k < n; ++k) {
        const NphResult r = devi_flush_staged(devs[k]);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    b->solve_ms = 0.0;
    b->passes_ms = 0.0;
    if (n == 0U) {
        return NPH_RESULT_OK;
    }
    const NphStepParams sp = in[0].params;
    for (u32 k = 1U; k < n; ++k) {
        if (memcmp(&in[k].params, &sp, sizeof sp) != 0) {
            return NPH_RESULT_INVALID_ARGUMENT; // one set of step numbers a batch
        }
    }
    // The learned warm start: one network for the batch (every world that
    // has one has the same), residuals wherever measured; no contact log.
    const NphDevice *teacher = NULL;
    int any_measure = 0;
    for (u32 k = 0U; k < n; ++k) {
        NphDevice *d = devs[k];
        if (d->log_cap > 0U) {
            return NPH_RESULT_UNSUPPORTED;
        }
        if (d->learn_on) {
            if (teacher == NULL) {
                teacher = d;
            } else if (d->learn_hash != teacher->learn_hash) {
                return NPH_RESULT_UNSUPPORTED;
            }
        }
        if (d->learn_on || d->measure) {
            any_measure = 1;
            const NphResult r = devi_reserve_stats(d);
            if (r != NPH_RESULT_OK) {
                return r;
            }
        }
    }
    // Sizes and each world's offsets in the batch's arrays.
    u64 A = 0U, G = 0U, J = 0U, W[T_COUNT];
    memset(W, 0, sizeof W);
    u64 shapes = 0U;
    u32 n_over = 0U;
    // Per world: its awake, group and joint offsets, its overflow slot, and
    // its first wave in each table.
    enum { O_A = 0, O_G, O_J, O_OVER, O_T, O_STRIDE = O_T + T_COUNT };
    b->offs.resize((size_t)O_STRIDE * n);
    u64 *off = b->offs.data();
    for (u32 k = 0U; k < n; ++k) {
        const NphSolveInput *x = &in[k];
        u64 *o = off + (size_t)O_STRIDE * k;
        o[O_A] = A;
        o[O_G] = G;
        o[O_J] = J;
        o[O_OVER] = n_over;
        for (u32 t = 0U; t < T_COUNT; ++t) {
            o[O_T + t] = W[t];
        }
        u32 g = 0U, j = 0U;
        for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
            const u32 per = c == NPH_GRAPH_OVERFLOW ? 1U : NPH_LANES;
            const u32 gc = (x->color_count[c] + per - 1U) / per;
            g += gc;
            j += x->joint_color_count[c];
            if (c < NPH_GRAPH_COLORS) {
                const u32 jc = x->joint_color_count[c];
                W[T_COLOR0 + c] += waves_of(((jc + 63U) & ~63U) + gc * NPH_LANES);
            }
        }
        if (x->color_count[NPH_GRAPH_OVERFLOW] > 0U ||
            x->joint_color_count[NPH_GRAPH_OVERFLOW] > 0U) {
            n_over += 1U;
        }
        A += x->awake_count;
        G += g;
        J += j;
        W[T_SLOTS] += waves_of(x->body_slots);
        W[T_BODIES] += waves_of(x->awake_count + 1U);
        W[T_GROUPS] += waves_of(g);
        W[T_LANES] += waves_of(g * NPH_LANES);
        W[T_JOINTS] += waves_of(j);
        shapes += devs[k]->cap_shapes;
    }
    if (A + n > 0x7FFFFFFFULL || G * NPH_LANES * NPH_PHYSICS_MANIFOLD_MAX_POINTS > 0x7FFFFFFFULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    const u64 lanes = G * NPH_LANES;
    const u64 point_lanes = lanes * NPH_PHYSICS_MANIFOLD_MAX_POINTS; // four rows a group at most
    NphResult r = dev_reserve(&b->sbodies, &b->cap_sbodies, A + n, s);
    if (r == NPH_RESULT_OK) {
        r = dev_reserve(&b->params, &b->cap_params, A > 0U ? A : 1U, s);
    }
    if (r == NPH_RESULT_OK && lanes > b->cap_lanes) {
        const u32 c = devi_grow_cap(b->cap_lanes, (u32)lanes);
        if ((r = devi_grow_buffer((void **)&b->lane_body_a, 0, (u64)c * 4U, s, 0)) ==
                NPH_RESULT_OK &&
            (r = devi_grow_buffer((void **)&b->lane_body_b, 0, (u64)c * 4U, s, 0)) ==
                NPH_RESULT_OK &&
            (r = devi_grow_buffer((void **)&b->lane_flags, 0, (u64)c * 4U, s, 0)) ==
                NPH_RESULT_OK &&
            (r = devi_grow_buffer((void **)&b->lanes, 0, (u64)c * NPH_L_FIELDS * 4U, s, 0)) ==
                NPH_RESULT_OK) {
            b->cap_lanes = c;
        }
    }
    if (r == NPH_RESULT_OK && point_lanes > b->cap_point_lanes) {
        const u32 c = devi_grow_cap(b->cap_point_lanes, (u32)point_lanes);
        if ((r = devi_grow_buffer((void **)&b->points, 0, (u64)c * NPH_P_FIELDS * 4U, s, 0)) ==
            NPH_RESULT_OK) {
            b->cap_point_lanes = c;
        }
    }
    const u32 nb_scan = (u32)((A + NPH_SCAN - 1U) / NPH_SCAN);
    if (r == NPH_RESULT_OK && A + 1U > b->cap_gbodies) {
        const u32 c = devi_grow_cap(b->cap_gbodies, (u32)(A + 1U));
        if ((r = devi_grow_buffer((void **)&b->gcount, 0, (u64)c * 4U, s, 0)) == NPH_RESULT_OK &&
            (r = devi_grow_buffer((void **)&b->gpos, 0, (u64)c * 4U, s, 0)) == NPH_RESULT_OK) {
            b->cap_gbodies = c;
        }
    }
    if (r == NPH_RESULT_OK) {
        r = dev_reserve(&b->bsum, &b->cap_bsum, nb_scan + 1U, s);
    }
    if (r == NPH_RESULT_OK) {
        r = dev_reserve(&b->grown, &b->cap_grown, shapes + 1U, s);
    }
    if (r == NPH_RESULT_OK) {
        r = dev_reserve(&b->wstart, &b->cap_wstart, n + 1U, s);
    }
    if (r == NPH_RESULT_OK && any_measure && lanes > b->cap_learn_lanes) {
        const u32 c = devi_grow_cap(b->cap_learn_lanes, (u32)lanes);
        const u64 rows = (u64)c * NPH_LEARNED_ROWS_PER_LANE;
        if ((r = devi_grow_buffer((void **)&b->lx, 0, rows * NPH_LEARNED_IN * 2U, s, 0)) ==
                NPH_RESULT_OK &&
            (r = devi_grow_buffer((void **)&b->ly, 0, rows * NPH_LEARNED_OUT * 4U, s, 0)) ==
                NPH_RESULT_OK &&
            (r = devi_grow_buffer((void **)&b->lres, 0, (u64)c * 4U, s, 0)) == NPH_RESULT_OK) {
            b->cap_learn_lanes = c;
        }
    }
    if (r != NPH_RESULT_OK) {
        return r;
    }
    // The stage: descriptors, ranges, lists, group tables, wave tables.
    u64 total_waves = 0U;
    for (u32 t = 0U; t < T_COUNT; ++t) {
        total_waves += W[t];
    }
    u64 at = 0U;
    const u64 o_ws = stage_take(&at, (u64)n * sizeof(NphStepData));
    const u64 o_rg = stage_take(&at, (u64)n * NPH_SLOTS * sizeof(NphRange));
    const u64 o_wg = stage_take(&at, (u64)n * 4U);
    const u64 o_ao = stage_take(&at, (u64)(n + 1U) * 4U);
    const u64 o_ov = stage_take(&at, (u64)(n_over + 1U) * 4U);
    const u64 o_aw = stage_take(&at, A * 4U);
    const u64 o_lc = stage_take(&at, lanes * 4U);
    const u64 o_gr = stage_take(&at, G * sizeof(NphGroup));
    const u64 o_sj = stage_take(&at, J * 4U);
    const u64 o_tab = stage_take(&at, total_waves * sizeof(NphWave));
    const u64 o_lw = stage_take(&at, (u64)n * sizeof(NphLearnWorld));
    r = stage_reserve(&b->solve, at, s);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    u8 *hs = b->solve.host;
    u8 *ds = b->solve.dev;
    NphStepData *ws = (NphStepData *)(hs + o_ws);
    NphRange *rg = (NphRange *)(hs + o_rg);
    u32 *wg = (u32 *)(hs + o_wg);
    u32 *ao = (u32 *)(hs + o_ao);
    u32 *ov = (u32 *)(hs + o_ov);
    u32 *aw = (u32 *)(hs + o_aw);
    u32 *lc = (u32 *)(hs + o_lc);
    NphGroup *gr = (NphGroup *)(hs + o_gr);
    u32 *sj = (u32 *)(hs + o_sj);
    NphLearnWorld *lws = (NphLearnWorld *)(hs + o_lw);
    NphWave *tabs[T_COUNT];
    {
        NphWave *t0 = (NphWave *)(hs + o_tab);
        for (u32 t = 0U; t < T_COUNT; ++t) {
            tabs[t] = t0;
            t0 += W[t];
        }
    }
    b->pool->run(n, [&](u32 k) {
        const NphSolveInput *x = &in[k];
        const NphDevice *d = devs[k];
        const u64 *o = off + (size_t)O_STRIDE * k;
        const u64 a_off = o[O_A], g_off = o[O_G], j_off = o[O_J];
        NphStepData *sd = &ws[k];
        memset(sd, 0, sizeof *sd);
        sd->bodies = d->bodies;
        sd->shapes = d->shapes;
        sd->contacts = d->contacts;
        sd->solver_of_body = d->solver_of_body;
        sd->awake_bodies = (const u32 *)(ds + o_aw) + a_off;
        sd->sbodies = b->sbodies + a_off + k;
        sd->params = b->params + a_off;
        sd->groups = (NphGroup *)(ds + o_gr) + g_off;
        sd->lane_contact = (u32 *)(ds + o_lc) + g_off * NPH_LANES;
        sd->lane_body_a = b->lane_body_a + g_off * NPH_LANES;
        sd->lane_body_b = b->lane_body_b + g_off * NPH_LANES;
        sd->lane_flags = b->lane_flags + g_off * NPH_LANES;
        sd->lanes = b->lanes + g_off * NPH_LANES;
        sd->points = b->points + g_off * NPH_LANES * NPH_PHYSICS_MANIFOLD_MAX_POINTS;
        sd->lane_stride = b->cap_lanes;
        sd->point_stride = b->cap_point_lanes;
        sd->body_count = x->awake_count;
        sd->joint_conf = d->joint_conf;
        sd->joint_state = d->joint_state;
        sd->step_joints = (const u32 *)(ds + o_sj) + j_off;
        ao[k] = (u32)a_off;
        memcpy(aw + a_off, x->awake_bodies, (size_t)x->awake_count * 4U);
        // Groups, colour by colour then the overflow, eight contacts a group.
        u32 g = 0U, at_c = 0U, j = 0U;
        for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
            const u32 per = c == NPH_GRAPH_OVERFLOW ? 1U : NPH_LANES;
            const u32 m = x->color_count[c];
            NphRange *rr = &rg[k * NPH_SLOTS + c];
            rr->first = g;
            rr->jfirst = j;
            rr->jcount = x->joint_color_count[c];
            for (u32 i = 0U; i < m; i += per) {
                const u32 lanes_here = nph_minu(per, m - i);
                NphGroup *grp = &gr[g_off + g];
                memset(grp, 0, sizeof *grp);
                grp->lane_count = lanes_here;
                u32 *lcg = lc + (g_off + g) * NPH_LANES;
                for (u32 l = 0U; l < NPH_LANES; ++l) {
                    lcg[l] = l < lanes_here ? x->graph_contacts[at_c + i + l] : NPH_NULL32;
                }
                g += 1U;
            }
            rr->count = g - rr->first;
            at_c += m;
            j += rr->jcount;
            if (c < NPH_GRAPH_COLORS) {
                const u32 threads = ((rr->jcount + 63U) & ~63U) + rr->count * NPH_LANES;
                (void)add_waves(tabs[T_COLOR0 + c], (u32)o[O_T + T_COLOR0 + c], k, threads);
            }
        }
        if (rg[k * NPH_SLOTS + NPH_GRAPH_OVERFLOW].count > 0U ||
            rg[k * NPH_SLOTS + NPH_GRAPH_OVERFLOW].jcount > 0U) {
            ov[o[O_OVER]] = k;
        }
        if (j > 0U) {
            memcpy(sj + j_off, x->graph_joints, (size_t)j * 4U);
        }
        sd->joint_count = j;
        wg[k] = g;
        NphLearnWorld *lw = &lws[k];
        memset(lw, 0, sizeof *lw);
        lw->gate = d->gate;
        lw->stats = d->wstats;
        lw->lane_base = (u32)(g_off * NPH_LANES);
        lw->groups = g;
        lw->learn = d->learn_on ? 1U : 0U;
        lw->measure = (d->learn_on || d->measure) ? 1U : 0U;
        (void)add_waves(tabs[T_SLOTS], (u32)o[O_T + T_SLOTS], k, x->body_slots);
        (void)add_waves(tabs[T_BODIES], (u32)o[O_T + T_BODIES], k, x->awake_count + 1U);
        (void)add_waves(tabs[T_GROUPS], (u32)o[O_T + T_GROUPS], k, g);
        (void)add_waves(tabs[T_LANES], (u32)o[O_T + T_LANES], k, g * NPH_LANES);
        (void)add_waves(tabs[T_JOINTS], (u32)o[O_T + T_JOINTS], k, j);
    });
    const u64 a_off = A;
    const u32 over = n_over;
    ao[n] = (u32)a_off;
    HIP_TRY(hipMemcpyAsync(ds, hs, at, hipMemcpyHostToDevice, s));

    PassPlan pl;
    memset(&pl, 0, sizeof pl);
    {
        const NphWave *t0 = (const NphWave *)(ds + o_tab);
        for (u32 t = 0U; t < T_COUNT; ++t) {
            pl.tab[t] = t0;
            pl.nw[t] = (u32)W[t];
            t0 += W[t];
        }
    }
    pl.ws = (const NphStepData *)(ds + o_ws);
    pl.ranges = (const NphRange *)(ds + o_rg);
    pl.over = (const u32 *)(ds + o_ov);
    pl.n_over = over;
    const NphStepData *dws = pl.ws;
    const u32 *dao = (const u32 *)(ds + o_ao);
    const u32 bb = b->tune.body_block;
#define GRID(t) dim3(devi_blocks(pl.nw[t] * NPH_WAVE, bb)), dim3(bb), 0, s, pl.tab[t], pl.nw[t]

    HIP_TRY(hipEventRecord(b->ev[0], s));
    hipLaunchKernelGGL(k_b_map_fill, GRID(T_SLOTS), dws);
    hipLaunchKernelGGL(k_b_map_set, GRID(T_BODIES), dws);
    hipLaunchKernelGGL(k_b_prepare_bodies, GRID(T_BODIES), dws, sp);
    if (J > 0U) {
        hipLaunchKernelGGL(k_b_joint_prepare, GRID(T_JOINTS), dws, sp);
    }
    if (G > 0U) {
        hipLaunchKernelGGL(k_b_group_points, GRID(T_GROUPS), dws);
        hipLaunchKernelGGL(k_b_group_scan, dim3(n), dim3(NPH_SCAN), 0, s, dws,
                           (const u32 *)(ds + o_wg));
        hipLaunchKernelGGL(k_b_lane_prepare, GRID(T_LANES), dws, sp);
        hipLaunchKernelGGL(k_b_group_impact, GRID(T_GROUPS), dws);
    }
    HIP_TRY(hipGetLastError());
    NphLearnArgs la;
    memset(&la, 0, sizeof la);
    const NphLearnWorld *dlw = (const NphLearnWorld *)(ds + o_lw);
    if (any_measure) {
        la.x = teacher != NULL ? b->lx : NULL;
        la.y = b->ly;
        la.res = b->lres;
        la.inv_h = sp.inv_h;
        if (teacher != NULL) {
            la.mode = teacher->learn_mode;
            la.gate_on = teacher->gate_on ? 1U : 0U;
            memcpy(la.shift, teacher->shift, sizeof la.shift);
            memcpy(la.scale, teacher->scale, sizeof la.scale);
            memcpy(la.out_scale, teacher->out_scale, sizeof la.out_scale);
        }
    }
    if (teacher != NULL && G > 0U) {
        hipLaunchKernelGGL(k_b_learned_features, GRID(T_LANES), dws, dlw, la);
        HIP_TRY(hipGetLastError());
        r = nph_amd_mlp_forward(b->lx, (u32)(lanes * NPH_LEARNED_ROWS_PER_LANE),
                                teacher->mlp_params, b->ly, (void *)s);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        hipLaunchKernelGGL(k_b_learned_apply, GRID(T_LANES), dws, dlw, la);
        HIP_TRY(hipGetLastError());
    }
    u32 timed = 0U;
    const int time_launches = b->time_kernels && !b->tune.use_graph;
    if (time_launches) {
        const u32 need = (3U * sp.substeps + 1U) * NPH_GRAPH_COLORS;
        while (b->kev_pairs < need) {
            hipEvent_t *ne = (hipEvent_t *)realloc(b->kev, (size_t)(2U * need) * sizeof *ne);
            if (ne == NULL) {
                return NPH_RESULT_OUT_OF_MEMORY;
            }
            b->kev = ne;
            for (u32 i = 2U * b->kev_pairs; i < 2U * need; ++i) {
                HIP_TRY(hipEventCreate(&b->kev[i]));
            }
            b->kev_pairs = need;
        }
    }
    HIP_TRY(hipEventRecord(b->ev[2], s));
    r = b->tune.use_graph ? launch_substeps_graph(b, &pl, &sp)
                          : launch_substeps(b, &pl, &sp, time_launches, &timed);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    HIP_TRY(hipEventRecord(b->ev[3], s));
    if (G > 0U) {
        hipLaunchKernelGGL(k_b_store, GRID(T_LANES), dws);
    }
    if (any_measure) {
        hipLaunchKernelGGL(k_b_learned_clear, dim3(devi_blocks(n, 64U)), dim3(64), 0, s, dlw, n);
        if (G > 0U) {
            hipLaunchKernelGGL(k_b_learned_residual, GRID(T_LANES), dws, dlw, la);
            if (teacher != NULL && la.gate_on != 0U) {
                hipLaunchKernelGGL(k_b_learned_threshold, dim3(devi_blocks(n, 64U)), dim3(64), 0,
                                   s, dlw, n);
                hipLaunchKernelGGL(k_b_learned_gate, GRID(T_LANES), dws, dlw, la);
            }
        }
        hipLaunchKernelGGL(k_b_learned_accumulate, dim3(1), dim3(256), 0, s, dlw, n, b->acc);
        HIP_TRY(hipGetLastError());
        b->acc_steps += 1U;
    }
    hipLaunchKernelGGL(k_b_finalize, GRID(T_BODIES), dws, sp);
    // Grown proxies: counts, one scan over the batch, placement.
    hipLaunchKernelGGL(k_b_grown_count, GRID(T_BODIES), dws, dao, b->gcount);
    if (A > 0U) {
        hipLaunchKernelGGL(k_scan_blocks, dim3(nb_scan), dim3(NPH_SCAN), 0, s, b->gcount, b->gpos,
                           (u32)A, b->bsum);
    }
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(NPH_SCAN), 0, s, b->bsum, nb_scan, b->gpos,
                       (u32)A);
    if (A > 0U) {
        hipLaunchKernelGGL(k_scan_add, dim3(nb_scan), dim3(NPH_SCAN), 0, s, b->gpos, (u32)A,
                           b->bsum);
    }
    hipLaunchKernelGGL(k_b_grown_scatter, GRID(T_BODIES), dws, dao, b->gcount, b->gpos, b->grown);
    hipLaunchKernelGGL(k_b_world_starts, dim3(devi_blocks(n + 1U, 256U)), dim3(256), 0, s, dao,
                       b->gpos, n, b->wstart);
    HIP_TRY(hipGetLastError());
#undef GRID
    HIP_TRY(hipEventRecord(b->ev[1], s));
    r = pinned_reserve(s, (void **)&b->h_wstart, &b->cap_h_wstart, (u64)(n + 1U) * 4U);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    HIP_TRY(hipMemcpyAsync(b->h_wstart, b->wstart, (u64)(n + 1U) * 4U, hipMemcpyDeviceToHost, s));
    b->pending_n = n;
    b->pending_timed = timed;
    return NPH_RESULT_OK;
}

NphResult nph_devb_solve_end(NphDeviceBatch *b, const NphGrown **grown, const u32 **starts)
{
    *grown = NULL;
    *starts = NULL;
    hipStream_t s = b->stream;
    const u32 n = b->pending_n;
    const u32 timed = b->pending_timed;
    b->pending_n = 0U;
    b->pending_timed = 0U;
    if (n == 0U) {
        b->wait_ms[1] = 0.0;
        return NPH_RESULT_OK;
    }
    NphResult r = NPH_RESULT_OK;
    const auto w0 = std::chrono::steady_clock::now();
    HIP_TRY(hipStreamSynchronize(s));
    b->wait_ms[1] =
        std::chrono::duration<f64, std::milli>(std::chrono::steady_clock::now() - w0).count();
    const u32 m = b->h_wstart[n];
    r = pinned_reserve(s, (void **)&b->h_grown, &b->cap_h_grown, (u64)(m + 1U) * sizeof(NphGrown));
    if (r != NPH_RESULT_OK) {
        return r;
    }
    if (m > 0U) {
        HIP_TRY(hipMemcpyAsync(b->h_grown, b->grown, (u64)m * sizeof(NphGrown),
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    for (u32 i = 0U; i < timed; ++i) {
        float kms = 0.0F;
        HIP_TRY(hipEventElapsedTime(&kms, b->kev[2U * i], b->kev[2U * i + 1U]));
        b->kernel_ms += kms;
        b->kernel_launches += 1U;
    }
    float ms = 0.0F;
    (void)hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
    b->solve_ms = ms;
    ms = 0.0F;
    (void)hipEventElapsedTime(&ms, b->ev[2], b->ev[3]);
    b->passes_ms = ms;
    *grown = b->h_grown;
    *starts = b->h_wstart;
    return NPH_RESULT_OK;
}
