// SPDX-License-Identifier: Apache-2.0
// Teams: T adjacent lanes of one 64-wide wavefront that work on one item
// (a pair of shapes) together, and the ordered collectives the narrow phase
// builds on. A team never spans two waves, so it needs no workgroup
// barrier: its lanes exchange values through cross-lane permutes (ballot,
// ds_bpermute) and through the team's own slice of LDS behind a wavefront
// fence. With T == 1 every collective is the identity, which is how the
// same code runs one item a thread (and on the host).
//
// Ordered results. The reference scans candidates in a fixed order and
// keeps the first best one; several of its tests compare values rounded
// differently from the values they rank (products instead of quotients),
// so "the best" is not always a total order. The collectives below keep
// that scan exactly:
//  * first_max / first_min: a maximum with ties to the lowest index, which
//    is what a scan with a strict comparison keeps when the comparison is
//    a total order on the values;
//  * replay: the candidates a chunk of lanes passes, visited in lane order
//    by every lane of the team, so a scan whose decision depends on the
//    state left by earlier candidates runs unchanged over them.
#ifndef NPH_TEAM_H
#define NPH_TEAM_H

#include "nph_math.h"

#if defined(__cplusplus)

#if defined(__HIPCC__)
#define NPH_TD __host__ __device__ inline
#else
#define NPH_TD inline
#endif

// An (index, value) pair as a scan leaves it: `index` is NPH_NULL32 while
// no candidate has passed.
struct NphPick {
    f32 value;
    u32 index;
};

template <u32 T> struct NphTeam;

// One lane: the scalar case, host and device.
template <> struct NphTeam<1U> {
    static constexpr u32 size = 1U;
    u32 lane;
    NPH_TD NphTeam() : lane(0U) {}
    NPH_TD u64 ballot(bool p) const { return p ? 1ULL : 0ULL; }
    NPH_TD f32 get(f32 v, u32) const { return v; }
    NPH_TD u32 get(u32 v, u32) const { return v; }
    NPH_TD NphVec3 get(NphVec3 v, u32) const { return v; }
    NPH_TD void sync() const {}
    NPH_TD u32 below(u64) const { return 0U; }
    NPH_TD NphPick first_max(NphPick p) const { return p; }
    NPH_TD NphPick first_min(NphPick p) const { return p; }
    NPH_TD u32 add(u32 v) const { return v; }
};

#if defined(__HIPCC__)
// T lanes of a wave, T a power of two up to 32; lanes [k T, (k + 1) T).
template <u32 T> struct NphTeam {
    static_assert(T == 2U || T == 4U || T == 8U || T == 16U || T == 32U, "team size");
    static constexpr u32 size = T;
    u32 lane;  // 0..T-1
    u32 shift; // the team's first lane in the wave
    __device__ inline NphTeam()
    {
        const u32 l = __lane_id();
        lane = l & (T - 1U);
        shift = l & ~(T - 1U);
    }
    // Bit k: lane k of the team passed (inactive lanes count as false).
    __device__ inline u64 ballot(bool p) const
    {
        return (__ballot(p ? 1 : 0) >> shift) & ((1ULL << T) - 1ULL);
    }
    __device__ inline f32 get(f32 v, u32 src) const { return __shfl(v, (int)src, (int)T); }
    __device__ inline u32 get(u32 v, u32 src) const
    {
        return (u32)__shfl((int)v, (int)src, (int)T);
    }
    __device__ inline NphVec3 get(NphVec3 v, u32 src) const
    {
        return v3(get(v.x, src), get(v.y, src), get(v.z, src));
    }
    // The team's LDS writes before this are seen by its reads after it.
    __device__ inline void sync() const
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // How many of the lanes below this one a ballot holds.
    __device__ inline u32 below(u64 mask) const
    {
        return (u32)__popcll(mask & ((1ULL << lane) - 1ULL));
    }
    // The largest value, ties to the lowest index; entries with index
    // NPH_NULL32 never win. Butterfly over the team: every lane ends with
    // the result.
    __device__ inline NphPick first_max(NphPick p) const
    {
        for (u32 m = T >> 1U; m > 0U; m >>= 1U) {
            const f32 ov = __shfl_xor(p.value, (int)m, (int)T);
            const u32 oi = (u32)__shfl_xor((int)p.index, (int)m, (int)T);
            if (oi != NPH_NULL32 &&
                (p.index == NPH_NULL32 || ov > p.value || (ov == p.value && oi < p.index))) {
                p.value = ov;
                p.index = oi;
            }
        }
        return p;
    }
    // The smallest value, ties to the lowest index.
    __device__ inline NphPick first_min(NphPick p) const
    {
        for (u32 m = T >> 1U; m > 0U; m >>= 1U) {
            const f32 ov = __shfl_xor(p.value, (int)m, (int)T);
            const u32 oi = (u32)__shfl_xor((int)p.index, (int)m, (int)T);
            if (oi != NPH_NULL32 &&
                (p.index == NPH_NULL32 || ov < p.value || (ov == p.value && oi < p.index))) {
                p.value = ov;
                p.index = oi;
            }
        }
        return p;
    }
    // The sum over the team (integers: exact in any order).
    __device__ inline u32 add(u32 v) const
    {
        for (u32 m = T >> 1U; m > 0U; m >>= 1U) {
            v += (u32)__shfl_xor((int)v, (int)m, (int)T);
        }
        return v;
    }
};
#endif

// A scan's running pick merged with a later chunk's (every index of the
// chunk above every index before it): a strictly larger value wins.
NPH_TD NphPick pick_later_max(NphPick run, NphPick chunk)
{
    if (chunk.index != NPH_NULL32 && (run.index == NPH_NULL32 || chunk.value > run.value)) {
        return chunk;
    }
    return run;
}

NPH_TD NphPick pick_later_min(NphPick run, NphPick chunk)
{
    if (chunk.index != NPH_NULL32 && (run.index == NPH_NULL32 || chunk.value < run.value)) {
        return chunk;
    }
    return run;
}

// The lowest set bit's index (mask != 0).
NPH_TD u32 nph_lowest_bit(u64 mask) { return (u32)__builtin_ctzll(mask); }

#endif // __cplusplus
#endif // NPH_TEAM_H
