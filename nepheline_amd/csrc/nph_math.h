// SPDX-License-Identifier: Apache-2.0
// f32/f64 vector, quaternion and matrix arithmetic of the world step, for
// the host C17 code and the HIP kernels alike.
//
// Bit parity with the reference's CPU step needs every IEEE operation in the
// reference's order, so each helper below states its operations explicitly
// (no FMA: the HIP sources are compiled with -ffp-contract=off; no fast
// math; min/max as compare-and-select).  The orders follow
// engine/math/include/nepheline/math/{scalar,vec,quat,mat,geometry}.h; the
// citation sits on each function.
#ifndef NPH_MATH_H
#define NPH_MATH_H

#include "../../include/nepheline_physics.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define NPH_HD __host__ __device__ static inline
#else
#define NPH_HD static inline
#endif

#define NPH_NULL32 0xFFFFFFFFu
#define NPH_F32_MAX 3.402823466e+38F
#define NPH_F32_INF (__builtin_inff())

// scalar.h:59-90
NPH_HD f32 nph_minf(f32 a, f32 b) { return a < b ? a : b; }
NPH_HD f32 nph_maxf(f32 a, f32 b) { return a > b ? a : b; }
NPH_HD f32 nph_clampf(f32 v, f32 lo, f32 hi) { return v < lo ? lo : (v > hi ? hi : v); }
NPH_HD f32 nph_saturate(f32 v) { return nph_clampf(v, 0.0F, 1.0F); }
NPH_HD f32 nph_absf(f32 x) { return __builtin_fabsf(x); }
NPH_HD f32 nph_sqrtf(f32 x) { return __builtin_sqrtf(x); }
// scalar.h:183-186: 1 / sqrt, two correctly rounded operations.
NPH_HD f32 nph_rsqrtf(f32 x) { return 1.0F / __builtin_sqrtf(x); }
NPH_HD u32 nph_minu(u32 a, u32 b) { return a < b ? a : b; }
NPH_HD u32 nph_maxu(u32 a, u32 b) { return a > b ? a : b; }
NPH_HD int nph_isfinitef(f32 x) { return __builtin_isfinite(x); }

NPH_HD u32 nph_f2u(f32 f)
{
    u32 u;
    memcpy(&u, &f, 4);
    return u;
}
NPH_HD f32 nph_u2f(u32 u)
{
    f32 f;
    memcpy(&f, &u, 4);
    return f;
}

// ---------------------------------------------------------------------------
// NphVec3 (vec.h:267-469)
// ---------------------------------------------------------------------------
NPH_HD NphVec3 v3(f32 x, f32 y, f32 z)
{
    NphVec3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
NPH_HD NphVec3 v3_zero(void) { return v3(0.0F, 0.0F, 0.0F); }
NPH_HD NphVec3 v3_splat(f32 s) { return v3(s, s, s); }
NPH_HD f32 v3_at(NphVec3 v, u32 i) { return i == 0U ? v.x : (i == 1U ? v.y : v.z); }
NPH_HD NphVec3 v3_add(NphVec3 a, NphVec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
NPH_HD NphVec3 v3_sub(NphVec3 a, NphVec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
NPH_HD NphVec3 v3_mul(NphVec3 a, NphVec3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
NPH_HD NphVec3 v3_scale(NphVec3 v, f32 s) { return v3(v.x * s, v.y * s, v.z * s); }
// vec.h:292-295: a + b s, per component.
NPH_HD NphVec3 v3_mul_add(NphVec3 a, NphVec3 b, f32 s)
{
    return v3(a.x + b.x * s, a.y + b.y * s, a.z + b.z * s);
}
NPH_HD NphVec3 v3_neg(NphVec3 v) { return v3(-v.x, -v.y, -v.z); }
// vec.h:300-303: (x x' + y y') + z z'.
NPH_HD f32 v3_dot(NphVec3 a, NphVec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
NPH_HD NphVec3 v3_cross(NphVec3 a, NphVec3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
NPH_HD f32 v3_len_sq(NphVec3 v) { return v3_dot(v, v); }
NPH_HD f32 v3_len(NphVec3 v) { return nph_sqrtf(v3_len_sq(v)); }
NPH_HD f32 v3_dist_sq(NphVec3 a, NphVec3 b) { return v3_len_sq(v3_sub(a, b)); }
NPH_HD f32 v3_dist(NphVec3 a, NphVec3 b) { return v3_len(v3_sub(a, b)); }
// vec.h:326-331
NPH_HD NphVec3 v3_normalize(NphVec3 v) { return v3_scale(v, nph_rsqrtf(v3_len_sq(v))); }
// vec.h:337-350
NPH_HD NphVec3 v3_normalize_or(NphVec3 v, NphVec3 fallback)
{
    const f32 len_sq = v3_len_sq(v);
    if (len_sq > 0.0F && len_sq <= NPH_F32_MAX) {
        return v3_scale(v, nph_rsqrtf(len_sq));
    }
    if (!(len_sq > NPH_F32_MAX) || !nph_isfinitef(v.x) || !nph_isfinitef(v.y) ||
        !nph_isfinitef(v.z)) {
        return fallback;
    }
    const f32 m = nph_maxf(nph_absf(v.x), nph_maxf(nph_absf(v.y), nph_absf(v.z)));
    const NphVec3 u = v3_scale(v, 1.0F / m);
    return v3_scale(u, nph_rsqrtf(v3_len_sq(u)));
}
NPH_HD NphVec3 v3_min(NphVec3 a, NphVec3 b)
{
    return v3(nph_minf(a.x, b.x), nph_minf(a.y, b.y), nph_minf(a.z, b.z));
}
NPH_HD NphVec3 v3_max(NphVec3 a, NphVec3 b)
{
    return v3(nph_maxf(a.x, b.x), nph_maxf(a.y, b.y), nph_maxf(a.z, b.z));
}
// vec.h:369-372
NPH_HD NphVec3 v3_clamp(NphVec3 v, NphVec3 lo, NphVec3 hi) { return v3_min(v3_max(v, lo), hi); }
NPH_HD NphVec3 v3_abs(NphVec3 v) { return v3(nph_absf(v.x), nph_absf(v.y), nph_absf(v.z)); }
// vec.h:356-360 with scalar.h's lerp (1 - t) a + t b.
NPH_HD NphVec3 v3_lerp(NphVec3 a, NphVec3 b, f32 t)
{
    return v3((1.0F - t) * a.x + t * b.x, (1.0F - t) * a.y + t * b.y, (1.0F - t) * a.z + t * b.z);
}
// vec.h:412-420
NPH_HD u32 v3_max_index(NphVec3 v)
{
    return v.x >= v.y ? (v.x >= v.z ? 0U : 2U) : (v.y >= v.z ? 1U : 2U);
}
NPH_HD u32 v3_max_abs_index(NphVec3 v) { return v3_max_index(v3_abs(v)); }
// vec.h:421-430: v - n (v.n)
NPH_HD NphVec3 v3_reject(NphVec3 v, NphVec3 n) { return v3_sub(v, v3_scale(n, v3_dot(v, n))); }
// vec.h:460-469 (Hughes and Moeller 1999)
NPH_HD NphVec3 v3_perpendicular(NphVec3 n)
{
    const NphVec3 a = v3_abs(n);
    const NphVec3 axis = (a.x <= a.y && a.x <= a.z) ? v3(1.0F, 0.0F, 0.0F)
                         : (a.y <= a.z)             ? v3(0.0F, 1.0F, 0.0F)
                                                    : v3(0.0F, 0.0F, 1.0F);
    return v3_normalize(v3_cross(n, axis));
}

// f64 (vec.h:899-994)
NPH_HD NphDVec3 dv3(f64 x, f64 y, f64 z)
{
    NphDVec3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
NPH_HD NphDVec3 dv3_add_v3(NphDVec3 a, NphVec3 b)
{
    return dv3(a.x + (f64)b.x, a.y + (f64)b.y, a.z + (f64)b.z);
}
// vec.h:991-994: (f32)(p - origin), per component.
NPH_HD NphVec3 dv3_relative(NphDVec3 p, NphDVec3 origin)
{
    return v3((f32)(p.x - origin.x), (f32)(p.y - origin.y), (f32)(p.z - origin.z));
}

// ---------------------------------------------------------------------------
// NphQuat (quat.h:30-239)
// ---------------------------------------------------------------------------
NPH_HD NphQuat q4(f32 x, f32 y, f32 z, f32 w)
{
    NphQuat q;
    q.x = x;
    q.y = y;
    q.z = z;
    q.w = w;
    return q;
}
NPH_HD NphQuat q_identity(void) { return q4(0.0F, 0.0F, 0.0F, 1.0F); }
NPH_HD NphVec3 q_xyz(NphQuat q) { return v3(q.x, q.y, q.z); }
NPH_HD NphQuat q_add(NphQuat a, NphQuat b) { return q4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
NPH_HD NphQuat q_scale(NphQuat q, f32 s) { return q4(q.x * s, q.y * s, q.z * s, q.w * s); }
NPH_HD NphQuat q_conj(NphQuat q) { return q4(-q.x, -q.y, -q.z, q.w); }
// quat.h:70-76 (Hamilton product a b)
NPH_HD NphQuat q_mul(NphQuat a, NphQuat b)
{
    return q4(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
              a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
              a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
// quat.h:81-88: ((x x' + y y') + z z') + w w'
NPH_HD f32 q_len_sq(NphQuat q) { return q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w; }
NPH_HD NphQuat q_normalize(NphQuat q) { return q_scale(q, nph_rsqrtf(q_len_sq(q))); }
// quat.h:109-113
NPH_HD NphQuat q_normalize_or_identity(NphQuat q)
{
    const f32 len_sq = q_len_sq(q);
    return len_sq > 0.0F ? q_scale(q, nph_rsqrtf(len_sq)) : q_identity();
}
// quat.h:157-164: v + 2 w (u x v) + 2 u x (u x v)
NPH_HD NphVec3 q_rotate(NphQuat q, NphVec3 v)
{
    const NphVec3 u = q_xyz(q);
    const NphVec3 t = v3_scale(v3_cross(u, v), 2.0F);
    return v3_add(v3_add(v, v3_scale(t, q.w)), v3_cross(u, t));
}
NPH_HD NphVec3 q_rotate_inv(NphQuat q, NphVec3 v) { return q_rotate(q_conj(q), v); }
// quat.h:171-185
NPH_HD NphVec3 q_axis_x(NphQuat q)
{
    return v3(1.0F - 2.0F * (q.y * q.y + q.z * q.z), 2.0F * (q.x * q.y + q.w * q.z),
              2.0F * (q.x * q.z - q.w * q.y));
}
NPH_HD NphVec3 q_axis_y(NphQuat q)
{
    return v3(2.0F * (q.x * q.y - q.w * q.z), 1.0F - 2.0F * (q.x * q.x + q.z * q.z),
              2.0F * (q.y * q.z + q.w * q.x));
}
NPH_HD NphVec3 q_axis_z(NphQuat q)
{
    return v3(2.0F * (q.x * q.z + q.w * q.y), 2.0F * (q.y * q.z - q.w * q.x),
              1.0F - 2.0F * (q.x * q.x + q.y * q.y));
}
// quat.h:234-239
NPH_HD NphQuat q_integrate(NphQuat q, NphVec3 omega, f32 dt)
{
    const NphQuat w = q4(omega.x, omega.y, omega.z, 0.0F);
    const NphQuat dq = q_scale(q_mul(w, q), 0.5F * dt);
    return q_normalize_or_identity(q_add(q, dq));
}

// ---------------------------------------------------------------------------
// NphMat3, column-major (mat.h:24-121, mat.c:14-29)
// ---------------------------------------------------------------------------
NPH_HD NphMat3 m3_cols(NphVec3 c0, NphVec3 c1, NphVec3 c2)
{
    NphMat3 m;
    m.c[0] = c0;
    m.c[1] = c1;
    m.c[2] = c2;
    return m;
}
NPH_HD NphMat3 m3_rows(NphVec3 r0, NphVec3 r1, NphVec3 r2)
{
    return m3_cols(v3(r0.x, r1.x, r2.x), v3(r0.y, r1.y, r2.y), v3(r0.z, r1.z, r2.z));
}
NPH_HD NphMat3 m3_zero(void) { return m3_cols(v3_zero(), v3_zero(), v3_zero()); }
NPH_HD NphMat3 m3_diag(NphVec3 d)
{
    return m3_cols(v3(d.x, 0.0F, 0.0F), v3(0.0F, d.y, 0.0F), v3(0.0F, 0.0F, d.z));
}
NPH_HD NphMat3 m3_from_quat(NphQuat q) { return m3_cols(q_axis_x(q), q_axis_y(q), q_axis_z(q)); }
NPH_HD NphMat3 m3_add(NphMat3 a, NphMat3 b)
{
    return m3_cols(v3_add(a.c[0], b.c[0]), v3_add(a.c[1], b.c[1]), v3_add(a.c[2], b.c[2]));
}
NPH_HD NphMat3 m3_scale(NphMat3 m, f32 s)
{
    return m3_cols(v3_scale(m.c[0], s), v3_scale(m.c[1], s), v3_scale(m.c[2], s));
}
// mat.h:103-107: (c0 x + c1 y) + c2 z
NPH_HD NphVec3 m3_mul_v3(NphMat3 m, NphVec3 v)
{
    return v3_add(v3_add(v3_scale(m.c[0], v.x), v3_scale(m.c[1], v.y)), v3_scale(m.c[2], v.z));
}
NPH_HD NphMat3 m3_mul(NphMat3 a, NphMat3 b)
{
    return m3_cols(m3_mul_v3(a, b.c[0]), m3_mul_v3(a, b.c[1]), m3_mul_v3(a, b.c[2]));
}
NPH_HD NphMat3 m3_transpose(NphMat3 m) { return m3_rows(m.c[0], m.c[1], m.c[2]); }
// mat.c:14-29: the adjugate's rows over the determinant; false when |det| < 1e-20.
NPH_HD int m3_inverse(NphMat3 m, NphMat3 *out)
{
    const NphVec3 r0 = v3_cross(m.c[1], m.c[2]);
    const NphVec3 r1 = v3_cross(m.c[2], m.c[0]);
    const NphVec3 r2 = v3_cross(m.c[0], m.c[1]);
    const f32 det = v3_dot(m.c[0], r0);
    if (nph_absf(det) < 1e-20F) {
        return 0;
    }
    const f32 inv_det = 1.0F / det;
    *out = m3_rows(v3_scale(r0, inv_det), v3_scale(r1, inv_det), v3_scale(r2, inv_det));
    return 1;
}

// ---------------------------------------------------------------------------
// Poses, planes, boxes (physics/types.h:60-120, math/geometry.h:77-292)
// ---------------------------------------------------------------------------
NPH_HD NphPhysicsPose pose_make(NphVec3 p, NphQuat q)
{
    NphPhysicsPose r;
    r.position = p;
    r.rotation = q;
    return r;
}
NPH_HD NphPhysicsPose pose_identity(void) { return pose_make(v3_zero(), q_identity()); }
NPH_HD NphVec3 pose_point(NphPhysicsPose t, NphVec3 p)
{
    return v3_add(q_rotate(t.rotation, p), t.position);
}
NPH_HD NphVec3 pose_vector(NphPhysicsPose t, NphVec3 v) { return q_rotate(t.rotation, v); }
NPH_HD NphVec3 pose_inv_point(NphPhysicsPose t, NphVec3 p)
{
    return q_rotate_inv(t.rotation, v3_sub(p, t.position));
}
// types.h: the pose of b in a's frame.
NPH_HD NphPhysicsPose pose_inv_mul(NphPhysicsPose a, NphPhysicsPose b)
{
    return pose_make(pose_inv_point(a, b.position), q_mul(q_conj(a.rotation), b.rotation));
}

NPH_HD NphPlane plane_make(NphVec3 n, f32 d)
{
    NphPlane p;
    p.normal = n;
    p.d = d;
    return p;
}
NPH_HD f32 plane_distance(NphPlane p, NphVec3 x) { return v3_dot(p.normal, x) + p.d; }

NPH_HD NphAABB aabb_make(NphVec3 lo, NphVec3 hi)
{
    NphAABB b;
    b.min = lo;
    b.max = hi;
    return b;
}
NPH_HD NphVec3 aabb_center(NphAABB b) { return v3_scale(v3_add(b.min, b.max), 0.5F); }
NPH_HD NphAABB aabb_grow(NphAABB b, f32 m)
{
    const NphVec3 mm = v3_splat(m);
    return aabb_make(v3_sub(b.min, mm), v3_add(b.max, mm));
}
NPH_HD NphAABB aabb_union(NphAABB a, NphAABB b)
{
    return aabb_make(v3_min(a.min, b.min), v3_max(a.max, b.max));
}
NPH_HD int aabb_overlaps(NphAABB a, NphAABB b)
{
    return a.min.x <= b.max.x && b.min.x <= a.max.x && a.min.y <= b.max.y && b.min.y <= a.max.y &&
           a.min.z <= b.max.z && b.min.z <= a.max.z;
}
NPH_HD int aabb_contains(NphAABB outer, NphAABB inner)
{
    return outer.min.x <= inner.min.x && outer.min.y <= inner.min.y &&
           outer.min.z <= inner.min.z && inner.max.x <= outer.max.x &&
           inner.max.y <= outer.max.y && inner.max.z <= outer.max.z;
}

// f64 to f32 rounded towards -inf / +inf (shape.c:71-89): one step of the
// f32 grid is added or removed through the bit pattern (nextafterf's effect
// for finite values).
NPH_HD f32 nph_next_down(f32 f)
{
    if (f != f || f == -NPH_F32_INF) {
        return f;
    }
    if (f == 0.0F) {
        return -1.40129846e-45F; // smallest negative subnormal
    }
    const u32 u = nph_f2u(f);
    return nph_u2f(f > 0.0F ? u - 1U : u + 1U);
}
NPH_HD f32 nph_next_up(f32 f)
{
    if (f != f || f == NPH_F32_INF) {
        return f;
    }
    if (f == 0.0F) {
        return 1.40129846e-45F;
    }
    const u32 u = nph_f2u(f);
    return nph_u2f(f > 0.0F ? u + 1U : u - 1U);
}
NPH_HD f32 nph_round_down(f64 x)
{
    f32 f = (f32)x;
    if ((f64)f > x) {
        f = nph_next_down(f);
    }
    return f;
}
NPH_HD f32 nph_round_up(f64 x)
{
    f32 f = (f32)x;
    if ((f64)f < x) {
        f = nph_next_up(f);
    }
    return f;
}
// shape.c:91-97: box `rel` around `origin`, rounded outward to f32.
NPH_HD NphAABB nph_world_box(NphDVec3 origin, NphAABB rel)
{
    NphAABB b;
    b.min = v3(nph_round_down(origin.x + (f64)rel.min.x), nph_round_down(origin.y + (f64)rel.min.y),
               nph_round_down(origin.z + (f64)rel.min.z));
    b.max = v3(nph_round_up(origin.x + (f64)rel.max.x), nph_round_up(origin.y + (f64)rel.max.y),
               nph_round_up(origin.z + (f64)rel.max.z));
    return b;
}

// physics_internal.h:286-291
NPH_HD u32 nph_mix_id(u32 a, u32 b)
{
    u32 h = a * 0x9E3779B1U;
    h ^= b + 0x7F4A7C15U + (h << 6U) + (h >> 2U);
    return h;
}

// ---------------------------------------------------------------------------
// atan2f (scalar.h:239-242 calls the C library's).  The reference's C library
// (glibc, flt-32) computes it by the method of Sun's fdlibm, float version:
// atanf reduces |x| onto one of four breakpoints (atan 0.5, 1, 1.5, inf, each
// a hi + lo pair), evaluates an 11-term odd polynomial split into odd and
// even parts, and returns the constant for |x| >= 2^25; atan2f handles the
// special operands, then atanf(|y/x|) and the quadrant with pi as hi + lo.
// Written here from that published method; it agrees with glibc 2.35's atanf
// on all 2^32 inputs and with its atan2f on the sampled pairs of the tests.
// ---------------------------------------------------------------------------
NPH_HD f32 nph_atanf(f32 x)
{
    const i32 hx = (i32)nph_f2u(x);
    const i32 ix = hx & 0x7fffffff;
    if (ix >= 0x4c000000) { // |x| >= 2^25, or NaN
        if (ix > 0x7f800000) {
            return x + x;
        }
        return hx > 0 ? 1.5707962513e+00F + 7.5497894159e-08F
                      : -1.5707962513e+00F - 7.5497894159e-08F;
    }
    i32 id;
    if (ix < 0x3ee00000) { // |x| < 7/16
        if (ix < 0x31000000) { // |x| < 2^-29
            return x;
        }
        id = -1;
    } else {
        x = nph_absf(x);
        if (ix < 0x3f980000) {     // |x| < 19/16
            if (ix < 0x3f300000) { // 7/16 <= |x| < 11/16
                id = 0;
                x = (2.0F * x - 1.0F) / (2.0F + x);
            } else { // 11/16 <= |x| < 19/16
                id = 1;
                x = (x - 1.0F) / (x + 1.0F);
            }
        } else if (ix < 0x401c0000) { // |x| < 39/16
            id = 2;
            x = (x - 1.5F) / (1.0F + 1.5F * x);
        } else { // 39/16 <= |x| < 2^25
            id = 3;
            x = -1.0F / x;
        }
    }
    const f32 z = x * x;
    const f32 w = z * z;
    const f32 s1 =
        z * (3.3333334327e-01F +
             w * (1.4285714924e-01F +
                  w * (9.0908870101e-02F +
                       w * (6.6610731184e-02F + w * (4.9768779427e-02F + w * 1.6285819933e-02F)))));
    const f32 s2 = w * (-2.0000000298e-01F +
                        w * (-1.1111110449e-01F +
                             w * (-7.6918758452e-02F +
                                  w * (-5.8335702866e-02F + w * -3.6531571299e-02F))));
    if (id < 0) {
        return x - x * (s1 + s2);
    }
    const f32 hi = id == 0 ? 4.6364760399e-01F
                   : id == 1 ? 7.8539812565e-01F
                   : id == 2 ? 9.8279368877e-01F
                             : 1.5707962513e+00F;
    const f32 lo = id == 0 ? 5.0121582440e-09F
                   : id == 1 ? 3.7748947079e-08F
                   : id == 2 ? 3.4473217170e-08F
                             : 7.5497894159e-08F;
    const f32 r = hi - ((x * (s1 + s2) - lo) - x);
    return hx < 0 ? -r : r;
}

NPH_HD f32 nph_atan2f(f32 y, f32 x)
{
    const f32 pi_o_4 = 7.8539818525e-01F;
    const f32 pi_o_2 = 1.5707963705e+00F;
    const f32 pi = 3.1415927410e+00F;
    const f32 pi_lo = -8.7422776573e-08F;
    const f32 tiny = 1.0e-30F;
    const i32 hx = (i32)nph_f2u(x);
    const i32 ix = hx & 0x7fffffff;
    const i32 hy = (i32)nph_f2u(y);
    const i32 iy = hy & 0x7fffffff;
    if (ix > 0x7f800000 || iy > 0x7f800000) {
        return x + y;
    }
    if (hx == 0x3f800000) { // x = 1
        return nph_atanf(y);
    }
    const i32 m = ((hy >> 31) & 1) | ((hx >> 30) & 2); // 2 sign(x) + sign(y)
    if (iy == 0) {
        return m <= 1 ? y : (m == 2 ? pi + tiny : -pi - tiny);
    }
    if (ix == 0) {
        return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
    }
    if (ix == 0x7f800000) {
        if (iy == 0x7f800000) {
            return m == 0   ? pi_o_4 + tiny
                   : m == 1 ? -pi_o_4 - tiny
                   : m == 2 ? 3.0F * pi_o_4 + tiny
                            : -3.0F * pi_o_4 - tiny;
        }
        return m == 0 ? 0.0F : (m == 1 ? -0.0F : (m == 2 ? pi + tiny : -pi - tiny));
    }
    if (iy == 0x7f800000) {
        return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
    }
    const i32 k = (iy - ix) >> 23;
    f32 z;
    if (k > 26) { // |y/x| > 2^26
        z = pi_o_2 + 0.5F * pi_lo;
    } else if (hx < 0 && k < -26) {
        z = 0.0F;
    } else {
        z = nph_atanf(nph_absf(y / x));
    }
    switch (m) {
    case 0:
        return z;
    case 1:
        return nph_u2f(nph_f2u(z) ^ 0x80000000U);
    case 2:
        return pi - (z - pi_lo);
    default:
        return (z - pi_lo) - pi;
    }
}

// ---------------------------------------------------------------------------
// Hull block accessors (hull.h:80-105)
// ---------------------------------------------------------------------------
NPH_HD const NphVec3 *hull_vertices(const NphPhysicsHull *h)
{
    return (const NphVec3 *)((const u8 *)h + h->vertices_offset);
}
NPH_HD const NphPlane *hull_planes(const NphPhysicsHull *h)
{
    return (const NphPlane *)((const u8 *)h + h->planes_offset);
}
NPH_HD const u16 *hull_faces(const NphPhysicsHull *h)
{
    return (const u16 *)((const u8 *)h + h->faces_offset);
}
NPH_HD const NphPhysicsHullEdge *hull_edges(const NphPhysicsHull *h)
{
    return (const NphPhysicsHullEdge *)((const u8 *)h + h->edges_offset);
}

#endif // NPH_MATH_H
