// SPDX-License-Identifier: Apache-2.0
// The learned warm start's network on gfx950: a 16 -> 64 -> 64 -> 16 MLP
// (14 contact features in, zero-padded; GELU; 3 outputs used, the normal
// and two friction impulses) with BF16 weights and activations and FP32
// accumulation on the matrix cores (v_mfma_f32_16x16x16_bf16).  Three
// kernels: inference, the loss gradient of a minibatch (forward and
// backward fused, weight gradients accumulated across a wave's tiles in
// MFMA accumulators), and Adam.
//
// A wave works on 16 rows (contacts) at a time.  For the 16x16x16 MFMA,
// lane l holds A[l % 16][4 (l / 16) + j], B[4 (l / 16) + j][l % 16] and
// D[4 (l / 16) + j][l % 16], j = 0..3.  An accumulator in D layout is
// therefore already a B operand.  Inference computes each layer
// transposed so that activations stay in registers as B operands and the
// weights, A operands, stay in VGPRs (k_mlp_forward).  The gradient kernel
// passes A operands through the wave's LDS tiles, with the weights rounded
// to BF16 into LDS once per block.
//
// Parameters are one FP32 array (the master copy Adam updates):
//   W1 [16][64], b1 [64], W2 [64][64], b2 [64], W3 [64][16], b3 [16]
// row-major (input index first).  Rows 14 and 15 of W1 and columns 3..15
// of W3 stay zero (their gradients are masked).

#include <hip/hip_runtime.h>

#include "../../include/nepheline_physics.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

#define IN 16
#define HID 64
#define OUT 16
#define NOUT 3
#define P_W1 0
#define P_B1 (P_W1 + IN * HID)
#define P_W2 (P_B1 + HID)
#define P_B2 (P_W2 + HID * HID)
#define P_W3 (P_B2 + HID)
#define P_B3 (P_W3 + HID * OUT)
#define P_COUNT (P_B3 + OUT)
#define WAVES 4
#define BLOCK (64 * WAVES)

static_assert(P_COUNT == NPH_AMD_MLP_PARAMS, "parameter layout");

__device__ static inline unsigned short bf16_of(float f)
{
    // Round to nearest even (finite inputs).
    unsigned u = __float_as_uint(f);
    u += 0x7FFFU + ((u >> 16U) & 1U);
    return (unsigned short)(u >> 16U);
}

// The standard normal distribution at x: its CDF by Hastings' rational
// approximation (Abramowitz and Stegun 26.2.17: for a >= 0,
// Phi(-a) = phi(a) (b1 t + b2 t^2 + ... + b5 t^5), t = 1 / (1 + p a),
// |error| < 7.5e-8) and its density phi. The exponential and the
// reciprocal run on the transcendental unit (v_exp_f32, v_rcp_f32), with
// no branch; against the exact CDF the result stays within 5e-7 (float32
// rounding included).
struct Gauss {
    float cdf;
    float pdf;
};

#define HASTINGS_P 0.2316419F
#define HASTINGS_B1 0.319381530F
#define HASTINGS_B2 -0.356563782F
#define HASTINGS_B3 1.781477937F
#define HASTINGS_B4 -1.821255978F
#define HASTINGS_B5 1.330274429F
#define INV_SQRT_2PI 0.39894228040F
#define HALF_LOG2E 0.72134752044F // exp(-a^2 / 2) = 2^(-a^2 log2(e) / 2)

__device__ static inline Gauss gauss(float x)
{
    const float a = fabsf(x);
    const float pdf = INV_SQRT_2PI * __builtin_amdgcn_exp2f(-HALF_LOG2E * a * a);
    const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(HASTINGS_P, a, 1.0F));
    float poly = __builtin_fmaf(t, HASTINGS_B5, HASTINGS_B4);
    poly = __builtin_fmaf(t, poly, HASTINGS_B3);
    poly = __builtin_fmaf(t, poly, HASTINGS_B2);
    poly = __builtin_fmaf(t, poly, HASTINGS_B1);
    const float below = pdf * (t * poly); // Phi(-a)
    return Gauss{x >= 0.0F ? 1.0F - below : below, pdf};
}

// GELU in its erf form, x Phi(x), and its derivative Phi(x) + x phi(x).
// With a = |x|: x Phi(x) = max(x, 0) - a Phi(-a), and the density's
// constant is folded into the polynomial (the inference kernel's form).
__device__ static inline float gelu(float x)
{
    const float a = fabsf(x);
    const float e = __builtin_amdgcn_exp2f(-HALF_LOG2E * a * a);
    const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(HASTINGS_P, a, 1.0F));
    float poly = __builtin_fmaf(t, INV_SQRT_2PI * HASTINGS_B5, INV_SQRT_2PI * HASTINGS_B4);
    poly = __builtin_fmaf(t, poly, INV_SQRT_2PI * HASTINGS_B3);
    poly = __builtin_fmaf(t, poly, INV_SQRT_2PI * HASTINGS_B2);
    poly = __builtin_fmaf(t, poly, INV_SQRT_2PI * HASTINGS_B1);
    return __builtin_fmaf(-a, e * (t * poly), x > 0.0F ? x : 0.0F);
}

__device__ static inline float gelu_grad(float x)
{
    const Gauss g = gauss(x);
    return g.cdf + x * g.pdf;
}

// The same two through the device library's erff and expf, for checking
// the approximation.
__device__ static inline float gelu_erf(float x) { return 0.5F * x * (1.0F + erff(x * 0.70710678118F)); }

__device__ static inline float gelu_grad_erf(float x)
{
    const float cdf = 0.5F * (1.0F + erff(x * 0.70710678118F));
    return cdf + x * (0.39894228040F * expf(-0.5F * x * x));
}

struct MlpShared {
    unsigned short w1[IN * HID];
    unsigned short w2[HID * HID];
    unsigned short w3[HID * OUT];
    float b1[HID];
    float b2[HID];
    float b3[OUT];
    unsigned short ta[WAVES][16 * HID]; // a wave's activations, [16 rows][64]
    unsigned short tb[WAVES][16 * HID];
    unsigned short tc[WAVES][16 * OUT];
};

__device__ static void load_params(MlpShared *s, const float *p)
{
    for (unsigned i = threadIdx.x; i < IN * HID; i += BLOCK) {
        s->w1[i] = bf16_of(p[P_W1 + i]);
    }
    for (unsigned i = threadIdx.x; i < HID * HID; i += BLOCK) {
        s->w2[i] = bf16_of(p[P_W2 + i]);
    }
    for (unsigned i = threadIdx.x; i < HID * OUT; i += BLOCK) {
        s->w3[i] = bf16_of(p[P_W3 + i]);
    }
    for (unsigned i = threadIdx.x; i < HID; i += BLOCK) {
        s->b1[i] = p[P_B1 + i];
        s->b2[i] = p[P_B2 + i];
    }
    for (unsigned i = threadIdx.x; i < OUT; i += BLOCK) {
        s->b3[i] = p[P_B3 + i];
    }
    __syncthreads();
}

// The wave's LDS writes visible to its own later reads.
__device__ static inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ static inline v4f mfma(v4s a, v4s b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// A[m][k] from a row-major [16][ld] tile, k block kb.
__device__ static inline v4s frag_a(const unsigned short *t, unsigned ld, unsigned kb, unsigned l)
{
    const unsigned short *r = t + (l & 15U) * ld + kb * 16U + 4U * (l >> 4U);
    return v4s{(short)r[0], (short)r[1], (short)r[2], (short)r[3]};
}

// A[m][k] = T[k][fb*16 + m] from a row-major [16][ld] tile (its transpose).
__device__ static inline v4s frag_at(const unsigned short *t, unsigned ld, unsigned fb, unsigned l)
{
    const unsigned c = fb * 16U + (l & 15U);
    const unsigned r0 = 4U * (l >> 4U);
    return v4s{(short)t[r0 * ld + c], (short)t[(r0 + 1U) * ld + c], (short)t[(r0 + 2U) * ld + c],
               (short)t[(r0 + 3U) * ld + c]};
}

// B[k][n] from a row-major [K][ldn] matrix, blocks kb, nb.
__device__ static inline v4s frag_b(const unsigned short *w, unsigned ldn, unsigned kb, unsigned nb,
                                    unsigned l)
{
    const unsigned k0 = kb * 16U + 4U * (l >> 4U);
    const unsigned n = nb * 16U + (l & 15U);
    return v4s{(short)w[k0 * ldn + n], (short)w[(k0 + 1U) * ldn + n],
               (short)w[(k0 + 2U) * ldn + n], (short)w[(k0 + 3U) * ldn + n]};
}

// B[k][n] = W[n][k] from a row-major [N][ldk] matrix (its transpose).
__device__ static inline v4s frag_bt(const unsigned short *w, unsigned ldk, unsigned kb, unsigned nb,
                                     unsigned l)
{
    const unsigned short *r = w + (nb * 16U + (l & 15U)) * ldk + kb * 16U + 4U * (l >> 4U);
    return v4s{(short)r[0], (short)r[1], (short)r[2], (short)r[3]};
}

__device__ static inline v4s to_b(v4f d)
{
    return v4s{(short)bf16_of(d[0]), (short)bf16_of(d[1]), (short)bf16_of(d[2]),
               (short)bf16_of(d[3])};
}

// The forward pass of 16 rows from r0 (rows past n are zero): pre-
// activations z1, z2 (bias added) kept, H1 in ta, H2 in tb, the output y.
__device__ static void forward_tile(const MlpShared *s, unsigned short *ta, unsigned short *tb,
                                    const unsigned short *x, unsigned r0, unsigned n, unsigned l,
                                    v4s *xa, v4f z1[4], v4f z2[4], v4f *y)
{
    const unsigned row = r0 + (l & 15U);
    if (row < n) {
        const unsigned short *xr = x + (unsigned long)row * IN + 4U * (l >> 4U);
        *xa = v4s{(short)xr[0], (short)xr[1], (short)xr[2], (short)xr[3]};
    } else {
        *xa = v4s{0, 0, 0, 0};
    }
    const v4f zero = {0.0F, 0.0F, 0.0F, 0.0F};
    for (unsigned nb = 0U; nb < 4U; ++nb) {
        z1[nb] = mfma(*xa, frag_b(s->w1, HID, 0U, nb, l), zero);
        const unsigned col = nb * 16U + (l & 15U);
        for (unsigned j = 0U; j < 4U; ++j) {
            z1[nb][j] += s->b1[col];
            ta[(4U * (l >> 4U) + j) * HID + col] = bf16_of(gelu(z1[nb][j]));
        }
    }
    wave_sync();
    for (unsigned nb = 0U; nb < 4U; ++nb) {
        z2[nb] = zero;
    }
    for (unsigned kb = 0U; kb < 4U; ++kb) {
        const v4s a = frag_a(ta, HID, kb, l);
        for (unsigned nb = 0U; nb < 4U; ++nb) {
            z2[nb] = mfma(a, frag_b(s->w2, HID, kb, nb, l), z2[nb]);
        }
    }
    for (unsigned nb = 0U; nb < 4U; ++nb) {
        const unsigned col = nb * 16U + (l & 15U);
        for (unsigned j = 0U; j < 4U; ++j) {
            z2[nb][j] += s->b2[col];
            tb[(4U * (l >> 4U) + j) * HID + col] = bf16_of(gelu(z2[nb][j]));
        }
    }
    wave_sync();
    v4f acc = zero;
    for (unsigned kb = 0U; kb < 4U; ++kb) {
        acc = mfma(frag_a(tb, HID, kb, l), frag_b(s->w3, OUT, kb, 0U, l), acc);
    }
    for (unsigned j = 0U; j < 4U; ++j) {
        acc[j] += s->b3[l & 15U];
    }
    *y = acc;
}

// GELU of four accumulators, rounded to BF16 (nearest even) two at a
// time by v_cvt_pk_bf16_f32.
__device__ static inline v4s activate(v4f z)
{
    const v4f g = {gelu(z[0]), gelu(z[1]), gelu(z[2]), gelu(z[3])};
    return __builtin_bit_cast(v4s, __builtin_convertvector(g, v4bf));
}

// A operand of W^T from a row-major [K][ldn] FP32 matrix W, rounded to
// BF16: A[m][k] = W[kb*16 + k][mb*16 + m], i.e. W[kb*16 + q + j][mb*16 + c].
__device__ static inline v4s frag_wt(const float *w, unsigned ldn, unsigned kb, unsigned mb,
                                     unsigned l)
{
    const float *p = w + (kb * 16U + 4U * (l >> 4U)) * ldn + mb * 16U + (l & 15U);
    return v4s{(short)bf16_of(p[0]), (short)bf16_of(p[ldn]), (short)bf16_of(p[2U * ldn]),
               (short)bf16_of(p[3U * ldn])};
}

// Inference: y[r][0..3] for r < n (x: n rows of 16 BF16), output 3 zero.
//
// Each layer is computed transposed, Z^T = W^T H^T, on 16 rows at a time:
// the weights are the A operands, held in VGPRs for the wave's whole run
// (24 fragments, 48 registers); the rows are the B operands. A D-layout
// accumulator (lane l: Z^T[q + j][c], q = 4 (l / 16), c = l % 16) biased,
// activated and rounded to BF16 is the next layer's B operand for the same
// 16 features as it stands, so the layers chain in registers with no LDS
// and no synchronisation. The input tile is the B operand of layer one as
// loaded (lane l: row c, features q..q+3: 8 contiguous bytes), and lanes
// 0..15 end holding outputs 0..3 of row c (one 16-byte store each).
__global__ void __launch_bounds__(BLOCK) k_mlp_forward(const unsigned short *x, unsigned n,
                                                       const float *params, float *y)
{
    const unsigned l = threadIdx.x & 63U;
    const unsigned c = l & 15U;
    const unsigned q = 4U * (l >> 4U);
    v4s w1[4], w2[4][4], w3[4];
    v4f b1[4], b2[4], b3; // biases in D layout: the accumulators' start
    for (unsigned mb = 0U; mb < 4U; ++mb) {
        w1[mb] = frag_wt(params + P_W1, HID, 0U, mb, l);
        w3[mb] = frag_wt(params + P_W3, OUT, mb, 0U, l);
        for (unsigned kb = 0U; kb < 4U; ++kb) {
            w2[mb][kb] = frag_wt(params + P_W2, HID, kb, mb, l);
        }
        for (unsigned j = 0U; j < 4U; ++j) {
            b1[mb][j] = params[P_B1 + mb * 16U + q + j];
            b2[mb][j] = params[P_B2 + mb * 16U + q + j];
        }
        b3[mb] = params[P_B3 + q + mb];
    }
    const unsigned waves = gridDim.x * WAVES;
    const unsigned tiles = (n + 15U) / 16U;
    for (unsigned t = blockIdx.x * WAVES + (threadIdx.x >> 6U); t < tiles; t += waves) {
        const unsigned row = t * 16U + c;
        v4s xb = v4s{0, 0, 0, 0};
        if (row < n) {
            xb = *(const v4s *)(x + (unsigned long)row * IN + q);
        }
        v4s h1[4];
        for (unsigned mb = 0U; mb < 4U; ++mb) {
            const v4f z = mfma(w1[mb], xb, b1[mb]);
            h1[mb] = activate(z);
        }
        v4f z2[4] = {b2[0], b2[1], b2[2], b2[3]};
        for (unsigned kb = 0U; kb < 4U; ++kb) {
            for (unsigned mb = 0U; mb < 4U; ++mb) {
                z2[mb] = mfma(w2[mb][kb], h1[kb], z2[mb]);
            }
        }
        v4s h2[4];
        for (unsigned mb = 0U; mb < 4U; ++mb) {
            h2[mb] = activate(z2[mb]);
        }
        v4f o = b3;
        for (unsigned kb = 0U; kb < 4U; ++kb) {
            o = mfma(w3[kb], h2[kb], o);
        }
        if (q == 0U && row < n) {
            *(float4 *)(y + (unsigned long)row * 4U) = make_float4(o[0], o[1], o[2], 0.0F);
        }
    }
}

// The GELU and its derivative at n points, by the approximation the
// network uses and through erff, for checking one against the other.
__global__ void __launch_bounds__(256) k_mlp_gelu(const float *x, unsigned n, float *out)
{
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float v = x[i];
        out[4UL * i] = gelu(v);
        out[4UL * i + 1U] = gelu_grad(v);
        out[4UL * i + 2U] = gelu_erf(v);
        out[4UL * i + 3U] = gelu_grad_erf(v);
    }
}

// The gradient of L = sum over rows < n and outputs < 3 of (y - t)^2 times
// `scale`, added into grad[P_COUNT]; sum of squares times `scale` into
// *loss.
__global__ void __launch_bounds__(BLOCK) k_mlp_grad(const unsigned short *x, const float *target,
                                                    unsigned n, const float *params, float scale,
                                                    float *grad, float *loss)
{
    __shared__ MlpShared s;
    load_params(&s, params);
    const unsigned l = threadIdx.x & 63U;
    const unsigned w = threadIdx.x >> 6U;
    const unsigned waves = gridDim.x * WAVES;
    const unsigned tiles = (n + 15U) / 16U;
    const v4f zero = {0.0F, 0.0F, 0.0F, 0.0F};
    v4f gw1[4], gw2[4][4], gw3[4];
    float gb1[4], gb2[4], gb3 = 0.0F, lsum = 0.0F;
    for (unsigned i = 0U; i < 4U; ++i) {
        gw1[i] = zero;
        gw3[i] = zero;
        gb1[i] = 0.0F;
        gb2[i] = 0.0F;
        for (unsigned k = 0U; k < 4U; ++k) {
            gw2[i][k] = zero;
        }
    }
    unsigned short *ta = s.ta[w];
    unsigned short *tb = s.tb[w];
    unsigned short *tc = s.tc[w];
    for (unsigned t = blockIdx.x * WAVES + w; t < tiles; t += waves) {
        const unsigned r0 = t * 16U;
        v4s xa;
        v4f z1[4], z2[4], y;
        forward_tile(&s, ta, tb, x, r0, n, l, &xa, z1, z2, &y);
        // dL/dy, in D layout: row 4 (l / 16) + j, column l % 16.
        const unsigned col = l & 15U;
        v4f dy = zero;
        for (unsigned j = 0U; j < 4U; ++j) {
            const unsigned row = r0 + 4U * (l >> 4U) + j;
            if (row < n && col < NOUT) {
                const float e = y[j] - target[(unsigned long)row * 4U + col];
                lsum += e * e;
                dy[j] = 2.0F * e * scale;
            }
            gb3 += dy[j];
            tc[(4U * (l >> 4U) + j) * OUT + col] = bf16_of(dy[j]);
        }
        const v4s dyb = to_b(dy);
        // dW3 = H2^T dY.
        for (unsigned fb = 0U; fb < 4U; ++fb) {
            gw3[fb] = mfma(frag_at(tb, HID, fb, l), dyb, gw3[fb]);
        }
        wave_sync();
        // dH2 = dY W3^T, then through the GELU.
        const v4s dya = frag_a(tc, OUT, 0U, l);
        v4f dz2[4];
        for (unsigned nb = 0U; nb < 4U; ++nb) {
            dz2[nb] = mfma(dya, frag_bt(s.w3, OUT, 0U, nb, l), zero);
            for (unsigned j = 0U; j < 4U; ++j) {
                dz2[nb][j] *= gelu_grad(z2[nb][j]);
                gb2[nb] += dz2[nb][j];
            }
        }
        // dW2 = H1^T dZ2.
        for (unsigned fb = 0U; fb < 4U; ++fb) {
            const v4s a = frag_at(ta, HID, fb, l);
            for (unsigned nb = 0U; nb < 4U; ++nb) {
                gw2[fb][nb] = mfma(a, to_b(dz2[nb]), gw2[fb][nb]);
            }
        }
        // dH1 = dZ2 W2^T (dZ2 to A layout through the tile H2 used).
        wave_sync();
        for (unsigned nb = 0U; nb < 4U; ++nb) {
            const unsigned c = nb * 16U + col;
            for (unsigned j = 0U; j < 4U; ++j) {
                tb[(4U * (l >> 4U) + j) * HID + c] = bf16_of(dz2[nb][j]);
            }
        }
        wave_sync();
        v4f dz1[4];
        for (unsigned nb = 0U; nb < 4U; ++nb) {
            dz1[nb] = zero;
        }
        for (unsigned kb = 0U; kb < 4U; ++kb) {
            const v4s a = frag_a(tb, HID, kb, l);
            for (unsigned nb = 0U; nb < 4U; ++nb) {
                dz1[nb] = mfma(a, frag_bt(s.w2, HID, kb, nb, l), dz1[nb]);
            }
        }
        for (unsigned nb = 0U; nb < 4U; ++nb) {
            for (unsigned j = 0U; j < 4U; ++j) {
                dz1[nb][j] *= gelu_grad(z1[nb][j]);
                gb1[nb] += dz1[nb][j];
            }
        }
        // dW1 = X^T dZ1: A[k][row] = X[row][k].
        v4s xt;
        {
            const unsigned r = r0 + 4U * (l >> 4U);
            short e[4];
            for (unsigned j = 0U; j < 4U; ++j) {
                e[j] = r + j < n ? (short)x[(unsigned long)(r + j) * IN + col] : (short)0;
            }
            xt = v4s{e[0], e[1], e[2], e[3]};
        }
        for (unsigned nb = 0U; nb < 4U; ++nb) {
            gw1[nb] = mfma(xt, to_b(dz1[nb]), gw1[nb]);
        }
        wave_sync();
    }
    // The wave's sums into the global gradient.
    const unsigned col = l & 15U;
    const unsigned r4 = 4U * (l >> 4U);
    for (unsigned nb = 0U; nb < 4U; ++nb) {
        for (unsigned j = 0U; j < 4U; ++j) {
            const unsigned k = r4 + j;
            if (k < 14U) {
                atomicAdd(&grad[P_W1 + k * HID + nb * 16U + col], gw1[nb][j]);
            }
        }
        atomicAdd(&grad[P_B1 + nb * 16U + col], gb1[nb]);
        atomicAdd(&grad[P_B2 + nb * 16U + col], gb2[nb]);
    }
    for (unsigned fb = 0U; fb < 4U; ++fb) {
        for (unsigned nb = 0U; nb < 4U; ++nb) {
            for (unsigned j = 0U; j < 4U; ++j) {
                atomicAdd(&grad[P_W2 + (fb * 16U + r4 + j) * HID + nb * 16U + col], gw2[fb][nb][j]);
            }
        }
        if (col < NOUT) {
            for (unsigned j = 0U; j < 4U; ++j) {
                atomicAdd(&grad[P_W3 + (fb * 16U + r4 + j) * OUT + col], gw3[fb][j]);
            }
        }
    }
    if (col < NOUT) {
        atomicAdd(&grad[P_B3 + col], gb3);
    }
    atomicAdd(loss, lsum * scale);
}

// Adam (Kingma and Ba 2015) on the FP32 master parameters; step t from 1.
__global__ void __launch_bounds__(256) k_mlp_adam(float *p, const float *g, float *m, float *v,
                                                  unsigned count, float lr, float beta1,
                                                  float beta2, float eps, unsigned t)
{
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        const float gi = g[i];
        const float mi = beta1 * m[i] + (1.0F - beta1) * gi;
        const float vi = beta2 * v[i] + (1.0F - beta2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / (1.0F - powf(beta1, (float)t));
        const float vh = vi / (1.0F - powf(beta2, (float)t));
        p[i] -= lr * mh / (sqrtf(vh) + eps);
    }
}

#define MLP_TRY(expr)                                                                              \
    do {                                                                                           \
        const hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                                    \
            fprintf(stderr, "nepheline_amd: %s failed: %s\n", #expr, hipGetErrorString(e_));       \
            return e_ == hipErrorOutOfMemory ? NPH_RESULT_OUT_OF_MEMORY : NPH_RESULT_DEVICE_LOST;  \
        }                                                                                          \
    } while (0)

// ---------------------------------------------------------------------------
// Training sets: rows of 16 features (slot 15 the carried-nothing flag) and
// 4 targets (normal and two friction impulses, the flag).
// ---------------------------------------------------------------------------
#define TBLOCK 256U
#define STAT_WORDS 48U // per block: 16 sums, 16 sums of squares, 3 + 3 for the targets

// Rows kept (flag >= min_flag), counted per block.
__global__ void __launch_bounds__(TBLOCK) k_keep_count(const float *f, unsigned n, float min_flag,
                                                       unsigned *block_counts)
{
    __shared__ unsigned c;
    if (threadIdx.x == 0U) {
        c = 0U;
    }
    __syncthreads();
    const unsigned i = blockIdx.x * TBLOCK + threadIdx.x;
    if (i < n && f[(unsigned long)i * IN + 15U] >= min_flag) {
        atomicAdd(&c, 1U);
    }
    __syncthreads();
    if (threadIdx.x == 0U) {
        block_counts[blockIdx.x] = c;
    }
}

// The kept rows, in order, after `base` rows of the set.
__global__ void __launch_bounds__(TBLOCK) k_keep_scatter(const float *f, const float *t, unsigned n,
                                                         float min_flag, const unsigned *block_first,
                                                         float *of, float *ot)
{
    __shared__ unsigned wave_count[TBLOCK / 64U];
    const unsigned i = blockIdx.x * TBLOCK + threadIdx.x;
    const unsigned lane = threadIdx.x & 63U;
    const unsigned w = threadIdx.x >> 6U;
    const int k = i < n && f[(unsigned long)i * IN + 15U] >= min_flag;
    const unsigned long long m = __ballot(k);
    if (lane == 0U) {
        wave_count[w] = (unsigned)__popcll(m);
    }
    __syncthreads();
    if (!k) {
        return;
    }
    unsigned pos = block_first[blockIdx.x] + (unsigned)__popcll(m & ((1ULL << lane) - 1ULL));
    for (unsigned j = 0U; j < w; ++j) {
        pos += wave_count[j];
    }
    for (unsigned j = 0U; j < IN; ++j) {
        of[(unsigned long)pos * IN + j] = f[(unsigned long)i * IN + j];
    }
    for (unsigned j = 0U; j < 4U; ++j) {
        ot[(unsigned long)pos * 4U + j] = t[(unsigned long)i * 4U + j];
    }
}

// The target a row trains for: impulses divided by the normal mass (slot 10).
__device__ static inline float target_of(const float *f, const float *t, unsigned long i,
                                         unsigned j)
{
    const float nm = f[i * IN + 10U];
    return nm > 0.0F ? t[i * 4U + j] / nm : 0.0F;
}

// Per block: sums and sums of squares of the features and targets.
__global__ void __launch_bounds__(TBLOCK) k_set_stats(const float *f, const float *t, unsigned n,
                                                      double *partials)
{
    __shared__ double s[STAT_WORDS][TBLOCK / 64U];
    const unsigned lane = threadIdx.x & 63U;
    const unsigned w = threadIdx.x >> 6U;
    const unsigned long i = (unsigned long)blockIdx.x * TBLOCK + threadIdx.x;
    for (unsigned j = 0U; j < STAT_WORDS; ++j) {
        double v = 0.0;
        if (i < n) {
            if (j < 16U) {
                v = f[i * IN + j];
            } else if (j < 32U) {
                v = (double)f[i * IN + j - 16U] * (double)f[i * IN + j - 16U];
            } else if (j < 35U) {
                v = target_of(f, t, i, j - 32U);
            } else if (j < 38U) {
                const double y = target_of(f, t, i, j - 35U);
                v = y * y;
            }
        }
        for (unsigned off = 32U; off > 0U; off >>= 1U) {
            v += __shfl_down(v, off, 64);
        }
        if (lane == 0U) {
            s[j][w] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < STAT_WORDS) {
        double v = 0.0;
        for (unsigned k = 0U; k < TBLOCK / 64U; ++k) {
            v += s[threadIdx.x][k];
        }
        partials[(unsigned long)blockIdx.x * STAT_WORDS + threadIdx.x] = v;
    }
}

struct Norm {
    float shift[16];
    float scale[16];
    float out_inv[4];
};

// Network inputs (BF16 of the normalised features, 14 used) and scaled
// targets.
__global__ void __launch_bounds__(TBLOCK) k_set_convert(const float *f, const float *t, unsigned n,
                                                        Norm nm, unsigned short *x, float *y)
{
    const unsigned long i = (unsigned long)blockIdx.x * TBLOCK + threadIdx.x;
    if (i >= n) {
        return;
    }
    for (unsigned j = 0U; j < IN; ++j) {
        x[i * IN + j] = j < 14U ? bf16_of((f[i * IN + j] - nm.shift[j]) * nm.scale[j])
                                : (unsigned short)0U;
    }
    for (unsigned j = 0U; j < 4U; ++j) {
        y[i * 4U + j] = j < NOUT ? target_of(f, t, i, j) * nm.out_inv[j] : 0.0F;
    }
}

// A minibatch: rows (a * (first + r) + b) mod n, an affine permutation of
// the set (a prime to n).
__global__ void __launch_bounds__(TBLOCK) k_gather(const unsigned short *x, const float *y,
                                                   unsigned n, unsigned long a, unsigned long b,
                                                   unsigned first, unsigned count,
                                                   unsigned short *bx, float *by)
{
    const unsigned r = blockIdx.x * TBLOCK + threadIdx.x;
    if (r >= count) {
        return;
    }
    const unsigned long src = (a * (unsigned long)(first + r) + b) % n;
    const uint4 *xs = (const uint4 *)(x + src * IN);
    uint4 *xd = (uint4 *)(bx + (unsigned long)r * IN);
    xd[0] = xs[0];
    xd[1] = xs[1];
    ((float4 *)by)[r] = ((const float4 *)y)[src];
}

static unsigned blocks_for_rows(unsigned n)
{
    const unsigned tiles = (n + 15U) / 16U;
    unsigned b = (tiles + 4U * WAVES - 1U) / (4U * WAVES); // about four tiles a wave
    return b < 1U ? 1U : (b > 1024U ? 1024U : b);
}

extern "C" {

NphResult nph_amd_mlp_forward(const void *x_bf16, u32 rows, const f32 *params, f32 *y,
                              void *stream)
{
    if (rows == 0U) {
        return NPH_RESULT_OK;
    }
    if (x_bf16 == NULL || params == NULL || y == NULL || ((uintptr_t)x_bf16 & 7U) != 0U ||
        ((uintptr_t)y & 15U) != 0U) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks_for_rows(rows)), dim3(BLOCK), 0,
                       (hipStream_t)stream, (const unsigned short *)x_bf16, rows, params, y);
    MLP_TRY(hipGetLastError());
    return NPH_RESULT_OK;
}

NphResult nph_amd_mlp_gelu(const f32 *x, u32 n, f32 *out, void *stream)
{
    if (n == 0U) {
        return NPH_RESULT_OK;
    }
    if (x == NULL || out == NULL) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_mlp_gelu, dim3((n + 255U) / 256U), dim3(256), 0, (hipStream_t)stream, x,
                       n, out);
    MLP_TRY(hipGetLastError());
    return NPH_RESULT_OK;
}

NphResult nph_amd_mlp_grad(const void *x_bf16, const f32 *target, u32 rows, const f32 *params,
                           f32 scale, f32 *grad, f32 *loss, void *stream)
{
    if (rows == 0U) {
        return NPH_RESULT_OK;
    }
    if (x_bf16 == NULL || target == NULL || params == NULL || grad == NULL || loss == NULL) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks_for_rows(rows)), dim3(BLOCK), 0,
                       (hipStream_t)stream, (const unsigned short *)x_bf16, target, rows, params,
                       scale, grad, loss);
    MLP_TRY(hipGetLastError());
    return NPH_RESULT_OK;
}

NphResult nph_amd_mlp_adam(f32 *params, const f32 *grad, f32 *m, f32 *v, u32 step, f32 lr,
                           void *stream)
{
    if (params == NULL || grad == NULL || m == NULL || v == NULL || step == 0U) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_mlp_adam, dim3((P_COUNT + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       params, grad, m, v, (unsigned)P_COUNT, lr, 0.9F, 0.999F, 1e-8F, step);
    MLP_TRY(hipGetLastError());
    return NPH_RESULT_OK;
}

void *nph_amd_device_alloc(u64 bytes)
{
    void *p = NULL;
    if (bytes == 0U || hipMalloc(&p, bytes) != hipSuccess) {
        return NULL;
    }
    return p;
}

void nph_amd_device_free(void *p)
{
    if (p != NULL) {
        (void)hipFree(p);
    }
}

NphResult nph_amd_device_copy(void *dst, const void *src, u64 bytes)
{
    if (bytes == 0U) {
        return NPH_RESULT_OK;
    }
    if (dst == NULL || src == NULL) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    MLP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    return NPH_RESULT_OK;
}

// ---------------------------------------------------------------------------
// The trainer
// ---------------------------------------------------------------------------
struct NphAmdTrainer {
    hipStream_t stream;
    u32 cap, rows;
    float *raw_x, *raw_t;   // cap rows as logged
    unsigned short *x;      // cap rows of network inputs
    float *y;               // cap rows of scaled targets
    unsigned short *bx;     // a minibatch
    float *by;
    u32 batch_cap;
    float *params, *grad, *m, *v, *loss;
    unsigned *blocks;       // per-block counts and firsts
    u32 cap_blocks;
    u32 *h_blocks;
    u32 step;
    u64 rng;
    int prepared;
    Norm norm;
};

static u64 splitmix(u64 *s)
{
    u64 z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30U)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27U)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31U);
}

static double unit(u64 *s) { return ((double)(splitmix(s) >> 11U) + 0.5) * (1.0 / 9007199254740992.0); }

// Normal deviates by Box and Muller.
static double gauss(u64 *s)
{
    const double u = unit(s);
    const double v = unit(s);
    return sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}

void nph_amd_trainer_destroy(NphAmdTrainer *t)
{
    if (t == NULL) {
        return;
    }
    (void)hipStreamSynchronize(t->stream);
    void *bufs[] = {t->raw_x, t->raw_t, t->x, t->y, t->bx, t->by, t->params, t->grad,
                    t->m, t->v, t->loss, t->blocks};
    for (u32 i = 0U; i < sizeof bufs / sizeof bufs[0]; ++i) {
        if (bufs[i] != NULL) {
            (void)hipFree(bufs[i]);
        }
    }
    free(t->h_blocks);
    (void)hipStreamDestroy(t->stream);
    free(t);
}

NphResult nph_amd_trainer_create(u32 capacity, u64 seed, NphAmdTrainer **out)
{
    if (out == NULL || capacity == 0U) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    *out = NULL;
    NphAmdTrainer *t = (NphAmdTrainer *)calloc(1, sizeof *t);
    if (t == NULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess) {
        free(t);
        return NPH_RESULT_DEVICE_LOST;
    }
    t->cap = capacity;
    t->rng = seed;
    const u64 c = capacity;
    if (hipMalloc((void **)&t->raw_x, c * IN * 4U) != hipSuccess ||
        hipMalloc((void **)&t->raw_t, c * 16U) != hipSuccess ||
        hipMalloc((void **)&t->params, P_COUNT * 4U) != hipSuccess ||
        hipMalloc((void **)&t->grad, P_COUNT * 4U) != hipSuccess ||
        hipMalloc((void **)&t->m, P_COUNT * 4U) != hipSuccess ||
        hipMalloc((void **)&t->v, P_COUNT * 4U) != hipSuccess ||
        hipMalloc((void **)&t->loss, 16U) != hipSuccess) {
        nph_amd_trainer_destroy(t);
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    // He-style initial weights; zero biases, zero rows 14, 15 of W1 and
    // columns 3..15 of W3.
    float *p = (float *)calloc(P_COUNT, sizeof(float));
    if (p == NULL) {
        nph_amd_trainer_destroy(t);
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    u64 s = seed ^ 0x5DEECE66DULL;
    for (u32 i = 0U; i < 14U; ++i) {
        for (u32 j = 0U; j < HID; ++j) {
            p[P_W1 + i * HID + j] = (float)(gauss(&s) * sqrt(2.0 / 14.0));
        }
    }
    for (u32 i = 0U; i < HID * HID; ++i) {
        p[P_W2 + i] = (float)(gauss(&s) * sqrt(2.0 / HID));
    }
    for (u32 i = 0U; i < HID; ++i) {
        for (u32 j = 0U; j < NOUT; ++j) {
            p[P_W3 + i * OUT + j] = (float)(gauss(&s) * sqrt(1.0 / HID));
        }
    }
    const hipError_t e = hipMemcpy(t->params, p, P_COUNT * 4U, hipMemcpyHostToDevice);
    free(p);
    if (e != hipSuccess || hipMemset(t->m, 0, P_COUNT * 4U) != hipSuccess ||
        hipMemset(t->v, 0, P_COUNT * 4U) != hipSuccess) {
        nph_amd_trainer_destroy(t);
        return NPH_RESULT_DEVICE_LOST;
    }
    *out = t;
    return NPH_RESULT_OK;
}

static NphResult reserve_blocks(NphAmdTrainer *t, u32 blocks)
{
    if (blocks <= t->cap_blocks) {
        return NPH_RESULT_OK;
    }
    if (t->blocks != NULL) {
        MLP_TRY(hipFree(t->blocks));
        t->blocks = NULL;
    }
    free(t->h_blocks);
    t->h_blocks = (u32 *)malloc((size_t)blocks * 4U);
    if (t->h_blocks == NULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    MLP_TRY(hipMalloc((void **)&t->blocks, (u64)blocks * STAT_WORDS * 8U));
    t->cap_blocks = blocks;
    return NPH_RESULT_OK;
}

NphResult nph_amd_trainer_add(NphAmdTrainer *t, const f32 *features, const f32 *targets, u32 rows,
                              f32 min_flag)
{
    if (t == NULL || (rows > 0U && (features == NULL || targets == NULL))) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    if (rows == 0U) {
        return NPH_RESULT_OK;
    }
    const u32 blocks = (rows + TBLOCK - 1U) / TBLOCK;
    NphResult r = reserve_blocks(t, blocks);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    hipLaunchKernelGGL(k_keep_count, dim3(blocks), dim3(TBLOCK), 0, t->stream, features, rows,
                       min_flag, t->blocks);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipMemcpyAsync(t->h_blocks, t->blocks, (u64)blocks * 4U, hipMemcpyDeviceToHost,
                           t->stream));
    MLP_TRY(hipStreamSynchronize(t->stream));
    // Each block's first row in the set; rows past the capacity are dropped.
    u32 at = t->rows;
    u32 kept = 0U;
    for (u32 b = 0U; b < blocks; ++b) {
        const u32 c = t->h_blocks[b];
        t->h_blocks[b] = at;
        at += c;
        kept += c;
    }
    if ((u64)t->rows + kept > t->cap) {
        return NPH_RESULT_FULL;
    }
    MLP_TRY(hipMemcpyAsync(t->blocks, t->h_blocks, (u64)blocks * 4U, hipMemcpyHostToDevice,
                           t->stream));
    hipLaunchKernelGGL(k_keep_scatter, dim3(blocks), dim3(TBLOCK), 0, t->stream, features, targets,
                       rows, min_flag, t->blocks, t->raw_x, t->raw_t);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipStreamSynchronize(t->stream));
    t->rows += kept;
    t->prepared = 0;
    return NPH_RESULT_OK;
}

u32 nph_amd_trainer_rows(const NphAmdTrainer *t) { return t == NULL ? 0U : t->rows; }

NphResult nph_amd_trainer_prepare(NphAmdTrainer *t)
{
    if (t == NULL || t->rows == 0U) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    const u32 n = t->rows;
    const u32 blocks = (n + TBLOCK - 1U) / TBLOCK;
    NphResult r = reserve_blocks(t, blocks);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    double *partials = (double *)t->blocks;
    hipLaunchKernelGGL(k_set_stats, dim3(blocks), dim3(TBLOCK), 0, t->stream, t->raw_x, t->raw_t, n,
                       partials);
    MLP_TRY(hipGetLastError());
    double *h = (double *)malloc((size_t)blocks * STAT_WORDS * 8U);
    if (h == NULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    MLP_TRY(hipMemcpyAsync(h, partials, (u64)blocks * STAT_WORDS * 8U, hipMemcpyDeviceToHost,
                           t->stream));
    MLP_TRY(hipStreamSynchronize(t->stream));
    double sum[STAT_WORDS];
    memset(sum, 0, sizeof sum);
    for (u32 b = 0U; b < blocks; ++b) {
        for (u32 j = 0U; j < STAT_WORDS; ++j) {
            sum[j] += h[(u64)b * STAT_WORDS + j];
        }
    }
    free(h);
    for (u32 j = 0U; j < 16U; ++j) {
        const double mean = sum[j] / n;
        const double var = sum[16U + j] / n - mean * mean;
        const double sd = var > 0.0 ? sqrt(var) : 0.0;
        t->norm.shift[j] = j < 14U ? (float)mean : 0.0F;
        // A constant feature (say, one friction for every body) enters as zero.
        t->norm.scale[j] = (j < 14U && sd > 1e-6 * (fabs(mean) + 1e-30)) ? (float)(1.0 / sd)
                                                                          : (j < 14U ? 1.0F : 0.0F);
    }
    // Targets scaled by their root mean square (zero stays zero).
    for (u32 j = 0U; j < 4U; ++j) {
        const double ms = j < NOUT ? sum[35U + j] / n : 0.0;
        t->norm.out_inv[j] = ms > 0.0 ? (float)(1.0 / sqrt(ms)) : 1.0F;
    }
    if (t->x == NULL) {
        MLP_TRY(hipMalloc((void **)&t->x, (u64)t->cap * IN * 2U));
        MLP_TRY(hipMalloc((void **)&t->y, (u64)t->cap * 16U));
    }
    hipLaunchKernelGGL(k_set_convert, dim3(blocks), dim3(TBLOCK), 0, t->stream, t->raw_x, t->raw_t,
                       n, t->norm, t->x, t->y);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipStreamSynchronize(t->stream));
    t->prepared = 1;
    return NPH_RESULT_OK;
}

static u64 gcd64(u64 a, u64 b)
{
    while (b != 0U) {
        const u64 r = a % b;
        a = b;
        b = r;
    }
    return a;
}

NphResult nph_amd_trainer_epoch(NphAmdTrainer *t, u32 batch, f32 lr, f32 *mean_loss)
{
    if (t == NULL || !t->prepared || batch == 0U || mean_loss == NULL) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    const u32 n = t->rows;
    const u32 b = batch < n ? batch : n;
    if (b > t->batch_cap) {
        if (t->bx != NULL) {
            MLP_TRY(hipFree(t->bx));
            MLP_TRY(hipFree(t->by));
            t->bx = NULL;
            t->by = NULL;
        }
        MLP_TRY(hipMalloc((void **)&t->bx, (u64)b * IN * 2U));
        MLP_TRY(hipMalloc((void **)&t->by, (u64)b * 16U));
        t->batch_cap = b;
    }
    // This epoch's order: a prime to n.
    u64 a = 1U;
    if (n > 1U) {
        do {
            a = 1U + splitmix(&t->rng) % (n - 1U);
        } while (gcd64(a, n) != 1U);
    }
    const u64 off = splitmix(&t->rng) % n;
    MLP_TRY(hipMemsetAsync(t->loss, 0, 4U, t->stream));
    const u32 batches = n / b;
    for (u32 k = 0U; k < batches; ++k) {
        hipLaunchKernelGGL(k_gather, dim3((b + TBLOCK - 1U) / TBLOCK), dim3(TBLOCK), 0, t->stream,
                           t->x, t->y, n, (unsigned long)a, (unsigned long)off, k * b, b, t->bx,
                           t->by);
        MLP_TRY(hipMemsetAsync(t->grad, 0, P_COUNT * 4U, t->stream));
        NphResult r = nph_amd_mlp_grad(t->bx, t->by, b, t->params, 1.0F / (3.0F * (float)b),
                                       t->grad, t->loss, (void *)t->stream);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        t->step += 1U;
        r = nph_amd_mlp_adam(t->params, t->grad, t->m, t->v, t->step, lr, (void *)t->stream);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    float total = 0.0F;
    MLP_TRY(hipMemcpyAsync(&total, t->loss, 4U, hipMemcpyDeviceToHost, t->stream));
    MLP_TRY(hipStreamSynchronize(t->stream));
    *mean_loss = total / (float)batches;
    return NPH_RESULT_OK;
}

NphResult nph_amd_trainer_loss(NphAmdTrainer *t, f32 *loss)
{
    if (t == NULL || !t->prepared || loss == NULL) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    MLP_TRY(hipMemsetAsync(t->loss, 0, 4U, t->stream));
    MLP_TRY(hipMemsetAsync(t->grad, 0, P_COUNT * 4U, t->stream));
    NphResult r = nph_amd_mlp_grad(t->x, t->y, t->rows, t->params,
                                   1.0F / (3.0F * (float)t->rows), t->grad, t->loss,
                                   (void *)t->stream);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    float total = 0.0F;
    MLP_TRY(hipMemcpyAsync(&total, t->loss, 4U, hipMemcpyDeviceToHost, t->stream));
    MLP_TRY(hipStreamSynchronize(t->stream));
    *loss = total;
    return NPH_RESULT_OK;
}

NphResult nph_amd_trainer_warm_start(const NphAmdTrainer *t, f32 *params, NphAmdWarmStart *out)
{
    if (t == NULL || !t->prepared || params == NULL || out == NULL) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    MLP_TRY(hipMemcpyAsync(params, t->params, P_COUNT * 4U, hipMemcpyDeviceToHost, t->stream));
    MLP_TRY(hipStreamSynchronize(t->stream));
    memset(out, 0, sizeof *out);
    out->params = params;
    for (u32 j = 0U; j < 16U; ++j) {
        out->feature_shift[j] = t->norm.shift[j];
        out->feature_scale[j] = t->norm.scale[j];
    }
    for (u32 j = 0U; j < 4U; ++j) {
        out->output_scale[j] = 1.0F / t->norm.out_inv[j];
    }
    out->mode = NPH_AMD_WARM_NEW_POINTS;
    out->gate = true;
    return NPH_RESULT_OK;
}

} // extern "C"
