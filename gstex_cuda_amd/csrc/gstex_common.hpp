// Shared device-side definitions for the gfx950 textured-2DGS rasteriser.
//
// Conventions fixed by the reference interface (kept identical on purpose):
//   quaternions (w, x, y, z) -> axes ax1, ax2, ax3 = columns of R  (helpers.cuh:166-185)
//   viewmat / c2w row-major 4x4                                     (helpers.cuh:113-131)
//   texture_dims rows {h, w, start}; texel (i, j) at row start + i*w + j of (X, C)
//                                                                   (texture_helpers.cuh:162-215)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gstex {

constexpr int kWave = 64;

// ---------------------------------------------------------------- per-Gaussian splat record
// One 128-byte record per Gaussian per view, written once by the record kernel
// and read (wave-uniformly) by the raster kernels.  All 32 slots are 32-bit.
enum RecordSlot : int {
    R_AX1 = 0,     // 0..2  first in-plane axis
    R_IQ1 = 3,     // is1 * is1, is1 = 1 / (scale.x * glob_scale)
    R_AX2 = 4,     // 4..6  second in-plane axis
    R_IQ2 = 7,     // is2 * is2, is2 = 1 / (scale.y * glob_scale)
    R_AX3 = 8,     // 8..10 splat normal
    R_OPAC = 11,
    R_MO = 12,     // 12..14 mean - camera origin (per view)
    R_PIX_X = 15,  // projected mean (blur term)
    R_RGB = 16,    // 16..18
    R_PIX_Y = 19,
    R_UV0 = 20,    // 20..21
    R_UMAP = 22,   // 22..24
    R_VMAP = 25,   // 25..27
    R_TEX_H = 28,  // int bits
    R_TEX_W = 29,  // int bits
    R_TEX_START = 30,
    R_NUM = 31,    // ax3 . (mean - camera origin): numerator of the ray/plane t
    kRecordFloats = 32
};

// Per-Gaussian scalars the backward raster kernel sums over pixels (one
// wave-reduced atomic per slot); the finalize kernel maps them to the
// reference's gradient tensors.  With vl1, vl2 the adjoints of the in-plane
// coordinates l1, l2, vu, vv those of the texture coordinates, vnum / vden
// those of the ray-plane numerator / denominator:
enum GradSlot : int {
    G_RGB = 0,    // 0..2   sum vis * dL/dimg
    G_OPAC = 3,   //        sum G * dL/dalpha
    G_P11 = 4,    //        sum vl1 * l1
    G_P12 = 5,    //        sum vl1 * l2
    G_P21 = 6,    //        sum vl2 * l1
    G_P22 = 7,    //        sum vl2 * l2
    G_U1 = 8,     //        sum vu * l1
    G_U2 = 9,     //        sum vu * l2
    G_V1 = 10,    //        sum vv * l1
    G_V2 = 11,    //        sum vv * l2
    G_VL1 = 12,   //        sum vl1
    G_VL2 = 13,   //        sum vl2
    G_VU = 14,    //        sum vu
    G_VV = 15,    //        sum vv
    G_VNUM = 16,  //        sum vnum
    G_AX3X = 17,  // 17..19 sum vden * ray + vis * dL/dnormal
    G_EXTRA = 20, // 20..31 texel gradients carried through the same reduction
    G_PIX = 30,   // 30..31 d/d projected mean (blur term)
    kGradFloats = 32
};

// Per-wave LDS accumulator for one Gaussian's texel gradients (h * w * C floats);
// covers a 16 x 16 chart with up to 4 channels.
constexpr int kTexAccFloats = 1024;

// settings bits, texture.cu:56-65 / 381-387
constexpr int kSetNearest = 1 << 2;
constexpr int kSetUvGrad = 1 << 8;
constexpr int kSetBlur = 1 << 9;
constexpr int kSetNdc = 1 << 10;
constexpr int kSetNormalFlip = 1 << 15;
constexpr int kSetAlphaViz = 1 << 16;      // alpha visualisation (texture.cu:56-65)
constexpr int kSetWhiteOutline = 1 << 24;  //   outline drawn white instead of hidden
constexpr int kSetOpacityThresh = 1 << 25; //   opacity < 0.5 hidden
// alpha bound = bits 17..21 / 8 (sigma threshold 0.5 bound^2); outline bound = bits 26..29 / 4
__host__ __device__ inline float viz_alpha_bound(int settings) {
    return (float)((settings >> 17) & 31) / 8.0f;
}
__host__ __device__ inline float viz_outline_bound(int settings) {
    return (float)((settings >> 26) & 15) / 4.0f;
}

// ---------------------------------------------------------------- small vector helpers
struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }

// R * p + t for a row-major 4x4 (helpers.cuh:124-131).
__device__ __forceinline__ V3 xform43(const float* m, V3 p) {
    return V3{m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
              m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}

// R^T * p (helpers.cuh:114-121).
__device__ __forceinline__ V3 xform43_rot_t(const float* m, V3 p) {
    return V3{m[0] * p.x + m[4] * p.y + m[8] * p.z,
              m[1] * p.x + m[5] * p.y + m[9] * p.z,
              m[2] * p.x + m[6] * p.y + m[10] * p.z};
}

// Pinhole projection with the reference's +1e-6 guard (helpers.cuh:145-152).
__device__ __forceinline__ void project_pix(float fx, float fy, float cx, float cy, V3 pv,
                                            float& u, float& v) {
    float rw = 1.f / (pv.z + 1e-6f);
    u = pv.x * rw * fx + cx;
    v = pv.y * rw * fy + cy;
}

// Quaternion (w, x, y, z), not normalised, to the three rotation axes.
__device__ __forceinline__ void quat_axes(float w, float x, float y, float z, V3& a1, V3& a2,
                                          V3& a3) {
    a1 = V3{1.f - 2.f * (y * y + z * z), 2.f * (x * y + w * z), 2.f * (x * z - w * y)};
    a2 = V3{2.f * (x * y - w * z), 1.f - 2.f * (x * x + z * z), 2.f * (y * z + w * x)};
    a3 = V3{2.f * (x * z + w * y), 2.f * (y * z - w * x), 1.f - 2.f * (x * x + y * y)};
}

// The same three functions with every product and sum rounded on its own (no
// fused multiply-add).  They feed the integer results of the projection, the
// tile rectangles and the depth sort keys, which then follow the written
// operation order bit for bit whatever the compiler would contract.
__device__ __forceinline__ V3 xform43_strict(const float* m, V3 p) {
#pragma clang fp contract(off)
    return V3{m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
              m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}

__device__ __forceinline__ void project_pix_strict(float fx, float fy, float cx, float cy, V3 pv,
                                                   float& u, float& v) {
#pragma clang fp contract(off)
    float rw = 1.f / (pv.z + 1e-6f);
    u = pv.x * rw * fx + cx;
    v = pv.y * rw * fy + cy;
}

__device__ __forceinline__ void quat_axes_strict(float w, float x, float y, float z, V3& a1,
                                                 V3& a2, V3& a3) {
#pragma clang fp contract(off)
    a1 = V3{1.f - 2.f * (y * y + z * z), 2.f * (x * y + w * z), 2.f * (x * z - w * y)};
    a2 = V3{2.f * (x * y - w * z), 1.f - 2.f * (x * x + z * z), 2.f * (y * z + w * x)};
    a3 = V3{2.f * (x * z + w * y), 2.f * (y * z - w * x), 1.f - 2.f * (x * x + y * y)};
}

// Adjoint of quat_axes: g1, g2, g3 are the gradients of the three axes.
__device__ __forceinline__ void quat_axes_vjp(float w, float x, float y, float z, V3 g1, V3 g2,
                                              V3 g3, float out[4]) {
    out[0] = 2.f * (x * (g2.z - g3.y) + y * (g3.x - g1.z) + z * (g1.y - g2.x));
    out[1] = 2.f * (-2.f * x * (g2.y + g3.z) + y * (g1.y + g2.x) + z * (g1.z + g3.x) +
                    w * (g2.z - g3.y));
    out[2] = 2.f * (x * (g1.y + g2.x) - 2.f * y * (g1.x + g3.z) + z * (g2.z + g3.y) +
                    w * (g3.x - g1.z));
    out[3] = 2.f * (x * (g1.z + g3.x) + y * (g2.z + g3.y) - 2.f * z * (g1.x + g2.y) +
                    w * (g1.y - g2.x));
}

// Two int16 values (clamped) in one 32-bit word: lo | hi << 16.
__host__ __device__ inline unsigned pack_i16(int lo, int hi) {
    lo = lo < -32768 ? -32768 : (lo > 32767 ? 32767 : lo);
    hi = hi < -32768 ? -32768 : (hi > 32767 ? 32767 : hi);
    return (unsigned)(lo & 0xffff) | ((unsigned)(hi & 0xffff) << 16);
}

// Float -> int32 truncation (v_cvt_i32_f32: toward zero, saturating, NaN -> 0).
__device__ __forceinline__ int f2i_trunc(float x) { return (int)x; }

}  // namespace gstex
