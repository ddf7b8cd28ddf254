// Device-side pieces shared by the raster forward and backward kernels:
// launch arguments, the per-pixel camera ray, the splat record view, the
// ray/splat intersection with its alpha, and the bilinear chart taps.
#pragma once

#include "gstex_common.hpp"
#include "gstex_kernels.hpp"

namespace gstex {

struct RasterArgs {
    int img_w, img_h, bw, tiles_x, tiles_y, channels, settings;
    int num_points, num_ids, num_texels;
    float fx, fy, cx, cy;
    const float* viewmat;
    const float* c2w;
    const float* background;
    const int* ids_sorted;
    const int* tile_bins;
    const float* records;
    const void* texture;   // float32 or bfloat16 texels (kernel template parameter)
    uint64_t* wave_masks;  // kMaxTileWaves planes of mask_plane words, or nullptr
    int mask_plane;
    const uint2* boxes;    // per-Gaussian screen boxes (culling), or nullptr
    const float4* ellipses;  // screen ellipses (x, y, A, B) for culling with blur off, or nullptr
    const float* ellipse_c;  // their C values
    bool texture_bf16;
    bool precise_edge;     // float64 plane hits for near-edge-on splats
};

// |n . ray| below which a splat counts as near-edge-on for precise_edge.
constexpr float kEdgeOnDen = 0.05f;

// Texel value idx of a float32 (TB = false) or bfloat16 (TB = true) texture;
// bfloat16 widens exactly to float32.
template <bool TB>
__device__ __forceinline__ float texel(const void* tex, int idx) {
    if constexpr (TB) {
        const unsigned short h = static_cast<const unsigned short*>(tex)[idx];
        return __uint_as_float((unsigned)h << 16);
    } else {
        return static_cast<const float*>(tex)[idx];
    }
}

struct FwdOut {
    float* out_img;
    float* out_depth;
    float* out_reg;
    float* out_texture;
    float* out_normal;
    float* final_Ts;
    int* final_idx;
    int* depth_idx;
    float* out_reg_s;
};

struct BwdIn {
    const float* final_Ts;
    const int* final_idx;
    const int* depth_idx;
    const float* final_s;
    const float* v_img;
    const float* v_depth;
    const float* v_reg;
    const float* v_alpha;
    const float* v_texture_out;
    const float* v_normal;
    float* grad_rec;   // (N, 32) accumulators
    float* v_texture;  // (X, C)
    // optional, three channels only: v_img and v_texture_out then point at
    // the forward's out_img and out_texture, and the image cotangents are the
    // L2 loss's, ((out_img + out_texture) - loss_target) * loss_scale
    const float* loss_target;
    float loss_scale;
};

// One pixel channel of that cotangent, each step rounded as a separate
// float32 operation (as gstex_image_loss_grad rounds it).
__device__ __forceinline__ float l2_cotangent(float img, float tex, float target, float scale) {
#pragma clang fp contract(off)
    const float d = img + tex;
    return (d - target) * scale;
}

// One view of a multi-view raster pass: its launch arguments, forward
// outputs and backward inputs, in a device table the queue kernels read.
struct ViewRaster {
    RasterArgs a;
    FwdOut o;
    BwdIn b;
};

// Header of a multi-view table (device copy; the host keeps its own).
struct ViewsHeader {
    int num_views, num_items, waves_per_tile;
    int head;  // the work queue's next index, reset before each pass
};

// A work item packs (view, tile): tile in the low kViewItemTileBits bits.
constexpr int kViewItemTileBits = 20;

__host__ __device__ inline int raster_block_threads(int bw) {
    return bw == 16 ? 256 : ((bw * bw + kWave - 1) / kWave) * kWave;
}

// Pixel of this lane inside its tile.  bw = 16 uses 8x8 quadrants per wave.
__device__ __forceinline__ void lane_pixel(int bw, int tid, int& lx, int& ly, bool& in_tile) {
    if (bw == 16) {
        const int w = tid >> 6, l = tid & 63;
        lx = (w & 1) * 8 + (l & 7);
        ly = (w >> 1) * 8 + (l >> 3);
        in_tile = true;
    } else {
        lx = tid % bw;
        ly = tid / bw;
        in_tile = tid < bw * bw;
    }
}

struct PixelRay {
    V3 o, r;
    float px, py, view_depth;
};

// Camera ray through the pixel centre (texture_helpers.cuh:336-352).
__device__ __forceinline__ PixelRay pixel_ray(const float* c2w, const float* vm, float fx,
                                              float fy, float cx, float cy, int j, int i) {
    PixelRay pr;
    pr.px = (float)j + 0.5f;
    pr.py = (float)i + 0.5f;
    pr.o = v3(c2w[3], c2w[7], c2w[11]);
    const float u = (pr.px - cx) / fx, v = (pr.py - cy) / fy;
    V3 r = xform43(c2w, v3(u, v, 1.f));
    r = sub3(r, pr.o);
    const float nrm = sqrtf(r.x * r.x + r.y * r.y + r.z * r.z);
    pr.r = v3(r.x / nrm, r.y / nrm, r.z / nrm);
    pr.view_depth = vm[8] * pr.r.x + vm[9] * pr.r.y + vm[10] * pr.r.z;
    return pr;
}

struct Splat {
    V3 a1, a2, a3, mo, rgb, umap, vmap;  // mo = mean - camera origin
    float iq1, iq2, num, opac, pix_x, pix_y, u0, v0;  // iq = squared inverse scale
    int h, w, start;
};

// Records and ids are read-only while a raster kernel runs.  Reading them
// through the constant address space lets a wave-uniform address become a
// scalar (SMEM) load even in kernels that issue atomics, which otherwise
// forces 128-byte records into vector registers.
#define GSTEX_CONST __attribute__((address_space(4)))

template <typename T>
__device__ __forceinline__ const GSTEX_CONST T* as_const_space(const T* p) {
    return (const GSTEX_CONST T*)p;
}

// A struct read through the constant address space (scalar loads for a
// wave-uniform address), dword by dword.
template <typename T>
__device__ __forceinline__ T load_const(const GSTEX_CONST T* p) {
    static_assert(sizeof(T) % 4 == 0, "dword-sized struct");
    T out;
    const GSTEX_CONST uint32_t* src = reinterpret_cast<const GSTEX_CONST uint32_t*>(p);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) dst[k] = src[k];
    return out;
}

// Adjacent float reads of one record merge into wide loads (SMEM dwordx8/x16
// for a uniform constant-space address, dwordx4 otherwise).
template <typename P>
__device__ __forceinline__ Splat load_splat(P q) {
    Splat s;
    s.a1 = v3(q[R_AX1], q[R_AX1 + 1], q[R_AX1 + 2]);
    s.iq1 = q[R_IQ1];
    s.a2 = v3(q[R_AX2], q[R_AX2 + 1], q[R_AX2 + 2]);
    s.iq2 = q[R_IQ2];
    s.a3 = v3(q[R_AX3], q[R_AX3 + 1], q[R_AX3 + 2]);
    s.opac = q[R_OPAC];
    s.mo = v3(q[R_MO], q[R_MO + 1], q[R_MO + 2]);
    s.pix_x = q[R_PIX_X];
    s.rgb = v3(q[R_RGB], q[R_RGB + 1], q[R_RGB + 2]);
    s.pix_y = q[R_PIX_Y];
    s.u0 = q[R_UV0];
    s.v0 = q[R_UV0 + 1];
    s.umap = v3(q[R_UMAP], q[R_UMAP + 1], q[R_UMAP + 2]);
    s.vmap = v3(q[R_VMAP], q[R_VMAP + 1], q[R_VMAP + 2]);
    s.h = __float_as_int(q[R_TEX_H]);
    s.w = __float_as_int(q[R_TEX_W]);
    s.start = __float_as_int(q[R_TEX_START]);
    s.num = q[R_NUM];
    return s;
}

// a / b for normal operands of moderate range (no overflow or denormal
// quotients): hardware reciprocal (1 ulp), quotient, and one residual
// correction q + (a - q b) r, which gives the IEEE quotient in nearly every
// case at four instructions instead of the ten of the full divide (whose
// scaling steps only matter near the exponent limits).
__device__ __forceinline__ float div_fast(float a, float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    const float q = a * r;
    return fmaf(fmaf(-q, b, a), r, q);
}

// Lane mask of a per-lane condition over the active lanes.  HIP's __ballot,
// __any and __all take an int, so the compare's lane mask is widened into a
// register and compared again (two VALU instructions per call); the builtin
// on a bool takes the mask as it is.
__device__ __forceinline__ uint64_t lanes_where(bool c) {
    return __builtin_amdgcn_ballot_w64(c);
}
__device__ __forceinline__ bool any_lane(bool c) { return lanes_where(c) != 0; }

// log2(e) rounded to float32 (0x3fb8aa3b), the factor of __expf
constexpr float kLog2E = 1.44269502f;

struct Pair {
    V3 delta;
    float den, t, l1, l2, G, Gb, alpha;
    bool blur_sel, skip;
};

// Ray/splat intersection and alpha for one (pixel, Gaussian) pair.  With
// precise_edge, a near-edge-on splat (|n . ray| < 0.05) has its plane hit
// (t, delta, l1, l2) solved in float64: there t = n.(m - o) / n.r and
// delta = o + t r - m cancel badly in float32.
__device__ __forceinline__ Pair pair_eval(const Splat& s, const PixelRay& pr, bool use_blur,
                                          bool precise_edge = false) {
    Pair p;
    float den = dot3(s.a3, pr.r);
    const float eps = 1e-6f;
    // |den| below eps becomes eps with den's sign, +eps for both zeros
    // (den + 0 is +0 for -0): den in [0, eps) -> eps, (-eps, 0) -> -eps
    den = fabsf(den) < eps ? copysignf(eps, den + 0.f) : den;
    p.den = den;
    if (precise_edge && fabsf(den) < kEdgeOnDen) {
        const double ax = s.a3.x, ay = s.a3.y, az = s.a3.z;
        const double rx = pr.r.x, ry = pr.r.y, rz = pr.r.z;
        double dd = ax * rx + ay * ry + az * rz;
        if (0.0 <= dd && dd < 1e-6) dd = 1e-6;
        else if (-1e-6 < dd && dd <= 0.0) dd = -1e-6;
        const double mx = s.mo.x, my = s.mo.y, mz = s.mo.z;
        const double t = (ax * mx + ay * my + az * mz) / dd;
        const double dx = t * rx - mx, dy = t * ry - my, dz = t * rz - mz;
        p.t = (float)t;
        p.delta = v3((float)dx, (float)dy, (float)dz);
        p.l1 = (float)(dx * s.a1.x + dy * s.a1.y + dz * s.a1.z);
        p.l2 = (float)(dx * s.a2.x + dy * s.a2.y + dz * s.a2.z);
    } else {
        // The numerator n . (m - o) and m - o are per splat and view (built
        // with the record); |den| >= 1e-6 keeps the quotient far from
        // overflow.  The hit offset o + t r - m is formed as t r - (m - o).
        p.t = div_fast(s.num, den);
        p.delta = v3(fmaf(p.t, pr.r.x, -s.mo.x), fmaf(p.t, pr.r.y, -s.mo.y),
                     fmaf(p.t, pr.r.z, -s.mo.z));
        p.l1 = dot3(p.delta, s.a1);
        p.l2 = dot3(p.delta, s.a2);
    }
    // exp(-sigma), sigma = q / 2, as v_exp of q * (-log2(e) / 2): the same
    // float product as (q / 2) * -log2(e) (halving is exact), one multiply
    const float q = s.iq1 * p.l1 * p.l1 + s.iq2 * p.l2 * p.l2;
    p.G = __builtin_amdgcn_exp2f(q * (-0.5f * kLog2E));
    p.Gb = 0.f;
    p.blur_sel = false;
    if (use_blur) {
        const float sigma = 0.5f * q;
        const float dx = s.pix_x - pr.px, dy = s.pix_y - pr.py;
        const float sigma_blur = 0.5f * 2.0f * (dx * dx + dy * dy);
        p.Gb = __expf(-sigma_blur);
        p.blur_sel = sigma_blur < sigma;
    }
    p.alpha = fminf(0.99f, s.opac * (p.blur_sel ? p.Gb : p.G));
    p.skip = p.t < 0.01f || p.t > 1000.0f || p.alpha < 1.f / 255.f;
    return p;
}

// lanes_where(p.skip), as a ballot per compare (each folds into its compare's
// lane mask; the ballot of their union would be widened into a register)
__device__ __forceinline__ uint64_t skip_lanes(const Pair& p) {
    return lanes_where(p.t < 0.01f) | lanes_where(p.t > 1000.0f) |
           lanes_where(p.alpha < 1.f / 255.f);
}

// Texels of one chart (wave-uniform), read through a raw buffer resource over
// the chart's h * w * C values: taps are byte offsets inside the chart, and a
// read past the resource returns 0 (the hardware range check), so no per-tap
// clamping is needed for memory safety.  A chart that does not lie inside the
// texture (start < 0, h or w outside [1, 2^24), start + h * w > num_texels) or
// that spans 2^31 bytes or more gets an empty resource: its texels read as 0
// and it receives no texel gradient (the reference indexes out of bounds
// there).  Descriptor word 3 is the gfx9-family raw-buffer format (32-bit data
// format, so the range check applies).
constexpr int kBufferWord3 = 0x00020000;

struct Chart {
    __amdgpu_buffer_rsrc_t rsrc;
    int ce;   // bytes per texel (C values)
    int wce;  // bytes per chart row (w texels)
    bool ok;
};

template <bool TB>
__device__ __forceinline__ Chart chart_of(const void* texture, const Splat& s, int C,
                                          int num_texels) {
    constexpr int E = TB ? 2 : 4;
    Chart ch;
    const long long n = (long long)s.h * (long long)s.w;
    ch.ce = C * E;
    ch.ok = s.start >= 0 && s.h > 0 && s.w > 0 && s.h < (1 << 24) && s.w < (1 << 24) &&
            (long long)s.start + n <= (long long)num_texels && n * ch.ce < 0x80000000ll;
    const char* base = static_cast<const char*>(texture) + (ch.ok ? (size_t)s.start * ch.ce : 0);
    ch.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(base), (short)0,
                                                ch.ok ? (int)(n * ch.ce) : 0, kBufferWord3);
    ch.wce = s.w * ch.ce;
    return ch;
}

// Texel value at byte offset off of a chart (bfloat16 widened).
template <bool TB>
__device__ __forceinline__ float chart_texel(const Chart& ch, int off) {
    if constexpr (TB) {
        const unsigned short h = __builtin_amdgcn_raw_buffer_load_b16(ch.rsrc, off, 0, 0);
        return __uint_as_float((unsigned)h << 16);
    } else {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ch.rsrc, off, 0, 0));
    }
}

// Bilinear (or nearest) taps of one chart at clamped (u, v) in [0, 1]: weights
// and the byte offsets of texels (i, j) = (iu, iv), (iu, jv), (ju, iv),
// (ju, jv) inside the chart, with the replicate-padded neighbour indices
// (texture_helpers.cuh:162-215).  For a chart inside the texture tu = h u
// lies in [0, h], so only the upper clamps can act.
// Also returns the texel row / column indices of the taps: rows iu, ju and
// columns iv, jv (tap k is at row k < 2 ? iu : ju, column k & 1 ? jv : iv).
__device__ __forceinline__ void chart_taps_ij(const Splat& s, const Chart& ch, float u, float v,
                                              bool bilinear, float wt[4], int off[4], float& fu,
                                              float& fv, int& iu, int& iv, int& ju, int& jv) {
    const float tu = (float)s.h * u, tv = (float)s.w * v;
    const int iu0 = f2i_trunc(tu), iv0 = f2i_trunc(tv);
    // tu, tv >= 0: the part above the truncation, exact (tu - floor(tu))
    fu = __builtin_amdgcn_fractf(tu);
    fv = __builtin_amdgcn_fractf(tv);
    iu = min(iu0, s.h - 1);
    iv = min(iv0, s.w - 1);
    ju = min(iu0 + 1, s.h - 1);
    jv = min(iv0 + 1, s.w - 1);
    const float wii = (1.f - fu) * (1.f - fv), wij = (1.f - fu) * fv, wji = fu * (1.f - fv),
                wjj = fu * fv;
    if (bilinear) {
        wt[0] = wii; wt[1] = wij; wt[2] = wji; wt[3] = wjj;
    } else {
        wt[0] = wt[1] = wt[2] = wt[3] = 0.f;
        if (wii >= wij && wii >= wji && wii >= wjj) wt[0] = 1.f;
        else if (wij >= wii && wij >= wji && wij >= wjj) wt[1] = 1.f;
        else if (wji >= wii && wji >= wij && wji >= wjj) wt[2] = 1.f;
        else wt[3] = 1.f;
    }
    // iu * wce < h * w * C * E < 2^31; iv < 2^24 and ce <= 256 fit the 24-bit multiply
    off[0] = iu * ch.wce + (int)__umul24((unsigned)iv, (unsigned)ch.ce);
    const int dv = jv != iv ? ch.ce : 0;
    off[1] = off[0] + dv;
    off[2] = off[0] + (ju != iu ? ch.wce : 0);
    off[3] = off[2] + dv;
}

__device__ __forceinline__ void chart_taps(const Splat& s, const Chart& ch, float u, float v,
                                           bool bilinear, float wt[4], int off[4], float& fu,
                                           float& fv) {
    int iu, iv, ju, jv;
    chart_taps_ij(s, ch, u, v, bilinear, wt, off, fu, fv, iu, iv, ju, jv);
}

__device__ __forceinline__ void splat_uv(const Splat& s, V3 delta, float& u, float& v) {
    u = s.u0 + dot3(s.umap, delta);
    v = s.v0 + dot3(s.vmap, delta);
    u = fminf(fmaxf(u, 0.f), 1.f);
    v = fminf(fmaxf(v, 0.f), 1.f);
}

// ------------------------------------------------------------------ wave primitives
constexpr int kDppQuadXor1 = 0xB1;      // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;      // quad_perm [2,3,0,1]
constexpr int kDppRowMirror = 0x140;    // lane i <-> 15 - i within 16
constexpr int kDppRowHalfMirror = 0x141; // lane i <-> 7 - i within 8
constexpr int kDppRowRor8 = 0x128;      // lane i <- lane (i + 8) mod 16 within 16 (= i ^ 8)

template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

__device__ __forceinline__ float swizzle_xor16(float v) {
    // ds_swizzle bit mode: and 0x1f, or 0, xor 0x10 (within each 32-lane half)
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
}

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}

// Wave-wide min / max of an int: the four DPP rounds cover each row of 16
// lanes (partner masks 15, 7, 2, 1), the swizzle pairs the rows of each half,
// and lanes 0 and 32 hold the two half results.
__device__ __forceinline__ int wave_min(int v) {
    v = min(v, dpp_i<kDppRowMirror>(v));
    v = min(v, dpp_i<kDppRowHalfMirror>(v));
    v = min(v, dpp_i<kDppQuadXor2>(v));
    v = min(v, dpp_i<kDppQuadXor1>(v));
    v = min(v, __builtin_amdgcn_ds_swizzle(v, 0x401F));
    return min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 32));
}

__device__ __forceinline__ int wave_max(int v) {
    v = max(v, dpp_i<kDppRowMirror>(v));
    v = max(v, dpp_i<kDppRowHalfMirror>(v));
    v = max(v, dpp_i<kDppQuadXor2>(v));
    v = max(v, dpp_i<kDppQuadXor1>(v));
    v = max(v, __builtin_amdgcn_ds_swizzle(v, 0x401F));
    return max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 32));
}

// Pairwise merge of lanes that sample the same texel quad (equal key, the
// first tap row; -1 for lanes with nothing to add): over lane bits 0, 3 and 1
// (horizontal, vertical, horizontal neighbours in an 8x8 quadrant), on a match
// the lower lane takes the pair's sum and the upper lane drops to zero, so the
// atomics of a pixel block land on few distinct lanes.  Per round
// q <- keep * q + take * partner(q).  Call with all lanes active.
struct QuadMerge {
    float keep1, take1, keep8, take8, keep2, take2;
};

__device__ __forceinline__ QuadMerge quad_merge_setup(int key, int lane) {
    const bool m1 = dpp_i<kDppQuadXor1>(key) == key;
    const bool m8 = dpp_i<kDppRowRor8>(key) == key;
    const bool m2 = dpp_i<kDppQuadXor2>(key) == key;
    QuadMerge m;
    m.keep1 = m1 && (lane & 1) ? 0.f : 1.f;
    m.take1 = m1 && !(lane & 1) ? 1.f : 0.f;
    m.keep8 = m8 && (lane & 8) ? 0.f : 1.f;
    m.take8 = m8 && !(lane & 8) ? 1.f : 0.f;
    m.keep2 = m2 && (lane & 2) ? 0.f : 1.f;
    m.take2 = m2 && !(lane & 2) ? 1.f : 0.f;
    return m;
}

__device__ __forceinline__ float quad_merge(float q, const QuadMerge& m) {
    q = fmaf(dpp<kDppQuadXor1>(q), m.take1, q * m.keep1);
    q = fmaf(dpp<kDppRowRor8>(q), m.take8, q * m.keep8);
    q = fmaf(dpp<kDppQuadXor2>(q), m.take2, q * m.keep2);
    return q;
}

// ------------------------------------------------------------------ wave reduction
// One halving round over lane bit 5 (bit 4) by a permlane swap: lanes with
// the bit clear end with a summed with the partner lane's a, lanes with it set
// with b summed with the partner's b.  The swap moves the halves each lane
// gives up in one instruction, so a pair costs a swap and an add.
__device__ __forceinline__ float swap32_sum(float a, float b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b),
                                                    false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float swap16_sum(float a, float b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b),
                                                    false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// Reduce-scatter of N <= 64 per-lane values over the wave in six halving
// rounds: permlane swaps over lane bits 5 and 4, then DPP lane permutes with
// partner masks 15, 7, 2 and 1 selecting on bits 3, 2, 1 and 0.  Each partner
// mask leaves the bits already selected on unchanged, and together they span
// all six lane bits, so every lane ends with the wave total of one value.  A
// round over n values pairs value k with value k + n - n/2 (the lane with the
// bit clear keeps the first, the lane with it set the second); with n odd the
// middle value is summed by both lanes and kept by both.  Which value a lane
// ends with, and whether it is the lowest lane holding it, is rs_slot<N>.
template <int N, int ROUND>
struct ReduceScatter {
    static constexpr int H = N / 2, NP = N - H;
    __device__ __forceinline__ static float run(const float (&v)[N], int lane) {
        float w[NP];
        if constexpr (ROUND == 0) {
#pragma unroll
            for (int k = 0; k < H; ++k) w[k] = swap32_sum(v[k], v[k + NP]);
            if constexpr (NP > H) w[H] = swap32_sum(v[H], v[H]);
        } else if constexpr (ROUND == 1) {
#pragma unroll
            for (int k = 0; k < H; ++k) w[k] = swap16_sum(v[k], v[k + NP]);
            if constexpr (NP > H) w[H] = swap16_sum(v[H], v[H]);
        } else {
            constexpr int CTRL = ROUND == 2   ? kDppRowMirror
                                 : ROUND == 3 ? kDppRowHalfMirror
                                 : ROUND == 4 ? kDppQuadXor2
                                              : kDppQuadXor1;
            const bool upper = (lane >> (5 - ROUND)) & 1;
#pragma unroll
            for (int k = 0; k < H; ++k) {
                const float lo = v[k] + dpp<CTRL>(v[k]);
                const float hi = v[k + NP] + dpp<CTRL>(v[k + NP]);
                w[k] = upper ? hi : lo;
            }
            if constexpr (NP > H) w[H] = v[H] + dpp<CTRL>(v[H]);
        }
        if constexpr (ROUND == 5) {
            static_assert(NP == 1, "at most 64 values");
            return w[0];
        } else {
            return ReduceScatter<NP, ROUND + 1>::run(w, lane);
        }
    }
};

template <int N>
__device__ __forceinline__ float wave_reduce_scatter(const float (&v)[N], int lane) {
    return ReduceScatter<N, 0>::run(v, lane);
}

// The value index lane ends with after wave_reduce_scatter<N>, and whether
// it is the lowest lane holding that index (bit 6 set when it is not):
// traced back from the last round to the first.
template <int N>
__device__ __forceinline__ int rs_slot(int lane) {
    int ns[6];
    ns[0] = N;
#pragma unroll
    for (int r = 1; r < 6; ++r) ns[r] = ns[r - 1] - ns[r - 1] / 2;
    int pos = 0, dup = 0;
#pragma unroll
    for (int r = 5; r >= 0; --r) {
        const int h = ns[r] / 2, np = ns[r] - h;
        const int bit = (lane >> (5 - r)) & 1;
        if (pos < h) pos += bit ? np : 0;
        else dup |= bit;
    }
    return pos | (dup << 6);
}

// Index of this wave within the workgroup, as a wave-uniform (SGPR) value:
// divergence analysis treats any function of threadIdx.x as per-lane, which
// would turn loads addressed by it into vector loads.
__device__ __forceinline__ int wave_index() {
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
}

__device__ __forceinline__ void tile_range(const RasterArgs& a, int tile, int& r0, int& r1) {
    const int2 rg = reinterpret_cast<const int2*>(a.tile_bins)[tile];
    r0 = min(max(rg.x, 0), a.num_ids);
    r1 = min(max(rg.y, r0), a.num_ids);
}

__device__ __forceinline__ int splat_id(const RasterArgs& a, int e) {
    const int g = __builtin_amdgcn_readfirstlane(as_const_space(a.ids_sorted)[e]);
    return min(max(g, 0), a.num_points - 1);
}

// Ties values to this program point in SGPRs (an empty asm with "s" operands),
// so the scalar loads producing them are issued together here instead of
// being sunk, one dependent batch at a time, into the branches that use them.
__device__ __forceinline__ void pin_s(float a, float b, float c, float d) {
    asm volatile("" ::"s"(a), "s"(b), "s"(c), "s"(d));
}

// Record of Gaussian g (wave-uniform) as scalar loads, all in flight at once.
__device__ __forceinline__ Splat splat_record(const RasterArgs& a, int g) {
    const Splat s = load_splat(as_const_space(a.records) + (size_t)g * kRecordFloats);
    pin_s(s.a1.x, s.a1.y, s.a1.z, s.iq1);
    pin_s(s.a2.x, s.a2.y, s.a2.z, s.iq2);
    pin_s(s.a3.x, s.a3.y, s.a3.z, s.opac);
    pin_s(s.mo.x, s.mo.y, s.mo.z, s.rgb.x);
    pin_s(s.rgb.y, s.rgb.z, s.u0, s.v0);
    pin_s(s.umap.x, s.umap.y, s.umap.z, s.vmap.x);
    pin_s(s.vmap.y, s.vmap.z, __int_as_float(s.h), __int_as_float(s.w));
    pin_s(__int_as_float(s.start), s.pix_x, s.pix_y, s.num);
    return s;
}

// Pixel-index rectangle [x0, x1] x [y0, y1] covered by this wave's in-image
// pixels (wave-uniform; empty when it has none).
struct WaveRect {
    int x0, x1, y0, y1;
};

__device__ __forceinline__ WaveRect wave_rect(bool inside, int j, int i) {
    WaveRect r;
    r.x0 = wave_min(inside ? j : 0x7fffffff);
    r.x1 = wave_max(inside ? j : -0x7fffffff);
    r.y0 = wave_min(inside ? i : 0x7fffffff);
    r.y1 = wave_max(inside ? i : -0x7fffffff);
    return r;
}

// Does a screen box (int16 pairs lo | hi << 16 for x and y) meet the rectangle?
__device__ __forceinline__ bool box_meets(uint2 bx, const WaveRect& r) {
    const int x0 = (int)(short)(bx.x & 0xffff), x1 = (int)(short)(bx.x >> 16);
    const int y0 = (int)(short)(bx.y & 0xffff), y1 = (int)(short)(bx.y >> 16);
    return (x1 >= r.x0) & (x0 <= r.x1) & (y1 >= r.y0) & (y0 <= r.y1);
}

// One 64-entry window of a tile list (entries base .. base + 63, base a
// multiple of 64): this lane's Gaussian id for entry base + lane (0 outside
// [lo, hi]), the wave mask of the entries in [lo, hi], and of those the ones
// whose screen box meets the wave rectangle (all of them without boxes or
// with cull = false).
// Whether the screen ellipse (p - c)^T [[A, B], [B, C]] (p - c) <= 1 meets
// the pixel centres of the wave rectangle (A <= 0: no ellipse, so yes).  The
// quadratic's minimum over the rectangle, padded by 0.1 pixel, is 0 when the
// centre is inside and otherwise lies on an edge, where it is a clamped 1-D
// minimum.  The argmins use the hardware reciprocal; evaluating the quadratic
// at a slightly moved point on the edge changes the value only to second
// order.
__device__ __forceinline__ bool ellipse_meets(float4 el, float C, const WaveRect& r) {
    if (!(el.z > 0.f)) return true;
    const float A = el.z, B = el.w;
    const float x0 = (float)r.x0 + 0.4f - el.x, x1 = (float)r.x1 + 0.6f - el.x;
    const float y0 = (float)r.y0 + 0.4f - el.y, y1 = (float)r.y1 + 0.6f - el.y;
    if (x0 <= 0.f && x1 >= 0.f && y0 <= 0.f && y1 >= 0.f) return true;
    const float ia = __builtin_amdgcn_rcpf(A), ic = __builtin_amdgcn_rcpf(C);
    auto on_x = [&](float x) {  // edge x = const
        const float y = fminf(fmaxf(-B * x * ic, y0), y1);
        return A * x * x + y * (2.f * B * x + C * y);
    };
    auto on_y = [&](float y) {  // edge y = const
        const float x = fminf(fmaxf(-B * y * ia, x0), x1);
        return C * y * y + x * (2.f * B * y + A * x);
    };
    const float q = fminf(fminf(on_x(x0), on_x(x1)), fminf(on_y(y0), on_y(y1)));
    return q <= 1.f;
}

// local_outline (texture_helpers.cuh:390-417): distance in pixels from pixel
// (j, i) to the nearest pixel of its 7 x 7 neighbourhood whose ray meets the
// splat plane outside sigma <= thresh (1000 when there is none), with the ray
// model and plane hit of the raster.
__device__ inline float outline_distance(const RasterArgs& a, const Splat& s, int j, int i,
                                         float thresh) {
    float best = 1000.f;
    for (int dx = -3; dx <= 3; ++dx)
        for (int dy = -3; dy <= 3; ++dy) {
            const PixelRay q = pixel_ray(a.c2w, a.viewmat, a.fx, a.fy, a.cx, a.cy, j + dx, i + dy);
            float den = dot3(s.a3, q.r);
            if (0.f <= den && den < 1e-6f) den = 1e-6f;
            else if (-1e-6f < den && den <= 0.f) den = -1e-6f;
            const float t = div_fast(s.num, den);
            const V3 d = v3(fmaf(t, q.r.x, -s.mo.x), fmaf(t, q.r.y, -s.mo.y),
                            fmaf(t, q.r.z, -s.mo.z));
            const float l1 = dot3(d, s.a1), l2 = dot3(d, s.a2);
            if (0.5f * (s.iq1 * l1 * l1 + s.iq2 * l2 * l2) > thresh)
                best = fminf(best, sqrtf((float)(dx * dx + dy * dy)));
        }
    return best;
}

struct Window {
    uint64_t in_range, cand;
    int gid;
};

__device__ __forceinline__ Window window_candidates(const RasterArgs& a, int base, int lo, int hi,
                                                   const WaveRect& rect, int lane, bool cull) {
    Window w;
    const int e = base + lane;
    const bool in_range = e >= lo && e <= hi;
    bool cand = in_range;
    w.gid = 0;
    if (in_range) {
        w.gid = min(max(a.ids_sorted[e], 0), a.num_points - 1);
        if (cull) {
            const uint2 bx = a.boxes[w.gid];  // one 8-byte load
            cand = box_meets(bx, rect);
            if (a.ellipses != nullptr) {
                const float4 el = a.ellipses[w.gid];
                const float ec = a.ellipse_c[w.gid];
                cand = cand && ellipse_meets(el, ec, rect);
            }
        }
    }
    w.in_range = lanes_where(in_range);
    w.cand = lanes_where(cand);
    return w;
}

inline RasterArgs make_raster_args(const RasterCall& c, const float* records) {
    RasterArgs a;
    a.img_w = c.img_w;
    a.img_h = c.img_h;
    a.bw = c.block_width;
    a.tiles_x = (c.img_w + c.block_width - 1) / c.block_width;
    a.tiles_y = (c.img_h + c.block_width - 1) / c.block_width;
    a.channels = c.channels;
    a.settings = c.settings;
    a.num_points = c.num_points;
    a.num_ids = c.num_ids;
    a.num_texels = c.num_texels;
    a.fx = c.fx;
    a.fy = c.fy;
    a.cx = c.cx;
    a.cy = c.cy;
    a.viewmat = c.viewmat;
    a.c2w = c.c2w;
    a.background = c.background;
    a.ids_sorted = c.ids_sorted;
    a.tile_bins = c.tile_bins;
    a.records = records;
    a.texture = c.texture;
    a.texture_bf16 = c.texture_bf16 != 0;
    a.precise_edge = raster_precise_edge_enabled();
    a.wave_masks = c.wave_masks;
    a.mask_plane = (int)raster_wave_mask_words(c.num_ids, a.tiles_x * a.tiles_y);
    // boxes hold int16 pixel indices, so culling is off for larger images
    // the alpha visualisation mode draws splats out to its own sigma bound,
    // beyond the alpha >= 1/255 region the culling bounds hold
    a.boxes = raster_culling_enabled() && c.num_points > 0 && c.img_w < 32000 &&
                      c.img_h < 32000 && !(c.settings & kSetAlphaViz)
                  ? record_boxes(records, c.num_points)
                  : nullptr;
    // the ellipses bound the Gaussian alpha only; with blur the box also
    // covers the blur disc
    const bool ell = a.boxes != nullptr && !(c.settings & kSetBlur);
    a.ellipses = ell ? record_conics(records, c.num_points) : nullptr;
    a.ellipse_c = ell ? record_conic_c(records, c.num_points) : nullptr;
    return a;
}

inline hipError_t build_raster_records(const RasterCall& c, float* records, hipStream_t s) {
    return launch_build_records(c.num_points, c.means, c.scales, c.glob_scale, c.quats, c.colors,
                                c.opacities, c.uv0, c.umap, c.vmap, c.texture_dims, c.viewmat,
                                c.c2w, c.fx, c.fy, c.cx, c.cy, records, s);
}

}  // namespace gstex
