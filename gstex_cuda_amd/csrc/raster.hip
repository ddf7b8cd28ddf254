// Textured 2D-Gaussian tile rasteriser for gfx950: forward compositing,
// back-to-front backward, and the per-Gaussian gradient finalize.
//
// Work decomposition: one workgroup per bw x bw screen tile; for bw = 16 the
// tile is four 8x8 quadrants, one wave64 each.  A wave walks its tile's
// depth-sorted Gaussian list on its own (no workgroup barriers): the list
// index, the Gaussian id and its 128-byte splat record are wave-uniform, so
// they come through scalar loads into SGPRs, and a wave retires as soon as
// all 64 of its pixels are saturated.
//
// Per-pixel semantics follow texture.cu:11-329 (forward) and 331-760
// (backward): ray/plane intersection, Gaussian alpha with the 0.99 clamp and
// 1/255 skip, termination at T <= 1e-4, colour + normal + bilinear texture
// compositing, median depth and the L2 depth distortion.
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
    const float* texture;
};

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
};

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
    V3 a1, a2, a3, m, rgb, umap, vmap;
    float is1, is2, opac, pix_x, pix_y, u0, v0;
    int h, w, start;
};

__device__ __forceinline__ Splat load_splat(const float* __restrict__ rec) {
    Splat s;
    const float4* q = reinterpret_cast<const float4*>(rec);
    const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5], q6 = q[6],
                 q7 = q[7];
    s.a1 = v3(q0.x, q0.y, q0.z);
    s.is1 = q0.w;
    s.a2 = v3(q1.x, q1.y, q1.z);
    s.is2 = q1.w;
    s.a3 = v3(q2.x, q2.y, q2.z);
    s.opac = q2.w;
    s.m = v3(q3.x, q3.y, q3.z);
    s.pix_x = q3.w;
    s.rgb = v3(q4.x, q4.y, q4.z);
    s.pix_y = q4.w;
    s.u0 = q5.x;
    s.v0 = q5.y;
    s.umap = v3(q5.z, q5.w, q6.x);
    s.vmap = v3(q6.y, q6.z, q6.w);
    s.h = __float_as_int(q7.x);
    s.w = __float_as_int(q7.y);
    s.start = __float_as_int(q7.z);
    return s;
}

struct Pair {
    V3 delta;
    float den, t, l1, l2, G, Gb, alpha;
    bool blur_sel, skip;
};

// Ray/splat intersection and alpha for one (pixel, Gaussian) pair.
__device__ __forceinline__ Pair pair_eval(const Splat& s, const PixelRay& pr, bool use_blur) {
    Pair p;
    float den = dot3(s.a3, pr.r);
    const float eps = 1e-6f;
    if (0.f <= den && den < eps) den = eps;
    else if (-eps < den && den <= 0.f) den = -eps;
    p.den = den;
    p.t = dot3(s.a3, sub3(s.m, pr.o)) / den;
    const V3 pos = v3(pr.o.x + p.t * pr.r.x, pr.o.y + p.t * pr.r.y, pr.o.z + p.t * pr.r.z);
    p.delta = sub3(pos, s.m);
    p.l1 = dot3(p.delta, s.a1);
    p.l2 = dot3(p.delta, s.a2);
    const float sigma = 0.5f * (s.is1 * s.is1 * p.l1 * p.l1 + s.is2 * s.is2 * p.l2 * p.l2);
    p.G = __expf(-sigma);
    p.Gb = 0.f;
    p.blur_sel = false;
    if (use_blur) {
        const float dx = s.pix_x - pr.px, dy = s.pix_y - pr.py;
        const float sigma_blur = 0.5f * 2.0f * (dx * dx + dy * dy);
        p.Gb = __expf(-sigma_blur);
        p.blur_sel = sigma_blur < sigma;
    }
    p.alpha = fminf(0.99f, s.opac * (p.blur_sel ? p.Gb : p.G));
    p.skip = p.t < 0.01f || p.t > 1000.0f || p.alpha < 1.f / 255.f;
    return p;
}

// Bilinear (or nearest) taps of one chart at clamped (u, v): weights and the
// texel rows (start + i*w + j), clamped into [0, X) for memory safety.
__device__ __forceinline__ void chart_taps(const Splat& s, float u, float v, bool bilinear,
                                           int num_texels, float wt[4], int row[4], float& fu,
                                           float& fv) {
    const float tu = (float)s.h * u, tv = (float)s.w * v;
    int iu = f2i_trunc(tu), iv = f2i_trunc(tv);
    fu = tu - (float)iu;
    fv = tv - (float)iv;
    const int ju = max(min(iu + 1, s.h - 1), 0), jv = max(min(iv + 1, s.w - 1), 0);
    iu = max(min(iu, s.h - 1), 0);
    iv = max(min(iv, s.w - 1), 0);
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
    const int lim = num_texels - 1;
    row[0] = min(max(s.start + iu * s.w + iv, 0), lim);
    row[1] = min(max(s.start + iu * s.w + jv, 0), lim);
    row[2] = min(max(s.start + ju * s.w + iv, 0), lim);
    row[3] = min(max(s.start + ju * s.w + jv, 0), lim);
}

__device__ __forceinline__ void splat_uv(const Splat& s, V3 delta, float& u, float& v) {
    u = s.u0 + dot3(s.umap, delta);
    v = s.v0 + dot3(s.vmap, delta);
    u = fminf(fmaxf(u, 0.f), 1.f);
    v = fminf(fmaxf(v, 0.f), 1.f);
}

__device__ __forceinline__ void tile_range(const RasterArgs& a, int tile, int& r0, int& r1) {
    const int2 rg = reinterpret_cast<const int2*>(a.tile_bins)[tile];
    r0 = min(max(rg.x, 0), a.num_ids);
    r1 = min(max(rg.y, r0), a.num_ids);
}

__device__ __forceinline__ int splat_id(const RasterArgs& a, int e) {
    const int g = __builtin_amdgcn_readfirstlane(a.ids_sorted[e]);
    return min(max(g, 0), a.num_points - 1);
}

// ------------------------------------------------------------------ forward
template <int CM>
__global__ __launch_bounds__(256) void k_raster_forward(RasterArgs a, FwdOut o) {
    const int tile = blockIdx.x;
    const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    int lx, ly;
    bool in_tile;
    lane_pixel(a.bw, threadIdx.x, lx, ly, in_tile);
    const int j = tx * a.bw + lx, i = ty * a.bw + ly;
    const bool inside = in_tile && i < a.img_h && j < a.img_w;
    const PixelRay pr = pixel_ray(a.c2w, a.viewmat, a.fx, a.fy, a.cx, a.cy, j, i);
    const bool use_blur = a.settings & kSetBlur, use_ndc = a.settings & kSetNdc;
    const bool flip = a.settings & kSetNormalFlip, bilinear = !(a.settings & kSetNearest);
    const int C = a.channels;

    int r0, r1;
    tile_range(a, tile, r0, r1);
    bool done = !inside;
    float T = 1.f;
    int cur_idx = 0, pdi = -1;
    float rgb0 = 0.f, rgb1 = 0.f, rgb2 = 0.f, n0 = 0.f, n1 = 0.f, n2 = 0.f;
    float pdepth = 0.f, preg = 0.f, s0 = 0.f, s1 = 0.f, s2 = 0.f;
    float tex[CM];
#pragma unroll
    for (int c = 0; c < CM; ++c) tex[c] = 0.f;

    for (int e = r0; e < r1; ++e) {
        if (!__any(!done)) break;
        const int g = splat_id(a, e);
        const Splat s = load_splat(a.records + (size_t)g * kRecordFloats);
        if (done) continue;
        const Pair p = pair_eval(s, pr, use_blur);
        const float next_T = T * (1.f - p.alpha);
        if (next_T <= 1e-4f) done = true;
        if (p.skip || done) continue;
        const float vis = p.alpha * T;
        rgb0 += s.rgb.x * vis;
        rgb1 += s.rgb.y * vis;
        rgb2 += s.rgb.z * vis;
        V3 hn = s.a3;
        if (flip && dot3(pr.r, s.a3) > 0.f) hn = v3(-hn.x, -hn.y, -hn.z);
        n0 += vis * hn.x;
        n1 += vis * hn.y;
        n2 += vis * hn.z;
        const float t_view = p.t * pr.view_depth;
        float u, v;
        splat_uv(s, p.delta, u, v);
        float wt[4], fu, fv;
        int row[4];
        chart_taps(s, u, v, bilinear, a.num_texels, wt, row, fu, fv);
#pragma unroll
        for (int c = 0; c < CM; ++c) {
            if (c < C) {
                const float val = wt[0] * a.texture[(size_t)row[0] * C + c] +
                                  wt[1] * a.texture[(size_t)row[1] * C + c] +
                                  wt[2] * a.texture[(size_t)row[2] * C + c] +
                                  wt[3] * a.texture[(size_t)row[3] * C + c];
                tex[c] += vis * val;
            }
        }
        if (T > 0.5f) {
            pdepth = t_view;
            pdi = e;
        }
        float dval = p.t;
        if (use_ndc) {
            const float tn = 0.01f, tf = 1000.0f;
            dval = (tf * t_view - tf * tn) / ((tf - tn) * t_view);
        }
        preg += vis * (dval * dval * s0 + s2 - 2.f * dval * s1);
        s0 += vis;
        s1 += vis * dval;
        s2 += vis * dval * dval;
        T = next_T;
        cur_idx = e;
    }
    if (!inside) return;
    const int pid = i * a.img_w + j;
    const float bg0 = a.background[0], bg1 = a.background[1], bg2 = a.background[2];
    o.final_Ts[pid] = T;
    o.final_idx[pid] = cur_idx;
    o.depth_idx[pid] = pdi;
    o.out_img[3 * pid + 0] = rgb0 + T * bg0;
    o.out_img[3 * pid + 1] = rgb1 + T * bg1;
    o.out_img[3 * pid + 2] = rgb2 + T * bg2;
    o.out_normal[3 * pid + 0] = n0;
    o.out_normal[3 * pid + 1] = n1;
    o.out_normal[3 * pid + 2] = n2;
    o.out_depth[pid] = pdepth;
    o.out_reg[pid] = preg;
    o.out_reg_s[3 * pid + 0] = s0;
    o.out_reg_s[3 * pid + 1] = s1;
    o.out_reg_s[3 * pid + 2] = s2;
#pragma unroll
    for (int c = 0; c < CM; ++c)
        if (c < C) o.out_texture[(size_t)pid * C + c] = tex[c];
}

// ------------------------------------------------------------------ wave reduction
// Reduce 32 per-lane values over the 64-lane wave by recursive halving: each
// exchange round halves the values a lane carries, so the whole reduction
// costs ~32 shuffles instead of 32 x 6.  On return lane L holds the wave sum
// of value index L >> 1.
__device__ __forceinline__ float wave_reduce32(float (&v)[kGradFloats], int lane) {
#pragma unroll
    for (int rnd = 0; rnd < 5; ++rnd) {
        const int half = 16 >> rnd;      // values kept after this round
        const int off = 32 >> rnd;       // partner distance
        const bool upper = (lane & off) != 0;
#pragma unroll
        for (int k = 0; k < half; ++k) {
            const float lo = v[k], hi = v[k + half];
            const float send = upper ? lo : hi;
            const float keep = upper ? hi : lo;
            v[k] = keep + __shfl_xor(send, off);
        }
    }
    return v[0] + __shfl_xor(v[0], 1);
}

// ------------------------------------------------------------------ backward
template <int CM>
__global__ __launch_bounds__(256) void k_raster_backward(RasterArgs a, BwdIn b) {
    const int tile = blockIdx.x;
    const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    const int lane = threadIdx.x & (kWave - 1);
    int lx, ly;
    bool in_tile;
    lane_pixel(a.bw, threadIdx.x, lx, ly, in_tile);
    const int j = tx * a.bw + lx, i = ty * a.bw + ly;
    const bool inside = in_tile && i < a.img_h && j < a.img_w;
    const PixelRay pr = pixel_ray(a.c2w, a.viewmat, a.fx, a.fy, a.cx, a.cy, j, i);
    const bool use_blur = a.settings & kSetBlur, use_ndc = a.settings & kSetNdc;
    const bool bilinear = !(a.settings & kSetNearest), prop_uv = a.settings & kSetUvGrad;
    const int C = a.channels;
    const int pid = inside ? i * a.img_w + j : 0;

    float T = 1.f, S0 = 0.f, S1 = 0.f, S2 = 0.f, vdep = 0.f, vreg = 0.f, vT = 0.f;
    V3 vo = v3(0.f, 0.f, 0.f), vn = v3(0.f, 0.f, 0.f);
    int dfin = -1, bfin = -1;
    float vtex[CM];
#pragma unroll
    for (int c = 0; c < CM; ++c) vtex[c] = 0.f;
    if (inside) {
        T = b.final_Ts[pid];
        S0 = b.final_s[3 * pid];
        S1 = b.final_s[3 * pid + 1];
        S2 = b.final_s[3 * pid + 2];
        dfin = b.depth_idx[pid];
        bfin = b.final_idx[pid];
        vo = v3(b.v_img[3 * pid], b.v_img[3 * pid + 1], b.v_img[3 * pid + 2]);
        vn = v3(b.v_normal[3 * pid], b.v_normal[3 * pid + 1], b.v_normal[3 * pid + 2]);
        vdep = b.v_depth[pid];
        vreg = b.v_reg[pid];
        vT = a.background[0] * vo.x + a.background[1] * vo.y + a.background[2] * vo.z -
             b.v_alpha[pid];
#pragma unroll
        for (int c = 0; c < CM; ++c)
            if (c < C) vtex[c] = b.v_texture_out[(size_t)pid * C + c];
    }
    int r0, r1;
    tile_range(a, tile, r0, r1);
    // Entries past the wave's last contributing index are dead for every lane.
    int wave_last = bfin;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) wave_last = max(wave_last, __shfl_xor(wave_last, off));
    wave_last = __builtin_amdgcn_readfirstlane(wave_last);
    const int e_hi = min(r1 - 1, wave_last);
    const float tn = 0.01f, tf = 1000.0f;

    for (int e = e_hi; e >= r0; --e) {
        const int g = splat_id(a, e);
        const Splat s = load_splat(a.records + (size_t)g * kRecordFloats);
        bool valid = inside && e <= bfin;
        Pair p;
        if (valid) {
            p = pair_eval(s, pr, use_blur);
            valid = !p.skip;
        }
        if (!__any(valid)) continue;
        float gr[kGradFloats];
#pragma unroll
        for (int k = 0; k < kGradFloats; ++k) gr[k] = 0.f;
        if (valid) {
            T *= 1.f / (1.f - p.alpha);
            const float vis = p.alpha * T;
            gr[G_RGB + 0] = vis * vo.x;
            gr[G_RGB + 1] = vis * vo.y;
            gr[G_RGB + 2] = vis * vo.z;
            float vvis = dot3(s.rgb, vo) + dot3(s.a3, vn);
            float u, v;
            splat_uv(s, p.delta, u, v);
            float wt[4], fu, fv;
            int row[4];
            chart_taps(s, u, v, bilinear, a.num_texels, wt, row, fu, fv);
            float vu = 0.f, vv = 0.f;
#pragma unroll
            for (int c = 0; c < CM; ++c) {
                if (c < C) {
                    const float c0 = a.texture[(size_t)row[0] * C + c];
                    const float c1 = a.texture[(size_t)row[1] * C + c];
                    const float c2 = a.texture[(size_t)row[2] * C + c];
                    const float c3 = a.texture[(size_t)row[3] * C + c];
                    const float val = wt[0] * c0 + wt[1] * c1 + wt[2] * c2 + wt[3] * c3;
                    const float vval = vis * vtex[c];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const float add = wt[k] * vval;
                        if (add != 0.f) unsafeAtomicAdd(b.v_texture + (size_t)row[k] * C + c, add);
                    }
                    if (bilinear && prop_uv) {
                        vu += (float)s.h * (vval * (-(1.f - fv) * c0 - fv * c1 + (1.f - fv) * c2 + fv * c3));
                        vv += (float)s.w * (vval * (-(1.f - fu) * c0 + (1.f - fu) * c1 - fu * c2 + fu * c3));
                    }
                    vvis += val * vtex[c];
                }
            }
            gr[G_UV0 + 0] = vu;
            gr[G_UV0 + 1] = vv;
            gr[G_UMAP + 0] = p.delta.x * vu;
            gr[G_UMAP + 1] = p.delta.y * vu;
            gr[G_UMAP + 2] = p.delta.z * vu;
            gr[G_VMAP + 0] = p.delta.x * vv;
            gr[G_VMAP + 1] = p.delta.y * vv;
            gr[G_VMAP + 2] = p.delta.z * vv;
            const V3 vd_uv = v3(s.umap.x * vu + s.vmap.x * vv, s.umap.y * vu + s.vmap.y * vv,
                                s.umap.z * vu + s.vmap.z * vv);

            float valpha = T * vvis - T * vT;
            float vTcur = p.alpha * vvis + (1.f - p.alpha) * vT;
            const float t_view = p.t * pr.view_depth;
            const float dval = use_ndc ? (tf * t_view - tf * tn) / ((tf - tn) * t_view) : p.t;
            const float vd_val = 2.f * (vis * dval * S0 - vis * S1) * vreg;
            const float vw = (dval * dval * S0 - 2.f * dval * S1 + S2) * vreg;
            valpha += vw * T;
            vTcur += vw * p.alpha;
            vT = vTcur;

            float vt = use_ndc ? 0.f : vd_val;
            float vtview = (e == dfin && dfin != -1) ? vdep : 0.f;
            if (use_ndc) vtview += (tf * tn) / ((tf - tn) * t_view * t_view) * vd_val;
            vt += pr.view_depth * vtview;

            const float gsel = p.blur_sel ? p.Gb : p.G;
            gr[G_OPAC] = gsel * valpha;
            const float vsigma = p.blur_sel ? 0.f : -s.opac * p.G * valpha;
            const float vl1 = vsigma * s.is1 * s.is1 * p.l1;
            const float vl2 = vsigma * s.is2 * s.is2 * p.l2;
            gr[G_INV_SX] = vsigma * s.is1 * p.l1 * p.l1;
            gr[G_INV_SY] = vsigma * s.is2 * p.l2 * p.l2;
            const V3 vdelta = v3(vl1 * s.a1.x + vl2 * s.a2.x + vd_uv.x,
                                 vl1 * s.a1.y + vl2 * s.a2.y + vd_uv.y,
                                 vl1 * s.a1.z + vl2 * s.a2.z + vd_uv.z);
            gr[G_AX1 + 0] = vl1 * p.delta.x;
            gr[G_AX1 + 1] = vl1 * p.delta.y;
            gr[G_AX1 + 2] = vl1 * p.delta.z;
            gr[G_AX2 + 0] = vl2 * p.delta.x;
            gr[G_AX2 + 1] = vl2 * p.delta.y;
            gr[G_AX2 + 2] = vl2 * p.delta.z;
            vt += dot3(vdelta, pr.r);
            const float vnum = vt / p.den;
            const float vden = -vt * p.t / p.den;
            const V3 mo = sub3(s.m, pr.o);
            gr[G_MEAN + 0] = -vdelta.x + vnum * s.a3.x;
            gr[G_MEAN + 1] = -vdelta.y + vnum * s.a3.y;
            gr[G_MEAN + 2] = -vdelta.z + vnum * s.a3.z;
            gr[G_AX3 + 0] = vnum * mo.x + vden * pr.r.x + vis * vn.x;
            gr[G_AX3 + 1] = vnum * mo.y + vden * pr.r.y + vis * vn.y;
            gr[G_AX3 + 2] = vnum * mo.z + vden * pr.r.z + vis * vn.z;
            if (use_blur) {
                const float vsb = p.blur_sel ? -s.opac * p.Gb * valpha : 0.f;
                gr[G_PIX + 0] = 2.0f * vsb * (s.pix_x - pr.px);
                gr[G_PIX + 1] = 2.0f * vsb * (s.pix_y - pr.py);
            }
        }
        const float tot = wave_reduce32(gr, lane);
        const int slot = lane >> 1;
        if ((lane & 1) == 0 && slot < kGradUsed && tot != 0.f)
            unsafeAtomicAdd(b.grad_rec + (size_t)g * kGradFloats + slot, tot);
    }
}

// Map the per-Gaussian accumulators onto the reference gradient tensors.
__global__ __launch_bounds__(256) void k_grad_finalize(
    int n, const float* __restrict__ grad_rec, const float* __restrict__ means,
    const float* __restrict__ scales, float glob_scale, const float* __restrict__ quats,
    const float* __restrict__ viewmat, float fx, float fy, float* __restrict__ v_colors,
    float* __restrict__ v_opacity, float* __restrict__ v_means, float* __restrict__ v_scales,
    float* __restrict__ v_quats, float* __restrict__ v_uv0, float* __restrict__ v_umap,
    float* __restrict__ v_vmap) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    float acc[kGradFloats];
    const float4* src = reinterpret_cast<const float4*>(grad_rec + (size_t)g * kGradFloats);
#pragma unroll
    for (int k = 0; k < kGradFloats / 4; ++k) {
        const float4 q = src[k];
        acc[4 * k] = q.x; acc[4 * k + 1] = q.y; acc[4 * k + 2] = q.z; acc[4 * k + 3] = q.w;
    }
    v_colors[3 * g] = acc[G_RGB];
    v_colors[3 * g + 1] = acc[G_RGB + 1];
    v_colors[3 * g + 2] = acc[G_RGB + 2];
    v_opacity[g] = acc[G_OPAC];
    V3 vm = v3(acc[G_MEAN], acc[G_MEAN + 1], acc[G_MEAN + 2]);
    if (acc[G_PIX] != 0.f || acc[G_PIX + 1] != 0.f) {
        // blur term: d projected mean / d mean (helpers.cuh:114-121,155-164)
        const V3 m = v3(means[3 * g], means[3 * g + 1], means[3 * g + 2]);
        const V3 pv = xform43(viewmat, m);
        const float rw = 1.f / (pv.z + 1e-6f);
        const float vpx = fx * acc[G_PIX], vpy = fy * acc[G_PIX + 1];
        const V3 vview = v3(vpx * rw, vpy * rw, -(vpx * pv.x + vpy * pv.y) * rw * rw);
        const V3 vmx = xform43_rot_t(viewmat, vview);
        vm = v3(vm.x + vmx.x, vm.y + vmx.y, vm.z + vmx.z);
    }
    v_means[3 * g] = vm.x;
    v_means[3 * g + 1] = vm.y;
    v_means[3 * g + 2] = vm.z;
    const float4 q = reinterpret_cast<const float4*>(quats)[g];
    float vq[4];
    quat_axes_vjp(q.x, q.y, q.z, q.w, v3(acc[G_AX1], acc[G_AX1 + 1], acc[G_AX1 + 2]),
                  v3(acc[G_AX2], acc[G_AX2 + 1], acc[G_AX2 + 2]),
                  v3(acc[G_AX3], acc[G_AX3 + 1], acc[G_AX3 + 2]), vq);
    reinterpret_cast<float4*>(v_quats)[g] = make_float4(vq[0], vq[1], vq[2], vq[3]);
    const float sx = scales[3 * g], sy = scales[3 * g + 1];
    const float is1 = 1.f / (sx * glob_scale), is2 = 1.f / (sy * glob_scale);
    v_scales[3 * g] = -acc[G_INV_SX] * is1 / sx;
    v_scales[3 * g + 1] = -acc[G_INV_SY] * is2 / sy;
    v_scales[3 * g + 2] = 0.f;
    v_uv0[2 * g] = acc[G_UV0];
    v_uv0[2 * g + 1] = acc[G_UV0 + 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v_umap[3 * g + k] = acc[G_UMAP + k];
        v_vmap[3 * g + k] = acc[G_VMAP + k];
    }
}

// ------------------------------------------------------------------ launchers
template <int CM>
static hipError_t launch_fwd_cm(const RasterArgs& a, const FwdOut& o, hipStream_t s) {
    const int tiles = a.tiles_x * a.tiles_y;
    k_raster_forward<CM><<<tiles, raster_block_threads(a.bw), 0, s>>>(a, o);
    return hipGetLastError();
}

template <int CM>
static hipError_t launch_bwd_cm(const RasterArgs& a, const BwdIn& b, hipStream_t s) {
    const int tiles = a.tiles_x * a.tiles_y;
    k_raster_backward<CM><<<tiles, raster_block_threads(a.bw), 0, s>>>(a, b);
    return hipGetLastError();
}

static RasterArgs make_args(const RasterCall& c, const float* records) {
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
    return a;
}

size_t raster_forward_workspace_bytes(int n) {
    return (size_t)n * kRecordFloats * sizeof(float) + 256;
}

size_t raster_backward_workspace_bytes(int n) {
    return 2 * (size_t)n * kRecordFloats * sizeof(float) + 256;
}

static hipError_t build_records(const RasterCall& c, float* records, hipStream_t s) {
    return launch_build_records(c.num_points, c.means, c.scales, c.glob_scale, c.quats, c.colors,
                                c.opacities, c.uv0, c.umap, c.vmap, c.texture_dims, c.viewmat,
                                c.fx, c.fy, c.cx, c.cy, records, s);
}

hipError_t raster_forward_stage(const RasterCall& c, const float* records, float* out_img,
                                float* out_depth, float* out_reg, float* out_texture,
                                float* out_normal, float* final_Ts, int* final_idx,
                                int* depth_idx, float* out_reg_s, hipStream_t s) {
    const RasterArgs a = make_args(c, records);
    const FwdOut o{out_img, out_depth, out_reg, out_texture, out_normal,
                   final_Ts, final_idx, depth_idx, out_reg_s};
    if (a.tiles_x * a.tiles_y <= 0) return hipSuccess;
    if (c.channels <= 1) return launch_fwd_cm<1>(a, o, s);
    if (c.channels <= 3) return launch_fwd_cm<3>(a, o, s);
    if (c.channels <= 4) return launch_fwd_cm<4>(a, o, s);
    if (c.channels <= 8) return launch_fwd_cm<8>(a, o, s);
    if (c.channels <= 16) return launch_fwd_cm<16>(a, o, s);
    return launch_fwd_cm<64>(a, o, s);
}

hipError_t raster_backward_stage(const RasterCall& c, const float* records,
                                 const float* final_Ts, const int* final_idx,
                                 const int* depth_idx, const float* final_s, const float* v_img,
                                 const float* v_depth, const float* v_reg, const float* v_alpha,
                                 const float* v_texture_out, const float* v_normal,
                                 float* grad_rec, float* v_texture, hipStream_t s) {
    const RasterArgs a = make_args(c, records);
    const BwdIn b{final_Ts, final_idx, depth_idx, final_s, v_img, v_depth, v_reg, v_alpha,
                  v_texture_out, v_normal, grad_rec, v_texture};
    if (a.tiles_x * a.tiles_y <= 0) return hipSuccess;
    if (c.channels <= 1) return launch_bwd_cm<1>(a, b, s);
    if (c.channels <= 3) return launch_bwd_cm<3>(a, b, s);
    if (c.channels <= 4) return launch_bwd_cm<4>(a, b, s);
    if (c.channels <= 8) return launch_bwd_cm<8>(a, b, s);
    if (c.channels <= 16) return launch_bwd_cm<16>(a, b, s);
    return launch_bwd_cm<64>(a, b, s);
}

hipError_t grad_finalize(int n, const float* grad_rec, const float* means, const float* scales,
                         float glob_scale, const float* quats, const float* viewmat, float fx,
                         float fy, float* v_colors, float* v_opacity, float* v_means,
                         float* v_scales, float* v_quats, float* v_uv0, float* v_umap,
                         float* v_vmap, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_grad_finalize<<<(n + 255) / 256, 256, 0, s>>>(n, grad_rec, means, scales, glob_scale, quats,
                                                   viewmat, fx, fy, v_colors, v_opacity, v_means,
                                                   v_scales, v_quats, v_uv0, v_umap, v_vmap);
    return hipGetLastError();
}

hipError_t raster_forward(const RasterCall& c, float* out_img, float* out_depth, float* out_reg,
                          float* out_texture, float* out_normal, float* final_Ts, int* final_idx,
                          int* depth_idx, float* out_reg_s, void* ws, size_t ws_bytes,
                          hipStream_t s) {
    if (ws_bytes < raster_forward_workspace_bytes(c.num_points)) return hipErrorInvalidValue;
    float* records = (float*)ws;
    hipError_t e = build_records(c, records, s);
    if (e != hipSuccess) return e;
    return raster_forward_stage(c, records, out_img, out_depth, out_reg, out_texture, out_normal,
                                final_Ts, final_idx, depth_idx, out_reg_s, s);
}

hipError_t raster_backward(const RasterCall& c, const float* final_Ts, const int* final_idx,
                           const int* depth_idx, const float* final_s, const float* v_img,
                           const float* v_depth, const float* v_reg, const float* v_alpha,
                           const float* v_texture_out, const float* v_normal, float* v_colors,
                           float* v_opacity, float* v_means, float* v_scales, float* v_quats,
                           float* v_uv0, float* v_umap, float* v_vmap, float* v_texture, void* ws,
                           size_t ws_bytes, hipStream_t s) {
    if (ws_bytes < raster_backward_workspace_bytes(c.num_points)) return hipErrorInvalidValue;
    float* records = (float*)ws;
    float* grad_rec = records + (size_t)c.num_points * kRecordFloats;
    hipError_t e = build_records(c, records, s);
    if (e != hipSuccess) return e;
    e = fill_f32(grad_rec, (size_t)c.num_points * kGradFloats, 0.f, s);
    if (e != hipSuccess) return e;
    e = fill_f32(v_texture, (size_t)c.num_texels * (size_t)c.channels, 0.f, s);
    if (e != hipSuccess) return e;
    e = raster_backward_stage(c, records, final_Ts, final_idx, depth_idx, final_s, v_img, v_depth,
                              v_reg, v_alpha, v_texture_out, v_normal, grad_rec, v_texture, s);
    if (e != hipSuccess) return e;
    return grad_finalize(c.num_points, grad_rec, c.means, c.scales, c.glob_scale, c.quats,
                         c.viewmat, c.fx, c.fy, v_colors, v_opacity, v_means, v_scales, v_quats,
                         v_uv0, v_umap, v_vmap, s);
}

}  // namespace gstex
