// Textured 2D-Gaussian tile rasteriser for gfx950: forward compositing.
//
// Work decomposition: one workgroup per bw x bw screen tile; for bw = 16 the
// tile is four 8x8 quadrants, one wave64 each.  A wave walks its tile's
// depth-sorted Gaussian list on its own (no workgroup barriers): the list
// index, the Gaussian id and its 128-byte splat record are wave-uniform, so
// they come through scalar loads into SGPRs, and a wave retires as soon as
// all 64 of its pixels are saturated.
//
// Per-pixel semantics follow texture.cu:11-329: ray/plane intersection,
// Gaussian alpha with the 0.99 clamp and 1/255 skip, termination at
// T <= 1e-4, colour + normal + bilinear texture compositing, median depth and
// the L2 depth distortion.  The backward lives in raster_bwd.hip.
#include "raster_common.hpp"

namespace gstex {

// MODE 0 compiles out the blur (settings bit 9), NDC distortion (bit 10), alpha visualisation (bits 16-29)
// and float64 edge-on code for the common settings; MODE 1 keeps all of it;
// MODE 2 is MODE 0 plus the float64 edge-on solve.  TB selects bfloat16
// texels.  CM <= 4 is the exact channel count; above 4, CM bounds a runtime
// count.
//
// raster_forward_wave runs one wave's share of a tile: wave `wave` of the
// tile's raster_block_threads(bw) / 64 waves (an 8x8 quadrant for bw = 16).
template <int CM, int MODE, bool TB>
__device__ __forceinline__ void raster_forward_wave(const RasterArgs& a, const FwdOut& o,
                                                    const int tile, const int wave) {
    constexpr bool GENERAL = MODE == 1, PRECISE = MODE != 0;
    const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    const int lane = threadIdx.x & (kWave - 1);
    int lx, ly;
    bool in_tile;
    lane_pixel(a.bw, wave * kWave + lane, lx, ly, in_tile);
    const int j = tx * a.bw + lx, i = ty * a.bw + ly;
    const bool inside = in_tile && i < a.img_h && j < a.img_w;
    const PixelRay pr = pixel_ray(a.c2w, a.viewmat, a.fx, a.fy, a.cx, a.cy, j, i);
    const bool use_blur = GENERAL && (a.settings & kSetBlur);
    const bool use_ndc = GENERAL && (a.settings & kSetNdc);
    const bool flip = a.settings & kSetNormalFlip, bilinear = !(a.settings & kSetNearest);
    // alpha visualisation (texture.cu:56-65,192-212,226-241): alpha 0.99 inside
    // sigma <= 0.5 bound^2, else 0; near the edge the splat is drawn white or not
    const bool viz = GENERAL && (a.settings & kSetAlphaViz);
    const float viz_bound = viz_alpha_bound(a.settings);
    const float viz_thresh = 0.5f * viz_bound * viz_bound;
    const float viz_outline = viz_outline_bound(a.settings);
    const bool viz_white = a.settings & kSetWhiteOutline, viz_opac = a.settings & kSetOpacityThresh;
    const int C = CM <= 4 ? CM : a.channels;
    constexpr int E = TB ? 2 : 4;

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

    // The list is walked in 64-entry windows aligned to multiples of 64: one
    // coalesced read of the window's ids, a ballot of the entries whose screen
    // box meets this wave's pixels, then the candidates front to back.  The
    // composited entries of each window form one word of the wave masks
    // (every window the walk reaches is written, zeros included).
    uint64_t* const mplane =
        a.wave_masks ? a.wave_masks + (size_t)wave * a.mask_plane : nullptr;
    const int mword0 = (r0 >> 6) + tile, mentry0 = r0 & ~63;
    // Culling rectangle: the wave's live pixels.  A pixel that is done can
    // take nothing more, so once pixels finish the rectangle shrinks to the
    // ones still live (recomputed when their count drops), and an entry that
    // meets none of them is passed over like any culled entry.
    WaveRect rect = wave_rect(inside, j, i);
    int live = __popcll(lanes_where(inside));

    // A pixel whose T (1 - 1/255) <= 1e-4 can composite nothing more: every
    // entry with alpha >= 1/255 ends it (rounding is monotonic), and ending
    // changes none of its outputs.  Such a pixel stops at once.  Every live
    // pixel then has T (1 - 1/255) > 1e-4, so an entry with alpha < 1/255 at
    // all of the wave's pixels (a culled one) can neither composite nor end
    // any of them, and culling stays on for the whole walk.
    constexpr float kKeep = 1.f - 1.f / 255.f;
    const bool cull = a.boxes != nullptr;
    // Copies of one entry in a row: the reference leaves the slots of the
    // Gaussians it counts but never writes (zero screen extent) zeroed, so they
    // all hold Gaussian 0 and sort to the front of tile 0.  Once an entry has
    // left every pixel of the wave as it was (each live pixel skipped it, or
    // it was culled), a copy of it does the same, so the windows that follow
    // and hold nothing but copies of it are passed over, eight per read (their
    // mask words written as zero).
    int dup_g = -1;
    bool dup_noop = false;
    for (int base = mentry0; base < r1; base += 64) {
        if (!any_lane(!done)) break;
        const Window w = window_candidates(a, base, r0, r1 - 1, rect, lane, cull);
        int ev_e = -1, ev_g = -1;
        bool ev_noop = false;
        uint64_t cand = w.cand;
        uint64_t mbits = 0;
        while (cand != 0) {
            if (!any_lane(!done)) break;
            const int bit = __ffsll((unsigned long long)cand) - 1;
            cand &= cand - 1;
            const int e = base + bit;
            const int g = __builtin_amdgcn_readlane(w.gid, bit);
            const Splat s = splat_record(a, g);
            bool comp = false;
            if (!done) {
                Pair p = pair_eval(s, pr, use_blur, PRECISE && a.precise_edge);
                if (viz) {
                    const float sigma = 0.5f * (s.iq1 * p.l1 * p.l1 + s.iq2 * p.l2 * p.l2);
                    p.alpha = (sigma > viz_thresh || (viz_opac && s.opac < 0.5f)) ? 0.f : 0.99f;
                    p.skip = p.t < 0.01f || p.t > 1000.0f || p.alpha < 1.f / 255.f;
                }
                const float next_T = T * (1.f - p.alpha);
                if (next_T <= 1e-4f) done = true;
                comp = !(p.skip || done);
                if (comp) {
                    const float vis = p.alpha * T;
                    // colour and texture weights: as given, or (visualisation,
                    // within the outline bound of the edge) white or nothing
                    float cr = s.rgb.x, cg = s.rgb.y, cb = s.rgb.z;
                    bool tex_on = true, tex_white = false;
                    if (viz && !(outline_distance(a, s, j, i, viz_thresh) > viz_outline)) {
                        tex_white = viz_white;
                        tex_on = viz_white;
                        cr = cg = cb = viz_white ? 1.f : 0.f;
                    }
                    rgb0 += cr * vis;
                    rgb1 += cg * vis;
                    rgb2 += cb * vis;
                    // normal facing the camera: the flip goes into the weight's
                    // sign (exact), so the axis stays a scalar operand
                    const float nvis = (flip && dot3(pr.r, s.a3) > 0.f) ? -vis : vis;
                    n0 += nvis * s.a3.x;
                    n1 += nvis * s.a3.y;
                    n2 += nvis * s.a3.z;
                    const float t_view = p.t * pr.view_depth;
                    float u, v;
                    splat_uv(s, p.delta, u, v);
                    const Chart ch = chart_of<TB>(a.texture, s, C, a.num_texels);
                    float wt[4], fu, fv;
                    int off[4];
                    chart_taps(s, ch, u, v, bilinear, wt, off, fu, fv);
                    // texels in chunks of up to 4 channels, all taps loaded before use
                    constexpr int CK = CM < 4 ? CM : 4;
#pragma unroll
                    for (int c0 = 0; c0 < CM; c0 += CK) {
                        if (c0 >= C) break;
                        float tv[4][CK];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
#pragma unroll
                            for (int cc = 0; cc < CK; ++cc)
                                tv[k][cc] = c0 + cc < C
                                                ? chart_texel<TB>(ch, off[k] + (c0 + cc) * E)
                                                : 0.f;
                        }
#pragma unroll
                        for (int cc = 0; cc < CK; ++cc) {
                            const float val = wt[0] * tv[0][cc] + wt[1] * tv[1][cc] +
                                              wt[2] * tv[2][cc] + wt[3] * tv[3][cc];
                            if (!viz) tex[c0 + cc] += vis * val;
                            else if (tex_on) tex[c0 + cc] += vis * (tex_white ? 1.f : val);
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
                    if (T * kKeep <= 1e-4f) done = true;
                }
            }
            const bool any_comp = any_lane(comp);
            if (any_comp) mbits |= 1ull << bit;
            ev_e = e;
            ev_g = g;
            ev_noop = !any_comp;
        }
        if (mplane && lane == 0) mplane[mword0 + ((base - mentry0) >> 6)] = mbits;
        {
            const int now = __popcll(lanes_where(!done));
            if (now < live) {
                live = now;
                rect = wave_rect(!done, j, i);
            }
        }
        // what a copy of the window's last entry would do: as the evaluation
        // of that entry found, or nothing if it was culled
        bool dup_run = false;
        if (w.in_range != 0) {
            const int last = 63 - __builtin_clzll(w.in_range);
            if (ev_e == base + last) {
                dup_g = ev_g;
                dup_noop = ev_noop;
            } else {
                dup_g = __builtin_amdgcn_readlane(w.gid, last);
                dup_noop = cull;
            }
            // a run of copies shows as two equal ids at the window's end
            // (distinct Gaussians otherwise), so the common case reads nothing
            dup_run = last > 0 && ((w.in_range >> (last - 1)) & 1ull) &&
                      __builtin_amdgcn_readlane(w.gid, last - 1) == dup_g;
        }
        // pass over the windows made only of copies of that entry, eight
        // windows per read
        if (dup_noop && dup_run && any_lane(!done)) {
            int nb = base + 64;
            while (nb < r1) {
                const int cnt = min(r1 - nb, 512);
                const int* ids = a.ids_sorted + nb + lane;
                unsigned same = 0;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const bool s_k = 64 * k + lane >= cnt || ids[64 * k] == dup_g;
                    if (!any_lane(!s_k)) same |= 1u << k;
                }
                const int lead = min(__builtin_ctz(~same), (cnt + 63) >> 6);
                if (mplane && lane < lead) mplane[mword0 + ((nb - mentry0) >> 6) + lane] = 0ull;
                nb += 64 * lead;
                if (lead < 8) break;
            }
            base = nb - 64;
        }
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

template <int CM, int MODE, bool TB>
__global__ __launch_bounds__(256) void k_raster_forward(RasterArgs a, FwdOut o) {
    raster_forward_wave<CM, MODE, TB>(a, o, blockIdx.x, wave_index());
}

// ------------------------------------------------------------------ multi-view queue
// Orders the (view, tile) items of a multi-view table longest list first (64
// length classes, two per octave; order inside a class is arbitrary): a
// counting sort in one workgroup.
__global__ __launch_bounds__(1024) void k_views_order(const ViewRaster* __restrict__ views,
                                                      ViewsHeader* __restrict__ hdr,
                                                      int* __restrict__ items) {
    __shared__ int hist[64];
    const int tid = threadIdx.x;
    if (tid < 64) hist[tid] = 0;
    __syncthreads();
    const int nv = hdr->num_views;
    auto cls = [&](const RasterArgs& a, int t) {
        int r0, r1;
        tile_range(a, t, r0, r1);
        const unsigned n = (unsigned)(r1 - r0) + 1u;
        const int lg = 31 - __builtin_clz(n);
        const int half = lg > 0 ? (int)((n >> (lg - 1)) & 1u) : 0;
        return 63 - (2 * lg + half);  // descending length
    };
    for (int v = 0; v < nv; ++v) {
        const RasterArgs& a = views[v].a;
        const int nt = a.tiles_x * a.tiles_y;
        for (int t = tid; t < nt; t += blockDim.x) atomicAdd(&hist[cls(a, t)], 1);
    }
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int k = 0; k < 64; ++k) {
            const int c = hist[k];
            hist[k] = run;
            run += c;
        }
        hdr->head = 0;
    }
    __syncthreads();
    for (int v = 0; v < nv; ++v) {
        const RasterArgs& a = views[v].a;
        const int nt = a.tiles_x * a.tiles_y;
        for (int t = tid; t < nt; t += blockDim.x) {
            const int pos = atomicAdd(&hist[cls(a, t)], 1);
            items[pos] = (v << kViewItemTileBits) | t;
        }
    }
}

// Takes the next work index of a queue pass (wave-uniform): lane 0 advances
// the header's head.
__device__ __forceinline__ int queue_next(ViewsHeader* hdr) {
    int k = 0;
    if ((threadIdx.x & (kWave - 1)) == 0) k = atomicAdd(&hdr->head, 1);
    return __builtin_amdgcn_readfirstlane(__shfl(k, 0));
}

// Persistent forward over every view of a table: each wave takes (view,
// tile, wave-in-tile) work from one queue ordered longest list first, so the
// longest tiles of all views start first and the rest fill the machine
// around them.  Every wave leaves once the queue is drained.
template <int CM, int MODE, bool TB>
__global__ __launch_bounds__(256) void k_raster_forward_queue(const ViewRaster* __restrict__ views,
                                                              ViewsHeader* __restrict__ hdr,
                                                              const int* __restrict__ items) {
    const int wpt = hdr->waves_per_tile, n_work = hdr->num_items * wpt;
    const int nv = hdr->num_views;
    for (;;) {
        const int k = queue_next(hdr);
        if (k >= n_work) break;
        const int it = k / wpt, wave = k - it * wpt;
        const int item = __builtin_amdgcn_readfirstlane(as_const_space(items)[it]);
        const int v = item >> kViewItemTileBits, tile = item & ((1 << kViewItemTileBits) - 1);
        if (v >= nv) continue;
        const GSTEX_CONST ViewRaster* vr = as_const_space(views) + v;
        const RasterArgs a = load_const(&vr->a);
        const FwdOut o = load_const(&vr->o);
        if (tile >= a.tiles_x * a.tiles_y) continue;
        raster_forward_wave<CM, MODE, TB>(a, o, tile, wave);
    }
}

// ------------------------------------------------------------------ launchers
template <int CM>
static hipError_t launch_fwd_cm(const RasterArgs& a, const FwdOut& o, hipStream_t s) {
    const int tiles = a.tiles_x * a.tiles_y;
    // 1: blur, NDC distortion or alpha visualisation (and the float64 edge-on
    // solve if on); 2: the plain path with the float64 edge-on solve
    const int mode = (a.settings & (kSetBlur | kSetNdc | kSetAlphaViz)) ? 1 : a.precise_edge ? 2 : 0;
    const dim3 grid(tiles), block(raster_block_threads(a.bw));
    if (a.texture_bf16) {
        if (mode == 1) k_raster_forward<CM, 1, true><<<grid, block, 0, s>>>(a, o);
        else if (mode == 2) k_raster_forward<CM, 2, true><<<grid, block, 0, s>>>(a, o);
        else k_raster_forward<CM, 0, true><<<grid, block, 0, s>>>(a, o);
    } else {
        if (mode == 1) k_raster_forward<CM, 1, false><<<grid, block, 0, s>>>(a, o);
        else if (mode == 2) k_raster_forward<CM, 2, false><<<grid, block, 0, s>>>(a, o);
        else k_raster_forward<CM, 0, false><<<grid, block, 0, s>>>(a, o);
    }
    return hipGetLastError();
}

template <int CM>
static hipError_t launch_fwd_queue_cm(const ViewRaster* views, ViewsHeader* hdr, const int* items,
                                      int mode, bool bf16, int blocks, hipStream_t s) {
    const dim3 grid(blocks), block(256);
    if (bf16) {
        if (mode == 1) k_raster_forward_queue<CM, 1, true><<<grid, block, 0, s>>>(views, hdr, items);
        else if (mode == 2) k_raster_forward_queue<CM, 2, true><<<grid, block, 0, s>>>(views, hdr, items);
        else k_raster_forward_queue<CM, 0, true><<<grid, block, 0, s>>>(views, hdr, items);
    } else {
        if (mode == 1) k_raster_forward_queue<CM, 1, false><<<grid, block, 0, s>>>(views, hdr, items);
        else if (mode == 2) k_raster_forward_queue<CM, 2, false><<<grid, block, 0, s>>>(views, hdr, items);
        else k_raster_forward_queue<CM, 0, false><<<grid, block, 0, s>>>(views, hdr, items);
    }
    return hipGetLastError();
}

hipError_t views_order(const ViewRaster* views, ViewsHeader* hdr, int* items, hipStream_t s) {
    k_views_order<<<1, 1024, 0, s>>>(views, hdr, items);
    return hipGetLastError();
}

hipError_t raster_forward_views(const ViewRaster* views, ViewsHeader* hdr, const int* items,
                                int channels, int settings, bool bf16, int blocks,
                                hipStream_t s) {
    const int mode = (settings & (kSetBlur | kSetNdc | kSetAlphaViz)) ? 1
                     : raster_precise_edge_enabled()                 ? 2
                                                                     : 0;
    if (channels == 1) return launch_fwd_queue_cm<1>(views, hdr, items, mode, bf16, blocks, s);
    if (channels == 2) return launch_fwd_queue_cm<2>(views, hdr, items, mode, bf16, blocks, s);
    if (channels == 3) return launch_fwd_queue_cm<3>(views, hdr, items, mode, bf16, blocks, s);
    if (channels == 4) return launch_fwd_queue_cm<4>(views, hdr, items, mode, bf16, blocks, s);
    if (channels <= 8) return launch_fwd_queue_cm<8>(views, hdr, items, mode, bf16, blocks, s);
    if (channels <= 16) return launch_fwd_queue_cm<16>(views, hdr, items, mode, bf16, blocks, s);
    return launch_fwd_queue_cm<64>(views, hdr, items, mode, bf16, blocks, s);
}

static bool g_raster_culling = true;
bool raster_culling_enabled() { return g_raster_culling; }
bool set_raster_culling(bool on) {
    const bool prev = g_raster_culling;
    g_raster_culling = on;
    return prev;
}

static bool g_precise_edge = false;
bool raster_precise_edge_enabled() { return g_precise_edge; }
bool set_raster_precise_edge(bool on) {
    const bool prev = g_precise_edge;
    g_precise_edge = on;
    return prev;
}

size_t raster_forward_workspace_bytes(int n) {
    return splat_records_bytes(n) + 256;
}

hipError_t raster_forward_stage(const RasterCall& c, const float* records, float* out_img,
                                float* out_depth, float* out_reg, float* out_texture,
                                float* out_normal, float* final_Ts, int* final_idx,
                                int* depth_idx, float* out_reg_s, hipStream_t s) {
    const RasterArgs a = make_raster_args(c, records);
    const FwdOut o{out_img, out_depth, out_reg, out_texture, out_normal,
                   final_Ts, final_idx, depth_idx, out_reg_s};
    if (a.tiles_x * a.tiles_y <= 0) return hipSuccess;
    if (c.channels == 1) return launch_fwd_cm<1>(a, o, s);
    if (c.channels == 2) return launch_fwd_cm<2>(a, o, s);
    if (c.channels == 3) return launch_fwd_cm<3>(a, o, s);
    if (c.channels == 4) return launch_fwd_cm<4>(a, o, s);
    if (c.channels <= 8) return launch_fwd_cm<8>(a, o, s);
    if (c.channels <= 16) return launch_fwd_cm<16>(a, o, s);
    return launch_fwd_cm<64>(a, o, s);
}

hipError_t raster_forward(const RasterCall& c, float* out_img, float* out_depth, float* out_reg,
                          float* out_texture, float* out_normal, float* final_Ts, int* final_idx,
                          int* depth_idx, float* out_reg_s, void* ws, size_t ws_bytes,
                          hipStream_t s) {
    if (ws_bytes < raster_forward_workspace_bytes(c.num_points)) return hipErrorInvalidValue;
    float* records = (float*)ws;
    hipError_t e = build_raster_records(c, records, s);
    if (e != hipSuccess) return e;
    return raster_forward_stage(c, records, out_img, out_depth, out_reg, out_texture, out_normal,
                                final_Ts, final_idx, depth_idx, out_reg_s, s);
}

}  // namespace gstex
