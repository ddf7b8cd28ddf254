// Texture edit (texture_edit.cu:11-227 semantics): splat a per-pixel edit
// canvas back onto the Gaussians' texel charts.
//
// For every pair the forward would composite (same alpha, skip and stop
// rules; blur from settings bit 0 as in the reference edit kernel) whose view
// depth lies in the pixel's [depth_lower, depth_upper], the bilinear weights
// of the pair's texel quad (always bilinear, replicate padding) times
// (r*a, g*a, b*a, a, 1) are added to channels 0..4 of the output, a being the
// pixel's edit alpha.  The walk is the forward's: quadrant waves, 64-entry
// windows with screen-box culling, SMEM records.  The values reach the output
// by the texel routes of the raster backward: wave reduce-scatters for one
// texel, a shared texel quad or a 3x3 texel window, else DPP-merged lanes into
// an LDS accumulator or direct float atomics.
#include "raster_common.hpp"

namespace gstex {

struct EditIn {
    const float* upd_img;      // (H, W, 3)
    const float* upd_alpha;    // (H, W)
    const float* depth_lower;  // (H, W)
    const float* depth_upper;  // (H, W)
    float* out;                // (X, C), C >= 5
};

template <bool BLUR>
__global__ __launch_bounds__(256) void k_texture_edit(RasterArgs a, EditIn ed) {
    const int tile = blockIdx.x;
    const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    const int lane = threadIdx.x & (kWave - 1);
    int lx, ly;
    bool in_tile;
    lane_pixel(a.bw, threadIdx.x, lx, ly, in_tile);
    const int j = tx * a.bw + lx, i = ty * a.bw + ly;
    const bool inside = in_tile && i < a.img_h && j < a.img_w;
    const PixelRay pr = pixel_ray(a.c2w, a.viewmat, a.fx, a.fy, a.cx, a.cy, j, i);
    const int C = a.channels;
    float val[5] = {0.f, 0.f, 0.f, 0.f, 1.f};
    float lo = 0.f, hi = -1.f;
    if (inside) {
        const int pid = i * a.img_w + j;
        const float al = ed.upd_alpha[pid];
        val[0] = ed.upd_img[3 * pid] * al;
        val[1] = ed.upd_img[3 * pid + 1] * al;
        val[2] = ed.upd_img[3 * pid + 2] * al;
        val[3] = al;
        lo = ed.depth_lower[pid];
        hi = ed.depth_upper[pid];
    }
    int r0, r1;
    tile_range(a, tile, r0, r1);
    bool done = !inside;
    float T = 1.f;
    // culling rectangle of the live pixels (shrinks as pixels finish, as in
    // k_raster_forward)
    WaveRect rect = wave_rect(inside, j, i);
    int live = __popcll(__ballot(inside));
    // wave-private texel accumulator, kept zero between Gaussians by the flush
    __shared__ float tex_acc[256 / kWave][kTexAccFloats];
    float* acc = tex_acc[wave_index()];
    for (int k = lane; k < kTexAccFloats; k += kWave) acc[k] = 0.f;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    const float inv_c4 = 1.f / (float)(4 * C);  // chart byte offset -> texel index
    // a pixel with T (1 - 1/255) <= 1e-4 can take no further entry (each one
    // with alpha >= 1/255 ends it), so it stops at once and culled entries
    // never matter (as in k_raster_forward)
    constexpr float kKeep = 1.f - 1.f / 255.f;
    const bool cull = a.boxes != nullptr;
    for (int base = r0 & ~63; base < r1; base += 64) {
        if (!__any(!done)) break;
        {
            const int now = __popcll(__ballot(!done));
            if (now < live) {
                live = now;
                rect = wave_rect(!done, j, i);
            }
        }
        const Window w = window_candidates(a, base, r0, r1 - 1, rect, lane, cull);
        uint64_t cand = w.cand;
        while (cand != 0) {
            if (!__any(!done)) break;
            const int bit = __ffsll((unsigned long long)cand) - 1;
            cand &= cand - 1;
            const int g = __builtin_amdgcn_readlane(w.gid, bit);
            const Splat s = splat_record(a, g);
            bool hit = false;
            float wt[4] = {0.f, 0.f, 0.f, 0.f};
            int row[4] = {0, 0, 0, 0};  // byte offsets inside the chart (float values)
            float fu = 0.f, fv = 0.f;
            int iu = 0, iv = 0, ju = 0, jv = 0;  // tap rows (u) and columns (v)
            // out has the texture's (X, C) float layout; charts outside it get nothing
            const Chart ch = chart_of<false>(ed.out, s, C, a.num_texels);
            if (!done) {
                const Pair p = pair_eval(s, pr, BLUR, a.precise_edge);
                const float next_T = T * (1.f - p.alpha);
                if (next_T <= 1e-4f) done = true;
                if (!(p.skip || done)) {
                    const float t_view = p.t * pr.view_depth;
                    if (t_view >= lo && t_view <= hi) {
                        hit = true;
                        float u, v;
                        splat_uv(s, p.delta, u, v);
                        chart_taps_ij(s, ch, u, v, true, wt, row, fu, fv, iu, iv, ju, jv);
                    }
                    T = next_T;
                    if (T * kKeep <= 1e-4f) done = true;
                }
            }
            if (!__any(hit) || !ch.ok) continue;
            float* const out_chart = ed.out + (size_t)s.start * C;
            // How the (taps x 5) values reach the output (wave-uniform choice,
            // as for the texel gradients of the raster backward):
            //   single  1x1 chart: 5 values in a wave reduce-scatter
            //   quad    every hitting lane samples the same texel quad: 20 values
            //   window  every hitting lane's taps lie in one 3x3 texel window:
            //           the 9 separable weight products x 5 values
            //   lds     charts of up to kTexAccFloats / 5 texels: merged lanes
            //           add into the wave's LDS accumulator, then flush
            //   direct  one atomic per tap and value
            // Lanes that do not hit have zero weights, so they add nothing.
            const int first = __ffsll((unsigned long long)__ballot(hit)) - 1;
            if (s.h * s.w == 1) {
                const float ws = (wt[0] + wt[1]) + (wt[2] + wt[3]);
                float v5[5];
#pragma unroll
                for (int k = 0; k < 5; ++k) v5[k] = ws * val[k];
                const float tot = wave_reduce_scatter(v5, lane);
                const int code = rs_slot<5>(lane);
                if (code < 64 && tot != 0.f) unsafeAtomicAdd(out_chart + code, tot);
                continue;
            }
            const int row0 = __builtin_amdgcn_readlane(row[0], first);
            if (!__any(hit && row[0] != row0)) {
                const int row1 = __builtin_amdgcn_readlane(row[1], first);
                const int row2 = __builtin_amdgcn_readlane(row[2], first);
                const int row3 = __builtin_amdgcn_readlane(row[3], first);
                float v20[20];
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int k = 0; k < 5; ++k) v20[5 * t + k] = wt[t] * val[k];
                const float tot = wave_reduce_scatter(v20, lane);
                const int code = rs_slot<20>(lane);
                if (code < 64 && tot != 0.f) {
                    const int t = code / 5, k = code - 5 * t;
                    const int rr = t == 0 ? row0 : t == 1 ? row1 : t == 2 ? row2 : row3;
                    unsafeAtomicAdd(out_chart + (rr >> 2) + k, tot);
                }
                continue;
            }
            {
                // window origin: the first hitting lane's tap row / column, one
                // lower when some hitting lane lies below it
                const int au = __builtin_amdgcn_readlane(iu, first);
                const int av = __builtin_amdgcn_readlane(iv, first);
                const int ou = au - (__any(hit && iu < au) ? 1 : 0);
                const int ov = av - (__any(hit && iv < av) ? 1 : 0);
                if (!__any(hit && (iu < ou || iu > ou + 1 || iv < ov || iv > ov + 1))) {
                    // tap (row r, column q) has weight A[r] * B[q]; rows iu, ju and
                    // columns iv, jv relative to the origin lie in 0..2
                    const float h1 = hit ? 1.f : 0.f;
                    const float wa0 = h1 * (1.f - fu), wa1 = h1 * fu, wb0 = 1.f - fv, wb1 = fv;
                    const int ri = iu - ou, rj = ju - ou, ci = iv - ov, cj = jv - ov;
                    float A[3], Bw[3];
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        A[r] = (ri == r ? wa0 : 0.f) + (rj == r ? wa1 : 0.f);
                        Bw[r] = (ci == r ? wb0 : 0.f) + (cj == r ? wb1 : 0.f);
                    }
                    float v45[45];
#pragma unroll
                    for (int r = 0; r < 3; ++r)
#pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            const float ab = A[r] * Bw[q];
#pragma unroll
                            for (int k = 0; k < 5; ++k) v45[(3 * r + q) * 5 + k] = ab * val[k];
                        }
                    const float tot = wave_reduce_scatter(v45, lane);
                    const int code = rs_slot<45>(lane);
                    if (code < 64 && tot != 0.f) {
                        const int rq = code / 5, k = code - 5 * rq;
                        const int r = rq / 3, q = rq - 3 * r;
                        // a window row or column past the chart edge has no weight
                        if (ou + r < s.h && ov + q < s.w)
                            unsafeAtomicAdd(out_chart + ((size_t)(ou + r) * s.w + (ov + q)) * C + k,
                                            tot);
                    }
                    continue;
                }
            }
            const QuadMerge qm = quad_merge_setup(hit ? row[0] : -1, lane);
            if (s.h * s.w * 5 <= kTexAccFloats) {
                // charts of up to kTexAccFloats / 5 texels: merged lanes add
                // into the wave's LDS accumulator (5 values per texel), then
                // the texels some lane touched go out with one atomic each
                int tix[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) tix[t] = (int)((float)row[t] * inv_c4 + 0.5f);
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    float q[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) q[t] = quad_merge(wt[t] * val[k], qm);
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        if (q[t] != 0.f) atomicAdd(acc + tix[t] * 5 + k, q[t]);
                }
                int tlo = 0x7fffffff, thi = -1;
                if (hit) {
                    tlo = min(min(tix[0], tix[1]), min(tix[2], tix[3]));
                    thi = max(max(tix[0], tix[1]), max(tix[2], tix[3]));
                }
                tlo = wave_min(tlo);
                thi = wave_max(thi);
                // every lane's LDS adds precede these reads in the wave's LDS queue
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                const int nv = (thi - tlo + 1) * 5;
                for (int idx = lane; idx < nv; idx += kWave) {
                    const int tx = tlo + idx / 5, k = idx - (idx / 5) * 5;
                    const float v = acc[tx * 5 + k];
                    if (v != 0.f) {
                        unsafeAtomicAdd(out_chart + (size_t)tx * C + k, v);
                        acc[tx * 5 + k] = 0.f;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            } else {
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    float q[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) q[t] = quad_merge(wt[t] * val[k], qm);
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        if (q[t] != 0.f) unsafeAtomicAdd(out_chart + (row[t] >> 2) + k, q[t]);
                }
            }
        }
    }
}

size_t texture_edit_workspace_bytes(int n) { return splat_records_bytes(n) + 256; }

hipError_t texture_edit(const RasterCall& c, const float* upd_img, const float* upd_alpha,
                        const float* depth_lower, const float* depth_upper, float* out,
                        void* ws, size_t ws_bytes, hipStream_t s) {
    if (ws_bytes < texture_edit_workspace_bytes(c.num_points)) return hipErrorInvalidValue;
    hipError_t e = fill_f32(out, (size_t)c.num_texels * (size_t)c.channels, 0.f, s);
    if (e != hipSuccess) return e;
    float* records = (float*)ws;
    if ((e = build_raster_records(c, records, s)) != hipSuccess) return e;
    RasterArgs a = make_raster_args(c, records);
    if (c.settings & 1) {  // the edit kernel's blur flag: the box covers the blur disc
        a.ellipses = nullptr;
        a.ellipse_c = nullptr;
    }
    const int tiles = a.tiles_x * a.tiles_y;
    if (tiles <= 0) return hipSuccess;
    const EditIn ed{upd_img, upd_alpha, depth_lower, depth_upper, out};
    if (c.settings & 1)
        k_texture_edit<true><<<tiles, raster_block_threads(a.bw), 0, s>>>(a, ed);
    else
        k_texture_edit<false><<<tiles, raster_block_threads(a.bw), 0, s>>>(a, ed);
    return hipGetLastError();
}

}  // namespace gstex
