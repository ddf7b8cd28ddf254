// Back-to-front adjoint of the tile rasteriser (texture.cu:331-760 semantics)
// and the per-Gaussian gradient finalize.
//
// Per (wave, Gaussian) the lanes' contributions are summed with butterfly
// reduce-scatters built from permlane swaps and DPP lane permutes (all VALU,
// no LDS traffic), after which a lane holds the wave total of one value and
// issues its atomic: one over the 20 per-Gaussian scalars from which the
// finalize kernel rebuilds every parameter gradient, one over the texel
// values of the chart's route.  Using delta = l1*ax1 + l2*ax2 (the ray hits
// the splat plane), each outer product sum_pixels v * delta becomes
// ax1 * sum(v*l1) + ax2 * sum(v*l2).
#include "raster_common.hpp"
#include "sh_basis.hpp"

namespace gstex {

// ------------------------------------------------------------------ backward
// MODE 0 compiles out the blur (settings bit 9), NDC distortion (bit 10)
// and float64 edge-on code for the common settings; MODE 1 keeps all of it;
// MODE 2 is MODE 0 plus the float64 edge-on solve.  TB selects bfloat16
// texels.  CM <= 4 is the exact channel count; above 4, CM bounds a runtime
// count.  The plain kernels for CM <= 4 are held to 96 VGPRs (5 waves per
// SIMD), which they need at most one register more than.
//
// Texel gradients of the window route over the R x Q texels from chart row
// ou, column ov (R, Q <= 3): window tap (row r, column q) has weight
// A[r] * B[q], and rows and columns from R, Q on carry zero weight on every
// valid lane.  One reduction of the R * Q * CM values; the lane holding a
// wave total adds it to its texel.
template <int R, int Q, int CM>
__device__ __forceinline__ void window_texel_add(const float (&A)[3], const float (&B)[3],
                                                 const float (&vval)[CM], int lane,
                                                 float* vtex_chart, int ou, int ov, int h,
                                                 int w) {
    float vals[R * Q * CM];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const float ab = A[r] * B[q];
#pragma unroll
            for (int c = 0; c < CM; ++c) vals[(Q * r + q) * CM + c] = ab * vval[c];
        }
    const float tot = wave_reduce_scatter(vals, lane);
    const int code = rs_slot<R * Q * CM>(lane);
    if (code < 64 && tot != 0.f) {
        const int rq = code / CM, c = code - CM * rq;
        const int r = rq / Q, q = rq - Q * r;
        // a window row or column past the chart edge carries no weight
        // (taps clamp inside the chart)
        if (ou + r < h && ov + q < w)
            unsafeAtomicAdd(vtex_chart + ((ou + r) * w + (ov + q)) * CM + c, tot);
    }
}

// Window weights along one axis: weight w0 at window index i and w1 at j,
// summed where they meet, i.e. W[r] = (i == r ? w0 : 0) + (j == r ? w1 : 0).
// On a lane whose taps lie in the window, i is 0 or 1 and j is i or i + 1
// (the second tap clamps onto the first at the chart edge), which leaves the
// four cases below; other lanes get finite weights of no meaning (their
// texel terms are zero).
__device__ __forceinline__ void window_weights(float w0, float w1, int i, int j, float (&W)[3]) {
    const bool same = j == i;
    const float first = same ? w0 + w1 : w0, second = same ? 0.f : w1;
    const bool at0 = i == 0;
    W[0] = at0 ? first : 0.f;
    W[1] = at0 ? second : first;
    W[2] = at0 ? 0.f : second;
}

// raster_backward_wave runs one wave's share of a tile (wave `wave` of the
// tile, as in the forward); `acc` is this hardware wave's LDS texel
// accumulator, zero on entry and left zero.
template <int CM, int MODE, bool TB>
__device__ __forceinline__ void raster_backward_wave(const RasterArgs& a, const BwdIn& b,
                                                     const int tile, const int wave,
                                                     float* const acc) {
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
    const bool bilinear = !(a.settings & kSetNearest), prop_uv = a.settings & kSetUvGrad;
    const int C = CM <= 4 ? CM : a.channels;
    constexpr int E = TB ? 2 : 4;  // bytes per texel value
    constexpr int kES = TB ? 1 : 2;  // chart byte offset >> kES = value index in the chart
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
#pragma unroll
        for (int c = 0; c < CM; ++c)
            if (c < C) vtex[c] = b.v_texture_out[(size_t)pid * C + c];
        if constexpr (CM == 3) {
            if (b.loss_target) {  // vo, vtex hold out_img, out_texture
                const float* tg = b.loss_target + 3 * pid;
                const float g0 = l2_cotangent(vo.x, vtex[0], tg[0], b.loss_scale);
                const float g1 = l2_cotangent(vo.y, vtex[1], tg[1], b.loss_scale);
                const float g2 = l2_cotangent(vo.z, vtex[2], tg[2], b.loss_scale);
                vo = v3(g0, g1, g2);
                vtex[0] = g0;
                vtex[1] = g1;
                vtex[2] = g2;
            }
        }
        vT = a.background[0] * vo.x + a.background[1] * vo.y + a.background[2] * vo.z -
             b.v_alpha[pid];
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
    const int last_e = inside ? bfin : -1;  // entries up to it may contribute

    // One list entry (wave-uniform index e, Gaussian g with record s).
    auto entry = [&](const int e, const int g, const Splat& s) {
        // The pair math and the taps run for every lane: they are finite for
        // any pixel ray, and tap offsets stay inside the chart.  A lane that
        // does not contribute keeps T and vT and gets zero seeds (vis, valpha,
        // vt, vu, vv), so every gradient value below is zero for it without a
        // branch or per-value selects.
        const Pair p = pair_eval(s, pr, use_blur, PRECISE && a.precise_edge);
        const bool valid = e <= last_e && !p.skip;
        const uint64_t vmask = lanes_where(e <= last_e) & ~skip_lanes(p);
        if (vmask == 0) return;

        // Transmittance before this Gaussian (1 - alpha >= 0.01) and the texel taps.
        const Chart ch = chart_of<TB>(a.texture, s, C, a.num_texels);
        // (a reciprocal and a product: about 1.5 ulp per entry, 3e-6 relative
        // after a thousand)
        if (valid) T = T * __builtin_amdgcn_rcpf(1.f - p.alpha);
        const float vis = valid ? p.alpha * T : 0.f;
        float u, v, wt[4], fu, fv;
        int row[4];  // byte offsets inside the chart
        int iu, iv, ju, jv;  // tap rows (u) and columns (v) of the chart
        splat_uv(s, p.delta, u, v);
        chart_taps_ij(s, ch, u, v, bilinear, wt, row, fu, fv, iu, iv, ju, jv);
        // How the texel gradients reach HBM (wave-uniform choice):
        //   single  1x1 chart, C <= 4: C values in the reduction
        //   quad    every valid lane samples the same texel quad, C <= 3:
        //           4 taps x C values in the reduction
        //   window  bilinear, C <= 4, every valid lane's taps inside one 3x3
        //           texel window: the 9 C separable weight products in the
        //           reduction
        //   lds     chart of <= kTexAccFloats values: LDS accumulation + flush
        //   direct  one atomic per tap
        const int first = __ffsll((unsigned long long)vmask) - 1;
        const int hw = ch.ok ? s.h * s.w : 0;
        const bool tex_single = hw == 1 && C >= 1 && C <= 4;
        const int row0 = __builtin_amdgcn_readlane(row[0], first);
        const bool quad32 = ch.ok && !tex_single && C >= 1 && C <= 3 &&
                            !(lanes_where(row[0] != row0) & vmask);
        // window origin (ou, ov): the anchor lane's tap row / column, one
        // lower when some valid lane lies below it (formed only where the
        // window route is open)
        int ou = 0, ov = 0;
        bool win3 = false;
        if (CM <= 4 && ch.ok && !tex_single && !quad32 && bilinear && C >= 1) {
            const int au = __builtin_amdgcn_readlane(iu, first);
            const int av = __builtin_amdgcn_readlane(iv, first);
            ou = au - ((lanes_where(iu < au) & vmask) ? 1 : 0);
            ov = av - ((lanes_where(iv < av) & vmask) ? 1 : 0);
            win3 = !(lanes_where(((uint32_t)(iu - ou) | (uint32_t)(iv - ov)) > 1u) & vmask);
        }
        const bool tex_lds = ch.ok && !tex_single && !quad32 && !win3 && C > 0 &&
                             hw * C <= kTexAccFloats;
        float* const vtex_chart = b.v_texture + (size_t)(ch.ok ? s.start : 0) * C;

        // The per-Gaussian values (GradSlot 0..19, then the two blur values
        // in the general kernels) and the texel terms of the routes.
        constexpr int NB = G_EXTRA + (GENERAL ? 2 : 0);
        float gr[NB];
        float vval[CM];  // vis * dL/dtexture_out per channel
        {
            gr[G_RGB + 0] = vis * vo.x;
            gr[G_RGB + 1] = vis * vo.y;
            gr[G_RGB + 2] = vis * vo.z;
            float vvis = dot3(s.rgb, vo) + dot3(s.a3, vn);
            float vu = 0.f, vv = 0.f;
            // D[k] = sum over channels of vtex[c] * (tap k's texel, channel c).
            // The sampled value's share of vvis is sum_k wt[k] D[k], and
            // d val / d(u, v) through the bilinear weights
            // (texture_helpers.cuh:262-299) takes tap differences, which are
            // differences of two D: coincident taps (equal texels) give equal
            // D and exactly zero.
            float D[4] = {0.f, 0.f, 0.f, 0.f};
            // texels in chunks of up to 4 channels, all taps loaded before use
            constexpr int CK = CM < 4 ? CM : 4;
#pragma unroll
            for (int cb = 0; cb < CM; cb += CK) {
                if (cb >= C) break;
                float tv[4][CK];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
#pragma unroll
                    for (int cc = 0; cc < CK; ++cc)
                        tv[k][cc] = cb + cc < C ? chart_texel<TB>(ch, row[k] + (cb + cc) * E)
                                                : 0.f;
                }
#pragma unroll
                for (int cc = 0; cc < CK; ++cc) {
                    const int c = cb + cc;
                    if (c >= C) break;
#pragma unroll
                    for (int k = 0; k < 4; ++k) D[k] = fmaf(tv[k][cc], vtex[c], D[k]);
                    vval[c] = vis * vtex[c];
                    if constexpr (CM > 4) {
                        if (tex_lds) {
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                const float add = wt[k] * vval[c];
                                if (add != 0.f) atomicAdd(acc + (row[k] >> kES) + c, add);
                            }
                        }
                    }
                    if (!tex_single && !quad32 && !win3 && !tex_lds && ch.ok) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const float add = wt[k] * vval[c];
                            if (add != 0.f) unsafeAtomicAdd(vtex_chart + (row[k] >> kES) + c, add);
                        }
                    }
                }
            }
            vvis += (wt[0] * D[0] + wt[1] * D[1]) + (wt[2] * D[2] + wt[3] * D[3]);
            if (bilinear && prop_uv) {
                vu = vis * (float)s.h * ((1.f - fv) * (D[2] - D[0]) + fv * (D[3] - D[1]));
                vv = vis * (float)s.w * ((1.f - fu) * (D[1] - D[0]) + fu * (D[3] - D[2]));
            }

            if (!valid) vu = vv = 0.f;  // NaN-safe zero seeds
            const float dvis = vvis - vT;
            float valpha = T * dvis;
            float vTcur = fmaf(p.alpha, dvis, vT);  // alpha vvis + (1 - alpha) vT
            const float t_view = p.t * pr.view_depth;
            const float dval = use_ndc ? (tf * t_view - tf * tn) / ((tf - tn) * t_view) : p.t;
            const float vd_val = 2.f * (vis * dval * S0 - vis * S1) * vreg;
            const float vw = (dval * dval * S0 - 2.f * dval * S1 + S2) * vreg;
            valpha += vw * T;
            vTcur += vw * p.alpha;
            if (valid) vT = vTcur;
            valpha = valid ? valpha : 0.f;

            float vt = use_ndc ? 0.f : vd_val;
            float vtview = (e == dfin && dfin != -1) ? vdep : 0.f;
            if (use_ndc) vtview += (tf * tn) / ((tf - tn) * t_view * t_view) * vd_val;
            vt += pr.view_depth * vtview;
            vt = valid ? vt : 0.f;

            gr[G_OPAC] = (p.blur_sel ? p.Gb : p.G) * valpha;
            const float vsigma = p.blur_sel ? 0.f : -s.opac * p.G * valpha;
            const float vl1 = vsigma * s.iq1 * p.l1;
            const float vl2 = vsigma * s.iq2 * p.l2;
            const V3 vdelta = v3(vl1 * s.a1.x + vl2 * s.a2.x + s.umap.x * vu + s.vmap.x * vv,
                                 vl1 * s.a1.y + vl2 * s.a2.y + s.umap.y * vu + s.vmap.y * vv,
                                 vl1 * s.a1.z + vl2 * s.a2.z + s.umap.z * vu + s.vmap.z * vv);
            vt += dot3(vdelta, pr.r);
            const float vnum = vt * __builtin_amdgcn_rcpf(p.den);
            const float vden = -vnum * p.t;
            gr[G_P11] = vl1 * p.l1;
            gr[G_P12] = vl1 * p.l2;
            gr[G_P21] = vl2 * p.l1;
            gr[G_P22] = vl2 * p.l2;
            gr[G_U1] = vu * p.l1;
            gr[G_U2] = vu * p.l2;
            gr[G_V1] = vv * p.l1;
            gr[G_V2] = vv * p.l2;
            gr[G_VL1] = vl1;
            gr[G_VL2] = vl2;
            gr[G_VU] = vu;
            gr[G_VV] = vv;
            gr[G_VNUM] = vnum;
            gr[G_AX3X + 0] = vden * pr.r.x + vis * vn.x;
            gr[G_AX3X + 1] = vden * pr.r.y + vis * vn.y;
            gr[G_AX3X + 2] = vden * pr.r.z + vis * vn.z;
            if constexpr (GENERAL) {
                float px = 0.f, py = 0.f;
                if (use_blur) {
                    const float vsb = p.blur_sel ? -s.opac * p.Gb * valpha : 0.f;
                    px = 2.0f * vsb * (s.pix_x - pr.px);
                    py = 2.0f * vsb * (s.pix_y - pr.py);
                }
                gr[G_EXTRA] = px;
                gr[G_EXTRA + 1] = py;
            }
        }

        // Reduce and add: lane L holds the wave total of value slot(L), and
        // the lowest lane holding a value adds it.  The per-Gaussian values
        // and the route's texel values are reduced separately (about the same
        // VALU as one reduction of both, with half the live registers).
        {
            float* const grec = b.grad_rec + (size_t)g * kGradFloats;
            const float tot = wave_reduce_scatter(gr, lane);
            const int code = rs_slot<NB>(lane);
            if (code < 64 && tot != 0.f)
                unsafeAtomicAdd(grec + (code < G_EXTRA ? code : G_PIX + (code - G_EXTRA)), tot);
        }
        if constexpr (CM <= 4) {
            if (tex_single) {
                float vals[CM];
                const float wsum = (wt[0] + wt[1]) + (wt[2] + wt[3]);
#pragma unroll
                for (int c = 0; c < CM; ++c) vals[c] = wsum * vval[c];
                const float tot = wave_reduce_scatter(vals, lane);
                const int code = rs_slot<CM>(lane);
                if (code < 64 && tot != 0.f) unsafeAtomicAdd(vtex_chart + code, tot);
                return;
            }
            if (quad32) {
                constexpr int QC = CM <= 3 ? CM : 3;
                float vals[4 * QC];
#pragma unroll
                for (int k = 0; k < 4; ++k)
#pragma unroll
                    for (int c = 0; c < QC; ++c) vals[QC * k + c] = wt[k] * vval[c];
                const float tot = wave_reduce_scatter(vals, lane);
                const int code = rs_slot<4 * QC>(lane);
                // the anchor lane's other tap rows (every valid lane's quad),
                // read here while every lane is active
                const int row1 = __builtin_amdgcn_readlane(row[1], first);
                const int row2 = __builtin_amdgcn_readlane(row[2], first);
                const int row3 = __builtin_amdgcn_readlane(row[3], first);
                if (code < 64 && tot != 0.f) {
                    const int k = code / QC, c = code - QC * k;
                    const int rr = k == 0 ? row0 : k == 1 ? row1 : k == 2 ? row2 : row3;
                    unsafeAtomicAdd(vtex_chart + (rr >> kES) + c, tot);
                }
                return;
            }
            if (win3) {
                // separable weights: tap (row r, column q) has weight
                // wa[r] * wb[q]; rows iu, ju and columns iv, jv relative to
                // the window origin lie in 0..2 (ju = iu at the chart edge)
                const int ri = iu - ou, rj = ju - ou, ci = iv - ov, cj = jv - ov;
                float A[3], Bw[3];
                window_weights(1.f - fu, fu, ri, rj, A);
                window_weights(1.f - fv, fv, ci, cj, Bw);
                // when no valid lane's taps reach window row (column) 2, its
                // weights are zero on every lane and the window is 2 x 3 (3 x 2);
                // not for 4 channels, whose kernel has no registers to spare
                if constexpr (CM <= 3) {
                    if (!(lanes_where(rj == 2) & vmask)) {
                        window_texel_add<2, 3, CM>(A, Bw, vval, lane, vtex_chart, ou, ov, s.h,
                                                   s.w);
                        return;
                    }
                    if (!(lanes_where(cj == 2) & vmask)) {
                        window_texel_add<3, 2, CM>(A, Bw, vval, lane, vtex_chart, ou, ov, s.h,
                                                   s.w);
                        return;
                    }
                }
                window_texel_add<3, 3, CM>(A, Bw, vval, lane, vtex_chart, ou, ov, s.h, s.w);
                return;
            }
            if (tex_lds) {
                // merge lanes on the same texel quad before the LDS adds
                const QuadMerge qm = quad_merge_setup(valid ? row[0] : -1, lane);
                int lofs[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) lofs[k] = row[k] >> kES;
#pragma unroll
                for (int c = 0; c < CM; ++c) {
                    if (c >= C) break;
                    float q[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) q[k] = quad_merge(wt[k] * vval[c], qm);
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        if (q[k] != 0.f) atomicAdd(acc + lofs[k] + c, q[k]);
                }
            }
        }
        if (tex_lds) {
            // Flush only the texel rows some valid lane touched.
            int rlo = 0x7fffffff, rhi = -1;
            if (valid) {
                rlo = min(min(row[0], row[1]), min(row[2], row[3]));
                rhi = max(max(row[0], row[1]), max(row[2], row[3]));
            }
            rlo = wave_min(rlo);
            rhi = wave_max(rhi);
            // every lane's LDS adds precede these reads in the wave's LDS queue
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            const int n_acc = (rhi >> kES) + C;  // past the last value of the last touched texel
            for (int k = (rlo >> kES) + lane; k < n_acc; k += kWave) {
                const float v = acc[k];
                if (v != 0.f) {
                    unsafeAtomicAdd(vtex_chart + k, v);
                    acc[k] = 0.f;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        }
    };

    // Entry order: back to front over [r0, e_hi], in 64-entry windows aligned
    // to multiples of 64 (one coalesced id read per window).  With wave masks
    // only the entries the forward composited for this wave are visited;
    // without them, the entries whose screen box meets the wave's pixels (a
    // culled entry has alpha < 1/255 at all of them, so it adds nothing).
    const int mword0 = (r0 >> 6) + tile, mentry0 = r0 & ~63;
    const WaveRect rect = wave_rect(inside, j, i);
    if (a.wave_masks != nullptr) {
        // The mask words of 64 windows are read at once, lane k holding window
        // top - k, so windows this wave never composited (long runs of them in
        // lists the forward culled or passed over) cost no read of their ids.
        const size_t plane = (size_t)wave * a.mask_plane + mword0;
        const uint64_t* mplane = a.wave_masks + plane;
        const GSTEX_CONST uint64_t* mplane_s = as_const_space(a.wave_masks) + plane;
        const int top = e_hi >= r0 ? ((e_hi & ~63) - mentry0) >> 6 : -1;
        for (int wt = top; wt >= 0; wt -= 64) {
            const int wi = wt - lane;
            const int wb = mentry0 + (wi << 6);
            const int lo = max(r0 - wb, 0), hi = min(e_hi - wb, 63);
            bool any = false;
            if (wi >= 0 && hi >= lo) any = (mplane[wi] & ((~0ull >> (63 - hi)) & (~0ull << lo))) != 0;
            uint64_t nz = lanes_where(any);
            while (nz != 0) {
                const int k = __ffsll((unsigned long long)nz) - 1;
                nz &= nz - 1;
                // the word again, through the scalar path: no vector register
                // holds it while the entries are processed
                const int base = mentry0 + ((wt - k) << 6);
                const Window w = window_candidates(a, base, r0, e_hi, rect, lane, false);
                uint64_t m = mplane_s[wt - k] & w.in_range;
                while (m != 0) {
                    const int bit = 63 - __builtin_clzll(m);
                    m &= ~(1ull << bit);
                    const int g = __builtin_amdgcn_readlane(w.gid, bit);
                    entry(base + bit, g, splat_record(a, g));
                }
            }
        }
        return;
    }
    for (int base = e_hi >= r0 ? (e_hi & ~63) : mentry0 - 64; base >= mentry0; base -= 64) {
        const Window w = window_candidates(a, base, r0, e_hi, rect, lane, a.boxes != nullptr);
        uint64_t m = w.cand;
        while (m != 0) {
            const int bit = 63 - __builtin_clzll(m);
            m &= ~(1ull << bit);
            const int g = __builtin_amdgcn_readlane(w.gid, bit);
            entry(base + bit, g, splat_record(a, g));
        }
    }
}

// Wave-private texel-gradient accumulator (one chart at a time), zeroed once
// and kept zero between Gaussians by the route's flush.
__device__ __forceinline__ float* zeroed_tex_acc(float (&tex_acc)[256 / kWave][kTexAccFloats]) {
    float* acc = tex_acc[wave_index()];
    for (int k = threadIdx.x & (kWave - 1); k < kTexAccFloats; k += kWave) acc[k] = 0.f;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    return acc;
}

// The plain kernels for CM <= 4 are held to 5 waves per SIMD.
template <int CM, int MODE, bool TB>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MODE == 0 && CM <= 4 ? 5 : 1)))
void k_raster_backward(RasterArgs a, BwdIn b) {
    __shared__ float tex_acc[256 / kWave][kTexAccFloats];
    float* const acc = zeroed_tex_acc(tex_acc);
    raster_backward_wave<CM, MODE, TB>(a, b, blockIdx.x, wave_index(), acc);
}

// Persistent backward over every view of a table, in the queue order the
// forward pass left in `items` (longest list first).
template <int CM, int MODE, bool TB>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MODE == 0 && CM <= 4 ? 5 : 1)))
void k_raster_backward_queue(const ViewRaster* __restrict__ views, ViewsHeader* __restrict__ hdr,
                             const int* __restrict__ items) {
    __shared__ float tex_acc[256 / kWave][kTexAccFloats];
    float* const acc = zeroed_tex_acc(tex_acc);
    const int wpt = hdr->waves_per_tile, n_work = hdr->num_items * wpt;
    const int nv = hdr->num_views;
    for (;;) {
        int k = 0;
        if ((threadIdx.x & (kWave - 1)) == 0) k = atomicAdd(&hdr->head, 1);
        k = __builtin_amdgcn_readfirstlane(__shfl(k, 0));
        if (k >= n_work) break;
        const int it = k / wpt, wave = k - it * wpt;
        const int item = __builtin_amdgcn_readfirstlane(as_const_space(items)[it]);
        const int v = item >> kViewItemTileBits, tile = item & ((1 << kViewItemTileBits) - 1);
        if (v >= nv) continue;
        const GSTEX_CONST ViewRaster* vr = as_const_space(views) + v;
        RasterArgs a = load_const(&vr->a);
        const BwdIn b = load_const(&vr->b);
        if (tile >= a.tiles_x * a.tiles_y) continue;
        // the backward differentiates the Gaussian alpha whatever the alpha
        // visualisation bits say; masks such a forward wrote mark other entries
        if (a.settings & kSetAlphaViz) a.wave_masks = nullptr;
        raster_backward_wave<CM, MODE, TB>(a, b, tile, wave, acc);
    }
}

// Rebuild the reference gradient tensors from the reduced scalars and the
// splat record (axes, mean, uv maps, inverse scales).
// The 22 reference gradient values of Gaussian g for one view, in the order
// colour (3), opacity, mean (3), scale (3), quaternion (4), uv0 (2), umap (3),
// vmap (3).
constexpr int kFinalFloats = 22;
// Gradient records of a block's 256 Gaussians, staged in LDS by contiguous
// 16-byte loads (row stride 33 floats: the lanes of a wave read distinct banks).
using GradRows = float[256][kGradFloats + 1];
__device__ __forceinline__ void stage_grad_rows(const float* __restrict__ grad_rec, int g0, int n,
                                                GradRows& rows) {
    const int cnt = min(256, n - g0), tid = threadIdx.x;
    const float4* src = reinterpret_cast<const float4*>(grad_rec + (size_t)g0 * kGradFloats);
#pragma unroll
    for (int k = 0; k < kGradFloats / 4; ++k) {
        const int f = k * 256 + tid, rr = f / (kGradFloats / 4), c = 4 * (f % (kGradFloats / 4));
        if (rr < cnt) {
            const float4 t = src[f];
            rows[rr][c] = t.x; rows[rr][c + 1] = t.y; rows[rr][c + 2] = t.z; rows[rr][c + 3] = t.w;
        }
    }
}

__device__ __forceinline__ void finalize_one(int g, const float* q_row,
                                             const float* __restrict__ records,
                                             const float* __restrict__ quats,
                                             const float* __restrict__ scales,
                                             const float* __restrict__ viewmat,
                                             const float* __restrict__ c2w, float fx, float fy,
                                             float (&out)[kFinalFloats]) {
    float q[kGradFloats];
#pragma unroll
    for (int k = 0; k < kGradFloats; ++k) q[k] = q_row[k];
    const Splat s = load_splat(records + (size_t)g * kRecordFloats);
    const V3 mo = s.mo;
    out[0] = q[G_RGB];
    out[1] = q[G_RGB + 1];
    out[2] = q[G_RGB + 2];
    out[3] = q[G_OPAC];
    const V3 ga1 = v3(s.a1.x * q[G_P11] + s.a2.x * q[G_P12], s.a1.y * q[G_P11] + s.a2.y * q[G_P12],
                      s.a1.z * q[G_P11] + s.a2.z * q[G_P12]);
    const V3 ga2 = v3(s.a1.x * q[G_P21] + s.a2.x * q[G_P22], s.a1.y * q[G_P21] + s.a2.y * q[G_P22],
                      s.a1.z * q[G_P21] + s.a2.z * q[G_P22]);
    const V3 ga3 = v3(mo.x * q[G_VNUM] + q[G_AX3X], mo.y * q[G_VNUM] + q[G_AX3X + 1],
                      mo.z * q[G_VNUM] + q[G_AX3X + 2]);
    V3 vm = v3(-(s.a1.x * q[G_VL1] + s.a2.x * q[G_VL2] + s.umap.x * q[G_VU] + s.vmap.x * q[G_VV]) +
                   s.a3.x * q[G_VNUM],
               -(s.a1.y * q[G_VL1] + s.a2.y * q[G_VL2] + s.umap.y * q[G_VU] + s.vmap.y * q[G_VV]) +
                   s.a3.y * q[G_VNUM],
               -(s.a1.z * q[G_VL1] + s.a2.z * q[G_VL2] + s.umap.z * q[G_VU] + s.vmap.z * q[G_VV]) +
                   s.a3.z * q[G_VNUM]);
    if (q[G_PIX] != 0.f || q[G_PIX + 1] != 0.f) {
        // blur term: d projected mean / d mean (helpers.cuh:114-121,155-164)
        const V3 o = v3(c2w[3], c2w[7], c2w[11]);
        const V3 pv = xform43(viewmat, v3(s.mo.x + o.x, s.mo.y + o.y, s.mo.z + o.z));
        const float rw = 1.f / (pv.z + 1e-6f);
        const float vpx = fx * q[G_PIX], vpy = fy * q[G_PIX + 1];
        const V3 vview = v3(vpx * rw, vpy * rw, -(vpx * pv.x + vpy * pv.y) * rw * rw);
        const V3 vmx = xform43_rot_t(viewmat, vview);
        vm = v3(vm.x + vmx.x, vm.y + vmx.y, vm.z + vmx.z);
    }
    out[4] = vm.x;
    out[5] = vm.y;
    out[6] = vm.z;
    // sigma = 0.5 (l1^2 is1^2 + ...) with is1 = 1 / (s_x g): d/ds_x = -(sum vl1 l1) / s_x
    out[7] = -q[G_P11] / scales[3 * g];
    out[8] = -q[G_P22] / scales[3 * g + 1];
    out[9] = 0.f;
    const float4 qt = reinterpret_cast<const float4*>(quats)[g];
    float vq[4];
    quat_axes_vjp(qt.x, qt.y, qt.z, qt.w, ga1, ga2, ga3, vq);
    out[10] = vq[0];
    out[11] = vq[1];
    out[12] = vq[2];
    out[13] = vq[3];
    out[14] = q[G_VU];
    out[15] = q[G_VV];
    out[16] = s.a1.x * q[G_U1] + s.a2.x * q[G_U2];
    out[17] = s.a1.y * q[G_U1] + s.a2.y * q[G_U2];
    out[18] = s.a1.z * q[G_U1] + s.a2.z * q[G_U2];
    out[19] = s.a1.x * q[G_V1] + s.a2.x * q[G_V2];
    out[20] = s.a1.y * q[G_V1] + s.a2.y * q[G_V2];
    out[21] = s.a1.z * q[G_V1] + s.a2.z * q[G_V2];
}

struct FinalOut {
    float *v_colors, *v_opacity, *v_means, *v_scales, *v_quats, *v_uv0, *v_umap, *v_vmap;
};

__device__ __forceinline__ void store_final(const FinalOut& o, int g,
                                            const float (&f)[kFinalFloats]) {
    o.v_colors[3 * g] = f[0];
    o.v_colors[3 * g + 1] = f[1];
    o.v_colors[3 * g + 2] = f[2];
    o.v_opacity[g] = f[3];
    o.v_means[3 * g] = f[4];
    o.v_means[3 * g + 1] = f[5];
    o.v_means[3 * g + 2] = f[6];
    o.v_scales[3 * g] = f[7];
    o.v_scales[3 * g + 1] = f[8];
    o.v_scales[3 * g + 2] = f[9];
    reinterpret_cast<float4*>(o.v_quats)[g] = make_float4(f[10], f[11], f[12], f[13]);
    o.v_uv0[2 * g] = f[14];
    o.v_uv0[2 * g + 1] = f[15];
    o.v_umap[3 * g] = f[16];
    o.v_umap[3 * g + 1] = f[17];
    o.v_umap[3 * g + 2] = f[18];
    o.v_vmap[3 * g] = f[19];
    o.v_vmap[3 * g + 1] = f[20];
    o.v_vmap[3 * g + 2] = f[21];
}

// Rebuild the reference gradient tensors from the reduced scalars and the
// splat record (axes, mean, uv maps, inverse scales).
__global__ __launch_bounds__(256) void k_grad_finalize(
    int n, const float* __restrict__ grad_rec, const float* __restrict__ records,
    const float* __restrict__ quats, const float* __restrict__ scales,
    const float* __restrict__ viewmat, const float* __restrict__ c2w, float fx, float fy,
    FinalOut o) {
    __shared__ GradRows rows;
    const int g0 = blockIdx.x * blockDim.x, g = g0 + threadIdx.x;
    stage_grad_rows(grad_rec, g0, n, rows);
    __syncthreads();
    if (g >= n) return;
    float f[kFinalFloats];
    finalize_one(g, rows[threadIdx.x], records, quats, scales, viewmat, c2w, fx, fy, f);
    store_final(o, g, f);
}

// k_grad_finalize followed by the SH adjoint of its colour gradient (what
// gstex_compute_sh_backward does with viewdirs = means - camera position,
// the same float32 subtraction): v_coeffs rows of 3 nb floats, bands up to
// D.  The block's LDS first holds its gradient records, then its SH rows
// (sh_rows_staged), so the dynamic LDS is the larger of the two.
template <int D>
__global__ __launch_bounds__(256) void k_grad_finalize_sh(
    int n, const float* __restrict__ grad_rec, const float* __restrict__ records,
    const float* __restrict__ quats, const float* __restrict__ scales,
    const float* __restrict__ viewmat, const float* __restrict__ c2w, float fx, float fy,
    FinalOut o, const float* __restrict__ means, int nb, float* __restrict__ v_coeffs) {
    extern __shared__ float lds[];
    GradRows& rows = *reinterpret_cast<GradRows*>(lds);
    const int g0 = blockIdx.x * blockDim.x, g = g0 + threadIdx.x;
    stage_grad_rows(grad_rec, g0, n, rows);
    __syncthreads();
    const bool valid = g < n;
    float b[25];
    b[0] = kShC0;
    float c0 = 0.f, c1 = 0.f, c2 = 0.f;
    if (valid) {
        float f[kFinalFloats];
        finalize_one(g, rows[threadIdx.x], records, quats, scales, viewmat, c2w, fx, fy, f);
        store_final(o, g, f);
        c0 = f[0];
        c1 = f[1];
        c2 = f[2];
        if constexpr (D >= 1) {
            float x, y, z;
            sh_basis<D>(means[3 * g] - c2w[3], means[3 * g + 1] - c2w[7],
                        means[3 * g + 2] - c2w[11], b, x, y, z);
        }
    }
    __syncthreads();  // the gradient records are consumed
    sh_rows_staged<D>(lds, nb, b, c0, c1, c2, g0, n, v_coeffs);
}

// The sum over V views of k_grad_finalize: view v's gradient records start at
// grad_rec + v n 32 and its splat records rec_stride bytes after view v - 1's;
// cameras are (V, 16) row-major matrices and (V, 2) focal lengths.  Each
// view's values are formed exactly as k_grad_finalize forms them and added in
// view order without fused multiply-add, so the result equals adding the
// per-view outputs one after another.  colors_views (V, n, 3), when given,
// receives each view's colour gradient (the input of the SH adjoint).
__global__ __launch_bounds__(256) void k_grad_finalize_views(
    int n, int views, const float* __restrict__ grad_rec, const char* __restrict__ records,
    size_t rec_stride, const float* __restrict__ quats, const float* __restrict__ scales,
    const float* __restrict__ viewmats, const float* __restrict__ c2ws,
    const float* __restrict__ focals, FinalOut o, float* __restrict__ colors_views) {
    __shared__ GradRows rows;
    const int g0 = blockIdx.x * blockDim.x, g = g0 + threadIdx.x;
    float acc[kFinalFloats];
    for (int v = 0; v < views; ++v) {
        if (v > 0) __syncthreads();  // the previous view's rows are consumed
        stage_grad_rows(grad_rec + (size_t)v * n * kGradFloats, g0, n, rows);
        __syncthreads();
        if (g >= n) continue;
        float f[kFinalFloats];
        finalize_one(g, rows[threadIdx.x],
                     reinterpret_cast<const float*>(records + (size_t)v * rec_stride), quats,
                     scales, viewmats + 16 * v, c2ws + 16 * v, focals[2 * v], focals[2 * v + 1],
                     f);
        if (colors_views) {
            float* cv = colors_views + ((size_t)v * n + g) * 3;
            cv[0] = f[0];
            cv[1] = f[1];
            cv[2] = f[2];
        }
        {
#pragma clang fp contract(off)
#pragma unroll
            for (int k = 0; k < kFinalFloats; ++k) acc[k] = v == 0 ? f[k] : acc[k] + f[k];
        }
    }
    if (g < n) store_final(o, g, acc);
}

// ------------------------------------------------------------------ launchers
template <int CM>
static hipError_t launch_bwd_cm(const RasterArgs& a, const BwdIn& b, hipStream_t s) {
    const int tiles = a.tiles_x * a.tiles_y;
    // 1: blur, NDC distortion or alpha visualisation (and the float64 edge-on
    // solve if on); 2: the plain path with the float64 edge-on solve
    const int mode = (a.settings & (kSetBlur | kSetNdc)) ? 1 : a.precise_edge ? 2 : 0;
    const dim3 grid(tiles), block(raster_block_threads(a.bw));
    if (a.texture_bf16) {
        if (mode == 1) k_raster_backward<CM, 1, true><<<grid, block, 0, s>>>(a, b);
        else if (mode == 2) k_raster_backward<CM, 2, true><<<grid, block, 0, s>>>(a, b);
        else k_raster_backward<CM, 0, true><<<grid, block, 0, s>>>(a, b);
    } else {
        if (mode == 1) k_raster_backward<CM, 1, false><<<grid, block, 0, s>>>(a, b);
        else if (mode == 2) k_raster_backward<CM, 2, false><<<grid, block, 0, s>>>(a, b);
        else k_raster_backward<CM, 0, false><<<grid, block, 0, s>>>(a, b);
    }
    return hipGetLastError();
}

template <int CM>
static hipError_t launch_bwd_queue_cm(const ViewRaster* views, ViewsHeader* hdr, const int* items,
                                      int mode, bool bf16, int blocks, hipStream_t s) {
    const dim3 grid(blocks), block(256);
    if (bf16) {
        if (mode == 1) k_raster_backward_queue<CM, 1, true><<<grid, block, 0, s>>>(views, hdr, items);
        else if (mode == 2) k_raster_backward_queue<CM, 2, true><<<grid, block, 0, s>>>(views, hdr, items);
        else k_raster_backward_queue<CM, 0, true><<<grid, block, 0, s>>>(views, hdr, items);
    } else {
        if (mode == 1) k_raster_backward_queue<CM, 1, false><<<grid, block, 0, s>>>(views, hdr, items);
        else if (mode == 2) k_raster_backward_queue<CM, 2, false><<<grid, block, 0, s>>>(views, hdr, items);
        else k_raster_backward_queue<CM, 0, false><<<grid, block, 0, s>>>(views, hdr, items);
    }
    return hipGetLastError();
}

hipError_t raster_backward_views(const ViewRaster* views, ViewsHeader* hdr, const int* items,
                                 int channels, int settings, bool bf16, int blocks,
                                 hipStream_t s) {
    const int mode = (settings & (kSetBlur | kSetNdc)) ? 1 : raster_precise_edge_enabled() ? 2 : 0;
    if (channels == 1) return launch_bwd_queue_cm<1>(views, hdr, items, mode, bf16, blocks, s);
    if (channels == 2) return launch_bwd_queue_cm<2>(views, hdr, items, mode, bf16, blocks, s);
    if (channels == 3) return launch_bwd_queue_cm<3>(views, hdr, items, mode, bf16, blocks, s);
    if (channels == 4) return launch_bwd_queue_cm<4>(views, hdr, items, mode, bf16, blocks, s);
    if (channels <= 8) return launch_bwd_queue_cm<8>(views, hdr, items, mode, bf16, blocks, s);
    if (channels <= 16) return launch_bwd_queue_cm<16>(views, hdr, items, mode, bf16, blocks, s);
    return launch_bwd_queue_cm<64>(views, hdr, items, mode, bf16, blocks, s);
}

size_t raster_backward_workspace_bytes(int n) {
    return splat_records_bytes(n) + (size_t)n * kGradRecordBytes + 512;
}

hipError_t raster_backward_stage(const RasterCall& c, const float* records,
                                 const float* final_Ts, const int* final_idx,
                                 const int* depth_idx, const float* final_s, const float* v_img,
                                 const float* v_depth, const float* v_reg, const float* v_alpha,
                                 const float* v_texture_out, const float* v_normal,
                                 float* grad_rec, float* v_texture, hipStream_t s,
                                 const float* loss_target, float loss_scale) {
    RasterArgs a = make_raster_args(c, records);
    // The backward differentiates the Gaussian alpha whatever the alpha
    // visualisation bits say (texture.cu:381-387 reads only bits 2, 8, 9, 10);
    // masks written by a visualisation forward would mark other entries.
    if (c.settings & kSetAlphaViz) a.wave_masks = nullptr;
    if (loss_target && c.channels != 3) return hipErrorInvalidValue;
    const BwdIn b{final_Ts, final_idx, depth_idx, final_s, v_img, v_depth, v_reg, v_alpha,
                  v_texture_out, v_normal, grad_rec, v_texture, loss_target, loss_scale};
    if (a.tiles_x * a.tiles_y <= 0) return hipSuccess;
    if (c.channels == 1) return launch_bwd_cm<1>(a, b, s);
    if (c.channels == 2) return launch_bwd_cm<2>(a, b, s);
    if (c.channels == 3) return launch_bwd_cm<3>(a, b, s);
    if (c.channels == 4) return launch_bwd_cm<4>(a, b, s);
    if (c.channels <= 8) return launch_bwd_cm<8>(a, b, s);
    if (c.channels <= 16) return launch_bwd_cm<16>(a, b, s);
    return launch_bwd_cm<64>(a, b, s);
}

hipError_t grad_finalize(int n, const float* grad_rec, const float* records, const float* quats,
                         const float* scales, const float* viewmat, const float* c2w, float fx,
                         float fy, float* v_colors, float* v_opacity, float* v_means,
                         float* v_scales, float* v_quats, float* v_uv0, float* v_umap,
                         float* v_vmap, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const FinalOut o{v_colors, v_opacity, v_means, v_scales, v_quats, v_uv0, v_umap, v_vmap};
    k_grad_finalize<<<(n + 255) / 256, 256, 0, s>>>(n, grad_rec, records, quats, scales, viewmat,
                                                   c2w, fx, fy, o);
    return hipGetLastError();
}

bool grad_finalize_sh_supported(int nb, const float* v_coeffs) {
    return (reinterpret_cast<uintptr_t>(v_coeffs) & 15) == 0 &&
           std::max(sizeof(GradRows), sh_rows_stage_bytes(nb)) <= 65536;
}

hipError_t grad_finalize_sh(int n, const float* grad_rec, const float* records, const float* quats,
                            const float* scales, const float* viewmat, const float* c2w,
                            float fx, float fy, float* v_colors, float* v_opacity,
                            float* v_means, float* v_scales, float* v_quats, float* v_uv0,
                            float* v_umap, float* v_vmap, const float* means, int nb,
                            int degrees_to_use, float* v_coeffs, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (!grad_finalize_sh_supported(nb, v_coeffs)) return hipErrorInvalidValue;
    const FinalOut o{v_colors, v_opacity, v_means, v_scales, v_quats, v_uv0, v_umap, v_vmap};
    const size_t lds = std::max(sizeof(GradRows), sh_rows_stage_bytes(nb));
    const int g = (n + 255) / 256;
#define GSTEX_FIN_SH(D)                                                                       \
    k_grad_finalize_sh<D><<<g, 256, lds, s>>>(n, grad_rec, records, quats, scales, viewmat, c2w, \
                                              fx, fy, o, means, nb, v_coeffs)
    switch (degrees_to_use) {
        case 0: GSTEX_FIN_SH(0); break;
        case 1: GSTEX_FIN_SH(1); break;
        case 2: GSTEX_FIN_SH(2); break;
        case 3: GSTEX_FIN_SH(3); break;
        default: GSTEX_FIN_SH(4); break;
    }
#undef GSTEX_FIN_SH
    return hipGetLastError();
}

hipError_t grad_finalize_views(int n, int views, const float* grad_rec, const void* records,
                               size_t rec_stride, const float* quats, const float* scales,
                               const float* viewmats, const float* c2ws, const float* focals,
                               float* v_colors, float* v_opacity, float* v_means, float* v_scales,
                               float* v_quats, float* v_uv0, float* v_umap, float* v_vmap,
                               float* colors_views, hipStream_t s) {
    if (n <= 0 || views <= 0) return hipSuccess;
    const FinalOut o{v_colors, v_opacity, v_means, v_scales, v_quats, v_uv0, v_umap, v_vmap};
    k_grad_finalize_views<<<(n + 255) / 256, 256, 0, s>>>(
        n, views, grad_rec, (const char*)records, rec_stride, quats, scales, viewmats, c2ws,
        focals, o, colors_views);
    return hipGetLastError();
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
    // gradient records after the splat records + cull boxes, 256-byte aligned
    const size_t rec_bytes = (splat_records_bytes(c.num_points) + 255) & ~(size_t)255;
    float* grad_rec = (float*)((char*)ws + rec_bytes);
    hipError_t e = build_raster_records(c, records, s);
    if (e != hipSuccess) return e;
    e = fill_f32(grad_rec, (size_t)c.num_points * kGradFloats, 0.f, s);
    if (e != hipSuccess) return e;
    e = fill_f32(v_texture, (size_t)c.num_texels * (size_t)c.channels, 0.f, s);
    if (e != hipSuccess) return e;
    e = raster_backward_stage(c, records, final_Ts, final_idx, depth_idx, final_s, v_img, v_depth,
                              v_reg, v_alpha, v_texture_out, v_normal, grad_rec, v_texture, s);
    if (e != hipSuccess) return e;
    return grad_finalize(c.num_points, grad_rec, records, c.quats, c.scales, c.viewmat, c.c2w,
                         c.fx, c.fy, v_colors, v_opacity, v_means, v_scales, v_quats, v_uv0,
                         v_umap, v_vmap, s);
}

}  // namespace gstex
