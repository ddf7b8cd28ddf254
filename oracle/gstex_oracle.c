/*
 * gstex CPU oracle — TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * A plain-C restatement of the reference's textured-2DGS rasteriser semantics:
 *   projection / AABB        get_aabb_2d.cu:11-89
 *   tile counts              get_aabb_2d.py:70-92
 *   intersect keys           forward.cu:13-71, helpers.cuh:37-51,75-92
 *   sort                     utils.py:159-160 (stable order, ties by position)
 *   bin edges                forward.cu:76-98
 *   SH forward / backward    sh.cuh:6-253
 *   texture sampling         texture_sample.cu:11-36, texture_helpers.cuh:162-237
 *   raster forward           texture.cu:11-329
 *   raster backward          texture.cu:331-760
 *
 * The floating-point body (oracle_body.h) is compiled twice: *_f32 keeps the
 * reference's FP32 operation order, *_f64 evaluates the same formulas in FP64.
 * Integer work (tile counts, keys, sort, bin edges) exists once and is exact.
 * Build with -ffp-contract=off so the FP32 restatement is not re-associated.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int img_w, img_h, block_width, channels, settings;
    double glob_scale, fx, fy, cx, cy;
    /* float arrays for the _f32 functions, double arrays for the _f64 ones */
    const void *means, *scales, *quats, *colors, *opacities, *uv0, *umap, *vmap, *texture;
    const void *viewmat, *c2w, *background;
    const int *texture_dims, *ids_sorted, *tile_bins;
    const unsigned char *tile_mask; /* optional: nonzero = render this tile (NULL = all) */
} orc_scene;

/* Float -> int32 the way GPU conversion instructions do it: truncate toward
 * zero, saturate out-of-range values, NaN -> 0 (the C cast is undefined there). */
static inline int orc_sat_i32(double x) {
    if (x != x) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return (-2147483647 - 1);
    return (int)x;
}

int orc_num_sh_bases(int degree) { /* sh.cuh:34-44 */
    if (degree == 0) return 1;
    if (degree == 1) return 4;
    if (degree == 2) return 9;
    if (degree == 3) return 16;
    return 25;
}

/* get_aabb_2d.py:70-92: number of tiles each AABB touches, in float32 as torch does. */
void orc_num_tiles_hit(int n, const float *centers, const float *extents, int img_h, int img_w,
                       int bw, int *out) {
    int tx = (img_w + bw - 1) / bw, ty = (img_h + bw - 1) / bw;
    float fbw = (float)bw;
    for (int i = 0; i < n; ++i) {
        int lo[2], hi[2], bound[2] = {tx, ty};
        for (int k = 0; k < 2; ++k) {
            float c = centers[2 * i + k], e = extents[2 * i + k];
            float a = (c - e) / fbw;
            float b = (c + e) / fbw + 1.0f;
            lo[k] = orc_sat_i32(floorf(a));
            hi[k] = orc_sat_i32(floorf(b));
            if (lo[k] < 0) lo[k] = 0;
            if (lo[k] > bound[k]) lo[k] = bound[k];
            if (hi[k] < 0) hi[k] = 0;
            if (hi[k] > bound[k]) hi[k] = bound[k];
        }
        out[i] = (hi[0] - lo[0]) * (hi[1] - lo[1]);
    }
}

/* helpers.cuh:37-51,75-92: tile rectangle [min, max) used when writing keys. */
static void orc_tile_rect(float cxp, float cyp, float ex, float ey, int tiles_x, int tiles_y, int bw,
                          int mn[2], int mx[2]) {
    float fbw = (float)bw;
    float tc[2] = {cxp / fbw, cyp / fbw}, te[2] = {ex / fbw, ey / fbw};
    int bound[2] = {tiles_x, tiles_y};
    for (int k = 0; k < 2; ++k) {
        int a = orc_sat_i32(tc[k] - te[k]);
        int b = orc_sat_i32(tc[k] + te[k] + 1.0f);
        a = a > 0 ? a : 0;
        b = b > 0 ? b : 0;
        mn[k] = a < bound[k] ? a : bound[k];
        mx[k] = b < bound[k] ? b : bound[k];
    }
}

/* helpers.cuh:53-73,94-111 (get_tile_bbox_wrapped): the rectangle unclamped,
 * with the lower corner one tile further when it is <= 0 (not below INT_MIN). */
static void orc_tile_rect_wrapped(float cxp, float cyp, float ex, float ey, int bw, int mn[2],
                                  int mx[2]) {
    float fbw = (float)bw;
    float tc[2] = {cxp / fbw, cyp / fbw}, te[2] = {ex / fbw, ey / fbw};
    for (int k = 0; k < 2; ++k) {
        int a = orc_sat_i32(tc[k] - te[k]);
        if (a <= 0 && a > INT32_MIN) a -= 1;
        mn[k] = a;
        mx[k] = orc_sat_i32(tc[k] + te[k] + 1.0f);
    }
}

/* Tile count of the wrapped rectangle (0 when max(extent) <= 1e-4). */
void orc_num_tiles_hit_wrapped(int n, const float *centers, const float *extents, int bw, int *out) {
    for (int g = 0; g < n; ++g) {
        float ex = extents[2 * g], ey = extents[2 * g + 1];
        out[g] = 0;
        if ((double)ex <= 1e-4 && (double)ey <= 1e-4) continue;
        int mn[2], mx[2];
        orc_tile_rect_wrapped(centers[2 * g], centers[2 * g + 1], ex, ey, bw, mn, mx);
        long long w = (long long)mx[0] - mn[0], h = (long long)mx[1] - mn[1];
        long long a = (w > 0 ? w : 0) * (h > 0 ? h : 0);
        out[g] = a > 0x7fffffffLL ? 0x7fffffff : (int)a;
    }
}

/* forward.cu:34-62 (wrapped=true): the wrapped rectangle row by row, tile
 * indices modulo the grid.  Each Gaussian writes at most its own count window
 * [cum[g-1], cum[g]); the reference writes the whole rectangle from cum[g-1],
 * past the window (into the next Gaussian's slots) when it holds more tiles. */
void orc_map_gaussian_to_intersects_wrapped(int n, int num_isects, const float *centers,
                                            const float *extents, const float *depths,
                                            const int *cum, int tiles_x, int tiles_y, int bw,
                                            int64_t *keys, int32_t *ids) {
    memset(keys, 0, sizeof(int64_t) * (size_t)num_isects);
    memset(ids, 0, sizeof(int32_t) * (size_t)num_isects);
    for (int g = 0; g < n; ++g) {
        float ex = extents[2 * g], ey = extents[2 * g + 1];
        if ((double)ex <= 1e-4 && (double)ey <= 1e-4) continue;
        int mn[2], mx[2];
        orc_tile_rect_wrapped(centers[2 * g], centers[2 * g + 1], ex, ey, bw, mn, mx);
        int cur = g == 0 ? 0 : cum[g - 1];
        int end = cum[g] < num_isects ? cum[g] : num_isects;
        int32_t dbits;
        memcpy(&dbits, depths + g, 4);
        int64_t did = (int64_t)dbits;
        for (int i = mn[1]; i < mx[1] && cur < end; ++i)
            for (int j = mn[0]; j < mx[0] && cur < end; ++j) {
                int ti = i % tiles_y, tj = j % tiles_x;
                if (ti < 0) ti += tiles_y;
                if (tj < 0) tj += tiles_x;
                if (cur >= 0) {
                    keys[cur] = ((int64_t)(ti * tiles_x + tj) << 32) | did;
                    ids[cur] = g;
                }
                ++cur;
            }
    }
}

/* forward.cu:13-71 (wrapped=false): keys (tile << 32 | depth bits) and ids,
 * written from cum_tiles_hit[g-1]; slots a Gaussian counts but never writes stay 0. */
void orc_map_gaussian_to_intersects(int n, int num_isects, const float *centers, const float *extents,
                                    const float *depths, const int *cum, int tiles_x, int tiles_y,
                                    int bw, int64_t *keys, int32_t *ids) {
    memset(keys, 0, sizeof(int64_t) * (size_t)num_isects);
    memset(ids, 0, sizeof(int32_t) * (size_t)num_isects);
    for (int g = 0; g < n; ++g) {
        float ex = extents[2 * g], ey = extents[2 * g + 1];
        if ((double)ex <= 1e-4 && (double)ey <= 1e-4) continue;
        int mn[2], mx[2];
        orc_tile_rect(centers[2 * g], centers[2 * g + 1], ex, ey, tiles_x, tiles_y, bw, mn, mx);
        int cur = g == 0 ? 0 : cum[g - 1];
        int32_t dbits;
        memcpy(&dbits, depths + g, 4);
        int64_t did = (int64_t)dbits;
        for (int i = mn[1]; i < mx[1]; ++i)
            for (int j = mn[0]; j < mx[0]; ++j) {
                int64_t tile = (int64_t)(i * tiles_x + j);
                if (cur >= 0 && cur < num_isects) {
                    keys[cur] = (tile << 32) | did;
                    ids[cur] = g;
                }
                ++cur;
            }
    }
}

/* Stable ascending sort of int64 keys carrying int32 ids (utils.py:159-160). */
void orc_sort_pairs(int m, const int64_t *keys_in, const int32_t *ids_in, int64_t *keys_out,
                    int32_t *ids_out) {
    int64_t *ka = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    int64_t *kb = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    int32_t *ia = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    int32_t *ib = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    memcpy(ka, keys_in, sizeof(int64_t) * (size_t)m);
    memcpy(ia, ids_in, sizeof(int32_t) * (size_t)m);
    for (int width = 1; width < m; width *= 2) { /* bottom-up merge sort: stable */
        for (int lo = 0; lo < m; lo += 2 * width) {
            int mid = lo + width < m ? lo + width : m;
            int hi = lo + 2 * width < m ? lo + 2 * width : m;
            int a = lo, b = mid, k = lo;
            while (a < mid && b < hi) {
                if (ka[b] < ka[a]) { kb[k] = ka[b]; ib[k++] = ia[b++]; }
                else { kb[k] = ka[a]; ib[k++] = ia[a++]; }
            }
            while (a < mid) { kb[k] = ka[a]; ib[k++] = ia[a++]; }
            while (b < hi) { kb[k] = ka[b]; ib[k++] = ia[b++]; }
        }
        int64_t *tk = ka; ka = kb; kb = tk;
        int32_t *ti = ia; ia = ib; ib = ti;
    }
    memcpy(keys_out, ka, sizeof(int64_t) * (size_t)m);
    memcpy(ids_out, ia, sizeof(int32_t) * (size_t)m);
    free(ka); free(kb); free(ia); free(ib);
}

/* forward.cu:76-98: [first, last+1) of each tile in the sorted keys; empty tiles stay (0, 0). */
void orc_tile_bin_edges(int m, const int64_t *keys_sorted, int num_tiles, int32_t *bins) {
    memset(bins, 0, sizeof(int32_t) * 2 * (size_t)num_tiles);
    for (int idx = 0; idx < m; ++idx) {
        int cur = (int32_t)(keys_sorted[idx] >> 32);
        if (cur < 0 || cur >= num_tiles) continue;
        if (idx == 0) bins[2 * cur] = 0;
        if (idx == m - 1) bins[2 * cur + 1] = m;
        if (idx == 0) continue;
        int prev = (int32_t)(keys_sorted[idx - 1] >> 32);
        if (prev != cur) {
            if (prev >= 0 && prev < num_tiles) bins[2 * prev + 1] = idx;
            bins[2 * cur] = idx;
        }
    }
}

/* ---- FP32 restatement ---- */
#define REAL float
#define IN float
#define SFX(name) name##_f32
#define SQRT sqrtf
#define EXP expf
#include "oracle_body.h"
#undef REAL
#undef IN
#undef SFX
#undef SQRT
#undef EXP

/* ---- FP64 evaluation of the same formulas ---- */
#define REAL double
#define IN double
#define SFX(name) name##_f64
#define SQRT sqrt
#define EXP exp
#include "oracle_body.h"
#undef REAL
#undef IN
#undef SFX
#undef SQRT
#undef EXP
