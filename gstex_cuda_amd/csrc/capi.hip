// Exported C ABI (declared in include/gstex_hip.h): argument validation,
// then the gfx950 launchers.  No entry point allocates device memory.
#include <stdarg.h>
#include <stdio.h>

#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/gstex_hip.h"
#include "gstex_common.hpp"
#include "gstex_kernels.hpp"
#include "raster_common.hpp"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return GSTEX_OK;
    if (e == hipErrorInvalidValue)
        return fail(GSTEX_ERR_SPACE, "%s: invalid value (workspace too small?): %s", what,
                    hipGetErrorString(e));
    return fail(GSTEX_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

#define REQUIRE(cond, ...)                              \
    do {                                                \
        if (!(cond)) return fail(GSTEX_ERR_ARG, __VA_ARGS__); \
    } while (0)

int tiles_of(int extent, int bw) { return (extent + bw - 1) / bw; }

// Tile grids the kernels can index: tile coordinates are packed as 16-bit
// halves in the binning emission, and tile indices (and the raster grid) are
// 32-bit ints.
#define REQUIRE_TILE_GRID(tx, ty)                                                           \
    REQUIRE((tx) < 65536 && (ty) < 65536 && (int64_t)(tx) * (int64_t)(ty) <= 0x7fffffffll,   \
            "tile grid %d x %d is too large (each side below 65536, at most 2^31 - 1 tiles)", \
            (tx), (ty))

int check_raster_common(int tiles_x, int tiles_y, int block_width, int img_width, int img_height,
                        int num_probs, int channels, int num_points, int num_ids, int num_texels,
                        int settings) {
    REQUIRE(block_width >= 2 && block_width <= 16,
            "block_width must be between 2 and 16, got %d", block_width);
    REQUIRE(img_width > 0 && img_height > 0, "image size must be positive, got %dx%d", img_width,
            img_height);
    REQUIRE_TILE_GRID(tiles_x, tiles_y);
    REQUIRE(tiles_x == tiles_of(img_width, block_width) &&
                tiles_y == tiles_of(img_height, block_width),
            "tile_bounds (%d, %d) do not match img_size (%d, %d) and block width %d", tiles_x,
            tiles_y, img_width, img_height, block_width);
    REQUIRE(num_probs == 1, "texture_info[1] (num_probs) must be 1, got %d", num_probs);
    REQUIRE(channels >= 0 && channels <= 64, "texture channels must be in [0, 64], got %d",
            channels);
    REQUIRE(num_points >= 0 && num_ids >= 0, "negative sizes");
    REQUIRE(channels == 0 || num_texels > 0, "texture has no rows");
    REQUIRE(settings >= 0 && settings < (1 << 30), "settings bits 30 and 31 are not defined");
    return GSTEX_OK;
}

}  // namespace

extern "C" {

const char* gstex_last_error(void) { return g_err; }

int gstex_abi_version(void) { return 1; }

int gstex_get_aabb_2d(int num_points, const float* means, const float* scales, float glob_scale,
                      const float* quats, const float* viewmat, float fx, float fy, float cx,
                      float cy, float* centers, float* extents, void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(means && scales && quats && viewmat && centers && extents, "null pointer argument");
    return check_hip(gstex::launch_get_aabb_2d(num_points, means, scales, glob_scale, quats,
                                               viewmat, fx, fy, cx, cy, centers, extents,
                                               as_stream(stream)),
                     "get_aabb_2d");
}

int gstex_num_tiles_hit(int num_points, const float* centers, const float* extents, int img_height,
                        int img_width, int block_width, int32_t* num_tiles_hit, void* stream) {
    REQUIRE(num_points >= 0 && block_width > 0, "invalid sizes");
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(centers && extents && num_tiles_hit, "null pointer argument");
    return check_hip(gstex::launch_num_tiles_hit(num_points, centers, extents, block_width,
                                                 tiles_of(img_width, block_width),
                                                 tiles_of(img_height, block_width), num_tiles_hit,
                                                 as_stream(stream)),
                     "num_tiles_hit");
}

int gstex_num_tiles_hit_wrapped(int num_points, const float* centers, const float* extents,
                                int block_width, int32_t* num_tiles_hit, void* stream) {
    REQUIRE(num_points >= 0 && block_width > 0, "invalid sizes");
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(centers && extents && num_tiles_hit, "null pointer argument");
    return check_hip(gstex::num_tiles_hit_wrapped(num_points, centers, extents, block_width,
                                                  num_tiles_hit, as_stream(stream)),
                     "num_tiles_hit_wrapped");
}

int gstex_project_gaussians(int num_points, const float* means, const float* scales,
                            float glob_scale, const float* quats, const float* viewmat, float fx,
                            float fy, float cx, float cy, int img_height, int img_width,
                            int block_width, float* centers, float* extents, float* depths,
                            int32_t* num_tiles_hit, void* stream) {
    REQUIRE(num_points >= 0 && block_width > 0, "invalid sizes");
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(means && scales && quats && viewmat && centers && extents && depths && num_tiles_hit,
            "null pointer argument");
    return check_hip(gstex::launch_project_fused(num_points, means, scales, glob_scale, quats,
                                                 viewmat, fx, fy, cx, cy, block_width,
                                                 tiles_of(img_width, block_width),
                                                 tiles_of(img_height, block_width), centers,
                                                 extents, depths, num_tiles_hit, nullptr,
                                                 as_stream(stream)),
                     "project_gaussians");
}

int gstex_project_gaussians_total(int num_points, const float* means, const float* scales,
                                  float glob_scale, const float* quats, const float* viewmat,
                                  float fx, float fy, float cx, float cy, int img_height,
                                  int img_width, int block_width, float* centers, float* extents,
                                  float* depths, int32_t* num_tiles_hit,
                                  uint32_t* num_intersects_out, void* stream) {
    REQUIRE(num_points >= 0 && block_width > 0, "invalid sizes");
    REQUIRE(num_intersects_out, "null pointer argument");
    REQUIRE(num_points == 0 || (means && scales && quats && viewmat && centers && extents &&
                                depths && num_tiles_hit),
            "null pointer argument");
    return check_hip(gstex::launch_project_fused(num_points, means, scales, glob_scale, quats,
                                                 viewmat, fx, fy, cx, cy, block_width,
                                                 tiles_of(img_width, block_width),
                                                 tiles_of(img_height, block_width), centers,
                                                 extents, depths, num_tiles_hit,
                                                 num_intersects_out, as_stream(stream)),
                     "project_gaussians_total");
}

int gstex_project_points(int num_points, const float* points, const float* viewmat, float fx,
                         float fy, float cx, float cy, int clip, float* pix, float* depths,
                         void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(points && viewmat && pix && depths, "null pointer argument");
    return check_hip(gstex::launch_project_points(num_points, points, viewmat, fx, fy, cx, cy,
                                                  clip, pix, depths, as_stream(stream)),
                     "project_points");
}

int gstex_compute_sh_forward(int num_points, int degree, int degrees_to_use, const float* viewdirs,
                             const float* coeffs, float* colors, void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    REQUIRE(degree >= 0 && degree <= 4, "SH degree must be in [0, 4], got %d", degree);
    REQUIRE(degrees_to_use >= 0 && degrees_to_use <= degree,
            "degrees_to_use must be in [0, degree], got %d", degrees_to_use);
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(viewdirs && coeffs && colors, "null pointer argument");
    return check_hip(gstex::launch_sh_forward(num_points, degree, degrees_to_use, viewdirs, coeffs,
                                              colors, as_stream(stream)),
                     "compute_sh_forward");
}

int gstex_compute_sh_forward_bf16(int num_points, int degree, int degrees_to_use,
                                  const float* viewdirs, const uint16_t* coeffs, float* colors,
                                  void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    REQUIRE(degree >= 0 && degree <= 4, "SH degree must be in [0, 4], got %d", degree);
    REQUIRE(degrees_to_use >= 0 && degrees_to_use <= degree,
            "degrees_to_use must be in [0, degree], got %d", degrees_to_use);
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(viewdirs && coeffs && colors, "null pointer argument");
    return check_hip(gstex::launch_sh_forward_bf16(num_points, degree, degrees_to_use, viewdirs,
                                                   coeffs, colors, as_stream(stream)),
                     "compute_sh_forward_bf16");
}

int gstex_compute_sh_forward_means(int num_points, int degree, int degrees_to_use,
                                   const float* means, const float* c2w, const void* coeffs,
                                   int coeffs_format, float* colors, void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    REQUIRE(degree >= 0 && degree <= 4, "SH degree must be in [0, 4], got %d", degree);
    REQUIRE(degrees_to_use >= 0 && degrees_to_use <= degree,
            "degrees_to_use must be in [0, degree], got %d", degrees_to_use);
    REQUIRE(coeffs_format == GSTEX_TEXTURE_F32 || coeffs_format == GSTEX_TEXTURE_BF16,
            "coeffs_format must be 0 (float32) or 1 (bfloat16), got %d", coeffs_format);
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(means && c2w && coeffs && colors, "null pointer argument");
    return check_hip(gstex::launch_sh_forward_means(num_points, degree, degrees_to_use, means,
                                                    c2w, coeffs,
                                                    coeffs_format == GSTEX_TEXTURE_BF16, colors,
                                                    as_stream(stream)),
                     "compute_sh_forward_means");
}

int gstex_compute_sh_backward(int num_points, int degree, int degrees_to_use,
                              const float* viewdirs, const float* v_colors, float* v_coeffs,
                              void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    REQUIRE(degree >= 0 && degree <= 4, "SH degree must be in [0, 4], got %d", degree);
    REQUIRE(degrees_to_use >= 0 && degrees_to_use <= degree,
            "degrees_to_use must be in [0, degree], got %d", degrees_to_use);
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(viewdirs && v_colors && v_coeffs, "null pointer argument");
    return check_hip(gstex::launch_sh_backward(num_points, degree, degrees_to_use, viewdirs,
                                               v_colors, v_coeffs, as_stream(stream)),
                     "compute_sh_backward");
}

int gstex_compute_sh_backward_views(int num_points, int num_views, int degree,
                                    int degrees_to_use, const float* viewdirs,
                                    const float* v_colors, float* v_coeffs, void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    REQUIRE(num_views >= 1, "num_views must be >= 1, got %d", num_views);
    REQUIRE(degree >= 0 && degree <= 4, "SH degree must be in [0, 4], got %d", degree);
    REQUIRE(degrees_to_use >= 0 && degrees_to_use <= degree,
            "degrees_to_use must be in [0, degree], got %d", degrees_to_use);
    REQUIRE((int64_t)num_points * num_views * 3 < (int64_t)1 << 40, "input too large");
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(viewdirs && v_colors && v_coeffs, "null pointer argument");
    return check_hip(gstex::launch_sh_backward_views(num_points, num_views, degree,
                                                     degrees_to_use, viewdirs, v_colors,
                                                     v_coeffs, as_stream(stream)),
                     "compute_sh_backward_views");
}

int gstex_map_gaussian_to_intersects(int num_points, int num_intersects, const float* centers,
                                     const float* extents, const float* depths,
                                     const int32_t* cum_tiles_hit, int tiles_x, int tiles_y,
                                     int block_width, int wrapped, int64_t* isect_ids,
                                     int32_t* gaussian_ids, void* stream) {
    REQUIRE(num_points >= 0 && num_intersects >= 0, "negative sizes");
    REQUIRE(block_width > 0 && tiles_x >= 0 && tiles_y >= 0, "invalid tile bounds");
    REQUIRE_TILE_GRID(tiles_x, tiles_y);
    REQUIRE(!wrapped || (tiles_x > 0 && tiles_y > 0), "wrapped binning needs a non-empty tile grid");
    if (num_intersects == 0) return GSTEX_OK;
    REQUIRE(isect_ids && gaussian_ids, "null pointer argument");
    REQUIRE(num_points == 0 || (centers && extents && depths && cum_tiles_hit),
            "null pointer argument");
    return check_hip(gstex::map_intersects(num_points, num_intersects, centers, extents, depths,
                                           cum_tiles_hit, tiles_x, tiles_y, block_width,
                                           wrapped != 0, isect_ids, gaussian_ids,
                                           as_stream(stream)),
                     "map_gaussian_to_intersects");
}

size_t gstex_sort_intersects_workspace_size(int num_intersects) {
    return num_intersects > 0 ? gstex::sort_isects_workspace_bytes(num_intersects) : 0;
}

int gstex_sort_intersects(int num_intersects, int end_bit, const int64_t* isect_ids,
                          const int32_t* gaussian_ids, int64_t* isect_ids_sorted,
                          int32_t* gaussian_ids_sorted, void* workspace, size_t workspace_bytes,
                          void* stream) {
    REQUIRE(num_intersects >= 0, "num_intersects must be >= 0");
    REQUIRE(end_bit >= 0 && end_bit <= 64, "end_bit must be in [0, 64], got %d", end_bit);
    if (num_intersects == 0) return GSTEX_OK;
    REQUIRE(isect_ids && gaussian_ids && isect_ids_sorted && gaussian_ids_sorted && workspace,
            "null pointer argument");
    REQUIRE(workspace_bytes >= gstex::sort_isects_workspace_bytes(num_intersects),
            "workspace too small: %zu < %zu", workspace_bytes,
            gstex::sort_isects_workspace_bytes(num_intersects));
    return check_hip(gstex::sort_isects(num_intersects, isect_ids, gaussian_ids, isect_ids_sorted,
                                        gaussian_ids_sorted, end_bit, workspace, workspace_bytes,
                                        as_stream(stream)),
                     "sort_intersects");
}

size_t gstex_sort_pairs_u32_workspace_size(int n) {
    return n > 0 ? gstex::sort_u32_pairs_workspace_bytes(n) : 0;
}

int gstex_sort_pairs_u32(int n, int end_bit, const uint32_t* keys, const uint32_t* values,
                         uint32_t* keys_sorted, uint32_t* values_sorted, void* workspace,
                         size_t workspace_bytes, void* stream) {
    REQUIRE(n >= 0, "n must be >= 0");
    REQUIRE(end_bit >= 0 && end_bit <= 32, "end_bit must be in [0, 32], got %d", end_bit);
    if (n == 0) return GSTEX_OK;
    REQUIRE(keys && values && keys_sorted && values_sorted && workspace, "null pointer argument");
    REQUIRE(workspace_bytes >= gstex::sort_u32_pairs_workspace_bytes(n),
            "workspace too small: %zu < %zu", workspace_bytes,
            gstex::sort_u32_pairs_workspace_bytes(n));
    return check_hip(gstex::sort_u32_pairs(n, keys, values, keys_sorted, values_sorted, end_bit,
                                           workspace, workspace_bytes, as_stream(stream)),
                     "sort_pairs_u32");
}

int gstex_get_tile_bin_edges(int num_intersects, const int64_t* isect_ids_sorted, int tiles_x,
                             int tiles_y, int32_t* tile_bins, void* stream) {
    REQUIRE(num_intersects >= 0 && tiles_x >= 0 && tiles_y >= 0, "negative sizes");
    REQUIRE_TILE_GRID(tiles_x, tiles_y);
    if (tiles_x * tiles_y == 0) return GSTEX_OK;
    REQUIRE(tile_bins && (num_intersects == 0 || isect_ids_sorted), "null pointer argument");
    return check_hip(gstex::tile_bin_edges(num_intersects, isect_ids_sorted, tiles_x * tiles_y,
                                           tile_bins, as_stream(stream)),
                     "get_tile_bin_edges");
}

size_t gstex_bin_gaussians_workspace_size(int num_points, int num_intersects) {
    if (num_points <= 0 || num_intersects <= 0) return 0;
    return gstex::bin_gaussians_workspace_bytes(num_points, num_intersects);
}

size_t gstex_bin_gaussians_depth_workspace_size(int num_points) {
    return gstex::bin_depth_workspace_bytes(num_points);
}

size_t gstex_bin_gaussians_tiles_workspace_size(int num_points, int num_intersects) {
    (void)num_points;
    return gstex::bin_tiles_workspace_bytes(num_intersects);
}

int gstex_bin_gaussians_depth(int num_points, const float* centers, const float* extents,
                              const float* depths, const int32_t* num_tiles_hit, int tiles_x,
                              int tiles_y, int block_width, int32_t* tile_bins, void* depth_ws,
                              size_t depth_ws_bytes, void* stream) {
    REQUIRE(num_points >= 0, "negative sizes");
    REQUIRE(block_width > 0 && tiles_x >= 0 && tiles_y >= 0, "invalid tile bounds");
    REQUIRE_TILE_GRID(tiles_x, tiles_y);
    if (tiles_x * tiles_y == 0) return GSTEX_OK;
    REQUIRE(tile_bins, "null pointer argument");
    if (num_points > 0) {
        REQUIRE(centers && extents && depths && num_tiles_hit && depth_ws,
                "null pointer argument");
        const size_t need = gstex::bin_depth_workspace_bytes(num_points);
        REQUIRE(depth_ws_bytes >= need, "workspace too small: %zu < %zu", depth_ws_bytes, need);
    }
    return check_hip(gstex::bin_depth(num_points, centers, extents, depths, num_tiles_hit,
                                      tiles_x, tiles_y, block_width, tile_bins, depth_ws,
                                      depth_ws_bytes, as_stream(stream)),
                     "bin_gaussians_depth");
}

int gstex_bin_gaussians_tiles(int num_points, int num_intersects,
                              const uint32_t* num_intersects_dev, int tiles_x, int tiles_y,
                              const void* depth_ws, size_t depth_ws_bytes,
                              int32_t* gaussian_ids_sorted, int32_t* tile_bins, void* tiles_ws,
                              size_t tiles_ws_bytes, void* stream) {
    REQUIRE(num_points >= 0 && num_intersects >= 0, "negative sizes");
    REQUIRE(tiles_x >= 0 && tiles_y >= 0, "invalid tile bounds");
    REQUIRE_TILE_GRID(tiles_x, tiles_y);
    if (tiles_x * tiles_y == 0) return GSTEX_OK;
    REQUIRE(tile_bins, "null pointer argument");
    if (num_points > 0 && num_intersects > 0) {
        REQUIRE(depth_ws && gaussian_ids_sorted && tiles_ws, "null pointer argument");
        const size_t need_d = gstex::bin_depth_workspace_bytes(num_points);
        const size_t need_t = gstex::bin_tiles_workspace_bytes(num_intersects);
        REQUIRE(depth_ws_bytes >= need_d, "depth workspace too small: %zu < %zu",
                depth_ws_bytes, need_d);
        REQUIRE(tiles_ws_bytes >= need_t, "workspace too small: %zu < %zu", tiles_ws_bytes,
                need_t);
    }
    return check_hip(gstex::bin_tiles(num_points, num_intersects, num_intersects_dev, tiles_x,
                                      tiles_y, depth_ws, depth_ws_bytes, gaussian_ids_sorted,
                                      tile_bins, tiles_ws, tiles_ws_bytes, as_stream(stream)),
                     "bin_gaussians_tiles");
}

int gstex_bin_gaussians(int num_points, int num_intersects, const float* centers,
                        const float* extents, const float* depths, const int32_t* num_tiles_hit,
                        int tiles_x, int tiles_y, int block_width, int32_t* gaussian_ids_sorted,
                        int32_t* tile_bins, void* workspace, size_t workspace_bytes, void* stream) {
    REQUIRE(num_points >= 0 && num_intersects >= 0, "negative sizes");
    REQUIRE(block_width > 0 && tiles_x >= 0 && tiles_y >= 0, "invalid tile bounds");
    REQUIRE_TILE_GRID(tiles_x, tiles_y);
    if (tiles_x * tiles_y == 0) return GSTEX_OK;
    REQUIRE(tile_bins, "null pointer argument");
    if (num_points > 0 && num_intersects > 0) {
        REQUIRE(centers && extents && depths && num_tiles_hit && gaussian_ids_sorted && workspace,
                "null pointer argument");
        const size_t need = gstex::bin_gaussians_workspace_bytes(num_points, num_intersects);
        REQUIRE(workspace_bytes >= need, "workspace too small: %zu < %zu", workspace_bytes, need);
    }
    return check_hip(gstex::bin_gaussians(num_points, num_intersects, centers, extents, depths,
                                          num_tiles_hit, tiles_x, tiles_y, block_width,
                                          gaussian_ids_sorted, tile_bins, nullptr, workspace,
                                          workspace_bytes, as_stream(stream)),
                     "bin_gaussians");
}

int gstex_count_intersects(int num_points, const int32_t* num_tiles_hit,
                           uint32_t* num_intersects_out, void* stream) {
    REQUIRE(num_points >= 0, "negative sizes");
    REQUIRE(num_intersects_out && (num_points == 0 || num_tiles_hit), "null pointer argument");
    return check_hip(gstex::count_intersects(num_points, num_tiles_hit, num_intersects_out,
                                             as_stream(stream)),
                     "count_intersects");
}

int gstex_bin_gaussians_capped(int num_points, int capacity, const float* centers,
                               const float* extents, const float* depths,
                               const int32_t* num_tiles_hit, int tiles_x, int tiles_y,
                               int block_width, int32_t* gaussian_ids_sorted, int32_t* tile_bins,
                               uint32_t* num_intersects_out, void* workspace,
                               size_t workspace_bytes, void* stream) {
    REQUIRE(num_points >= 0 && capacity >= 0, "negative sizes");
    REQUIRE(block_width > 0 && tiles_x >= 0 && tiles_y >= 0, "invalid tile bounds");
    REQUIRE_TILE_GRID(tiles_x, tiles_y);
    if (tiles_x * tiles_y == 0) return GSTEX_OK;
    REQUIRE(tile_bins && num_intersects_out, "null pointer argument");
    if (num_points > 0) {
        REQUIRE(centers && extents && depths && num_tiles_hit && workspace &&
                    (capacity == 0 || gaussian_ids_sorted),
                "null pointer argument");
        const size_t need = gstex::bin_gaussians_workspace_bytes(num_points, capacity > 0 ? capacity : 1);
        REQUIRE(workspace_bytes >= need, "workspace too small: %zu < %zu", workspace_bytes, need);
    }
    return check_hip(gstex::bin_gaussians(num_points, capacity, centers, extents, depths,
                                          num_tiles_hit, tiles_x, tiles_y, block_width,
                                          gaussian_ids_sorted, tile_bins, num_intersects_out,
                                          workspace, workspace_bytes, as_stream(stream)),
                     "bin_gaussians_capped");
}

size_t gstex_texture_forward_workspace_size(int num_points) {
    return gstex::raster_forward_workspace_bytes(num_points > 0 ? num_points : 0);
}

size_t gstex_texture_backward_workspace_size(int num_points) {
    return gstex::raster_backward_workspace_bytes(num_points > 0 ? num_points : 0);
}

int gstex_texture_forward(int tiles_x, int tiles_y, int block_width, int img_width, int img_height,
                          int num_charts, int num_probs, int channels,
                          const int32_t* texture_dims, const int32_t* gaussian_ids_sorted,
                          int num_ids, const int32_t* tile_bins, const float* colors,
                          const float* opacities, const float* means, const float* scales,
                          int num_points, float glob_scale, const float* quats, const float* uv0,
                          const float* umap, const float* vmap, const float* texture,
                          int num_texels, const float* viewmat, const float* c2w, float fx,
                          float fy, float cx, float cy, int settings, const float* background,
                          float* out_img, float* out_depth, float* out_reg, float* out_texture,
                          float* out_normal, float* final_Ts, int32_t* final_idx,
                          int32_t* depth_idx, float* out_reg_s, void* workspace,
                          size_t workspace_bytes, void* stream) {
    (void)num_charts;
    int rc = check_raster_common(tiles_x, tiles_y, block_width, img_width, img_height, num_probs,
                                 channels, num_points, num_ids, num_texels, settings);
    if (rc) return rc;
    REQUIRE(out_img && out_depth && out_reg && out_normal && final_Ts && final_idx && depth_idx &&
                out_reg_s && (channels == 0 || out_texture),
            "null output pointer");
    REQUIRE(tile_bins && viewmat && c2w && background, "null pointer argument");
    REQUIRE(num_ids == 0 || num_points > 0, "gaussian_ids_sorted is non-empty but num_points is 0");
    REQUIRE(num_ids == 0 || gaussian_ids_sorted, "null pointer argument");
    if (num_points > 0)
        REQUIRE(texture_dims && colors && opacities && means && scales && quats && uv0 && umap &&
                    vmap && (channels == 0 || texture),
                "null pointer argument");
    REQUIRE(workspace_bytes >= gstex::raster_forward_workspace_bytes(num_points),
            "workspace too small: %zu < %zu", workspace_bytes,
            gstex::raster_forward_workspace_bytes(num_points));
    gstex::RasterCall c{img_width, img_height, block_width, channels, settings, num_points, num_ids,
                        num_texels, glob_scale, fx, fy, cx, cy, texture_dims, gaussian_ids_sorted,
                        tile_bins, colors, opacities, means, scales, quats, uv0, umap, vmap,
                        texture, viewmat, c2w, background};
    return check_hip(gstex::raster_forward(c, out_img, out_depth, out_reg, out_texture, out_normal,
                                           final_Ts, final_idx, depth_idx, out_reg_s, workspace,
                                           workspace_bytes, as_stream(stream)),
                     "texture_forward");
}

int gstex_texture_backward(int img_height, int img_width, int block_width, int num_charts,
                           int num_probs, int channels, const int32_t* texture_dims,
                           const int32_t* gaussian_ids_sorted, int num_ids,
                           const int32_t* tile_bins, const float* colors, const float* opacities,
                           const float* means, const float* scales, int num_points,
                           float glob_scale, const float* quats, const float* uv0,
                           const float* umap, const float* vmap, const float* texture,
                           int num_texels, const float* viewmat, const float* c2w, float fx,
                           float fy, float cx, float cy, int settings, const float* background,
                           const float* final_Ts, const int32_t* final_idx,
                           const int32_t* depth_idx, const float* final_s, const float* v_output,
                           const float* v_output_depth, const float* v_output_reg,
                           const float* v_output_alpha, const float* v_output_texture,
                           const float* v_output_normal, float* v_colors, float* v_opacity,
                           float* v_means, float* v_scales, float* v_quats, float* v_uv0,
                           float* v_umap, float* v_vmap, float* v_texture, void* workspace,
                           size_t workspace_bytes, void* stream) {
    (void)num_charts;
    if (block_width < 2 || block_width > 16)
        return fail(GSTEX_ERR_ARG, "block_width must be between 2 and 16, got %d", block_width);
    const int tiles_x = tiles_of(img_width, block_width), tiles_y = tiles_of(img_height, block_width);
    int rc = check_raster_common(tiles_x, tiles_y, block_width, img_width, img_height, num_probs,
                                 channels, num_points, num_ids, num_texels, settings);
    if (rc) return rc;
    REQUIRE(tile_bins && viewmat && c2w && background && final_Ts && final_idx && depth_idx &&
                final_s && v_output && v_output_depth && v_output_reg && v_output_alpha &&
                v_output_normal && (channels == 0 || v_output_texture),
            "null pointer argument");
    REQUIRE(num_ids == 0 || num_points > 0, "gaussian_ids_sorted is non-empty but num_points is 0");
    REQUIRE(num_ids == 0 || gaussian_ids_sorted, "null pointer argument");
    if (num_points > 0)
        REQUIRE(texture_dims && colors && opacities && means && scales &&
                    quats && uv0 && umap && vmap && v_colors && v_opacity && v_means && v_scales &&
                    v_quats && v_uv0 && v_umap && v_vmap && (channels == 0 || (texture && v_texture)),
                "null pointer argument");
    REQUIRE(workspace_bytes >= gstex::raster_backward_workspace_bytes(num_points),
            "workspace too small: %zu < %zu", workspace_bytes,
            gstex::raster_backward_workspace_bytes(num_points));
    gstex::RasterCall c{img_width, img_height, block_width, channels, settings, num_points, num_ids,
                        num_texels, glob_scale, fx, fy, cx, cy, texture_dims, gaussian_ids_sorted,
                        tile_bins, colors, opacities, means, scales, quats, uv0, umap, vmap,
                        texture, viewmat, c2w, background};
    return check_hip(gstex::raster_backward(c, final_Ts, final_idx, depth_idx, final_s, v_output,
                                            v_output_depth, v_output_reg, v_output_alpha,
                                            v_output_texture, v_output_normal, v_colors, v_opacity,
                                            v_means, v_scales, v_quats, v_uv0, v_umap, v_vmap,
                                            v_texture, workspace, workspace_bytes,
                                            as_stream(stream)),
                     "texture_backward");
}

int gstex_set_raster_culling(int enabled) { return gstex::set_raster_culling(enabled != 0) ? 1 : 0; }

int gstex_set_precise_edge_on(int enabled) {
    return gstex::set_raster_precise_edge(enabled != 0) ? 1 : 0;
}

size_t gstex_texture_edit_workspace_size(int num_points) {
    return gstex::texture_edit_workspace_bytes(num_points > 0 ? num_points : 0);
}

int gstex_texture_edit(int tiles_x, int tiles_y, int block_width, int img_width, int img_height,
                       int num_charts, int num_probs, int channels, int texture_total_size,
                       const int32_t* texture_dims, const float* updated_img,
                       const float* updated_alpha, const float* depth_lower,
                       const float* depth_upper, const int32_t* gaussian_ids_sorted, int num_ids,
                       const int32_t* tile_bins, const float* opacities, const float* means,
                       const float* scales, int num_points, float glob_scale, const float* quats,
                       const float* uv0, const float* umap, const float* vmap,
                       const float* viewmat, const float* c2w, float fx, float fy, float cx,
                       float cy, int settings, const float* background, float* updated_texture,
                       void* workspace, size_t workspace_bytes, void* stream) {
    (void)num_charts;
    (void)background;
    REQUIRE(channels >= 5,
            "texture_edit writes 5 channels (r*a, g*a, b*a, a, weight): texture_info[2] must be "
            ">= 5, got %d",
            channels);
    REQUIRE(texture_total_size > 0, "texture_total_size must be positive, got %d",
            texture_total_size);
    int rc = check_raster_common(tiles_x, tiles_y, block_width, img_width, img_height, num_probs,
                                 channels, num_points, num_ids, texture_total_size, 0);
    if (rc) return rc;
    REQUIRE(updated_texture && updated_img && updated_alpha && depth_lower && depth_upper &&
                tile_bins && viewmat && c2w,
            "null pointer argument");
    REQUIRE(num_ids == 0 || (num_points > 0 && gaussian_ids_sorted), "null pointer argument");
    if (num_points > 0)
        REQUIRE(texture_dims && opacities && means && scales && quats && uv0 && umap && vmap,
                "null pointer argument");
    REQUIRE(workspace_bytes >= gstex::texture_edit_workspace_bytes(num_points),
            "workspace too small: %zu < %zu", workspace_bytes,
            gstex::texture_edit_workspace_bytes(num_points));
    gstex::RasterCall c{img_width, img_height, block_width, channels, settings, num_points, num_ids,
                        texture_total_size, glob_scale, fx, fy, cx, cy, texture_dims,
                        gaussian_ids_sorted, tile_bins, nullptr, opacities, means, scales, quats,
                        uv0, umap, vmap, nullptr, viewmat, c2w, nullptr};
    return check_hip(gstex::texture_edit(c, updated_img, updated_alpha, depth_lower, depth_upper,
                                         updated_texture, workspace, workspace_bytes,
                                         as_stream(stream)),
                     "texture_edit");
}

size_t gstex_splat_records_size(int num_points) {
    return gstex::splat_records_bytes(num_points);
}

int gstex_build_splat_records(int num_points, const float* means, const float* scales,
                              float glob_scale, const float* quats, const float* colors,
                              const float* opacities, const float* uv0, const float* umap,
                              const float* vmap, const int32_t* texture_dims,
                              const float* viewmat, const float* c2w, float fx, float fy,
                              float cx, float cy, void* records, void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(means && scales && quats && colors && opacities && uv0 && umap && vmap &&
                texture_dims && viewmat && c2w && records,
            "null pointer argument");
    REQUIRE(((uintptr_t)records & 15) == 0, "records must be 16-byte aligned");
    return check_hip(gstex::launch_build_records(num_points, means, scales, glob_scale, quats,
                                                 colors, opacities, uv0, umap, vmap, texture_dims,
                                                 viewmat, c2w, fx, fy, cx, cy, (float*)records,
                                                 as_stream(stream)),
                     "build_splat_records");
}

static gstex::RasterCall stage_call(int img_width, int img_height, int block_width, int channels,
                                    int settings, int num_points, const int32_t* ids, int num_ids,
                                    const int32_t* tile_bins, const void* texture,
                                    int num_texels, int texture_format, const float* viewmat,
                                    const float* c2w, float fx, float fy, float cx, float cy,
                                    const float* background, uint64_t* wave_masks) {
    return gstex::RasterCall{img_width, img_height, block_width, channels, settings, num_points,
                             num_ids, num_texels, 1.f, fx, fy, cx, cy, nullptr, ids, tile_bins,
                             nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr, texture, viewmat, c2w, background, wave_masks,
                             texture_format == GSTEX_TEXTURE_BF16};
}

size_t gstex_raster_wave_masks_size(int num_ids, int tiles_x, int tiles_y) {
    if (num_ids < 0 || tiles_x <= 0 || tiles_y <= 0) return 0;
    return gstex::kMaxTileWaves * gstex::raster_wave_mask_words(num_ids, tiles_x * tiles_y) *
           sizeof(uint64_t);
}

int gstex_raster_forward(int img_width, int img_height, int block_width, int channels,
                         int settings, const void* records, int num_points,
                         const int32_t* gaussian_ids_sorted, int num_ids,
                         const int32_t* tile_bins, const void* texture, int num_texels,
                         int texture_format,
                         const float* viewmat, const float* c2w, float fx, float fy, float cx,
                         float cy, const float* background, float* out_img, float* out_depth,
                         float* out_reg, float* out_texture, float* out_normal, float* final_Ts,
                         int32_t* final_idx, int32_t* depth_idx, float* out_reg_s,
                         uint64_t* wave_masks, size_t wave_masks_bytes, void* stream) {
    const int tx = img_width > 0 && block_width > 0 ? tiles_of(img_width, block_width) : 0;
    const int ty = img_height > 0 && block_width > 0 ? tiles_of(img_height, block_width) : 0;
    int rc = check_raster_common(tx, ty, block_width, img_width, img_height, 1, channels,
                                 num_points, num_ids, num_texels, settings);
    if (rc) return rc;
    REQUIRE(tile_bins && viewmat && c2w && background && (num_ids == 0 || gaussian_ids_sorted) &&
                (num_points == 0 || records) && (channels == 0 || texture),
            "null pointer argument");
    REQUIRE(out_img && out_depth && out_reg && out_normal && final_Ts && final_idx && depth_idx &&
                out_reg_s && (channels == 0 || out_texture),
            "null output pointer");
    REQUIRE(num_ids == 0 || num_points > 0, "gaussian_ids_sorted is non-empty but num_points is 0");
    REQUIRE(!wave_masks || wave_masks_bytes >= gstex_raster_wave_masks_size(num_ids, tx, ty),
            "wave_masks_bytes is below gstex_raster_wave_masks_size");
    REQUIRE(texture_format == GSTEX_TEXTURE_F32 || texture_format == GSTEX_TEXTURE_BF16,
            "texture_format must be GSTEX_TEXTURE_F32 (0) or GSTEX_TEXTURE_BF16 (1), got %d",
            texture_format);
    const gstex::RasterCall c =
        stage_call(img_width, img_height, block_width, channels, settings, num_points,
                   gaussian_ids_sorted, num_ids, tile_bins, texture, num_texels, texture_format,
                   viewmat, c2w, fx,
                   fy, cx, cy, background, wave_masks);
    return check_hip(gstex::raster_forward_stage(c, (const float*)records, out_img, out_depth,
                                                 out_reg, out_texture, out_normal, final_Ts,
                                                 final_idx, depth_idx, out_reg_s,
                                                 as_stream(stream)),
                     "raster_forward");
}

static int raster_backward_impl(int img_width, int img_height, int block_width, int channels,
                                int settings, const void* records, int num_points,
                                const int32_t* gaussian_ids_sorted, int num_ids,
                                const int32_t* tile_bins, const void* texture, int num_texels,
                                int texture_format, const float* viewmat, const float* c2w,
                                float fx, float fy, float cx, float cy, const float* background,
                                const float* final_Ts, const int32_t* final_idx,
                                const int32_t* depth_idx, const float* final_s,
                                const float* v_output, const float* v_output_depth,
                                const float* v_output_reg, const float* v_output_alpha,
                                const float* v_output_texture, const float* v_output_normal,
                                float* grad_records, float* v_texture,
                                const uint64_t* wave_masks, size_t wave_masks_bytes,
                                void* stream, const float* loss_target, float loss_scale) {
    const int tx = img_width > 0 && block_width > 0 ? tiles_of(img_width, block_width) : 0;
    const int ty = img_height > 0 && block_width > 0 ? tiles_of(img_height, block_width) : 0;
    int rc = check_raster_common(tx, ty, block_width, img_width, img_height, 1, channels,
                                 num_points, num_ids, num_texels, settings);
    if (rc) return rc;
    REQUIRE(tile_bins && viewmat && c2w && background && (num_ids == 0 || gaussian_ids_sorted) &&
                (num_points == 0 || (records && grad_records)) &&
                (channels == 0 || (texture && v_texture && v_output_texture)),
            "null pointer argument");
    REQUIRE(final_Ts && final_idx && depth_idx && final_s && v_output && v_output_depth &&
                v_output_reg && v_output_alpha && v_output_normal,
            "null pointer argument");
    REQUIRE(num_ids == 0 || num_points > 0, "gaussian_ids_sorted is non-empty but num_points is 0");
    REQUIRE(!wave_masks || wave_masks_bytes >= gstex_raster_wave_masks_size(num_ids, tx, ty),
            "wave_masks_bytes is below gstex_raster_wave_masks_size");
    REQUIRE(texture_format == GSTEX_TEXTURE_F32 || texture_format == GSTEX_TEXTURE_BF16,
            "texture_format must be GSTEX_TEXTURE_F32 (0) or GSTEX_TEXTURE_BF16 (1), got %d",
            texture_format);
    const gstex::RasterCall c =
        stage_call(img_width, img_height, block_width, channels, settings, num_points,
                   gaussian_ids_sorted, num_ids, tile_bins, texture, num_texels, texture_format,
                   viewmat, c2w, fx,
                   fy, cx, cy, background, const_cast<uint64_t*>(wave_masks));
    return check_hip(gstex::raster_backward_stage(c, (const float*)records, final_Ts, final_idx,
                                                  depth_idx, final_s, v_output, v_output_depth,
                                                  v_output_reg, v_output_alpha, v_output_texture,
                                                  v_output_normal, grad_records, v_texture,
                                                  as_stream(stream), loss_target, loss_scale),
                     "raster_backward");
}

int gstex_raster_backward(int img_width, int img_height, int block_width, int channels,
                          int settings, const void* records, int num_points,
                          const int32_t* gaussian_ids_sorted, int num_ids,
                          const int32_t* tile_bins, const void* texture, int num_texels,
                          int texture_format,
                          const float* viewmat, const float* c2w, float fx, float fy, float cx,
                          float cy, const float* background, const float* final_Ts,
                          const int32_t* final_idx, const int32_t* depth_idx,
                          const float* final_s, const float* v_output,
                          const float* v_output_depth, const float* v_output_reg,
                          const float* v_output_alpha, const float* v_output_texture,
                          const float* v_output_normal, float* grad_records, float* v_texture,
                          const uint64_t* wave_masks, size_t wave_masks_bytes,
                          void* stream) {
    return raster_backward_impl(img_width, img_height, block_width, channels, settings, records,
                                num_points, gaussian_ids_sorted, num_ids, tile_bins, texture,
                                num_texels, texture_format, viewmat, c2w, fx, fy, cx, cy,
                                background, final_Ts, final_idx, depth_idx, final_s, v_output,
                                v_output_depth, v_output_reg, v_output_alpha, v_output_texture,
                                v_output_normal, grad_records, v_texture, wave_masks,
                                wave_masks_bytes, stream, nullptr, 0.f);
}

int gstex_raster_backward_l2(int img_width, int img_height, int block_width, int channels,
                             int settings, const void* records, int num_points,
                             const int32_t* gaussian_ids_sorted, int num_ids,
                             const int32_t* tile_bins, const void* texture, int num_texels,
                             int texture_format, const float* viewmat, const float* c2w,
                             float fx, float fy, float cx, float cy, const float* background,
                             const float* final_Ts, const int32_t* final_idx,
                             const int32_t* depth_idx, const float* final_s,
                             const float* out_img, const float* out_texture,
                             const float* target, float scale, const float* v_output_depth,
                             const float* v_output_reg, const float* v_output_alpha,
                             const float* v_output_normal, float* grad_records,
                             float* v_texture, const uint64_t* wave_masks,
                             size_t wave_masks_bytes, void* stream) {
    REQUIRE(channels == 3, "gstex_raster_backward_l2 needs 3 channels, got %d", channels);
    REQUIRE(out_img && out_texture && target, "null pointer argument");
    return raster_backward_impl(img_width, img_height, block_width, channels, settings, records,
                                num_points, gaussian_ids_sorted, num_ids, tile_bins, texture,
                                num_texels, texture_format, viewmat, c2w, fx, fy, cx, cy,
                                background, final_Ts, final_idx, depth_idx, final_s, out_img,
                                v_output_depth, v_output_reg, v_output_alpha, out_texture,
                                v_output_normal, grad_records, v_texture, wave_masks,
                                wave_masks_bytes, stream, target, scale);
}

int gstex_finalize_gradients(int num_points, const float* grad_records, const void* records,
                             const float* quats, const float* scales, const float* viewmat,
                             const float* c2w, float fx, float fy, float* v_colors,
                             float* v_opacity, float* v_means, float* v_scales, float* v_quats,
                             float* v_uv0, float* v_umap, float* v_vmap, void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(grad_records && records && scales && quats && viewmat && c2w && v_colors &&
                v_opacity && v_means && v_scales && v_quats && v_uv0 && v_umap && v_vmap,
            "null pointer argument");
    return check_hip(gstex::grad_finalize(num_points, grad_records, (const float*)records, quats,
                                          scales, viewmat, c2w, fx, fy, v_colors, v_opacity,
                                          v_means, v_scales, v_quats, v_uv0, v_umap, v_vmap,
                                          as_stream(stream)),
                     "finalize_gradients");
}

int gstex_finalize_gradients_sh(int num_points, const float* grad_records, const void* records,
                                const float* quats, const float* scales, const float* viewmat,
                                const float* c2w, float fx, float fy, const float* means,
                                int sh_degree, int degrees_to_use, float* v_colors,
                                float* v_opacity, float* v_means, float* v_scales,
                                float* v_quats, float* v_uv0, float* v_umap, float* v_vmap,
                                float* v_sh_coeffs, void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    REQUIRE(sh_degree >= 0 && sh_degree <= 3,
            "gstex_finalize_gradients_sh takes SH degrees 0..3, got %d", sh_degree);
    REQUIRE(degrees_to_use >= 0 && degrees_to_use <= sh_degree,
            "degrees_to_use must be in 0..%d, got %d", sh_degree, degrees_to_use);
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(grad_records && records && scales && quats && viewmat && c2w && means && v_colors &&
                v_opacity && v_means && v_scales && v_quats && v_uv0 && v_umap && v_vmap &&
                v_sh_coeffs,
            "null pointer argument");
    const int nb = (sh_degree + 1) * (sh_degree + 1);
    REQUIRE(gstex::grad_finalize_sh_supported(nb, v_sh_coeffs),
            "v_sh_coeffs must be 16-byte aligned");
    return check_hip(gstex::grad_finalize_sh(num_points, grad_records, (const float*)records,
                                             quats, scales, viewmat, c2w, fx, fy, v_colors,
                                             v_opacity, v_means, v_scales, v_quats, v_uv0,
                                             v_umap, v_vmap, means, nb, degrees_to_use,
                                             v_sh_coeffs, as_stream(stream)),
                     "finalize_gradients_sh");
}

int gstex_finalize_gradients_views(int num_points, int num_views, const float* grad_records,
                                   const void* records, size_t records_stride,
                                   const float* quats, const float* scales,
                                   const float* viewmats, const float* c2ws,
                                   const float* focals, float* v_colors, float* v_opacity,
                                   float* v_means, float* v_scales, float* v_quats,
                                   float* v_uv0, float* v_umap, float* v_vmap,
                                   float* colors_views, void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    REQUIRE(num_views >= 1, "num_views must be >= 1, got %d", num_views);
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(records_stride >= gstex::splat_records_bytes(num_points) && records_stride % 16 == 0,
            "records_stride must be a multiple of 16 of at least gstex_splat_records_size(N)");
    REQUIRE(grad_records && records && scales && quats && viewmats && c2ws && focals &&
                v_colors && v_opacity && v_means && v_scales && v_quats && v_uv0 && v_umap &&
                v_vmap,
            "null pointer argument");
    return check_hip(gstex::grad_finalize_views(num_points, num_views, grad_records, records,
                                                records_stride, quats, scales, viewmats, c2ws,
                                                focals, v_colors, v_opacity, v_means, v_scales,
                                                v_quats, v_uv0, v_umap, v_vmap, colors_views,
                                                as_stream(stream)),
                     "finalize_gradients_views");
}

int gstex_fill_f32(float* dst, size_t n, float v, void* stream) {
    if (n == 0) return GSTEX_OK;
    REQUIRE(dst, "null pointer argument");
    return check_hip(gstex::fill_f32(dst, n, v, as_stream(stream)), "fill_f32");
}

int gstex_texture_sample_forward(int num_charts, int num_probs, int channels,
                                 const int32_t* texture_dims, const float* uvs, int num_queries,
                                 const float* texture, int num_texels, float* output,
                                 void* stream) {
    (void)num_charts;
    (void)num_probs;
    REQUIRE(num_queries >= 0 && channels >= 0 && num_texels >= 0, "negative sizes");
    if (num_queries == 0 || channels == 0) return GSTEX_OK;
    REQUIRE(num_texels > 0, "texture has no rows");
    REQUIRE(texture_dims && uvs && texture && output, "null pointer argument");
    return check_hip(gstex::sample_forward(num_queries, channels, num_texels, texture_dims, uvs,
                                           texture, output, as_stream(stream)),
                     "texture_sample_forward");
}

size_t gstex_texture_sample_backward_workspace_size(int num_queries) {
    return gstex::sample_backward_workspace_bytes(num_queries);
}

int gstex_texture_sample_backward(int num_charts, int num_probs, int channels,
                                  const int32_t* texture_dims, const float* uvs, int num_queries,
                                  const float* texture, int num_texels, const float* v_output,
                                  float* v_texture, void* workspace, size_t workspace_bytes,
                                  void* stream) {
    (void)num_charts;
    (void)num_probs;
    (void)texture;
    REQUIRE(num_queries >= 0 && channels >= 0 && num_texels >= 0, "negative sizes");
    if (num_texels == 0 || channels == 0) return GSTEX_OK;
    REQUIRE(v_texture && (num_queries == 0 || (texture_dims && uvs && v_output)),
            "null pointer argument");
    REQUIRE(num_queries == 0 || workspace, "null workspace");
    REQUIRE(workspace_bytes >= gstex::sample_backward_workspace_bytes(num_queries),
            "workspace too small: %zu < %zu", workspace_bytes,
            gstex::sample_backward_workspace_bytes(num_queries));
    return check_hip(gstex::sample_backward(num_queries, channels, num_texels, texture_dims, uvs,
                                            v_output, v_texture, workspace, workspace_bytes,
                                            as_stream(stream)),
                     "texture_sample_backward");
}

int gstex_preprocess_gaussians(int num_points, const float* log_scales, const float* quats_raw,
                               const float* mapping, const float* rgb_logits,
                               const float* opacity_logits, float* scales, float* quats,
                               float* uv0, float* umap, float* vmap, float* colors,
                               float* opacities, void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(log_scales && quats_raw && mapping && opacity_logits && scales && quats && uv0 &&
                umap && vmap && opacities && (!rgb_logits || colors),
            "null pointer argument");
    REQUIRE(((uintptr_t)quats_raw % 16) == 0 && ((uintptr_t)mapping % 16) == 0 &&
                ((uintptr_t)quats % 16) == 0,
            "quats and mapping must be 16-byte aligned");
    return check_hip(gstex::preprocess_gaussians(num_points, log_scales, quats_raw, mapping,
                                                 rgb_logits, opacity_logits, scales, quats, uv0,
                                                 umap, vmap, colors, opacities,
                                                 as_stream(stream)),
                     "preprocess_gaussians");
}

int gstex_preprocess_gaussians_backward(
    int num_points, const float* log_scales, const float* quats_raw, const float* mapping,
    const float* rgb_logits, const float* opacity_logits, const float* v_scales,
    const float* v_quats, const float* v_uv0, const float* v_umap, const float* v_vmap,
    const float* v_colors, const float* v_opacities, float* v_log_scales, float* v_quats_raw,
    float* v_mapping, float* v_rgb_logits, float* v_opacity_logits, void* stream) {
    REQUIRE(num_points >= 0, "num_points must be >= 0");
    if (num_points == 0) return GSTEX_OK;
    REQUIRE(log_scales && quats_raw && mapping && opacity_logits && v_scales && v_quats &&
                v_uv0 && v_umap && v_vmap && v_opacities && v_log_scales && v_quats_raw &&
                v_opacity_logits && (!rgb_logits || (v_colors && v_rgb_logits)),
            "null pointer argument");
    REQUIRE(((uintptr_t)quats_raw % 16) == 0 && ((uintptr_t)mapping % 16) == 0 &&
                ((uintptr_t)v_quats % 16) == 0 && ((uintptr_t)v_quats_raw % 16) == 0 &&
                ((uintptr_t)v_mapping % 16) == 0,
            "quats, mapping and their gradients must be 16-byte aligned");
    return check_hip(gstex::preprocess_gaussians_backward(
                         num_points, log_scales, quats_raw, mapping, rgb_logits, opacity_logits,
                         v_scales, v_quats, v_uv0, v_umap, v_vmap, v_colors, v_opacities,
                         v_log_scales, v_quats_raw, v_mapping, v_rgb_logits, v_opacity_logits,
                         as_stream(stream)),
                     "preprocess_gaussians_backward");
}

int gstex_image_loss_grad(int num_pixels, int channels, int with_img, const float* out_img,
                          const float* out_texture, const float* target, float scale,
                          const float* out_normal, float normal_scale, float* v_img,
                          float* v_texture_out, float* v_normal, void* stream) {
    REQUIRE(num_pixels >= 0, "num_pixels must be >= 0");
    REQUIRE(channels >= 1 && channels <= 64, "channels must be in [1, 64], got %d", channels);
    if (num_pixels == 0) return GSTEX_OK;
    REQUIRE(out_texture && target && v_img && v_texture_out && (out_img || !with_img),
            "null pointer argument");
    REQUIRE(!out_normal || v_normal, "v_normal is required with out_normal");
    return check_hip(gstex::image_loss_grad(num_pixels, channels, with_img != 0, out_img,
                                            out_texture, target, scale, out_normal,
                                            normal_scale, v_img, v_texture_out, v_normal,
                                            as_stream(stream)),
                     "image_loss_grad");
}

int gstex_adam_step(size_t n, float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                    double lr, double beta1, double beta2, double eps, int step, int grad_mode,
                    void* values, int values_format, void* stream) {
    if (n == 0) return GSTEX_OK;
    REQUIRE(param && grad && exp_avg && exp_avg_sq, "null pointer argument");
    REQUIRE(step >= 1, "step must be >= 1 (the first step is 1), got %d", step);
    REQUIRE(beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0,
            "betas must be in [0, 1)");
    REQUIRE(grad_mode == GSTEX_GRAD_PLAIN || grad_mode == GSTEX_GRAD_SIGMOID,
            "grad_mode must be GSTEX_GRAD_PLAIN or GSTEX_GRAD_SIGMOID");
    REQUIRE(values_format == GSTEX_TEXTURE_F32 || values_format == GSTEX_TEXTURE_BF16,
            "values_format must be GSTEX_TEXTURE_F32 or GSTEX_TEXTURE_BF16");
    return check_hip(gstex::adam_step(n, param, grad, exp_avg, exp_avg_sq, lr, beta1, beta2, eps,
                                      step, grad_mode, values,
                                      values_format == GSTEX_TEXTURE_BF16, as_stream(stream)),
                     "adam_step");
}

}  // extern "C"

// ---------------------------------------------------------------- multi-view raster pass
namespace {

struct ViewsHost {
    int num_views, num_items, channels, settings, bf16;
};
std::mutex g_views_mu;
std::unordered_map<const void*, ViewsHost> g_views;

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
size_t views_hdr_bytes() { return align256(sizeof(gstex::ViewsHeader)); }
size_t views_table_bytes(int v) { return align256((size_t)v * sizeof(gstex::ViewRaster)); }

int persistent_blocks() {
    static int blocks[256] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 256) return 2048;
    if (blocks[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        // more workgroups than can be resident: the extra ones start once
        // others leave and find the queue drained
        blocks[dev] = cus * 8;
    }
    return blocks[dev];
}

int views_lookup(const void* table, ViewsHost& h) {
    std::lock_guard<std::mutex> lock(g_views_mu);
    auto it = g_views.find(table);
    REQUIRE(it != g_views.end(), "views table %p was not set up by gstex_views_setup", table);
    h = it->second;
    return GSTEX_OK;
}

}  // namespace

extern "C" {

size_t gstex_views_table_size(int num_views, int total_tiles) {
    if (num_views <= 0 || total_tiles < 0) return 0;
    return views_hdr_bytes() + views_table_bytes(num_views) + align256((size_t)total_tiles * 4);
}

int gstex_views_setup(int num_views, const gstex_view* views, int block_width, int channels,
                      int settings, int num_points, const void* texture, int num_texels,
                      int texture_format, const float* background, float* v_texture, void* table,
                      size_t table_bytes, void* stream) {
    REQUIRE(num_views >= 1 && num_views < 2048, "num_views must be in [1, 2047], got %d",
            num_views);
    REQUIRE(views && table && background, "null pointer argument");
    REQUIRE(texture_format == GSTEX_TEXTURE_F32 || texture_format == GSTEX_TEXTURE_BF16,
            "texture_format must be GSTEX_TEXTURE_F32 or GSTEX_TEXTURE_BF16");
    REQUIRE(channels == 0 || texture, "null texture");
    REQUIRE(channels == 0 || v_texture, "null v_texture");
    REQUIRE(((uintptr_t)table % 256) == 0, "the views table must be 256-byte aligned");
    std::vector<gstex::ViewRaster> host((size_t)num_views);
    long long total = 0;
    for (int v = 0; v < num_views; ++v) {
        const gstex_view& w = views[v];
        const int tx = tiles_of(w.img_width > 0 ? w.img_width : 1, block_width > 0 ? block_width : 1);
        const int ty = tiles_of(w.img_height > 0 ? w.img_height : 1, block_width > 0 ? block_width : 1);
        int rc = check_raster_common(tx, ty, block_width, w.img_width, w.img_height, 1, channels,
                                     num_points, w.num_ids, num_texels, settings);
        if (rc) {
            char msg[sizeof(g_err)];
            snprintf(msg, sizeof(msg), "%s", g_err);
            return fail(rc, "view %d: %s", v, msg);
        }
        REQUIRE((long long)tx * ty < (1ll << gstex::kViewItemTileBits),
                "view %d: %d x %d tiles exceed the work-item packing", v, tx, ty);
        REQUIRE(w.viewmat && w.c2w && w.tile_bins && w.out_img && w.out_depth && w.out_reg &&
                    w.out_normal && w.final_Ts && w.final_idx && w.depth_idx && w.out_reg_s &&
                    (channels == 0 || w.out_texture),
                "view %d: null pointer argument", v);
        REQUIRE(w.num_ids == 0 || (w.gaussian_ids_sorted && num_points > 0),
                "view %d: null gaussian_ids_sorted", v);
        REQUIRE(num_points == 0 || w.records, "view %d: null records", v);
        REQUIRE(w.v_img && w.v_depth && w.v_reg && w.v_alpha && w.v_normal && w.grad_records &&
                    (channels == 0 || w.v_texture_out),
                "view %d: null backward input", v);
        const size_t mask_need = (size_t)gstex::kMaxTileWaves *
                                 gstex::raster_wave_mask_words(w.num_ids, tx * ty) * 8;
        REQUIRE(!w.wave_masks || w.wave_masks_bytes >= mask_need,
                "view %d: wave masks of %zu bytes, need %zu", v, w.wave_masks_bytes, mask_need);
        gstex::RasterCall c{w.img_width, w.img_height, block_width, channels, settings,
                            num_points, w.num_ids, num_texels, 1.f, w.fx, w.fy, w.cx, w.cy,
                            nullptr, w.gaussian_ids_sorted, w.tile_bins, nullptr, nullptr,
                            nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, texture,
                            w.viewmat, w.c2w, background,
                            static_cast<uint64_t*>(w.wave_masks),
                            texture_format == GSTEX_TEXTURE_BF16};
        gstex::ViewRaster& r = host[(size_t)v];
        r.a = gstex::make_raster_args(c, static_cast<const float*>(w.records));
        r.o = gstex::FwdOut{w.out_img, w.out_depth, w.out_reg, w.out_texture, w.out_normal,
                            w.final_Ts, w.final_idx, w.depth_idx, w.out_reg_s};
        r.b = gstex::BwdIn{w.final_Ts, w.final_idx, w.depth_idx, w.out_reg_s, w.v_img,
                           w.v_depth, w.v_reg, w.v_alpha, w.v_texture_out, w.v_normal,
                           w.grad_records, v_texture};
        total += (long long)tx * ty;
    }
    REQUIRE(total < (1ll << 30), "too many tiles over all views: %lld", total);
    const size_t need = gstex_views_table_size(num_views, (int)total);
    REQUIRE(table_bytes >= need, "views table too small: %zu < %zu", table_bytes, need);
    gstex::ViewsHeader hdr{num_views, (int)total, gstex::raster_block_threads(block_width) / 64,
                           0};
    char* base = static_cast<char*>(table);
    hipStream_t s = as_stream(stream);
    hipError_t e = hipMemcpyAsync(base, &hdr, sizeof(hdr), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(base + views_hdr_bytes(), host.data(),
                           host.size() * sizeof(gstex::ViewRaster), hipMemcpyHostToDevice, s);
    // the host copies go out of scope on return
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    int rc = check_hip(e, "views_setup");
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_views_mu);
    g_views[table] = ViewsHost{num_views, (int)total, channels, settings,
                               texture_format == GSTEX_TEXTURE_BF16 ? 1 : 0};
    return GSTEX_OK;
}

int gstex_views_release(void* table) {
    std::lock_guard<std::mutex> lock(g_views_mu);
    g_views.erase(table);
    return GSTEX_OK;
}

int gstex_views_raster_forward(void* table, void* stream) {
    ViewsHost h;
    int rc = views_lookup(table, h);
    if (rc) return rc;
    char* base = static_cast<char*>(table);
    auto* hdr = reinterpret_cast<gstex::ViewsHeader*>(base);
    auto* views = reinterpret_cast<const gstex::ViewRaster*>(base + views_hdr_bytes());
    auto* items = reinterpret_cast<int*>(base + views_hdr_bytes() + views_table_bytes(h.num_views));
    hipStream_t s = as_stream(stream);
    rc = check_hip(gstex::views_order(views, hdr, items, s), "views_raster_forward: order");
    if (rc) return rc;
    return check_hip(gstex::raster_forward_views(views, hdr, items, h.channels, h.settings,
                                                 h.bf16 != 0, persistent_blocks(), s),
                     "views_raster_forward");
}

int gstex_views_raster_backward(void* table, void* stream) {
    ViewsHost h;
    int rc = views_lookup(table, h);
    if (rc) return rc;
    char* base = static_cast<char*>(table);
    auto* hdr = reinterpret_cast<gstex::ViewsHeader*>(base);
    auto* views = reinterpret_cast<const gstex::ViewRaster*>(base + views_hdr_bytes());
    auto* items = reinterpret_cast<int*>(base + views_hdr_bytes() + views_table_bytes(h.num_views));
    hipStream_t s = as_stream(stream);
    rc = check_hip(gstex::fill_i32(&hdr->head, 1, 0, s), "views_raster_backward: queue reset");
    if (rc) return rc;
    return check_hip(gstex::raster_backward_views(views, hdr, items, h.channels, h.settings,
                                                  h.bf16 != 0, persistent_blocks(), s),
                     "views_raster_backward");
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free forward
// gstex_texture_forward_noalloc keeps the forward's splat records (its only
// scratch) in this file-scope device array instead of a caller workspace.
// Each device the library runs on holds one copy until the process ends; its
// device address comes from hipGetSymbolAddress, once per device, and only
// that address is handed to kernels (in host code the array's name is a host
// stand-in, not a device pointer).  The records are written in full before
// they are read, so nothing in it needs clearing between calls.
static __device__ __align__(256) unsigned char scratch[GSTEX_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxScratchDevices = 256;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of this device's scratch array (looked up on the first call
// made with the device current).
int scratch_of_current_device(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return check_hip(e, "texture_forward_noalloc: hipGetDevice");
    REQUIRE(dev >= 0 && dev < kMaxScratchDevices, "texture_forward_noalloc: device %d is out of "
            "range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return check_hip(e, "texture_forward_noalloc: hipGetSymbolAddress");
        REQUIRE(p != nullptr, "texture_forward_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GSTEX_OK;
}

}  // namespace

extern "C" {

int gstex_texture_forward_noalloc(
    int tiles_x, int tiles_y, int block_width, int img_width, int img_height, int num_charts,
    int num_probs, int channels, const int32_t* texture_dims, const int32_t* gaussian_ids_sorted,
    int num_ids, const int32_t* tile_bins, const float* colors, const float* opacities,
    const float* means, const float* scales, int num_points, float glob_scale,
    const float* quats, const float* uv0, const float* umap, const float* vmap,
    const float* texture, int num_texels, const float* viewmat, const float* c2w, float fx,
    float fy, float cx, float cy, int settings, const float* background, float* out_img,
    float* out_depth, float* out_reg, float* out_texture, float* out_normal, float* final_Ts,
    int32_t* final_idx, int32_t* depth_idx, float* out_reg_s, void* stream) {
    const size_t per_point = gstex::splat_records_bytes(1);
    const size_t max_points = (GSTEX_NOALLOC_SCRATCH_BYTES - 256) / per_point;
    REQUIRE(num_points >= 0 && (size_t)num_points <= max_points,
            "texture_forward_noalloc: %d Gaussians need %zu bytes of scratch; its fixed scratch "
            "of %zu bytes holds at most %zu Gaussians (use gstex_texture_forward with a "
            "workspace)",
            num_points, gstex::raster_forward_workspace_bytes(num_points > 0 ? num_points : 0),
            (size_t)GSTEX_NOALLOC_SCRATCH_BYTES, max_points);
    void* ws = nullptr;
    const int rc = scratch_of_current_device(&ws);
    if (rc) return rc;
    return gstex_texture_forward(tiles_x, tiles_y, block_width, img_width, img_height, num_charts,
                                 num_probs, channels, texture_dims, gaussian_ids_sorted, num_ids,
                                 tile_bins, colors, opacities, means, scales, num_points,
                                 glob_scale, quats, uv0, umap, vmap, texture, num_texels, viewmat,
                                 c2w, fx, fy, cx, cy, settings, background, out_img, out_depth,
                                 out_reg, out_texture, out_normal, final_Ts, final_idx, depth_idx,
                                 out_reg_s, ws, GSTEX_NOALLOC_SCRATCH_BYTES, stream);
}

}  // extern "C"
