// Host-side launchers of the gfx950 kernels (internal; the exported C ABI is
// in include/gstex_hip.h and capi.hip).  Every launcher enqueues on the given
// stream, allocates nothing, and returns the launch status.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <algorithm>

namespace gstex {

// projection / records
hipError_t launch_get_aabb_2d(int n, const float* means, const float* scales, float glob_scale,
                              const float* quats, const float* viewmat, float fx, float fy,
                              float cx, float cy, float* centers, float* extents,
                              hipStream_t stream);
hipError_t launch_project_fused(int n, const float* means, const float* scales, float glob_scale,
                                const float* quats, const float* viewmat, float fx, float fy,
                                float cx, float cy, int bw, int tiles_x, int tiles_y,
                                float* centers, float* extents, float* depths, int* tiles_hit,
                                uint32_t* total, hipStream_t stream);
hipError_t launch_num_tiles_hit(int n, const float* centers, const float* extents, int bw,
                                int tiles_x, int tiles_y, int* out, hipStream_t stream);
hipError_t launch_project_points(int n, const float* points, const float* viewmat, float fx,
                                 float fy, float cx, float cy, int clip, float* pix,
                                 float* depths, hipStream_t stream);
// Splat records: n records of kRecordFloats floats, followed by n cull boxes
// (uint2: int16 pixel-index ranges x0 | x1 << 16, y0 | y1 << 16).
hipError_t launch_build_records(int n, const float* means, const float* scales, float glob_scale,
                                const float* quats, const float* colors, const float* opacities,
                                const float* uv0, const float* umap, const float* vmap,
                                const int* texture_dims, const float* viewmat, const float* c2w,
                                float fx, float fy, float cx, float cy, float* records,
                                hipStream_t stream);

// SH
hipError_t launch_sh_forward(int n, int degree, int degrees_to_use, const float* viewdirs,
                             const float* coeffs, float* colors, hipStream_t stream);
hipError_t launch_sh_forward_bf16(int n, int degree, int degrees_to_use, const float* viewdirs,
                                  const unsigned short* coeffs, float* colors,
                                  hipStream_t stream);
// SH colours at the directions means - c2w[:3, 3] (float32 or bfloat16 coefficients).
hipError_t launch_sh_forward_means(int n, int degree, int degrees_to_use, const float* means,
                                   const float* c2w, const void* coeffs, bool bf16,
                                   float* colors, hipStream_t stream);
hipError_t launch_sh_backward(int n, int degree, int degrees_to_use, const float* viewdirs,
                              const float* v_colors, float* v_coeffs, hipStream_t stream);
hipError_t launch_sh_backward_views(int n, int views, int degree, int degrees_to_use,
                                    const float* viewdirs, const float* v_colors,
                                    float* v_coeffs, hipStream_t stream);
int sh_num_bases(int degree);

// fills
hipError_t fill_f32(float* p, size_t n, float v, hipStream_t s);
hipError_t fill_i32(int* p, size_t n, int v, hipStream_t s);
hipError_t fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);

// binning
hipError_t map_intersects(int n, int num_isects, const float* centers, const float* extents,
                          const float* depths, const int* cum, int tiles_x, int tiles_y, int bw,
                          bool wrapped, int64_t* keys, int* ids, hipStream_t s);
hipError_t num_tiles_hit_wrapped(int n, const float* centers, const float* extents, int bw,
                                 int* out, hipStream_t s);
size_t sort_isects_workspace_bytes(int n);
hipError_t sort_isects(int n, const int64_t* keys_in, const int* ids_in, int64_t* keys_out,
                       int* ids_out, int end_bit, void* ws, size_t ws_bytes, hipStream_t s);
hipError_t tile_bin_edges(int m, const int64_t* keys_sorted, int num_tiles, int* bins,
                          hipStream_t s);
// Stable LSD radix sort of (uint32 key, uint32 value) pairs over key bits [0, end_bit).
size_t sort_u32_pairs_workspace_bytes(int n);
hipError_t sort_u32_pairs(int n, const uint32_t* keys_in, const uint32_t* vals_in,
                          uint32_t* keys_out, uint32_t* vals_out, int end_bit, void* ws,
                          size_t ws_bytes, hipStream_t s);
size_t bin_gaussians_workspace_bytes(int n, int m);
hipError_t count_intersects(int n, const int* counts, unsigned* out, hipStream_t s);
// Depth-first binning into a list of capacity cap >= the intersect total m =
// sum(max(counts, 0)), which is computed on the device (and copied to *m_out
// when given).  Slots past m hold sentinel entries outside every bin; entries
// past cap are dropped (m > cap on return tells the caller to grow the list).
hipError_t bin_gaussians(int n, int cap, const float* centers, const float* extents,
                         const float* depths, const int* counts, int tiles_x, int tiles_y, int bw,
                         int* ids_sorted, int* tile_bins, unsigned* m_out, void* ws,
                         size_t ws_bytes, hipStream_t s);
// The same in two phases: the depth phase (n only; clears the tile bins),
// then the tile phase, which reads the depth phase's workspace and takes the
// intersect total cap from m_total on the device (nullptr: cap itself).
size_t bin_depth_workspace_bytes(int n);
size_t bin_tiles_workspace_bytes(int cap);
hipError_t bin_depth(int n, const float* centers, const float* extents, const float* depths,
                     const int* counts, int tiles_x, int tiles_y, int bw, int* tile_bins,
                     void* ws, size_t ws_bytes, hipStream_t s);
hipError_t bin_tiles(int n, int cap, const unsigned* m_total, int tiles_x, int tiles_y,
                     const void* depth_ws, size_t depth_ws_bytes, int* ids_sorted,
                     int* tile_bins, void* ws, size_t ws_bytes, hipStream_t s);

// raster
struct RasterCall {
    int img_w, img_h, block_width, channels, settings;
    int num_points, num_ids, num_texels;
    float glob_scale, fx, fy, cx, cy;
    const int* texture_dims;
    const int* ids_sorted;
    const int* tile_bins;
    const float* colors;
    const float* opacities;
    const float* means;
    const float* scales;
    const float* quats;
    const float* uv0;
    const float* umap;
    const float* vmap;
    const void* texture;  // (X, C) texels: float32, or bfloat16 when texture_bf16
    const float* viewmat;
    const float* c2w;
    const float* background;
    // Optional per-wave record of which list entries the forward composited
    // (see raster_wave_mask_words); nullptr disables it.
    uint64_t* wave_masks;
    int texture_bf16;
};

// Wave masks: for every tile wave (at most kMaxTileWaves per tile) one bit per
// entry of the tile's list, set when the forward composited that entry for at
// least one of the wave's pixels.  The backward visits only set bits.  Plane
// w holds wave w of every tile; tile t's entries [r0, r1) occupy words
// (r0 >> 6) + t .. ((r1 - 1) >> 6) + t, so tiles never share a word.
constexpr int kMaxTileWaves = 4;
inline size_t raster_wave_mask_words(int num_ids, int num_tiles) {
    return (size_t)(num_ids >> 6) + (size_t)num_tiles + 1;
}

// Process-wide switch for the per-wave culling by screen boxes (default on).
bool raster_culling_enabled();
bool set_raster_culling(bool on);
// Process-wide switch: float64 ray/plane hits for near-edge-on splats (default off).
bool raster_precise_edge_enabled();
bool set_raster_precise_edge(bool on);
size_t raster_forward_workspace_bytes(int n);
size_t raster_backward_workspace_bytes(int n);
hipError_t raster_forward(const RasterCall& c, float* out_img, float* out_depth, float* out_reg,
                          float* out_texture, float* out_normal, float* final_Ts, int* final_idx,
                          int* depth_idx, float* out_reg_s, void* ws, size_t ws_bytes,
                          hipStream_t s);
hipError_t raster_backward(const RasterCall& c, const float* final_Ts, const int* final_idx,
                           const int* depth_idx, const float* final_s, const float* v_img,
                           const float* v_depth, const float* v_reg, const float* v_alpha,
                           const float* v_texture_out, const float* v_normal, float* v_colors,
                           float* v_opacity, float* v_means, float* v_scales, float* v_quats,
                           float* v_uv0, float* v_umap, float* v_vmap, float* v_texture, void* ws,
                           size_t ws_bytes, hipStream_t s);

// per-stage raster entry points (records built by launch_build_records); with
// loss_target (three channels) the backward takes out_img / out_texture in
// place of v_img / v_texture_out and forms the L2 image-loss cotangents itself
hipError_t raster_forward_stage(const RasterCall& c, const float* records, float* out_img,
                                float* out_depth, float* out_reg, float* out_texture,
                                float* out_normal, float* final_Ts, int* final_idx,
                                int* depth_idx, float* out_reg_s, hipStream_t s);
hipError_t raster_backward_stage(const RasterCall& c, const float* records,
                                 const float* final_Ts, const int* final_idx,
                                 const int* depth_idx, const float* final_s, const float* v_img,
                                 const float* v_depth, const float* v_reg, const float* v_alpha,
                                 const float* v_texture_out, const float* v_normal,
                                 float* grad_rec, float* v_texture, hipStream_t s,
                                 const float* loss_target = nullptr, float loss_scale = 0.f);
hipError_t grad_finalize(int n, const float* grad_rec, const float* records, const float* quats,
                         const float* scales, const float* viewmat, const float* c2w, float fx,
                         float fy, float* v_colors, float* v_opacity, float* v_means,
                         float* v_scales, float* v_quats, float* v_uv0, float* v_umap,
                         float* v_vmap, hipStream_t s);
// grad_finalize followed by the SH adjoint of the colour gradient (what
// compute_sh_backward gives with viewdirs = means - c2w translation), for a
// coefficient stride of nb bases (at most 16) into 16-byte aligned v_coeffs.
bool grad_finalize_sh_supported(int nb, const float* v_coeffs);
hipError_t grad_finalize_sh(int n, const float* grad_rec, const float* records, const float* quats,
                            const float* scales, const float* viewmat, const float* c2w,
                            float fx, float fy, float* v_colors, float* v_opacity,
                            float* v_means, float* v_scales, float* v_quats, float* v_uv0,
                            float* v_umap, float* v_vmap, const float* means, int nb,
                            int degrees_to_use, float* v_coeffs, hipStream_t s);
hipError_t grad_finalize_views(int n, int views, const float* grad_rec, const void* records,
                               size_t rec_stride, const float* quats, const float* scales,
                               const float* viewmats, const float* c2ws, const float* focals,
                               float* v_colors, float* v_opacity, float* v_means, float* v_scales,
                               float* v_quats, float* v_uv0, float* v_umap, float* v_vmap,
                               float* colors_views, hipStream_t s);
// Per-view splat data, one buffer: n records (128 B), n screen ellipses
// (float4 centre x, y, A, B: 16 B), n screen boxes (8 B), n ellipse C values
// (4 B).  Each array starts at a multiple of its element size.
constexpr size_t kRecordBytes = 128;
constexpr size_t kCullConicBytes = 16;
constexpr size_t kCullBoxBytes = 8;
constexpr size_t kCullConicCBytes = 4;
inline size_t splat_records_bytes(int n) {
    return n > 0 ? (size_t)n * (kRecordBytes + kCullConicBytes + kCullBoxBytes + kCullConicCBytes)
                 : 0;
}
__host__ __device__ inline float4* record_conics(float* records, int n) {
    return reinterpret_cast<float4*>(reinterpret_cast<char*>(records) + (size_t)n * kRecordBytes);
}
__host__ __device__ inline const float4* record_conics(const float* records, int n) {
    return record_conics(const_cast<float*>(records), n);
}
__host__ __device__ inline uint2* record_boxes(float* records, int n) {
    return reinterpret_cast<uint2*>(reinterpret_cast<char*>(records) +
                                    (size_t)n * (kRecordBytes + kCullConicBytes));
}
__host__ __device__ inline const uint2* record_boxes(const float* records, int n) {
    return record_boxes(const_cast<float*>(records), n);
}
__host__ __device__ inline float* record_conic_c(float* records, int n) {
    return reinterpret_cast<float*>(reinterpret_cast<char*>(records) +
                                    (size_t)n * (kRecordBytes + kCullConicBytes + kCullBoxBytes));
}
__host__ __device__ inline const float* record_conic_c(const float* records, int n) {
    return record_conic_c(const_cast<float*>(records), n);
}
constexpr size_t kGradRecordBytes = 128;

// training-step glue (example.py:126-143 preprocessing, torch.optim.Adam)
hipError_t preprocess_gaussians(int n, const float* log_scales, const float* quats_raw,
                                const float* mapping, const float* rgb_logits,
                                const float* opacity_logits, float* scales, float* quats,
                                float* uv0, float* umap, float* vmap, float* colors,
                                float* opacities, hipStream_t s);
hipError_t preprocess_gaussians_backward(
    int n, const float* log_scales, const float* quats_raw, const float* mapping,
    const float* rgb_logits, const float* opacity_logits, const float* v_scales,
    const float* v_quats, const float* v_uv0, const float* v_umap, const float* v_vmap,
    const float* v_colors, const float* v_opacities, float* v_log_scales, float* v_quats_raw,
    float* v_mapping, float* v_rgb_logits, float* v_opacity_logits, hipStream_t s);
hipError_t image_loss_grad(int npix, int C, int with_img, const float* out_img,
                           const float* out_tex, const float* target, float scale,
                           const float* out_normal, float normal_scale, float* v_img,
                           float* v_tex, float* v_normal, hipStream_t s);
hipError_t adam_step(size_t n, float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                     double lr, double beta1, double beta2, double eps, int step, int grad_mode,
                     void* values, int values_bf16, hipStream_t s);

// multi-view raster pass (the table types are in raster_common.hpp)
struct ViewRaster;
struct ViewsHeader;
hipError_t views_order(const ViewRaster* views, ViewsHeader* hdr, int* items, hipStream_t s);
hipError_t raster_forward_views(const ViewRaster* views, ViewsHeader* hdr, const int* items,
                                int channels, int settings, bool bf16, int blocks,
                                hipStream_t s);
hipError_t raster_backward_views(const ViewRaster* views, ViewsHeader* hdr, const int* items,
                                 int channels, int settings, bool bf16, int blocks,
                                 hipStream_t s);

// texture edit (texture_edit.cu semantics; out is X x C with C >= 5, zeroed here)
size_t texture_edit_workspace_bytes(int n);
hipError_t texture_edit(const RasterCall& c, const float* upd_img, const float* upd_alpha,
                        const float* depth_lower, const float* depth_upper, float* out,
                        void* ws, size_t ws_bytes, hipStream_t s);

// sampling
hipError_t sample_forward(int q, int C, int num_texels, const int* dims, const float* uvs,
                          const float* texture, float* out, hipStream_t s);
size_t sample_backward_workspace_bytes(int q);
hipError_t sample_backward(int q, int C, int num_texels, const int* dims, const float* uvs,
                           const float* v_out, float* v_texture, void* ws, size_t ws_bytes,
                           hipStream_t s);

}  // namespace gstex
