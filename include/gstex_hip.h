/*
 * gstex_hip — C ABI of the gfx950 textured 2D-Gaussian rasteriser.
 *
 * Every entry point takes device pointers, plain sizes and scalars, and a
 * hipStream_t passed as void* (NULL = the legacy default stream).  Nothing
 * allocates device memory: outputs and workspaces are caller-owned, with
 * workspace sizes given by the *_workspace_size functions.  Each call returns
 * 0 on success or a nonzero code; gstex_last_error() then describes the most
 * recent failure of the calling thread.  Launches are asynchronous.
 *
 * Layouts are the reference's (row-major, contiguous):
 *   means/scales (N,3) f32, quats (N,4) f32 (w,x,y,z), colors (N,3),
 *   opacities (N,1), uv0 (N,1,2), umap/vmap (N,1,3), texture_dims (N,3) i32
 *   {h, w, start}, texture (X,C) f32, viewmat/c2w (4,4) row-major,
 *   images (H,W,k) f32, tile_bins (T,2) i32.
 *
 * Each declaration cites the reference interface it replaces
 * (paths relative to gstex_cuda/ in the reference tree).
 */
#ifndef GSTEX_HIP_H
#define GSTEX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSTEX_OK 0
#define GSTEX_ERR_ARG 1    /* invalid argument (shape, size, null pointer, setting) */
#define GSTEX_ERR_HIP 2    /* a HIP runtime / launch error */
#define GSTEX_ERR_SPACE 3  /* workspace too small */

/* Texel storage formats of the per-stage raster entry points (bfloat16 texels
 * are widened exactly to float32; compositing and gradients stay float32). */
#define GSTEX_TEXTURE_F32 0
#define GSTEX_TEXTURE_BF16 1

/* Description of the last failure on this thread (empty string if none). */
const char *gstex_last_error(void);
/* ABI version of this library. */
int gstex_abi_version(void);

/* Replaces get_aabb_2d_tensor (cuda/csrc/get_aabb_2d.cu:91-126), bound as
 * _C.get_aabb_2d (cuda/csrc/ext.cpp:17).  centers/extents: (N,2) f32. */
int gstex_get_aabb_2d(int num_points, const float *means, const float *scales, float glob_scale,
                      const float *quats, const float *viewmat, float fx, float fy, float cx,
                      float cy, float *centers, float *extents, void *stream);

/* Replaces the torch function get_num_tiles_hit_2d (get_aabb_2d.py:70-92). */
int gstex_num_tiles_hit(int num_points, const float *centers, const float *extents, int img_height,
                        int img_width, int block_width, int32_t *num_tiles_hit, void *stream);

/* Tile count of each Gaussian's wrapped rectangle (0 when max(extent) <= 1e-4),
 * the counts wrapped binning needs; no counterpart in the reference, whose
 * get_num_tiles_hit_2d (get_aabb_2d.py:70-92) clamps to the image. */
int gstex_num_tiles_hit_wrapped(int num_points, const float *centers, const float *extents,
                                int block_width, int32_t *num_tiles_hit, void *stream);

/* Fused caller-side projection (example.py:146-152): get_aabb_2d plus the
 * view-space depth of project_points (get_aabb_2d.py:22-32) plus
 * get_num_tiles_hit_2d (get_aabb_2d.py:70-92), one pass over the Gaussians. */
int gstex_project_gaussians(int num_points, const float *means, const float *scales,
                            float glob_scale, const float *quats, const float *viewmat, float fx,
                            float fy, float cx, float cy, int img_height, int img_width,
                            int block_width, float *centers, float *extents, float *depths,
                            int32_t *num_tiles_hit, void *stream);
/* gstex_project_gaussians plus the total of compute_cumulative_intersects
 * (utils.py:40-59): sum(num_tiles_hit) written to *num_intersects_out (device
 * memory) by the same launch, for callers that read it back to size the
 * list. */
int gstex_project_gaussians_total(int num_points, const float *means, const float *scales,
                                  float glob_scale, const float *quats, const float *viewmat,
                                  float fx, float fy, float cx, float cy, int img_height,
                                  int img_width, int block_width, float *centers, float *extents,
                                  float *depths, int32_t *num_tiles_hit,
                                  uint32_t *num_intersects_out, void *stream);

/* Replaces the torch function project_points (get_aabb_2d.py:22-32): pixel
 * coordinates (N,2) and view-space depth (N,) of points (N,3) under the first
 * three rows of a row-major viewmat; clip != 0 clamps view z to [0.01, 1000]
 * first (and returns the clamped depth). */
int gstex_project_points(int num_points, const float *points, const float *viewmat, float fx,
                         float fy, float cx, float cy, int clip, float *pix, float *depths,
                         void *stream);

/* Replaces compute_sh_forward_tensor (cuda/csrc/bindings.cu:18-43).
 * coeffs (N, bases(degree), 3); degrees_to_use <= degree. */
int gstex_compute_sh_forward(int num_points, int degree, int degrees_to_use, const float *viewdirs,
                             const float *coeffs, float *colors, void *stream);

/* compute_sh_forward with bfloat16 coefficients (N, D, 3), widened exactly to
 * float32; the arithmetic and the colours are float32.  No reference
 * counterpart: bfloat16 storage of SH features. */
int gstex_compute_sh_forward_bf16(int num_points, int degree, int degrees_to_use,
                                  const float *viewdirs, const uint16_t *coeffs, float *colors,
                                  void *stream);

/* compute_sh_forward at the view directions means - c2w[:3, 3] (c2w 4 x 4
 * row-major; the same float32 subtraction as forming the directions first),
 * with float32 (coeffs_format 0) or bfloat16 (1) coefficients, so the
 * directions need no pass of their own.  No reference counterpart (its
 * trainer forms the directions with torch); for training steps with SH
 * colours. */
int gstex_compute_sh_forward_means(int num_points, int degree, int degrees_to_use,
                                   const float *means, const float *c2w, const void *coeffs,
                                   int coeffs_format, float *colors, void *stream);

/* Replaces compute_sh_backward_tensor (cuda/csrc/bindings.cu:45-75).
 * v_coeffs (N, bases(degree), 3) is fully written. */
int gstex_compute_sh_backward(int num_points, int degree, int degrees_to_use,
                              const float *viewdirs, const float *v_colors, float *v_coeffs,
                              void *stream);

/* The sum over num_views views of gstex_compute_sh_backward, for batched
 * multi-view steps.  No counterpart in the reference, whose trainer renders
 * one view per step (example.py:227-255).  viewdirs and v_colors are
 * (num_views, N, 3) stacks; v_coeffs (N, bases(degree), 3) receives
 * sum_v d colours_v / d coeffs, each coefficient written once.  The per-view
 * products are added in view order without fused multiply-add, so the result
 * equals summing the per-view gstex_compute_sh_backward outputs in order. */
int gstex_compute_sh_backward_views(int num_points, int num_views, int degree,
                                    int degrees_to_use, const float *viewdirs,
                                    const float *v_colors, float *v_coeffs, void *stream);

/* Replaces map_gaussian_to_intersects_tensor (cuda/csrc/bindings.cu:77-121).
 * isect_ids (I,) int64 and gaussian_ids (I,) int32 are fully written
 * (zero where a Gaussian counts a slot but writes none).  wrapped != 0 takes the
 * unclamped rectangle of get_tile_bbox_wrapped (helpers.cuh:53-73,94-111) with
 * tile indices modulo the grid (forward.cu:34-62).  Each Gaussian writes at
 * most its own count window [cum[g-1], cum[g]); the reference writes its whole
 * rectangle there, past the window when it holds more tiles than the count.
 * Use gstex_num_tiles_hit_wrapped for counts that hold every wrapped tile. */
int gstex_map_gaussian_to_intersects(int num_points, int num_intersects, const float *centers,
                                     const float *extents, const float *depths,
                                     const int32_t *cum_tiles_hit, int tiles_x, int tiles_y,
                                     int block_width, int wrapped, int64_t *isect_ids,
                                     int32_t *gaussian_ids, void *stream);

/* Replaces torch.sort + torch.gather in bin_and_sort_gaussians (utils.py:159-160):
 * stable ascending LSD radix sort of the int64 keys over bits [0, end_bit)
 * (end_bit <= 64; keys must be non-negative when end_bit < 64), carrying the
 * int32 ids. */
size_t gstex_sort_intersects_workspace_size(int num_intersects);
int gstex_sort_intersects(int num_intersects, int end_bit, const int64_t *isect_ids,
                          const int32_t *gaussian_ids, int64_t *isect_ids_sorted,
                          int32_t *gaussian_ids_sorted, void *workspace, size_t workspace_bytes,
                          void *stream);

/* No counterpart in the reference: the sort that binning and the
 * texture_sample adjoint use, for callers that sort their own 32-bit keys.
 * Stable ascending radix sort of (uint32 key, uint32 value) pairs over key
 * bits [0, end_bit) (end_bit <= 32; higher key bits are carried, not
 * compared), one launch per 8-bit digit plus one histogram launch. */
size_t gstex_sort_pairs_u32_workspace_size(int n);
int gstex_sort_pairs_u32(int n, int end_bit, const uint32_t *keys, const uint32_t *values,
                         uint32_t *keys_sorted, uint32_t *values_sorted, void *workspace,
                         size_t workspace_bytes, void *stream);

/* Replaces get_tile_bin_edges_tensor (cuda/csrc/bindings.cu:123-141).
 * tile_bins (tiles_x*tiles_y, 2) int32 is fully written. */
int gstex_get_tile_bin_edges(int num_intersects, const int64_t *isect_ids_sorted, int tiles_x,
                             int tiles_y, int32_t *tile_bins, void *stream);

/* Replaces the total of compute_cumulative_intersects (utils.py:40-59,
 * num_intersects = cum_tiles_hit[-1]) where only the total is needed:
 * sum(max(num_tiles_hit, 0)) written to *num_intersects_out (device memory);
 * the caller reads it back. */
int gstex_count_intersects(int num_points, const int32_t *num_tiles_hit,
                           uint32_t *num_intersects_out, void *stream);

/* Replaces bin_and_sort_gaussians (utils.py:106-162) where only the sorted ids
 * and tile bins are consumed (texture.py:207-222): identical outputs, computed
 * by sorting the Gaussians by depth once and stable-sorting the emitted
 * intersects by tile.  num_intersects = sum(num_tiles_hit). */
size_t gstex_bin_gaussians_workspace_size(int num_points, int num_intersects);
int gstex_bin_gaussians(int num_points, int num_intersects, const float *centers,
                        const float *extents, const float *depths, const int32_t *num_tiles_hit,
                        int tiles_x, int tiles_y, int block_width, int32_t *gaussian_ids_sorted,
                        int32_t *tile_bins, void *workspace, size_t workspace_bytes, void *stream);
/* gstex_bin_gaussians as two calls (no counterpart in the
 * reference; the same outputs), for a caller that reads the intersect total
 * back to the host: gstex_bin_gaussians_depth needs only the Gaussians (entry
 * counts, the depth sort, the emission offsets) and can be queued before the
 * read-back; gstex_bin_gaussians_tiles then takes num_intersects (the list
 * length) and the depth phase's workspace, unchanged in between, and writes
 * the sorted ids and the bins.  The depth phase clears tile_bins.  When
 * num_intersects_dev is given it holds num_intersects on the device (as
 * written by gstex_project_gaussians_total or gstex_count_intersects), which
 * saves a launch.  Workspaces: gstex_bin_gaussians_depth_workspace_size(N)
 * and gstex_bin_gaussians_tiles_workspace_size(N, num_intersects) bytes. */
size_t gstex_bin_gaussians_depth_workspace_size(int num_points);
size_t gstex_bin_gaussians_tiles_workspace_size(int num_points, int num_intersects);
int gstex_bin_gaussians_depth(int num_points, const float *centers, const float *extents,
                              const float *depths, const int32_t *num_tiles_hit, int tiles_x,
                              int tiles_y, int block_width, int32_t *tile_bins, void *depth_ws,
                              size_t depth_ws_bytes, void *stream);
int gstex_bin_gaussians_tiles(int num_points, int num_intersects,
                              const uint32_t *num_intersects_dev, int tiles_x, int tiles_y,
                              const void *depth_ws, size_t depth_ws_bytes,
                              int32_t *gaussian_ids_sorted, int32_t *tile_bins, void *tiles_ws,
                              size_t tiles_ws_bytes, void *stream);
/* Like gstex_bin_gaussians, with no host-side intersect count (no host sync,
 * so a caller can capture the step in a graph): the list has room for
 * `capacity` entries, the intersect total sum(num_tiles_hit) is computed on
 * the device and written to *num_intersects_out (device memory), slots past it
 * lie outside every bin, and entries past the capacity are dropped -- a total
 * above `capacity` on return means the binning must be redone with a larger
 * list.  Workspace: gstex_bin_gaussians_workspace_size(num_points, capacity). */
int gstex_bin_gaussians_capped(int num_points, int capacity, const float *centers,
                               const float *extents, const float *depths,
                               const int32_t *num_tiles_hit, int tiles_x, int tiles_y,
                               int block_width, int32_t *gaussian_ids_sorted, int32_t *tile_bins,
                               uint32_t *num_intersects_out, void *workspace,
                               size_t workspace_bytes, void *stream);

/* Replaces texture_forward_tensor (cuda/csrc/texture.cu:762-901).  The nine
 * outputs are fully written.  num_ids = length of gaussian_ids_sorted,
 * num_texels = rows of texture, channels = texture_info[2]; num_probs must be 1. */
size_t gstex_texture_forward_workspace_size(int num_points);
int gstex_texture_forward(int tiles_x, int tiles_y, int block_width, int img_width, int img_height,
                          int num_charts, int num_probs, int channels,
                          const int32_t *texture_dims, const int32_t *gaussian_ids_sorted,
                          int num_ids, const int32_t *tile_bins, const float *colors,
                          const float *opacities, const float *means, const float *scales,
                          int num_points, float glob_scale, const float *quats, const float *uv0,
                          const float *umap, const float *vmap, const float *texture,
                          int num_texels, const float *viewmat, const float *c2w, float fx,
                          float fy, float cx, float cy, int settings, const float *background,
                          float *out_img, float *out_depth, float *out_reg, float *out_texture,
                          float *out_normal, float *final_Ts, int32_t *final_idx,
                          int32_t *depth_idx, float *out_reg_s, void *workspace,
                          size_t workspace_bytes, void *stream);

/* ---- Multi-view raster pass ------------------------------------------------
 * No counterpart in the reference, whose trainer renders one view per launch
 * (texture.py:187-193, texture.cu:864).  For a batch of views of one scene
 * (shared Gaussians, texels and settings; each view with its own camera,
 * splat records from gstex_build_splat_records, sorted lists, wave masks,
 * outputs and backward inputs), one forward launch and one backward launch
 * raster every tile of every view.  The (view, tile) work is ordered longest
 * list first and taken by persistent waves from one queue, so the longest
 * tiles of all views start first.  Results equal gstex_raster_forward /
 * gstex_raster_backward per view (texel-gradient and gradient-record sums
 * in another order).
 *
 * gstex_views_setup validates the views and writes their launch table into
 * `table` (gstex_views_table_size bytes, 256-byte aligned, caller-owned);
 * it synchronises the stream and must not be captured.  The table stays
 * valid while every buffer it names does; gstex_views_release forgets it.
 * gstex_views_raster_forward / _backward only launch (graph-capturable);
 * the backward adds into each view's grad_records and into v_texture, which
 * the caller zeroes first, and needs the forward of the same lists before it. */
typedef struct gstex_view {
    int img_width, img_height;
    float fx, fy, cx, cy;
    const float *viewmat, *c2w;        /* (4,4) row-major, device */
    const void *records;               /* gstex_splat_records_size(num_points) bytes */
    const int32_t *gaussian_ids_sorted;
    int num_ids;                       /* length of gaussian_ids_sorted */
    const int32_t *tile_bins;          /* (tiles, 2) */
    void *wave_masks;                  /* optional: gstex_raster_wave_masks_size bytes */
    size_t wave_masks_bytes;
    float *out_img, *out_depth, *out_reg, *out_texture, *out_normal, *final_Ts;
    int32_t *final_idx, *depth_idx;
    float *out_reg_s;
    const float *v_img, *v_depth, *v_reg, *v_alpha, *v_texture_out, *v_normal;
    float *grad_records;               /* (num_points, 32), see gstex_raster_backward */
} gstex_view;
size_t gstex_views_table_size(int num_views, int total_tiles);
int gstex_views_setup(int num_views, const gstex_view *views, int block_width, int channels,
                      int settings, int num_points, const void *texture, int num_texels,
                      int texture_format, const float *background, float *v_texture, void *table,
                      size_t table_bytes, void *stream);
int gstex_views_release(void *table);
int gstex_views_raster_forward(void *table, void *stream);
int gstex_views_raster_backward(void *table, void *stream);

/* No counterpart in the reference.  gstex_texture_forward without a
 * workspace argument, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the splat records live in a fixed device array of
 * GSTEX_NOALLOC_SCRATCH_BYTES inside the library, one per device, held until
 * the process ends.  It calls no allocating routine.  num_points above
 * (GSTEX_NOALLOC_SCRATCH_BYTES - 256) / 156 = 645,275 is rejected with
 * GSTEX_ERR_ARG.  Arguments, outputs and results are those of
 * gstex_texture_forward; the device of `stream` must be current.  Every call
 * on one device shares that device's array, so calls that overlap on one
 * device (other streams or threads) race. */
#define GSTEX_NOALLOC_SCRATCH_BYTES (96u << 20)
int gstex_texture_forward_noalloc(int tiles_x, int tiles_y, int block_width, int img_width,
                                  int img_height, int num_charts, int num_probs, int channels,
                                  const int32_t *texture_dims, const int32_t *gaussian_ids_sorted,
                                  int num_ids, const int32_t *tile_bins, const float *colors,
                                  const float *opacities, const float *means, const float *scales,
                                  int num_points, float glob_scale, const float *quats,
                                  const float *uv0, const float *umap, const float *vmap,
                                  const float *texture, int num_texels, const float *viewmat,
                                  const float *c2w, float fx, float fy, float cx, float cy,
                                  int settings, const float *background, float *out_img,
                                  float *out_depth, float *out_reg, float *out_texture,
                                  float *out_normal, float *final_Ts, int32_t *final_idx,
                                  int32_t *depth_idx, float *out_reg_s, void *stream);

/* Replaces texture_backward_tensor (cuda/csrc/texture.cu:903-1053).  The nine
 * gradient outputs are fully written (v_scales[:,2] = 0). */
size_t gstex_texture_backward_workspace_size(int num_points);
int gstex_texture_backward(int img_height, int img_width, int block_width, int num_charts,
                           int num_probs, int channels, const int32_t *texture_dims,
                           const int32_t *gaussian_ids_sorted, int num_ids,
                           const int32_t *tile_bins, const float *colors, const float *opacities,
                           const float *means, const float *scales, int num_points,
                           float glob_scale, const float *quats, const float *uv0,
                           const float *umap, const float *vmap, const float *texture,
                           int num_texels, const float *viewmat, const float *c2w, float fx,
                           float fy, float cx, float cy, int settings, const float *background,
                           const float *final_Ts, const int32_t *final_idx,
                           const int32_t *depth_idx, const float *final_s, const float *v_output,
                           const float *v_output_depth, const float *v_output_reg,
                           const float *v_output_alpha, const float *v_output_texture,
                           const float *v_output_normal, float *v_colors, float *v_opacity,
                           float *v_means, float *v_scales, float *v_quats, float *v_uv0,
                           float *v_umap, float *v_vmap, float *v_texture, void *workspace,
                           size_t workspace_bytes, void *stream);

/* ---- Per-stage render path ------------------------------------------------
 * texture_forward_tensor / texture_backward_tensor (cuda/csrc/texture.cu:762-1053)
 * split into their stages, so a training step builds the per-Gaussian splat
 * records once for both passes and each kernel can be timed alone:
 *   gstex_build_splat_records -> gstex_raster_forward -> gstex_raster_backward
 *   -> gstex_finalize_gradients.
 * records: gstex_splat_records_size(N) bytes, 256-byte aligned: N 128-byte
 * splat records followed by N 8-byte screen boxes, outside of which a Gaussian
 * cannot pass the alpha test (the raster kernels skip it for waves whose pixels
 * all lie outside; gstex_set_raster_culling(0) turns that off).
 * grad_records: N x 32 float32 accumulators (the raster backward adds into them,
 * so the caller zeroes them and v_texture first). */
size_t gstex_splat_records_size(int num_points);
int gstex_build_splat_records(int num_points, const float *means, const float *scales,
                              float glob_scale, const float *quats, const float *colors,
                              const float *opacities, const float *uv0, const float *umap,
                              const float *vmap, const int32_t *texture_dims,
                              const float *viewmat, const float *c2w, float fx, float fy,
                              float cx, float cy, void *records, void *stream);
/* Process-wide switch for the per-wave culling by screen boxes (1 = on, the
 * default; 0 = every list entry is evaluated).  Results are identical either
 * way; returns the previous setting. */
int gstex_set_raster_culling(int enabled);
/* Process-wide switch (default 0): when 1, the raster and edit kernels solve
 * the ray/plane hit of near-edge-on splats (|n . ray| < 0.05) in float64.
 * This departs from the reference's float32 arithmetic (to the float64
 * oracle's benefit) for those pairs only; returns the previous setting.  No
 * reference counterpart. */
int gstex_set_precise_edge_on(int enabled);
/* Wave masks (optional, for the per-stage forward/backward pair): one bit per
 * (tile wave, list entry), written by gstex_raster_forward and read by
 * gstex_raster_backward, which then visits only the entries the forward
 * composited for that wave.  Size in bytes for a binning with num_ids entries;
 * the forward writes every word the backward reads, so no clearing is needed.
 * Pass the same buffer, ids and tile_bins to both calls, or NULL to both;
 * wave_masks_bytes is the buffer's size (checked against the size above). */
size_t gstex_raster_wave_masks_size(int num_ids, int tiles_x, int tiles_y);
int gstex_raster_forward(int img_width, int img_height, int block_width, int channels,
                         int settings, const void *records, int num_points,
                         const int32_t *gaussian_ids_sorted, int num_ids,
                         const int32_t *tile_bins, const void *texture, int num_texels,
                         int texture_format,
                         const float *viewmat, const float *c2w, float fx, float fy, float cx,
                         float cy, const float *background, float *out_img, float *out_depth,
                         float *out_reg, float *out_texture, float *out_normal, float *final_Ts,
                         int32_t *final_idx, int32_t *depth_idx, float *out_reg_s,
                         uint64_t *wave_masks, size_t wave_masks_bytes, void *stream);
int gstex_raster_backward(int img_width, int img_height, int block_width, int channels,
                          int settings, const void *records, int num_points,
                          const int32_t *gaussian_ids_sorted, int num_ids,
                          const int32_t *tile_bins, const void *texture, int num_texels,
                          int texture_format,
                          const float *viewmat, const float *c2w, float fx, float fy, float cx,
                          float cy, const float *background, const float *final_Ts,
                          const int32_t *final_idx, const int32_t *depth_idx,
                          const float *final_s, const float *v_output,
                          const float *v_output_depth, const float *v_output_reg,
                          const float *v_output_alpha, const float *v_output_texture,
                          const float *v_output_normal, float *grad_records, float *v_texture,
                          const uint64_t *wave_masks, size_t wave_masks_bytes,
                          void *stream);
/* gstex_raster_backward for the L2 image loss mean((out_img + out_texture -
 * target)^2) of a training step: in place of v_output and v_output_texture it
 * takes the forward's out_img and out_texture (H x W x 3 each), the target
 * and scale (2 / (H W 3) for that mean), and forms the image and texture
 * cotangents ((out_img + out_texture) - target) * scale per pixel channel,
 * each step rounded as gstex_image_loss_grad with with_img = 1 rounds it, so
 * that the separate loss pass and its buffers are not needed.  Three
 * channels only.  No counterpart in the reference (its trainer forms the
 * loss gradient with torch autograd, example.py:188-210). */
int gstex_raster_backward_l2(int img_width, int img_height, int block_width, int channels,
                             int settings, const void *records, int num_points,
                             const int32_t *gaussian_ids_sorted, int num_ids,
                             const int32_t *tile_bins, const void *texture, int num_texels,
                             int texture_format, const float *viewmat, const float *c2w,
                             float fx, float fy, float cx, float cy, const float *background,
                             const float *final_Ts, const int32_t *final_idx,
                             const int32_t *depth_idx, const float *final_s,
                             const float *out_img, const float *out_texture,
                             const float *target, float scale, const float *v_output_depth,
                             const float *v_output_reg, const float *v_output_alpha,
                             const float *v_output_normal, float *grad_records,
                             float *v_texture, const uint64_t *wave_masks,
                             size_t wave_masks_bytes, void *stream);
int gstex_finalize_gradients(int num_points, const float *grad_records, const void *records,
                             const float *quats, const float *scales, const float *viewmat,
                             const float *c2w, float fx, float fy, float *v_colors,
                             float *v_opacity, float *v_means, float *v_scales, float *v_quats,
                             float *v_uv0, float *v_umap, float *v_vmap, void *stream);
/* gstex_finalize_gradients followed by gstex_compute_sh_backward(num_points,
 * sh_degree, degrees_to_use, viewdirs, v_colors, v_sh_coeffs) with viewdirs =
 * means - the camera position (c2w[:3, 3], the same float32 subtraction), in
 * one pass: the same outputs, and v_colors is written too.  No counterpart in
 * the reference (whose trainer applies the SH adjoint through torch
 * autograd); for training steps with SH colours.  sh_degree 0..3 (v_sh_coeffs
 * (num_points, (sh_degree + 1)^2, 3), 16-byte aligned). */
int gstex_finalize_gradients_sh(int num_points, const float *grad_records, const void *records,
                                const float *quats, const float *scales, const float *viewmat,
                                const float *c2w, float fx, float fy, const float *means,
                                int sh_degree, int degrees_to_use, float *v_colors,
                                float *v_opacity, float *v_means, float *v_scales,
                                float *v_quats, float *v_uv0, float *v_umap, float *v_vmap,
                                float *v_sh_coeffs, void *stream);
/* The sum over num_views views of gstex_finalize_gradients, for batched
 * multi-view steps (no counterpart in the reference, whose trainer renders one
 * view per step, example.py:227-255).  grad_records holds num_views stacked
 * N x 32 blocks; view v's splat records start records_stride bytes after view
 * v - 1's (records_stride >= gstex_splat_records_size(N), a multiple of 16);
 * viewmats and c2ws are (num_views, 4, 4) row-major, focals (num_views, 2)
 * holds fx, fy; all are device pointers.  The outputs receive the sums over
 * the views, formed as gstex_finalize_gradients forms each view's values and
 * added in view order without fused multiply-add (equal to summing the
 * per-view outputs in order).  colors_views (num_views, N, 3), if not NULL,
 * receives each view's colour gradient, the input of
 * gstex_compute_sh_backward_views. */
int gstex_finalize_gradients_views(int num_points, int num_views, const float *grad_records,
                                   const void *records, size_t records_stride,
                                   const float *quats, const float *scales,
                                   const float *viewmats, const float *c2ws,
                                   const float *focals, float *v_colors, float *v_opacity,
                                   float *v_means, float *v_scales, float *v_quats,
                                   float *v_uv0, float *v_umap, float *v_vmap,
                                   float *colors_views, void *stream);
/* Fills n float32 values with v (a plain kernel, safe inside captured graphs). */
int gstex_fill_f32(float *dst, size_t n, float v, void *stream);

/* Replaces texture_edit_tensor (cuda/csrc/texture_edit.cu:229-346), bound as
 * gstex_cuda.cuda.texture_edit.  updated_texture (texture_total_size x C,
 * C = texture_info[2] >= 5) is zeroed, then channels 0..4 receive the edit
 * (r*a, g*a, b*a, a, weight) of every composited pair whose view depth lies in
 * the pixel's [depth_lower, depth_upper].  updated_img (H,W,3), updated_alpha,
 * depth_lower, depth_upper (H,W).  settings bit 0 is the edit kernel's blur
 * switch (texture_edit.cu:51).  The reference writes past the output when
 * C < 5; this entry point rejects that.  background is unused, as in the
 * reference. */
size_t gstex_texture_edit_workspace_size(int num_points);
int gstex_texture_edit(int tiles_x, int tiles_y, int block_width, int img_width, int img_height,
                       int num_charts, int num_probs, int channels, int texture_total_size,
                       const int32_t *texture_dims, const float *updated_img,
                       const float *updated_alpha, const float *depth_lower,
                       const float *depth_upper, const int32_t *gaussian_ids_sorted, int num_ids,
                       const int32_t *tile_bins, const float *opacities, const float *means,
                       const float *scales, int num_points, float glob_scale, const float *quats,
                       const float *uv0, const float *umap, const float *vmap,
                       const float *viewmat, const float *c2w, float fx, float fy, float cx,
                       float cy, int settings, const float *background, float *updated_texture,
                       void *workspace, size_t workspace_bytes, void *stream);

/* Replaces texture_sample_forward_tensor (cuda/csrc/texture_sample.cu:72-105). */
int gstex_texture_sample_forward(int num_charts, int num_probs, int channels,
                                 const int32_t *texture_dims, const float *uvs, int num_queries,
                                 const float *texture, int num_texels, float *output,
                                 void *stream);

/* Replaces texture_sample_backward_tensor (cuda/csrc/texture_sample.cu:107-141)
 * with the adjoint w.r.t. the texels computed from v_output (the reference
 * kernel reads v_texture there, texture_sample.cu:58).  v_texture fully written.
 * The queries are sorted by chart in the workspace
 * (gstex_texture_sample_backward_workspace_size(num_queries) bytes). */
size_t gstex_texture_sample_backward_workspace_size(int num_queries);
int gstex_texture_sample_backward(int num_charts, int num_probs, int channels,
                                  const int32_t *texture_dims, const float *uvs, int num_queries,
                                  const float *texture, int num_texels, const float *v_output,
                                  float *v_texture, void *workspace, size_t workspace_bytes,
                                  void *stream);

/* ---- Training-step glue ------------------------------------------------
 * The reference does these steps with torch ops in its training loop
 * (example.py:126-143 preprocessing, example.py:231-233,277 torch.optim.Adam);
 * they have no native counterpart there. */

/* Raw parameters -> render inputs (example.py:126-143): scales (N,3) =
 * exp(log_scales.xy) with z = 1e-5 * mean(xy); quats (N,4) = q / |q|; uv0 (N,2),
 * umap, vmap (N,3) from mapping (N,4) = (uv0, log uv scale, theta) and the
 * rotation axes; colors (N,3) = sigmoid(rgb_logits) when rgb_logits is given
 * (NULL: colours come from elsewhere, e.g. SH, and colors is not written);
 * opacities (N,1) = sigmoid(opacity_logits).  quats_raw and mapping must be
 * 16-byte aligned. */
int gstex_preprocess_gaussians(int num_points, const float *log_scales, const float *quats_raw,
                               const float *mapping, const float *rgb_logits,
                               const float *opacity_logits, float *scales, float *quats,
                               float *uv0, float *umap, float *vmap, float *colors,
                               float *opacities, void *stream);

/* Adjoint of gstex_preprocess_gaussians: gradients of the render inputs ->
 * gradients of the raw parameters (log_scales.z gets 0: scales.z is detached).
 * v_mapping may be NULL (mapping frozen, as in example.py:117); v_rgb_logits is
 * written only when rgb_logits is given.  v_quats_raw and v_mapping 16-byte
 * aligned. */
int gstex_preprocess_gaussians_backward(
    int num_points, const float *log_scales, const float *quats_raw, const float *mapping,
    const float *rgb_logits, const float *opacity_logits, const float *v_scales,
    const float *v_quats, const float *v_uv0, const float *v_umap, const float *v_vmap,
    const float *v_colors, const float *v_opacities, float *v_log_scales, float *v_quats_raw,
    float *v_mapping, float *v_rgb_logits, float *v_opacity_logits, void *stream);

/* Cotangents of the image losses of a training step in one pass (no
 * counterpart in the reference, which forms them with torch autograd,
 * example.py:188-210).  Per pixel, with t = out_texture[:3] (missing channels
 * 0): d = (with_img ? out_img + t : t) - target; v_img = with_img ? scale d : 0;
 * v_texture_out[:3] = scale d, its other channels 0.  When out_normal is
 * given, v_normal = normal_scale * out_normal with normal_scale subtracted
 * from z (the gradient of the normal term, normal_scale = 2 weight / pixels).
 * Each step is rounded as the separate float32 add, subtract and multiply. */
int gstex_image_loss_grad(int num_pixels, int channels, int with_img, const float *out_img,
                          const float *out_texture, const float *target, float scale,
                          const float *out_normal, float normal_scale, float *v_img,
                          float *v_texture_out, float *v_normal, void *stream);

/* One torch.optim.Adam step (no weight decay, no amsgrad) over n float32 values,
 * step >= 1 counting from 1.  grad_mode GSTEX_GRAD_SIGMOID treats grad as the
 * gradient of sigmoid(param).  values (may be NULL) receives sigmoid(updated
 * param) in values_format (GSTEX_TEXTURE_F32 / GSTEX_TEXTURE_BF16), e.g. the
 * texels the next render reads. */
#define GSTEX_GRAD_PLAIN 0
#define GSTEX_GRAD_SIGMOID 1
int gstex_adam_step(size_t n, float *param, const float *grad, float *exp_avg, float *exp_avg_sq,
                    double lr, double beta1, double beta2, double eps, int step, int grad_mode,
                    void *values, int values_format, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GSTEX_HIP_H */
