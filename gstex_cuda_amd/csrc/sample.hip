// Bilinear queries into the jagged per-Gaussian texel tensor and the adjoint
// w.r.t. the texels.  Forward semantics: texture_sample.cu:11-36 (no uv clamp,
// replicate padding on the integer coordinates).  Coordinates the reference
// would read out of bounds (h*u <= -1) are clamped to the first texel.
#include "gstex_common.hpp"
#include "gstex_kernels.hpp"

namespace gstex {

__device__ __forceinline__ void sample_taps(int h, int w, int start, float u, float v, int num_texels,
                                            float wt[4], int row[4]) {
    const float tu = (float)h * u, tv = (float)w * v;
    int iu = f2i_trunc(tu), iv = f2i_trunc(tv);
    const float fu = tu - (float)iu, fv = tv - (float)iv;
    const int ju = max(min(iu + 1, h - 1), 0), jv = max(min(iv + 1, w - 1), 0);
    iu = max(min(iu, h - 1), 0);
    iv = max(min(iv, w - 1), 0);
    wt[0] = (1.f - fu) * (1.f - fv);
    wt[1] = (1.f - fu) * fv;
    wt[2] = fu * (1.f - fv);
    wt[3] = fu * fv;
    const int lim = num_texels - 1;
    row[0] = min(max(start + iu * w + iv, 0), lim);
    row[1] = min(max(start + iu * w + jv, 0), lim);
    row[2] = min(max(start + ju * w + iv, 0), lim);
    row[3] = min(max(start + ju * w + jv, 0), lim);
}

__global__ __launch_bounds__(256) void k_sample_forward(int q, int C, int num_texels,
                                                        const int* __restrict__ dims,
                                                        const float* __restrict__ uvs,
                                                        const float* __restrict__ texture,
                                                        float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q) return;
    float wt[4];
    int row[4];
    sample_taps(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], uvs[2 * i], uvs[2 * i + 1],
                num_texels, wt, row);
    for (int c = 0; c < C; ++c)
        out[(size_t)i * C + c] = wt[0] * texture[(size_t)row[0] * C + c] +
                                 wt[1] * texture[(size_t)row[1] * C + c] +
                                 wt[2] * texture[(size_t)row[2] * C + c] +
                                 wt[3] * texture[(size_t)row[3] * C + c];
}

// v_texture += sum over queries of weight * v_output (v_texture pre-zeroed).
__global__ __launch_bounds__(256) void k_sample_backward(int q, int C, int num_texels,
                                                         const int* __restrict__ dims,
                                                         const float* __restrict__ uvs,
                                                         const float* __restrict__ v_out,
                                                         float* __restrict__ v_texture) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q) return;
    float wt[4];
    int row[4];
    sample_taps(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], uvs[2 * i], uvs[2 * i + 1],
                num_texels, wt, row);
    // taps that clamp onto one texel (every tap of a 1x1 chart, the far
    // edge of larger ones) are added once with their summed weight; a
    // zero weight masks the lane off the atomic
#pragma unroll
    for (int k = 1; k < 4; ++k) {
#pragma unroll
        for (int j = 0; j < k; ++j) {
            if (wt[k] != 0.f && wt[j] != 0.f && row[k] == row[j]) {
                wt[j] += wt[k];
                wt[k] = 0.f;
            }
        }
    }
    for (int c = 0; c < C; ++c) {
        const float g = v_out[(size_t)i * C + c];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float add = wt[k] * g;
            if (add != 0.f) unsafeAtomicAdd(v_texture + (size_t)row[k] * C + c, add);
        }
    }
}

hipError_t sample_forward(int q, int C, int num_texels, const int* dims, const float* uvs,
                          const float* texture, float* out, hipStream_t s) {
    if (q <= 0 || C <= 0) return hipSuccess;
    k_sample_forward<<<(q + 255) / 256, 256, 0, s>>>(q, C, num_texels, dims, uvs, texture, out);
    return hipGetLastError();
}

hipError_t sample_backward(int q, int C, int num_texels, const int* dims, const float* uvs,
                           const float* v_out, float* v_texture, hipStream_t s) {
    hipError_t e = fill_f32(v_texture, (size_t)num_texels * (size_t)C, 0.f, s);
    if (e != hipSuccess || q <= 0 || C <= 0) return e;
    k_sample_backward<<<(q + 255) / 256, 256, 0, s>>>(q, C, num_texels, dims, uvs, v_out,
                                                     v_texture);
    return hipGetLastError();
}

}  // namespace gstex
