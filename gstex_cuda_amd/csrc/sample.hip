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

// Taps that clamp onto one texel (every tap of a 1x1 chart, the far edge of
// larger ones) are added once with their summed weight; a zero weight marks
// a tap with nothing to add.
__device__ __forceinline__ void merge_taps(float wt[4], const int row[4]) {
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
}

// Query keys for the adjoint: the bucket of 2^kSampleBucketShift texels that
// holds the query's chart start (clamped into the texture), and its index.
// With C <= 4 the query is also packed into one 32-byte record (u, v, start,
// h | w << 16, the C cotangents), read in query order here so that the
// sorted pass gathers one record per query instead of three arrays.  A chart
// side of 2^16 or more does not fit the packing; such a record is marked
// (h | w << 16 = 0xffffffff) and that query reads the original arrays.
constexpr int kSampleBucketShift = 8;
constexpr uint32_t kSampleUnpacked = 0xffffffffu;

__global__ __launch_bounds__(256) void k_sample_keys(int q, int C, int num_texels,
                                                     const int* __restrict__ dims,
                                                     const float* __restrict__ uvs,
                                                     const float* __restrict__ v_out,
                                                     uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ idx,
                                                     float4* __restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q) return;
    const int h = dims[3 * i], w = dims[3 * i + 1], start = dims[3 * i + 2];
    keys[i] = (uint32_t)min(max(start, 0), num_texels) >> kSampleBucketShift;
    idx[i] = (uint32_t)i;
    if (rec != nullptr) {
        const bool fits = h >= 0 && w >= 0 && h < 65536 && w < 65536;
        const uint32_t hw = fits ? ((uint32_t)h | ((uint32_t)w << 16)) : kSampleUnpacked;
        float g[4] = {0.f, 0.f, 0.f, 0.f};
        for (int c = 0; c < C; ++c) g[c] = v_out[(size_t)i * C + c];
        rec[2 * (size_t)i] = make_float4(uvs[2 * i], uvs[2 * i + 1], __int_as_float(start),
                                         __uint_as_float(hw));
        rec[2 * (size_t)i + 1] = make_float4(g[0], g[1], g[2], g[3]);
    }
}

// v_texture += sum over queries of weight * v_output, over the queries sorted
// by chart-start bucket: a workgroup takes kSampleRange consecutive sorted
// queries, whose charts lie in one contiguous run of texels starting at the
// first query's bucket.  Their taps are added into an LDS window over that run
// (taps past the window go straight to HBM), and the window is flushed with
// one atomic per value it received.  Each thread takes kSamplePer
// consecutive queries, so the lanes of a wave fall on different charts.
constexpr int kSamplePer = 16;
constexpr int kSampleRange = 256 * kSamplePer;
constexpr int kSampleWindow = 8192;  // floats (32 KiB)

__global__ __launch_bounds__(256) void k_sample_backward_sorted(
    int q, int C, int num_texels, const uint32_t* __restrict__ keys_sorted,
    const uint32_t* __restrict__ idx_sorted, const float4* __restrict__ rec,
    const int* __restrict__ dims, const float* __restrict__ uvs, const float* __restrict__ v_out,
    float* __restrict__ v_texture) {
    __shared__ float win[kSampleWindow];
    __shared__ int used;
    const int first = blockIdx.x * kSampleRange;
    const int last = min(first + kSampleRange, q);
    for (int k = threadIdx.x; k < kSampleWindow; k += blockDim.x) win[k] = 0.f;
    if (threadIdx.x == 0) used = 0;
    __syncthreads();
    const long long base = ((long long)keys_sorted[first] << kSampleBucketShift) * C;
    int hi = 0;  // past the last window value this thread touched
    const int k0 = first + threadIdx.x * kSamplePer;
    for (int k = k0; k < min(k0 + kSamplePer, last); ++k) {
        const int i = (int)idx_sorted[k];
        int h, w, start;
        float u, v, g[4];
        bool packed = false;
        if (rec != nullptr) {
            const float4 a = rec[2 * (size_t)i];
            const float4 b = rec[2 * (size_t)i + 1];
            const uint32_t hw = __float_as_uint(a.w);
            packed = hw != kSampleUnpacked;
            u = a.x;
            v = a.y;
            start = __float_as_int(a.z);
            h = (int)(hw & 0xffffu);
            w = (int)(hw >> 16);
            g[0] = b.x;
            g[1] = b.y;
            g[2] = b.z;
            g[3] = b.w;
        }
        if (!packed) {
            h = dims[3 * i];
            w = dims[3 * i + 1];
            start = dims[3 * i + 2];
            u = uvs[2 * i];
            v = uvs[2 * i + 1];
        }
        float wt[4];
        int row[4];
        sample_taps(h, w, start, u, v, num_texels, wt, row);
        merge_taps(wt, row);
        for (int c = 0; c < C; ++c) {
            const float gc = packed ? g[c & 3] : v_out[(size_t)i * C + c];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float add = wt[t] * gc;
                if (add == 0.f) continue;
                const long long at = (long long)row[t] * C + c;
                const long long off = at - base;
                if (off >= 0 && off < kSampleWindow) {
                    atomicAdd(win + off, add);
                    hi = max(hi, (int)off + 1);
                } else {
                    unsafeAtomicAdd(v_texture + at, add);
                }
            }
        }
    }
    atomicMax(&used, hi);
    __syncthreads();
    const int n = used;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const float val = win[k];
        if (val != 0.f) unsafeAtomicAdd(v_texture + base + k, val);
    }
}

hipError_t sample_forward(int q, int C, int num_texels, const int* dims, const float* uvs,
                          const float* texture, float* out, hipStream_t s) {
    if (q <= 0 || C <= 0) return hipSuccess;
    k_sample_forward<<<(q + 255) / 256, 256, 0, s>>>(q, C, num_texels, dims, uvs, texture, out);
    return hipGetLastError();
}

static size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t sample_backward_workspace_bytes(int q) {
    if (q <= 0) return 256;
    return 4 * up256(sizeof(uint32_t) * (size_t)q) + up256(32 * (size_t)q) +
           sort_u32_pairs_workspace_bytes(q) + 256;
}

hipError_t sample_backward(int q, int C, int num_texels, const int* dims, const float* uvs,
                           const float* v_out, float* v_texture, void* ws, size_t ws_bytes,
                           hipStream_t s) {
    hipError_t e = fill_f32(v_texture, (size_t)num_texels * (size_t)C, 0.f, s);
    if (e != hipSuccess || q <= 0 || C <= 0) return e;
    if (ws_bytes < sample_backward_workspace_bytes(q)) return hipErrorInvalidValue;
    char* p = static_cast<char*>(ws);
    const size_t a = up256(sizeof(uint32_t) * (size_t)q);
    uint32_t* keys = (uint32_t*)p;
    uint32_t* idx = (uint32_t*)(p + a);
    uint32_t* keys_s = (uint32_t*)(p + 2 * a);
    uint32_t* idx_s = (uint32_t*)(p + 3 * a);
    float4* rec = (float4*)(p + 4 * a);
    void* rws = p + 4 * a + up256(32 * (size_t)q);
    const size_t rws_bytes = ws_bytes - 4 * a - up256(32 * (size_t)q);
    // records only when the cotangents fit them
    if (C > 4) rec = nullptr;
    k_sample_keys<<<(q + 255) / 256, 256, 0, s>>>(q, C, num_texels, dims, uvs, v_out, keys, idx,
                                                  rec);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int end_bit = 32 - __builtin_clz(((unsigned)num_texels >> kSampleBucketShift) | 1u);
    e = sort_u32_pairs(q, keys, idx, keys_s, idx_s, end_bit, rws, rws_bytes, s);
    if (e != hipSuccess) return e;
    k_sample_backward_sorted<<<(q + kSampleRange - 1) / kSampleRange, 256, 0, s>>>(
        q, C, num_texels, keys_s, idx_s, rec, dims, uvs, v_out, v_texture);
    return hipGetLastError();
}

}  // namespace gstex
