// Real spherical harmonics up to degree 4: colours from per-Gaussian
// coefficients and the adjoint w.r.t. the coefficients.  Basis constants and
// band grouping follow sh.cuh:6-210 (the reference's sign convention).
//
// Kernels are instantiated per evaluated degree so every basis value and
// coefficient index is a compile-time constant (no scratch); the coefficient
// stride (bases of the stored degree) stays a runtime value.
#include "gstex_common.hpp"
#include "gstex_kernels.hpp"

namespace gstex {

constexpr float kShC0 = 0.28209479177387814f;
constexpr float kShC1 = 0.4886025119029199f;
constexpr float kShC2_0 = 1.0925484305920792f, kShC2_1 = -1.0925484305920792f,
                kShC2_2 = 0.31539156525252005f, kShC2_3 = -1.0925484305920792f,
                kShC2_4 = 0.5462742152960396f;
constexpr float kShC3_0 = -0.5900435899266435f, kShC3_1 = 2.890611442640554f,
                kShC3_2 = -0.4570457994644658f, kShC3_3 = 0.3731763325901154f,
                kShC3_4 = -0.4570457994644658f, kShC3_5 = 1.445305721320277f,
                kShC3_6 = -0.5900435899266435f;
constexpr float kShC4_0 = 2.5033429417967046f, kShC4_1 = -1.7701307697799304f,
                kShC4_2 = 0.9461746957575601f, kShC4_3 = -0.6690465435572892f,
                kShC4_4 = 0.10578554691520431f, kShC4_5 = -0.6690465435572892f,
                kShC4_6 = 0.47308734787878004f, kShC4_7 = -1.7701307697799304f,
                kShC4_8 = 0.6258357354491761f;

__host__ __device__ inline int num_sh_bases(int degree) {
    return degree <= 0 ? 1 : degree == 1 ? 4 : degree == 2 ? 9 : degree == 3 ? 16 : 25;
}

// Basis values 1..24 at the normalised direction for bands up to D.
template <int D>
__device__ __forceinline__ void sh_basis(float dx, float dy, float dz, float (&b)[25], float& x,
                                         float& y, float& z) {
    const float norm = sqrtf(dx * dx + dy * dy + dz * dz);
    x = dx / norm;
    y = dy / norm;
    z = dz / norm;
    const float xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    b[0] = kShC0;
    b[1] = -kShC1 * y;
    b[2] = kShC1 * z;
    b[3] = -kShC1 * x;
    if constexpr (D >= 2) {
        b[4] = kShC2_0 * xy;
        b[5] = kShC2_1 * yz;
        b[6] = kShC2_2 * (2.f * zz - xx - yy);
        b[7] = kShC2_3 * xz;
        b[8] = kShC2_4 * (xx - yy);
    }
    if constexpr (D >= 3) {
        b[9] = kShC3_0 * y * (3.f * xx - yy);
        b[10] = kShC3_1 * xy * z;
        b[11] = kShC3_2 * y * (4.f * zz - xx - yy);
        b[12] = kShC3_3 * z * (2.f * zz - 3.f * xx - 3.f * yy);
        b[13] = kShC3_4 * x * (4.f * zz - xx - yy);
        b[14] = kShC3_5 * z * (xx - yy);
        b[15] = kShC3_6 * x * (xx - 3.f * yy);
    }
    if constexpr (D >= 4) {
        b[16] = kShC4_0 * xy * (xx - yy);
        b[17] = kShC4_1 * yz * (3.f * xx - yy);
        b[18] = kShC4_2 * xy * (7.f * zz - 1.f);
        b[19] = kShC4_3 * yz * (7.f * zz - 3.f);
        b[20] = kShC4_4 * (zz * (35.f * zz - 30.f) + 3.f);
        b[21] = kShC4_5 * xz * (7.f * zz - 3.f);
        b[22] = kShC4_6 * (xx - yy) * (7.f * zz - 1.f);
        b[23] = kShC4_7 * xz * (xx - 3.f * yy);
        b[24] = kShC4_8 * (xx * (xx - 3.f * yy) - yy * (3.f * xx - yy));
    }
}

// Coefficient k of a float32 or bfloat16 (uint16 bits, widened exactly) array.
__device__ __forceinline__ float coef(const float* c, int k) { return c[k]; }
__device__ __forceinline__ float coef(const unsigned short* c, int k) {
    return __uint_as_float((unsigned)c[k] << 16);
}

template <int LO, int HI, typename CT>
__device__ __forceinline__ float band_sum(const float (&b)[25], const CT* c, int ch) {
    float s = b[LO] * coef(c, LO * 3 + ch);
#pragma unroll
    for (int k = LO + 1; k < HI; ++k) s += b[k] * coef(c, k * 3 + ch);
    return s;
}

// One lane per Gaussian; coefficients are (N, nb, 3) float32 or bfloat16,
// evaluated up to degree D in float32.
template <int D, typename CT>
__global__ __launch_bounds__(256) void k_sh_forward(int n, int nb, const float* __restrict__ viewdirs,
                                                    const CT* __restrict__ coeffs,
                                                    float* __restrict__ colors) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const CT* c = coeffs + (size_t)i * nb * 3;
    float acc[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) acc[ch] = kShC0 * coef(c, ch);
    if constexpr (D >= 1) {
        float b[25], x, y, z;
        sh_basis<D>(viewdirs[3 * i], viewdirs[3 * i + 1], viewdirs[3 * i + 2], b, x, y, z);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            acc[ch] += kShC1 * (-y * coef(c, 3 + ch) + z * coef(c, 6 + ch) - x * coef(c, 9 + ch));
            if constexpr (D >= 2) acc[ch] += band_sum<4, 9>(b, c, ch);
            if constexpr (D >= 3) acc[ch] += band_sum<9, 16>(b, c, ch);
            if constexpr (D >= 4) acc[ch] += band_sum<16, 25>(b, c, ch);
        }
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) colors[3 * i + ch] = acc[ch];
}

// v_coeffs[k][c] = basis_k * v_colors[c] for k < bases(D), zero above
// (sh.cuh:234-253); the output is fully written, so no memset.
// With staging (v_coeffs 16-byte aligned and dynamic LDS of 256 x (3 nb + 1)
// floats), a block's rows go through LDS (rows padded to an odd word count,
// so the row writes do not collide in banks) and leave as 16-byte stores of
// consecutive addresses: the block's rows are one contiguous span.  Without
// it each thread stores its own row.
template <int D>
__global__ __launch_bounds__(256) void k_sh_backward(int n, int nb, const float* __restrict__ viewdirs,
                                                     const float* __restrict__ v_colors,
                                                     float* __restrict__ v_coeffs, bool staged) {
    extern __shared__ float stage[];
    const int i0 = blockIdx.x * blockDim.x, i = i0 + threadIdx.x;
    const bool valid = i < n;
    if (!staged && !valid) return;
    constexpr int used = D <= 0 ? 1 : D == 1 ? 4 : D == 2 ? 9 : D == 3 ? 16 : 25;
    float b[25];
    b[0] = kShC0;
    float g0 = 0.f, g1 = 0.f, g2 = 0.f;
    if (valid) {
        if constexpr (D >= 1) {
            float x, y, z;
            sh_basis<D>(viewdirs[3 * i], viewdirs[3 * i + 1], viewdirs[3 * i + 2], b, x, y, z);
        }
        g0 = v_colors[3 * i];
        g1 = v_colors[3 * i + 1];
        g2 = v_colors[3 * i + 2];
    }
    const int R = 3 * nb;
    float* out = staged ? stage + threadIdx.x * (R + 1) : v_coeffs + (size_t)i * R;
#pragma unroll
    for (int k = 0; k < used; ++k) {
        out[3 * k] = b[k] * g0;
        out[3 * k + 1] = b[k] * g1;
        out[3 * k + 2] = b[k] * g2;
    }
    for (int k = 3 * used; k < R; ++k) out[k] = 0.f;
    if (!staged) return;
    __syncthreads();
    const int cnt = min(n - i0, (int)blockDim.x);
    const int total = cnt * R;
    float* dst = v_coeffs + (size_t)i0 * R;
    // position q of the span is row r, column c; both advance without a
    // division per element
    const int step = 4 * (int)blockDim.x, step_r = step / R, step_c = step - step_r * R;
    int q = 4 * (int)threadIdx.x, r = q / R, c = q - r * R;
    for (; q < total; q += step) {
        float v[4];
        int rr = r, cc = c;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            v[j] = q + j < total ? stage[rr * (R + 1) + cc] : 0.f;
            if (++cc == R) {
                cc = 0;
                ++rr;
            }
        }
        if (q + 4 <= total) {
            *reinterpret_cast<float4*>(dst + q) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
            for (int j = 0; q + j < total; ++j) dst[q + j] = v[j];
        }
        r += step_r;
        c += step_c;
        if (c >= R) {
            c -= R;
            ++r;
        }
    }
}

// The sum over V views of k_sh_backward: viewdirs and v_colors are (V, n, 3)
// stacks, v_coeffs receives sum_v basis(dir_v) x v_colors_v.  The products are
// rounded on their own and added in view order (no fused multiply-add), so
// the result equals adding the per-view outputs one after another; each
// coefficient is written once instead of once per view and once per sum.
template <int D>
__global__ __launch_bounds__(256) void k_sh_backward_views(int n, int nb, int views,
                                                           const float* __restrict__ viewdirs,
                                                           const float* __restrict__ v_colors,
                                                           float* __restrict__ v_coeffs) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr int used = D <= 0 ? 1 : D == 1 ? 4 : D == 2 ? 9 : D == 3 ? 16 : 25;
    float acc[3 * used];
    for (int v = 0; v < views; ++v) {
        const size_t o = ((size_t)v * n + i) * 3;
        float b[25];
        b[0] = kShC0;
        if constexpr (D >= 1) {
            float x, y, z;
            sh_basis<D>(viewdirs[o], viewdirs[o + 1], viewdirs[o + 2], b, x, y, z);
        }
        const float g0 = v_colors[o], g1 = v_colors[o + 1], g2 = v_colors[o + 2];
#pragma unroll
        for (int k = 0; k < used; ++k) {
            const float p0 = b[k] * g0, p1 = b[k] * g1, p2 = b[k] * g2;
            acc[3 * k] = v == 0 ? p0 : acc[3 * k] + p0;
            acc[3 * k + 1] = v == 0 ? p1 : acc[3 * k + 1] + p1;
            acc[3 * k + 2] = v == 0 ? p2 : acc[3 * k + 2] + p2;
        }
    }
    float* out = v_coeffs + (size_t)i * nb * 3;
#pragma unroll
    for (int k = 0; k < 3 * used; ++k) out[k] = acc[k];
    for (int k = 3 * used; k < 3 * nb; ++k) out[k] = 0.f;
}

template <typename CT>
static hipError_t sh_forward_t(int n, int degree, int degrees_to_use, const float* viewdirs,
                               const CT* coeffs, float* colors, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int nb = num_sh_bases(degree), g = (n + 255) / 256;
    switch (degrees_to_use) {
        case 0: k_sh_forward<0, CT><<<g, 256, 0, stream>>>(n, nb, viewdirs, coeffs, colors); break;
        case 1: k_sh_forward<1, CT><<<g, 256, 0, stream>>>(n, nb, viewdirs, coeffs, colors); break;
        case 2: k_sh_forward<2, CT><<<g, 256, 0, stream>>>(n, nb, viewdirs, coeffs, colors); break;
        case 3: k_sh_forward<3, CT><<<g, 256, 0, stream>>>(n, nb, viewdirs, coeffs, colors); break;
        default: k_sh_forward<4, CT><<<g, 256, 0, stream>>>(n, nb, viewdirs, coeffs, colors); break;
    }
    return hipGetLastError();
}

hipError_t launch_sh_forward(int n, int degree, int degrees_to_use, const float* viewdirs,
                             const float* coeffs, float* colors, hipStream_t stream) {
    return sh_forward_t<float>(n, degree, degrees_to_use, viewdirs, coeffs, colors, stream);
}

hipError_t launch_sh_forward_bf16(int n, int degree, int degrees_to_use, const float* viewdirs,
                                  const unsigned short* coeffs, float* colors,
                                  hipStream_t stream) {
    return sh_forward_t<unsigned short>(n, degree, degrees_to_use, viewdirs, coeffs, colors,
                                        stream);
}

hipError_t launch_sh_backward(int n, int degree, int degrees_to_use, const float* viewdirs,
                              const float* v_colors, float* v_coeffs, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int nb = num_sh_bases(degree), g = (n + 255) / 256;
    // staged rows up to 64 KB of LDS per block (degree <= 3)
    const size_t lds = (size_t)256 * (3 * nb + 1) * sizeof(float);
    const bool staged = (reinterpret_cast<uintptr_t>(v_coeffs) & 15) == 0 && lds <= 65536;
    const size_t sh = staged ? lds : 0;
    switch (degrees_to_use) {
        case 0: k_sh_backward<0><<<g, 256, sh, stream>>>(n, nb, viewdirs, v_colors, v_coeffs, staged); break;
        case 1: k_sh_backward<1><<<g, 256, sh, stream>>>(n, nb, viewdirs, v_colors, v_coeffs, staged); break;
        case 2: k_sh_backward<2><<<g, 256, sh, stream>>>(n, nb, viewdirs, v_colors, v_coeffs, staged); break;
        case 3: k_sh_backward<3><<<g, 256, sh, stream>>>(n, nb, viewdirs, v_colors, v_coeffs, staged); break;
        default: k_sh_backward<4><<<g, 256, sh, stream>>>(n, nb, viewdirs, v_colors, v_coeffs, staged); break;
    }
    return hipGetLastError();
}

hipError_t launch_sh_backward_views(int n, int views, int degree, int degrees_to_use,
                                    const float* viewdirs, const float* v_colors,
                                    float* v_coeffs, hipStream_t stream) {
    if (n <= 0 || views <= 0) return hipSuccess;
    const int nb = num_sh_bases(degree), g = (n + 255) / 256;
    switch (degrees_to_use) {
        case 0: k_sh_backward_views<0><<<g, 256, 0, stream>>>(n, nb, views, viewdirs, v_colors, v_coeffs); break;
        case 1: k_sh_backward_views<1><<<g, 256, 0, stream>>>(n, nb, views, viewdirs, v_colors, v_coeffs); break;
        case 2: k_sh_backward_views<2><<<g, 256, 0, stream>>>(n, nb, views, viewdirs, v_colors, v_coeffs); break;
        case 3: k_sh_backward_views<3><<<g, 256, 0, stream>>>(n, nb, views, viewdirs, v_colors, v_coeffs); break;
        default: k_sh_backward_views<4><<<g, 256, 0, stream>>>(n, nb, views, viewdirs, v_colors, v_coeffs); break;
    }
    return hipGetLastError();
}

int sh_num_bases(int degree) { return num_sh_bases(degree); }

}  // namespace gstex
