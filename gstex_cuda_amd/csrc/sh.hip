// Real spherical harmonics up to degree 4: colours from per-Gaussian
// coefficients and the adjoint w.r.t. the coefficients.  Basis constants and
// band grouping follow sh.cuh:6-210 (the reference's sign convention).
//
// Kernels are instantiated per evaluated degree so every basis value and
// coefficient index is a compile-time constant (no scratch); the coefficient
// stride (bases of the stored degree) stays a runtime value.
#include "gstex_common.hpp"
#include "gstex_kernels.hpp"
#include "sh_basis.hpp"

namespace gstex {

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
// With c2w (4 x 4 row-major camera-to-world), dirs holds the means and the
// direction of Gaussian i is means[i] - c2w[:3, 3] (one float32 subtraction
// per component).  With staging (float32 coefficients, 16-byte aligned, and
// dynamic LDS of 256 x (3 nb + 1) floats), the block's coefficient rows, one
// contiguous span, come in as 16-byte loads of consecutive addresses and
// each thread reads its row from LDS (rows padded to an odd word count).
template <int D, typename CT>
__global__ __launch_bounds__(256) void k_sh_forward(int n, int nb, const float* __restrict__ dirs,
                                                    const float* __restrict__ c2w,
                                                    const CT* __restrict__ coeffs,
                                                    float* __restrict__ colors, bool staged) {
    extern __shared__ float stage[];
    const int i0 = blockIdx.x * blockDim.x, i = i0 + threadIdx.x;
    const CT* c = coeffs + (size_t)i * nb * 3;
    if constexpr (sizeof(CT) == 4) {
        if (staged) {
            const int R = 3 * nb, cnt = min(n - i0, (int)blockDim.x), total = cnt * R;
            const float* src = reinterpret_cast<const float*>(coeffs) + (size_t)i0 * R;
            const int step = 4 * (int)blockDim.x, step_r = step / R, step_c = step - step_r * R;
            int q = 4 * (int)threadIdx.x, r = q / R, cc = q - r * R;
            for (; q < total; q += step) {
                float v[4];
                if (q + 4 <= total) {
                    const float4 t = *reinterpret_cast<const float4*>(src + q);
                    v[0] = t.x;
                    v[1] = t.y;
                    v[2] = t.z;
                    v[3] = t.w;
                } else {
                    for (int j = 0; j < 4; ++j) v[j] = q + j < total ? src[q + j] : 0.f;
                }
                int rr = r, c2 = cc;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (q + j < total) stage[rr * (R + 1) + c2] = v[j];
                    if (++c2 == R) {
                        c2 = 0;
                        ++rr;
                    }
                }
                r += step_r;
                cc += step_c;
                if (cc >= R) {
                    cc -= R;
                    ++r;
                }
            }
            __syncthreads();
            c = reinterpret_cast<const CT*>(stage + threadIdx.x * (R + 1));
        }
    }
    if (i >= n) return;
    float acc[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) acc[ch] = kShC0 * coef(c, ch);
    if constexpr (D >= 1) {
        float b[25], x, y, z;
        float dx = dirs[3 * i], dy = dirs[3 * i + 1], dz = dirs[3 * i + 2];
        if (c2w) {
            dx -= c2w[3];
            dy -= c2w[7];
            dz -= c2w[11];
        }
        sh_basis<D>(dx, dy, dz, b, x, y, z);
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
    if (staged) {
        sh_rows_staged<D>(stage, nb, b, g0, g1, g2, i0, n, v_coeffs);
        return;
    }
    float* out = v_coeffs + (size_t)i * 3 * nb;
#pragma unroll
    for (int k = 0; k < used; ++k) {
        out[3 * k] = b[k] * g0;
        out[3 * k + 1] = b[k] * g1;
        out[3 * k + 2] = b[k] * g2;
    }
    for (int k = 3 * used; k < 3 * nb; ++k) out[k] = 0.f;
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
static hipError_t sh_forward_t(int n, int degree, int degrees_to_use, const float* dirs,
                               const float* c2w, const CT* coeffs, float* colors,
                               hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int nb = num_sh_bases(degree), g = (n + 255) / 256;
    const size_t lds = sh_rows_stage_bytes(nb);
    const bool staged = sizeof(CT) == 4 && (reinterpret_cast<uintptr_t>(coeffs) & 15) == 0 &&
                        lds <= 65536;
    const size_t sh = staged ? lds : 0;
#define GSTEX_SHF(D) \
    k_sh_forward<D, CT><<<g, 256, sh, stream>>>(n, nb, dirs, c2w, coeffs, colors, staged)
    switch (degrees_to_use) {
        case 0: GSTEX_SHF(0); break;
        case 1: GSTEX_SHF(1); break;
        case 2: GSTEX_SHF(2); break;
        case 3: GSTEX_SHF(3); break;
        default: GSTEX_SHF(4); break;
    }
#undef GSTEX_SHF
    return hipGetLastError();
}

hipError_t launch_sh_forward(int n, int degree, int degrees_to_use, const float* viewdirs,
                             const float* coeffs, float* colors, hipStream_t stream) {
    return sh_forward_t<float>(n, degree, degrees_to_use, viewdirs, nullptr, coeffs, colors,
                               stream);
}

hipError_t launch_sh_forward_bf16(int n, int degree, int degrees_to_use, const float* viewdirs,
                                  const unsigned short* coeffs, float* colors,
                                  hipStream_t stream) {
    return sh_forward_t<unsigned short>(n, degree, degrees_to_use, viewdirs, nullptr, coeffs,
                                        colors, stream);
}

hipError_t launch_sh_forward_means(int n, int degree, int degrees_to_use, const float* means,
                                   const float* c2w, const void* coeffs, bool bf16,
                                   float* colors, hipStream_t stream) {
    if (bf16)
        return sh_forward_t<unsigned short>(n, degree, degrees_to_use, means, c2w,
                                            static_cast<const unsigned short*>(coeffs), colors,
                                            stream);
    return sh_forward_t<float>(n, degree, degrees_to_use, means, c2w,
                               static_cast<const float*>(coeffs), colors, stream);
}

hipError_t launch_sh_backward(int n, int degree, int degrees_to_use, const float* viewdirs,
                              const float* v_colors, float* v_coeffs, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int nb = num_sh_bases(degree), g = (n + 255) / 256;
    // staged rows up to 64 KB of LDS per block (degree <= 3)
    const size_t lds = sh_rows_stage_bytes(nb);
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
