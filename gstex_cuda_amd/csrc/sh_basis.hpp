// Real spherical-harmonic basis up to degree 4 (constants and band grouping
// of sh.cuh:6-210, the reference's sign convention), shared by the SH kernels
// and the gradient finalize that applies the SH adjoint.
#pragma once
#include <hip/hip_runtime.h>

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

// The SH adjoint rows of a block of 256 Gaussians: thread t's row (3 nb
// floats) is basis_k * (g0, g1, g2) for k < bases(D) and zeros above
// (sh.cuh:234-253).  The rows go through LDS (stage: 256 x (3 nb + 1)
// floats, rows padded to an odd word count so the row writes do not collide
// in banks) and leave as 16-byte stores of consecutive addresses, since the
// block's rows are one contiguous span of out (16-byte aligned).  Threads past
// n hold no row; every thread of the block must call this.
template <int D>
__device__ __forceinline__ void sh_rows_staged(float* stage, int nb, const float (&b)[25],
                                               float g0, float g1, float g2, int i0, int n,
                                               float* __restrict__ out) {
    constexpr int used = D <= 0 ? 1 : D == 1 ? 4 : D == 2 ? 9 : D == 3 ? 16 : 25;
    const int R = 3 * nb;
    float* row = stage + threadIdx.x * (R + 1);
#pragma unroll
    for (int k = 0; k < used; ++k) {
        row[3 * k] = b[k] * g0;
        row[3 * k + 1] = b[k] * g1;
        row[3 * k + 2] = b[k] * g2;
    }
    for (int k = 3 * used; k < R; ++k) row[k] = 0.f;
    __syncthreads();
    const int cnt = min(n - i0, (int)blockDim.x);
    const int total = cnt * R;
    float* dst = out + (size_t)i0 * R;
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

// LDS bytes of sh_rows_staged for a block of 256 (staged only up to 64 KB).
inline size_t sh_rows_stage_bytes(int nb) { return (size_t)256 * (3 * nb + 1) * sizeof(float); }

}  // namespace gstex
