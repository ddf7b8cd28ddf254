// Real spherical harmonics up to degree 4: colours from per-Gaussian
// coefficients and the adjoint w.r.t. the coefficients.  Basis constants and
// band grouping follow sh.cuh:6-210 (the reference's sign convention).
#include "gstex_common.hpp"
#include "gstex_kernels.hpp"

namespace gstex {

__constant__ float kShC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                               -1.0925484305920792f, 0.5462742152960396f};
__constant__ float kShC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                               0.3731763325901154f,  -0.4570457994644658f, 1.445305721320277f,
                               -0.5900435899266435f};
__constant__ float kShC4[9] = {2.5033429417967046f,  -1.7701307697799304f, 0.9461746957575601f,
                               -0.6690465435572892f, 0.10578554691520431f, -0.6690465435572892f,
                               0.47308734787878004f, -1.7701307697799304f, 0.6258357354491761f};
constexpr float kShC0 = 0.28209479177387814f;
constexpr float kShC1 = 0.4886025119029199f;

__host__ __device__ inline int num_sh_bases(int degree) {
    return degree <= 0 ? 1 : degree == 1 ? 4 : degree == 2 ? 9 : degree == 3 ? 16 : 25;
}

// Basis values b[1..24] (b[0] = C0) for the normalised direction.
__device__ __forceinline__ void sh_basis(int degree, float dx, float dy, float dz, float* b) {
    const float norm = sqrtf(dx * dx + dy * dy + dz * dz);
    const float x = dx / norm, y = dy / norm, z = dz / norm;
    const float xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    b[1] = -kShC1 * y;
    b[2] = kShC1 * z;
    b[3] = -kShC1 * x;
    if (degree < 2) return;
    b[4] = kShC2[0] * xy;
    b[5] = kShC2[1] * yz;
    b[6] = kShC2[2] * (2.f * zz - xx - yy);
    b[7] = kShC2[3] * xz;
    b[8] = kShC2[4] * (xx - yy);
    if (degree < 3) return;
    b[9] = kShC3[0] * y * (3.f * xx - yy);
    b[10] = kShC3[1] * xy * z;
    b[11] = kShC3[2] * y * (4.f * zz - xx - yy);
    b[12] = kShC3[3] * z * (2.f * zz - 3.f * xx - 3.f * yy);
    b[13] = kShC3[4] * x * (4.f * zz - xx - yy);
    b[14] = kShC3[5] * z * (xx - yy);
    b[15] = kShC3[6] * x * (xx - 3.f * yy);
    if (degree < 4) return;
    b[16] = kShC4[0] * xy * (xx - yy);
    b[17] = kShC4[1] * yz * (3.f * xx - yy);
    b[18] = kShC4[2] * xy * (7.f * zz - 1.f);
    b[19] = kShC4[3] * yz * (7.f * zz - 3.f);
    b[20] = kShC4[4] * (zz * (35.f * zz - 30.f) + 3.f);
    b[21] = kShC4[5] * xz * (7.f * zz - 3.f);
    b[22] = kShC4[6] * (xx - yy) * (7.f * zz - 1.f);
    b[23] = kShC4[7] * xz * (xx - 3.f * yy);
    b[24] = kShC4[8] * (xx * (xx - 3.f * yy) - yy * (3.f * xx - yy));
}

// One lane per Gaussian; coefficients are (N, bases(degree), 3).
__global__ __launch_bounds__(256) void k_sh_forward(int n, int degree, int degrees_to_use,
                                                    const float* __restrict__ viewdirs,
                                                    const float* __restrict__ coeffs,
                                                    float* __restrict__ colors) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nb = num_sh_bases(degree);
    const float* c = coeffs + (size_t)i * nb * 3;
    float acc[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) acc[ch] = kShC0 * c[ch];
    if (degrees_to_use >= 1) {
        const float dx = viewdirs[3 * i], dy = viewdirs[3 * i + 1], dz = viewdirs[3 * i + 2];
        float b[25];
        sh_basis(degrees_to_use, dx, dy, dz, b);
        const float norm = sqrtf(dx * dx + dy * dy + dz * dz);
        const float x = dx / norm, y = dy / norm, z = dz / norm;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch)
            acc[ch] += kShC1 * (-y * c[3 + ch] + z * c[6 + ch] - x * c[9 + ch]);
        const int lo[3] = {4, 9, 16}, hi[3] = {9, 16, 25};
        for (int band = 0; band < 3; ++band) {
            if (degrees_to_use < band + 2) break;
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                float s = b[lo[band]] * c[lo[band] * 3 + ch];
                for (int k = lo[band] + 1; k < hi[band]; ++k) s += b[k] * c[k * 3 + ch];
                acc[ch] += s;
            }
        }
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) colors[3 * i + ch] = acc[ch];
}

// v_coeffs[k][c] = basis_k * v_colors[c] for k < bases(degrees_to_use), zero
// above (sh.cuh:234-253; the output is fully written here, so no memset).
__global__ __launch_bounds__(256) void k_sh_backward(int n, int degree, int degrees_to_use,
                                                     const float* __restrict__ viewdirs,
                                                     const float* __restrict__ v_colors,
                                                     float* __restrict__ v_coeffs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nb = num_sh_bases(degree);
    const int used = num_sh_bases(degrees_to_use) < nb ? num_sh_bases(degrees_to_use) : nb;
    float b[25];
    b[0] = kShC0;
    if (degrees_to_use >= 1)
        sh_basis(degrees_to_use, viewdirs[3 * i], viewdirs[3 * i + 1], viewdirs[3 * i + 2], b);
    const float g0 = v_colors[3 * i], g1 = v_colors[3 * i + 1], g2 = v_colors[3 * i + 2];
    float* out = v_coeffs + (size_t)i * nb * 3;
    for (int k = 0; k < nb; ++k) {
        const float bk = k < used ? b[k] : 0.f;
        out[3 * k] = bk * g0;
        out[3 * k + 1] = bk * g1;
        out[3 * k + 2] = bk * g2;
    }
}

hipError_t launch_sh_forward(int n, int degree, int degrees_to_use, const float* viewdirs,
                             const float* coeffs, float* colors, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_sh_forward<<<(n + 255) / 256, 256, 0, stream>>>(n, degree, degrees_to_use, viewdirs, coeffs,
                                                     colors);
    return hipGetLastError();
}

hipError_t launch_sh_backward(int n, int degree, int degrees_to_use, const float* viewdirs,
                              const float* v_colors, float* v_coeffs, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_sh_backward<<<(n + 255) / 256, 256, 0, stream>>>(n, degree, degrees_to_use, viewdirs,
                                                      v_colors, v_coeffs);
    return hipGetLastError();
}

int sh_num_bases(int degree) { return num_sh_bases(degree); }

}  // namespace gstex
