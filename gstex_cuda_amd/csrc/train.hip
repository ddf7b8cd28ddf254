// Training-step glue around the raster path: the per-Gaussian parameter
// preprocessing of a GStex training step (raw parameters -> the reference's
// render inputs), its adjoint, the image-loss cotangents and a fused Adam
// update.
//
// Preprocessing follows example.py:126-143:
//   scales.xy = exp(log_scales.xy), scales.z = 1e-5 * mean(scales.xy) (no gradient)
//   quats     = q / |q|;  R = normalised rotation matrix (_torch_impl.py:116-133)
//   uv0       = mapping[0:2];  uvscale = exp(mapping[2]);  theta = mapping[3]
//   umap      = uvscale * ( ax1 cos(theta) + ax2 sin(theta))
//   vmap      = uvscale * (-ax1 sin(theta) + ax2 cos(theta))
//   colors    = sigmoid(rgb logits), opacities = sigmoid(opacity logits)
// (texture values = sigmoid(texture logits) come from the Adam kernel's value
// output, see k_adam).  The adam update follows torch.optim.Adam's multi-tensor
// path (example.py:231-233,277 use it with default betas and eps):
//   m = lerp(m, g, 1 - b1);  v = v * b2 + (1 - b2) g g
//   p = p - (lr / bc1) * m / (sqrt(v) / sqrt(bc2) + eps),  bc_k = 1 - b_k^step
#include <cstdint>

#include "gstex_common.hpp"
#include "gstex_kernels.hpp"

namespace gstex {

__device__ __forceinline__ float sigmoidf(float x) { return 1.f / (1.f + expf(-x)); }

// Axes (columns 0 and 1 of R) of a unit quaternion (w, x, y, z).
__device__ __forceinline__ void unit_quat_axes(const float q[4], V3& a1, V3& a2) {
    V3 a3;
    quat_axes(q[0], q[1], q[2], q[3], a1, a2, a3);
}

struct PreIn {
    const float *log_scales, *quats_raw, *mapping, *rgb_logits, *opacity_logits;
};

__device__ __forceinline__ void load_unit_quat(const float* quats_raw, int i, float qn[4],
                                               float& norm) {
    const float4 q = reinterpret_cast<const float4*>(quats_raw)[i];
    norm = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    qn[0] = q.x / norm;
    qn[1] = q.y / norm;
    qn[2] = q.z / norm;
    qn[3] = q.w / norm;
}

__global__ __launch_bounds__(256) void k_preprocess(int n, PreIn in, float* __restrict__ scales,
                                                    float* __restrict__ quats,
                                                    float* __restrict__ uv0,
                                                    float* __restrict__ umap,
                                                    float* __restrict__ vmap,
                                                    float* __restrict__ colors,
                                                    float* __restrict__ opacities) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float sx = expf(in.log_scales[3 * i]), sy = expf(in.log_scales[3 * i + 1]);
    scales[3 * i] = sx;
    scales[3 * i + 1] = sy;
    scales[3 * i + 2] = 1e-5f * ((sx + sy) / 2.f);
    float qn[4], norm;
    load_unit_quat(in.quats_raw, i, qn, norm);
    reinterpret_cast<float4*>(quats)[i] = make_float4(qn[0], qn[1], qn[2], qn[3]);
    V3 a1, a2;
    unit_quat_axes(qn, a1, a2);
    const float4 mp = reinterpret_cast<const float4*>(in.mapping)[i];
    uv0[2 * i] = mp.x;
    uv0[2 * i + 1] = mp.y;
    const float uvs = expf(mp.z), c = cosf(mp.w), s = sinf(mp.w);
    umap[3 * i] = uvs * (a1.x * c + a2.x * s);
    umap[3 * i + 1] = uvs * (a1.y * c + a2.y * s);
    umap[3 * i + 2] = uvs * (a1.z * c + a2.z * s);
    vmap[3 * i] = uvs * (-a1.x * s + a2.x * c);
    vmap[3 * i + 1] = uvs * (-a1.y * s + a2.y * c);
    vmap[3 * i + 2] = uvs * (-a1.z * s + a2.z * c);
    if (colors) {
        colors[3 * i] = sigmoidf(in.rgb_logits[3 * i]);
        colors[3 * i + 1] = sigmoidf(in.rgb_logits[3 * i + 1]);
        colors[3 * i + 2] = sigmoidf(in.rgb_logits[3 * i + 2]);
    }
    opacities[i] = sigmoidf(in.opacity_logits[i]);
}

struct PreGrad {
    const float *v_scales, *v_quats, *v_uv0, *v_umap, *v_vmap, *v_colors, *v_opacities;
};

// Adjoint of k_preprocess: reference-layout gradients -> raw-parameter
// gradients.  scales.z is detached, so log_scales.z gets 0.  The rotation
// reaches umap / vmap through ax1, ax2; its adjoint joins the gradient of the
// unit quaternion, which the normalisation maps back by (g - qn (qn . g)) / |q|.
__global__ __launch_bounds__(256) void k_preprocess_backward(
    int n, PreIn in, PreGrad g, float* __restrict__ v_log_scales, float* __restrict__ v_quats_raw,
    float* __restrict__ v_mapping, float* __restrict__ v_rgb_logits,
    float* __restrict__ v_opacity_logits) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float sx = expf(in.log_scales[3 * i]), sy = expf(in.log_scales[3 * i + 1]);
    v_log_scales[3 * i] = g.v_scales[3 * i] * sx;
    v_log_scales[3 * i + 1] = g.v_scales[3 * i + 1] * sy;
    v_log_scales[3 * i + 2] = 0.f;
    float qn[4], norm;
    load_unit_quat(in.quats_raw, i, qn, norm);
    V3 a1, a2;
    unit_quat_axes(qn, a1, a2);
    const float4 mp = reinterpret_cast<const float4*>(in.mapping)[i];
    const float uvs = expf(mp.z), c = cosf(mp.w), s = sinf(mp.w);
    const V3 gu = v3(g.v_umap[3 * i], g.v_umap[3 * i + 1], g.v_umap[3 * i + 2]);
    const V3 gv = v3(g.v_vmap[3 * i], g.v_vmap[3 * i + 1], g.v_vmap[3 * i + 2]);
    const V3 g1 = v3(uvs * (c * gu.x - s * gv.x), uvs * (c * gu.y - s * gv.y),
                     uvs * (c * gu.z - s * gv.z));
    const V3 g2 = v3(uvs * (s * gu.x + c * gv.x), uvs * (s * gu.y + c * gv.y),
                     uvs * (s * gu.z + c * gv.z));
    float gq[4];
    quat_axes_vjp(qn[0], qn[1], qn[2], qn[3], g1, g2, v3(0.f, 0.f, 0.f), gq);
    const float4 vq = reinterpret_cast<const float4*>(g.v_quats)[i];
    gq[0] += vq.x;
    gq[1] += vq.y;
    gq[2] += vq.z;
    gq[3] += vq.w;
    const float d = qn[0] * gq[0] + qn[1] * gq[1] + qn[2] * gq[2] + qn[3] * gq[3];
    reinterpret_cast<float4*>(v_quats_raw)[i] =
        make_float4((gq[0] - qn[0] * d) / norm, (gq[1] - qn[1] * d) / norm,
                    (gq[2] - qn[2] * d) / norm, (gq[3] - qn[3] * d) / norm);
    if (v_mapping) {
        const V3 um = v3(uvs * (a1.x * c + a2.x * s), uvs * (a1.y * c + a2.y * s),
                         uvs * (a1.z * c + a2.z * s));
        const V3 vm = v3(uvs * (-a1.x * s + a2.x * c), uvs * (-a1.y * s + a2.y * c),
                         uvs * (-a1.z * s + a2.z * c));
        reinterpret_cast<float4*>(v_mapping)[i] =
            make_float4(g.v_uv0[2 * i], g.v_uv0[2 * i + 1], dot3(gu, um) + dot3(gv, vm),
                        dot3(gu, vm) - dot3(gv, um));
    }
    if (v_rgb_logits) {
        for (int k = 0; k < 3; ++k) {
            const float y = sigmoidf(in.rgb_logits[3 * i + k]);
            v_rgb_logits[3 * i + k] = g.v_colors[3 * i + k] * (1.f - y) * y;
        }
    }
    const float yo = sigmoidf(in.opacity_logits[i]);
    v_opacity_logits[i] = g.v_opacities[i] * (1.f - yo) * yo;
}

// float -> bfloat16 bits, round to nearest even (NaN stays NaN).
__device__ __forceinline__ unsigned short bf16_rne(float x) {
    const unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// One Adam step over n values.  grad_mode 1 treats grad as the gradient of
// sigmoid(param) and chains it to the logit (g * (1 - s) * s, s = sigmoid of
// the old param).  values (if given) receives sigmoid(new param) as float32
// (format 0) or bfloat16 (format 1): the next step's render input.
__global__ __launch_bounds__(256) void k_adam(size_t n, float* __restrict__ param,
                                              const float* __restrict__ grad,
                                              float* __restrict__ exp_avg,
                                              float* __restrict__ exp_avg_sq, float one_m_b1,
                                              float b2, float one_m_b2, float bc2_sqrt,
                                              float eps, float neg_step_size, int grad_mode,
                                              void* __restrict__ values, int values_bf16) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float p = param[i], gr = grad[i];
        if (grad_mode == 1) {
            const float y = sigmoidf(p);
            gr = gr * (1.f - y) * y;
        }
        float m = exp_avg[i];
        m = m + one_m_b1 * (gr - m);
        float v = exp_avg_sq[i] * b2;
        v = v + one_m_b2 * gr * gr;
        const float denom = sqrtf(v) / bc2_sqrt + eps;
        p = p + neg_step_size * (m / denom);
        param[i] = p;
        exp_avg[i] = m;
        exp_avg_sq[i] = v;
        if (values) {
            const float y = sigmoidf(p);
            if (values_bf16) static_cast<unsigned short*>(values)[i] = bf16_rne(y);
            else static_cast<float*>(values)[i] = y;
        }
    }
}

hipError_t preprocess_gaussians(int n, const float* log_scales, const float* quats_raw,
                                const float* mapping, const float* rgb_logits,
                                const float* opacity_logits, float* scales, float* quats,
                                float* uv0, float* umap, float* vmap, float* colors,
                                float* opacities, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const PreIn in{log_scales, quats_raw, mapping, rgb_logits, opacity_logits};
    k_preprocess<<<(n + 255) / 256, 256, 0, s>>>(n, in, scales, quats, uv0, umap, vmap,
                                                 rgb_logits ? colors : nullptr, opacities);
    return hipGetLastError();
}

hipError_t preprocess_gaussians_backward(
    int n, const float* log_scales, const float* quats_raw, const float* mapping,
    const float* rgb_logits, const float* opacity_logits, const float* v_scales,
    const float* v_quats, const float* v_uv0, const float* v_umap, const float* v_vmap,
    const float* v_colors, const float* v_opacities, float* v_log_scales, float* v_quats_raw,
    float* v_mapping, float* v_rgb_logits, float* v_opacity_logits, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const PreIn in{log_scales, quats_raw, mapping, rgb_logits, opacity_logits};
    const PreGrad g{v_scales, v_quats, v_uv0, v_umap, v_vmap, v_colors, v_opacities};
    k_preprocess_backward<<<(n + 255) / 256, 256, 0, s>>>(
        n, in, g, v_log_scales, v_quats_raw, v_mapping, rgb_logits ? v_rgb_logits : nullptr,
        v_opacity_logits);
    return hipGetLastError();
}

// Per pixel: d = (with_img ? out_img + tex[:3] : tex[:3]) - target, rounded
// as torch's add / sub / mul steps are; v_img = with_img ? scale d : 0,
// v_texture_out[:3] = scale d and 0 above; v_normal = normal_scale n with
// normal_scale subtracted from z (when out_normal is given).
__global__ __launch_bounds__(256) void k_image_loss_grad(
    int npix, int C, int with_img, const float* __restrict__ out_img,
    const float* __restrict__ out_tex, const float* __restrict__ target, float scale,
    const float* __restrict__ out_normal, float normal_scale, float* __restrict__ v_img,
    float* __restrict__ v_tex, float* __restrict__ v_normal) {
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float t = c < C ? out_tex[(size_t)p * C + c] : 0.f;
        const float img = with_img ? out_img[3 * p + c] + t : t;
        const float g = (img - target[3 * p + c]) * scale;
        v_img[3 * p + c] = with_img ? g : 0.f;
        if (c < C) v_tex[(size_t)p * C + c] = g;
    }
    for (int c = 3; c < C; ++c) v_tex[(size_t)p * C + c] = 0.f;
    if (out_normal) {
        v_normal[3 * p] = out_normal[3 * p] * normal_scale;
        v_normal[3 * p + 1] = out_normal[3 * p + 1] * normal_scale;
        v_normal[3 * p + 2] = out_normal[3 * p + 2] * normal_scale - normal_scale;
    }
}

// The same for three texture channels, over the flat (3 npix) arrays: four
// consecutive floats per thread, 16-byte loads and stores.
__global__ __launch_bounds__(256) void k_image_loss_grad3(
    size_t nflat, int with_img, const float* __restrict__ out_img,
    const float* __restrict__ out_tex, const float* __restrict__ target, float scale,
    const float* __restrict__ out_normal, float normal_scale, float* __restrict__ v_img,
    float* __restrict__ v_tex, float* __restrict__ v_normal) {
#pragma clang fp contract(off)
    const size_t k0 = 4 * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (k0 >= nflat) return;
    if (k0 + 4 <= nflat) {
        const float4 t = *reinterpret_cast<const float4*>(out_tex + k0);
        const float4 tg = *reinterpret_cast<const float4*>(target + k0);
        float4 d = t;
        if (with_img) {
            const float4 im = *reinterpret_cast<const float4*>(out_img + k0);
            d = make_float4(im.x + t.x, im.y + t.y, im.z + t.z, im.w + t.w);
        }
        const float4 g = make_float4((d.x - tg.x) * scale, (d.y - tg.y) * scale,
                                     (d.z - tg.z) * scale, (d.w - tg.w) * scale);
        *reinterpret_cast<float4*>(v_tex + k0) = g;
        *reinterpret_cast<float4*>(v_img + k0) = with_img ? g : make_float4(0.f, 0.f, 0.f, 0.f);
        if (out_normal) {
            const float4 nv = *reinterpret_cast<const float4*>(out_normal + k0);
            float r[4] = {nv.x * normal_scale, nv.y * normal_scale, nv.z * normal_scale,
                          nv.w * normal_scale};
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if ((k0 + j) % 3 == 2) r[j] = r[j] - normal_scale;  // z of its pixel
            *reinterpret_cast<float4*>(v_normal + k0) = make_float4(r[0], r[1], r[2], r[3]);
        }
        return;
    }
    for (size_t k = k0; k < nflat; ++k) {
        const float t = out_tex[k];
        const float d = with_img ? out_img[k] + t : t;
        const float g = (d - target[k]) * scale;
        v_tex[k] = g;
        v_img[k] = with_img ? g : 0.f;
        if (out_normal) {
            const float r = out_normal[k] * normal_scale;
            v_normal[k] = k % 3 == 2 ? r - normal_scale : r;
        }
    }
}

hipError_t image_loss_grad(int npix, int C, int with_img, const float* out_img,
                           const float* out_tex, const float* target, float scale,
                           const float* out_normal, float normal_scale, float* v_img,
                           float* v_tex, float* v_normal, hipStream_t s) {
    if (npix <= 0) return hipSuccess;
    auto aligned = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    if (C == 3 && aligned(out_tex) && aligned(target) && aligned(v_tex) && aligned(v_img) &&
        (!with_img || aligned(out_img)) && (!out_normal || (aligned(out_normal) &&
                                                            aligned(v_normal)))) {
        const size_t nflat = 3 * (size_t)npix;
        const size_t threads = (nflat + 3) / 4;
        k_image_loss_grad3<<<(unsigned)((threads + 255) / 256), 256, 0, s>>>(
            nflat, with_img, out_img, out_tex, target, scale, out_normal, normal_scale, v_img,
            v_tex, v_normal);
        return hipGetLastError();
    }
    k_image_loss_grad<<<(npix + 255) / 256, 256, 0, s>>>(npix, C, with_img, out_img, out_tex,
                                                         target, scale, out_normal, normal_scale,
                                                         v_img, v_tex, v_normal);
    return hipGetLastError();
}

hipError_t adam_step(size_t n, float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                     double lr, double beta1, double beta2, double eps, int step, int grad_mode,
                     void* values, int values_bf16, hipStream_t s) {
    if (n == 0) return hipSuccess;
    // host-side scalars in double, as torch computes them in Python
    const double bc1 = 1.0 - pow(beta1, (double)step), bc2 = 1.0 - pow(beta2, (double)step);
    const float neg_step_size = (float)(-(lr / bc1));
    const float bc2_sqrt = (float)sqrt(bc2);
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    k_adam<<<(unsigned)blocks, 256, 0, s>>>(n, param, grad, exp_avg, exp_avg_sq,
                                            (float)(1.0 - beta1), (float)beta2,
                                            (float)(1.0 - beta2), bc2_sqrt, (float)eps,
                                            neg_step_size, grad_mode, values, values_bf16);
    return hipGetLastError();
}

}  // namespace gstex
