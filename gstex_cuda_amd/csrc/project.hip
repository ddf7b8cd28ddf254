// Per-Gaussian kernels: screen AABB, depth, tile counts and the splat records
// consumed by the raster kernels.  One lane per Gaussian; these kernels stream
// N x (40..128) bytes and are HBM-bound.
#include "gstex_common.hpp"
#include "gstex_kernels.hpp"

namespace gstex {

// Screen AABB of the four corners mean +- 3*g*s.x*ax1 +- 3*g*s.y*ax2, each
// projected with view z clamped to >= 0.01; a mean at view z <= 0.01 gives
// (projected mean, zero extent).  Semantics of get_aabb_2d.cu:11-89.  No
// multiply-add is fused here: the AABB and depth become tile counts and sort
// keys, so they keep the written float32 operation order exactly.
__device__ __forceinline__ void aabb_one(const float* __restrict__ means,
                                         const float* __restrict__ scales, float glob_scale,
                                         const float* __restrict__ quats, const float* vm,
                                         float fx, float fy, float cx, float cy, int i,
                                         float& ccx, float& ccy, float& ex, float& ey, float& depth) {
#pragma clang fp contract(off)
    const float t_near = 0.01f;
    V3 p = v3(means[3 * i], means[3 * i + 1], means[3 * i + 2]);
    V3 pv = xform43_strict(vm, p);
    depth = pv.z;
    const bool clipped = pv.z <= t_near;
    const float4 q = reinterpret_cast<const float4*>(quats)[i];
    V3 a1, a2, a3;
    quat_axes_strict(q.x, q.y, q.z, q.w, a1, a2, a3);
    const float ell = 3.0f * glob_scale;
    const float e1 = ell * scales[3 * i], e2 = ell * scales[3 * i + 1];
    const V3 d1 = v3(e1 * a1.x, e1 * a1.y, e1 * a1.z);
    const V3 d2 = v3(e2 * a2.x, e2 * a2.y, e2 * a2.z);
    float mxx = -INFINITY, mxy = -INFINITY, mnx = INFINITY, mny = INFINITY;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float s1 = (c & 2) ? -1.f : 1.f, s2 = (c & 1) ? -1.f : 1.f;
        V3 corner = v3(p.x + s1 * d1.x + s2 * d2.x, p.y + s1 * d1.y + s2 * d2.y,
                       p.z + s1 * d1.z + s2 * d2.z);
        V3 cv = xform43_strict(vm, corner);
        cv.z = fmaxf(cv.z, t_near);
        float u, v;
        project_pix_strict(fx, fy, cx, cy, cv, u, v);
        mxx = fmaxf(mxx, u);
        mxy = fmaxf(mxy, v);
        mnx = fminf(mnx, u);
        mny = fminf(mny, v);
    }
    if (clipped) {
        project_pix_strict(fx, fy, cx, cy, pv, ccx, ccy);
        ex = 0.f;
        ey = 0.f;
    } else {
        ccx = 0.5f * (mxx + mnx);
        ccy = 0.5f * (mxy + mny);
        ex = 0.5f * (mxx - mnx);
        ey = 0.5f * (mxy - mny);
    }
}

// Tile count of one AABB with the float32 arithmetic of get_aabb_2d.py:70-92.
__device__ __forceinline__ int tiles_hit_one(float ccx, float ccy, float ex, float ey, int bw,
                                             int tiles_x, int tiles_y) {
    const float fbw = (float)bw;
    int lox = f2i_trunc(floorf((ccx - ex) / fbw));
    int loy = f2i_trunc(floorf((ccy - ey) / fbw));
    int hix = f2i_trunc(floorf((ccx + ex) / fbw + 1.f));
    int hiy = f2i_trunc(floorf((ccy + ey) / fbw + 1.f));
    lox = min(max(lox, 0), tiles_x);
    loy = min(max(loy, 0), tiles_y);
    hix = min(max(hix, 0), tiles_x);
    hiy = min(max(hiy, 0), tiles_y);
    return (hix - lox) * (hiy - loy);
}

__global__ __launch_bounds__(256) void k_get_aabb_2d(int n, const float* __restrict__ means,
                                                     const float* __restrict__ scales,
                                                     float glob_scale,
                                                     const float* __restrict__ quats,
                                                     const float* __restrict__ viewmat, float fx,
                                                     float fy, float cx, float cy,
                                                     float* __restrict__ centers,
                                                     float* __restrict__ extents) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float ccx, ccy, ex, ey, depth;
    aabb_one(means, scales, glob_scale, quats, viewmat, fx, fy, cx, cy, i, ccx, ccy, ex, ey, depth);
    reinterpret_cast<float2*>(centers)[i] = make_float2(ccx, ccy);
    reinterpret_cast<float2*>(extents)[i] = make_float2(ex, ey);
}

// Centres, extents, depths and tile counts in one pass; with `total`, also
// sum(tiles_hit) added into *total (zeroed beforehand).
__global__ __launch_bounds__(256) void k_project_fused(
    int n, const float* __restrict__ means, const float* __restrict__ scales, float glob_scale,
    const float* __restrict__ quats, const float* __restrict__ viewmat, float fx, float fy,
    float cx, float cy, int bw, int tiles_x, int tiles_y, float* __restrict__ centers,
    float* __restrict__ extents, float* __restrict__ depths, int* __restrict__ tiles_hit,
    uint32_t* __restrict__ total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int hit = 0;
    if (i < n) {
        float ccx, ccy, ex, ey, depth;
        aabb_one(means, scales, glob_scale, quats, viewmat, fx, fy, cx, cy, i, ccx, ccy, ex, ey,
                 depth);
        reinterpret_cast<float2*>(centers)[i] = make_float2(ccx, ccy);
        reinterpret_cast<float2*>(extents)[i] = make_float2(ex, ey);
        depths[i] = depth;
        hit = tiles_hit_one(ccx, ccy, ex, ey, bw, tiles_x, tiles_y);
        tiles_hit[i] = hit;
    }
    if (total) {
        // the intersect total: a workgroup sum, one atomic per workgroup
        __shared__ uint32_t wsum[256 / kWave];
        uint32_t v = (uint32_t)max(hit, 0);
#pragma unroll
        for (int off = kWave / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
#pragma unroll
            for (int w = 0; w < 256 / kWave; ++w) t += wsum[w];
            if (t) atomicAdd(total, t);
        }
    }
}

__global__ __launch_bounds__(256) void k_num_tiles_hit(int n, const float* __restrict__ centers,
                                                       const float* __restrict__ extents, int bw,
                                                       int tiles_x, int tiles_y,
                                                       int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 c = reinterpret_cast<const float2*>(centers)[i];
    const float2 e = reinterpret_cast<const float2*>(extents)[i];
    out[i] = tiles_hit_one(c.x, c.y, e.x, e.y, bw, tiles_x, tiles_y);
}

__global__ __launch_bounds__(256) void k_project_points(int n, const float* __restrict__ points,
                                                        const float* __restrict__ viewmat,
                                                        float fx, float fy, float cx, float cy,
                                                        int clip, float* __restrict__ pix,
                                                        float* __restrict__ depths) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 pv = xform43_strict(viewmat, v3(points[3 * i], points[3 * i + 1], points[3 * i + 2]));
    if (clip) pv.z = fminf(fmaxf(pv.z, 0.01f), 1000.0f);
    float u, v;
    project_pix_strict(fx, fy, cx, cy, pv, u, v);
    reinterpret_cast<float2*>(pix)[i] = make_float2(u, v);
    depths[i] = pv.z;
}

// Splat records: axes from the (unnormalised) quaternion exactly as the
// raster loop of texture.cu:166-169 derives them, plus every per-Gaussian
// field the raster needs, in one 128-byte line.
// Inverse of a 3x3 matrix by the adjugate; false when (nearly) singular.
__device__ bool inv3d(const double M[3][3], double Mi[3][3]) {
    const double c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
    const double c01 = M[0][2] * M[2][1] - M[0][1] * M[2][2];
    const double c02 = M[0][1] * M[1][2] - M[0][2] * M[1][1];
    const double c10 = M[1][2] * M[2][0] - M[1][0] * M[2][2];
    const double c11 = M[0][0] * M[2][2] - M[0][2] * M[2][0];
    const double c12 = M[0][2] * M[1][0] - M[0][0] * M[1][2];
    const double c20 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
    const double c21 = M[0][1] * M[2][0] - M[0][0] * M[2][1];
    const double c22 = M[0][0] * M[1][1] - M[0][1] * M[1][0];
    const double det = M[0][0] * c00 + M[0][1] * c10 + M[0][2] * c20;
    double scale = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) scale = fmax(scale, fabs(M[i][j]));
    if (!(fabs(det) > 1e-12 * scale * scale * scale)) return false;
    const double id = 1.0 / det;
    Mi[0][0] = c00 * id; Mi[0][1] = c01 * id; Mi[0][2] = c02 * id;
    Mi[1][0] = c10 * id; Mi[1][1] = c11 * id; Mi[1][2] = c12 * id;
    Mi[2][0] = c20 * id; Mi[2][1] = c21 * id; Mi[2][2] = c22 * id;
    return true;
}

// Screen box (pixel index ranges, int16 pairs) outside of which the Gaussian
// cannot pass the raster's alpha test, so a wave whose pixels all lie outside
// may skip it.  alpha = min(0.99, opac * exp(-sigma)) >= 1/255 needs
// sigma <= ln(255 opac): on the splat plane that is the ellipse
// (l1/s1)^2 + (l2/s2)^2 <= 2 ln(255 opac), inside the rectangle
// m +- r s1 ax1 +- r s2 ax2.  A pixel whose ray hits plane point P is the
// projection of P in the ray model of the raster (c2w, fx, fy, cx, cy), so the
// box of the four projected corners holds every such pixel when all corners
// lie in front of the camera.  The blur alpha (settings bit 9) adds a disc of
// radius sqrt(ln(255 opac)) pixels around the projected mean.  The radius is
// padded by 2% and the box by 2 pixels.  No box (the full range) is given when
// the bound does not hold: non-orthonormal axes, a corner behind or at the
// camera, the camera within the plane's clamp distance, a singular c2w, or a
// non-finite input; an empty box when opac < 1/255 (never visible).
__device__ uint2 cull_box(V3 m, V3 a1, V3 a2, V3 a3, float s1, float s2, float opac,
                          const float* c2w, float fx, float fy, float cx, float cy, float pu,
                          float pv, bool& bounded) {
    bounded = false;
    const uint2 all = make_uint2(pack_i16(-32768, 32767), pack_i16(-32768, 32767));
    const uint2 none = make_uint2(pack_i16(32767, -32768), pack_i16(32767, -32768));
    // non-finite or zero scales / opacity give NaN alphas (0.99 after the clamp)
    if (!(fabsf(opac) < 3e38f) || !(fabsf(s1) < 3e38f) || !(fabsf(s2) < 3e38f) ||
        !(fabsf(1.f / s1) < 3e38f) || !(fabsf(1.f / s2) < 3e38f))
        return all;
    if (opac < 1.f / 255.f) return none;
    const float o11 = dot3(a1, a1) - 1.f, o22 = dot3(a2, a2) - 1.f, o33 = dot3(a3, a3) - 1.f;
    const float o12 = dot3(a1, a2), o13 = dot3(a1, a3), o23 = dot3(a2, a3);
    const float tol = 1e-4f;
    if (!(fabsf(o11) < tol && fabsf(o22) < tol && fabsf(o33) < tol && fabsf(o12) < tol &&
          fabsf(o13) < tol && fabsf(o23) < tol))
        return all;
    const V3 o = v3(c2w[3], c2w[7], c2w[11]);
    // ray/plane denominators are clamped at 1e-6 and t is kept only up to 1000,
    // so a plane passing within 1e-3 of the camera is left unbounded
    if (!(fabsf(dot3(a3, sub3(m, o))) > 2e-3f)) return all;
    // inverse of the c2w rotation block (adjugate / determinant)
    const float r00 = c2w[0], r01 = c2w[1], r02 = c2w[2];
    const float r10 = c2w[4], r11 = c2w[5], r12 = c2w[6];
    const float r20 = c2w[8], r21 = c2w[9], r22 = c2w[10];
    const float c00 = r11 * r22 - r12 * r21, c01 = r02 * r21 - r01 * r22, c02 = r01 * r12 - r02 * r11;
    const float c10 = r12 * r20 - r10 * r22, c11 = r00 * r22 - r02 * r20, c12 = r02 * r10 - r00 * r12;
    const float c20 = r10 * r21 - r11 * r20, c21 = r01 * r20 - r00 * r21, c22 = r00 * r11 - r01 * r10;
    const float det = r00 * c00 + r01 * c10 + r02 * c20;
    if (!(fabsf(det) > 1e-12f)) return all;
    const float id = 1.f / det;
    const float r = 1.02f * sqrtf(2.f * logf(255.f * opac)) + 1e-3f;
    float umin = 3e38f, umax = -3e38f, vmin = 3e38f, vmax = -3e38f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float e1 = (k & 1) ? r * s1 : -r * s1, e2 = (k & 2) ? r * s2 : -r * s2;
        const V3 d = v3(m.x + e1 * a1.x + e2 * a2.x - o.x, m.y + e1 * a1.y + e2 * a2.y - o.y,
                        m.z + e1 * a1.z + e2 * a2.z - o.z);
        const float qx = id * (c00 * d.x + c01 * d.y + c02 * d.z);
        const float qy = id * (c10 * d.x + c11 * d.y + c12 * d.z);
        const float qz = id * (c20 * d.x + c21 * d.y + c22 * d.z);
        if (!(qz > 1e-6f)) return all;
        const float u = fx * qx / qz + cx - 0.5f, v = fy * qy / qz + cy - 0.5f;
        umin = fminf(umin, u); umax = fmaxf(umax, u);
        vmin = fminf(vmin, v); vmax = fmaxf(vmax, v);
    }
    // blur disc around the projected mean (pixel index coordinates)
    const float rb = sqrtf(logf(255.f * opac)) + 0.05f;
    umin = fminf(umin, pu - 0.5f - rb); umax = fmaxf(umax, pu - 0.5f + rb);
    vmin = fminf(vmin, pv - 0.5f - rb); vmax = fmaxf(vmax, pv - 0.5f + rb);
    if (!(umin > -30000.f && umax < 30000.f && vmin > -30000.f && vmax < 30000.f)) {
        // clamp far boxes (still conservative)
        umin = fmaxf(umin, -30000.f); vmin = fmaxf(vmin, -30000.f);
        umax = fminf(umax, 30000.f); vmax = fminf(vmax, 30000.f);
        if (!(umin <= umax && vmin <= vmax)) return all;
    }
    bounded = true;
    return make_uint2(pack_i16((int)floorf(umin) - 2, (int)ceilf(umax) + 2),
                      pack_i16((int)floorf(vmin) - 2, (int)ceilf(vmax) + 2));
}

// Screen ellipse holding every pixel whose ray can pass the alpha test (blur
// off), for a per-wave test tighter than the box.  On the splat plane that
// region is the ellipse (l1 / (r s1))^2 + (l2 / (r s2))^2 <= 1 with r as in
// cull_box.  Plane coordinates (l1, l2, 1) map to homogeneous pixel
// coordinates by G = K Rinv [a1 a2 m-o] (the raster's ray model: pixel
// (fx x / z + cx, fy y / z + cy) for camera-frame point (x, y, z)), so pixel p
// lies on a ray line meeting the ellipse iff
//   [p 1] G^-T diag(q1, q2, -1) G^-1 [p 1]^T <= 0,   q = 1 / (r s)^2.
// When that conic is a bounded ellipse it is returned as centre (x, y) and
// normalised shape (A, B, C): (p - c)^T [[A, B], [B, C]] (p - c) <= 1, in
// continuous pixel coordinates (pixel index j has centre j + 0.5).  Float64
// throughout.  A = -1 means "use the box only": cull_box found no bound, the
// conic is not an ellipse (the region reaches the camera's vanishing line),
// its eigenvalue ratio exceeds 4e4 (a float32 evaluation of the quadratic
// would lose the 2% radius margin to cancellation), or a value is not finite.
// The raster's own float32 error, mapped to the screen, is about 1e-7 * focal
// pixels, which the 0.1-pixel pad of the wave rectangle covers.
__device__ void cull_conic(V3 m, V3 a1, V3 a2, float s1, float s2, float opac, const float* c2w,
                           float fx, float fy, float cx, float cy, float4& ell, float& ell_c) {
    ell = make_float4(0.f, 0.f, -1.f, 0.f);
    ell_c = 0.f;
    double R[3][3] = {{c2w[0], c2w[1], c2w[2]}, {c2w[4], c2w[5], c2w[6]},
                      {c2w[8], c2w[9], c2w[10]}};
    double Ri[3][3];
    if (!inv3d(R, Ri)) return;
    const double mo[3] = {(double)m.x - c2w[3], (double)m.y - c2w[7], (double)m.z - c2w[11]};
    const double cols[3][3] = {{a1.x, a1.y, a1.z}, {a2.x, a2.y, a2.z}, {mo[0], mo[1], mo[2]}};
    double G[3][3];
    for (int j = 0; j < 3; ++j) {
        double x[3];
        for (int i = 0; i < 3; ++i)
            x[i] = Ri[i][0] * cols[j][0] + Ri[i][1] * cols[j][1] + Ri[i][2] * cols[j][2];
        G[0][j] = fx * x[0] + cx * x[2];
        G[1][j] = fy * x[1] + cy * x[2];
        G[2][j] = x[2];
    }
    double Gi[3][3];
    if (!inv3d(G, Gi)) return;
    const double r = 1.02 * sqrt(2.0 * log(255.0 * (double)opac)) + 1e-3;
    const double d[3] = {1.0 / (r * s1 * r * s1), 1.0 / (r * s2 * r * s2), -1.0};
    double Q[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Q[i][j] = d[0] * Gi[0][i] * Gi[0][j] + d[1] * Gi[1][i] * Gi[1][j] +
                      d[2] * Gi[2][i] * Gi[2][j];
    const double det = Q[0][0] * Q[1][1] - Q[0][1] * Q[0][1];
    if (!(Q[0][0] > 0.0 && Q[1][1] > 0.0 && det > 1e-4 * Q[0][0] * Q[1][1])) return;
    const double c0 = -(Q[1][1] * Q[0][2] - Q[0][1] * Q[1][2]) / det;
    const double c1 = -(Q[0][0] * Q[1][2] - Q[0][1] * Q[0][2]) / det;
    const double f = Q[2][2] + Q[0][2] * c0 + Q[1][2] * c1;
    if (!(f < 0.0)) return;
    const double A = Q[0][0] / -f, B = Q[0][1] / -f, C = Q[1][1] / -f;
    if (!(fabs(c0) < 1e7 && fabs(c1) < 1e7 && A < 1e30 && C < 1e30 && fabs(B) < 1e30 &&
          A > 1e-30 && C > 1e-30))
        return;
    ell = make_float4((float)c0, (float)c1, (float)A, (float)B);
    ell_c = (float)C;
}

__global__ __launch_bounds__(256) void k_build_records(
    int n, const float* __restrict__ means, const float* __restrict__ scales, float glob_scale,
    const float* __restrict__ quats, const float* __restrict__ colors,
    const float* __restrict__ opacities, const float* __restrict__ uv0,
    const float* __restrict__ umap, const float* __restrict__ vmap,
    const int* __restrict__ texture_dims, const float* __restrict__ viewmat,
    const float* __restrict__ c2w, float fx, float fy, float cx, float cy,
    float* __restrict__ records, uint2* __restrict__ boxes, float4* __restrict__ ellipses,
    float* __restrict__ ellipse_c) {
    // records go out through LDS: each thread writes its 32 floats there
    // (row stride 33 floats, so the 64 lanes of a wave hit distinct banks) and
    // the block's 256 records leave as contiguous 16-byte stores
    __shared__ float srec[256][kRecordFloats + 1];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int i0 = blockIdx.x * blockDim.x, tid = threadIdx.x;
    if (i < n) {
    const float4 q = reinterpret_cast<const float4*>(quats)[i];
    V3 a1, a2, a3;
    quat_axes(q.x, q.y, q.z, q.w, a1, a2, a3);
    const V3 m = v3(means[3 * i], means[3 * i + 1], means[3 * i + 2]);
    const V3 pv = xform43(viewmat, m);
    float pu, pvv;
    project_pix(fx, fy, cx, cy, pv, pu, pvv);
    float r[kRecordFloats];
    r[R_AX1 + 0] = a1.x; r[R_AX1 + 1] = a1.y; r[R_AX1 + 2] = a1.z;
    const float is1 = 1.f / (scales[3 * i] * glob_scale);
    const float is2 = 1.f / (scales[3 * i + 1] * glob_scale);
    r[R_IQ1] = is1 * is1;
    r[R_AX2 + 0] = a2.x; r[R_AX2 + 1] = a2.y; r[R_AX2 + 2] = a2.z;
    r[R_IQ2] = is2 * is2;
    r[R_AX3 + 0] = a3.x; r[R_AX3 + 1] = a3.y; r[R_AX3 + 2] = a3.z;
    r[R_OPAC] = opacities[i];
    const V3 o = v3(c2w[3], c2w[7], c2w[11]);
    const V3 mo = sub3(m, o);
    r[R_MO + 0] = mo.x; r[R_MO + 1] = mo.y; r[R_MO + 2] = mo.z;
    r[R_PIX_X] = pu;
    // colours may be absent (texture_edit does not read them)
    r[R_RGB + 0] = colors ? colors[3 * i] : 0.f;
    r[R_RGB + 1] = colors ? colors[3 * i + 1] : 0.f;
    r[R_RGB + 2] = colors ? colors[3 * i + 2] : 0.f;
    r[R_PIX_Y] = pvv;
    r[R_UV0 + 0] = uv0[2 * i]; r[R_UV0 + 1] = uv0[2 * i + 1];
    r[R_UMAP + 0] = umap[3 * i]; r[R_UMAP + 1] = umap[3 * i + 1]; r[R_UMAP + 2] = umap[3 * i + 2];
    r[R_VMAP + 0] = vmap[3 * i]; r[R_VMAP + 1] = vmap[3 * i + 1]; r[R_VMAP + 2] = vmap[3 * i + 2];
    r[R_TEX_H] = __int_as_float(texture_dims[3 * i]);
    r[R_TEX_W] = __int_as_float(texture_dims[3 * i + 1]);
    r[R_TEX_START] = __int_as_float(texture_dims[3 * i + 2]);
    r[R_NUM] = dot3(a3, mo);
#pragma unroll
    for (int k = 0; k < kRecordFloats; ++k) srec[tid][k] = r[k];
    const float s1 = scales[3 * i] * glob_scale, s2 = scales[3 * i + 1] * glob_scale;
    bool bounded;
    boxes[i] = cull_box(m, a1, a2, a3, s1, s2, r[R_OPAC], c2w, fx, fy, cx, cy, pu, pvv, bounded);
    float4 ell = make_float4(0.f, 0.f, -1.f, 0.f);
    float ell_c = 0.f;
    if (bounded) cull_conic(m, a1, a2, s1, s2, r[R_OPAC], c2w, fx, fy, cx, cy, ell, ell_c);
    ellipses[i] = ell;
    ellipse_c[i] = ell_c;
    }
    __syncthreads();
    const int cnt = min(256, n - i0);
    float4* dst = reinterpret_cast<float4*>(records + (size_t)i0 * kRecordFloats);
#pragma unroll
    for (int k = 0; k < kRecordFloats / 4; ++k) {
        const int f = k * 256 + tid;  // 16-byte word of the block's records
        const int rr = f / (kRecordFloats / 4), c = 4 * (f % (kRecordFloats / 4));
        if (rr < cnt)
            dst[f] = make_float4(srec[rr][c], srec[rr][c + 1], srec[rr][c + 2], srec[rr][c + 3]);
    }
}

hipError_t launch_get_aabb_2d(int n, const float* means, const float* scales, float glob_scale,
                              const float* quats, const float* viewmat, float fx, float fy,
                              float cx, float cy, float* centers, float* extents,
                              hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_get_aabb_2d<<<(n + 255) / 256, 256, 0, stream>>>(n, means, scales, glob_scale, quats, viewmat,
                                                      fx, fy, cx, cy, centers, extents);
    return hipGetLastError();
}

hipError_t launch_project_fused(int n, const float* means, const float* scales, float glob_scale,
                                const float* quats, const float* viewmat, float fx, float fy,
                                float cx, float cy, int bw, int tiles_x, int tiles_y,
                                float* centers, float* extents, float* depths, int* tiles_hit,
                                uint32_t* total, hipStream_t stream) {
    if (total) {
        hipError_t e = fill_u32(total, 1, 0u, stream);
        if (e != hipSuccess) return e;
    }
    if (n <= 0) return hipSuccess;
    k_project_fused<<<(n + 255) / 256, 256, 0, stream>>>(n, means, scales, glob_scale, quats,
                                                        viewmat, fx, fy, cx, cy, bw, tiles_x,
                                                        tiles_y, centers, extents, depths,
                                                        tiles_hit, total);
    return hipGetLastError();
}

hipError_t launch_num_tiles_hit(int n, const float* centers, const float* extents, int bw,
                                int tiles_x, int tiles_y, int* out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_num_tiles_hit<<<(n + 255) / 256, 256, 0, stream>>>(n, centers, extents, bw, tiles_x, tiles_y,
                                                        out);
    return hipGetLastError();
}

hipError_t launch_project_points(int n, const float* points, const float* viewmat, float fx,
                                 float fy, float cx, float cy, int clip, float* pix,
                                 float* depths, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_project_points<<<(n + 255) / 256, 256, 0, stream>>>(n, points, viewmat, fx, fy, cx, cy, clip,
                                                         pix, depths);
    return hipGetLastError();
}

hipError_t launch_build_records(int n, const float* means, const float* scales, float glob_scale,
                                const float* quats, const float* colors, const float* opacities,
                                const float* uv0, const float* umap, const float* vmap,
                                const int* texture_dims, const float* viewmat, const float* c2w,
                                float fx, float fy, float cx, float cy, float* records,
                                hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_build_records<<<(n + 255) / 256, 256, 0, stream>>>(
        n, means, scales, glob_scale, quats, colors, opacities, uv0, umap, vmap, texture_dims,
        viewmat, c2w, fx, fy, cx, cy, records, record_boxes(records, n),
        record_conics(records, n), record_conic_c(records, n));
    return hipGetLastError();
}

}  // namespace gstex
