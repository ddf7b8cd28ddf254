/*
 * CPU oracle body — TEST INFRASTRUCTURE ONLY.
 *
 * Included twice by gstex_oracle.c: once with REAL=float (the reference's own
 * FP32 arithmetic, same operation order as the CUDA kernels) and once with
 * REAL=double (the FP64 oracle used for accuracy claims).  Each function names
 * the reference file:line range it restates.  Nothing in the product path may
 * link or call this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker.
 *
 * Conventions copied from the reference because the interface requires them:
 *   - quaternions are (w, x, y, z); the rotation matrix is stored as three
 *     axes ax1, ax2, ax3 (the columns of R), helpers.cuh:166-185;
 *   - viewmat / c2w are row-major 4x4, helpers.cuh:113-131;
 *   - texture_dims rows are {h, w, start}, texel (i, j) of a chart lives at
 *     row (start + i*w + j) of the (X, C) texture, texture_helpers.cuh:162-215.
 */

#ifndef REAL
#error "REAL must be defined"
#endif

/* Input arrays of the scene, typed IN (float for _f32, double for _f64). */
#define SF(field) ((const IN *)S->field)

static inline REAL SFX(dot3)(const REAL a[3], const REAL b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* helpers.cuh:166-185 (glm column-major: axis k = column k). */
static void SFX(quat_axes)(const IN *qf, REAL ax1[3], REAL ax2[3], REAL ax3[3]) {
    REAL w = qf[0], x = qf[1], y = qf[2], z = qf[3];
    ax1[0] = (REAL)1 - (REAL)2 * (y * y + z * z);
    ax1[1] = (REAL)2 * (x * y + w * z);
    ax1[2] = (REAL)2 * (x * z - w * y);
    ax2[0] = (REAL)2 * (x * y - w * z);
    ax2[1] = (REAL)1 - (REAL)2 * (x * x + z * z);
    ax2[2] = (REAL)2 * (y * z + w * x);
    ax3[0] = (REAL)2 * (x * z + w * y);
    ax3[1] = (REAL)2 * (y * z - w * x);
    ax3[2] = (REAL)1 - (REAL)2 * (x * x + y * y);
}

/* helpers.cuh:187-228: VJP of quat_to_rotmat; vR[k] is column (axis) k. */
static void SFX(quat_axes_vjp)(const IN *qf, REAL vR[3][3], REAL vq[4]) {
    REAL w = qf[0], x = qf[1], y = qf[2], z = qf[3];
    vq[0] = (REAL)2 * (x * (vR[1][2] - vR[2][1]) + y * (vR[2][0] - vR[0][2]) +
                       z * (vR[0][1] - vR[1][0]));
    vq[1] = (REAL)2 * (-(REAL)2 * x * (vR[1][1] + vR[2][2]) + y * (vR[0][1] + vR[1][0]) +
                       z * (vR[0][2] + vR[2][0]) + w * (vR[1][2] - vR[2][1]));
    vq[2] = (REAL)2 * (x * (vR[0][1] + vR[1][0]) - (REAL)2 * y * (vR[0][0] + vR[2][2]) +
                       z * (vR[1][2] + vR[2][1]) + w * (vR[2][0] - vR[0][2]));
    vq[3] = (REAL)2 * (x * (vR[0][2] + vR[2][0]) + y * (vR[1][2] + vR[2][1]) -
                       (REAL)2 * z * (vR[0][0] + vR[1][1]) + w * (vR[0][1] - vR[1][0]));
}

/* helpers.cuh:124-131: R * p + t for a row-major 4x4 matrix. */
static void SFX(xform43)(const IN *m, const REAL p[3], REAL out[3]) {
    out[0] = (REAL)m[0] * p[0] + (REAL)m[1] * p[1] + (REAL)m[2] * p[2] + (REAL)m[3];
    out[1] = (REAL)m[4] * p[0] + (REAL)m[5] * p[1] + (REAL)m[6] * p[2] + (REAL)m[7];
    out[2] = (REAL)m[8] * p[0] + (REAL)m[9] * p[1] + (REAL)m[10] * p[2] + (REAL)m[11];
}

/* helpers.cuh:114-121: R^T * p. */
static void SFX(xform43_rot_t)(const IN *m, const REAL p[3], REAL out[3]) {
    out[0] = (REAL)m[0] * p[0] + (REAL)m[4] * p[1] + (REAL)m[8] * p[2];
    out[1] = (REAL)m[1] * p[0] + (REAL)m[5] * p[1] + (REAL)m[9] * p[2];
    out[2] = (REAL)m[2] * p[0] + (REAL)m[6] * p[1] + (REAL)m[10] * p[2];
}

/* helpers.cuh:145-152. */
static void SFX(project_pix)(REAL fx, REAL fy, const REAL pv[3], REAL cx, REAL cy, REAL out[2]) {
    REAL rw = (REAL)1 / (pv[2] + (REAL)1e-6f);
    REAL px = pv[0] * rw, py = pv[1] * rw;
    out[0] = px * fx + cx;
    out[1] = py * fy + cy;
}

/* helpers.cuh:155-164. */
static void SFX(project_pix_vjp)(REAL fx, REAL fy, const REAL pv[3], const REAL vxy[2], REAL out[3]) {
    REAL rw = (REAL)1 / (pv[2] + (REAL)1e-6f);
    REAL vpx = fx * vxy[0], vpy = fy * vxy[1];
    out[0] = vpx * rw;
    out[1] = vpy * rw;
    out[2] = -(vpx * pv[0] + vpy * pv[1]) * rw * rw;
}

/* get_aabb_2d.cu:11-89: screen AABB of the four +-3*g*s corners. */
void SFX(orc_get_aabb_2d)(int n, const IN *means, const IN *scales, double glob_scale,
                          const IN *quats, const IN *viewmat, double fx, double fy,
                          double cx, double cy, REAL *centers, REAL *extents) {
    const REAL t_near = (REAL)0.01f;
    for (int i = 0; i < n; ++i) {
        REAL p[3] = {means[3 * i], means[3 * i + 1], means[3 * i + 2]};
        REAL pv[3];
        SFX(xform43)(viewmat, p, pv);
        int clipped = pv[2] <= t_near;
        REAL xy[2];
        SFX(project_pix)((REAL)fx, (REAL)fy, pv, (REAL)cx, (REAL)cy, xy);
        REAL a1[3], a2[3], a3[3];
        SFX(quat_axes)(quats + 4 * i, a1, a2, a3);
        REAL ell = (REAL)3.0f * (REAL)glob_scale;
        REAL sx = scales[3 * i], sy = scales[3 * i + 1];
        REAL mx[2] = {-INFINITY, -INFINITY}, mn[2] = {INFINITY, INFINITY};
        for (int c = 0; c < 4; ++c) {
            REAL s1 = (c & 2) ? -(REAL)1 : (REAL)1; /* corner00, 01, 10, 11 */
            REAL s2 = (c & 1) ? -(REAL)1 : (REAL)1;
            REAL q[3], qv[3], pix[2];
            for (int k = 0; k < 3; ++k) {
                REAL t1 = ell * sx * a1[k], t2 = ell * sy * a2[k];
                q[k] = (s1 > 0 ? p[k] + t1 : p[k] - t1);
                q[k] = (s2 > 0 ? q[k] + t2 : q[k] - t2);
            }
            SFX(xform43)(viewmat, q, qv);
            if (qv[2] < t_near) qv[2] = t_near; /* clip_and_project, helpers.cuh:250-257 */
            SFX(project_pix)((REAL)fx, (REAL)fy, qv, (REAL)cx, (REAL)cy, pix);
            for (int k = 0; k < 2; ++k) {
                if (pix[k] > mx[k]) mx[k] = pix[k];
                if (pix[k] < mn[k]) mn[k] = pix[k];
            }
        }
        if (clipped) {
            centers[2 * i] = xy[0];
            centers[2 * i + 1] = xy[1];
            extents[2 * i] = 0;
            extents[2 * i + 1] = 0;
        } else {
            centers[2 * i] = (REAL)0.5f * (mx[0] + mn[0]);
            centers[2 * i + 1] = (REAL)0.5f * (mx[1] + mn[1]);
            extents[2 * i] = (REAL)0.5f * (mx[0] - mn[0]);
            extents[2 * i + 1] = (REAL)0.5f * (mx[1] - mn[1]);
        }
    }
}

/* sh.cuh:6-31 constants (textbook real SH normalisation, sign convention of the reference). */
static const REAL SFX(SH_C0) = (REAL)0.28209479177387814;
static const REAL SFX(SH_C1) = (REAL)0.4886025119029199;
static const REAL SFX(SH_C2)[5] = {(REAL)1.0925484305920792, (REAL)-1.0925484305920792,
                                   (REAL)0.31539156525252005, (REAL)-1.0925484305920792,
                                   (REAL)0.5462742152960396};
static const REAL SFX(SH_C3)[7] = {(REAL)-0.5900435899266435, (REAL)2.890611442640554,
                                   (REAL)-0.4570457994644658, (REAL)0.3731763325901154,
                                   (REAL)-0.4570457994644658, (REAL)1.445305721320277,
                                   (REAL)-0.5900435899266435};
static const REAL SFX(SH_C4)[9] = {(REAL)2.5033429417967046, (REAL)-1.7701307697799304,
                                   (REAL)0.9461746957575601, (REAL)-0.6690465435572892,
                                   (REAL)0.10578554691520431, (REAL)-0.6690465435572892,
                                   (REAL)0.47308734787878004, (REAL)-1.7701307697799304,
                                   (REAL)0.6258357354491761};

/* Basis values for degree <= 4 at the normalised direction (sh.cuh:61-116). */
static int SFX(sh_basis)(int degree, const IN *dir, REAL *b) {
    b[0] = SFX(SH_C0);
    if (degree < 1) return 1;
    REAL dx = dir[0], dy = dir[1], dz = dir[2];
    REAL norm = SQRT(dx * dx + dy * dy + dz * dz);
    REAL x = dx / norm, y = dy / norm, z = dz / norm;
    REAL xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    b[1] = -SFX(SH_C1) * y;
    b[2] = SFX(SH_C1) * z;
    b[3] = -SFX(SH_C1) * x;
    if (degree < 2) return 4;
    b[4] = SFX(SH_C2)[0] * xy;
    b[5] = SFX(SH_C2)[1] * yz;
    b[6] = SFX(SH_C2)[2] * ((REAL)2 * zz - xx - yy);
    b[7] = SFX(SH_C2)[3] * xz;
    b[8] = SFX(SH_C2)[4] * (xx - yy);
    if (degree < 3) return 9;
    b[9] = SFX(SH_C3)[0] * y * ((REAL)3 * xx - yy);
    b[10] = SFX(SH_C3)[1] * xy * z;
    b[11] = SFX(SH_C3)[2] * y * ((REAL)4 * zz - xx - yy);
    b[12] = SFX(SH_C3)[3] * z * ((REAL)2 * zz - (REAL)3 * xx - (REAL)3 * yy);
    b[13] = SFX(SH_C3)[4] * x * ((REAL)4 * zz - xx - yy);
    b[14] = SFX(SH_C3)[5] * z * (xx - yy);
    b[15] = SFX(SH_C3)[6] * x * (xx - (REAL)3 * yy);
    if (degree < 4) return 16;
    b[16] = SFX(SH_C4)[0] * xy * (xx - yy);
    b[17] = SFX(SH_C4)[1] * yz * ((REAL)3 * xx - yy);
    b[18] = SFX(SH_C4)[2] * xy * ((REAL)7 * zz - (REAL)1);
    b[19] = SFX(SH_C4)[3] * yz * ((REAL)7 * zz - (REAL)3);
    b[20] = SFX(SH_C4)[4] * (zz * ((REAL)35 * zz - (REAL)30) + (REAL)3);
    b[21] = SFX(SH_C4)[5] * xz * ((REAL)7 * zz - (REAL)3);
    b[22] = SFX(SH_C4)[6] * (xx - yy) * ((REAL)7 * zz - (REAL)1);
    b[23] = SFX(SH_C4)[7] * xz * (xx - (REAL)3 * yy);
    b[24] = SFX(SH_C4)[8] * (xx * (xx - (REAL)3 * yy) - yy * ((REAL)3 * xx - yy));
    return 25;
}

/* sh.cuh:212-232 + 46-118: colors = sum_k basis_k * coeff[k] for k < bases(degrees_to_use),
 * summed per degree band in the reference's grouping. */
void SFX(orc_sh_forward)(int n, int degree, int degrees_to_use, const IN *viewdirs,
                         const IN *coeffs, REAL *colors) {
    int nb = orc_num_sh_bases(degree);
    REAL b[25];
    for (int i = 0; i < n; ++i) {
        int used = SFX(sh_basis)(degrees_to_use, viewdirs + 3 * i, b);
        const IN *c = coeffs + (size_t)i * nb * 3;
        for (int ch = 0; ch < 3; ++ch) {
            REAL acc = SFX(SH_C0) * (REAL)c[ch];
            if (used >= 4) {
                /* the reference factors C1 out of the degree-1 band (sh.cuh:77-79) */
                REAL dx = viewdirs[3 * i], dy = viewdirs[3 * i + 1], dz = viewdirs[3 * i + 2];
                REAL norm = SQRT(dx * dx + dy * dy + dz * dz);
                REAL x = dx / norm, y = dy / norm, z = dz / norm;
                acc += SFX(SH_C1) * (-y * (REAL)c[3 + ch] + z * (REAL)c[6 + ch] - x * (REAL)c[9 + ch]);
            }
            int band_lo[3] = {4, 9, 16}, band_hi[3] = {9, 16, 25};
            for (int band = 0; band < 3; ++band) {
                if (used < band_hi[band]) break;
                REAL s = 0;
                for (int k = band_lo[band]; k < band_hi[band]; ++k) {
                    REAL term = b[k] * (REAL)c[k * 3 + ch];
                    s = (k == band_lo[band]) ? term : s + term;
                }
                acc += s;
            }
            colors[3 * i + ch] = acc;
        }
    }
}

/* sh.cuh:234-253 + 120-210; entries above bases(degrees_to_use) stay zero (bindings.cu:62-63). */
void SFX(orc_sh_backward)(int n, int degree, int degrees_to_use, const IN *viewdirs,
                          const IN *v_colors, REAL *v_coeffs) {
    int nb = orc_num_sh_bases(degree);
    REAL b[25];
    for (int i = 0; i < n; ++i) {
        int used = SFX(sh_basis)(degrees_to_use, viewdirs + 3 * i, b);
        REAL *vc = v_coeffs + (size_t)i * nb * 3;
        for (int k = 0; k < nb * 3; ++k) vc[k] = 0;
        for (int k = 0; k < used; ++k)
            for (int ch = 0; ch < 3; ++ch) vc[k * 3 + ch] = b[k] * (REAL)v_colors[3 * i + ch];
    }
}

/* texture_helpers.cuh:162-237 (bilinear, replicate pad) on a jagged chart.
 * Out-of-range integer coordinates that the reference would read out of bounds
 * (tcoord <= -1) are clamped to 0; elsewhere the arithmetic is the reference's. */
static void SFX(chart_taps)(int h, int w, int si, int C, int ch, REAL u, REAL v, int bilinear,
                            REAL wt[4], long idx[4]) {
    REAL tu = (REAL)h * u, tv = (REAL)w * v;
    int iu = orc_sat_i32(tu), iv = orc_sat_i32(tv);
    REAL fu = tu - (REAL)iu, fv = tv - (REAL)iv;
    int ju = iu + 1 < h - 1 ? iu + 1 : h - 1;
    int jv = iv + 1 < w - 1 ? iv + 1 : w - 1;
    iu = iu < h - 1 ? iu : h - 1;
    iv = iv < w - 1 ? iv : w - 1;
    if (iu < 0) iu = 0;
    if (iv < 0) iv = 0;
    if (ju < 0) ju = 0;
    if (jv < 0) jv = 0;
    REAL wii = ((REAL)1 - fu) * ((REAL)1 - fv);
    REAL wij = ((REAL)1 - fu) * fv;
    REAL wji = fu * ((REAL)1 - fv);
    REAL wjj = fu * fv;
    if (bilinear) {
        wt[0] = wii; wt[1] = wij; wt[2] = wji; wt[3] = wjj;
    } else {
        wt[0] = wt[1] = wt[2] = wt[3] = 0;
        if (wii >= wij && wii >= wji && wii >= wjj) wt[0] = 1;
        else if (wij >= wii && wij >= wji && wij >= wjj) wt[1] = 1;
        else if (wji >= wii && wji >= wij && wji >= wjj) wt[2] = 1;
        else wt[3] = 1;
    }
    idx[0] = ((long)si + (long)iu * w + iv) * C + ch;
    idx[1] = ((long)si + (long)iu * w + jv) * C + ch;
    idx[2] = ((long)si + (long)ju * w + iv) * C + ch;
    idx[3] = ((long)si + (long)ju * w + jv) * C + ch;
}

/* texture_helpers.cuh:262-299: d value / d (u, v) through the bilinear weights. */
static void SFX(chart_uv_grad)(int h, int w, REAL u, REAL v, const REAL cv[4], REAL v_val, REAL vuv[2]) {
    REAL tu = (REAL)h * u, tv = (REAL)w * v;
    int iu = orc_sat_i32(tu), iv = orc_sat_i32(tv);
    REAL fu = tu - (REAL)iu, fv = tv - (REAL)iv;
    REAL gu = v_val * (-((REAL)1 - fv) * cv[0] - fv * cv[1] + ((REAL)1 - fv) * cv[2] + fv * cv[3]);
    REAL gv = v_val * (-((REAL)1 - fu) * cv[0] + ((REAL)1 - fu) * cv[1] - fu * cv[2] + fu * cv[3]);
    vuv[0] += (REAL)h * gu;
    vuv[1] += (REAL)w * gv;
}

/* texture_sample.cu:11-36: one bilinear query per row, no uv clamp. */
void SFX(orc_texture_sample)(int q, int C, const int *dims, const IN *uvs, const IN *texture,
                             REAL *out) {
    for (int i = 0; i < q; ++i) {
        int h = dims[3 * i], w = dims[3 * i + 1], si = dims[3 * i + 2];
        for (int ch = 0; ch < C; ++ch) {
            REAL wt[4];
            long idx[4];
            SFX(chart_taps)(h, w, si, C, ch, uvs[2 * i], uvs[2 * i + 1], 1, wt, idx);
            out[(size_t)i * C + ch] = wt[0] * (REAL)texture[idx[0]] + wt[1] * (REAL)texture[idx[1]] +
                                      wt[2] * (REAL)texture[idx[2]] + wt[3] * (REAL)texture[idx[3]];
        }
    }
}

/* Correct adjoint of orc_texture_sample w.r.t. the texture (the reference's own
 * backward at texture_sample.cu:38-70 reads v_texture instead of v_output). */
void SFX(orc_texture_sample_vjp)(int q, int C, const int *dims, const IN *uvs,
                                 const IN *v_out, REAL *v_texture) {
    for (int i = 0; i < q; ++i) {
        int h = dims[3 * i], w = dims[3 * i + 1], si = dims[3 * i + 2];
        for (int ch = 0; ch < C; ++ch) {
            REAL wt[4];
            long idx[4];
            SFX(chart_taps)(h, w, si, C, ch, uvs[2 * i], uvs[2 * i + 1], 1, wt, idx);
            for (int k = 0; k < 4; ++k) v_texture[idx[k]] += wt[k] * (REAL)v_out[(size_t)i * C + ch];
        }
    }
}

/* texture_helpers.cuh:336-352. */
static void SFX(get_ray)(const IN *c2w, REAL fx, REAL fy, REAL cx, REAL cy, REAL px, REAL py,
                         REAL o[3], REAL r[3]) {
    o[0] = c2w[3]; o[1] = c2w[7]; o[2] = c2w[11];
    REAL uvw[3] = {(px - cx) / fx, (py - cy) / fy, (REAL)1};
    SFX(xform43)(c2w, uvw, r);
    r[0] = r[0] - o[0]; r[1] = r[1] - o[1]; r[2] = r[2] - o[2];
    REAL nrm = SQRT(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    r[0] = r[0] / nrm; r[1] = r[1] / nrm; r[2] = r[2] / nrm;
}

/* texture_helpers.cuh:302-313: returns the clamped denominator too. */
static REAL SFX(compute_tt)(const REAL o[3], const REAL r[3], const REAL m[3], const REAL n[3], REAL *den_out) {
    REAL diff[3] = {m[0] - o[0], m[1] - o[1], m[2] - o[2]};
    REAL den = SFX(dot3)(n, r);
    const REAL eps = (REAL)1e-6f;
    if ((REAL)0 <= den && den < eps) den = eps;
    else if (-eps < den && den <= (REAL)0) den = -eps;
    if (den_out) *den_out = den;
    return SFX(dot3)(n, diff) / den;
}

/* Geometry of one (pixel, Gaussian) pair, texture.cu:158-221 / 517-559. */
typedef struct {
    REAL ax1[3], ax2[3], ax3[3], mean[3], delta[3];
    REAL t, l1, l2, alpha, exp_sigma, exp_sigma_blur, sigma_1, sigma_2, sigma_factor;
    REAL pview[3], xy[2];
    REAL non_blur, blur;
    REAL sx, sy, opac;
} SFX(pair_t);

/* mode 0 = forward formula (texture.cu:179-200), 1 = backward formula (texture.cu:535-555). */
static void SFX(pair_geom)(const orc_scene *S, int g, const REAL o[3], const REAL r[3], REAL px, REAL py,
                           int use_blur, int mode, SFX(pair_t) *P) {
    SFX(quat_axes)(SF(quats) + 4 * g, P->ax1, P->ax2, P->ax3);
    for (int k = 0; k < 3; ++k) P->mean[k] = SF(means)[3 * g + k];
    P->t = SFX(compute_tt)(o, r, P->mean, P->ax3, NULL);
    for (int k = 0; k < 3; ++k) P->delta[k] = (o[k] + P->t * r[k]) - P->mean[k];
    P->l1 = SFX(dot3)(P->delta, P->ax1);
    P->l2 = SFX(dot3)(P->delta, P->ax2);
    P->sx = SF(scales)[3 * g];
    P->sy = SF(scales)[3 * g + 1];
    P->opac = SF(opacities)[g];
    REAL gs = (REAL)S->glob_scale;
    REAL sigma;
    if (mode == 0) {
        REAL is1 = (REAL)1 / (P->sx * gs), is2 = (REAL)1 / (P->sy * gs);
        sigma = (REAL)0.5f * (is1 * is1 * P->l1 * P->l1 + is2 * is2 * P->l2 * P->l2);
    } else {
        P->sigma_factor = (REAL)0.5f / (gs * gs);
        P->sigma_1 = P->l1 * P->l1 / (P->sx * P->sx);
        P->sigma_2 = P->l2 * P->l2 / (P->sy * P->sy);
        sigma = P->sigma_factor * (P->sigma_1 + P->sigma_2);
    }
    SFX(xform43)(SF(viewmat), P->mean, P->pview);
    SFX(project_pix)((REAL)S->fx, (REAL)S->fy, P->pview, (REAL)S->cx, (REAL)S->cy, P->xy);
    REAL dx = P->xy[0] - px, dy = P->xy[1] - py;
    REAL sigma_blur = (REAL)0.5f * (REAL)2.0f * (dx * dx + dy * dy);
    P->non_blur = 1;
    P->blur = 0;
    if (use_blur && sigma_blur < sigma) {
        P->blur = 1;
        P->non_blur = 0;
    }
    P->exp_sigma = EXP(-sigma);
    P->exp_sigma_blur = EXP(-sigma_blur);
    REAL a = P->opac * (P->non_blur * P->exp_sigma + P->blur * P->exp_sigma_blur);
    P->alpha = a < (REAL)0.99f ? a : (REAL)0.99f;
}

static int SFX(pair_skip)(const SFX(pair_t) *P) {
    return P->t < (REAL)0.01f || P->t > (REAL)1000.0f || P->alpha < (REAL)1.f / (REAL)255.f;
}

/* local_outline (texture_helpers.cuh:390-417) with get_pixel_to_texel
 * (:355-366): distance in pixels to the nearest pixel of the 7 x 7
 * neighbourhood whose plane hit lies outside sigma <= thresh (1000 if none). */
static REAL SFX(local_outline)(const orc_scene *S, const SFX(pair_t) *P, REAL px, REAL py,
                               REAL thresh) {
    REAL gs = (REAL)S->glob_scale, s1 = P->sx * gs, s2 = P->sy * gs, best = (REAL)1000.0f;
    for (int dx = -3; dx <= 3; ++dx)
        for (int dy = -3; dy <= 3; ++dy) {
            REAL o[3], r[3], d[3];
            SFX(get_ray)(SF(c2w), (REAL)S->fx, (REAL)S->fy, (REAL)S->cx, (REAL)S->cy,
                         px + (REAL)dx, py + (REAL)dy, o, r);
            REAL t = SFX(compute_tt)(o, r, P->mean, P->ax3, NULL);
            for (int k = 0; k < 3; ++k) d[k] = (o[k] + t * r[k]) - P->mean[k];
            REAL as1 = SFX(dot3)(d, P->ax1) / s1, as2 = SFX(dot3)(d, P->ax2) / s2;
            if ((REAL)0.5f * (as1 * as1 + as2 * as2) > thresh) {
                REAL dis = SQRT((REAL)(dx * dx + dy * dy));
                if (best > dis) best = dis;
            }
        }
    return best;
}

/* texture.cu:11-329: front-to-back compositing of one image. */
void SFX(orc_texture_forward)(const orc_scene *S, REAL *out_img, REAL *out_depth, REAL *out_reg,
                              REAL *out_texture, REAL *out_normal, REAL *final_Ts, int *final_idx,
                              int *depth_idx, REAL *out_reg_s) {
    const int W = S->img_w, H = S->img_h, bw = S->block_width, C = S->channels;
    const int tiles_x = (W + bw - 1) / bw, tiles_y = (H + bw - 1) / bw;
    const int use_blur = (S->settings >> 9) & 1, use_ndc = (S->settings >> 10) & 1;
    const int viz = (S->settings >> 15) & 1, bilinear = !((S->settings >> 2) & 1);
    /* alpha visualisation mode, texture.cu:56-65 */
    const int aviz = (S->settings >> 16) & 1, white = (S->settings >> 24) & 1;
    const int opac_thresh = (S->settings >> 25) & 1;
    const REAL abound = (REAL)(float)((S->settings >> 17) & 31) / (REAL)8.0f;
    const REAL obound = (REAL)(float)((S->settings >> 26) & 15) / (REAL)4.0f;
    const REAL sthresh = (REAL)0.5f * abound * abound;
    REAL *tex = (REAL *)malloc(sizeof(REAL) * (C > 0 ? C : 1));
    for (int ty = 0; ty < tiles_y; ++ty)
        for (int tx = 0; tx < tiles_x; ++tx) {
            int tile = ty * tiles_x + tx;
            if (S->tile_mask && !S->tile_mask[tile]) continue;
            int rx = S->tile_bins[2 * tile], ry = S->tile_bins[2 * tile + 1];
            for (int ly = 0; ly < bw; ++ly)
                for (int lx = 0; lx < bw; ++lx) {
                    int i = ty * bw + ly, j = tx * bw + lx;
                    if (i >= H || j >= W) continue;
                    REAL px = (REAL)((float)j + 0.5), py = (REAL)((float)i + 0.5);
                    REAL o[3], r[3];
                    SFX(get_ray)(SF(c2w), (REAL)S->fx, (REAL)S->fy, (REAL)S->cx, (REAL)S->cy, px, py, o, r);
                    REAL view_depth = (REAL)SF(viewmat)[8] * r[0] + (REAL)SF(viewmat)[9] * r[1] +
                                      (REAL)SF(viewmat)[10] * r[2];
                    int done = 0, cur_idx = 0, pdi = -1;
                    REAL T = 1, rgb[3] = {0, 0, 0}, nrm[3] = {0, 0, 0}, pdepth = 0, preg = 0;
                    REAL ts[3] = {0, 0, 0};
                    for (int c = 0; c < C; ++c) tex[c] = 0;
                    for (int e = rx; e < ry && !done; ++e) {
                        int g = S->ids_sorted[e];
                        SFX(pair_t) P;
                        SFX(pair_geom)(S, g, o, r, px, py, use_blur, 0, &P);
                        REAL pixel_dis = (REAL)1000.0f;
                        if (aviz) { /* texture.cu:201-212 */
                            REAL gs = (REAL)S->glob_scale;
                            REAL is1 = (REAL)1 / (P.sx * gs), is2 = (REAL)1 / (P.sy * gs);
                            REAL sigma = (REAL)0.5f * (is1 * is1 * P.l1 * P.l1 + is2 * is2 * P.l2 * P.l2);
                            P.alpha = (REAL)0.99f;
                            if (sigma > sthresh) P.alpha = 0;
                            if (opac_thresh && P.opac < (REAL)0.5f) P.alpha = 0;
                            pixel_dis = SFX(local_outline)(S, &P, px, py, sthresh);
                        }
                        int skip = SFX(pair_skip)(&P);
                        REAL next_T = T * ((REAL)1 - P.alpha);
                        if (next_T <= (REAL)1e-4f) done = 1;
                        if (skip || done) continue;
                        REAL vis = P.alpha * T;
                        /* texture.cu:226-235,268-273: colour and texture as given, or
                         * within the outline bound white (bit 24) or nothing */
                        int plain = !aviz || pixel_dis > obound, outline_white = !plain && white;
                        for (int k = 0; k < 3; ++k) {
                            if (plain) rgb[k] = rgb[k] + (REAL)SF(colors)[3 * g + k] * vis;
                            else if (outline_white) rgb[k] = rgb[k] + vis;
                        }
                        REAL hn[3] = {P.ax3[0], P.ax3[1], P.ax3[2]};
                        if (viz && SFX(dot3)(r, P.ax3) > 0) {
                            hn[0] = -hn[0]; hn[1] = -hn[1]; hn[2] = -hn[2];
                        }
                        for (int k = 0; k < 3; ++k) nrm[k] += vis * hn[k];
                        REAL t_view = P.t * view_depth;
                        /* get_uv + clamp_coord, texture_helpers.cuh:19-60 */
                        REAL u = (REAL)SF(uv0)[2 * g] + ((REAL)SF(umap)[3 * g] * P.delta[0] + (REAL)SF(umap)[3 * g + 1] * P.delta[1] + (REAL)SF(umap)[3 * g + 2] * P.delta[2]);
                        REAL v = (REAL)SF(uv0)[2 * g + 1] + ((REAL)SF(vmap)[3 * g] * P.delta[0] + (REAL)SF(vmap)[3 * g + 1] * P.delta[1] + (REAL)SF(vmap)[3 * g + 2] * P.delta[2]);
                        if (u <= 0) u = 0;
                        if (u >= 1) u = 1;
                        if (v <= 0) v = 0;
                        if (v >= 1) v = 1;
                        const int *d = S->texture_dims + 3 * g;
                        for (int c = 0; c < C; ++c) {
                            REAL wt[4];
                            long idx[4];
                            SFX(chart_taps)(d[0], d[1], d[2], C, c, u, v, bilinear, wt, idx);
                            REAL val = wt[0] * (REAL)SF(texture)[idx[0]] + wt[1] * (REAL)SF(texture)[idx[1]] +
                                       wt[2] * (REAL)SF(texture)[idx[2]] + wt[3] * (REAL)SF(texture)[idx[3]];
                            if (plain) tex[c] += vis * val;
                            else if (outline_white) tex[c] += vis;
                        }
                        if (T > (REAL)0.5f) {
                            pdepth = t_view;
                            pdi = e;
                        }
                        REAL dval = P.t;
                        if (use_ndc) {
                            const REAL tn = (REAL)0.01f, tf = (REAL)1000.0f;
                            dval = (tf * t_view - tf * tn) / ((tf - tn) * t_view);
                        }
                        /* distortion, helpers.cuh:259-264 */
                        preg += vis * (dval * dval * ts[0] + ts[2] - (REAL)2 * dval * ts[1]);
                        ts[0] += vis;
                        ts[1] += vis * dval;
                        ts[2] += vis * dval * dval;
                        T = next_T;
                        cur_idx = e;
                    }
                    int pid = i * W + j;
                    final_Ts[pid] = T;
                    final_idx[pid] = cur_idx;
                    depth_idx[pid] = pdi;
                    for (int k = 0; k < 3; ++k) {
                        out_img[3 * pid + k] = rgb[k] + T * (REAL)SF(background)[k];
                        out_normal[3 * pid + k] = nrm[k];
                        out_reg_s[3 * pid + k] = ts[k];
                    }
                    out_depth[pid] = pdepth;
                    out_reg[pid] = preg;
                    for (int c = 0; c < C; ++c) out_texture[(size_t)pid * C + c] = tex[c];
                }
        }
    free(tex);
}

/* texture.cu:331-760: back-to-front adjoint, with T recovered by division.
 * Gradients are accumulated into zero-initialised arrays in the reference
 * layouts (v_scales z stays zero). */
void SFX(orc_texture_backward)(const orc_scene *S, const REAL *final_Ts, const int *final_idx,
                               const int *depth_idx, const REAL *final_s, const IN *v_img,
                               const IN *v_depth, const IN *v_reg, const IN *v_alpha_out,
                               const IN *v_tex_out, const IN *v_normal, REAL *v_colors,
                               REAL *v_opacity, REAL *v_means, REAL *v_scales, REAL *v_quats,
                               REAL *v_uv0, REAL *v_umap, REAL *v_vmap, REAL *v_texture) {
    const int W = S->img_w, H = S->img_h, bw = S->block_width, C = S->channels;
    const int tiles_x = (W + bw - 1) / bw, tiles_y = (H + bw - 1) / bw;
    const int use_blur = (S->settings >> 9) & 1, use_ndc = (S->settings >> 10) & 1;
    const int bilinear = !((S->settings >> 2) & 1), prop_uv = (S->settings >> 8) & 1;
    const REAL fx = S->fx, fy = S->fy;
    for (int ty = 0; ty < tiles_y; ++ty)
        for (int tx = 0; tx < tiles_x; ++tx) {
            int tile = ty * tiles_x + tx;
            if (S->tile_mask && !S->tile_mask[tile]) continue;
            int rx = S->tile_bins[2 * tile], ry = S->tile_bins[2 * tile + 1];
            for (int ly = 0; ly < bw; ++ly)
                for (int lx = 0; lx < bw; ++lx) {
                    int i = ty * bw + ly, j = tx * bw + lx;
                    if (i >= H || j >= W) continue;
                    int pid = i * W + j;
                    REAL px = (REAL)((float)j + 0.5), py = (REAL)((float)i + 0.5);
                    REAL o[3], r[3];
                    SFX(get_ray)(SF(c2w), (REAL)S->fx, (REAL)S->fy, (REAL)S->cx, (REAL)S->cy, px, py, o, r);
                    REAL view_depth = (REAL)SF(viewmat)[8] * r[0] + (REAL)SF(viewmat)[9] * r[1] +
                                      (REAL)SF(viewmat)[10] * r[2];
                    REAL T = final_Ts[pid];
                    const REAL *tsf = final_s + 3 * pid;
                    int dfin = depth_idx[pid], bfin = final_idx[pid];
                    REAL vo[3] = {v_img[3 * pid], v_img[3 * pid + 1], v_img[3 * pid + 2]};
                    REAL vn[3] = {v_normal[3 * pid], v_normal[3 * pid + 1], v_normal[3 * pid + 2]};
                    REAL vdep = v_depth[pid], vreg = v_reg[pid];
                    REAL vT = (REAL)SF(background)[0] * vo[0] + (REAL)SF(background)[1] * vo[1] +
                              (REAL)SF(background)[2] * vo[2] - (REAL)v_alpha_out[pid];
                    for (int e = ry - 1; e >= rx; --e) {
                        if (e > bfin) continue;
                        int g = S->ids_sorted[e];
                        SFX(pair_t) P;
                        SFX(pair_geom)(S, g, o, r, px, py, use_blur, 1, &P);
                        if (SFX(pair_skip)(&P)) continue;
                        T *= (REAL)1 / ((REAL)1 - P.alpha);
                        REAL vis = P.alpha * T;
                        const IN *rgb = SF(colors) + 3 * g;
                        for (int k = 0; k < 3; ++k) v_colors[3 * g + k] += vis * vo[k];
                        REAL vvis = (REAL)rgb[0] * vo[0] + (REAL)rgb[1] * vo[1] + (REAL)rgb[2] * vo[2];
                        vvis += SFX(dot3)(P.ax3, vn);
                        REAL vdiff[3] = {0, 0, 0};
                        {
                            const IN *um = SF(umap) + 3 * g, *vm = SF(vmap) + 3 * g, *u0 = SF(uv0) + 2 * g;
                            REAL u = (REAL)u0[0] + ((REAL)um[0] * P.delta[0] + (REAL)um[1] * P.delta[1] + (REAL)um[2] * P.delta[2]);
                            REAL v = (REAL)u0[1] + ((REAL)vm[0] * P.delta[0] + (REAL)vm[1] * P.delta[1] + (REAL)vm[2] * P.delta[2]);
                            if (u <= 0) u = 0;
                            if (u >= 1) u = 1;
                            if (v <= 0) v = 0;
                            if (v >= 1) v = 1;
                            const int *d = S->texture_dims + 3 * g;
                            REAL vuv[2] = {0, 0};
                            for (int c = 0; c < C; ++c) {
                                REAL wt[4], cv[4];
                                long idx[4];
                                SFX(chart_taps)(d[0], d[1], d[2], C, c, u, v, bilinear, wt, idx);
                                for (int k = 0; k < 4; ++k) cv[k] = SF(texture)[idx[k]];
                                REAL val = wt[0] * cv[0] + wt[1] * cv[1] + wt[2] * cv[2] + wt[3] * cv[3];
                                REAL vt = (REAL)v_tex_out[(size_t)pid * C + c];
                                REAL vval = vis * vt;
                                for (int k = 0; k < 4; ++k) v_texture[idx[k]] += wt[k] * vval;
                                REAL vuvd[2] = {0, 0};
                                if (bilinear) SFX(chart_uv_grad)(d[0], d[1], u, v, cv, vval, vuvd);
                                if (prop_uv) {
                                    vuv[0] += vuvd[0];
                                    vuv[1] += vuvd[1];
                                }
                                vvis += val * vt;
                            }
                            /* get_uv_vjp, texture_helpers.cuh:62-80 */
                            v_uv0[2 * g] += vuv[0];
                            v_uv0[2 * g + 1] += vuv[1];
                            for (int k = 0; k < 3; ++k) {
                                v_umap[3 * g + k] += P.delta[k] * vuv[0];
                                v_vmap[3 * g + k] += P.delta[k] * vuv[1];
                            }
                            if (prop_uv)
                                for (int k = 0; k < 3; ++k)
                                    vdiff[k] += (REAL)um[k] * vuv[0] + (REAL)vm[k] * vuv[1];
                        }
                        REAL valpha = T * vvis - T * vT;
                        REAL vTcur = P.alpha * vvis + ((REAL)1 - P.alpha) * vT;
                        REAL vt_local = 0, vtndc = 0, vtview = 0;
                        REAL t_view = P.t * view_depth;
                        const REAL tn = (REAL)0.01f, tf = (REAL)1000.0f;
                        REAL t_ndc = (tf * t_view - tf * tn) / ((tf - tn) * t_view);
                        {
                            /* distortion_vjp, helpers.cuh:266-269, with the final sums */
                            REAL val = use_ndc ? t_ndc : P.t;
                            REAL vval = (REAL)2 * (vis * val * tsf[0] - vis * tsf[1]) * vreg;
                            REAL vw = (val * val * tsf[0] - (REAL)2 * val * tsf[1] + tsf[2]) * vreg;
                            if (use_ndc) vtndc += vval;
                            else vt_local += vval;
                            valpha += vw * T;
                            vTcur += vw * P.alpha;
                        }
                        vT = vTcur;
                        REAL vsigma = -P.non_blur * P.opac * P.exp_sigma * valpha;
                        REAL vsc[2] = {-(REAL)2 * P.sigma_factor * P.sigma_1 * vsigma / P.sx,
                                       -(REAL)2 * P.sigma_factor * P.sigma_2 * vsigma / P.sy};
                        REAL vl1 = (REAL)2 * P.sigma_factor * P.l1 * vsigma / (P.sx * P.sx);
                        REAL vl2 = (REAL)2 * P.sigma_factor * P.l2 * vsigma / (P.sy * P.sy);
                        vt_local += vl1 * SFX(dot3)(r, P.ax1) + vl2 * SFX(dot3)(r, P.ax2);
                        vt_local += SFX(dot3)(r, vdiff);
                        if (e == dfin && dfin != -1) vtview += vdep;
                        vtview += (tf * tn) / ((tf - tn) * t_view * t_view) * vtndc;
                        vt_local += view_depth * vtview;
                        REAL vsigma_blur = -P.blur * P.opac * P.exp_sigma_blur * valpha;
                        REAL vxy[2] = {(REAL)2.0f * vsigma_blur * (P.xy[0] - px), (REAL)2.0f * vsigma_blur * (P.xy[1] - py)};
                        REAL vpv[3], vmx[3];
                        SFX(project_pix_vjp)(fx, fy, P.pview, vxy, vpv);
                        SFX(xform43_rot_t)(SF(viewmat), vpv, vmx);
                        /* compute_tt_vjp, texture_helpers.cuh:315-334 */
                        REAL den;
                        SFX(compute_tt)(o, r, P.mean, P.ax3, &den);
                        REAL diff[3] = {P.mean[0] - o[0], P.mean[1] - o[1], P.mean[2] - o[2]};
                        REAL vmtt[3], vntt[3];
                        for (int k = 0; k < 3; ++k) {
                            vmtt[k] = P.ax3[k] * vt_local / den;
                            vntt[k] = (diff[k] - r[k] * P.t) * vt_local / den;
                        }
                        for (int k = 0; k < 3; ++k)
                            v_means[3 * g + k] += -(P.ax1[k] * vl1 + P.ax2[k] * vl2) + vmtt[k] + vmx[k] - vdiff[k];
                        REAL vR[3][3];
                        for (int k = 0; k < 3; ++k) {
                            vR[0][k] = P.delta[k] * vl1;
                            vR[1][k] = P.delta[k] * vl2;
                            vR[2][k] = vntt[k] + vis * vn[k];
                        }
                        REAL vq[4];
                        SFX(quat_axes_vjp)(SF(quats) + 4 * g, vR, vq);
                        for (int k = 0; k < 4; ++k) v_quats[4 * g + k] += vq[k];
                        v_scales[3 * g] += vsc[0];
                        v_scales[3 * g + 1] += vsc[1];
                        v_opacity[g] += (P.non_blur * P.exp_sigma + P.blur * P.exp_sigma_blur) * valpha;
                    }
                }
        }
}

/* texture_edit.cu:11-227: for every pair the forward would composite (same
 * alpha, skip and stop rules, blur from settings bit 0) whose view depth
 * t * view_depth lies in [depth_lower, depth_upper] of the pixel, add the
 * bilinear weights (texture_helpers.cuh:162-250, always bilinear, replicate
 * padding) times (r*a, g*a, b*a, a, 1) -- a the pixel's edit alpha -- into
 * channels 0..4 of out (X x C texels, C >= 5, zero-initialised by the caller).
 * The pixel's T advances by every composited pair, in or out of range. */
void SFX(orc_texture_edit)(const orc_scene *S, const IN *upd_img, const IN *upd_alpha,
                           const IN *depth_lower, const IN *depth_upper, REAL *out) {
    const int W = S->img_w, H = S->img_h, bw = S->block_width, C = S->channels;
    const int tiles_x = (W + bw - 1) / bw, tiles_y = (H + bw - 1) / bw;
    const int use_blur = S->settings & 1;
    for (int ty = 0; ty < tiles_y; ++ty)
        for (int tx = 0; tx < tiles_x; ++tx) {
            int tile = ty * tiles_x + tx;
            if (S->tile_mask && !S->tile_mask[tile]) continue;
            int rx = S->tile_bins[2 * tile], ry = S->tile_bins[2 * tile + 1];
            for (int ly = 0; ly < bw; ++ly)
                for (int lx = 0; lx < bw; ++lx) {
                    int i = ty * bw + ly, j = tx * bw + lx;
                    if (i >= H || j >= W) continue;
                    int pid = i * W + j;
                    REAL px = (REAL)((float)j + 0.5), py = (REAL)((float)i + 0.5);
                    REAL o[3], r[3];
                    SFX(get_ray)(SF(c2w), (REAL)S->fx, (REAL)S->fy, (REAL)S->cx, (REAL)S->cy, px, py, o, r);
                    REAL view_depth = (REAL)SF(viewmat)[8] * r[0] + (REAL)SF(viewmat)[9] * r[1] +
                                      (REAL)SF(viewmat)[10] * r[2];
                    REAL a = (REAL)upd_alpha[pid];
                    REAL val[5] = {(REAL)upd_img[3 * pid] * a, (REAL)upd_img[3 * pid + 1] * a,
                                   (REAL)upd_img[3 * pid + 2] * a, a, (REAL)1};
                    REAL lo = (REAL)depth_lower[pid], hi = (REAL)depth_upper[pid];
                    int done = 0;
                    REAL T = 1;
                    for (int e = rx; e < ry && !done; ++e) {
                        int g = S->ids_sorted[e];
                        SFX(pair_t) P;
                        SFX(pair_geom)(S, g, o, r, px, py, use_blur, 0, &P);
                        int skip = SFX(pair_skip)(&P);
                        REAL next_T = T * ((REAL)1 - P.alpha);
                        if (next_T <= (REAL)1e-4f) done = 1;
                        if (skip || done) continue;
                        REAL t_view = P.t * view_depth;
                        if (t_view >= lo && t_view <= hi) {
                            REAL u = (REAL)SF(uv0)[2 * g] + ((REAL)SF(umap)[3 * g] * P.delta[0] + (REAL)SF(umap)[3 * g + 1] * P.delta[1] + (REAL)SF(umap)[3 * g + 2] * P.delta[2]);
                            REAL v = (REAL)SF(uv0)[2 * g + 1] + ((REAL)SF(vmap)[3 * g] * P.delta[0] + (REAL)SF(vmap)[3 * g + 1] * P.delta[1] + (REAL)SF(vmap)[3 * g + 2] * P.delta[2]);
                            if (u <= 0) u = 0;
                            if (u >= 1) u = 1;
                            if (v <= 0) v = 0;
                            if (v >= 1) v = 1;
                            const int *d = S->texture_dims + 3 * g;
                            for (int k = 0; k < 5; ++k) {
                                REAL wt[4];
                                long idx[4];
                                SFX(chart_taps)(d[0], d[1], d[2], C, k, u, v, 1, wt, idx);
                                for (int q = 0; q < 4; ++q) out[idx[q]] += wt[q] * val[k];
                            }
                        }
                        T = next_T;
                    }
                }
        }
}

#undef SF
