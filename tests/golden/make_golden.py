"""Writes tests/golden/kat.json: known-answer vectors derived from closed forms
and library routines, independently of the oracle's code.

The reference ships no golden vectors or fixtures for this path (its tests use
unseeded random data compared against its own torch code), so the vectors
here come from the mathematics the reference implements:

  sh        real spherical harmonics from scipy.special.sph_harm_y
            (Condon-Shortley phase; Y_lm = sqrt2 Im Y_l^|m| for m<0,
            sqrt2 Re Y_l^m for m>0), degree <= 4, at fixed directions
  aabb      a camera-facing splat with an identity pose: the +-3 sigma corners
            project to centre f*x/z + c, extent 3*g*s*f/z
  splat1    one camera-facing splat: alpha at a pixel is
            opac * exp(-0.5 * (dx^2/sx^2 + dy^2/sy^2)) with dx = (px-cx)*z/f,
            colour = alpha*c + (1-alpha)*bg, depth = z, normal = (0,0,alpha)
  splat2    two camera-facing splats on the same axis: front-to-back
            compositing alpha1*c1 + (1-alpha1)*alpha2*c2 + T*bg, and the L2
            distortion alpha1*(1-alpha1)*alpha2*(t1-t2)^2 with t the ray length
  bilinear  a texel grid that is affine in (i, j) is reproduced exactly by
            bilinear sampling inside the grid, and clamped (replicate-padded)
            outside it
  splat_tex one camera-facing splat (identity, and turned in its plane by a
            z-axis quaternion) over a 4 x 5 chart whose three channels are
            affine in the texel indices: at a pixel the ray meets the plane
            z = m_z at offset d = (m_z (px - cx)/f - m_x, m_z (py - cy)/f - m_y, 0)
            from the mean, l1 = d.ax1, l2 = d.ax2, u = u0 + umap.d and
            v = v0 + vmap.d clamped to [0, 1] (umap and vmap mix x and y), and
            the bilinear lookup is A + B min(h u, h - 1) + C min(w v, w - 1);
            out_texture = alpha * that, out_img = alpha * colour + (1 - alpha) bg
Run: python tests/golden/make_golden.py
"""
import json
import math
import os

import numpy as np
from scipy.special import sph_harm_y

HERE = os.path.dirname(os.path.abspath(__file__))


def r32(x):
    """Round inputs to float32 first: the renderer's inputs are float32 tensors."""
    if isinstance(x, (list, tuple)):
        return [r32(v) for v in x]
    return float(np.float32(x))


def real_sh(l, m, d):
    theta = math.acos(max(-1.0, min(1.0, d[2])))
    phi = math.atan2(d[1], d[0])
    y = complex(sph_harm_y(l, abs(m), theta, phi))
    if m < 0:
        return math.sqrt(2) * y.imag
    if m == 0:
        return y.real
    return math.sqrt(2) * y.real


def sh_vectors():
    dirs = [[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, -1.0, 0.0],
            [0.3, -0.5, 0.8124038404635961], [-0.6, 0.64, -0.48], [0.1, 0.2, -0.9746794344808963]]
    out = []
    for d in dirs:
        n = math.sqrt(sum(x * x for x in d))
        u = [x / n for x in d]
        out.append({"dir": d, "basis": [real_sh(l, m, u) for l in range(5) for m in range(-l, l + 1)]})
    return out


def aabb_vectors():
    cases = []
    f, cx, cy = 200.0, 64.0, 48.0
    for case in [(0.1, -0.2, 3.0, 0.05, 0.02, 1.0), (-0.3, 0.25, 5.5, 0.1, 0.1, 0.5),
                 (0.0, 0.0, 2.0, 0.01, 0.07, 2.0)]:
        x, y, z, sx, sy, g = r32(list(case))
        cases.append({"mean": [x, y, z], "scale": [sx, sy, 0.0], "glob_scale": g,
                      "fx": f, "fy": f, "cx": cx, "cy": cy,
                      "center": [f * x / z + cx, f * y / z + cy],
                      "extent": [3 * g * sx * f / z, 3 * g * sy * f / z]})
    return cases


def splat1_vectors():
    """One camera-facing splat (identity quaternion) rendered on a small image."""
    W, H, f = 32, 24, 40.0
    cx, cy = 16.0, 12.0
    mean = r32([0.05, -0.03, 2.5])
    sx, sy, opac = r32([0.12, 0.08, 0.7])
    color, bg, tex = r32([0.9, 0.2, 0.4]), r32([0.1, 0.3, 0.5]), r32([0.25, 0.5, 0.75])
    img = np.zeros((H, W, 3))
    alpha_map = np.zeros((H, W))
    for i in range(H):
        for j in range(W):
            px, py = j + 0.5, i + 0.5
            dx = (px - cx) / f * mean[2] - mean[0]
            dy = (py - cy) / f * mean[2] - mean[1]
            a = min(0.99, opac * math.exp(-0.5 * (dx * dx / sx ** 2 + dy * dy / sy ** 2)))
            if a < 1.0 / 255.0:
                a = 0.0
            alpha_map[i, j] = a
            for k in range(3):
                img[i, j, k] = a * color[k] + (1 - a) * bg[k]
    return {"W": W, "H": H, "f": f, "cx": cx, "cy": cy, "mean": mean, "scale": [sx, sy, 0.0],
            "opacity": opac, "color": color, "background": bg, "texel": tex,
            "img": img.tolist(), "alpha": alpha_map.tolist()}


def splat2_vectors():
    """Two camera-facing splats centred on the optical axis; pixel at the centre."""
    f, cx, cy, W, H = 50.0, 8.0, 8.0, 16, 16
    z1, z2 = 2.0, 3.5
    o1, o2 = 0.6, 0.5
    c1, c2, bg = [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]
    s = r32(0.3)
    o1 = r32(o1)
    res = []
    for (i, j) in [(7, 7), (8, 8), (7, 9)]:
        ux, uy = (j + 0.5 - cx) / f, (i + 0.5 - cy) / f
        alphas, ts = [], []
        for z, o in ((z1, o1), (z2, o2)):
            dx, dy = ux * z, uy * z
            alphas.append(min(0.99, o * math.exp(-0.5 * (dx * dx + dy * dy) / s ** 2)))
            ts.append(z * math.sqrt(1 + ux * ux + uy * uy))
        a1, a2 = alphas
        T = (1 - a1) * (1 - a2)
        col = [a1 * c1[k] + (1 - a1) * a2 * c2[k] + T * bg[k] for k in range(3)]
        reg = a1 * (1 - a1) * a2 * (ts[0] - ts[1]) ** 2
        res.append({"pixel": [i, j], "img": col, "T": T, "reg": reg, "depth": z1})
    return {"W": W, "H": H, "f": f, "cx": cx, "cy": cy, "z": [z1, z2], "opacity": [o1, o2],
            "color": [c1, c2], "background": bg, "scale": s, "pixels": res}


def bilinear_vectors():
    rng = np.random.default_rng(7)
    charts = [(3, 5), (4, 4), (1, 6), (7, 2)]
    a = [0.3, -0.2, 1.1]
    b = [0.05, 0.4, -0.7]
    c = [0.5, 0.25, 2.0]
    rows, dims, start = [], [], 0
    for (h, w) in charts:
        for i in range(h):
            for j in range(w):
                rows.append([a[k] * i + b[k] * j + c[k] for k in range(3)])
        dims.append([h, w, start])
        start += h * w
    queries = []
    for _ in range(40):
        ci = int(rng.integers(len(charts)))
        h, w = charts[ci]
        u, v = float(rng.uniform(0, 1)), float(rng.uniform(0, 1))
        ti = min(h * u, h - 1.0)
        tj = min(w * v, w - 1.0)
        queries.append({"dims": dims[ci], "uv": [u, v],
                        "value": [a[k] * ti + b[k] * tj + c[k] for k in range(3)]})
    return {"texture": rows, "queries": queries}


def splat_tex_vectors():
    """One camera-facing splat over an affine 4 x 5 x 3 chart (h = 4 indexes u,
    w = 5 indexes v; texel (i, j) is row i * w + j)."""
    W, H, f, cx, cy = 48, 40, 80.0, 24.0, 20.0
    h, w = 4, 5
    # dyadic coefficients: every texel value is exact in float32
    A = [0.125, 0.875, 0.3125]
    B = [0.1875, -0.09375, 0.0]     # per row step (u)
    C = [0.046875, 0.03125, -0.0625]  # per column step (v)
    rows = [[A[k] + B[k] * i + C[k] * j for k in range(3)] for i in range(h) for j in range(w)]
    assert rows == r32(rows)
    cases = []
    for theta in (0.0, 0.6):
        mean = r32([0.04, -0.03, 2.5])
        sx, sy, opac = r32([0.15, 0.10, 0.8])
        q = r32([math.cos(theta / 2), 0.0, 0.0, math.sin(theta / 2)])
        qw, qz = q[0], q[3]
        ax1 = [1 - 2 * qz * qz, 2 * qw * qz, 0.0]       # columns of R(q), unnormalised q
        ax2 = [-2 * qw * qz, 1 - 2 * qz * qz, 0.0]
        uv0 = r32([0.45, 0.55])
        umap = r32([2.0, 0.8, 0.0])
        vmap = r32([-0.6, 2.5, 0.0])
        color, bg = r32([0.9, 0.2, 0.4]), r32([0.1, 0.3, 0.5])
        tex_out = np.zeros((H, W, 3))
        img = np.zeros((H, W, 3))
        alpha_map = np.zeros((H, W))
        check = np.ones((H, W), bool)
        for i in range(H):
            for j in range(W):
                px, py = j + 0.5, i + 0.5
                d = [mean[2] * (px - cx) / f - mean[0], mean[2] * (py - cy) / f - mean[1], 0.0]
                l1 = sum(a * b for a, b in zip(d, ax1))
                l2 = sum(a * b for a, b in zip(d, ax2))
                a = min(0.99, opac * math.exp(-0.5 * (l1 * l1 / sx ** 2 + l2 * l2 / sy ** 2)))
                # float rounding may flip the 1/255 test right at the threshold
                check[i, j] = abs(a - 1.0 / 255.0) > 1e-5
                if a < 1.0 / 255.0:
                    a = 0.0
                u = min(max(uv0[0] + sum(x * y for x, y in zip(umap, d)), 0.0), 1.0)
                v = min(max(uv0[1] + sum(x * y for x, y in zip(vmap, d)), 0.0), 1.0)
                ti, tj = min(h * u, h - 1.0), min(w * v, w - 1.0)
                alpha_map[i, j] = a
                for k in range(3):
                    tex_out[i, j, k] = a * (A[k] + B[k] * ti + C[k] * tj)
                    img[i, j, k] = a * color[k] + (1 - a) * bg[k]
        cases.append({"W": W, "H": H, "f": f, "cx": cx, "cy": cy, "mean": mean,
                      "scale": [sx, sy, 0.0], "quat": q, "opacity": opac, "color": color,
                      "background": bg, "uv0": uv0, "umap": umap, "vmap": vmap,
                      "dims": [h, w, 0], "texture": rows, "out_texture": tex_out.tolist(),
                      "img": img.tolist(), "alpha": alpha_map.tolist(),
                      "check": check.astype(int).tolist()})
    return cases


def main():
    data = {"sh": sh_vectors(), "aabb": aabb_vectors(), "splat1": splat1_vectors(),
            "splat2": splat2_vectors(), "bilinear": bilinear_vectors(),
            "splat_tex": splat_tex_vectors()}
    with open(os.path.join(HERE, "kat.json"), "w") as fh:
        json.dump(data, fh, indent=1)


if __name__ == "__main__":
    main()
