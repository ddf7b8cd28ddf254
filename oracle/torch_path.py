"""PyTorch CPU path of one textured-2DGS view, forward + backward by autograd.

BASELINE AND TEST INFRASTRUCTURE ONLY.  ``bench.py``'s cpu_baseline leg times
it on the host cores, and ``tests/test_torch_path.py`` checks its outputs and
autograd gradients against the C oracle (and so cross-checks the oracle's
hand-written backward against automatic differentiation).  The product never
imports it; the GPU path fails loudly without its HIP library.

It computes what the reference's CUDA path computes (not the shortcuts of the
reference's own torch loop, ``_torch_impl.py:196-381``: alpha clamp 0.99
rather than 0.999, the stop rule, the blur and NDC options, non-square
images), written here as dense tensor algebra over (tile, list entry, pixel)
triples so that the host's cores all work on it:

* ``project``: screen AABB of the four +-3 g s corners, clip and projection
  (``get_aabb_2d.cu:11-89``, ``helpers.cuh:124-152,240-257``), view depths
  (``get_aabb_2d.py:22-32``).
* ``bin_gaussians``: tile counts (``get_aabb_2d.py:70-92``), tile rectangles
  and keys ``tile << 32 | depth bits`` (``forward.cu:13-71``, zero slots kept),
  a stable sort (``utils.py:159-160``) and bin edges (``forward.cu:76-98``).
* ``render_chunk``: the per-pixel front-to-back compositing of
  ``texture.cu:129-328``.  Per pair: ray/plane hit t = n.(m-o)/n.r with the
  +-1e-6 clamp, sigma, alpha = min(0.99, opac exp(-sigma)) (or the blur
  alpha), the skip test, the stop test T (1 - alpha) <= 1e-4 (which ends the
  pixel before that entry and is tested on skipped entries too), colour,
  normal, bilinear texels with replicate padding (``texture_helpers.cuh:
  19-24,37-60,162-237``), median depth while T > 0.5 and the distortion sums
  (``helpers.cuh:259-264``).  The running transmittance is an exclusive
  cumulative product along the list; the first stop and everything after it
  are masked out, exactly as the sequential loop never reaches them.
* SH colours up to degree 3 (``sh.cuh:46-116`` basis, real spherical
  harmonics with the standard normalisation constants).

The loss is separable over pixels, so a frame is processed in chunks of tiles
(or of one tile's pixels), each running forward and backward before the next:
memory stays bounded and the gradients accumulate in the leaf tensors.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Dict, Optional

import torch

EPS_DEN = 1e-6
T_NEAR, T_FAR = 0.01, 1000.0
ALPHA_MIN = 1.0 / 255.0
ALPHA_MAX = 0.99
T_STOP = 1e-4

SET_NEAREST = 1 << 2
SET_UV_GRAD = 1 << 8
SET_BLUR = 1 << 9
SET_NDC = 1 << 10


# ------------------------------------------------------------------ geometry
def quat_axes(q: torch.Tensor):
    """Columns of the rotation matrix of (w, x, y, z) quaternions, not
    normalised (``helpers.cuh:166-185``)."""
    w, x, y, z = q.unbind(-1)
    a1 = torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)], -1)
    a2 = torch.stack([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)], -1)
    a3 = torch.stack([2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)], -1)
    return a1, a2, a3


def _xform(vm: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
    """Row-major 4x4 ``vm`` applied to points (..., 3), summed left to right."""
    cols = [p[..., 0], p[..., 1], p[..., 2]]
    rows = []
    for r in range(3):
        acc = vm[r, 0] * cols[0] + vm[r, 1] * cols[1] + vm[r, 2] * cols[2] + vm[r, 3]
        rows.append(acc)
    return torch.stack(rows, -1)


def _project(pv: torch.Tensor, fx, fy, cx, cy):
    rw = 1.0 / (pv[..., 2] + 1e-6)
    return pv[..., 0] * rw * fx + cx, pv[..., 1] * rw * fy + cy


@torch.no_grad()
def project(means, scales, quats, viewmat, fx, fy, cx, cy, glob_scale=1.0):
    """Screen AABB centres and half extents (N, 2), and view depths (N,)."""
    pv = _xform(viewmat, means)
    depths = pv[:, 2].clone()
    clipped = pv[:, 2] <= T_NEAR
    mx, my = _project(pv, fx, fy, cx, cy)
    a1, a2, _ = quat_axes(quats)
    ell = 3.0 * glob_scale
    t1 = ell * scales[:, :1] * a1
    t2 = ell * scales[:, 1:2] * a2
    us, vs = [], []
    for s1 in (1.0, -1.0):
        for s2 in (1.0, -1.0):
            q = (means + t1 if s1 > 0 else means - t1)
            q = (q + t2 if s2 > 0 else q - t2)
            qv = _xform(viewmat, q)
            qv = torch.cat([qv[:, :2], qv[:, 2:].clamp(min=T_NEAR)], 1)
            u, v = _project(qv, fx, fy, cx, cy)
            us.append(u)
            vs.append(v)
    us, vs = torch.stack(us, 1), torch.stack(vs, 1)
    umax, umin = us.max(1).values, us.min(1).values
    vmax, vmin = vs.max(1).values, vs.min(1).values
    centers = torch.stack([0.5 * (umax + umin), 0.5 * (vmax + vmin)], 1)
    extents = torch.stack([0.5 * (umax - umin), 0.5 * (vmax - vmin)], 1)
    centers = torch.where(clipped[:, None], torch.stack([mx, my], 1), centers)
    extents = torch.where(clipped[:, None], torch.zeros_like(extents), extents)
    return centers, extents, depths


@torch.no_grad()
def bin_gaussians(centers, extents, depths, img_w, img_h, bw):
    """(ids_sorted (I,) int32, tile_bins (T, 2) int32) with the reference's
    key order: tile, then depth bits, ties in Gaussian order."""
    tx, ty = (img_w + bw - 1) // bw, (img_h + bw - 1) // bw
    # counts, get_num_tiles_hit_2d
    tl = torch.floor((centers - extents) / bw).to(torch.int64)
    br = torch.floor((centers + extents) / bw + 1).to(torch.int64)
    lim = torch.tensor([tx, ty])
    cnt2 = torch.minimum(torch.maximum(br, torch.zeros_like(br)), lim) - \
        torch.minimum(torch.maximum(tl, torch.zeros_like(tl)), lim)
    counts = cnt2[:, 0] * cnt2[:, 1]
    total = int(counts.sum())
    # written rectangles, get_tile_bbox (float division by bw, truncation)
    tc, te = centers / float(bw), extents / float(bw)
    lo = torch.minimum(torch.maximum(torch.trunc(tc - te).to(torch.int64), torch.zeros(1, dtype=torch.int64)), lim)
    hi = torch.minimum(torch.maximum(torch.trunc(tc + te + 1).to(torch.int64), torch.zeros(1, dtype=torch.int64)), lim)
    write = ~((extents[:, 0] <= 1e-4) & (extents[:, 1] <= 1e-4))
    wx = (hi[:, 0] - lo[:, 0]).clamp(min=0)
    wy = (hi[:, 1] - lo[:, 1]).clamp(min=0)
    nrect = torch.where(write, torch.minimum(wx * wy, counts), torch.zeros_like(counts))
    g = torch.repeat_interleave(torch.arange(centers.shape[0]), nrect)
    first = torch.cumsum(nrect, 0) - nrect
    k = torch.arange(g.numel()) - first[g]
    row, col = k // wx[g].clamp(min=1), k % wx[g].clamp(min=1)
    tile = (lo[g, 1] + row) * tx + (lo[g, 0] + col)
    dbits = depths.view(torch.int32).to(torch.int64)
    keys = torch.zeros(total, dtype=torch.int64)
    ids = torch.zeros(total, dtype=torch.int32)
    slot = (torch.cumsum(counts, 0) - counts)[g] + k
    keys[slot] = (tile << 32) | dbits[g]
    ids[slot] = g.to(torch.int32)
    ks, perm = torch.sort(keys, stable=True)
    ids_sorted = ids[perm]
    tiles = ks >> 32
    edges = torch.searchsorted(tiles, torch.arange(tx * ty + 1, dtype=torch.int64))
    bins = torch.stack([edges[:-1], edges[1:]], 1).to(torch.int32)
    return ids_sorted, bins


# ------------------------------------------------------------------ SH
_SH = (0.28209479177387814, 0.4886025119029199, 1.0925484305920792, 0.31539156525252005,
       0.5462742152960396, 0.5900435899266435, 2.890611442640554, 0.4570457994644658,
       0.3731763325901154, 1.445305721320277)


def sh_colors(coeffs: torch.Tensor, dirs: torch.Tensor, degree: int) -> torch.Tensor:
    """sum_k basis_k(normalize(dir)) coeffs[:, k] for degrees 0..3 (no offset or
    clamp): the real spherical harmonics in the sign convention of sh.cuh."""
    d = dirs / torch.linalg.vector_norm(dirs, dim=-1, keepdim=True)
    x, y, z = d.unbind(-1)
    b = [torch.full_like(x, _SH[0])]
    if degree >= 1:
        b += [-_SH[1] * y, _SH[1] * z, -_SH[1] * x]
    if degree >= 2:
        xx, yy, zz, xy, yz, xz = x * x, y * y, z * z, x * y, y * z, x * z
        b += [_SH[2] * xy, -_SH[2] * yz, _SH[3] * (2 * zz - xx - yy), -_SH[2] * xz,
              _SH[4] * (xx - yy)]
    if degree >= 3:
        b += [-_SH[5] * y * (3 * xx - yy), _SH[6] * xy * z, -_SH[7] * y * (4 * zz - xx - yy),
              _SH[8] * z * (2 * zz - 3 * xx - 3 * yy), -_SH[7] * x * (4 * zz - xx - yy),
              _SH[9] * z * (xx - yy), -_SH[5] * x * (xx - 3 * yy)]
    basis = torch.stack(b, -1)  # (N, D)
    return (basis[:, :, None] * coeffs[:, : basis.shape[1], :]).sum(1)


# ------------------------------------------------------------------ raster
@dataclass
class Camera:
    viewmat: torch.Tensor
    c2w: torch.Tensor
    fx: float
    fy: float
    cx: float
    cy: float
    img_w: int
    img_h: int


def _pixel_rays(cam: Camera, px: torch.Tensor, py: torch.Tensor):
    """Unit rays through pixel centres px + 0.5, py + 0.5
    (``texture_helpers.cuh:336-352``) and their view-space depth factor."""
    ft = cam.c2w.dtype
    u = (px.to(ft) + 0.5 - cam.cx) / cam.fx
    v = (py.to(ft) + 0.5 - cam.cy) / cam.fy
    d = torch.stack([u, v, torch.ones_like(u)], -1)
    R = cam.c2w[:3, :3]
    r = d[..., 0:1] * R[:, 0] + d[..., 1:2] * R[:, 1] + d[..., 2:3] * R[:, 2]
    r = r / torch.sqrt((r * r).sum(-1, keepdim=True))
    vd = (cam.viewmat[2, :3] * r).sum(-1)
    return r, vd


def render_chunk(p: Dict[str, torch.Tensor], cam: Camera, ids_sorted, tile_bins, tiles,
                 pix0: int, pix1: int, bw: int, settings: int, glob_scale: float = 1.0):
    """Outputs of the pixels [pix0, pix1) (row-major inside the tile) of the
    given tiles: dict of (B, P, ...) tensors plus the pixel coordinates and
    the inside mask.  Differentiable in every float tensor of ``p``."""
    tx = (cam.img_w + bw - 1) // bw
    B = tiles.numel()
    r0 = tile_bins[tiles, 0].long()
    L = (tile_bins[tiles, 1] - tile_bins[tiles, 0]).long()
    Lmax = max(int(L.max()), 1)
    k = torch.arange(Lmax)
    ev = k[None, :] < L[:, None]                                   # (B, L)
    e = (r0[:, None] + k[None, :]).clamp(max=max(ids_sorted.numel() - 1, 0))
    gid = torch.where(ev, ids_sorted[e].long(), torch.zeros_like(e)) if ids_sorted.numel() \
        else torch.zeros_like(e)

    pl = torch.arange(pix0, pix1)
    lx, ly = pl % bw, pl // bw
    tyy, txx = tiles // tx, tiles % tx
    px = txx[:, None] * bw + lx[None, :]                           # (B, P)
    py = tyy[:, None] * bw + ly[None, :]
    inside = (px < cam.img_w) & (py < cam.img_h) & (lx < bw)[None, :]
    ray, vd = _pixel_rays(cam, px, py)                             # (B, P, 3), (B, P)

    a1, a2, a3 = quat_axes(p["quats"][gid])                        # (B, L, 3)
    o = cam.c2w[:3, 3]
    mo = p["means"][gid] - o
    sx = p["scales"][gid][..., 0] * glob_scale
    sy = p["scales"][gid][..., 1] * glob_scale
    opac = p["opacities"][gid][..., 0]

    den = torch.einsum("blk,bpk->blp", a3, ray)
    den = torch.where((den >= 0) & (den < EPS_DEN), torch.full_like(den, EPS_DEN), den)
    den = torch.where((den > -EPS_DEN) & (den <= 0), torch.full_like(den, -EPS_DEN), den)
    num = (a3 * mo).sum(-1)
    t = num[..., None] / den                                       # (B, L, P)
    delta = t[..., None] * ray[:, None, :, :] - mo[:, :, None, :]  # (B, L, P, 3)
    l1 = (delta * a1[:, :, None, :]).sum(-1)
    l2 = (delta * a2[:, :, None, :]).sum(-1)
    is1, is2 = 1.0 / sx, 1.0 / sy
    sigma = 0.5 * ((is1 * is1)[..., None] * l1 * l1 + (is2 * is2)[..., None] * l2 * l2)
    G = torch.exp(-sigma)
    if settings & SET_BLUR:
        pv = _xform(cam.viewmat, p["means"][gid])
        mpx, mpy = _project(pv, cam.fx, cam.fy, cam.cx, cam.cy)
        dx = mpx[..., None] - (px.to(mpx.dtype) + 0.5)[:, None, :]
        dy = mpy[..., None] - (py.to(mpy.dtype) + 0.5)[:, None, :]
        sb = 0.5 * 2.0 * (dx * dx + dy * dy)
        sel = sb < sigma
        G = torch.where(sel, torch.exp(-sb), G)
    alpha = torch.clamp(opac[..., None] * G, max=ALPHA_MAX)
    live = ev[:, :, None] & inside[:, None, :]
    skip = (t < T_NEAR) | (t > T_FAR) | (alpha < ALPHA_MIN) | ~live
    one_m = 1.0 - alpha
    f = torch.where(skip, torch.ones_like(alpha), one_m)
    T = torch.cat([torch.ones_like(f[:, :1]), torch.cumprod(f, 1)[:, :-1]], 1)
    with torch.no_grad():
        done = (T * one_m <= T_STOP) & live
        active = torch.cumsum(done.to(torch.int32), 1) == 0
        contrib = active & ~skip
    vis = torch.where(contrib, alpha * T, torch.zeros_like(alpha))
    final_T = torch.where(contrib, one_m, torch.ones_like(one_m)).prod(1)   # (B, P)

    rgb = p["colors"][gid]
    img = (vis[..., None] * rgb[:, :, None, :]).sum(1) + final_T[..., None] * p["background"]
    normal = (vis[..., None] * a3[:, :, None, :]).sum(1)

    # texels
    dl = delta if settings & SET_UV_GRAD else delta.detach()
    uv0 = p["uv0"][gid][..., 0, :]
    um, vm = p["umap"][gid][..., 0, :], p["vmap"][gid][..., 0, :]
    if not settings & SET_UV_GRAD:
        uv0, um, vm = uv0.detach(), um.detach(), vm.detach()
    # clamped to [0, 1] in value; the gradient passes the clamp unchanged, as
    # get_uv_vjp (texture_helpers.cuh:62-80) ignores it
    u = uv0[..., 0:1] + (dl * um[:, :, None, :]).sum(-1)
    v = uv0[..., 1:2] + (dl * vm[:, :, None, :]).sum(-1)
    u = u + (u.detach().clamp(0.0, 1.0) - u.detach())
    v = v + (v.detach().clamp(0.0, 1.0) - v.detach())
    dims = p["texture_dims"][gid].long()
    h, w, st = dims[..., 0:1], dims[..., 1:2], dims[..., 2:3]
    tu, tv = h.to(u.dtype) * u, w.to(v.dtype) * v
    iu0, iv0 = torch.trunc(tu.detach()), torch.trunc(tv.detach())
    fu, fv = tu - iu0, tv - iv0
    iu0, iv0 = iu0.long(), iv0.long()
    iu, ju = torch.minimum(iu0, h - 1), torch.minimum(iu0 + 1, h - 1)
    iv, jv = torch.minimum(iv0, w - 1), torch.minimum(iv0 + 1, w - 1)
    wts = [(1 - fu) * (1 - fv), (1 - fu) * fv, fu * (1 - fv), fu * fv]
    if settings & SET_NEAREST:
        W4 = torch.stack([x.detach() for x in wts], -1)
        pick = torch.zeros_like(W4)
        pick.scatter_(-1, W4.argmax(-1, keepdim=True), 1.0)
        wts = list(pick.unbind(-1))
    X = p["texture"].shape[0]
    idx = [st + iu * w + iv, st + iu * w + jv, st + ju * w + iv, st + ju * w + jv]
    tex = p["texture"]
    val = 0
    for wk, ik in zip(wts, idx):
        val = val + wk[..., None] * tex[ik.clamp(0, max(X - 1, 0))]
    out_tex = (vis[..., None] * val).sum(1)

    # median depth and distortion
    tv_ = t * vd[:, None, :]
    with torch.no_grad():
        med = contrib & (T > 0.5)
        last = (med.to(torch.int64) * (k[None, :, None] + 1)).max(1).values   # (B, P), 0 = none
    depth = torch.gather(tv_, 1, (last - 1).clamp(min=0)[:, None, :])[:, 0, :]
    depth = torch.where(last > 0, depth, torch.zeros_like(depth))
    if settings & SET_NDC:
        dval = (T_FAR * tv_ - T_FAR * T_NEAR) / ((T_FAR - T_NEAR) * tv_)
    else:
        dval = t
    w0, w1, w2 = vis, vis * dval, vis * dval * dval
    S0 = torch.cumsum(w0, 1) - w0
    S1 = torch.cumsum(w1, 1) - w1
    S2 = torch.cumsum(w2, 1) - w2
    reg = (vis * (dval * dval * S0 + S2 - 2.0 * dval * S1)).sum(1)
    return dict(img=img, depth=depth, reg=reg, alpha=1.0 - final_T, texture=out_tex,
                normal=normal, final_T=final_T, px=px, py=py, inside=inside,
                depth_idx=torch.where(last > 0, r0[:, None] + last - 1, torch.full_like(last, -1)))


def _chunks(tile_bins, tiles, bw, pair_budget):
    """Groups of tiles of similar list length (and pixel ranges for tiles whose
    list alone exceeds the budget) of at most ``pair_budget`` pairs each."""
    L = (tile_bins[tiles, 1] - tile_bins[tiles, 0]).long()
    order = torch.argsort(L)
    tiles, L = tiles[order], L[order]
    npix = bw * bw
    out, cur = [], []
    for t, l in zip(tiles.tolist(), L.tolist()):
        l = max(l, 1)
        if l * npix > pair_budget:
            step = max(1, pair_budget // l)
            for a in range(0, npix, step):
                out.append((torch.tensor([t]), a, min(npix, a + step)))
            continue
        if cur and (len(cur) + 1) * l * npix > pair_budget:
            out.append((torch.tensor(cur), 0, npix))
            cur = []
        cur.append(t)
    if cur:
        out.append((torch.tensor(cur), 0, npix))
    return out


OUTPUT_KEYS = ("img", "depth", "reg", "alpha", "texture", "normal")


def render_backward(p: Dict[str, torch.Tensor], cam: Camera, ids_sorted, tile_bins, bw: int,
                    settings: int, cotangents: Optional[Dict[str, torch.Tensor]] = None,
                    target: Optional[torch.Tensor] = None, tiles: Optional[torch.Tensor] = None,
                    pair_budget: int = 4_000_000, glob_scale: float = 1.0):
    """Render the tiles (all by default) chunk by chunk and backpropagate
    either sum(output * cotangent) over the six outputs (``cotangents``, full
    (H, W, ...) images) or the L2 loss mean((img + texture - target)^2) over
    the whole frame (``target``).  Returns (outputs as (H, W, ...) images,
    loss).  Gradients accumulate in the ``.grad`` of the leaf tensors."""
    H, W = cam.img_h, cam.img_w
    tx, ty = (W + bw - 1) // bw, (H + bw - 1) // bw
    if tiles is None:
        tiles = torch.arange(tx * ty)
    C = p["texture"].shape[1]
    z = dict(dtype=p["means"].dtype)
    out = dict(img=torch.zeros(H, W, 3, **z), depth=torch.zeros(H, W, **z),
               reg=torch.zeros(H, W, **z), alpha=torch.zeros(H, W, **z),
               texture=torch.zeros(H, W, C, **z), normal=torch.zeros(H, W, 3, **z),
               depth_idx=torch.full((H, W), -1, dtype=torch.int64))
    loss_total = 0.0
    denom = float(H * W * 3)
    for tl, a, b in _chunks(tile_bins, tiles, bw, pair_budget):
        r = render_chunk(p, cam, ids_sorted, tile_bins, tl, a, b, bw, settings, glob_scale)
        ins = r["inside"]
        yy, xx = r["py"][ins], r["px"][ins]
        loss = 0.0
        if cotangents is not None:
            for key in OUTPUT_KEYS:
                ct = cotangents.get(key)
                if ct is None:
                    continue
                val = r[key][ins]
                loss = loss + (val * ct[yy, xx].reshape(val.shape)).sum()
        else:
            d = r["img"][ins] + r["texture"][ins][:, :3] - target[yy, xx]
            loss = (d * d).sum() / denom
        if torch.is_tensor(loss) and loss.requires_grad:
            loss.backward()
        loss_total += float(loss.detach()) if torch.is_tensor(loss) else float(loss)
        with torch.no_grad():
            for key in OUTPUT_KEYS:
                out[key][yy, xx] = r[key][ins].detach()
            out["depth_idx"][yy, xx] = r["depth_idx"][ins]
    return out, loss_total


def leaf_params(scene, dtype=torch.float32) -> Dict[str, torch.Tensor]:
    """CPU leaf tensors (requires_grad for every float parameter) of a scene
    object with numpy fields (means, scales, quats, opacities, colors, uv0,
    umap, vmap, texture, texture_dims, background[, sh_coeffs])."""
    p = {}
    for k in ("means", "scales", "quats", "opacities", "colors", "uv0", "umap", "vmap", "texture"):
        p[k] = torch.tensor(getattr(scene, k), dtype=dtype).requires_grad_(True)
    p["texture_dims"] = torch.tensor(scene.texture_dims, dtype=torch.int64)
    p["background"] = torch.tensor(scene.background, dtype=dtype)
    sh = getattr(scene, "sh_coeffs", None)
    if sh is not None and getattr(scene, "sh_degree", -1) >= 0:
        p["sh_coeffs"] = torch.tensor(sh, dtype=dtype).requires_grad_(True)
    return p


def camera_of(scene) -> Camera:
    return Camera(viewmat=torch.tensor(scene.viewmat), c2w=torch.tensor(scene.c2w),
                  fx=float(scene.fx), fy=float(scene.fy), cx=float(scene.cx), cy=float(scene.cy),
                  img_w=int(scene.img_w), img_h=int(scene.img_h))


def preprocess(p, cam: Camera, sh_degree: int, bw: int):
    """SH colours (when the scene has coefficients), projection and binning of
    one view: returns (params with 'colors' set, ids_sorted, tile_bins)."""
    q = dict(p)
    if "sh_coeffs" in p:
        dirs = p["means"].detach() - cam.c2w[:3, 3]
        q["colors"] = sh_colors(p["sh_coeffs"], dirs, sh_degree)
    centers, extents, depths = project(p["means"].detach(), p["scales"].detach(),
                                       p["quats"].detach(), cam.viewmat, cam.fx, cam.fy,
                                       cam.cx, cam.cy)
    ids, bins = bin_gaussians(centers, extents, depths, cam.img_w, cam.img_h, bw)
    return q, ids, bins


def train_iteration(p, cam: Camera, target: torch.Tensor, sh_degree: int, bw: int,
                    settings: int, tiles: Optional[torch.Tensor] = None,
                    pair_budget: int = 4_000_000):
    """One fwd+bwd iteration of the L2 photometric loss (img + texture vs
    target): SH, projection, binning, chunked render with autograd.  The SH
    colours are a non-leaf tensor, so their graph is kept across chunks and
    released by one final backward of the accumulated colour gradient."""
    for t in p.values():
        if torch.is_tensor(t) and t.grad is not None:
            t.grad = None
    q, ids, bins = preprocess(p, cam, sh_degree, bw)
    if "sh_coeffs" in p:
        col = q["colors"]
        leaf_col = col.detach().requires_grad_(True)
        q["colors"] = leaf_col
        out, loss = render_backward(q, cam, ids, bins, bw, settings, target=target, tiles=tiles,
                                    pair_budget=pair_budget)
        if leaf_col.grad is not None:
            col.backward(leaf_col.grad)
    else:
        out, loss = render_backward(q, cam, ids, bins, bw, settings, target=target, tiles=tiles,
                                    pair_budget=pair_budget)
    return out, loss, ids, bins


def host_cpu() -> Dict[str, object]:
    """Thread count torch uses, os.cpu_count(), the CPUs this process may run
    on, and the CPU model string."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"threads": torch.get_num_threads(), "os_cpu_count": os.cpu_count(),
            "affinity_cpus": affinity, "cpu_model": model}
