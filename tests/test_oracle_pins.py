"""Pins of the CPU oracle (CPU only).

Each check ties an oracle function to something it does not itself compute:
scipy's spherical harmonics, closed-form renders of camera-facing splats,
numpy's stable sort and searchsorted, geometric tile overlap, adjoint (dot
product) identities, and FP64 central finite differences of the FP64 forward.
"""
import numpy as np
import pytest

from helpers import (camera_facing_scene, frac_close, load_kat, oracle_scene, psnr,
                     tile_bounds, view_depths)

KAT = load_kat()


# ------------------------------------------------------------------ SH
def _basis(O, d, precision):
    d = np.asarray(d, np.float64).reshape(1, 3)
    out = []
    for k in range(25):
        c = np.zeros((1, 25, 3))
        c[0, k, 0] = 1.0
        out.append(O.sh_forward(4, 4, d, c, precision)[0, 0])
    return np.array(out)


def test_sh_basis_matches_scipy(oracle):
    for case in KAT["sh"]:
        b64 = _basis(oracle, case["dir"], "f64")
        np.testing.assert_allclose(b64, case["basis"], rtol=0, atol=1e-12)
        b32 = _basis(oracle, case["dir"], "f32")
        np.testing.assert_allclose(b32, case["basis"], rtol=0, atol=2e-6)


@pytest.mark.parametrize("degree,used", [(0, 0), (1, 1), (3, 2), (3, 3), (4, 4)])
def test_sh_backward_is_adjoint_of_forward(oracle, degree, used):
    rng = np.random.default_rng(degree * 10 + used)
    n, nb = 64, (degree + 1) ** 2
    dirs = rng.standard_normal((n, 3))
    coeffs = rng.standard_normal((n, nb, 3))
    vcol = rng.standard_normal((n, 3))
    col = oracle.sh_forward(degree, used, dirs, coeffs, "f64")
    vco = oracle.sh_backward(degree, used, dirs, vcol, "f64")
    lhs = float((col * vcol).sum())
    rhs = float((coeffs * vco).sum())
    assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(lhs))
    # bands above degrees_to_use get no gradient and contribute no colour
    assert np.all(vco[:, (used + 1) ** 2:, :] == 0)


def test_sh_degree0_is_c0_times_coeff(oracle):
    c = np.random.default_rng(0).uniform(0, 1, (8, 1, 3))
    out = oracle.sh_forward(0, 0, np.ones((8, 3)), c, "f64")
    np.testing.assert_allclose(out, 0.28209479177387814 * c[:, 0, :], rtol=1e-15, atol=0)


# ------------------------------------------------------------------ projection
def test_aabb_camera_facing_closed_form(oracle):
    eye = np.eye(4)
    for case in KAT["aabb"]:
        m = np.array([case["mean"]])
        s = np.array([case["scale"]])
        q = np.array([[1.0, 0, 0, 0]])
        for prec, tol in (("f64", 1e-9), ("f32", 2e-4)):
            c, e = oracle.get_aabb_2d(m, s, case["glob_scale"], q, eye, case["fx"], case["fy"],
                                      case["cx"], case["cy"], prec)
            # the corners sit at the mean's depth, so the box is exact up to the 1e-6 guard
            np.testing.assert_allclose(c[0], case["center"], rtol=1e-5, atol=tol)
            np.testing.assert_allclose(e[0], case["extent"], rtol=1e-5, atol=tol)


def test_aabb_rotation_swaps_extents(oracle):
    eye = np.eye(4)
    m = np.array([[0.1, 0.2, 3.0]])
    s = np.array([[0.05, 0.02, 0.0]])
    q90 = np.array([[np.cos(np.pi / 4), 0, 0, np.sin(np.pi / 4)]])  # 90 deg about z
    _, e0 = oracle.get_aabb_2d(m, s, 1.0, np.array([[1.0, 0, 0, 0]]), eye, 100, 100, 50, 50, "f64")
    _, e1 = oracle.get_aabb_2d(m, s, 1.0, q90, eye, 100, 100, 50, 50, "f64")
    np.testing.assert_allclose(e1[0], e0[0][::-1], rtol=1e-9)


def test_aabb_clipped_mean_gives_zero_extent(oracle):
    eye = np.eye(4)
    m = np.array([[0.3, -0.2, 0.005], [0.3, -0.2, -2.0]])
    s = np.full((2, 3), 0.1)
    q = np.tile([1.0, 0, 0, 0], (2, 1))
    c, e = oracle.get_aabb_2d(m, s, 1.0, q, eye, 100, 100, 50, 40, "f64")
    assert np.all(e == 0)
    z = m[:, 2] + 1e-6
    np.testing.assert_allclose(c[:, 0], 100 * m[:, 0] / z + 50, rtol=1e-12)
    np.testing.assert_allclose(c[:, 1], 100 * m[:, 1] / z + 40, rtol=1e-12)


# ------------------------------------------------------------------ binning
def _random_boxes(rng, n, W, H):
    c = np.stack([rng.uniform(-20, W + 20, n), rng.uniform(-20, H + 20, n)], 1).astype(np.float32)
    e = rng.uniform(0.0, 40.0, (n, 2)).astype(np.float32)
    e[rng.uniform(size=n) < 0.1] = 0.0  # some clipped (zero-extent) boxes
    return c, e


@pytest.mark.parametrize("bw", [16, 8])
def test_num_tiles_hit_is_geometric_overlap(oracle, bw):
    rng = np.random.default_rng(bw)
    W, H = 200, 120
    c, e = _random_boxes(rng, 500, W, H)
    got = oracle.num_tiles_hit(c, e, H, W, bw)
    tx, ty = tile_bounds(W, H, bw)[:2]
    lo = c.astype(np.float64) - e
    hi = c.astype(np.float64) + e
    want = np.ones(len(c), np.int64)
    for k, T in ((0, tx), (1, ty)):
        t = np.arange(T)
        ov = (t[None, :] * bw <= hi[:, k:k + 1]) & ((t[None, :] + 1) * bw > lo[:, k:k + 1])
        want *= ov.sum(1)
    # boxes whose edges land exactly on a tile boundary are excluded (float rounding)
    on_edge = np.any((np.abs(lo / bw - np.round(lo / bw)) < 1e-5) |
                     (np.abs(hi / bw - np.round(hi / bw)) < 1e-5), 1)
    assert np.array_equal(got[~on_edge], want[~on_edge])


def test_intersects_cover_exactly_the_overlapped_tiles(oracle):
    rng = np.random.default_rng(3)
    W, H, bw = 200, 120, 16
    tb = tile_bounds(W, H, bw)
    c, e = _random_boxes(rng, 300, W, H)
    d = rng.uniform(0.5, 30.0, 300).astype(np.float32)
    nt = oracle.num_tiles_hit(c, e, H, W, bw)
    cum = np.cumsum(nt).astype(np.int32)
    keys, ids = oracle.map_gaussian_to_intersects(c, e, d, cum, tb, bw)
    for g in range(len(c)):
        lo = 0 if g == 0 else cum[g - 1]
        seg_k, seg_i = keys[lo:cum[g]], ids[lo:cum[g]]
        if e[g, 0] <= 1e-4 and e[g, 1] <= 1e-4:
            assert np.all(seg_k == 0) and np.all(seg_i == 0)  # counted, never written
            continue
        tiles = set()
        for ty in range(tb[1]):
            for tx in range(tb[0]):
                if (tx * bw <= c[g, 0] + e[g, 0] and (tx + 1) * bw > c[g, 0] - e[g, 0] and
                        ty * bw <= c[g, 1] + e[g, 1] and (ty + 1) * bw > c[g, 1] - e[g, 1]):
                    tiles.add(ty * tb[0] + tx)
        assert set((seg_k >> 32).tolist()) == tiles
        assert np.all(seg_i == g)
        assert np.all((seg_k & 0xFFFFFFFF) == d[g:g + 1].view(np.uint32)[0])


def test_zero_slot_quirk_sorts_first_in_tile0(oracle):
    # a clipped Gaussian whose projected centre is inside the image is counted
    # (one tile) but writes nothing: its slot stays key 0 / id 0
    c = np.array([[40.0, 40.0], [5.0, 5.0], [60.0, 10.0]], np.float32)
    e = np.array([[10.0, 10.0], [0.0, 0.0], [3.0, 3.0]], np.float32)
    d = np.array([4.0, 0.005, 2.0], np.float32)
    keys, ids, ks, isd, bins = oracle.bin_and_sort(c, e, d, oracle.num_tiles_hit(c, e, 64, 64, 16),
                                                   (4, 4, 1), 16)
    assert (keys == 0).sum() == 1
    assert ks[0] == 0 and isd[0] == 0
    assert bins[0, 0] == 0 and bins[0, 1] >= 1


def test_wrapped_binning_hand_case(oracle):
    # forward.cu:34-62 with get_tile_bbox_wrapped (helpers.cuh:53-73,94-111) on a
    # 4 x 3 tile grid, bw 16.  Gaussian 0: centre (5, 5), extent (10, 3) spans
    # tiles x, y in [-1, 1) (a lower corner <= 0 moves one tile further), which
    # wrap to x 3, 0 and y 2, 0: row by row tiles 11, 8, 3, 0.  Gaussian 1:
    # centre (40, 20), extent (2, 2) is the single tile (2, 1) = 6.
    c = np.array([[5.0, 5.0], [40.0, 20.0]], np.float32)
    e = np.array([[10.0, 3.0], [2.0, 2.0]], np.float32)
    d = np.array([2.0, 3.0], np.float32)
    dbits = d.view(np.int32).astype(np.int64)
    cnt = oracle.num_tiles_hit(c, e, 48, 64, 16, wrapped=True)
    assert cnt.tolist() == [4, 1]
    assert oracle.num_tiles_hit(c, e, 48, 64, 16).tolist() == [1, 1]
    keys, ids = oracle.map_gaussian_to_intersects(c, e, d, np.cumsum(cnt).astype(np.int32),
                                                  (4, 3, 1), 16, wrapped=True)
    tiles = np.array([11, 8, 3, 0, 6], np.int64)
    assert np.array_equal(keys, (tiles << 32) | dbits[[0, 0, 0, 0, 1]])
    assert ids.tolist() == [0, 0, 0, 0, 1]
    # with the clamped counts each Gaussian fills only its own window
    keys, ids = oracle.map_gaussian_to_intersects(c, e, d, np.array([1, 2], np.int32),
                                                  (4, 3, 1), 16, wrapped=True)
    assert np.array_equal(keys, (np.array([11, 6], np.int64) << 32) | dbits)


def test_sort_matches_numpy_stable(oracle):
    rng = np.random.default_rng(11)
    keys = (rng.integers(0, 40, 5000).astype(np.int64) << 32) | rng.integers(0, 50, 5000)
    ids = rng.integers(0, 1 << 30, 5000).astype(np.int32)
    ks, is_ = oracle.sort_pairs(keys, ids)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ks, keys[order])
    assert np.array_equal(is_, ids[order])


def test_bin_edges_match_searchsorted(oracle):
    rng = np.random.default_rng(5)
    T = 37
    tiles = np.sort(rng.integers(0, T, 3000))
    keys = (tiles.astype(np.int64) << 32) | rng.integers(0, 1000, 3000)
    bins = oracle.tile_bin_edges(keys, T)
    lo = np.searchsorted(tiles, np.arange(T), "left")
    hi = np.searchsorted(tiles, np.arange(T), "right")
    empty = lo == hi
    assert np.array_equal(bins[~empty, 0], lo[~empty])
    assert np.array_equal(bins[~empty, 1], hi[~empty])
    assert np.all(bins[empty] == 0)


# ------------------------------------------------------------------ sampling
def test_bilinear_reproduces_affine_texels(oracle):
    bl = KAT["bilinear"]
    tex = np.array(bl["texture"])
    dims = np.array([q["dims"] for q in bl["queries"]], np.int32)
    uvs = np.array([q["uv"] for q in bl["queries"]])
    want = np.array([q["value"] for q in bl["queries"]])
    np.testing.assert_allclose(oracle.texture_sample(dims, uvs, tex, "f64"), want, atol=1e-12)
    np.testing.assert_allclose(oracle.texture_sample(dims, uvs, tex, "f32"), want, atol=1e-5)


def test_bilinear_vjp_is_adjoint(oracle):
    rng = np.random.default_rng(2)
    dims = np.array([[3, 4, 0], [2, 2, 12], [5, 1, 16]], np.int32)
    tex = rng.standard_normal((21, 4))
    q = rng.integers(0, 3, 200)
    uvs = rng.uniform(0, 1.2, (200, 2))
    v = rng.standard_normal((200, 4))
    out = oracle.texture_sample(dims[q], uvs, tex, "f64")
    vt = oracle.texture_sample_vjp(dims[q], uvs, v, 21, "f64")
    assert abs(float((out * v).sum()) - float((tex * vt).sum())) < 1e-10


# ------------------------------------------------------------------ raster forward
def _splat1_scene(k):
    m = k["mean"]
    return camera_facing_scene([m], [k["scale"]], [k["opacity"]], [k["color"]], k["W"], k["H"],
                               k["f"], k["cx"], k["cy"], texel=k["texel"],
                               background=k["background"])


def test_single_splat_closed_form(oracle):
    k = KAT["splat1"]
    s = _splat1_scene(k)
    for prec, tol in (("f64", 1e-9), ("f32", 2e-5)):
        f = oracle.texture_forward(oracle_scene(oracle, s), prec)
        alpha = np.array(k["alpha"])
        np.testing.assert_allclose(f["out_img"], np.array(k["img"]), atol=tol)
        np.testing.assert_allclose(1 - f["final_Ts"], alpha, atol=tol)
        np.testing.assert_allclose(f["out_normal"][..., 2], alpha, atol=tol)
        np.testing.assert_allclose(f["out_texture"], alpha[..., None] * np.array(k["texel"]), atol=tol)
        hit = alpha > 0
        np.testing.assert_allclose(f["out_depth"][hit], k["mean"][2], rtol=1e-5)
        assert np.all(f["out_depth"][~hit] == 0) and np.all(f["depth_idx"][~hit] == -1)
        assert np.all(f["out_reg"] == 0)


@pytest.mark.parametrize("case", [0, 1])
def test_textured_splat_closed_form(oracle, case):
    # tests/golden/make_golden.py splat_tex: one camera-facing splat (case 1
    # turned in its plane) over a 4 x 5 chart affine in its texel indices;
    # the expected texture is alpha * (A + B min(h u, h-1) + C min(w v, w-1))
    # with u, v from the closed-form plane hit, so a swapped u/v, umap/vmap,
    # h/w or axis fails it
    from helpers import textured_splat_scene

    k = KAT["splat_tex"][case]
    s = textured_splat_scene(k)
    chk = np.array(k["check"], bool)
    for prec, tol in (("f64", 1e-9), ("f32", 2e-5)):
        f = oracle.texture_forward(oracle_scene(oracle, s), prec)
        np.testing.assert_allclose(f["out_texture"][chk], np.array(k["out_texture"])[chk],
                                   atol=tol)
        np.testing.assert_allclose(f["out_img"][chk], np.array(k["img"])[chk], atol=tol)
        np.testing.assert_allclose((1 - f["final_Ts"])[chk], np.array(k["alpha"])[chk], atol=tol)


@pytest.mark.parametrize("outline_q,white", [(0, False), (6, False), (6, True), (11, True)])
def test_alpha_visualisation_single_splat_geometry(oracle, outline_q, white):
    # texture.cu:56-65,192-241 with local_outline (texture_helpers.cuh:390-417):
    # a camera-facing circular splat at depth d with scale s is drawn with
    # alpha 0.99 inside radius R = B s f / d pixels (alpha bound B); a pixel
    # whose 7 x 7 neighbourhood reaches outside the circle within the outline
    # bound is drawn white (bit 24) or not at all.  Expected values come from
    # circle geometry on the pixel grid, not from the oracle's ray solves.
    from helpers import camera_facing_scene

    W = H = 64
    f, d, sc, B = 64.0, 4.0, 0.3, 2.5
    R = B * sc * f / d  # 12 px; centres at half-integer offsets never lie on it
    col, tex, bg = [0.2, 0.4, 0.6], [0.3, 0.5, 0.7], [0.1, 0.1, 0.1]
    s = camera_facing_scene([[0.0, 0.0, d]], [[sc, sc, 1e-3]], [1.0], [col], W, H, f, W / 2,
                            H / 2, texel=tex, background=bg)
    settings = (1 << 8) | (1 << 16) | (int(B * 8) << 17) | (outline_q << 26) | (white << 24)
    jj, ii = np.meshgrid(np.arange(W) + 0.5 - W / 2, np.arange(H) + 0.5 - H / 2)
    inside = jj ** 2 + ii ** 2 <= R * R
    dis = np.full((H, W), 1000.0)
    for dx in range(-3, 4):
        for dy in range(-3, 4):
            out = (jj + dx) ** 2 + (ii + dy) ** 2 > R * R
            dis = np.where(out, np.minimum(dis, np.hypot(dx, dy)), dis)
    plain = dis > outline_q / 4.0
    a = float(np.float32(0.99))  # the reference's 0.99f
    rgb = np.where(plain[..., None], a * np.array(col), a if white else 0.0)
    img = np.where(inside[..., None], rgb + (1 - a) * np.array(bg), np.array(bg))
    texo = np.where(inside[..., None],
                    np.where(plain[..., None], a * np.array(tex), a if white else 0.0), 0.0)
    outlined = int((~plain & inside).sum())
    assert inside.sum() > 400 and (outlined == 0 if outline_q == 0 else outlined > 50)
    for prec, tol in (("f64", 1e-9), ("f32", 2e-5)):
        f_ = oracle.texture_forward(oracle_scene(oracle, s, settings=settings), prec)
        np.testing.assert_allclose(f_["out_img"], img, atol=tol)
        np.testing.assert_allclose(f_["out_texture"], texo, atol=tol)
        np.testing.assert_allclose(f_["final_Ts"], np.where(inside, 1 - a, 1.0), atol=tol)


def test_two_splats_compositing_and_distortion(oracle):
    k = KAT["splat2"]
    n = 2
    s = camera_facing_scene([[0, 0, z] for z in k["z"]], [[k["scale"], k["scale"], 0]] * n,
                            k["opacity"], k["color"], k["W"], k["H"], k["f"], k["cx"], k["cy"],
                            background=k["background"])
    f = oracle.texture_forward(oracle_scene(oracle, s), "f64")
    for px in k["pixels"]:
        i, j = px["pixel"]
        np.testing.assert_allclose(f["out_img"][i, j], px["img"], atol=1e-10)
        np.testing.assert_allclose(f["final_Ts"][i, j], px["T"], atol=1e-10)
        np.testing.assert_allclose(f["out_reg"][i, j], px["reg"], rtol=1e-8)
        np.testing.assert_allclose(f["out_depth"][i, j], px["depth"], rtol=1e-10)


def test_forward_invariants_random_scene(oracle):
    from gstex_cuda_amd.synthetic import make_scene

    s = make_scene(1, seed=21, num_points=400, img_w=96, img_h=80)
    sc = oracle_scene(oracle, s, colors=np.ones_like(s.colors))
    sc.background = np.zeros(3, np.float32)
    f = oracle.texture_forward(sc, "f64")
    T = f["final_Ts"]
    np.testing.assert_allclose(f["out_img"], np.repeat((1 - T)[..., None], 3, -1), atol=1e-12)
    np.testing.assert_allclose(f["out_reg_s"][..., 0], 1 - T, atol=1e-12)
    assert np.all(f["out_reg"] >= -1e-12)
    assert np.all((T > 1e-4) & (T <= 1.0))
    has = f["depth_idx"] >= 0
    assert np.all(f["depth_idx"][has] <= f["final_idx"][has])


def test_f32_restatement_tracks_f64(oracle):
    from gstex_cuda_amd.synthetic import make_scene

    s = make_scene(1)
    sc = oracle_scene(oracle, s)
    f32 = oracle.texture_forward(sc, "f32")
    f64 = oracle.texture_forward(sc, "f64")
    assert psnr(f32["out_img"], f64["out_img"]) > 80
    assert frac_close(f32["final_idx"], f64["final_idx"], 0) > 0.999


# ------------------------------------------------------------------ raster backward vs FD
def _fd_scene(seed):
    """Small scene away from the reference's non-differentiable clamps: alpha
    below the 0.99 cap and uv clamped only where alpha is already below 1/255."""
    from gstex_cuda_amd.synthetic import SceneSpec, make_scene

    spec = SceneSpec(num_points=10, img_w=40, img_h=36, focal=40.0, z_range=(2.0, 3.0),
                     texel_side=3, uv_scale=12.0, background=(0.2, 0.1, 0.3))
    s = make_scene(seed=seed, spec=spec)
    rng = np.random.default_rng(seed + 100)
    s.means[:, :2] *= 0.6
    s.scales[:, :2] = rng.uniform(0.08, 0.25, (10, 2)).astype(np.float32)
    s.opacities[:] = rng.uniform(0.2, 0.85, (10, 1)).astype(np.float32)
    # rebuild the uv maps for the new scales: |u - 0.5| <= 0.5 up to 6 sigma, where
    # alpha < 1/255 already (the reference passes gradient through the uv clamp)
    from gstex_cuda_amd.synthetic import quat_axes_np

    a1, a2, _ = quat_axes_np(s.quats)
    s.umap = (a1 / (12.0 * s.scales[:, :1])).astype(np.float32).reshape(10, 1, 3)
    s.vmap = (a2 / (12.0 * s.scales[:, 1:2])).astype(np.float32).reshape(10, 1, 3)
    return s


@pytest.mark.parametrize("seed,settings", [(1, 1 << 8), (2, 1 << 8), (3, 0),
                                           (4, (1 << 8) | (1 << 10)), (5, (1 << 8) | (1 << 9))])
def test_backward_matches_central_differences(oracle, seed, settings):
    s = _fd_scene(seed)
    sc = oracle_scene(oracle, s, settings=settings)
    H, W, C = s.img_h, s.img_w, s.texture.shape[1]
    rng = np.random.default_rng(seed)
    w = dict(img=rng.standard_normal((H, W, 3)), depth=rng.standard_normal((H, W)),
             reg=rng.standard_normal((H, W)), alpha=rng.standard_normal((H, W)),
             tex=rng.standard_normal((H, W, C)), normal=rng.standard_normal((H, W, 3)))
    if not settings & (1 << 8):
        # without bit 8 the reference detaches the texture lookup from the
        # geometry and the uv map (texture.cu:629-641); keep the texture out of the loss
        w["tex"][:] = 0.0

    params = {k: np.array(getattr(sc, k), np.float64) for k in
              ("means", "scales", "quats", "colors", "opacities", "uv0", "umap", "vmap", "texture")}

    def loss():
        for k, v in params.items():
            setattr(sc, k, v)
        f = oracle.texture_forward(sc, "f64")
        return float((w["img"] * f["out_img"]).sum() + (w["depth"] * f["out_depth"]).sum() +
                     (w["reg"] * f["out_reg"]).sum() + (w["alpha"] * (1 - f["final_Ts"])).sum() +
                     (w["tex"] * f["out_texture"]).sum() + (w["normal"] * f["out_normal"]).sum()), f

    _, fwd = loss()
    g = oracle.texture_backward(sc, fwd, w["img"], w["depth"], w["reg"], w["alpha"], w["tex"],
                                w["normal"], "f64")
    names = {"means": "v_means", "scales": "v_scales", "quats": "v_quats", "colors": "v_colors",
             "opacities": "v_opacity", "uv0": "v_uv0", "umap": "v_umap", "vmap": "v_vmap",
             "texture": "v_texture"}
    h = 1e-6
    for p, gname in names.items():
        arr = params[p]
        ana = g[gname].reshape(arr.shape)
        flat = arr.reshape(-1)
        idx = range(flat.size) if flat.size <= 120 else rng.choice(flat.size, 120, replace=False)
        num = np.zeros(flat.size)
        for i in idx:
            x0 = flat[i]
            flat[i] = x0 + h
            lp, _ = loss()
            flat[i] = x0 - h
            lm, _ = loss()
            flat[i] = x0
            num[i] = (lp - lm) / (2 * h)
        sel = np.zeros(flat.size, bool)
        sel[list(idx)] = True
        a, n_ = ana.reshape(-1)[sel], num[sel]
        if p == "scales":  # the third scale component is unused by the renderer
            assert np.all(ana.reshape(-1, 3)[:, 2] == 0)
        scale = max(np.abs(n_).max(), 1e-8)
        err = np.abs(a - n_) / scale
        assert np.mean(err < 1e-4) >= 0.98 and np.max(err) < 5e-2, (p, float(err.max()))


# ---------------------------------------------------------------- texture_edit
def _edit_scene(s):
    s.texture = np.zeros((s.texture.shape[0], 5), np.float32)
    return s


def test_texture_edit_single_splat_counts_pixels(oracle):
    """One camera-facing splat with a 1x1 chart: the weight channel is the number
    of pixels whose alpha passes 1/255 (closed form: the ellipse
    (x/s)^2 + (y/s)^2 <= 2 ln(255 opac) on the plane), and the colour channels
    are sums of r*a, g*a, b*a over those pixels."""
    W = H = 48
    f, z, sc, op = 48.0, 3.0, 0.2, 0.8
    s = _edit_scene(camera_facing_scene([[0.0, 0.0, z]], [[sc, sc, 0.0]], [op], [[1, 1, 1]], W,
                                        H, f, W / 2, H / 2))
    rng = np.random.default_rng(1)
    img = rng.random((H, W, 3)).astype(np.float32)
    al = rng.random((H, W)).astype(np.float32)
    lo = np.zeros((H, W), np.float32)
    hi = np.full((H, W), 1e9, np.float32)
    out = oracle.texture_edit(oracle_scene(oracle, s), img, al, lo, hi, "f64")
    # ray through pixel (j, i) hits the plane z at (x, y) = z * ((j+.5-cx)/f, (i+.5-cy)/f)
    jj, ii = np.meshgrid(np.arange(W), np.arange(H))
    x = z * (jj + 0.5 - W / 2) / f
    y = z * (ii + 0.5 - H / 2) / f
    sig = 0.5 * (x * x + y * y) / (sc * sc)
    hit = np.minimum(0.99, op * np.exp(-sig)) >= 1 / 255
    assert out[0, 4] == hit.sum()
    al64, img64 = al.astype(np.float64), img.astype(np.float64)
    np.testing.assert_allclose(out[0, 3], al64[hit].sum(), rtol=1e-12)
    for k in range(3):
        np.testing.assert_allclose(out[0, k], (img64[..., k] * al64)[hit].sum(), rtol=1e-12)


def test_texture_edit_channel_identities_and_depth_windows(oracle):
    """White canvas: channels 0-2 equal channel 3; opaque canvas: channel 3 equals
    channel 4; an empty depth window gives zero; two windows split at a depth
    add up to the full window."""
    from gstex_cuda_amd import synthetic

    s = _edit_scene(synthetic.make_scene(1, num_points=300, img_w=64, img_h=48))
    sc = oracle_scene(oracle, s, settings=0)
    H, W = s.img_h, s.img_w
    rng = np.random.default_rng(2)
    img = rng.random((H, W, 3)).astype(np.float32)
    al = rng.random((H, W)).astype(np.float32)
    lo = np.zeros((H, W), np.float32)
    hi = np.full((H, W), 1e9, np.float32)
    full = oracle.texture_edit(sc, img, al, lo, hi, "f64")
    assert full[:, 4].sum() > 0
    white = oracle.texture_edit(sc, np.ones_like(img), al, lo, hi, "f64")
    for k in range(3):
        np.testing.assert_allclose(white[:, k], white[:, 3], rtol=1e-12)
    opaque = oracle.texture_edit(sc, img, np.ones_like(al), lo, hi, "f64")
    np.testing.assert_allclose(opaque[:, 3], opaque[:, 4], rtol=1e-12)
    none = oracle.texture_edit(sc, img, al, hi, hi, "f64")
    assert np.all(none == 0)
    d = np.full((H, W), 4.0, np.float32)
    near = oracle.texture_edit(sc, img, al, lo, d, "f64")
    far = oracle.texture_edit(sc, img, al, np.nextafter(d, np.float32(np.inf)), hi, "f64")
    np.testing.assert_allclose(near + far, full, rtol=1e-12, atol=1e-12)


def test_saturated_pixel_stops_exactly():
    """A pixel with fl(T (1 - 1/255)) <= 1e-4 is ended by every entry with
    alpha >= 1/255 (texture.cu:213-221: the end test precedes compositing), so
    the raster stops it at once; a live pixel is never ended by an entry with
    alpha < 1/255, so culled entries never matter.  Checked in float32 on
    random pairs and on the float neighbourhood of the boundary."""
    f = np.float32
    keep = f(1) - f(1) / f(255)
    lim = f(1e-4)
    amin = f(1) / f(255)
    rng = np.random.default_rng(0)
    # every float T in a band around the boundary, with the extreme alphas
    T = np.arange(0, 120000, dtype=np.int64) + np.float32(9.99e-5).view(np.int32)
    T = T.astype(np.int32).view(np.float32)
    stopped = (T * keep) <= lim
    assert stopped.any() and (~stopped).any()
    for alpha in (amin, np.nextafter(amin, f(1)), f(0.5), f(0.99)):
        assert np.all((T[stopped] * (f(1) - alpha)) <= lim)
    below = np.nextafter(amin, f(0))
    for alpha in (below, f(1e-3), f(0)):
        assert np.all((T[~stopped] * (f(1) - alpha)) > lim)
    # random pairs over the whole ranges
    Tr = rng.uniform(9.9e-5, 1.02e-4, 2_000_000).astype(f)
    ar = rng.uniform(float(amin), 0.99, 2_000_000).astype(f)
    st = (Tr * keep) <= lim
    assert np.all((Tr[st] * (f(1) - ar[st])) <= lim)
    ab = rng.uniform(0.0, float(below), 2_000_000).astype(f)
    assert np.all((Tr[~st] * (f(1) - ab[~st])) > lim)
