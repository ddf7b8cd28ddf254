"""GPU parity: every HIP entry point, through the C ABI, against the oracle on
the same seeded inputs.

Tolerances (FP32 kernels vs the FP64 oracle unless stated):
  projection    centres/extents within 1e-4 px + 1e-5 relative
  integer work  tile counts, intersect keys, ids, sorted order, bin edges:
                bit-exact given the same float inputs
  SH, sampling  1e-5 absolute
  raster fwd    >= 99.9% of pixels within 1e-4 of the FP64 oracle, image
                PSNR >= 60 dB, final_idx equal on >= 99.9% of pixels (a pair
                whose alpha sits on the 1/255 or 1e-4 thresholds may round the
                other way in FP32)
  raster bwd    per gradient tensor: relative L2 error <= 2e-3 and 99% of
                entries within 2e-3 x the tensor's max magnitude
"""
import numpy as np
import pytest
import torch

from gpu_helpers import (DEV, gpu_backward, gpu_forward, grad_close, log, np_, oracle_with,
                         project_and_bin, random_grads, raster_inputs, to_dev)
from helpers import camera_facing_scene, frac_close, load_kat, psnr, tile_bounds, view_depths

pytestmark = pytest.mark.gpu

KAT = load_kat()


def scene(workload=1, **kw):
    from gstex_cuda_amd.synthetic import make_scene

    return make_scene(workload, **kw)


# ------------------------------------------------------------------ projection
@pytest.mark.parametrize("n", [1, 1000, 70001])
def test_aabb_matches_oracle(oracle, n):
    from gstex_cuda_amd import cuda as _C

    s = scene(2, num_points=n, seed=n)
    log("aabb", n)
    c, e = _C.get_aabb_2d(to_dev(s.means), to_dev(s.scales), 1.0, to_dev(s.quats),
                          to_dev(s.viewmat), s.fx, s.fy, s.cx, s.cy)
    c64, e64 = oracle.get_aabb_2d(s.means, s.scales, 1.0, s.quats, s.viewmat, s.fx, s.fy, s.cx,
                                  s.cy, "f64")
    np.testing.assert_allclose(np_(c), c64, rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(np_(e), e64, rtol=1e-5, atol=1e-4)
    # the fused projection returns the same boxes, the view depth and the tile counts
    pb = project_and_bin(s)
    assert torch.equal(pb["centers"], c) and torch.equal(pb["extents"], e)
    np.testing.assert_allclose(np_(pb["depths"]), s.means[:, 2], rtol=1e-6)
    want = oracle.num_tiles_hit(np_(c), np_(e), s.img_h, s.img_w, 16)
    assert np.array_equal(np_(pb["num_tiles_hit"]), want)


def test_project_points_and_clip():
    from gstex_cuda_amd.get_aabb_2d import project_points

    pts = np.array([[0.1, 0.2, 2.0], [0.0, 0.0, 0.001], [1.0, -1.0, 2000.0]], np.float32)
    vm = to_dev(np.eye(4, dtype=np.float32))
    pix, d = project_points(to_dev(pts), vm, (100.0, 100.0, 50.0, 40.0))
    np.testing.assert_allclose(np_(d), pts[:, 2], rtol=1e-7)
    np.testing.assert_allclose(np_(pix)[0], [100 * 0.1 / 2.0 + 50, 100 * 0.2 / 2.0 + 40], rtol=1e-6)
    _, dc = project_points(to_dev(pts), vm[:3], (100.0, 100.0, 50.0, 40.0), clip=True)
    np.testing.assert_allclose(np_(dc), [2.0, 0.01, 1000.0], rtol=1e-7)


# ------------------------------------------------------------------ binning
def _binning_case(oracle, s, bw):
    from gstex_cuda_amd import cuda as _C
    from gstex_cuda_amd.utils import bin_and_sort_gaussians, compute_cumulative_intersects

    pb = project_and_bin(s, bw)
    c, e, d, nt = pb["centers"], pb["extents"], pb["depths"], pb["num_tiles_hit"]
    m, cum = compute_cumulative_intersects(nt)
    tb = tile_bounds(s.img_w, s.img_h, bw)
    log("binning", s.num_points, s.img_w, s.img_h, bw, "I =", m)
    keys, ids, ks, isd, bins = bin_and_sort_gaussians(s.num_points, m, c, e, d, cum, tb, bw)
    o_keys, o_ids = oracle.map_gaussian_to_intersects(np_(c), np_(e), np_(d), np_(cum), tb, bw)
    assert np.array_equal(np_(keys), o_keys)
    assert np.array_equal(np_(ids), o_ids)
    o_ks, o_is = oracle.sort_pairs(o_keys, o_ids)
    assert np.array_equal(np_(ks), o_ks)
    assert np.array_equal(np_(isd), o_is)
    o_bins = oracle.tile_bin_edges(o_ks, tb[0] * tb[1])
    assert np.array_equal(np_(bins), o_bins)
    # depth-first binning used by the renderer gives the same ids and bins
    assert np.array_equal(np_(pb["ids"]), o_is)
    assert np.array_equal(np_(pb["bins"]), o_bins)
    return m


@pytest.mark.parametrize("bw", [16, 8, 5])
def test_binning_bit_exact(oracle, bw):
    s = scene(1, num_points=3000, seed=40 + bw, img_w=200, img_h=136)
    assert _binning_case(oracle, s, bw) > 0


def test_binning_zero_slots_and_ties(oracle):
    # clipped Gaussians with centres inside the image (counted, never written)
    # and many equal depths (ties broken by Gaussian id)
    s = scene(1, num_points=2000, seed=9)
    s.means[:50, 2] = -1.5   # behind the camera
    s.means[:50, :2] *= -0.1
    s.means[100:600, 2] = 3.0  # identical depth bits
    _binning_case(oracle, s, 16)


def test_binning_1080p_200k(oracle):
    # full-resolution 1080p with 200k Gaussians: exact against the oracle
    s = scene(2, num_points=200_000)
    assert _binning_case(oracle, s, 16) > 1_000_000


def test_binning_unpacked_keys(oracle):
    # 1080p at 8-pixel tiles (32400 tiles, 15 key bits) with 300k Gaussians
    # (19 id bits): tile and id do not fit one 32-bit key, so the list is
    # sorted as (tile, id) pairs, over two passes of 8 and 7 bits
    s = scene(2, num_points=300_000, seed=21)
    assert _binning_case(oracle, s, 8) > 1_000_000


@pytest.mark.parametrize("num_points", [5000, 70_000])
def test_binning_wide_tile_rows(oracle, num_points):
    # an 8200 x 24 image at 2-pixel tiles: 4100 tiles per row, past the
    # 4095-tile rows whose rectangles the fused emission stages as one word,
    # so the two-word staging runs (packed keys with 5000 Gaussians, (tile, id)
    # pairs with 70000)
    s = scene(1, num_points=num_points, seed=31, img_w=8200, img_h=24)
    assert _binning_case(oracle, s, 2) > 0


@pytest.mark.parametrize("bw,num_points", [(16, 3000), (8, 300_000)])
def test_binning_two_phase_equals_one_call(bw, num_points):
    # gstex_bin_gaussians_depth + gstex_bin_gaussians_tiles (the exact-size
    # step's path) give gstex_bin_gaussians' ids and bins, with the total
    # passed on the device or taken from num_intersects; packed keys at bw 16,
    # (tile, id) pairs at 1080p with 8-pixel tiles
    from gstex_cuda_amd import _lib

    s = scene(2 if num_points > 100_000 else 1, num_points=num_points, seed=5 + bw)
    pb = project_and_bin(s, bw)
    n, m = s.num_points, pb["num_intersects"]
    tx, ty, _ = tile_bounds(s.img_w, s.img_h, bw)
    lib = _lib.lib()
    dws = torch.empty(lib.gstex_bin_gaussians_depth_workspace_size(n), dtype=torch.uint8,
                      device=DEV)
    tws = torch.empty(lib.gstex_bin_gaussians_tiles_workspace_size(n, m), dtype=torch.uint8,
                      device=DEV)
    m_dev = torch.tensor([m], dtype=torch.int32, device=DEV)
    st = torch.cuda.current_stream().cuda_stream
    for dev_total in (m_dev, None):
        ids = torch.full((m,), -1, dtype=torch.int32, device=DEV)
        bins = torch.full((tx * ty, 2), -5, dtype=torch.int32, device=DEV)
        assert lib.gstex_bin_gaussians_depth(
            n, pb["centers"].data_ptr(), pb["extents"].data_ptr(), pb["depths"].data_ptr(),
            pb["num_tiles_hit"].data_ptr(), tx, ty, bw, bins.data_ptr(), dws.data_ptr(),
            dws.numel(), st) == 0
        assert lib.gstex_bin_gaussians_tiles(
            n, m, None if dev_total is None else dev_total.data_ptr(), tx, ty, dws.data_ptr(),
            dws.numel(), ids.data_ptr(), bins.data_ptr(), tws.data_ptr(), tws.numel(), st) == 0
        torch.cuda.synchronize()
        assert np.array_equal(np_(ids), np_(pb["ids"]))
        assert np.array_equal(np_(bins), np_(pb["bins"]))
    # a workspace below the queried size is refused
    assert lib.gstex_bin_gaussians_tiles(n, m, None, tx, ty, dws.data_ptr(), dws.numel(),
                                         ids.data_ptr(), bins.data_ptr(), tws.data_ptr(),
                                         tws.numel() - 1, st) != 0


def test_count_intersects(oracle):
    # gstex_count_intersects: the intersect total of compute_cumulative_intersects
    # at full C2 size, and for no Gaussians
    from gstex_cuda_amd import cuda as _C

    s = scene(2)
    pb = project_and_bin(s)
    c, e = oracle.get_aabb_2d(s.means, s.scales, 1.0, s.quats, s.viewmat, s.fx, s.fy, s.cx, s.cy)
    want = int(oracle.num_tiles_hit(c, e, s.img_h, s.img_w, 16).astype(np.int64).sum())
    assert _C.count_intersects(pb["num_tiles_hit"]) == want == pb["num_intersects"]
    assert _C.count_intersects(torch.zeros((0,), dtype=torch.int32, device=DEV)) == 0
    # the exact-size step takes the total from gstex_project_gaussians_total
    from gstex_cuda_amd.pipeline import scene_to_step

    step = scene_to_step(s, DEV)
    step.forward_pre()
    torch.cuda.synchronize()
    assert step.num_intersects == want
    assert np.array_equal(np_(step.tiles_hit), np_(pb["num_tiles_hit"]))


@pytest.mark.parametrize("view", [0, 3, 6, 13])
def test_projection_bit_exact_rotated_cameras(oracle, view):
    # orbit cameras with general rotations (products and sums that a fused
    # multiply-add would round differently): AABBs, depths, tile counts and the
    # binning equal the oracle's bit for bit, Gaussians behind the camera included
    from gstex_cuda_amd import synthetic

    s = scene(4, num_points=200_000)
    s = synthetic.view_scene(s, synthetic.make_orbit_views(16, radius=8.0, seed=4)[view])
    pb = project_and_bin(s)
    c, e = oracle.get_aabb_2d(s.means, s.scales, 1.0, s.quats, s.viewmat, s.fx, s.fy, s.cx, s.cy)
    assert np.array_equal(np_(pb["centers"]), c)
    assert np.array_equal(np_(pb["extents"]), e)
    assert np.array_equal(np_(pb["depths"]), view_depths(s.means, s.viewmat))
    nt = oracle.num_tiles_hit(c, e, s.img_h, s.img_w, 16)
    assert np.array_equal(np_(pb["num_tiles_hit"]), nt)
    log(f"view {view} {s.img_w}x{s.img_h} I={int(nt.sum())} behind="
        f"{int((np_(pb['depths']) < 0).sum())}")
    _binning_case(oracle, s, 16)


def test_binning_screen_sized_splats(oracle):
    # the twelve nearest splats scaled up to 5K-8K tiles each (nine cover all
    # 8160 tiles of the 1080p screen): one Gaussian's entries span several
    # emission blocks and the first blocks hold only a few Gaussians
    s = scene(2, num_points=20_000, seed=77)
    depth = (s.means @ s.viewmat[:3, :3].T + s.viewmat[:3, 3])[:, 2]
    s.scales[np.argsort(depth)[:12]] *= 60.0
    assert _binning_case(oracle, s, 16) > 500_000


@pytest.mark.parametrize("counts", ["wrapped", "clamped"])
def test_binning_wrapped(oracle, counts):
    # forward.cu:34-62: Gaussians spread past every image edge; wrapped counts
    # fill each wrapped rectangle, clamped counts cut it to the window
    from gstex_cuda_amd import cuda as _C
    from gstex_cuda_amd.utils import bin_and_sort_gaussians, compute_cumulative_intersects

    s = scene(1, num_points=3000, seed=61, img_w=200, img_h=136)
    s.means[:, :2] *= 1.6
    bw = 16
    pb = project_and_bin(s, bw)
    c, e, d = pb["centers"], pb["extents"], pb["depths"]
    tb = tile_bounds(s.img_w, s.img_h, bw)
    nt = _C.num_tiles_hit(c, e, s.img_h, s.img_w, bw, wrapped=counts == "wrapped")
    o_nt = oracle.num_tiles_hit(np_(c), np_(e), s.img_h, s.img_w, bw, wrapped=counts == "wrapped")
    assert np.array_equal(np_(nt), o_nt)
    m, cum = compute_cumulative_intersects(nt)
    log(f"wrapped binning ({counts} counts): I = {m}")
    keys, ids, ks, isd, bins = bin_and_sort_gaussians(s.num_points, m, c, e, d, cum, tb, bw,
                                                      wrapped=True)
    o_keys, o_ids = oracle.map_gaussian_to_intersects(np_(c), np_(e), np_(d), np_(cum), tb, bw,
                                                      wrapped=True)
    assert np.array_equal(np_(keys), o_keys)
    assert np.array_equal(np_(ids), o_ids)
    o_ks, o_is = oracle.sort_pairs(o_keys, o_ids)
    assert np.array_equal(np_(ks), o_ks)
    assert np.array_equal(np_(isd), o_is)
    assert np.array_equal(np_(bins), oracle.tile_bin_edges(o_ks, tb[0] * tb[1]))
    # Gaussians crossing an edge really do wrap: some tiles get entries from
    # the far side of the image
    assert (np_(nt) > oracle.num_tiles_hit(np_(c), np_(e), s.img_h, s.img_w, bw)).any() == (
        counts == "wrapped")


def test_sort_full_64bit_signed():
    from gstex_cuda_amd import cuda as _C

    rng = np.random.default_rng(3)
    keys = rng.integers(-(1 << 62), 1 << 62, 100_000, dtype=np.int64)
    keys[:1000] = keys[1000:2000]  # duplicates
    ids = np.arange(100_000, dtype=np.int32)
    ks, is_ = _C.sort_intersects(to_dev(keys), to_dev(ids), end_bit=64)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(np_(ks), keys[order])
    assert np.array_equal(np_(is_), ids[order])


@pytest.mark.parametrize("n,end_bit,kind", [
    (1, 32, "random"), (777, 7, "random"), (300_000, 32, "random"), (300_000, 13, "few"),
    (300_000, 25, "high_bits"), (5_000_000, 14, "few"), (5_000_000, 32, "random"),
    (4_194_303, 32, "sorted"), (4_194_304, 9, "reversed"), (65_537, 0, "random")])
def test_sort_pairs_u32_stable(n, end_bit, kind):
    # the one-sweep sort: tiles of 2048 (n < 4M) and 4096 items, digit widths
    # 7 to 8, one to four passes, duplicate-heavy keys (stability), keys with
    # bits above end_bit (carried, not compared) and presorted inputs
    from gstex_cuda_amd import cuda as _C

    rng = np.random.default_rng(n + end_bit)
    hi = 1 << end_bit
    if kind == "few":
        keys = rng.integers(0, min(hi, 40), n, dtype=np.uint64).astype(np.uint32)
    elif kind == "sorted":
        keys = np.sort(rng.integers(0, hi, n, dtype=np.uint64)).astype(np.uint32)
    elif kind == "reversed":
        keys = np.sort(rng.integers(0, hi, n, dtype=np.uint64))[::-1].astype(np.uint32)
    else:
        keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    if kind == "high_bits":
        keys |= rng.integers(0, 1 << (32 - end_bit), n, dtype=np.uint64).astype(np.uint32) << end_bit
    vals = rng.permutation(n).astype(np.uint32)
    log(f"n={n} end_bit={end_bit} {kind}")
    ks, vs = _C.sort_pairs_u32(to_dev(keys.view(np.int32)), to_dev(vals.view(np.int32)),
                               end_bit=end_bit)
    order = np.argsort(keys & np.uint32(hi - 1 if end_bit < 32 else 0xFFFFFFFF), kind="stable")
    if end_bit == 0:
        order = np.arange(n)
    assert np.array_equal(np_(ks).view(np.uint32), keys[order])
    assert np.array_equal(np_(vs).view(np.uint32), vals[order])


def test_empty_image_and_no_intersects(oracle):
    from gstex_cuda_amd.texture import texture_gaussians

    s = scene(1, num_points=100, seed=3)
    s.means[:, 2] = -5.0  # everything behind the camera, outside the image
    s.means[:, :2] = 50.0
    pb = project_and_bin(s)
    assert pb["num_intersects"] == 0
    r = {k: v.clone().requires_grad_(k in ("colors", "means")) for k, v in
         raster_inputs(s).items()}
    out = texture_gaussians((100, 1, 3), r["texture_dims"], pb["centers"], pb["extents"],
                            pb["depths"], pb["num_tiles_hit"], r["colors"], r["opacities"],
                            r["means"], r["scales"], 1.0, r["quats"], r["uv0"], r["umap"],
                            r["vmap"], r["texture"], r["viewmat"], r["c2w"], s.fx, s.fy, s.cx,
                            s.cy, s.img_h, s.img_w, 16, s.settings, r["background"])
    img = np_(out[0])
    assert np.allclose(img, s.background)
    (out[0].sum() + out[4].sum()).backward()
    assert torch.count_nonzero(r["colors"].grad) == 0


# ------------------------------------------------------------------ SH
@pytest.mark.parametrize("degree,used", [(0, 0), (1, 1), (2, 1), (3, 3), (4, 4), (4, 2)])
def test_sh_matches_oracle(oracle, degree, used):
    from gstex_cuda_amd import cuda as _C

    rng = np.random.default_rng(degree * 7 + used)
    n, nb = 5000, (degree + 1) ** 2
    dirs = rng.standard_normal((n, 3)).astype(np.float32)
    co = (0.3 * rng.standard_normal((n, nb, 3))).astype(np.float32)
    vc = rng.standard_normal((n, 3)).astype(np.float32)
    col = _C.compute_sh_forward(n, degree, used, to_dev(dirs), to_dev(co))
    np.testing.assert_allclose(np_(col), oracle.sh_forward(degree, used, dirs, co, "f64"), atol=1e-5)
    vco = _C.compute_sh_backward(n, degree, used, to_dev(dirs), to_dev(vc))
    np.testing.assert_allclose(np_(vco), oracle.sh_backward(degree, used, dirs, vc, "f64"), atol=1e-5)


@pytest.mark.parametrize("degree,used,bf16", [(3, 3, False), (3, 3, True), (4, 2, False),
                                               (1, 1, True), (0, 0, False)])
def test_sh_forward_means_equals_directions_first(degree, used, bf16):
    # gstex_compute_sh_forward_means: the colours of compute_sh_forward at
    # means - c2w[:3, 3], bit for bit, float32 and bfloat16 coefficients
    from gstex_cuda_amd import _lib

    rng = np.random.default_rng(degree * 11 + used)
    n, nb = 4099, (degree + 1) ** 2
    means = to_dev(rng.standard_normal((n, 3)).astype(np.float32) * 3.0)
    c2w = to_dev(np.array([[0.8, 0.0, 0.6, 0.3], [0.0, 1.0, 0.0, -1.2], [-0.6, 0.0, 0.8, 4.5],
                           [0.0, 0.0, 0.0, 1.0]], np.float32))
    co = to_dev((0.3 * rng.standard_normal((n, nb, 3))).astype(np.float32))
    if bf16:
        co = co.to(torch.bfloat16)
    lib = _lib.lib()
    st = torch.cuda.current_stream().cuda_stream
    got = torch.full((n, 3), 7.0, device=DEV)
    assert lib.gstex_compute_sh_forward_means(n, degree, used, means.data_ptr(), c2w.data_ptr(),
                                              co.data_ptr(), int(bf16), got.data_ptr(), st) == 0
    dirs = (means - c2w[:3, 3]).contiguous()
    want = torch.full((n, 3), -7.0, device=DEV)
    fn = lib.gstex_compute_sh_forward_bf16 if bf16 else lib.gstex_compute_sh_forward
    assert fn(n, degree, used, dirs.data_ptr(), co.data_ptr(), want.data_ptr(), st) == 0
    torch.cuda.synchronize()
    assert torch.equal(got, want)


@pytest.mark.parametrize("degree,used,views", [(3, 3, 5), (4, 2, 3), (0, 0, 2), (2, 2, 1)])
def test_sh_backward_views_sums_single_views(oracle, degree, used, views):
    # gstex_compute_sh_backward_views equals the per-view adjoints added in
    # view order (bit for bit) and the float64 oracle's sum (within 1e-5)
    from gstex_cuda_amd import cuda as _C
    from gstex_cuda_amd._lib import call

    rng = np.random.default_rng(degree * 11 + views)
    n, nb = 4099, (degree + 1) ** 2
    dirs = rng.standard_normal((views, n, 3)).astype(np.float32)
    vc = rng.standard_normal((views, n, 3)).astype(np.float32)
    out = torch.full((n, nb, 3), float("nan"), device=DEV)
    d_dirs, d_vc = to_dev(dirs), to_dev(vc)  # held for the duration of the call
    call("gstex_compute_sh_backward_views", n, views, degree, used, d_dirs.data_ptr(),
         d_vc.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    acc = None
    for v in range(views):
        one = _C.compute_sh_backward(n, degree, used, to_dev(dirs[v]), to_dev(vc[v]))
        acc = one.clone() if acc is None else acc.add_(one)
    assert torch.equal(out, acc)
    want = sum(oracle.sh_backward(degree, used, dirs[v], vc[v], "f64") for v in range(views))
    np.testing.assert_allclose(np_(out), want, atol=1e-5 * views)


def test_finalize_views_sums_single_views():
    # gstex_finalize_gradients_views equals the per-view gstex_finalize_gradients
    # outputs added in view order, bit for bit, blur (camera-dependent) slots
    # included, and hands back each view's colour gradient
    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd._lib import call, lib

    s = synthetic.make_scene(1, seed=5, num_points=3001, img_w=160, img_h=120)
    views = synthetic.make_views(s, 3, seed=9)
    n, V = s.num_points, len(views)
    r = raster_inputs(s)
    st = torch.cuda.current_stream().cuda_stream
    size = lib().gstex_splat_records_size(n)
    stride = (size + 255) // 256 * 256
    recs = torch.zeros((V * stride,), dtype=torch.uint8, device=DEV)
    rng = np.random.default_rng(2)
    grad = to_dev(rng.standard_normal((V, n, 32)).astype(np.float32))
    grad[:, : n // 2, 30:32] = 0.0  # half the Gaussians without the blur term
    vms = torch.stack([to_dev(v.viewmat) for v in views]).contiguous()
    c2ws = torch.stack([to_dev(v.c2w) for v in views]).contiguous()
    foc = to_dev(np.array([[v.fx, v.fy] for v in views], np.float32))
    shapes = dict(v_colors=(n, 3), v_opacity=(n, 1), v_means=(n, 3), v_scales=(n, 3),
                  v_quats=(n, 4), v_uv0=(n, 1, 2), v_umap=(n, 1, 3), v_vmap=(n, 1, 3))
    acc = None
    for i, v in enumerate(views):
        base = recs.data_ptr() + i * stride
        call("gstex_build_splat_records", n, r["means"].data_ptr(), r["scales"].data_ptr(), 1.0,
             r["quats"].data_ptr(), r["colors"].data_ptr(), r["opacities"].data_ptr(),
             r["uv0"].data_ptr(), r["umap"].data_ptr(), r["vmap"].data_ptr(),
             r["texture_dims"].data_ptr(), vms[i].data_ptr(), c2ws[i].data_ptr(), v.fx, v.fy,
             v.cx, v.cy, base, st)
        one = {k: torch.empty(sh, device=DEV) for k, sh in shapes.items()}
        call("gstex_finalize_gradients", n, grad[i].data_ptr(), base, r["quats"].data_ptr(),
             r["scales"].data_ptr(), vms[i].data_ptr(), c2ws[i].data_ptr(), v.fx, v.fy,
             *[one[k].data_ptr() for k in shapes], st)
        acc = one if acc is None else {k: acc[k].add_(one[k]) for k in shapes}
        if i == 0:
            first_colors = one["v_colors"].clone()
    out = {k: torch.full(sh, float("nan"), device=DEV) for k, sh in shapes.items()}
    cviews = torch.full((V, n, 3), float("nan"), device=DEV)
    call("gstex_finalize_gradients_views", n, V, grad.data_ptr(), recs.data_ptr(), stride,
         r["quats"].data_ptr(), r["scales"].data_ptr(), vms.data_ptr(), c2ws.data_ptr(),
         foc.data_ptr(), *[out[k].data_ptr() for k in shapes], cviews.data_ptr(), st)
    torch.cuda.synchronize()
    for k in shapes:
        assert torch.equal(out[k], acc[k]), k
    assert torch.equal(cviews[0], first_colors)
    assert torch.equal(cviews.sum(0), cviews[0] + cviews[1] + cviews[2])


def test_sh_adam_converges_like_reference_test():
    # tests/test_sh.py of the reference: degree 4, one point, 1000 Adam steps to colour 0.5
    from gstex_cuda_amd import sh

    g = torch.Generator(device="cpu").manual_seed(0)
    vd = torch.randn(1, 3, generator=g)
    vd = (vd / vd.norm(dim=-1, keepdim=True)).to(DEV)
    coeffs = torch.rand(1, 25, 3, generator=g).to(DEV).requires_grad_(True)
    opt = torch.optim.Adam([coeffs], lr=1e-2)
    gt = torch.full((1, 3), 0.5, device=DEV)
    for _ in range(1000):
        opt.zero_grad()
        col = sh.spherical_harmonics(4, vd, coeffs)
        loss = ((col - gt) ** 2).mean()
        loss.backward()
        opt.step()
    torch.testing.assert_close(sh.spherical_harmonics(4, vd, coeffs), gt)


# ------------------------------------------------------------------ sampling
def test_texture_sample_reference_scenario(oracle):
    # tests/test_sample.py of the reference, seeded: 5 charts of 2..7 x 2..8 texels,
    # 10 channels, 100 queries in [0, 1)
    from gstex_cuda_amd.texture_sample import texture_sample

    rng = np.random.default_rng(1)
    sz = 5
    h = rng.integers(0, 6, sz) + 2
    w = rng.integers(0, 7, sz) + 2
    start = np.concatenate([[0], np.cumsum(h * w)[:-1]])
    dims = np.stack([h, w, start], 1).astype(np.int32)
    tex = rng.uniform(0, 1, (int((h * w).sum()), 10)).astype(np.float32)
    uvs = rng.uniform(0, 1, (100, 2)).astype(np.float32)
    qd = dims[rng.integers(0, sz, 100)]
    t = to_dev(tex).requires_grad_(True)
    out = texture_sample((sz, 1, 10), to_dev(qd), t, to_dev(uvs))
    np.testing.assert_allclose(np_(out), oracle.texture_sample(qd, uvs, tex, "f64"), atol=1e-5)
    v = rng.standard_normal((100, 10)).astype(np.float32)
    out.backward(to_dev(v))
    np.testing.assert_allclose(np_(t.grad), oracle.texture_sample_vjp(qd, uvs, v, tex.shape[0], "f64"),
                               atol=1e-5)


def test_texture_sample_adjoint_one_texel_charts(oracle):
    # charts of 1x1, 1x2, 2x1 and 2x2 texels: most taps clamp onto a shared
    # texel and 4096 queries pile onto 8 charts
    from gstex_cuda_amd import cuda as _C

    rng = np.random.default_rng(7)
    h = np.array([1, 1, 2, 2, 1, 1, 2, 1])
    w = np.array([1, 2, 1, 2, 1, 1, 2, 1])
    start = np.concatenate([[0], np.cumsum(h * w)[:-1]])
    dims = np.stack([h, w, start], 1).astype(np.int32)
    X = int((h * w).sum())
    tex = rng.uniform(0, 1, (X, 3)).astype(np.float32)
    qd = dims[rng.integers(0, len(h), 4096)]
    uvs = rng.uniform(0, 1, (4096, 2)).astype(np.float32)
    v = rng.standard_normal((4096, 3)).astype(np.float32)
    got = _C.texture_sample_backward((len(h), 1, 3), to_dev(qd), to_dev(uvs), to_dev(tex),
                                     to_dev(v))
    want = oracle.texture_sample_vjp(qd, uvs, v, X, "f64")
    # sums of about 500 terms of size 1 in float32, in atomic order
    np.testing.assert_allclose(np_(got), want, atol=2e-4)
    fwd = _C.texture_sample_forward((len(h), 1, 3), to_dev(qd), to_dev(uvs), to_dev(tex))
    np.testing.assert_allclose(np_(fwd), oracle.texture_sample(qd, uvs, tex, "f64"), atol=1e-5)


@pytest.mark.parametrize("C,max_side,n_charts,q", [(3, 16, 20000, 300000), (40, 16, 3000, 60000),
                                                   (1, 4, 50, 70000)])
def test_texture_sample_adjoint_sorted_windows(oracle, C, max_side, n_charts, q):
    # the adjoint sorts the queries by chart and accumulates each workgroup's
    # run of charts in an LDS window: many workgroups, charts of 1x1 up to
    # 16x16, channel counts whose charts overflow the window (40 x 256 values)
    # and heavily shared charts (50 charts, 70k queries), uv outside [0, 1]
    from gstex_cuda_amd import cuda as _C

    rng = np.random.default_rng(C)
    h = rng.integers(1, max_side + 1, n_charts)
    w = rng.integers(1, max_side + 1, n_charts)
    start = np.concatenate([[0], np.cumsum(h * w)[:-1]])
    dims = np.stack([h, w, start], 1).astype(np.int32)
    X = int((h * w).sum())
    tex = rng.uniform(0, 1, (X, C)).astype(np.float32)
    qd = dims[rng.integers(0, n_charts, q)]
    uvs = rng.uniform(-0.2, 1.2, (q, 2)).astype(np.float32)
    v = rng.standard_normal((q, C)).astype(np.float32)
    got = np_(_C.texture_sample_backward((n_charts, 1, C), to_dev(qd), to_dev(uvs), to_dev(tex),
                                         to_dev(v)))
    want = oracle.texture_sample_vjp(qd, uvs, v, X, "f64")
    scale = max(float(np.abs(want).max()), 1.0)
    log(f"C={C} X={X} q={q}: max err {np.abs(got - want).max():.2e} (scale {scale:.1f})")
    np.testing.assert_allclose(got, want, atol=2e-6 * scale * (q / n_charts) ** 0.5 + 1e-5)
    assert np.isfinite(got).all()


def test_texture_sample_affine_kat():
    from gstex_cuda_amd import cuda as _C

    bl = KAT["bilinear"]
    tex = np.array(bl["texture"], np.float32)
    dims = np.array([q["dims"] for q in bl["queries"]], np.int32)
    uvs = np.array([q["uv"] for q in bl["queries"]], np.float32)
    out = _C.texture_sample_forward((4, 1, 3), to_dev(dims), to_dev(uvs), to_dev(tex))
    np.testing.assert_allclose(np_(out), [q["value"] for q in bl["queries"]], atol=2e-5)


# ------------------------------------------------------------------ raster forward
def _check_forward(g, o, label, min_psnr=60.0):
    img_err = np.abs(np_(g["out_img"]) - o["out_img"])
    p = psnr(np_(g["out_img"]), o["out_img"])
    fc = frac_close(np_(g["out_img"]), o["out_img"], 1e-4)
    fi = frac_close(np_(g["final_idx"]), o["final_idx"], 0)
    log(label, "psnr %.1f" % p, "frac1e-4 %.5f" % fc, "maxerr %.2e" % img_err.max(), "idx %.5f" % fi)
    assert p >= min_psnr and fc >= 0.999 and fi >= 0.999
    for k in ("out_texture", "out_normal", "out_depth", "out_reg", "final_Ts"):
        f = frac_close(np_(g[k]), o[k], 1e-4, 1e-4)
        assert f >= 0.999, (label, k, f)
    assert frac_close(np_(g["depth_idx"]), o["depth_idx"], 0) >= 0.999


@pytest.mark.parametrize("case", ["c1", "nonsquare_bw16", "bw8", "bw5", "nearest", "blur",
                                  "ndc_noflip", "viz_normal_flip"])
def test_forward_matches_oracle(oracle, case):
    bw, settings = 16, 1 << 8
    if case == "c1":
        s = scene(1)
    else:
        s = scene(1, num_points=1500, seed=sum(map(ord, case)), img_w=200, img_h=136)
    if case == "bw8":
        bw = 8
    if case == "bw5":
        bw = 5
    if case == "nearest":
        settings |= 1 << 2
    if case == "blur":
        settings |= 1 << 9
    if case == "ndc_noflip":
        settings = 1 << 10
    if case == "viz_normal_flip":
        settings |= 1 << 15
    pb = project_and_bin(s, bw)
    g, _ = gpu_forward(s, pb["ids"], pb["bins"], settings=settings, bw=bw)
    o = oracle.texture_forward(oracle_with(oracle, s, pb["ids"], pb["bins"], settings=settings,
                                           bw=bw), "f64")
    _check_forward(g, o, case)


@pytest.mark.parametrize("channels", [1, 2, 4, 7, 16, 33])
def test_forward_channel_counts(oracle, channels):
    s = scene(1, num_points=800, seed=channels, img_w=96, img_h=80)
    rng = np.random.default_rng(channels)
    s.texture = rng.uniform(0, 1, (s.texture.shape[0], channels)).astype(np.float32)
    pb = project_and_bin(s)
    g, _ = gpu_forward(s, pb["ids"], pb["bins"])
    o = oracle.texture_forward(oracle_with(oracle, s, pb["ids"], pb["bins"]), "f64")
    _check_forward(g, o, "C=%d" % channels)


def test_forward_closed_form_single_splat():
    k = KAT["splat1"]
    s = camera_facing_scene([k["mean"]], [k["scale"]], [k["opacity"]], [k["color"]], k["W"],
                            k["H"], k["f"], k["cx"], k["cy"], texel=k["texel"],
                            background=k["background"])
    pb = project_and_bin(s)
    g, _ = gpu_forward(s, pb["ids"], pb["bins"])
    np.testing.assert_allclose(np_(g["out_img"]), np.array(k["img"]), atol=2e-6)
    np.testing.assert_allclose(1 - np_(g["final_Ts"]), np.array(k["alpha"]), atol=2e-6)


@pytest.mark.parametrize("case", [0, 1])
def test_forward_closed_form_textured_splat(case):
    # the splat_tex known answer (closed-form plane hits over an affine 4 x 5
    # chart with u and v mixing x and y, clamped at both ends)
    from helpers import textured_splat_scene

    k = KAT["splat_tex"][case]
    s = textured_splat_scene(k)
    pb = project_and_bin(s)
    g, _ = gpu_forward(s, pb["ids"], pb["bins"])
    chk = np.array(k["check"], bool)
    np.testing.assert_allclose(np_(g["out_texture"])[chk], np.array(k["out_texture"])[chk],
                               atol=2e-5)
    np.testing.assert_allclose(np_(g["out_img"])[chk], np.array(k["img"])[chk], atol=2e-5)


def test_forward_invariants_1080p_500k(oracle):
    # 1080p, 500k Gaussians: sum of weights + T = 1 everywhere, and a sample of
    # tiles rendered exactly by the oracle
    s = scene(2)
    s.colors[:] = 1.0
    pb = project_and_bin(s)
    log("1080p intersects", pb["num_intersects"])
    g, _ = gpu_forward(s, pb["ids"], pb["bins"])
    T = np_(g["final_Ts"])
    np.testing.assert_allclose(np_(g["out_img"]), np.repeat((1 - T)[..., None], 3, -1), atol=2e-5)
    assert np.all((T > 1e-4) & (T <= 1.0))
    assert np.all(np_(g["out_reg"]) >= -1e-6)
    tiles = tile_bounds(s.img_w, s.img_h, 16)
    nt = tiles[0] * tiles[1]
    counts = np.diff(np_(pb["bins"]), axis=1)[:, 0]
    rng = np.random.default_rng(0)
    pick = np.concatenate([rng.choice(nt, 12, replace=False), np.argsort(counts)[-4:]])
    mask = np.zeros(nt, np.uint8)
    mask[pick] = 1
    sc = oracle_with(oracle, s, pb["ids"], pb["bins"])
    sc.tile_mask = mask
    o = oracle.texture_forward(sc, "f64")
    sel = np.zeros((s.img_h, s.img_w), bool)
    for t in pick:
        ty, tx = divmod(int(t), tiles[0])
        sel[ty * 16:(ty + 1) * 16, tx * 16:(tx + 1) * 16] = True
    fc = frac_close(np_(g["out_texture"])[sel], o["out_texture"][sel], 1e-4)
    log("1080p tile sample frac", fc, "busiest tile entries", int(counts.max()))
    assert fc >= 0.999


# ------------------------------------------------------------------ raster backward
@pytest.mark.parametrize("case", ["c1", "bw8", "no_uv_grad", "ndc", "blur", "nearest", "C7"])
def test_backward_matches_oracle(oracle, case):
    bw, settings = 16, 1 << 8
    s = scene(1, num_points=600 if case != "c1" else 1000, seed=len(case) * 13,
              img_w=128 if case != "c1" else 256, img_h=96 if case != "c1" else 256)
    if case == "bw8":
        bw = 8
    if case == "no_uv_grad":
        settings = 0
    if case == "ndc":
        settings |= 1 << 10
    if case == "blur":
        settings |= 1 << 9
    if case == "nearest":
        settings |= 1 << 2
    if case == "C7":
        s.texture = np.random.default_rng(7).uniform(0, 1, (s.texture.shape[0], 7)).astype(np.float32)
    pb = project_and_bin(s, bw)
    g, _ = gpu_forward(s, pb["ids"], pb["bins"], settings=settings, bw=bw)
    sc = oracle_with(oracle, s, pb["ids"], pb["bins"], settings=settings, bw=bw)
    o = oracle.texture_forward(sc, "f64")
    rng = np.random.default_rng(5)
    grads = random_grads(rng, s.img_h, s.img_w, s.texture.shape[1], 1e-3)
    gg = gpu_backward(s, pb["ids"], pb["bins"], g, grads, settings=settings, bw=bw)
    og = oracle.texture_backward(sc, o, grads["img"], grads["depth"], grads["reg"], grads["alpha"],
                                 grads["tex"], grads["normal"], "f64")
    bad = []
    for k in og:
        ok, l2, frac = grad_close(np_(gg[k]), og[k])
        log(case, k, "relL2 %.2e" % l2, "frac %.4f" % frac)
        if not ok:
            bad.append((k, l2, frac))
    assert not bad, bad


@pytest.mark.parametrize("side,channels", [(1, 3), (1, 5), (16, 7), (16, 3), (0, 3)])
def test_backward_texel_gradient_routes(oracle, side, channels):
    """Texel gradients leave the backward by three routes chosen per Gaussian:
    1x1 charts with <= 4 channels ride the wave reduction, charts of up to 1024
    values go through the LDS accumulator, larger ones atomically per tap.
    side 0 = mixed 1..16 grids sized by projected area."""
    from gstex_cuda_amd.synthetic import SceneSpec, make_scene

    spec = SceneSpec(num_points=700, img_w=128, img_h=96, focal=96.0, z_range=(2.0, 8.0),
                     texel_side=max(side, 1), texel_budget=60_000 if side == 0 else 0,
                     channels=channels)
    s = make_scene(seed=side * 10 + channels, spec=spec)
    pb = project_and_bin(s)
    g, _ = gpu_forward(s, pb["ids"], pb["bins"])
    sc = oracle_with(oracle, s, pb["ids"], pb["bins"])
    o = oracle.texture_forward(sc, "f64")
    grads = random_grads(np.random.default_rng(side), s.img_h, s.img_w, channels, 1e-3)
    gg = gpu_backward(s, pb["ids"], pb["bins"], g, grads)
    og = oracle.texture_backward(sc, o, grads["img"], grads["depth"], grads["reg"], grads["alpha"],
                                 grads["tex"], grads["normal"], "f64")
    for k in ("v_texture", "v_uv0", "v_umap", "v_means", "v_colors"):
        ok, l2, frac = grad_close(np_(gg[k]), og[k])
        log("texel route", side, channels, k, "relL2 %.2e" % l2, "frac %.4f" % frac)
        assert ok, (k, l2, frac)


def test_autograd_end_to_end(oracle):
    from gstex_cuda_amd.texture import texture_gaussians

    s = scene(1, num_points=700, seed=77, img_w=160, img_h=112)
    pb = project_and_bin(s)
    r = raster_inputs(s)
    leaves = {k: r[k].clone().requires_grad_(True) for k in
              ("colors", "opacities", "means", "scales", "quats", "uv0", "umap", "vmap", "texture")}
    r.update(leaves)
    out = texture_gaussians(s.texture_info, r["texture_dims"], pb["centers"], pb["extents"],
                            pb["depths"], pb["num_tiles_hit"], r["colors"], r["opacities"],
                            r["means"], r["scales"], 1.0, r["quats"], r["uv0"], r["umap"],
                            r["vmap"], r["texture"], r["viewmat"], r["c2w"], s.fx, s.fy, s.cx,
                            s.cy, s.img_h, s.img_w, 16, 1 << 8, r["background"])
    target = to_dev(s.target)
    loss = ((out[0] + out[4] - target) ** 2).mean() + 100 * out[2].mean() + 0.05 * (
        out[5][..., 0] ** 2 + out[5][..., 1] ** 2 + (1 - out[5][..., 2]) ** 2).mean()
    loss.backward()
    sc = oracle_with(oracle, s, pb["ids"], pb["bins"])
    o = oracle.texture_forward(sc, "f64")
    P = s.img_h * s.img_w
    diff = (o["out_img"] + o["out_texture"] - s.target) * 2 / (P * 3)
    vn = np.zeros((s.img_h, s.img_w, 3))
    vn[..., 0] = 0.05 * 2 * o["out_normal"][..., 0] / P
    vn[..., 1] = 0.05 * 2 * o["out_normal"][..., 1] / P
    vn[..., 2] = -0.05 * 2 * (1 - o["out_normal"][..., 2]) / P
    z = np.zeros((s.img_h, s.img_w))
    og = oracle.texture_backward(sc, o, diff, z, np.full_like(z, 100.0 / P), z, diff, vn, "f64")
    pairs = dict(colors="v_colors", opacities="v_opacity", means="v_means", scales="v_scales",
                 quats="v_quats", uv0="v_uv0", umap="v_umap", vmap="v_vmap", texture="v_texture")
    for k, gk in pairs.items():
        ok, l2, frac = grad_close(np_(leaves[k].grad), og[gk].reshape(leaves[k].shape))
        log("autograd", k, "relL2 %.2e" % l2, "frac %.4f" % frac)
        assert ok, (k, l2, frac)
