"""Overlay preparation (rgb_to_yuv / prepare_yuv, crates/render/src/project_compositor.rs:1098-1185)
and the 4:2:0 canvas scaler (scale_yuv420_to_canvas_cpu, crates/render/src/compositor.rs:89-305)
of the oracle, pinned by the reference's own known answers (compositor.rs:351-369,
project_compositor.rs:1606-1653), by the BT.709 definition and by an independent float64
restatement of the image crate's resampler. CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

from mmrecode_amd import abi

# ---------------------------------------------------------------- helpers shared with the GPU tests


def scale_planes(rng, width, height, kind="random"):
    """A 4:2:0 frame: random samples, or the reference's constant test frame (200 / 100 / 150)."""
    shapes = [(height, width), (height // 2, width // 2), (height // 2, width // 2)]
    if kind == "constant":
        return [np.full(s, v, np.uint8) for s, v in zip(shapes, (200, 100, 150))]
    if kind == "smooth":   # a gradient with noise: overshoot near the clamps stays rare
        out = []
        for s in shapes:
            yy, xx = np.mgrid[0:s[0], 0:s[1]]
            base = 128 + 100 * np.sin(xx / 7.0) * np.cos(yy / 5.0)
            out.append(np.clip(base + rng.normal(0, 8, s), 0, 255).astype(np.uint8))
        return out
    return [rng.integers(0, 256, s, dtype=np.uint8) for s in shapes]


def scale_job(src, src_strides, width, height, dst, canvas_w, canvas_h, mode, rng_range, filt):
    j = abi.new(abi.ScaleJob, width=width, height=height, canvas_width=canvas_w, canvas_height=canvas_h, mode=mode,
                range=rng_range, filter=filt)
    for k in range(3):
        j.src[k] = src[k]
        j.src_stride[k] = src_strides[k]
        j.dst[k] = dst[k]
        j.dst_stride[k] = canvas_w if k == 0 else canvas_w // 2
    return j


def oracle_scale(oracle, planes, canvas_w, canvas_h, mode, rng_range=abi.RANGE_LIMITED, filt=abi.SCALE_FILTER_LANCZOS3):
    height, width = planes[0].shape
    out = [np.zeros((canvas_h, canvas_w), np.uint8), np.zeros((canvas_h // 2, canvas_w // 2), np.uint8),
           np.zeros((canvas_h // 2, canvas_w // 2), np.uint8)]
    planes = [np.ascontiguousarray(p) for p in planes]
    j = scale_job([p.ctypes.data for p in planes], [p.shape[1] for p in planes], width, height,
                  [o.ctypes.data for o in out], canvas_w, canvas_h, mode, rng_range, filt)
    assert oracle.oracle_scale_yuv420(C.byref(j)) == 0
    return out


OVERLAY_OUTPUTS = ("y_limited", "y_full", "u_limited", "v_limited", "u_full", "v_full", "chroma_alpha")


def overlay_shapes(width, height):
    cw, ch = (width + 1) // 2, (height + 1) // 2
    return {n: ((height, width) if n.startswith("y_") else (ch, cw)) for n in OVERLAY_OUTPUTS}


def oracle_prepare(oracle, rgba):
    height, width = rgba.shape[:2]
    rgba = np.ascontiguousarray(rgba)
    out = {n: np.zeros(s, np.uint8) for n, s in overlay_shapes(width, height).items()}
    j = abi.new(abi.OverlayPrepareJob, rgba=rgba.ctypes.data, width=width, height=height,
                **{n: out[n].ctypes.data for n in OVERLAY_OUTPUTS})
    assert oracle.oracle_overlay_prepare(C.byref(j)) == 0
    return out


def random_overlay(rng, width, height):
    """Straight-alpha RGBA with transparent, opaque and partial cells."""
    rgba = rng.integers(0, 256, (height, width, 4), dtype=np.uint8)
    a = rgba[..., 3]
    a[rng.random((height, width)) < 0.3] = 0
    a[rng.random((height, width)) < 0.3] = 255
    return rgba


# ---------------------------------------------------------------- the reference's known answers


def test_fit_preserves_aspect_and_adds_black_bars(oracle):
    """compositor.rs:351-359: a 4x4 frame fit onto an 8x4 limited-range canvas."""
    out = oracle_scale(oracle, scale_planes(None, 4, 4, "constant"), 8, 4, abi.SCALE_FIT)
    assert list(out[0][0, 0:2]) == [16, 16]
    assert list(out[0][0, 2:6]) == [200] * 4
    assert list(out[0][0, 6:8]) == [16, 16]
    assert out[1][0, 0] == 128


def test_stretch_and_native_have_explicit_canvas_behavior(oracle):
    """compositor.rs:361-369: stretch resamples a constant frame to the same constant; native pads."""
    frame = scale_planes(None, 4, 4, "constant")
    stretched = oracle_scale(oracle, frame, 8, 4, abi.SCALE_STRETCH)
    assert (stretched[0] == 200).all()
    native = oracle_scale(oracle, frame, 8, 4, abi.SCALE_NATIVE)
    assert list(native[0][0, 0:2]) == [16, 16]
    assert list(native[0][0, 2:6]) == [200] * 4


def test_red_overlay_raises_luma_and_cr(oracle):
    """project_compositor.rs:1606-1653: an opaque red overlay composited over limited black raises
    Y above 16 and Cr above 128 (the prepared planes it blends)."""
    rgba = np.zeros((4, 4, 4), np.uint8)
    rgba[..., 0], rgba[..., 3] = 255, 255
    out = oracle_prepare(oracle, rgba)
    assert (out["y_limited"] > 16).all() and (out["v_limited"] > 128).all()
    assert (out["chroma_alpha"] == 255).all()


# ---------------------------------------------------------------- BT.709 (mathematics)


def _bt709(rgb, limited):
    kr, kb = 0.2126, 0.0722
    kg = 1 - kr - kb
    r, g, b = (rgb[..., i].astype(np.float64) for i in range(3))
    y = kr * r + kg * g + kb * b
    u = (b - y) / (2 * (1 - kb))
    v = (r - y) / (2 * (1 - kr))
    if limited:
        return 16 + y * 219 / 255, 128 + u * 224 / 255, 128 + v * 224 / 255
    return y, 128 + u, 128 + v


def test_rgb_to_yuv_is_bt709(oracle):
    """Each of the six outputs stays within 1.5 of the BT.709 definition (Kr 0.2126, Kb 0.0722;
    limited range 219 / 224 steps) over a grid of colours: the 8-bit integer coefficients round the
    exact ones, so a wrong coefficient, sign or channel order is off by tens."""
    vals = np.arange(0, 256, 15, dtype=np.uint8)
    rgb = np.stack(np.meshgrid(vals, vals, vals, indexing="ij"), -1).reshape(-1, 3)
    lim = np.zeros((len(rgb), 3), np.uint8)
    full = np.zeros((len(rgb), 3), np.uint8)
    L, F = (C.c_uint8 * 3)(), (C.c_uint8 * 3)()
    for i, (r, g, b) in enumerate(rgb):
        oracle.oracle_rgb_to_yuv(int(r), int(g), int(b), L, F)
        lim[i], full[i] = list(L), list(F)
    for got, limited, lo, hi in ((lim, True, (16, 16, 16), (235, 240, 240)), (full, False, (0, 0, 0), (255, 255, 255))):
        want = _bt709(rgb, limited)
        for c in range(3):
            err = np.abs(got[:, c] - np.clip(want[c], lo[c], hi[c]))
            assert err.max() <= 1.5, (limited, c, err.max())
    # grey stays on the neutral axis
    for v in (0, 128, 255):
        oracle.oracle_rgb_to_yuv(v, v, v, L, F)
        assert abs(L[1] - 128) <= 1 and abs(L[2] - 128) <= 1 and F[1] == 128 and F[2] == 128


@pytest.mark.parametrize("size", [(8, 6), (7, 5), (1, 1), (2, 3), (33, 17)])
def test_prepare_yuv_cells(oracle, size):
    """Per-pixel luma is rgb_to_yuv's; per 2x2 cell (1 or 2 samples wide / tall at an odd edge) the
    alpha is the rounded mean and U / V the alpha-weighted mean, within half a step; 128 for a
    transparent cell."""
    width, height = size
    rng = np.random.default_rng(width * 100 + height)
    rgba = random_overlay(rng, width, height)
    rgba[0:2, 0:2, 3] = 0                      # one transparent cell
    out = oracle_prepare(oracle, rgba)
    L, F = (C.c_uint8 * 3)(), (C.c_uint8 * 3)()
    conv = np.zeros((height, width, 6), np.float64)
    for y in range(height):
        for x in range(width):
            oracle.oracle_rgb_to_yuv(*(int(v) for v in rgba[y, x, :3]), L, F)
            conv[y, x] = list(L) + list(F)
    assert (out["y_limited"] == conv[..., 0]).all() and (out["y_full"] == conv[..., 3]).all()
    for cy in range((height + 1) // 2):
        for cx in range((width + 1) // 2):
            cell = rgba[2 * cy:2 * cy + 2, 2 * cx:2 * cx + 2]
            a = cell[..., 3].astype(np.float64)
            assert abs(out["chroma_alpha"][cy, cx] - a.mean()) <= 0.5
            for name, ch in (("u_limited", 1), ("v_limited", 2), ("u_full", 4), ("v_full", 5)):
                got = out[name][cy, cx]
                if a.sum() == 0:
                    assert got == 128
                else:
                    want = (conv[2 * cy:2 * cy + 2, 2 * cx:2 * cx + 2, ch] * a).sum() / a.sum()
                    assert abs(got - want) <= 0.5, (name, cy, cx)


# ---------------------------------------------------------------- the resampler


def _taps(oracle, filt, n, m, o):
    w = (C.c_float * 4096)()
    left = C.c_int32()
    count = oracle.oracle_scale_taps(filt, n, m, o, C.byref(left), w, 4096)
    assert count > 0
    return left.value, np.array(w[:count], np.float64)


def test_triangle_taps_closed_form(oracle):
    """FilterType::Triangle halving 4 samples to 2: centre 1.0, support 2, taps 0..2 at distances
    0.25, 0.25, 0.75 (scaled by 1/2) -> weights 3/7, 3/7, 1/7; doubling 2 samples to 4 puts output 1
    at weights 3/4, 1/4."""
    left, w = _taps(oracle, abi.SCALE_FILTER_TRIANGLE, 4, 2, 0)
    assert left == 0 and np.allclose(w, [3 / 7, 3 / 7, 1 / 7], atol=1e-7)
    left, w = _taps(oracle, abi.SCALE_FILTER_TRIANGLE, 2, 4, 1)
    assert left == 0 and np.allclose(w, [0.75, 0.25], atol=1e-7)


def test_triangle_upscale_known_answer(oracle):
    """A [0, 255] row doubled with Triangle is [0, 63.75, 191.25, 255] before rounding."""
    frame = [np.array([[0, 255], [0, 255]], np.uint8), np.full((1, 1), 90, np.uint8), np.full((1, 1), 170, np.uint8)]
    out = oracle_scale(oracle, frame, 4, 2, abi.SCALE_STRETCH, filt=abi.SCALE_FILTER_TRIANGLE)
    assert (out[0] == [[0, 64, 191, 255]] * 2).all()
    assert (out[1] == 90).all() and (out[2] == 170).all()


def _lanczos3_f64(x):
    x = np.asarray(x, np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.where(x == 0, 1.0, np.sin(np.pi * x) / (np.pi * x))
        s3 = np.where(x == 0, 1.0, np.sin(np.pi * x / 3) / (np.pi * x / 3))
    return np.where(np.abs(x) < 3, s * s3, 0.0)


def _taps_f64(filt, n, m, o):
    """The resampler's taps in float64, from the published algorithm."""
    ratio = n / m
    sratio = max(ratio, 1.0)
    support = (1.0 if filt == abi.SCALE_FILTER_TRIANGLE else 3.0) * sratio
    c = (o + 0.5) * ratio
    left = min(max(math.floor(c - support), 0), n - 1)
    right = min(max(math.ceil(c + support), left + 1), n)
    x = (np.arange(left, right) - (c - 0.5)) / sratio
    w = np.maximum(1 - np.abs(x), 0) if filt == abi.SCALE_FILTER_TRIANGLE else _lanczos3_f64(x)
    return left, w / w.sum()


@pytest.mark.parametrize("n,m", [(4, 8), (10, 7), (720, 1280), (1080, 480), (96, 17), (3, 12)])
@pytest.mark.parametrize("filt", [abi.SCALE_FILTER_LANCZOS3, abi.SCALE_FILTER_TRIANGLE])
def test_taps_match_float64_restatement(oracle, n, m, filt):
    """Tap windows identical and weights within 2e-6 of a float64 evaluation, for every output of
    up- and down-scaling passes; the weights sum to 1 and are mirror images for mirrored outputs."""
    for o in range(m):
        left, w = _taps(oracle, filt, n, m, o)
        l64, w64 = _taps_f64(filt, n, m, o)
        assert left == l64 and len(w) == len(w64), (o, left, l64, len(w), len(w64))
        assert np.abs(w - w64).max() <= 2e-6, o
        assert abs(w.sum() - 1) <= 1e-5
    for o in range(m // 2):
        _, a = _taps(oracle, filt, n, m, o)
        _, b = _taps(oracle, filt, n, m, m - 1 - o)
        if len(a) == len(b):
            assert np.abs(a - b[::-1]).max() <= 2e-6, o


def _resize_f64(plane, nw, nh, filt):
    h, w = plane.shape
    tmp = np.zeros((nh, w))
    for oy in range(nh):
        left, wt = _taps_f64(filt, h, nh, oy)
        tmp[oy] = (plane[left:left + len(wt)].astype(np.float64) * wt[:, None]).sum(0)
    out = np.zeros((nh, nw))
    for ox in range(nw):
        left, wt = _taps_f64(filt, w, nw, ox)
        out[:, ox] = (tmp[:, left:left + len(wt)] * wt[None, :]).sum(1)
    return np.clip(out, 0, 255)


def _crop_pad(scaled, W, H, bg):
    nh, nw = scaled.shape
    cw, ch = min(nw, W), min(nh, H)
    sx, sy, dx, dy = (nw - cw) // 2, (nh - ch) // 2, (W - cw) // 2, (H - ch) // 2
    out = np.full((H, W), float(bg))
    out[dy:dy + ch, dx:dx + cw] = scaled[sy:sy + ch, sx:sx + cw]
    return out


def _scaled_dims(oracle, sw, sh, cw, ch, mode):
    ow, oh = C.c_uint32(), C.c_uint32()
    oracle.oracle_scaled_dimensions(sw, sh, cw, ch, mode, C.byref(ow), C.byref(oh))
    return ow.value, oh.value


SCALE_CASES = [
    # (width, height, canvas_w, canvas_h, mode)
    (64, 48, 128, 72, abi.SCALE_FIT),
    (64, 48, 128, 72, abi.SCALE_FILL),
    (48, 64, 40, 40, abi.SCALE_FIT),
    (100, 36, 60, 80, abi.SCALE_FILL),
    (30, 20, 64, 16, abi.SCALE_STRETCH),
    (40, 30, 24, 50, abi.SCALE_NATIVE),
    (16, 16, 16, 16, abi.SCALE_FIT),       # no resize at all
    (18, 10, 36, 10, abi.SCALE_STRETCH),   # only the width changes
    (1024, 64, 16, 16, abi.SCALE_FIT),     # 64x downscale: a tile's taps exceed one-pass capacity
    (640, 40, 32, 24, abi.SCALE_STRETCH),  # 20x horizontal, 1.7x vertical downscale
]


@pytest.mark.parametrize("case", SCALE_CASES)
@pytest.mark.parametrize("filt", [abi.SCALE_FILTER_LANCZOS3, abi.SCALE_FILTER_TRIANGLE])
def test_scale_matches_float64_restatement(oracle, case, filt):
    """Every canvas sample within 1 of the float64 resampler (clamped, rounded) with its crop / pad,
    and at least 95% identical (f32 weights and sums move values that sit at a rounding tie, which
    the Triangle filter's rational weights make common; truncating instead of rounding would leave
    about half the samples off by one); the padding is exact."""
    width, height, cw, ch, mode = case
    rng = np.random.default_rng(width * 1000 + cw + mode)
    planes = scale_planes(rng, width, height, "smooth")
    for rng_range, black in ((abi.RANGE_LIMITED, 16), (abi.RANGE_FULL, 0)):
        out = oracle_scale(oracle, planes, cw, ch, mode, rng_range, filt)
        sw, sh = _scaled_dims(oracle, width, height, cw, ch, mode)
        for k in range(3):
            d = 1 if k == 0 else 2
            src = planes[k]
            if (sw // d, sh // d) == src.shape[::-1]:
                scaled = src.astype(np.float64)
            else:
                scaled = np.floor(_resize_f64(src, sw // d, sh // d, filt) + 0.5)
            want = _crop_pad(scaled, cw // d, ch // d, black if k == 0 else 128)
            diff = np.abs(out[k].astype(np.float64) - want)
            assert diff.max() <= 1, (k, diff.max())
            assert (diff == 0).mean() >= 0.95, (k, (diff == 0).mean())


@pytest.mark.parametrize("sw,sh,cw,ch", [(1920, 1080, 1280, 720), (1280, 720, 1080, 1920), (720, 576, 1920, 1080),
                                         (100, 3, 64, 64), (3, 100, 64, 64), (640, 480, 642, 480)])
def test_scaled_dimensions(oracle, sw, sh, cw, ch):
    """FIT fits inside the canvas touching two sides, FILL covers it touching two sides, both keep
    the aspect ratio to within one even step and give even sizes of at least 2; STRETCH is the
    canvas and NATIVE the frame."""
    fw, fh = _scaled_dims(oracle, sw, sh, cw, ch, abi.SCALE_FIT)
    assert fw <= cw and fh <= ch and (fw == cw or fh == ch)
    lw, lh = _scaled_dims(oracle, sw, sh, cw, ch, abi.SCALE_FILL)
    assert lw >= cw and lh >= ch and (lw == cw or lh == ch)
    for w, h in ((fw, fh), (lw, lh)):
        assert w % 2 == 0 and h % 2 == 0 and w >= 2 and h >= 2
        if w == cw:
            assert abs(h - sh * cw / sw) <= 1
        else:
            assert abs(w - sw * ch / sh) <= 1
    assert _scaled_dims(oracle, sw, sh, cw, ch, abi.SCALE_STRETCH) == (cw, ch)
    assert _scaled_dims(oracle, sw, sh, cw, ch, abi.SCALE_NATIVE) == (sw, sh)
