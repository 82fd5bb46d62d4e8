"""Pins of the oracle's bmfrPre and bmfrPost (bmfrPre.comp:5-98, bmfrPost.comp:5-124)
against what the shaders' definitions fix, on the CPU.

Each input is built so that an output names where it came from: the noisy colour
of pixel (x, y) is (x, y, 1), so the colour features of the padded grid tell which
image pixel every grid position read; a depth ramp normalises to a known ramp; a
weight array with one non-zero feature turns the weighted sum into that feature.
A shared mistake in the pixel offsets, the mirror, the block-local coordinates,
the blend weight, the layer ping-pong, the remodulation or the gamma fails one of
these, independently of the GPU kernels that are compared with the oracle.
"""
import numpy as np
import pytest

# pixelOffsets (bmfrGeneral.comp:38), fractions of the block edge
PIXEL_OFFSETS = [(.7, .85), (.95, .5), (.43, .76), (.97, .03), (.37, .58), (.03, .36), (.81, .46), (0, .78),
                 (.36, -.08), (-.06, 0), (.95, .1), (.85, .61), (.06, .1), (.43, .16), (0, .5), (.73, .38)]


def half(a):
    return np.asarray(a, np.float32).astype(np.float16).view(np.uint16)


def unhalf(a):
    return np.asarray(a, np.uint16).view(np.float16).astype(np.float32)


def shift(bs, frame):
    """ivec2(vec2(BLOCK) * pixelOffsets[frame % 16]): an fp32 product converted to int,
    i.e. truncated towards zero (-0.08 * 32 = -2.56 -> -2, not -3)."""
    ox, oy = PIXEL_OFFSETS[frame % 16]
    return int(np.float32(bs) * np.float32(ox)), int(np.float32(bs) * np.float32(oy))


def reflect(a, s):
    """mirror() of bmfrGeneral.comp:93-97 as a whole-array reflection."""
    return np.where(a < 0, -a - 1, np.where(a >= s, 2 * s - a - 1, a))


def source_of_grid(w, h, bs, frame):
    """For every padded grid position (gy, gx): the image pixel it reads and whether
    it lies inside the image unshifted-unmirrored."""
    wp, hp = (w // bs + 2) * bs, (h // bs + 2) * bs
    ox, oy = shift(bs, frame)
    gy, gx = np.mgrid[0:hp, 0:wp]
    ax, ay = gx - ox, gy - oy
    ix, iy = reflect(ax, w), reflect(ay, h)
    return ix, iy, (ix == ax) & (iy == ay)


def tagged_noisy(w, h):
    """Noisy RGBA16F whose colour is (x, y, 1): exact in fp16 below 2048."""
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    return half(np.stack([xx, yy, np.ones_like(xx), np.ones_like(xx)], -1))


def flat_gbuffer(w, h, depth=None):
    d = np.full((h, w), 3.0, np.float32) if depth is None else depth.astype(np.float32)
    n = np.zeros((h, w, 2), np.float32)  # theta 0: normal (0, 0, 1)
    return d, n


@pytest.mark.parametrize("w,h,bs", [(96, 72, 32), (77, 71, 32), (131, 65, 32), (50, 37, 16), (40, 23, 8)])
@pytest.mark.parametrize("frame", [0, 3, 8, 9, 15])
def test_pre_reads_the_shifted_mirrored_pixel(oracle_lib, w, h, bs, frame):
    """bmfrPre.comp:16-17: grid position g reads mirror(g - trunc(bs * pixelOffsets[f % 16]))."""
    depth, normal = flat_gbuffer(w, h)
    feats = oracle_lib.bmfr_pre(w, h, bs, frame, depth, normal, tagged_noisy(w, h))
    ix, iy, _ = source_of_grid(w, h, bs, frame)
    assert np.array_equal(unhalf(feats[10]), ix.astype(np.float32))
    assert np.array_equal(unhalf(feats[11]), iy.astype(np.float32))
    assert np.all(unhalf(feats[12]) == 1.0)
    assert np.all(unhalf(feats[0]) == 1.0)  # the constant feature


def test_pre_offsets_truncate_towards_zero(oracle_lib):
    """Frame 8 shifts by (trunc(11.52), trunc(-2.56)) = (11, -2) at 32-pixel blocks: the
    first grid row reads image row 2, the first column image column 11 mirrored."""
    w, h, bs = 96, 72, 32
    depth, normal = flat_gbuffer(w, h)
    feats = oracle_lib.bmfr_pre(w, h, bs, 8, depth, normal, tagged_noisy(w, h))
    assert unhalf(feats[11])[0, 0] == 2.0           # ay = 0 - (-2) = 2
    assert unhalf(feats[10])[0, 0] == 10.0          # ax = 0 - 11 = -11 -> mirror -> 10
    assert unhalf(feats[10])[0, 11] == 0.0          # ax = 0


@pytest.mark.parametrize("bs", [32, 16, 8])
def test_pre_block_local_coordinates(oracle_lib, bs):
    """POSITION_DEPTH (bmfrPre.comp:36-43): x, y = local invocation id / (block - 1), and
    their squares, the same in every block."""
    w, h = 3 * bs + 5, 2 * bs + 7
    depth, normal = flat_gbuffer(w, h)
    feats = oracle_lib.bmfr_pre(w, h, bs, 2, depth, normal, tagged_noisy(w, h))
    wp, hp = (w // bs + 2) * bs, (h // bs + 2) * bs
    gy, gx = np.mgrid[0:hp, 0:wp]
    fx = (gx % bs).astype(np.float32) / np.float32(bs - 1)
    fy = (gy % bs).astype(np.float32) / np.float32(bs - 1)
    assert np.array_equal(feats[4], half(fx)) and np.array_equal(feats[5], half(fy))
    assert np.array_equal(feats[7], half(fx * fx)) and np.array_equal(feats[8], half(fy * fy))


def test_pre_depth_normalises_per_block(oracle_lib):
    """A depth ramp d = 5 + 0.25 x normalises, per 32x32 block of the shifted grid, to
    (d - min) / (max - min + 1e-6) over the block: on blocks that read 32 consecutive
    unmirrored columns that is k / 31 for local column k (to 1e-6 relative), and its
    square; a constant depth gives 0."""
    w, h, bs, frame = 160, 96, 32, 5
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    depth = 5.0 + 0.25 * xx
    _, normal = flat_gbuffer(w, h)
    feats = oracle_lib.bmfr_pre(w, h, bs, frame, depth, normal, tagged_noisy(w, h))
    ix, iy, inside = source_of_grid(w, h, bs, frame)
    wp, hp = ix.shape[1], ix.shape[0]
    checked = 0
    for by in range(hp // bs):
        for bx in range(wp // bs):
            sl = np.s_[by * bs:(by + 1) * bs, bx * bs:(bx + 1) * bs]
            cols = ix[sl][0]
            if not np.all(np.diff(cols) == 1):
                continue  # a block that crosses a mirror edge
            z = unhalf(feats[6][sl])
            want = np.broadcast_to(np.arange(bs, dtype=np.float64) / 31.0, z.shape)
            assert np.allclose(z, want, rtol=0, atol=2.0 ** -11), (bx, by)
            assert np.allclose(unhalf(feats[9][sl]), want ** 2, rtol=0, atol=2.0 ** -11)
            checked += 1
    assert checked >= 4
    const = oracle_lib.bmfr_pre(w, h, bs, frame, np.full((h, w), 7.5, np.float32), normal, tagged_noisy(w, h))
    assert np.all(unhalf(const[6]) == 0.0)


def test_pre_normal_decode_is_spherical(oracle_lib):
    """normal = (cos phi sin theta, sin phi sin theta, cos theta) (bmfrPre.comp:26-29),
    against numpy's float64 spherical coordinates, to fp16 rounding."""
    w, h, bs = 96, 72, 32
    r = np.random.default_rng(11)
    theta = r.uniform(0, np.pi, (h, w)).astype(np.float32)
    phi = r.uniform(-np.pi, np.pi, (h, w)).astype(np.float32)
    depth = np.full((h, w), 2.0, np.float32)
    feats = oracle_lib.bmfr_pre(w, h, bs, 4, depth, np.stack([theta, phi], -1), tagged_noisy(w, h))
    ix, iy, _ = source_of_grid(w, h, bs, 4)
    t, p = theta[iy, ix].astype(np.float64), phi[iy, ix].astype(np.float64)
    want = [np.cos(p) * np.sin(t), np.sin(p) * np.sin(t), np.cos(t)]
    for k in range(3):
        assert np.allclose(unhalf(feats[1 + k]), want[k], rtol=0, atol=2.0 ** -11 + 1e-6)


# ---------------------------------------------------------------------------
# bmfrPost
# ---------------------------------------------------------------------------
def post_inputs(w, h, spp_byte=255, motion=None, albedo=255):
    depth, normal = flat_gbuffer(w, h)
    alb = np.full((h, w, 4), albedo, np.uint8)
    if motion is None:
        motion = np.full((h, w, 2), -1.0, np.float32)
    spp = np.full((h, w), spp_byte, np.uint8)
    return dict(depth=depth, normal=normal, albedo=alb, motion=half(motion), spp=spp,
                noisy=tagged_noisy(w, h))


def run_post(oracle_lib, w, h, bs, frame, inp, weights, denoised=None):
    if denoised is None:
        denoised = np.zeros((2, h, w, 4), np.uint16)
    return oracle_lib.bmfr_post(w, h, bs, frame, inp["depth"], inp["normal"], inp["albedo"], inp["motion"],
                                inp["spp"], inp["noisy"], weights, denoised)


def block_weights(w, h, bs, fn):
    """weights[by][bx][3 * feature + channel] = fn(bx, by, feature, channel)."""
    nbx, nby = w // bs + 2, h // bs + 2
    wt = np.zeros((nby, nbx, 30), np.float32)
    for by in range(nby):
        for bx in range(nbx):
            for f in range(10):
                for c in range(3):
                    wt[by, bx, 3 * f + c] = fn(bx, by, f, c)
    return wt


@pytest.mark.parametrize("frame", [0, 8, 9])
def test_post_block_weights_and_local_features(oracle_lib, frame):
    """With only the constant feature weighted, each pixel's output is its block's
    weight: the block is the one of the padded grid position g = pixel + shift, i.e.
    (g // bs).  With only feature 4 (x = local column / (bs - 1)) weighted 1, the output
    is that pixel's local column in the shifted grid / 31."""
    w, h, bs = 96, 72, 32
    inp = post_inputs(w, h)
    wt = block_weights(w, h, bs, lambda bx, by, f, c: (0.25 + 0.5 * bx + 0.125 * by if c == 0 else
                                                       1.0 + 0.25 * by if c == 1 else 0.5) if f == 0 else 0.0)
    den, _ = run_post(oracle_lib, w, h, bs, frame, inp, wt)
    out = unhalf(den[(frame & 1) ^ 1])
    ox, oy = shift(bs, frame)
    yy, xx = np.mgrid[0:h, 0:w]
    gx, gy = xx + ox, yy + oy
    # a pixel is written from its unmirrored grid position; with a negative shift the
    # first column (frame 9) or rows (frame 8) have none and keep their old contents
    covered = (gx >= 0) & (gy >= 0)
    assert covered.mean() < 1.0 if min(ox, oy) < 0 else covered.all()
    bx, by = gx // bs, gy // bs
    assert np.array_equal(out[..., 0], np.where(covered, unhalf(half(0.25 + 0.5 * bx + 0.125 * by)), 0))
    assert np.array_equal(out[..., 1], np.where(covered, unhalf(half(1.0 + 0.25 * by)), 0))
    assert np.array_equal(out[..., 2], np.where(covered, 0.5, 0)) and np.array_equal(out[..., 3], covered * 1.0)
    wt = block_weights(w, h, bs, lambda bx, by, f, c: 1.0 if f == 4 else 0.0)
    den, _ = run_post(oracle_lib, w, h, bs, frame, inp, wt)
    lx = (gx % bs).astype(np.float32)
    want = np.where(covered, unhalf(half(lx / np.float32(bs - 1))), 0)
    assert np.array_equal(unhalf(den[(frame & 1) ^ 1])[..., 0], want)


def test_post_non_finite_weights_and_clamp(oracle_lib):
    """bmfrPost.comp:96-104: NaN and Inf weights count as 0; the sum is clamped to
    [0, 10]."""
    w, h, bs = 96, 72, 32
    inp = post_inputs(w, h)
    wt = block_weights(w, h, bs, lambda bx, by, f, c: 0.0)
    wt[..., 0], wt[..., 1], wt[..., 2] = np.nan, np.inf, 0.75      # constant feature
    wt[..., 3 * 3 + 0] = 2.0                                        # normal.z (= 1) into red
    wt[..., 3 * 3 + 2] = -np.inf                                    # and into blue
    den, _ = run_post(oracle_lib, w, h, bs, 0, inp, wt)
    out = unhalf(den[1])
    # kept, +Inf would clamp to 10 and -Inf to 0
    assert np.all(out[..., 0] == 2.0) and np.all(out[..., 1] == 0.0) and np.all(out[..., 2] == 0.75)
    wt = block_weights(w, h, bs, lambda bx, by, f, c: (25.0, -3.0, 9.5)[c] if f == 0 else 0.0)
    den, _ = run_post(oracle_lib, w, h, bs, 0, inp, wt)
    out = unhalf(den[1])
    assert np.all(out[..., 0] == 10.0) and np.all(out[..., 1] == 0.0) and np.all(out[..., 2] == 9.5)


@pytest.mark.parametrize("frame", [1, 2, 7])
def test_post_temporal_blend(oracle_lib, frame):
    """bmfrPost.comp:109-118: on frame 0 or a rejected motion vector the blend weight is
    1; otherwise max(1 / spp, 0.1) with spp = R8 UNORM * 256, against the previous
    denoised layer frame & 1 sampled bilinearly at the motion uv (a texel centre returns
    the texel), written to layer (frame & 1) ^ 1."""
    w, h, bs = 64, 32, 8
    r = np.random.default_rng(frame)
    prev = half(r.uniform(0, 3, (2, h, w, 4)))
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    centre = np.stack([(xx + 0.5) / w, (yy + 0.5) / h], -1)  # exact in fp16 for these sizes
    assert np.array_equal(unhalf(half(centre)), centre)
    wt = block_weights(w, h, bs, lambda bx, by, f, c: (1.5, 0.5, 2.0)[c] if f == 0 else 0.0)
    cur = np.array([1.5, 0.5, 2.0], np.float32)
    hist = unhalf(prev[frame & 1])[..., :3]
    for spp_byte in (1, 3, 9, 255):
        pspp = np.float32(spp_byte) / np.float32(255) * np.float32(256)
        alpha = max(np.float32(1) / pspp, np.float32(0.1))
        inp = post_inputs(w, h, spp_byte=spp_byte, motion=centre)
        den, _ = run_post(oracle_lib, w, h, bs, frame, inp, wt, prev)
        want = alpha * cur + (np.float32(1) - alpha) * hist
        assert np.allclose(unhalf(den[(frame & 1) ^ 1])[..., :3], want, rtol=2.0 ** -10, atol=0), spp_byte
        assert np.array_equal(den[frame & 1], prev[frame & 1])  # the layer read is not written
    # rejected motion (negative u) and frame 0: no history
    inp = post_inputs(w, h, spp_byte=255, motion=np.full((h, w, 2), -1.0, np.float32))
    den, _ = run_post(oracle_lib, w, h, bs, frame, inp, wt, prev)
    assert np.all(unhalf(den[(frame & 1) ^ 1])[..., :3] == cur)
    inp = post_inputs(w, h, spp_byte=255, motion=centre)
    den, _ = run_post(oracle_lib, w, h, bs, 0, inp, wt, prev)
    assert np.all(unhalf(den[1])[..., :3] == cur)


@pytest.mark.parametrize("albedo", [0, 51, 200, 255])
def test_post_remodulation_and_gamma(oracle_lib, albedo):
    """bmfrPost.comp:120-123 on a constant image: final = clamp(((albedo / 255 + 1e-6) *
    denoised) ^ 0.454545, 0, 1) stored as B8G8R8A8 (blue byte first)."""
    w, h, bs = 96, 72, 32
    inp = post_inputs(w, h, albedo=albedo)
    col = np.array([0.35, 1.7, 0.06], np.float32)
    wt = block_weights(w, h, bs, lambda bx, by, f, c: col[c] if f == 0 else 0.0)
    den, final = run_post(oracle_lib, w, h, bs, 0, inp, wt)
    d = col.astype(np.float64)  # frame 0: alpha 1; remodulation uses the fp32 sum, not its fp16 store
    t = np.clip(((albedo / 255.0 + 1e-6) * d) ** 0.454545, 0, 1)
    want = np.floor(t * 255 + 0.5)
    got = final.reshape(-1, 4)
    assert np.all(np.abs(got[:, 2].astype(int) - want[0]) <= 1)
    assert np.all(np.abs(got[:, 1].astype(int) - want[1]) <= 1)
    assert np.all(np.abs(got[:, 0].astype(int) - want[2]) <= 1)
    assert np.all(got[:, 3] == 255)
