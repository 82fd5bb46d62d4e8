"""ctypes wrapper of the CPU oracle ``oracle/liboracle.so``.

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` - never by the product package.
Build with ``make -C oracle``.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.orc_rng_sequence.argtypes = [C.c_uint32] * 4 + [C.c_void_p]
        L.orc_fn_probe.argtypes = [C.c_int32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_wang_hash.argtypes = [C.c_uint32]
        L.orc_wang_hash.restype = C.c_uint32
        L.orc_scene_create.argtypes = [C.c_void_p]
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_scene_light_count.argtypes = [C.c_void_p]
        L.orc_scene_light_count.restype = C.c_uint32
        L.orc_scene_strength_sum.argtypes = [C.c_void_p]
        L.orc_scene_strength_sum.restype = C.c_float
        L.orc_scene_triangle_count.argtypes = [C.c_void_p]
        L.orc_scene_triangle_count.restype = C.c_uint32
        L.orc_scene_lights.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_intersect.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
        L.orc_trace.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                C.c_uint32, C.c_uint32, C.c_int32] + [C.c_void_p] * 7
        vp = C.c_void_p
        L.orc_accumulate.argtypes = [C.c_uint32] * 3 + [vp] * 9
        L.orc_taa.argtypes = [C.c_uint32] * 3 + [vp] * 4
        L.orc_convert.argtypes = [C.c_uint32] * 2 + [vp] * 2
        L.orc_bmfr_pre.argtypes = [C.c_uint32] * 4 + [vp] * 4
        L.orc_bmfr_fit.argtypes = [C.c_uint32] * 4 + [C.c_int32, vp, vp]
        L.orc_bmfr_lsq64.argtypes = [C.c_uint32] * 4 + [vp, vp]
        L.orc_bmfr_block_matrix.argtypes = [C.c_uint32] * 4 + [vp, C.c_uint32, C.c_uint32, vp]
        L.orc_bfr.argtypes = [C.c_uint32] * 4 + [C.c_int32] + [vp] * 9
        L.orc_bfr_blend.argtypes = [C.c_uint32] * 3 + [vp] * 6
        L.orc_bmfr_post.argtypes = [C.c_uint32] * 4 + [vp] * 9
        L.orc_brdf_probe.argtypes = [vp, C.c_uint32, vp, vp, vp, vp]
        L.orc_sample_brdf_probe.argtypes = [vp, vp, C.c_uint32, C.c_uint32, vp]
        L.orc_sample_light_probe.argtypes = [vp, C.c_uint32, vp, vp, C.c_uint32, C.c_uint32, vp]
        L.orc_shade_probe.argtypes = [vp] * 4
        L.orc_shade_probe.restype = C.c_int32
        _lib = L
    return _lib


def _p(a):
    """Address of an array's data; the caller keeps the array alive across the call."""
    return None if a is None else a.ctypes.data


def rng_sequence(width: int, height: int, frame: int, count: int) -> np.ndarray:
    out = np.zeros((height, width, count), dtype=np.uint32)
    lib().orc_rng_sequence(width, height, frame, count, _p(out))
    return out


FN_KINDS = {"log2": 0, "exp2": 1, "pow": 2, "pow5": 3, "sin": 4, "cos": 5}


def fn_probe(kind: str, x, y=None) -> np.ndarray:
    """The project's transcendental `kind` (pbrt_fn.hpp, shared by the device and the
    FP32 mirror) evaluated on float32 arrays."""
    x = np.ascontiguousarray(x, np.float32)
    yy = np.ascontiguousarray(y if y is not None else np.zeros_like(x), np.float32)
    out = np.zeros_like(x)
    lib().orc_fn_probe(FN_KINDS[kind], x.size, _p(x), _p(yy), _p(out))
    return out


def wang_hash(x: int) -> int:
    return int(lib().orc_wang_hash(x))


class OracleScene:
    """The oracle's own packing of a pbrt_scene_desc (same input the product gets)."""

    def __init__(self, desc):
        self.handle = lib().orc_scene_create(C.addressof(desc))

    @property
    def light_count(self) -> int:
        return int(lib().orc_scene_light_count(self.handle))

    @property
    def strength_sum(self) -> float:
        return float(lib().orc_scene_strength_sum(self.handle))

    @property
    def triangle_count(self) -> int:
        return int(lib().orc_scene_triangle_count(self.handle))

    def lights(self) -> np.ndarray:
        out = np.zeros((self.light_count, 32), dtype=np.float32)
        lib().orc_scene_lights(self.handle, _p(out))
        return out

    def intersect(self, rays: np.ndarray, brute: bool = False, any_hit: bool = False):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[0]
        ip = np.zeros((n, 2), dtype=np.int32)
        tuv = np.zeros((n, 3), dtype=np.float64)
        lib().orc_intersect(self.handle, _p(rays), n, int(brute), int(any_hit), _p(ip), _p(tuv))
        return ip, tuv

    def trace(self, width, height, inv_view, inv_proj, frame=0, sample=0, max_depth=2, min_depth=0,
              light_mode=0, output=0, precision=32, brute=False, rows=None, threads=8, prev_final=None):
        """Trace a frame (or rows [y0, y1)); returns a dict of images (H x W [x C])."""
        y0, y1 = rows if rows is not None else (0, height)
        params = np.array([width, height, max_depth, min_depth, light_mode, output, 1, frame, sample],
                          dtype=np.uint32)
        depth = np.zeros((height, width), dtype=np.float32)
        normal = np.zeros((height, width, 2), dtype=np.float32)
        material = np.zeros((height, width, 4), dtype=np.uint8)
        albedo = np.zeros((height, width, 4), dtype=np.uint8)
        illum = np.zeros((height, width, 4), dtype=np.float32) if prev_final is None else \
            np.ascontiguousarray(prev_final, dtype=np.float32).copy()
        radiance = np.zeros((height, width, 3), dtype=np.float64)
        stats = np.zeros(2, dtype=np.uint64)
        iv = np.ascontiguousarray(inv_view, dtype=np.float32)
        ip = np.ascontiguousarray(inv_proj, dtype=np.float32)
        lib().orc_trace(self.handle, _p(params), _p(iv), _p(ip), precision, int(brute), y0, y1, threads,
                        _p(depth), _p(normal), _p(material), _p(albedo), _p(illum), _p(radiance), _p(stats))
        return dict(depth=depth, normal=normal, material=material, albedo=albedo, illumination=illum,
                    radiance=radiance, rays=int(stats[0]), shadow_rays=int(stats[1]))

    def sample_light(self, mode, pos, normal, count, frame=1):
        """`count` draws of sampleLight (lighting.glsl:18-227) at `pos` with `normal`:
        dict of ok (bool), col (n,3), l (n,3), pdf, tmax."""
        out = np.zeros((count, 9), np.float64)
        pos, normal = np.ascontiguousarray(pos, np.float64), np.ascontiguousarray(normal, np.float64)
        lib().orc_sample_light_probe(self.handle, mode, _p(pos), _p(normal), count, frame, _p(out))
        return dict(ok=out[:, 0] > 0, col=out[:, 1:4], l=out[:, 4:7], pdf=out[:, 7], tmax=out[:, 8])

    def shade(self, origin, direction):
        """The surface the closest-hit shader (ptClosesthit.rchit:18-92) builds at the
        closest hit of one ray, or None on a miss."""
        out = np.zeros(34, np.float64)
        o, d = np.ascontiguousarray(origin, np.float64), np.ascontiguousarray(direction, np.float64)
        tri = lib().orc_shade_probe(self.handle, _p(o), _p(d), _p(out))
        if tri < 0:
            return None
        return dict(tri=tri, position=out[0:3], normal=out[3:6], diffuse=out[6:9], r0=out[9:12], r90=out[12],
                    alpha_roughness=out[13], metalness=out[14], emissive=out[15:18],
                    basis=out[18:27].reshape(3, 3), perceptual_roughness=out[27], t=out[28], u=out[29],
                    v=out[30], specular=out[31:34])

    def __del__(self):
        try:
            if self.handle:
                lib().orc_scene_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


# --------------------------------------------------------------------------
# Single shading functions (brdf.glsl, sampling.glsl), FP64
# --------------------------------------------------------------------------
def surface(perceptual_roughness=0.5, metalness=0.0, alpha_roughness=None, r0=(0.04,) * 3, r90=(1.0,) * 3,
            diffuse=(0.5,) * 3, specular=(0.04,) * 3, normal=(0, 0, 1), basis=None):
    """The 27-double surface record of the probes; basis rows = tangent, bitangent, normal."""
    if alpha_roughness is None:
        alpha_roughness = perceptual_roughness ** 2
    n = np.asarray(normal, np.float64)
    if basis is None:
        t = np.cross(n, [0.0, 1.0, 0.0] if abs(n[0]) > 0.9 else [1.0, 0.0, 0.0])
        t /= np.linalg.norm(t)
        b = np.cross(n, t)
        basis = np.stack([t, b, n])
    return np.concatenate([[perceptual_roughness, metalness, alpha_roughness], r0, r90, diffuse, specular, n,
                           np.asarray(basis, np.float64).ravel()]).astype(np.float64)


def brdf(surf, v, l, h=None):
    """BRDF (brdf.glsl:132-164), pdf_brdf(l, h) and the specular weight for direction
    pairs v, l of shape (3,) or (n, 3), with h = normalize(v + l) unless given:
    (brdf (n,3), pdf (n), spec_weight (n))."""
    arrs = [np.atleast_2d(np.asarray(x, np.float64)) for x in ((v, l) if h is None else (v, l, h))]
    arrs = [np.ascontiguousarray(x) for x in np.broadcast_arrays(*arrs)]
    n = arrs[0].shape[0]
    out = np.zeros((n, 5), np.float64)
    lib().orc_brdf_probe(_p(surf), n, _p(arrs[0]), _p(arrs[1]), _p(arrs[2]) if h is not None else None, _p(out))
    return out[:, :3], out[:, 3], out[:, 4]


def sample_brdf(surf, v, count, frame=1):
    """`count` draws of sampleBRDF (sampling.glsl:11-78): l (n,3), pdf (n), brdf (n,3)."""
    out = np.zeros((count, 7), np.float64)
    v = np.ascontiguousarray(v, np.float64)
    lib().orc_sample_brdf_probe(_p(surf), _p(v), count, frame, _p(out))
    return out[:, 0:3], out[:, 3], out[:, 4:7]


# --------------------------------------------------------------------------
# Screen-space passes
# --------------------------------------------------------------------------
def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def accumulate(width, height, frame, inv_proj, inv_view, prev_view, prev_origin, illumination, depth,
               prev_depth, prev_spp, prev_accumulated):
    """accumulator.comp; returns (motion RG16F bits, spp R8, accumulated RGBA16F bits)."""
    consts = np.concatenate([_c(inv_proj, np.float32), _c(inv_view, np.float32), _c(prev_view, np.float32),
                             _c(prev_origin, np.float32)])
    motion = np.zeros((height, width, 2), np.uint16)
    spp = np.zeros((height, width), np.uint8)
    acc = np.zeros((height, width, 4), np.uint16)
    ill = _c(illumination, np.float32)
    dep = _c(depth, np.float32)
    pd = _c(prev_depth, np.float32)
    ps = _c(prev_spp, np.uint8)
    pa = _c(prev_accumulated, np.uint16)
    lib().orc_accumulate(width, height, frame, _p(consts), _p(ill), _p(dep), _p(pd), _p(ps), _p(pa),
                         _p(motion), _p(spp), _p(acc))
    return motion, spp, acc


def taa(width, height, frame, motion, denoised_bgra, history):
    """taa.comp + history copy; returns (final B8G8R8A8, new history)."""
    hist = _c(history, np.uint8).copy()
    final = np.zeros((height, width, 4), np.uint8)
    mo = _c(motion, np.uint16)
    de = _c(denoised_bgra, np.uint8)
    lib().orc_taa(width, height, frame, _p(mo), _p(de), _p(hist), _p(final))
    return final, hist


def convert(width, height, src):
    out = np.zeros((height, width, 4), np.uint8)
    s = _c(src, np.float32)
    lib().orc_convert(width, height, _p(s), _p(out))
    return out


def bmfr_geometry(width, height, bs):
    wp = (width // bs + 2) * bs
    hp = (height // bs + 2) * bs
    return wp, hp, wp // bs, hp // bs


def _check_mirror(width, height, bs):
    """mirror() reflects once (bmfrGeneral.comp:93-97): the padded grid shifted by the
    most negative pixelOffsets must stay within one image size, or the passes would
    read outside the images (the product rejects such sizes)."""
    wp, hp, _, _ = bmfr_geometry(width, height, bs)
    if wp - int(np.float32(bs) * np.float32(-0.06)) > 2 * width or \
            hp - int(np.float32(bs) * np.float32(-0.08)) > 2 * height:
        raise ValueError(f"{width}x{height} with {bs}-pixel blocks mirrors outside the image")


def bmfr_pre(width, height, bs, frame, depth, normal, noisy):
    _check_mirror(width, height, bs)
    wp, hp, _, _ = bmfr_geometry(width, height, bs)
    feats = np.zeros((13, hp, wp), np.uint16)
    d, n, z = _c(depth, np.float32), _c(normal, np.float32), _c(noisy, np.uint16)
    lib().orc_bmfr_pre(width, height, bs, frame, _p(d), _p(n), _p(z), _p(feats))
    return feats


def bmfr_fit(width, height, bs, frame, features, subgroup=32):
    _, _, bx, by = bmfr_geometry(width, height, bs)
    w = np.zeros((by, bx, 30), np.float32)
    f = _c(features, np.uint16)
    lib().orc_bmfr_fit(width, height, bs, frame, subgroup, _p(f), _p(w))
    return w


def bmfr_lsq64(width, height, bs, frame, features):
    _, _, bx, by = bmfr_geometry(width, height, bs)
    w = np.zeros((by, bx, 30), np.float64)
    f = _c(features, np.uint16)
    lib().orc_bmfr_lsq64(width, height, bs, frame, _p(f), _p(w))
    return w


def bmfr_block_matrix(width, height, bs, frame, features, bx, by):
    out = np.zeros((bs * bs, 13), np.float32)
    f = _c(features, np.uint16)
    lib().orc_bmfr_block_matrix(width, height, bs, frame, _p(f), bx, by, _p(out))
    return out


def bmfr_post(width, height, bs, frame, depth, normal, albedo, motion, spp, noisy, weights, denoised):
    """bmfrPost; denoised is the (2, H, W, 4) RGBA16F array (updated copy returned)."""
    _check_mirror(width, height, bs)
    den = _c(denoised, np.uint16).copy()
    final = np.zeros((height, width, 4), np.uint8)
    args = [_c(depth, np.float32), _c(normal, np.float32), _c(albedo, np.uint8), _c(motion, np.uint16),
            _c(spp, np.uint8), _c(noisy, np.uint16), _c(weights, np.float32)]
    lib().orc_bmfr_post(width, height, bs, frame, *[_p(a) for a in args], _p(den), _p(final))
    return den, final


def bfr(width, height, bs, frame, depth, normal, albedo, motion, spp, noisy, denoised, iterations=40):
    """bfr.comp: returns (denoised (2, H, W, 4) RGBA16F copy, final BGRA8, alpha [by][bx][21])."""
    den = _c(denoised, np.uint16).copy()
    final = np.zeros((height, width, 4), np.uint8)
    nbx, nby = width // bs + 2, height // bs + 2
    weights = np.zeros((nby, nbx, 21), np.float32)
    args = [_c(depth, np.float32), _c(normal, np.float32), _c(albedo, np.uint8), _c(motion, np.uint16),
            _c(spp, np.uint8), _c(noisy, np.uint16)]
    lib().orc_bfr(width, height, bs, frame, iterations, *[_p(a) for a in args], _p(den), _p(final), _p(weights))
    return den, final, weights


def bfr_blend(width, height, radius, average, average_sq, den8, den16, den32):
    """bfrBlender.comp: BGRA8 blend of the 8/16/32 outputs."""
    final = np.zeros((height, width, 4), np.uint8)
    args = [_c(average, np.uint16), _c(average_sq, np.uint16), _c(den8, np.uint8), _c(den16, np.uint8),
            _c(den32, np.uint8)]
    lib().orc_bfr_blend(width, height, radius, *[_p(a) for a in args], _p(final))
    return final
