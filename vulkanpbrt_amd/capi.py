"""ctypes view of the C ABI in ``include/pbrt_c.h``, ``include/pbrt_scenegen.h`` and
``include/pbrt_io.h``.

The library ``libpbrt.so`` is built in-tree by :mod:`vulkanpbrt_amd.build`.
Loading fails loudly when it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libpbrt.so"

PBRT_OK = 0
PBRT_ERR_INVALID_ARG = -1
PBRT_ERR_UNSUPPORTED = -2
PBRT_ERR_OUT_OF_MEMORY = -3
PBRT_ERR_HIP = -4
PBRT_ERR_INTERNAL = -5

# enums (pbrt_c.h)
PBRT_BVH_SAH_HOST = 0
PBRT_BVH_LBVH_GPU = 1
PBRT_BVH_SBVH_HOST = 2
PBRT_BVH_PLOC_GPU = 3
LIGHT_SAMPLE_SURFACE_STRENGTH = 0
LIGHT_SAMPLE_LIGHT_STRENGTH = 1
LIGHT_SAMPLE_UNIFORM = 2
LIGHT_SAMPLE_AUTO = 0xFFFFFFFF
OUTPUT_DEMOD_ILLUMINATION_FLOAT = 0
OUTPUT_FINAL_IMAGE = 1
TRACER_MEGAKERNEL = 0
TRACER_WAVEFRONT = 1
TRACER_PERSISTENT = 2


class PushConstants(C.Structure):
    _fields_ = [
        ("view_inverse", C.c_float * 16),
        ("proj_inverse", C.c_float * 16),
        ("prev_view", C.c_float * 16),
        ("frame_number", C.c_uint32),
        ("sample_number", C.c_uint32),
    ]


class Light(C.Structure):
    _fields_ = [(n, C.c_float * 4) for n in (
        "v0_type", "v1_strength", "v2_angle", "dir_angle2",
        "col_ambient", "col_diffuse", "col_specular", "strengths")]


class Material(C.Structure):
    _fields_ = [
        ("ambient_roughness", C.c_float * 4),
        ("diffuse_ior", C.c_float * 4),
        ("specular_dissolve", C.c_float * 4),
        ("transmittance_illum", C.c_float * 4),
        ("emission_texture_id", C.c_float * 4),
        ("category_id", C.c_uint32),
        ("pad", C.c_float * 3),
    ]


class TextureDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba8", C.c_void_p)]


class MeshDesc(C.Structure):
    _fields_ = [
        ("vertex_count", C.c_uint32),
        ("positions", C.c_void_p),
        ("normals", C.c_void_p),
        ("uvs", C.c_void_p),
        ("index_count", C.c_uint32),
        ("index_stride", C.c_uint32),
        ("indices", C.c_void_p),
        ("material", Material),
        ("diffuse_texture", C.c_int32),
        ("normal_texture", C.c_int32),
        ("specular_texture", C.c_int32),
        ("emissive_texture", C.c_int32),
    ]


class InstanceDesc(C.Structure):
    _fields_ = [("transform", C.c_float * 16), ("mesh", C.c_uint32), ("pad", C.c_uint32 * 3)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("mesh_count", C.c_uint32),
        ("meshes", C.POINTER(MeshDesc)),
        ("instance_count", C.c_uint32),
        ("instances", C.POINTER(InstanceDesc)),
        ("texture_count", C.c_uint32),
        ("textures", C.POINTER(TextureDesc)),
        ("light_count", C.c_uint32),
        ("lights", C.POINTER(Light)),
        ("bvh_builder", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("triangle_count", C.c_uint32),
        ("instance_count", C.c_uint32),
        ("mesh_count", C.c_uint32),
        ("light_count", C.c_uint32),
        ("light_strength_sum", C.c_float),
        ("bvh_node_count", C.c_uint32),
        ("bvh_depth", C.c_uint32),
        ("light_sampling", C.c_uint32),
        ("device_bytes", C.c_uint64),
        ("bvh_builder", C.c_uint32),
        ("bvh_build_ms", C.c_float),
    ]


class TraceParams(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "width", "height", "max_recursion_depth", "min_recursion_depth",
        "light_sampling", "output", "write_gbuffer", "tracer", "block_threads", "queue_paths")]


TUNING_MAX = 16


class TraceTuning(C.Structure):
    _fields_ = [("best", TraceParams), ("candidates", C.c_uint32), ("tracer", C.c_uint32 * TUNING_MAX),
                ("block_threads", C.c_uint32 * TUNING_MAX), ("queue_paths", C.c_uint32 * TUNING_MAX),
                ("ms", C.c_float * TUNING_MAX)]


class TraceImages(C.Structure):
    _fields_ = [
        ("depth", C.c_void_p),
        ("normal", C.c_void_p),
        ("material", C.c_void_p),
        ("albedo", C.c_void_p),
        ("illumination", C.c_void_p),
        ("workspace", C.c_void_p),
        ("workspace_bytes", C.c_size_t),
    ]


class TraceTiming(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("extend_ms", "shade_ms", "shadow_ms", "megakernel_ms", "other_ms")] + [
        (n, C.c_uint64) for n in ("extend_launches", "shade_launches", "shadow_launches", "megakernel_launches",
                                  "frames")]


class CameraPose(C.Structure):
    _fields_ = [
        ("eye", C.c_float * 3),
        ("center", C.c_float * 3),
        ("up", C.c_float * 3),
        ("fovy_degrees", C.c_float),
        ("near_distance", C.c_float),
        ("far_distance", C.c_float),
    ]


class PhongMaterial(C.Structure):
    """vsg::PhongMaterial (pbrt_phong_material)."""
    _fields_ = [("ambient", C.c_float * 4), ("diffuse", C.c_float * 4), ("specular", C.c_float * 4),
                ("emissive", C.c_float * 4), ("transmissive", C.c_float * 3), ("shininess", C.c_float),
                ("alpha_mask", C.c_float), ("alpha_mask_cutoff", C.c_float), ("index_of_refraction", C.c_float),
                ("category_id", C.c_uint32)]


class PbrMaterial(C.Structure):
    """vsg::PbrMaterial (pbrt_pbr_material)."""
    _fields_ = [("base_color", C.c_float * 4), ("emissive", C.c_float * 4), ("diffuse", C.c_float * 4),
                ("specular", C.c_float * 4), ("transmission", C.c_float * 3), ("metallic", C.c_float),
                ("roughness", C.c_float), ("alpha_mask", C.c_float), ("alpha_mask_cutoff", C.c_float),
                ("index_of_refraction", C.c_float), ("category_id", C.c_uint32)]


class CameraMatrices(C.Structure):
    """CameraMatrices of RenderIO.hpp:27-33 (pbrt_camera_matrices), column-major."""
    _fields_ = [("view", C.c_float * 16), ("inv_view", C.c_float * 16), ("proj", C.c_float * 16),
                ("inv_proj", C.c_float * 16), ("has_proj", C.c_uint32), ("pad", C.c_uint32 * 3)]


class LoadOptions(C.Structure):
    _fields_ = [("instance_count", C.c_uint32), ("bvh_builder", C.c_uint32),
                ("instance_transforms", C.POINTER(C.c_float))]


class LoadedSceneStats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("mesh_count", "triangle_count", "vertex_count", "texture_count",
                                          "emissive_mesh_count", "instance_count")] + [
        ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3), ("parse_ms", C.c_double),
        ("file_bytes", C.c_uint64)]


class AccumulateParams(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("frame_number", C.c_uint32), ("reserved", C.c_uint32),
        ("inv_proj", C.c_float * 16), ("inv_view", C.c_float * 16), ("prev_view", C.c_float * 16),
        ("prev_origin", C.c_float * 4),
    ]


class AccumulateImages(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "illumination", "depth", "prev_depth", "prev_spp", "prev_accumulated", "motion", "spp", "accumulated")]


class TaaImages(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("motion", "denoised", "history", "final_image")]


BMFR_FIT_HOUSEHOLDER = 0
BMFR_FIT_GRAM_F16 = 1
BMFR_FIT_GRAM_F64 = 2
BMFR_NOALLOC_SCRATCH_BYTES = 8 << 20  # PBRT_BMFR_NOALLOC_SCRATCH_BYTES


class BmfrParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("block_size", C.c_uint32),
                ("frame_number", C.c_uint32), ("fit", C.c_uint32), ("fit_threads", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


class BmfrTuning(C.Structure):
    _fields_ = [("best", BmfrParams), ("candidates", C.c_uint32), ("fit_threads", C.c_uint32 * 4),
                ("ms", C.c_float * 4)]


class BmfrImages(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "depth", "normal", "albedo", "motion", "spp", "noisy", "denoised", "final_image", "workspace")] + [
        ("workspace_bytes", C.c_size_t)]


class BmfrBatchFrame(C.Structure):
    _fields_ = [("params", BmfrParams), ("images", BmfrImages)]


class BfrParams(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("width", "height", "block_size", "frame_number")]


class BfrImages(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "depth", "normal", "albedo", "motion", "spp", "noisy", "denoised", "final_image", "weights")]


class BfrBlendImages(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "average", "average_squared", "denoised8", "denoised16", "denoised32", "final_image")]


class CnnImages(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("count", C.c_uint32)] + [
        (n, C.c_void_p) for n in ("noisy", "depth", "normal", "albedo", "spp", "motion", "denoised", "final_image",
                                  "target", "workspace")] + [("workspace_bytes", C.c_size_t)]


# name -> (restype, argtypes); every symbol declared in include/*.h
SIGNATURES = {
    "pbrt_abi_version": (C.c_int32, []),
    "pbrt_last_error": (C.c_char_p, []),
    "pbrt_scene_create": (C.c_int32, [C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    "pbrt_scene_destroy": (C.c_int32, [C.c_void_p]),
    "pbrt_scene_get_info": (C.c_int32, [C.c_void_p, C.POINTER(SceneInfo)]),
    "pbrt_scene_get_lights": (C.c_int32, [C.c_void_p, C.POINTER(Light), C.c_uint32]),
    "pbrt_radix_sort_scratch_size": (C.c_int32, [C.c_uint32, C.POINTER(C.c_size_t)]),
    "pbrt_radix_sort_pairs": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                          C.c_void_p, C.c_size_t, C.c_void_p]),
    "pbrt_trace": (C.c_int32, [C.c_void_p, C.POINTER(TraceParams), C.POINTER(PushConstants),
                               C.POINTER(TraceImages), C.c_void_p, C.c_void_p]),
    "pbrt_trace_workspace_size": (C.c_int32, [C.c_uint32, C.c_uint32, C.POINTER(C.c_size_t)]),
    "pbrt_autotune_trace": (C.c_int32, [C.c_void_p, C.POINTER(TraceParams), C.POINTER(PushConstants),
                                        C.POINTER(TraceImages), C.c_uint32, C.POINTER(TraceTuning), C.c_void_p]),
    "pbrt_trace_timing_begin": (C.c_int32, [C.c_void_p]),
    "pbrt_trace_timing_end": (C.c_int32, [C.c_void_p, C.POINTER(TraceTiming)]),
    "pbrt_intersect": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pbrt_occluded": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "pbrt_rng_sequence": (C.c_int32, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "pbrt_accumulate": (C.c_int32, [C.POINTER(AccumulateParams), C.POINTER(AccumulateImages), C.c_void_p]),
    "pbrt_copy_back": (C.c_int32, [C.c_uint32, C.c_uint32] + [C.c_void_p] * 7),
    "pbrt_taa": (C.c_int32, [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(TaaImages), C.c_void_p]),
    "pbrt_convert": (C.c_int32, [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pbrt_bmfr_workspace_size": (C.c_int32, [C.POINTER(BmfrParams), C.POINTER(C.c_size_t)]),
    "pbrt_bmfr": (C.c_int32, [C.POINTER(BmfrParams), C.POINTER(BmfrImages), C.c_void_p]),
    "pbrt_bmfr_noalloc": (C.c_int32, [C.POINTER(BmfrParams), C.POINTER(BmfrImages), C.c_void_p]),
    "pbrt_autotune_bmfr": (C.c_int32, [C.POINTER(BmfrParams), C.POINTER(BmfrImages), C.c_uint32,
                                        C.POINTER(BmfrTuning), C.c_void_p]),
    "pbrt_bmfr_batched_workspace_size": (C.c_int32, [C.POINTER(BmfrBatchFrame), C.c_uint32, C.POINTER(C.c_size_t)]),
    "pbrt_bmfr_batched": (C.c_int32, [C.POINTER(BmfrBatchFrame), C.c_uint32, C.c_void_p, C.c_size_t, C.c_void_p]),
    "pbrt_running_mean": (C.c_int32, [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "pbrt_cnn_create_layers": (C.c_int32, [C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "pbrt_cnn_layers": (C.c_int32, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "pbrt_cnn_create_compact": (C.c_int32, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "pbrt_cnn_is_compact": (C.c_int32, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "pbrt_cnn_layers_parameter_count": (C.c_int32, [C.c_uint32, C.POINTER(C.c_uint32)]),
    "pbrt_frame_graph_create": (C.c_int32, [C.POINTER(C.c_void_p)]),
    "pbrt_frame_graph_destroy": (C.c_int32, [C.c_void_p]),
    "pbrt_frame_graph_begin": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "pbrt_frame_graph_abort": (C.c_int32, [C.c_void_p]),
    "pbrt_frame_graph_end": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "pbrt_cnn_create": (C.c_int32, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "pbrt_cnn_destroy": (C.c_int32, [C.c_void_p]),
    "pbrt_cnn_parameter_count": (C.c_int32, [C.POINTER(C.c_uint32)]),
    "pbrt_cnn_get_parameters": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "pbrt_cnn_set_parameters": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "pbrt_cnn_workspace_size": (C.c_int32, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_size_t)]),
    "pbrt_cnn_denoise": (C.c_int32, [C.c_void_p, C.POINTER(CnnImages), C.c_void_p]),
    "pbrt_cnn_gradients": (C.c_int32, [C.c_void_p, C.POINTER(CnnImages), C.c_void_p, C.c_void_p, C.c_void_p]),
    "pbrt_cnn_train_step": (C.c_int32, [C.c_void_p, C.POINTER(CnnImages), C.c_float, C.c_void_p, C.c_void_p]),
    "pbrt_bfr": (C.c_int32, [C.POINTER(BfrParams), C.POINTER(BfrImages), C.c_void_p]),
    "pbrt_bfr_blend": (C.c_int32, [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(BfrBlendImages), C.c_void_p]),
    "pbrt_generate_scene": (C.c_int32, [C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "pbrt_generated_scene_desc": (C.c_int32, [C.c_void_p, C.POINTER(SceneDesc)]),
    "pbrt_generated_scene_camera": (C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(CameraPose)]),
    "pbrt_generated_scene_destroy": (C.c_int32, [C.c_void_p]),
    "pbrt_pack_phong_material": (C.c_int32, [C.POINTER(PhongMaterial), C.POINTER(Material), C.POINTER(C.c_uint32)]),
    "pbrt_pack_pbr_material": (C.c_int32, [C.POINTER(PbrMaterial), C.POINTER(Material), C.POINTER(C.c_uint32)]),
    "pbrt_load_obj": (C.c_int32, [C.c_char_p, C.POINTER(LoadOptions), C.POINTER(C.c_void_p)]),
    "pbrt_loaded_scene_desc": (C.c_int32, [C.c_void_p, C.POINTER(SceneDesc)]),
    "pbrt_loaded_scene_stats_get": (C.c_int32, [C.c_void_p, C.POINTER(LoadedSceneStats)]),
    "pbrt_loaded_scene_destroy": (C.c_int32, [C.c_void_p]),
    "pbrt_image_load_rgba8": (C.c_int32, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                          C.POINTER(C.POINTER(C.c_uint8))]),
    "pbrt_io_free": (None, [C.c_void_p]),
    "pbrt_image_write_png": (C.c_int32, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32]),
    "pbrt_matrices_read": (C.c_int32, [C.c_char_p, C.POINTER(C.POINTER(CameraMatrices)), C.POINTER(C.c_uint32)]),
    "pbrt_matrices_write": (C.c_int32, [C.c_char_p, C.POINTER(CameraMatrices), C.c_uint32]),
    "pbrt_pfm_write": (C.c_int32, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32]),
    "pbrt_pfm_read": (C.c_int32, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                  C.POINTER(C.POINTER(C.c_float))]),
    "pbrt_normals_to_spherical": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "pbrt_spherical_to_normals": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "pbrt_albedo_from_float": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "pbrt_albedo_from_half": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "pbrt_unorm_to_float": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "pbrt_depth_to_position": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(CameraMatrices),
                                           C.c_void_p]),
    "pbrt_position_to_depth": (C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(CameraMatrices), C.c_void_p]),
}

_lib: C.CDLL | None = None


class PbrtError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"pbrt status {status}: {message}")
        self.status = status


def lib() -> C.CDLL:
    """Load ``libpbrt.so`` (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m vulkanpbrt_amd.build` "
                "(there is no CPU fallback for the HIP path)")
        # With torch imported, libamdhip64.so.7 is torch's copy and this library binds
        # to it (one HIP runtime per process, as it must be: torch's device pointers
        # and streams are passed in).  That runtime finds no device when its first
        # call comes from here, so torch initialises it first.
        torch = sys.modules.get("torch")
        if torch is not None:
            try:
                torch.cuda.is_available()
            except Exception:
                pass
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != PBRT_OK:
        raise PbrtError(status, lib().pbrt_last_error().decode())


def ptr(t) -> int:
    """Device (torch tensor) or host (numpy array) address as an int."""
    if t is None:
        return 0
    if hasattr(t, "data_ptr"):
        return int(t.data_ptr())
    return int(t.ctypes.data)


# --------------------------------------------------------------------------
# Scenes
# --------------------------------------------------------------------------
def _np_from(ptr_value, count, dtype):
    if not ptr_value or count == 0:
        return None
    buf = (C.c_byte * (count * np.dtype(dtype).itemsize)).from_address(ptr_value)
    return np.frombuffer(buf, dtype=dtype).copy()


def desc_arrays(d: SceneDesc) -> dict:
    """Host copies of a scene description's arrays (input of the oracle)."""
    meshes = []
    for i in range(d.mesh_count):
        m = d.meshes[i]
        idt = np.uint16 if m.index_stride == 2 else np.uint32
        meshes.append(dict(
            positions=_np_from(m.positions, 3 * m.vertex_count, np.float32),
            normals=_np_from(m.normals, 3 * m.vertex_count, np.float32),
            uvs=_np_from(m.uvs, 2 * m.vertex_count, np.float32),
            indices=_np_from(m.indices, m.index_count, idt),
            material=np.frombuffer(bytes(m.material), dtype=np.uint8).copy(),
            textures=np.array([m.diffuse_texture, m.normal_texture, m.specular_texture,
                               m.emissive_texture], dtype=np.int32),
        ))
    instances = [(np.array(d.instances[i].transform, dtype=np.float32), d.instances[i].mesh)
                 for i in range(d.instance_count)]
    textures = []
    for i in range(d.texture_count):
        t = d.textures[i]
        textures.append(_np_from(t.rgba8, 4 * t.width * t.height, np.uint8).reshape(t.height, t.width, 4))
    lights = np.frombuffer(bytes((Light * d.light_count).from_address(C.addressof(d.lights.contents)))
                           if d.light_count else b"", dtype=np.float32).reshape(-1, 32).copy()
    return dict(meshes=meshes, instances=instances, textures=textures, lights=lights)


class GeneratedScene:
    """A seed-generated workload (pbrt_generate_scene)."""

    def __init__(self, config: int, seed: int):
        h = C.c_void_p()
        check(lib().pbrt_generate_scene(config, seed, C.byref(h)))
        self.handle = h
        self.config = config
        self.seed = seed
        self.desc = SceneDesc()
        check(lib().pbrt_generated_scene_desc(self.handle, C.byref(self.desc)))

    def camera(self, frame: int = 0) -> CameraPose:
        pose = CameraPose()
        check(lib().pbrt_generated_scene_camera(self.handle, frame, C.byref(pose)))
        return pose

    def arrays(self) -> dict:
        """Host copies of the scene description (input of the oracle)."""
        return desc_arrays(self.desc)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().pbrt_generated_scene_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class LoadedScene:
    """A Wavefront OBJ/MTL scene read by pbrt_load_obj (the assimp front-end of the
    reference), ready for :class:`Scene`."""

    def __init__(self, path, transforms=None, bvh_builder: int = PBRT_BVH_SAH_HOST):
        opt = LoadOptions()
        opt.bvh_builder = bvh_builder
        self._transforms = None
        if transforms is not None:
            t = np.ascontiguousarray(np.asarray(transforms, dtype=np.float32).reshape(-1, 16))
            self._transforms = t
            opt.instance_count = t.shape[0]
            opt.instance_transforms = t.ctypes.data_as(C.POINTER(C.c_float))
        h = C.c_void_p()
        check(lib().pbrt_load_obj(str(path).encode(), C.byref(opt), C.byref(h)))
        self.handle = h
        self.desc = SceneDesc()
        check(lib().pbrt_loaded_scene_desc(self.handle, C.byref(self.desc)))

    def stats(self) -> LoadedSceneStats:
        st = LoadedSceneStats()
        check(lib().pbrt_loaded_scene_stats_get(self.handle, C.byref(st)))
        return st

    def arrays(self) -> dict:
        return desc_arrays(self.desc)

    def close(self):
        if getattr(self, "handle", None):
            lib().pbrt_loaded_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_image(path) -> np.ndarray:
    """An image file (PNG, TGA, PPM/PGM, PAM) as an (H, W, 4) uint8 RGBA array."""
    w, h = C.c_uint32(), C.c_uint32()
    px = C.POINTER(C.c_uint8)()
    check(lib().pbrt_image_load_rgba8(str(path).encode(), C.byref(w), C.byref(h), C.byref(px)))
    try:
        return np.ctypeslib.as_array(px, shape=(h.value, w.value, 4)).copy()
    finally:
        lib().pbrt_io_free(C.cast(px, C.c_void_p))


def write_png(path, pixels: np.ndarray, bgra: bool = False) -> None:
    """Write an (H, W, 4) uint8 array as an 8-bit RGBA PNG (`bgra` for B8G8R8A8 pixels)."""
    a = np.ascontiguousarray(pixels, dtype=np.uint8)
    assert a.ndim == 3 and a.shape[2] == 4
    check(lib().pbrt_image_write_png(str(path).encode(), a.shape[1], a.shape[0], a.ctypes.data, int(bgra)))


class Scene:
    """A packed, BVH-built scene resident in HBM (pbrt_scene_create)."""

    def __init__(self, desc: SceneDesc, keepalive=None):
        h = C.c_void_p()
        check(lib().pbrt_scene_create(C.byref(desc), C.byref(h)))
        self.handle = h
        self._keepalive = keepalive

    def info(self) -> SceneInfo:
        inf = SceneInfo()
        check(lib().pbrt_scene_get_info(self.handle, C.byref(inf)))
        return inf

    def lights(self) -> np.ndarray:
        n = self.info().light_count
        arr = (Light * n)()
        check(lib().pbrt_scene_get_lights(self.handle, arr, n))
        return np.frombuffer(bytes(arr), dtype=np.float32).reshape(n, 32).copy()

    def close(self):
        if getattr(self, "handle", None):
            lib().pbrt_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
