"""Helpers shared by GPU tests: device images and C-ABI calls on torch's stream."""
from __future__ import annotations

import ctypes as C

import numpy as np

from vulkanpbrt_amd import camera, capi


def stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def frame_push_constants(gen, width, height, frame, prev_frame=None, sample=0):
    view, proj = camera.pose_matrices(gen.camera(frame), width, height)
    pview, _ = camera.pose_matrices(gen.camera(prev_frame if prev_frame is not None else frame), width, height)
    return camera.push_constants(view, proj, pview, frame, sample), view, proj


def gbuffer(torch, width, height):
    dev = "cuda"
    return dict(
        depth=torch.zeros((height, width), dtype=torch.float32, device=dev),
        normal=torch.zeros((height, width, 2), dtype=torch.float32, device=dev),
        material=torch.zeros((height, width, 4), dtype=torch.uint8, device=dev),
        albedo=torch.zeros((height, width, 4), dtype=torch.uint8, device=dev),
        illumination=torch.zeros((height, width, 4), dtype=torch.float32, device=dev),
    )


_workspaces = {}


def trace_workspace(torch, width, height):
    key = (width, height)
    if key not in _workspaces:
        n = C.c_size_t()
        capi.check(capi.lib().pbrt_trace_workspace_size(width, height, C.byref(n)))
        _workspaces[key] = torch.zeros(n.value, dtype=torch.uint8, device="cuda")
    return _workspaces[key]


TRACERS = ("megakernel", "persistent", "wavefront")


def tracer_kind(name):
    """C ABI tracer enum of a tracer name used in test parameters."""
    return {"megakernel": capi.TRACER_MEGAKERNEL, "persistent": capi.TRACER_PERSISTENT,
            "wavefront": capi.TRACER_WAVEFRONT}[name]


def trace(torch, scene, pc, img, width, height, max_depth=2, tracer=capi.TRACER_MEGAKERNEL,
          output=capi.OUTPUT_DEMOD_ILLUMINATION_FLOAT, light_sampling=capi.LIGHT_SAMPLE_AUTO, counters=None,
          block_threads=0, queue_paths=0, min_depth=0):
    p = capi.TraceParams(width, height, max_depth, min_depth, light_sampling, output, 1, tracer, block_threads,
                         queue_paths)
    ws = trace_workspace(torch, width, height) if tracer == capi.TRACER_WAVEFRONT else None
    ti = capi.TraceImages(*(capi.ptr(img[k]) for k in ("depth", "normal", "material", "albedo", "illumination")),
                          capi.ptr(ws), ws.numel() if ws is not None else 0)
    capi.check(capi.lib().pbrt_trace(scene.handle, C.byref(p), C.byref(pc), C.byref(ti),
                                     capi.ptr(counters) or None, stream(torch)))


def to_np(img):
    return {k: v.cpu().numpy() for k, v in img.items()}


def camera_rays(pc, width, height):
    """Primary rays of createRay (camera.glsl:6-16) in float64 -> float32, n x 8."""
    iv = np.array(pc.view_inverse[:], np.float64).reshape(4, 4).T
    ip = np.array(pc.proj_inverse[:], np.float64).reshape(4, 4).T
    ys, xs = np.mgrid[0:height, 0:width]
    cx = (xs + 0.5) / width * 2 - 1
    cy = (ys + 0.5) / height * 2 - 1
    d = np.stack([cx, cy, np.ones_like(cx), np.ones_like(cx)], -1) @ ip.T
    d = d[..., :3] / np.linalg.norm(d[..., :3], axis=-1, keepdims=True)
    dw = np.concatenate([d, np.zeros_like(d[..., :1])], -1) @ iv.T
    o = iv[:3, 3]
    rays = np.zeros((height * width, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 1e-3
    rays[:, 4:7] = dw[..., :3].reshape(-1, 3)
    rays[:, 7] = 1e4
    return rays
