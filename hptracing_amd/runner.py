"""Python face of the GPU renderer: ``TraceRunner`` mirrors hp::TraceRunner
(src/hp/trace_runner.h:21-66) over the C ABI of libhptracing.so.

``TraceRunner(scene)`` uploads a ``KDTree`` to HBM once (the reference's
constructor, trace_runner.cpp:41-108); ``run(...)`` takes the argument list
of ``TraceRunner::run`` (trace_runner.h:53-65) and returns the RGBA8 image
that ``get_image`` produces.  ``cast`` exposes the ray-cast stage alone.
Errors raise ``HptError`` (the reference throws cl::Error).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _native as N
from .scene import KDTree


def _f3(v: Sequence[float]):
    return (C.c_float * 3)(*[float(x) for x in v])


def make_view(view_p, top_dir, right_dir, width, height, sample_x, sample_y,
              supersample_x=1, supersample_y=1, background_color=(0, 0, 0), sample=10,
              depth=6, disable_diffuse=0, brightness=1.0, seed=1) -> N.View:
    v = N.View()
    v.view, v.up, v.right = _f3(view_p), _f3(top_dir), _f3(right_dir)
    v.width, v.height = float(width), float(height)
    v.sample_x, v.sample_y = int(sample_x), int(sample_y)
    v.supersample_x, v.supersample_y = int(supersample_x), int(supersample_y)
    v.background = _f3(background_color)
    v.sample, v.depth, v.disable_diffuse = int(sample), int(depth), int(disable_diffuse)
    v.brightness = float(brightness)
    v.seed = int(seed) & ((1 << 64) - 1)
    return v


class TraceRunner:
    """Device-resident scene plus render / cast entry points."""

    def __init__(self, scene: KDTree):
        self.scene = scene
        self._h = C.c_void_p()
        d = scene.desc()
        N.check_device(N.device().hpt_scene_create(C.byref(d), C.byref(self._h)))
        st = N.SceneStats()
        N.check_device(N.device().hpt_scene_stats_get(self._h, C.byref(st)))
        self.stats = {k: getattr(st, k) for k, _ in N.SceneStats._fields_}
        self.last_stats: Optional[N.RenderStats] = None

    @property
    def handle(self):
        return self._h

    def get_tuning(self) -> dict:
        t = N.Tuning()
        N.check_device(N.device().hpt_scene_get_tuning(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in N.Tuning._fields_}

    def set_tuning(self, **kw) -> None:
        cur = self.get_tuning()
        cur.update(kw)
        t = N.Tuning(**cur)
        N.check_device(N.device().hpt_scene_set_tuning(self._h, C.byref(t)))

    def autotune(self, v: N.View, max_rays: int = 0) -> dict:
        """Pick the fastest tuning on renders of `v` with one launch group's
        passes (min(sample, 16)), shrunk to at most max_rays camera rays x
        passes (0: the library's default, 2^24)."""
        t = N.Tuning()
        N.check_device(N.device().hpt_autotune(self._h, C.byref(v), max_rays, C.byref(t)))
        return {k: getattr(t, k) for k, _ in N.Tuning._fields_}

    def autotune_cast(self, org, dir, max_rays: int = 0) -> dict:
        """Pick the fastest tuning for hpt_cast on a ray batch ((n,3) float32,
        numpy or torch); uses at most max_rays of the rays (0: 2^22)."""
        import torch
        org_t = torch.as_tensor(org, dtype=torch.float32).cuda().contiguous()
        dir_t = torch.as_tensor(dir, dtype=torch.float32).cuda().contiguous()
        torch.cuda.synchronize()
        t = N.Tuning()
        N.check_device(N.device().hpt_autotune_cast(self._h, org_t.data_ptr(), dir_t.data_ptr(),
                                                    org_t.shape[0], max_rays, C.byref(t)))
        return {k: getattr(t, k) for k, _ in N.Tuning._fields_}

    def set_arithmetic(self, rocm_opencl: bool) -> None:
        """dot/cross of the intersection test: the fused forms of ROCm's
        OpenCL builtins (default; the reference kernel's bits) or the literal
        formulas with every product and sum rounded."""
        N.check_device(N.device().hpt_scene_set_arithmetic(self._h, 1 if rocm_opencl else 0))

    def run(self, view_p, top_dir, right_dir, width, height, sample_x, sample_y,
            supersample_x=1, supersample_y=1, background_color=(0, 0, 0), sample=10,
            depth=6, disable_diffuse=0, brightness=1.0, seed=1, return_accum=False):
        v = make_view(view_p, top_dir, right_dir, width, height, sample_x, sample_y,
                      supersample_x, supersample_y, background_color, sample, depth,
                      disable_diffuse, brightness, seed)
        return self.render(v, return_accum=return_accum)

    def render(self, v: N.View, return_accum=False, time_kernels=False, count_work=False,
               pass_at_a_time=False):
        rgba = np.zeros((v.sample_y, v.sample_x, 4), np.uint8)
        R = v.sample_x * v.supersample_x * v.sample_y * v.supersample_y
        accum = np.zeros((R, 3), np.float32) if return_accum else None
        stats = N.RenderStats()
        flags = (N.HPT_RENDER_TIME_KERNELS if time_kernels else 0) | \
                (N.HPT_RENDER_COUNT_WORK if count_work else 0) | \
                (N.HPT_RENDER_PASS_AT_A_TIME if pass_at_a_time else 0)
        N.check_device(N.device().hpt_render_ex(
            self._h, C.byref(v), rgba.ctypes.data, accum.ctypes.data if return_accum else None,
            C.byref(stats), flags))
        self.last_stats = stats
        return (rgba, accum) if return_accum else rgba

    def render_batch(self, views, return_accum=False, use_graph=True):
        """Render several views in shared launches (hpt_batch_render).
        Returns a list of RGBA images (and accumulators)."""
        n = len(views)
        arr = (N.View * n)(*views)
        imgs = [np.zeros((v.sample_y, v.sample_x, 4), np.uint8) for v in views]
        accs = [np.zeros((v.sample_x * v.supersample_x * v.sample_y * v.supersample_y, 3),
                         np.float32) for v in views] if return_accum else None
        rgba_ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in imgs])
        acc_ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in accs]) if return_accum else None
        stats = N.RenderStats()
        flags = 0 if use_graph else N.HPT_RENDER_NO_GRAPH
        N.check_device(N.device().hpt_batch_render(self._h, arr, n, rgba_ptrs, acc_ptrs,
                                                   C.byref(stats), flags))
        self.last_stats = stats
        return (imgs, accs) if return_accum else imgs

    def cast(self, org, dir, counted=False):
        """Cast rays given as (n,3) float32 arrays (numpy or torch, any device);
        returns (tri_id int32, t float32[, counts (n,4) int32]) as torch CUDA
        tensors.  Runs on torch's current stream."""
        import torch
        dev = torch.device("cuda")
        org_t = torch.as_tensor(org, dtype=torch.float32).to(dev).contiguous()
        dir_t = torch.as_tensor(dir, dtype=torch.float32).to(dev).contiguous()
        n = org_t.shape[0]
        if org_t.shape != (n, 3) or dir_t.shape != (n, 3):
            raise ValueError("org and dir must be (n, 3)")
        ids = torch.empty(n, dtype=torch.int32, device=dev)
        t = torch.empty(n, dtype=torch.float32, device=dev)
        stream = torch.cuda.current_stream().cuda_stream
        if counted:
            counts = torch.empty((n, N.HPT_CAST_COUNTS), dtype=torch.int32, device=dev)
            N.check_device(N.device().hpt_cast_counted(
                self._h, org_t.data_ptr(), dir_t.data_ptr(), n, ids.data_ptr(), t.data_ptr(),
                counts.data_ptr(), stream))
            return ids, t, counts
        N.check_device(N.device().hpt_cast(
            self._h, org_t.data_ptr(), dir_t.data_ptr(), n, ids.data_ptr(), t.data_ptr(), stream))
        return ids, t

    def close(self):
        if self._h:
            N.device().hpt_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RenderServer:
    """Continuous-batching server over a resident scene (hpt_server_*).

    ``submit(view)`` queues a request and returns its id; ``wait(id)``
    blocks until the image is in the returned buffer.  Requests submitted
    while others render join them at the next pass.  A request submitted
    with ``progressive=True`` also publishes its image after every macro
    pass; ``peek(id)`` returns the latest one and its pass count."""

    def __init__(self, runner: TraceRunner, device_budget_bytes: int = 4 << 30):
        self.runner = runner
        self._h = C.c_void_p()
        opts = N.ServerOpts(device_budget_bytes)
        N.check_device(N.device().hpt_server_create(runner.handle, C.byref(opts),
                                                    C.byref(self._h)))
        self._buffers = {}

    def submit(self, v: N.View, return_accum=False, progressive=False) -> int:
        img = np.zeros((v.sample_y, v.sample_x, 4), np.uint8)
        acc = np.zeros((v.sample_x * v.supersample_x * v.sample_y * v.supersample_y, 3),
                       np.float32) if return_accum else None
        rid = C.c_int64()
        flags = N.HPT_SERVE_PROGRESSIVE if progressive else 0
        N.check_device(N.device().hpt_server_submit_ex(
            self._h, C.byref(v), img.ctypes.data, acc.ctypes.data if acc is not None else None,
            flags, C.byref(rid)))
        self._buffers[rid.value] = (img, acc)
        return rid.value

    def peek(self, rid: int, min_passes: int = 0, timeout_s: float = -1.0):
        """Latest published image of a progressive request once it has at
        least `min_passes` passes (or has finished): (rgba or None, passes).
        With a timeout that expires first, the latest image so far."""
        img, _ = self._buffers[rid]
        out = np.zeros_like(img)
        passes = C.c_int32(0)
        rc = N.device().hpt_server_peek(self._h, rid, min_passes, timeout_s, out.ctypes.data,
                                        C.byref(passes))
        if rc not in (N.HPT_OK, N.HPT_ERR_TIMEOUT):
            N.check_device(rc)
        return (out if passes.value > 0 else None), passes.value

    def wait(self, rid: int, timeout_s: float = -1.0):
        """Returns (rgba, accum or None, info dict)."""
        info = N.RequestInfo()
        N.check_device(N.device().hpt_server_wait(self._h, rid, timeout_s, C.byref(info)))
        img, acc = self._buffers.pop(rid)
        return img, acc, {"queue_ms": info.queue_ms, "latency_ms": info.latency_ms,
                          "camera_rays": info.camera_rays}

    def stats(self) -> dict:
        st = N.ServerStats()
        N.check_device(N.device().hpt_server_stats_get(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in N.ServerStats._fields_}

    def close(self):
        if self._h:
            N.device().hpt_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
