"""Python host mirror of the reference's operator interface for the hot path.

Each class mirrors the Java host driver the C ABI replaces (same names, same
argument meaning); device memory is torch tensors on the context's device
and every call goes through libmambart.so.  There is no CPU fallback: a
missing library or a non-zero status raises.

  Context             wrapper.core.OpenCLConfiguration (one queue per device)
  CNormalBVH          src/cl/scene/CNormalBVH.java:45-148 (host builder)
  Scene               CMesh buffers + CNormalBVH nodes/bounds on the device
  intersect_primitives / intersect_occlusion
                      IntersectPrimitives / intersectOcclusion
                      (src/cl/kernel/NormalBVH.cl:94-123,319-348)
  PrefixSumInteger    src/cl/algorithms/PrefixSumInteger.java:43-117
  CCompact            src/cl/algorithms/CCompact.java:30-89
  CAdaptiveEnvironment src/cl/algorithms/CAdaptiveEnvironment.java:37-145
  CEnvironment        src/cl/algorithms/CEnvironment.java:46-60
  bsdf_setup / bsdf_sample / bsdf_eval
                      setupBsdf / SampleBsdf / EvaluateBsdf (Bsdf.cl)
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import call

LG_SIZE = 1600 * 1600


def _ptr(x) -> Optional[int]:
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        assert x.flags["C_CONTIGUOUS"], "numpy array must be C-contiguous"
        return x.ctypes.data
    return x.data_ptr()


def _torch():
    import torch
    return torch


def _need(x, name: str, n: int, tail=(), dtype="float32"):
    """Raises unless x is a contiguous device tensor of shape (n, *tail) and
    the given dtype: the C ABI reads fixed strides (float4 rays, float[3]
    directions ...) and cannot see a tensor's shape."""
    torch = _torch()
    want = getattr(torch, dtype)
    if not isinstance(x, torch.Tensor) or x.dtype != want or not x.is_contiguous() or x.device.type != "cuda" \
            or tuple(x.shape) != (n, *tail):
        got = f"{type(x).__name__} {getattr(x, 'dtype', '')} {tuple(getattr(x, 'shape', ()))}"
        raise ValueError(f"{name}: need a contiguous {dtype} device tensor of shape {(n, *tail)}, got {got}")


def _need_rows(x, name: str, tail=(), dtype="float32"):
    """Raises unless x is a contiguous device tensor of shape (m, *tail), m >= 1."""
    torch = _torch()
    want = getattr(torch, dtype)
    ok = isinstance(x, torch.Tensor) and x.dtype == want and x.is_contiguous() and x.device.type == "cuda" \
        and x.dim() == 1 + len(tail) and x.shape[0] >= 1 and tuple(x.shape[1:]) == tuple(tail)
    if not ok:
        got = f"{type(x).__name__} {getattr(x, 'dtype', '')} {tuple(getattr(x, 'shape', ()))}"
        raise ValueError(f"{name}: need a contiguous {dtype} device tensor of shape (m, *{tuple(tail)}), got {got}")


def _check_job_inputs(job, materials, image=None, light_faces=None, num_lights=None):
    """What the bounce loop reads and writes through raw pointers: the 36-word
    material records (float32), the light list (int32, with its device count)
    and, if given, the float4 image of w*h pixels it writes."""
    _need_rows(materials, "materials", (36,))
    if (light_faces is None) != (num_lights is None):
        raise ValueError("light_faces and num_lights go together")
    if light_faces is not None:
        _need_rows(light_faces, "light_faces", (), "int32")
        if not (isinstance(num_lights, _torch().Tensor) and num_lights.numel() == 1):
            raise ValueError("num_lights: need a one-element int32 device tensor")
        _need(num_lights.reshape(1), "num_lights", 1, (), "int32")
    if image is not None:
        _need(image, "image", int(job.width) * int(job.height), (4,))


# ---------------------------------------------------------------------- context
class Context:
    """One device, one in-order stream (torch's current stream by default)."""

    def __init__(self, device: int = 0, stream=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("libmambart needs a HIP device; none is visible")
        self.device = device
        self.torch_device = torch.device("cuda", device)
        if stream is None:
            stream = torch.cuda.current_stream(self.torch_device)
        self.stream = stream
        h = C.c_void_p()
        # the library issues its work on torch's stream (handle 0 = null stream),
        # so torch copies and library kernels are ordered without extra syncs
        call("mt_ctx_create", device, C.c_void_p(stream.cuda_stream), 0, C.byref(h))
        self.handle = h
        # library objects made on this context; the C ABI requires them to be
        # destroyed before the context, so close() closes them first
        self._deps = weakref.WeakSet()

    def _own(self, obj) -> None:
        self._deps.add(obj)

    def sync(self) -> None:
        call("mt_ctx_sync", self.handle)

    def check_scans(self) -> None:
        call("mt_scan_check", self.handle)

    def close(self) -> None:
        if self.handle:
            for d in list(getattr(self, "_deps", ())):
                d.close()
            _lib.load().mt_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # helpers -------------------------------------------------------------
    def empty(self, shape, dtype):
        return _torch().empty(shape, dtype=dtype, device=self.torch_device)

    def to_device(self, arr: np.ndarray):
        return _torch().from_numpy(np.ascontiguousarray(arr)).to(self.torch_device)

    def scalar_i32(self, value: int):
        return _torch().tensor([value], dtype=_torch().int32, device=self.torch_device)


# ---------------------------------------------------------------------- host-side
class CNormalBVH:
    """CNormalBVH.build on the host (C++, multi-threaded, identical numbering)."""

    @staticmethod
    def build(points: np.ndarray, faces: np.ndarray, nthreads: int = 0):
        points = np.ascontiguousarray(points, dtype=np.float32)
        faces = np.ascontiguousarray(faces, dtype=np.int32)
        n = faces.shape[0]
        nodes = np.zeros((2 * n - 1, 7), dtype=np.int32)
        bounds = np.zeros((2 * n - 1, 8), dtype=np.float32)
        depth = C.c_int32(0)
        call("mt_bvh_build", n, _ptr(points), _ptr(faces), _ptr(nodes), _ptr(bounds),
             C.byref(depth), nthreads)
        return nodes, bounds, int(depth.value)

    @staticmethod
    def build_device(ctx: "Context", points, faces):
        """Same tree on the device from device points/faces; returns device
        (nodes int32[2N-1,7], bounds float32[2N-1,8]) and the depth."""
        torch = _torch()
        n = int(faces.shape[0])
        nodes = ctx.empty((2 * n - 1, 7), torch.int32)
        bounds = ctx.empty((2 * n - 1, 8), torch.float32)
        depth = C.c_int32(0)
        call("mt_bvh_build_device", ctx.handle, n, _ptr(points), _ptr(faces), _ptr(nodes), _ptr(bounds),
             C.byref(depth))
        return nodes, bounds, int(depth.value)


def gen_soup(seed: int, n: int, c_lo: float, c_hi: float, e_lo: float, e_hi: float,
             first_vertex: int = 0, mat: int = 0):
    points = np.zeros((3 * n, 4), dtype=np.float32)
    faces = np.zeros((n, 10), dtype=np.int32)
    call("mt_gen_soup", seed, n, c_lo, c_hi, e_lo, e_hi, first_vertex, mat, _ptr(points), _ptr(faces))
    return points, faces


def gen_room(mat: int = 0):
    points = np.zeros((8, 4), dtype=np.float32)
    faces = np.zeros((12, 10), dtype=np.int32)
    call("mt_gen_room", mat, _ptr(points), _ptr(faces))
    return points, faces


def gen_camera_rays(pos, look, up, vfov_deg: float, w: int, h: int):
    o = np.zeros((w * h, 4), dtype=np.float32)
    d = np.zeros((w * h, 4), dtype=np.float32)
    p3 = lambda v: np.ascontiguousarray(v, dtype=np.float32)
    pos, look, up = p3(pos), p3(look), p3(up)
    call("mt_gen_camera_rays", _ptr(pos), _ptr(look), _ptr(up), vfov_deg, w, h, _ptr(o), _ptr(d))
    return o, d


def gen_random_rays(seed: int, first: int, n: int, lo, hi, nthreads: int = 0):
    o = np.zeros((n, 4), dtype=np.float32)
    d = np.zeros((n, 4), dtype=np.float32)
    lo = np.ascontiguousarray(lo, dtype=np.float32)
    hi = np.ascontiguousarray(hi, dtype=np.float32)
    call("mt_gen_random_rays", seed, first, n, _ptr(lo), _ptr(hi), _ptr(o), _ptr(d), nthreads)
    return o, d


def gen_spheres(seed: int, count: int, mu: float = -2.5, sigma: float = 0.5, c_lo: float = -5.0,
                c_hi: float = 5.0, ground_y: float = -5.0, g: float = 50.0):
    nv = 2112 * count + 4
    nt = 4096 * count + 2
    points = np.zeros((nv, 4), dtype=np.float32)
    normals = np.zeros((nv, 4), dtype=np.float32)
    faces = np.zeros((nt, 10), dtype=np.int32)
    call("mt_gen_spheres", seed, count, mu, sigma, c_lo, c_hi, ground_y, g, _ptr(points),
         _ptr(normals), _ptr(faces))
    return points, normals, faces


# ---------------------------------------------------------------------- scene
class Scene:
    """Scene arrays (reference flat layout) resident on the device."""

    def __init__(self, ctx: Context, points: np.ndarray, faces: np.ndarray, nodes: np.ndarray,
                 bounds: np.ndarray, max_depth: int, normals: Optional[np.ndarray] = None,
                 uvs: Optional[np.ndarray] = None):
        torch = _torch()
        self.ctx = ctx
        self.points = ctx.to_device(np.asarray(points, dtype=np.float32))
        self.faces = ctx.to_device(np.asarray(faces, dtype=np.int32))
        self.nodes = ctx.to_device(np.asarray(nodes, dtype=np.int32))
        self.bounds = ctx.to_device(np.asarray(bounds, dtype=np.float32))
        self.normals = None if normals is None else ctx.to_device(np.asarray(normals, dtype=np.float32))
        self.uvs = None if uvs is None else ctx.to_device(np.asarray(uvs, dtype=np.float32))
        self.num_tris = int(faces.shape[0])
        desc = _lib.SceneDesc(
            _ptr(self.points), _ptr(self.normals), _ptr(self.uvs), _ptr(self.faces),
            _ptr(self.nodes), _ptr(self.bounds), self.num_tris, int(points.shape[0]), int(max_depth))
        h = C.c_void_p()
        torch.cuda.synchronize(ctx.torch_device)
        call("mt_scene_create", ctx.handle, C.byref(desc), C.byref(h))
        self.handle = h
        ctx._own(self)

    def close(self) -> None:
        if getattr(self, "handle", None):
            _lib.load().mt_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def intersect_primitives(ctx: Context, scene: Scene, o_tmin, d_tmax, count=None, mode: int = 0,
                         stats: bool = False):
    """Closest hit: returns (prim int32[n], t float32[n], stats int32[n,3] or None)."""
    torch = _torch()
    n = o_tmin.shape[0]
    _need(o_tmin, "o_tmin", n, (4,))
    _need(d_tmax, "d_tmax", n, (4,))
    if count is None:
        count = ctx.scalar_i32(n)
    prim = ctx.empty((n,), torch.int32)
    t = ctx.empty((n,), torch.float32)
    st = ctx.empty((n, 3), torch.int32) if stats else None
    call("mt_trace_closest", ctx.handle, scene.handle, _ptr(o_tmin), _ptr(d_tmax), _ptr(count), n,
         _ptr(prim), _ptr(t), _ptr(st), mode)
    return prim, t, st


def intersect_occlusion(ctx: Context, scene: Scene, o_tmin, d_tmax, count=None, mode: int = 0,
                        stats: bool = False):
    """Any hit: returns (occluded int32[n], stats or None)."""
    torch = _torch()
    n = o_tmin.shape[0]
    _need(o_tmin, "o_tmin", n, (4,))
    _need(d_tmax, "d_tmax", n, (4,))
    if count is None:
        count = ctx.scalar_i32(n)
    occ = ctx.empty((n,), torch.int32)
    st = ctx.empty((n, 3), torch.int32) if stats else None
    call("mt_trace_any", ctx.handle, scene.handle, _ptr(o_tmin), _ptr(d_tmax), _ptr(count), n,
         _ptr(occ), _ptr(st), mode)
    return occ, st


def autotune_trace(ctx: Context, scene: Scene, o_tmin, d_tmax, any_hit: bool = False, count=None, reps: int = 2):
    """Fastest packed traversal mode for this batch (mt_trace_autotune):
    returns (mode, ms per launch); every mode gives the same results."""
    torch = _torch()
    n = o_tmin.shape[0]
    _need(o_tmin, "o_tmin", n, (4,))
    _need(d_tmax, "d_tmax", n, (4,))
    if count is None:
        count = ctx.scalar_i32(n)
    prim = t = occ = None
    if any_hit:
        occ = ctx.empty((n,), torch.int32)
    else:
        prim = ctx.empty((n,), torch.int32)
        t = ctx.empty((n,), torch.float32)
    mode = C.c_int32(-1)
    ms = C.c_float(0.0)
    call("mt_trace_autotune", ctx.handle, scene.handle, _ptr(o_tmin), _ptr(d_tmax), _ptr(count), n,
         1 if any_hit else 0, _ptr(prim), _ptr(t), _ptr(occ), int(reps), C.byref(mode), C.byref(ms))
    return int(mode.value), float(ms.value)


def shade_hits(ctx: Context, scene: Scene, o_tmin, d_tmax, prim, t, count=None):
    torch = _torch()
    n = o_tmin.shape[0]
    if count is None:
        count = ctx.scalar_i32(n)
    p = ctx.empty((n, 4), torch.float32)
    nn = ctx.empty((n, 4), torch.float32)
    ng = ctx.empty((n, 4), torch.float32)
    uv = ctx.empty((n, 2), torch.float32)
    mat = ctx.empty((n,), torch.int32)
    call("mt_shade_hits", ctx.handle, scene.handle, _ptr(o_tmin), _ptr(d_tmax), _ptr(prim), _ptr(t),
         _ptr(count), n, _ptr(p), _ptr(nn), _ptr(ng), _ptr(uv), _ptr(mat))
    return p, nn, ng, uv, mat


RAY_RECORD_WORDS = 20           # Ray, Util.cl:129-138 (80 bytes)
INTERSECTION_RECORD_WORDS = 24  # Intersection, Util.cl:141-151 (96 bytes)


def ray_records(o_tmin, d_tmax):
    """Reference Ray records (float32 [n, 20] host array) for SoA rays, filled
    as initGlobalRay does (Util.cl:577-600): o, d (w 0), inv_d = 1/d (w 0),
    sign = inv_d < 0, extra 0, tMin and tMax from the w lanes."""
    o_tmin = np.asarray(o_tmin, np.float32)
    d_tmax = np.asarray(d_tmax, np.float32)
    n = o_tmin.shape[0]
    rec = np.zeros((n, RAY_RECORD_WORDS), np.float32)
    rec[:, 0:3] = o_tmin[:, :3]
    rec[:, 4:7] = d_tmax[:, :3]
    with np.errstate(divide="ignore"):
        inv = (np.float32(1.0) / d_tmax[:, :3]).astype(np.float32)
    rec[:, 8:11] = inv
    rec.view(np.int32)[:, 12:15] = (inv < 0).astype(np.int32)
    rec[:, 18] = o_tmin[:, 3]
    rec[:, 19] = d_tmax[:, 3]
    return rec


def intersect_records(ctx: Context, scene: Scene, rays, isects, count=None, mode: int = 0):
    """IntersectPrimitives over reference records (mt_intersect_primitives):
    rays float32 [n, 20] (tMax updated on hit), isects [n, 24] (float32 or
    int32 view of the 96-byte Intersection), both device tensors, in place."""
    n = rays.shape[0]
    _need(rays, "rays", n, (RAY_RECORD_WORDS,))
    _need(isects, "isects", n, (INTERSECTION_RECORD_WORDS,), dtype=str(isects.dtype).split(".")[-1]
          if str(isects.dtype) in ("torch.int32", "torch.float32") else "float32")
    if count is None:
        count = ctx.scalar_i32(n)
    call("mt_intersect_primitives", ctx.handle, scene.handle, _ptr(rays), _ptr(isects), _ptr(count), n, mode)
    return rays, isects


def intersect_occlusion_records(ctx: Context, scene: Scene, rays, hits, count=None, mode: int = 0):
    """intersectOcclusion over reference Ray records (mt_intersect_occlusion):
    rays float32 [n, 20] device tensor, hits int32 [n] written below count."""
    n = rays.shape[0]
    _need(rays, "rays", n, (RAY_RECORD_WORDS,))
    _need(hits, "hits", n, (), dtype="int32")
    if count is None:
        count = ctx.scalar_i32(n)
    call("mt_intersect_occlusion", ctx.handle, scene.handle, _ptr(rays), _ptr(hits), _ptr(count), n, mode)
    return hits


def shadow_rays(ctx: Context, o_tmin, d_tmax, prim, t, light_xyz, count=None):
    torch = _torch()
    n = o_tmin.shape[0]
    if count is None:
        count = ctx.scalar_i32(n)
    so = ctx.empty((n, 4), torch.float32)
    sd = ctx.empty((n, 4), torch.float32)
    valid = ctx.empty((n,), torch.int32)
    light = np.ascontiguousarray(light_xyz, dtype=np.float32)
    call("mt_shadow_rays", ctx.handle, _ptr(o_tmin), _ptr(d_tmax), _ptr(prim), _ptr(t), _ptr(light),
         _ptr(count), n, _ptr(so), _ptr(sd), _ptr(valid))
    return so, sd, valid


# ---------------------------------------------------------------------- D
class PrefixSumInteger:
    """PrefixSumInteger (PrefixSumInteger.java:43-117): exclusive scan + total."""

    def __init__(self, ctx: Context, cdata):
        torch = _torch()
        self.ctx = ctx
        self.cdata = cdata
        self.length = int(cdata.shape[0])
        self.clength = ctx.scalar_i32(self.length)
        self.cprefixsum = ctx.empty((self.length,), torch.int32)
        self.ctotal = ctx.scalar_i32(0)

    def execute(self) -> None:
        call("mt_scan_excl_i32", self.ctx.handle, _ptr(self.cdata), _ptr(self.cprefixsum),
             _ptr(self.clength), self.length, _ptr(self.ctotal))

    def getPrefixSum(self):
        return self.cprefixsum

    def getCTotal(self):
        return self.ctotal

    def getCount(self) -> int:
        self.ctx.check_scans()
        return int(self.ctotal.item())


class PrefixSumFloat:
    """PrefixSumFloat (PrefixSumFloat.java:49-117): exclusive float scan + total."""

    def __init__(self, ctx: Context, cdata):
        torch = _torch()
        self.ctx = ctx
        self.cdata = cdata
        self.length = int(cdata.shape[0])
        self.clength = ctx.scalar_i32(self.length)
        self.cprefixsum = ctx.empty((self.length,), torch.float32)
        self.ctotal = torch.zeros((1,), dtype=torch.float32, device=ctx.torch_device)

    def execute(self) -> None:
        call("mt_scan_excl_f32", self.ctx.handle, _ptr(self.cdata), _ptr(self.cprefixsum),
             _ptr(self.clength), self.length, _ptr(self.ctotal))

    def getPrefixSum(self):
        return self.cprefixsum

    def getCTotal(self):
        return self.ctotal


class CImage:
    """CImage (CImage.java:60-150): accumulation buffer -> tone-mapped ARGB frame."""

    def __init__(self, ctx: Context, width: int, height: int):
        torch = _torch()
        self.ctx = ctx
        self.width, self.height = width, height
        self.frameSize = width * height
        self.cFrameAccum = ctx.empty((self.frameSize, 4), torch.float32)
        self.cFrameBuffer = ctx.empty((self.frameSize, 4), torch.float32)
        self.cFrameARGB = ctx.empty((self.frameSize,), torch.int32)
        self.cTotalLogLuminance = torch.zeros((1,), dtype=torch.float32, device=ctx.torch_device)
        self.cTotalNumber = ctx.scalar_i32(0)
        self.frameCount = 1.0
        self.gamma = 2.2
        self.exposure = 0.18
        self.initBuffers()

    def initBuffers(self) -> None:
        call("mt_image_init", self.ctx.handle, _ptr(self.cFrameAccum), _ptr(self.cFrameARGB), self.frameSize)
        self.frameCount = 1.0

    def processImage(self) -> None:
        # the log-luminance totals are recomputed only for the first frames
        call("mt_image_process", self.ctx.handle, _ptr(self.cFrameAccum), float(self.frameCount), self.frameSize,
             float(self.exposure), float(self.gamma), 1 if self.frameCount < 3 else 0,
             _ptr(self.cFrameBuffer), _ptr(self.cFrameARGB), _ptr(self.cTotalLogLuminance),
             _ptr(self.cTotalNumber))

    def setGamma(self, value: float) -> None:
        self.gamma = float(value)

    def setExposure(self, value: float) -> None:
        self.exposure = float(value)

    def getFrameARGB(self):
        return self.cFrameARGB


def compact(ctx: Context, flags, flag_mode: int = 0, length=None, out_index=True,
            payloads: Sequence = ()):
    """Stable compaction; returns (out_index or None, d_count, [dst tensors])."""
    torch = _torch()
    n = int(flags.shape[0])
    if length is None:
        length = ctx.scalar_i32(n)
    idx = ctx.empty((n,), torch.int32) if out_index else None
    count = ctx.scalar_i32(0)
    arr = (_lib.Payload * max(1, len(payloads)))()
    dsts = []
    for j, src in enumerate(payloads):
        dst = torch.empty_like(src)
        eb = src.element_size() * int(np.prod(src.shape[1:], dtype=np.int64))
        arr[j] = _lib.Payload(_ptr(src), _ptr(dst), eb)
        dsts.append(dst)
    call("mt_compact", ctx.handle, _ptr(flags), flag_mode, _ptr(length), n, _ptr(idx), _ptr(count),
         arr, len(payloads))
    return idx, count, dsts


class CCompact:
    """CCompact (CCompact.java:30-89): compacts hit slots and their pixel ids.

    init(ctx, isectBuffer, pixels, totalCount), as the reference's init:
    isectBuffer is either the Intersection records themselves (a [n, 24]
    float32/int32 device tensor, 96 bytes per record; execute() compacts
    them in place with mt_compact_isects, tail records re-initialised) or,
    for SoA callers, the per-slot hit flags (int32 [n]; execute() moves the
    pixel ids and any `extra` payload streams).  Survivors keep slot order;
    totalCount receives their number, like the seven reference launches."""

    def init(self, ctx: Context, isectBuffer, pixels, totalCount, extra: Sequence = ()):
        self.ctx = ctx
        self.pixels = pixels
        self.totalCount = totalCount
        self.extra = list(extra)
        n = int(isectBuffer.shape[0])
        self.records = isectBuffer.dim() == 2 and isectBuffer.shape[1] == INTERSECTION_RECORD_WORDS
        if self.records:
            _need(pixels, "pixels", n, (), dtype="int32")
            self.isects = isectBuffer
            self.temp = _torch().empty_like(isectBuffer)   # CCompact.tempIsectBuffer
        else:
            self.hit = isectBuffer
        self.length = ctx.scalar_i32(n)

    def execute(self) -> None:
        ctx = self.ctx
        if self.records:
            n = int(self.isects.shape[0])
            call("mt_compact_isects", ctx.handle, _ptr(self.isects), _ptr(self.pixels), _ptr(self.temp),
                 _ptr(self.length), n, _ptr(self.totalCount))
            return
        streams = [self.pixels, *self.extra]
        _, cnt, dsts = compact(ctx, self.hit, 0, length=self.length, out_index=False, payloads=streams)
        for src, dst in zip(streams, dsts):
            src.copy_(dst)
        self.totalCount.copy_(cnt)


def mesh_lights(ctx: Context, scene: Scene, materials):
    torch = _torch()
    out = ctx.empty((scene.num_tris,), torch.int32)
    cnt = ctx.scalar_i32(0)
    call("mt_mesh_lights", ctx.handle, scene.handle, _ptr(materials), _ptr(out), _ptr(cnt))
    return out, cnt


# ---------------------------------------------------------------------- C
class CAdaptiveEnvironment:
    """CAdaptiveEnvironment (CAdaptiveEnvironment.java:37-145) over the 1600^2 grid."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = C.c_void_p()
        call("mt_lightgrid_create", ctx.handle, C.byref(h))
        self.handle = h
        ctx._own(self)
        torch = _torch()
        a, f, s = C.c_void_p(), C.c_void_p(), C.c_void_p()
        call("mt_lightgrid_arrays", h, C.byref(a), C.byref(f), C.byref(s))
        self._ptrs = (a.value, f.value, s.value)

    def read_arrays(self):
        """(accum, func, sat) copied to host numpy arrays (1600 x 1600 each)."""
        out = []
        for p in self._ptrs:
            a = np.empty((1600, 1600), dtype=np.float32)
            call("mt_buf_download", self.ctx.handle, _ptr(a), C.c_void_p(p), a.nbytes)
            out.append(a)
        return tuple(out)

    def write_accum(self, accum: np.ndarray) -> None:
        """Host -> grid accum (what accumLightGrid gathers on the device, Light.cl:35-38)."""
        a = np.ascontiguousarray(accum, dtype=np.float32).reshape(-1)
        assert a.size == LG_SIZE
        call("mt_buf_upload", self.ctx.handle, C.c_void_p(self._ptrs[0]), _ptr(a), a.nbytes)

    def device_arrays(self):
        """Raw device pointers (accum, func, sat) for other library calls."""
        return self._ptrs

    def clearBuffers(self) -> None:
        call("mt_lightgrid_clear", self.ctx.handle, self.handle)

    def setHDRLuminance(self, tile: np.ndarray) -> None:
        tile = np.ascontiguousarray(tile, dtype=np.float32).reshape(-1)
        assert tile.size == 512
        self.clearBuffers()
        call("mt_lightgrid_set_tile_func", self.ctx.handle, self.handle, _ptr(tile))

    def update(self) -> None:
        call("mt_lightgrid_update", self.ctx.handle, self.handle)

    def buildSAT(self) -> None:
        call("mt_lightgrid_build_sat", self.ctx.handle, self.handle)

    def setCameraPosition(self, pos) -> None:
        """CAdaptiveEnvironment.setCameraPosition (CAdaptiveEnvironment.java:61-67)."""
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.float32).reshape(-1)[:3])
        call("mt_lightgrid_set_camera", self.handle, _ptr(p))

    def sample(self, cell, u):
        torch = _torch()
        n = int(cell.shape[0])
        txy = self.ctx.empty((n, 2), torch.int32)
        uv = self.ctx.empty((n, 2), torch.float32)
        pdf = self.ctx.empty((n,), torch.float32)
        call("mt_lightgrid_sample", self.ctx.handle, self.handle, n, _ptr(cell), _ptr(u), _ptr(txy),
             _ptr(uv), _ptr(pdf))
        return txy, uv, pdf

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().mt_lightgrid_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lightgrid_cells(ctx: Context, dirs):
    torch = _torch()
    n = int(dirs.shape[0])
    cell = ctx.empty((n,), torch.int32)
    txy = ctx.empty((n, 2), torch.int32)
    call("mt_lightgrid_cells", ctx.handle, n, _ptr(dirs), _ptr(cell), _ptr(txy))
    return cell, txy


def lightgrid_seed_tile(lum: np.ndarray, w: int, h: int, samples: int = 500) -> np.ndarray:
    """The 16 x 32 tile function CAdaptiveEnvironment.setHDRLuminance seeds every
    cell with (CAdaptiveEnvironment.java:73-77: SAT.sampleToGrid(500, 16, 32),
    un-vendored): here the expected histogram of `samples` draws from the
    environment's luminance distribution, i.e. samples * (luminance summed over
    each tile's texel rectangle) / (total), entry i = tile (i % 16, i / 16).
    Rectangles as getSubRegionFromUnitBound (SAT.cl:28-44)."""
    lum = np.asarray(lum, np.float64).reshape(h, w)
    tile = np.zeros(512, np.float64)
    for ty in range(32):
        y0, y1 = int(np.float32(h) * np.float32(ty / np.float32(32))), int(np.float32(h) * np.float32((ty + 1) / np.float32(32)))
        for tx in range(16):
            x0, x1 = int(np.float32(w) * np.float32(tx / np.float32(16))), int(np.float32(w) * np.float32((tx + 1) / np.float32(16)))
            tile[ty * 16 + tx] = lum[y0:y1, x0:x1].sum()
    return (samples * tile / tile.sum()).astype(np.float32)


def sat_tiles(ctx: Context, func, nu: int, nv: int, rx: int, ry: int):
    sat = _torch().empty_like(func)
    call("mt_sat_tiles", ctx.handle, _ptr(func), _ptr(sat), nu, nv, rx, ry)
    return sat


class CEnvironment:
    """CEnvironment.setEnvironmentMap (CEnvironment.java:46-60): env RGB, luminance, SAT."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.handle = None
        ctx._own(self)

    def setEnvironmentMap(self, rgba: np.ndarray) -> None:
        torch = _torch()
        h, w = rgba.shape[0], rgba.shape[1]
        self.w, self.h = int(w), int(h)
        self.rgba = self.ctx.to_device(np.asarray(rgba, dtype=np.float32).reshape(h * w, 4))
        self.lum = self.ctx.empty((h * w,), torch.float32)
        self.sat = self.ctx.empty((h * w,), torch.float32)
        call("mt_env_lum_sat", self.ctx.handle, _ptr(self.rgba), self.w, self.h, _ptr(self.lum),
             _ptr(self.sat))
        self.close()
        hd = C.c_void_p()
        call("mt_env_create", self.ctx.handle, _ptr(self.rgba), _ptr(self.lum), _ptr(self.sat), self.w,
             self.h, C.byref(hd))
        self.handle = hd

    def illuminate(self, u):
        torch = _torch()
        n = int(u.shape[0])
        d = self.ctx.empty((n, 4), torch.float32)
        c = self.ctx.empty((n, 4), torch.float32)
        pw = self.ctx.empty((n,), torch.float32)
        tx = self.ctx.empty((n,), torch.int32)
        call("mt_env_illuminate", self.ctx.handle, self.handle, n, _ptr(u), _ptr(d), _ptr(c), _ptr(pw),
             _ptr(tx))
        return d, c, pw, tx

    def radiance(self, dirs):
        torch = _torch()
        n = int(dirs.shape[0])
        c = self.ctx.empty((n, 4), torch.float32)
        pw = self.ctx.empty((n,), torch.float32)
        tx = self.ctx.empty((n,), torch.int32)
        call("mt_env_radiance", self.ctx.handle, self.handle, n, _ptr(dirs), _ptr(c), _ptr(pw), _ptr(tx))
        return c, pw, tx

    def attachLightGrid(self, grid: Optional["CAdaptiveEnvironment"]) -> None:
        """EnvironmentLight.lightGrid (Light.cl:15-33): the grid renders with
        MT_RENDER_LIGHTGRID_ENV sample and accumulate into (None detaches)."""
        call("mt_env_attach_lightgrid", self.handle, grid.handle if grid is not None else None)
        self._grid = grid  # the environment reads the grid's arrays while attached

    def illuminate_adaptive(self, grid: "CAdaptiveEnvironment", hit_p, u):
        """illuminateEnvironmentLight1 (mt_env_illuminate_adaptive): hit_p, u float4
        [n] device tensors -> (dir, contrib, pdfw, lum, grid_index)."""
        torch = _torch()
        n = int(u.shape[0])
        _need(hit_p, "hit_p", n, (4,))
        _need(u, "u", n, (4,))
        d = self.ctx.empty((n, 4), torch.float32)
        c = self.ctx.empty((n, 4), torch.float32)
        pw = self.ctx.empty((n,), torch.float32)
        lum = self.ctx.empty((n,), torch.float32)
        gi = self.ctx.empty((n,), torch.int32)
        call("mt_env_illuminate_adaptive", self.ctx.handle, self.handle, grid.handle, n, _ptr(hit_p), _ptr(u),
             _ptr(d), _ptr(c), _ptr(pw), _ptr(lum), _ptr(gi))
        return d, c, pw, lum, gi

    def radiance_adaptive(self, grid: "CAdaptiveEnvironment", dirs, hit_p, accumulate: bool = True):
        """getRadianceEnvironmentLight1 (mt_env_radiance_adaptive), adding each
        luminance into the grid's accum when accumulate -> (contrib, pdfa, lum, grid_index)."""
        torch = _torch()
        n = int(dirs.shape[0])
        _need(dirs, "dirs", n, (4,))
        _need(hit_p, "hit_p", n, (4,))
        c = self.ctx.empty((n, 4), torch.float32)
        pa = self.ctx.empty((n,), torch.float32)
        lum = self.ctx.empty((n,), torch.float32)
        gi = self.ctx.empty((n,), torch.int32)
        call("mt_env_radiance_adaptive", self.ctx.handle, self.handle, grid.handle, n, _ptr(dirs), _ptr(hit_p),
             1 if accumulate else 0, _ptr(c), _ptr(pa), _ptr(lum), _ptr(gi))
        return c, pa, lum, gi

    def close(self):
        if self.handle:
            _lib.load().mt_env_destroy(self.handle)
            self.handle = None


def sat_sample_region(ctx: Context, func, sat, nu, nv, minx, miny, sx, sy, u):
    torch = _torch()
    n = int(u.shape[0])
    uv = ctx.empty((n, 2), torch.float32)
    off = ctx.empty((n, 2), torch.int32)
    pdf = ctx.empty((n,), torch.float32)
    call("mt_sat_sample_region", ctx.handle, _ptr(func), _ptr(sat), nu, nv, minx, miny, sx, sy, n, _ptr(u),
         _ptr(uv), _ptr(off), _ptr(pdf))
    return uv, off, pdf


# ---------------------------------------------------------------------- B
def bsdf_setup(ctx: Context, ray_d, n, ng, mat, materials):
    """ray_d, n, ng: float32[cnt, 4] (xyz used); mat: int32[cnt] (-1 = miss)."""
    torch = _torch()
    cnt = int(ray_d.shape[0])
    for x, nm in ((ray_d, "ray_d"), (n, "n"), (ng, "ng")):
        _need(x, f"bsdf_setup {nm}", cnt, (4,))
    _need(mat, "bsdf_setup mat", cnt, (), "int32")
    rec = ctx.empty((cnt, _lib.MT_BSDF_WORDS), torch.float32)
    mid = ctx.empty((cnt,), torch.int32)
    call("mt_bsdf_setup", ctx.handle, cnt, _ptr(ray_d), _ptr(n), _ptr(ng), _ptr(mat), _ptr(materials),
         _ptr(rec), _ptr(mid))
    return rec, mid


def bsdf_sample(ctx: Context, rec, mat_id, materials, u):
    torch = _torch()
    cnt = int(rec.shape[0])
    _need(rec, "bsdf_sample rec", cnt, (_lib.MT_BSDF_WORDS,))
    _need(mat_id, "bsdf_sample mat_id", cnt, (), "int32")
    _need(u, "bsdf_sample u", cnt, (3,))
    f = ctx.empty((cnt, 3), torch.float32)
    d = ctx.empty((cnt, 3), torch.float32)
    pdf = ctx.empty((cnt,), torch.float32)
    cs = ctx.empty((cnt,), torch.float32)
    call("mt_bsdf_sample", ctx.handle, cnt, _ptr(rec), _ptr(mat_id), _ptr(materials), _ptr(u), _ptr(f),
         _ptr(d), _ptr(pdf), _ptr(cs))
    return f, d, pdf, cs


def bsdf_eval(ctx: Context, rec, mat_id, materials, dirs):
    torch = _torch()
    cnt = int(rec.shape[0])
    _need(rec, "bsdf_eval rec", cnt, (_lib.MT_BSDF_WORDS,))
    _need(mat_id, "bsdf_eval mat_id", cnt, (), "int32")
    _need(dirs, "bsdf_eval dirs", cnt, (3,))
    f = ctx.empty((cnt, 3), torch.float32)
    pdf = ctx.empty((cnt,), torch.float32)
    cs = ctx.empty((cnt,), torch.float32)
    call("mt_bsdf_eval", ctx.handle, cnt, _ptr(rec), _ptr(mat_id), _ptr(materials), _ptr(dirs), _ptr(f),
         _ptr(pdf), _ptr(cs))
    return f, pdf, cs


# ---------------------------------------------------------------------- bounce loop
def render(ctx: Context, scene: Scene, materials, env: Optional["CEnvironment"], job, light_faces=None,
           num_lights=None, stats: bool = True):
    """CDeviceGI.loop for one job (see mt_render); returns (image float32[h*w,4], stats or None)."""
    torch = _torch()
    _check_job_inputs(job, materials, None, light_faces, num_lights)
    desc = render_desc(job)
    image = ctx.empty((job.width * job.height, 4), torch.float32)
    st = _lib.RenderStats() if stats else None
    call("mt_render", ctx.handle, scene.handle, _ptr(materials), env.handle if env is not None else None,
         _ptr(light_faces), _ptr(num_lights), C.byref(desc), _ptr(image), C.byref(st) if stats else None)
    return image, st


class RenderGraph:
    """One job's bounce loop captured once (mt_render_graph) and replayed per
    frame, as the reference's render thread re-runs CDeviceGI.loop every frame
    (CDeviceGI.java:214-271, 424-431).  launch(seed) writes the same image
    `render` writes for a job with that seed into `self.image`."""

    def __init__(self, ctx: Context, scene: Scene, materials, env: Optional["CEnvironment"], job,
                 light_faces=None, num_lights=None):
        torch = _torch()
        _check_job_inputs(job, materials, None, light_faces, num_lights)
        self.ctx = ctx
        self.image = ctx.empty((job.width * job.height, 4), torch.float32)
        # the graph keeps the addresses of everything it reads
        self._keep = (scene, materials, env, light_faces, num_lights)
        h = C.c_void_p()
        call("mt_render_graph_create", ctx.handle, scene.handle, _ptr(materials),
             env.handle if env is not None else None, _ptr(light_faces), _ptr(num_lights),
             C.byref(render_desc(job)), _ptr(self.image), C.byref(h))
        self.handle = h
        ctx._own(self)

    def launch(self, seed: int) -> None:
        call("mt_render_graph_launch", self.handle, int(seed) & 0xFFFFFFFF)

    def stats(self) -> "_lib.RenderStats":
        st = _lib.RenderStats()
        call("mt_render_graph_stats", self.handle, C.byref(st))
        return st

    def close(self) -> None:
        if self.handle:
            _lib.load().mt_render_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_desc(job) -> "_lib.RenderDesc":
    """mt_render_desc of a job (any object with the RenderJob fields)."""
    desc = _lib.RenderDesc()
    desc.pos[:] = list(job.pos)
    desc.look[:] = list(job.look)
    desc.up[:] = list(job.up)
    desc.vfov_deg = job.vfov_deg
    desc.width, desc.height, desc.spp = job.width, job.height, job.spp
    desc.max_bounces, desc.rr_start, desc.seed = job.max_bounces, job.rr_start, job.seed
    desc.mode, desc.flags = job.mode, job.flags
    return desc


class Batch:
    """Multi-job wavefront engine (mt_batch): jobs share one set of queues.

    The reference renders one job at a time per device (CDeviceGI.start ->
    execute -> loop, CDeviceGI.java:190-204, 424-431); here any number of
    jobs are in flight together and each job's image equals its plain
    `render` bit for bit.
    """

    def __init__(self, ctx: Context, queue_capacity: int = 1 << 22, arena_bytes: int = 8 << 30,
                 max_jobs: int = 1024, policy: int = _lib.MT_BATCH_OLDEST_FIRST):
        self.ctx = ctx
        h = C.c_void_p()
        call("mt_batch_create", ctx.handle, int(queue_capacity), int(arena_bytes), int(max_jobs), C.byref(h))
        self.handle = h
        ctx._own(self)
        call("mt_batch_set_policy", h, int(policy))
        self._keep = {}  # tag -> objects the device reads until the job completes
        self._done = (C.c_int64 * 256)()

    @staticmethod
    def job_bytes(job) -> int:
        b = C.c_uint64(0)
        call("mt_batch_job_bytes", C.byref(render_desc(job)), C.byref(b))
        return int(b.value)

    def submit(self, tag: int, scene: Scene, materials, env: Optional["CEnvironment"], job, image,
               light_faces=None, num_lights=None) -> None:
        """Queues a job; `image` (device float32[h*w,4]) is written when it completes.
        A tag may not be reused while its job is still in flight."""
        if int(tag) in self._keep:
            raise ValueError(f"tag {tag} is still in flight")
        _check_job_inputs(job, materials, image, light_faces, num_lights)
        d = _lib.JobDesc()
        d.scene = scene.handle
        d.materials = _ptr(materials)
        d.env = env.handle if env is not None else None
        d.light_faces = _ptr(light_faces)
        d.d_num_lights = _ptr(num_lights)
        d.render = render_desc(job)
        d.image = _ptr(image)
        d.tag = int(tag)
        call("mt_batch_submit", self.handle, C.byref(d))
        self._keep[int(tag)] = (scene, materials, env, image, light_faces, num_lights)

    def step(self):
        """One wavefront iteration; returns (completed tags, busy)."""
        n = C.c_int32(0)
        busy = C.c_int32(0)
        call("mt_batch_step", self.handle, len(self._done), self._done, C.byref(n), C.byref(busy))
        tags = [int(self._done[i]) for i in range(n.value)]
        for t in tags:
            self._keep.pop(t, None)
        return tags, bool(busy.value)

    def run(self):
        """Steps until every submitted job is complete; returns tags in completion order."""
        order = []
        busy = True
        while busy:
            tags, busy = self.step()
            order.extend(tags)
        return order

    def stats(self) -> "_lib.BatchStats":
        s = _lib.BatchStats()
        call("mt_batch_get_stats", self.handle, C.byref(s))
        return s

    def cancel(self, tag: int) -> bool:
        """Cancels a pending or running job (mt_batch_cancel); its tag is
        reported by a later step once its paths have drained, its image is
        not written.  Returns False when no pending or running job has the tag."""
        found = C.c_int32(0)
        call("mt_batch_cancel", self.handle, int(tag), C.byref(found))
        return bool(found.value)

    def set_iterations(self, iterations: int) -> None:
        """Wavefront iterations per step (mt_batch_set_iterations)."""
        call("mt_batch_set_iterations", self.handle, int(iterations))

    def set_graphs(self, on: bool) -> None:
        """Replay each step's iterations as a captured hipGraph (default) or
        launch their kernels one by one (mt_batch_set_graphs)."""
        call("mt_batch_set_graphs", self.handle, 1 if on else 0)

    def set_trace_mode(self, mode: int) -> None:
        """Closest-hit traversal of the shared queue: _lib.MT_TRACE_PACKED_STACK
        (one ray per lane), _lib.MT_TRACE_PACKED_POOL (wave pools) or
        _lib.MT_BATCH_TRACE_SPLIT (pools for the continuing paths, one ray per
        lane for the new camera rays) (mt_batch_set_trace_mode)."""
        call("mt_batch_set_trace_mode", self.handle, int(mode))

    def close(self) -> None:
        if self.handle:
            _lib.load().mt_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------- formats
class OBJParser:
    """OBJParser.readString (un-vendored; used at TracerAPI.java:321-322) into
    the reference's mesh arrays: (points float32[V,4], normals float32[N,4],
    uvs float32[T,2], faces int32[F,10], number of material groups)."""

    def readString(self, text):
        data = text.encode() if isinstance(text, str) else bytes(text)
        cnt = _lib.ObjCounts()
        call("mt_obj_parse", data, len(data), C.byref(cnt), None, None, None, None)
        points = np.zeros((cnt.num_points, 4), np.float32)
        normals = np.zeros((cnt.num_normals, 4), np.float32)
        uvs = np.zeros((cnt.num_uvs, 2), np.float32)
        faces = np.zeros((cnt.num_tris, 10), np.int32)
        if cnt.num_points == 0:
            return points, normals, uvs, faces, 0
        call("mt_obj_parse", data, len(data), C.byref(cnt), _ptr(points), _ptr(normals) if cnt.num_normals else None,
             _ptr(uvs) if cnt.num_uvs else None, _ptr(faces))
        return points, normals, uvs, faces, int(cnt.num_groups)

    def read(self, path):
        with open(path, "rb") as f:
            return self.readString(f.read())


class HDRBitmapReader:
    """HDRBitmapReader (un-vendored; used at UserInterfaceFXMLController.java:
    176-179): Radiance RGBE bytes -> float32[h, w, 4] (a = 1)."""

    def decode(self, data: bytes) -> np.ndarray:
        w, h = C.c_int32(0), C.c_int32(0)
        call("mt_hdr_decode", data, len(data), C.byref(w), C.byref(h), None)
        out = np.zeros((h.value, w.value, 4), np.float32)
        call("mt_hdr_decode", data, len(data), C.byref(w), C.byref(h), _ptr(out))
        return out

    def load(self, path) -> np.ndarray:
        with open(path, "rb") as f:
            return self.decode(f.read())
