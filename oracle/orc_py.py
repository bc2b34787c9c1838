"""ORACLE — test infrastructure only.

numpy front-end of liboracle.so, the plain-C CPU restatement of the
reference kernels (see orc.h).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module, and only as the checker or
the timed CPU baseline; the product library never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> str:
    res = subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, I32, I64, F = C.c_void_p, C.c_int32, C.c_int64, C.c_float
        sig = {
            "orc_bvh_build": [I32, P, P, P, P, P],
            "orc_xm_eval": [I32, I64, P, P, P],
            "orc_lg_fold": [P, P],
            "orc_trace_closest": [P, P, P, P, I64, P, P, P, P, P, P, I32],
            "orc_trace_any": [P, P, P, P, I64, P, P, P, P, P, I32],
            "orc_trace_brute": [P, P, I32, I64, P, P, P, P],
            "orc_shade_hits": [P, P, P, P, I64, P, P, P, P, P, P, P, P, P],
            "orc_shadow_rays": [I64, P, P, P, P, P, P, P, P],
            "orc_scan_excl_i32": [P, I64, P, P],
            "orc_compact_stable": [P, I64, P, P],
            "orc_sat_tiles": [P, P, I32, I32, I32, I32],
            "orc_lg_set_tile_func": [P, P],
            "orc_lg_sample": [P, P, I32, P, P, P, P, P],
            "orc_lg_cell_from_dirs": [I32, P, P, P],
            "orc_env_lum": [P, I32, I32, P],
            "orc_env_sat": [P, I32, I32, P],
            "orc_sat_sample_region": [P, P, I32, I32, I32, I32, I32, I32, I32, P, P, P, P],
            "orc_env_illuminate": [P, P, P, I32, I32, I32, P, P, P, P, P],
            "orc_env_radiance": [P, P, P, I32, I32, I32, P, P, P, P],
            "orc_env_illuminate_adaptive": [P, P, P, I32, I32, P, P, P, I32, P, P, P, P, P, P, P, P],
            "orc_env_radiance_adaptive": [P, P, P, I32, I32, P, P, I32, P, P, P, P, P, P],
            "orc_bsdf_setup": [I32, P, P, P, P, P, P, P],
            "orc_bsdf_sample": [I32, P, P, P, P, P, P, P, P],
            "orc_bsdf_eval": [I32, P, P, P, P, P, P, P],
            "orc_ggx_D": [P, F, F],
            "orc_ggx_lambda": [P, F, F],
            "orc_ggx_sample_vndf": [P, F, F, F, F, P],
            "orc_ggx_pdf_vndf": [P, P, F, F],
            "orc_wang_hash": [I32],
            "orc_get_random": [P],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int32
        L.orc_ggx_D.restype = C.c_float
        L.orc_ggx_lambda.restype = C.c_float
        L.orc_ggx_pdf_vndf.restype = C.c_float
        L.orc_ggx_sample_vndf.restype = None
        L.orc_wang_hash.restype = C.c_uint32
        L.orc_get_random.restype = C.c_float
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ok(rc, name):
    if rc != 0:
        raise RuntimeError(f"{name} returned {rc}")


# ---------------------------------------------------------------- A
def bvh_build(points, faces):
    points, faces = _f32(points), _i32(faces)
    n = faces.shape[0]
    nodes = np.zeros((2 * n - 1, 7), np.int32)
    bounds = np.zeros((2 * n - 1, 8), np.float32)
    d = C.c_int32(0)
    _ok(lib().orc_bvh_build(n, _p(points), _p(faces), _p(nodes), _p(bounds), C.byref(d)), "orc_bvh_build")
    return nodes, bounds, int(d.value)


def trace_closest(points, faces, nodes, bounds, o, d, nthreads=8, ray_stats=False):
    points, faces, nodes, bounds, o, d = _f32(points), _i32(faces), _i32(nodes), _f32(bounds), _f32(o), _f32(d)
    n = o.shape[0]
    prim = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    rs = np.zeros((n, 3), np.int32) if ray_stats else None
    tot = np.zeros(3, np.int64)
    _ok(lib().orc_trace_closest(_p(points), _p(faces), _p(nodes), _p(bounds), n, _p(o), _p(d), _p(prim), _p(t),
                                _p(rs), _p(tot), nthreads), "orc_trace_closest")
    return prim, t, tot, rs


def trace_any(points, faces, nodes, bounds, o, d, nthreads=8, ray_stats=False):
    points, faces, nodes, bounds, o, d = _f32(points), _i32(faces), _i32(nodes), _f32(bounds), _f32(o), _f32(d)
    n = o.shape[0]
    occ = np.zeros(n, np.int32)
    rs = np.zeros((n, 3), np.int32) if ray_stats else None
    tot = np.zeros(3, np.int64)
    _ok(lib().orc_trace_any(_p(points), _p(faces), _p(nodes), _p(bounds), n, _p(o), _p(d), _p(occ), _p(rs),
                            _p(tot), nthreads), "orc_trace_any")
    return occ, tot, rs


def trace_brute(points, faces, o, d):
    points, faces, o, d = _f32(points), _i32(faces), _f32(o), _f32(d)
    n = o.shape[0]
    prim = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    _ok(lib().orc_trace_brute(_p(points), _p(faces), faces.shape[0], n, _p(o), _p(d), _p(prim), _p(t)),
        "orc_trace_brute")
    return prim, t


def shade_hits(points, normals, uvs, faces, o, d, prim, t):
    points, faces, o, d, prim, t = _f32(points), _i32(faces), _f32(o), _f32(d), _i32(prim), _f32(t)
    normals = None if normals is None else _f32(normals)
    uvs = None if uvs is None else _f32(uvs)
    n = o.shape[0]
    p = np.zeros((n, 4), np.float32)
    nn = np.zeros((n, 4), np.float32)
    ng = np.zeros((n, 4), np.float32)
    uv = np.zeros((n, 2), np.float32)
    mat = np.zeros(n, np.int32)
    _ok(lib().orc_shade_hits(_p(points), _p(normals), _p(uvs), _p(faces), n, _p(o), _p(d), _p(prim), _p(t),
                             _p(p), _p(nn), _p(ng), _p(uv), _p(mat)), "orc_shade_hits")
    return p, nn, ng, uv, mat


def shadow_rays(o, d, prim, t, light):
    o, d, prim, t, light = _f32(o), _f32(d), _i32(prim), _f32(t), _f32(light)
    n = o.shape[0]
    so = np.zeros((n, 4), np.float32)
    sd = np.zeros((n, 4), np.float32)
    valid = np.zeros(n, np.int32)
    _ok(lib().orc_shadow_rays(n, _p(o), _p(d), _p(prim), _p(t), _p(light), _p(so), _p(sd), _p(valid)),
        "orc_shadow_rays")
    return so, sd, valid


# ---------------------------------------------------------------- D
def scan_excl(x):
    x = _i32(x)
    out = np.zeros_like(x)
    tot = C.c_int32(0)
    _ok(lib().orc_scan_excl_i32(_p(x), x.size, _p(out), C.byref(tot)), "orc_scan_excl_i32")
    return out, int(tot.value)


def scan_excl_f32(x):
    """orc_scan_excl_f32: the device float scan's exact summation order."""
    x = _f32(x)
    out = np.zeros_like(x)
    tot = C.c_float(0)
    f = lib().orc_scan_excl_f32
    f.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    f.restype = C.c_int32
    _ok(f(_p(x), x.size, _p(out), C.byref(tot)), "orc_scan_excl_f32")
    return out, float(tot.value)


def image_process(accum, frame_count, exposure, gamma, argb=None, totals=None):
    """orc_image_process: CImage.processImage on float32 accum[n,4].  `totals`
    = (total_log, count) to reuse (update_totals off) or None to compute.
    Returns (frame[n,4], argb[n], (total_log, count))."""
    accum = _f32(accum)
    n = accum.shape[0]
    frame = np.zeros((n, 4), np.float32)
    argb = np.full(n, image_init_argb(), np.int32) if argb is None else _i32(argb).copy()
    tl = C.c_float(0.0 if totals is None else totals[0])
    tc = C.c_int32(0 if totals is None else totals[1])
    f = lib().orc_image_process
    f.argtypes = [C.c_void_p, C.c_float, C.c_int32, C.c_float, C.c_float, C.c_int32, C.c_void_p, C.c_void_p,
                  C.c_void_p, C.c_void_p]
    f.restype = C.c_int32
    _ok(f(_p(accum), frame_count, n, exposure, gamma, 1 if totals is None else 0, _p(frame), _p(argb),
          C.byref(tl), C.byref(tc)), "orc_image_process")
    return frame, argb, (float(tl.value), int(tc.value))


def image_init_argb():
    f = lib().orc_image_init_argb
    f.restype = C.c_int32
    return int(f())


def argb(r, g, b, a):
    f = lib().orc_argb
    f.argtypes = [C.c_float] * 4
    f.restype = C.c_int32
    return int(f(r, g, b, a))


def compact_stable(pred):
    pred = _i32(pred)
    out = np.zeros(pred.size, np.int32)
    cnt = C.c_int32(0)
    _ok(lib().orc_compact_stable(_p(pred), pred.size, _p(out), C.byref(cnt)), "orc_compact_stable")
    return out[: cnt.value], int(cnt.value)


# ---------------------------------------------------------------- C
def sat_tiles(func, nu, nv, rx, ry):
    func = _f32(func)
    sat = np.zeros_like(func)
    _ok(lib().orc_sat_tiles(_p(func), _p(sat), nu, nv, rx, ry), "orc_sat_tiles")
    return sat


def lg_set_tile_func(tile):
    func = np.zeros(1600 * 1600, np.float32)
    tile = _f32(tile).reshape(-1)
    _ok(lib().orc_lg_set_tile_func(_p(func), _p(tile)), "orc_lg_set_tile_func")
    return func


def lg_sample(func, sat, cell, u):
    func, sat, cell, u = _f32(func), _f32(sat), _i32(cell), _f32(u)
    n = cell.size
    txy = np.zeros((n, 2), np.int32)
    uv = np.zeros((n, 2), np.float32)
    pdf = np.zeros(n, np.float32)
    _ok(lib().orc_lg_sample(_p(func), _p(sat), n, _p(cell), _p(u), _p(txy), _p(uv), _p(pdf)), "orc_lg_sample")
    return txy, uv, pdf


def lg_cells(dirs):
    dirs = _f32(dirs)
    n = dirs.shape[0]
    cell = np.zeros(n, np.int32)
    txy = np.zeros((n, 2), np.int32)
    _ok(lib().orc_lg_cell_from_dirs(n, _p(dirs), _p(cell), _p(txy)), "orc_lg_cell_from_dirs")
    return cell, txy


def env_lum_sat(rgba, w, h):
    rgba = _f32(rgba)
    lum = np.zeros(w * h, np.float32)
    sat = np.zeros(w * h, np.float32)
    _ok(lib().orc_env_lum(_p(rgba), w, h, _p(lum)), "orc_env_lum")
    _ok(lib().orc_env_sat(_p(lum), w, h, _p(sat)), "orc_env_sat")
    return lum, sat


def sat_sample_region(func, sat, nu, nv, minx, miny, sx, sy, u):
    func, sat, u = _f32(func), _f32(sat), _f32(u)
    n = u.shape[0]
    uv = np.zeros((n, 2), np.float32)
    off = np.zeros((n, 2), np.int32)
    pdf = np.zeros(n, np.float32)
    _ok(lib().orc_sat_sample_region(_p(func), _p(sat), nu, nv, minx, miny, sx, sy, n, _p(u), _p(uv), _p(off),
                                    _p(pdf)), "orc_sat_sample_region")
    return uv, off, pdf


def env_illuminate(rgba, lum, sat, w, h, u):
    rgba, lum, sat, u = _f32(rgba), _f32(lum), _f32(sat), _f32(u)
    n = u.shape[0]
    d = np.zeros((n, 4), np.float32)
    c = np.zeros((n, 4), np.float32)
    pw = np.zeros(n, np.float32)
    tx = np.zeros(n, np.int32)
    _ok(lib().orc_env_illuminate(_p(rgba), _p(lum), _p(sat), w, h, n, _p(u), _p(d), _p(c), _p(pw), _p(tx)),
        "orc_env_illuminate")
    return d, c, pw, tx


def env_radiance(rgba, lum, sat, w, h, dirs):
    rgba, lum, sat, dirs = _f32(rgba), _f32(lum), _f32(sat), _f32(dirs)
    n = dirs.shape[0]
    c = np.zeros((n, 4), np.float32)
    pw = np.zeros(n, np.float32)
    tx = np.zeros(n, np.int32)
    _ok(lib().orc_env_radiance(_p(rgba), _p(lum), _p(sat), w, h, n, _p(dirs), _p(c), _p(pw), _p(tx)),
        "orc_env_radiance")
    return c, pw, tx


def env_illuminate_adaptive(rgba, lum, sat, w, h, lg_func, lg_sat, cam, hit, u):
    """illuminateEnvironmentLight1: (dir, contrib, pdfw, lum, grid index, texel xy in the map)."""
    rgba, lum, sat, lg_func, lg_sat = _f32(rgba), _f32(lum), _f32(sat), _f32(lg_func), _f32(lg_sat)
    cam, hit, u = _f32(cam), _f32(hit), _f32(u)
    n = u.shape[0]
    d = np.zeros((n, 4), np.float32)
    c = np.zeros((n, 4), np.float32)
    pw = np.zeros(n, np.float32)
    lo = np.zeros(n, np.float32)
    gi = np.zeros(n, np.int32)
    txy = np.zeros((n, 2), np.int32)
    _ok(lib().orc_env_illuminate_adaptive(_p(rgba), _p(lum), _p(sat), w, h, _p(lg_func), _p(lg_sat), _p(cam), n,
                                          _p(hit), _p(u), _p(d), _p(c), _p(pw), _p(lo), _p(gi), _p(txy)),
        "orc_env_illuminate_adaptive")
    return d, c, pw, lo, gi, txy


def env_radiance_adaptive(rgba, lum, sat, w, h, cam, dirs, hit, accum=None):
    """getRadianceEnvironmentLight1: (contrib, pdfa, lum, grid index); adds lum
    into `accum` (float32[2 560 000], in place) in index order when given."""
    rgba, lum, sat, cam, dirs, hit = _f32(rgba), _f32(lum), _f32(sat), _f32(cam), _f32(dirs), _f32(hit)
    if accum is not None:
        assert accum.dtype == np.float32 and accum.flags.c_contiguous
    n = dirs.shape[0]
    c = np.zeros((n, 4), np.float32)
    pa = np.zeros(n, np.float32)
    lo = np.zeros(n, np.float32)
    gi = np.zeros(n, np.int32)
    _ok(lib().orc_env_radiance_adaptive(_p(rgba), _p(lum), _p(sat), w, h, _p(accum) if accum is not None else None,
                                        _p(cam), n, _p(dirs), _p(hit), _p(c), _p(pa), _p(lo), _p(gi)),
        "orc_env_radiance_adaptive")
    return c, pa, lo, gi


# ---------------------------------------------------------------- B
def bsdf_setup(ray_d, n, ng, mat, materials):
    ray_d, n, ng, mat, materials = _f32(ray_d), _f32(n), _f32(ng), _i32(mat), _f32(materials)
    cnt = ray_d.shape[0]
    rec = np.zeros((cnt, 20), np.float32)
    mid = np.zeros(cnt, np.int32)
    _ok(lib().orc_bsdf_setup(cnt, _p(ray_d), _p(n), _p(ng), _p(mat), _p(materials), _p(rec), _p(mid)),
        "orc_bsdf_setup")
    return rec, mid


def bsdf_sample(rec, mid, materials, u):
    rec, mid, materials, u = _f32(rec), _i32(mid), _f32(materials), _f32(u)
    cnt = rec.shape[0]
    f = np.zeros((cnt, 3), np.float32)
    d = np.zeros((cnt, 3), np.float32)
    pdf = np.zeros(cnt, np.float32)
    cs = np.zeros(cnt, np.float32)
    _ok(lib().orc_bsdf_sample(cnt, _p(rec), _p(mid), _p(materials), _p(u), _p(f), _p(d), _p(pdf), _p(cs)),
        "orc_bsdf_sample")
    return f, d, pdf, cs


def bsdf_eval(rec, mid, materials, dirs):
    rec, mid, materials, dirs = _f32(rec), _i32(mid), _f32(materials), _f32(dirs)
    cnt = rec.shape[0]
    f = np.zeros((cnt, 3), np.float32)
    pdf = np.zeros(cnt, np.float32)
    cs = np.zeros(cnt, np.float32)
    _ok(lib().orc_bsdf_eval(cnt, _p(rec), _p(mid), _p(materials), _p(dirs), _p(f), _p(pdf), _p(cs)),
        "orc_bsdf_eval")
    return f, pdf, cs


def ggx_D(wh, ax, ay):
    wh = _f32(wh)
    return float(lib().orc_ggx_D(_p(wh), ax, ay))


def ggx_lambda(w, ax, ay):
    w = _f32(w)
    return float(lib().orc_ggx_lambda(_p(w), ax, ay))


def ggx_sample_vndf(v, ax, ay, u1, u2):
    out = np.zeros(3, np.float32)
    v = _f32(v)
    lib().orc_ggx_sample_vndf(_p(v), ax, ay, u1, u2, _p(out))
    return out


def ggx_pdf_vndf(v, h, ax, ay):
    v, h = _f32(v), _f32(h)
    return float(lib().orc_ggx_pdf_vndf(_p(v), _p(h), ax, ay))


def wang_hash(seed: int) -> int:
    return int(lib().orc_wang_hash(int(np.int32(np.uint32(seed & 0xFFFFFFFF)))))


def get_random(state):
    st = np.ascontiguousarray(state, dtype=np.int32)
    r = lib().orc_get_random(_p(st))
    return float(r), st


# ---------------------------------------------------------------- bounce loop
def camera_basis(pos, look, up, vfov_deg, w, h):
    cam = np.zeros(14, np.float32)
    L = lib()
    L.orc_camera_basis.argtypes = [C.c_void_p] * 3 + [C.c_float, C.c_int32, C.c_int32, C.c_void_p]
    pos, look, up = _f32(pos), _f32(look), _f32(up)  # keep the arrays alive across the call
    _ok(L.orc_camera_basis(_p(pos), _p(look), _p(up), vfov_deg, w, h, _p(cam)),
        "orc_camera_basis")
    return cam


class LightGrid:
    """The oracle's light grid (LightGrid.cl:9-37; CAdaptiveEnvironment.java):
    accum / func / sat float32[1600 * 1600], the camera position, and `pend`,
    the fixed-point accumulation (uint64, units of 2^-20) of the guided
    render mode, which fold() adds into accum at the end of a frame."""

    def __init__(self, tile=None, cam=(0.0, 0.0, 0.0)):
        self.accum = np.zeros(1600 * 1600, np.float32)
        self.func = np.zeros(1600 * 1600, np.float32)
        self.sat = np.zeros(1600 * 1600, np.float32)
        self.pend = np.zeros(1600 * 1600, np.uint64)
        self.cam = np.asarray(cam, np.float32).copy()
        if tile is not None:
            self.set_tile_func(tile)

    def set_tile_func(self, tile):
        """setToAllFunc + the row/column SATs (CAdaptiveEnvironment.setHDRLuminance :69-88)."""
        self.func = lg_set_tile_func(tile)
        self.sat = sat_tiles(self.func, 1600, 1600, 16, 32)

    def update(self):
        """CAdaptiveEnvironment.update (:135-145): func <- accum, then the SATs."""
        self.func = self.accum.copy()
        self.sat = sat_tiles(self.func, 1600, 1600, 16, 32)

    def fold(self):
        _ok(lib().orc_lg_fold(_p(self.pend), _p(self.accum)), "orc_lg_fold")


def render_paths(points, normals, faces, nodes, bounds, materials, env, job, pids, light_faces=None,
                 nthreads=8, grid=None, grid_mode=0):
    """Traces the listed path ids one at a time; returns (L float32[max_pid+1,4], rays[2]).
    env = (rgba[h*w,4], lum, sat, w, h) or None.  grid (a LightGrid) with
    grid_mode 1 or 2: the light-grid guided environment (reference pdf /
    tile pick's pdf applied); its pend gains the paths' accumulation."""
    L = lib()
    P, I32, I64, U32 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32
    L.orc_render_paths_lg.argtypes = [P, P, P, P, P, P, P, P, P, I32, I32, P, I32, P, I32, I32, I32, I32, I32,
                                      U32, I32, I64, P, P, P, I32, P, P, P, P, I32]
    L.orc_render_paths_lg.restype = C.c_int32
    cam = camera_basis(job.pos, job.look, job.up, job.vfov_deg, job.width, job.height)
    pids = _i32(pids)
    out = np.zeros((int(pids.max()) + 1 if pids.size else 0, 4), np.float32)
    rays = np.zeros(2, np.int64)
    if env is None:
        rgba = lum = sat = None
        ew = eh = 0
    else:
        rgba, lum, sat, ew, eh = env
        rgba, lum, sat = _f32(rgba), _f32(lum), _f32(sat)
    normals = None if normals is None else _f32(normals)
    lf = None if light_faces is None else _i32(light_faces)
    points, faces, nodes, bounds, materials = _f32(points), _i32(faces), _i32(nodes), _f32(bounds), _f32(materials)
    g = grid if grid_mode else None
    _ok(L.orc_render_paths_lg(_p(points), _p(normals), _p(faces), _p(nodes), _p(bounds),
                              _p(materials), _p(rgba), _p(lum), _p(sat), ew, eh, _p(lf),
                              0 if lf is None else lf.size, _p(cam), job.width, job.height, job.spp,
                              job.max_bounces, job.rr_start, job.seed, job.flags, pids.size, _p(pids), _p(out),
                              _p(rays), nthreads, _p(g.func) if g else None, _p(g.sat) if g else None,
                              _p(g.cam) if g else None, _p(g.pend) if g else None, int(grid_mode) if g else 0),
        "orc_render_paths_lg")
    return out, rays


def path_pixel(pids, w, h, spp):
    """(x, y) of each path id under the bounce loop's convention (orc_path_pixel)."""
    L = lib()
    L.orc_path_pixel.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32)]
    L.orc_path_pixel.restype = None
    x, y = C.c_int32(), C.c_int32()
    out = np.empty((len(pids), 2), np.int32)
    for i, p in enumerate(pids):
        L.orc_path_pixel(int(p), w, h, spp, C.byref(x), C.byref(y))
        out[i] = (x.value, y.value)
    return out


def pixel_first_paths(w, h, spp):
    """First path id of every pixel (row-major), the inverse of path_pixel:
    pixel rank k in bands of 8 rows walked column by column, times spp."""
    y, x = np.divmod(np.arange(w * h, dtype=np.int64), w)
    fb = h // 8
    band = y // 8
    k = np.where(band < fb, band * 8 * w + x * 8 + (y % 8), fb * 8 * w + x * (h - 8 * fb) + (y - 8 * fb))
    return k * spp


def resolve(L_paths, w, h, spp):
    """Per-pixel mean over samples in sample order, float32 (k_resolve's order)."""
    first = pixel_first_paths(w, h, spp)
    acc = np.zeros((w * h, 3), np.float32)
    for s in range(spp):
        acc = (acc + L_paths[first + s, :3]).astype(np.float32)
    return (acc * np.float32(1.0 / spp)).astype(np.float32)


XM_SIN, XM_COS, XM_ACOS, XM_ATAN2, XM_POW5 = range(5)


def xm_eval(kind, a, b=None):
    """The shared transcendental functions (mt_xmath.h) on float32 arrays."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(a if b is None else b, np.float32)
    out = np.empty_like(a)
    _ok(lib().orc_xm_eval(kind, a.size, _p(a), _p(b), _p(out)), "orc_xm_eval")
    return out
