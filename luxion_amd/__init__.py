"""luxion_amd -- MI355X-native wavefront path tracer (Python host binding).

The renderer is the native library libluxion_mi355x.so (gfx950 HIP kernels
behind the C ABI in include/luxion_mi355x.h); this module is a thin ctypes
binding.  It mirrors the reference renderer's host interface
(seabiscuit-iv/luxion):

  reference                                  here
  Scene(file, envmap)  (src/scene/scene.cpp)  Scene.load_glb(ctx, file, envmap) / Scene(ctx, SceneData)
  pathtraceInit(Scene*)                       Render(scene, dmis=..., emis=...)
  pathtrace(pbo, frame, iter)                 Render.pathtrace(iter)
  copyImageToHost()                           Render.copy_image_to_host()  (per-pixel radiance sum)
  -dmis / -emis / -e                          dmis= / emis= / envmap=

There is no CPU fallback: every call goes to the GPU library and raises
LuxionError if it is missing or fails.
"""
import ctypes as C

import numpy as np

from . import _ffi as F
from ._ffi import LuxionError, check, load
from . import scenes

__all__ = ["Context", "Scene", "Render", "LuxionError", "scenes", "load", "trace_closest", "trace_occluded"]


class Context:
    """One HIP device + stream with an HBM budget (0 = all free memory)."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        self.lib = load()
        self.h = C.c_void_p()
        check(self.lib.lx_context_create(device, hbm_budget_bytes, C.byref(self.h)))

    def close(self):
        if self.h:
            self.lib.lx_context_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def bytes_in_use(self):
        return self.lib.lx_context_bytes_in_use(self.h)

    def tuning(self):
        t = F.lx_tuning()
        check(self.lib.lx_tuning_get(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in F.lx_tuning._fields_}

    def autotune(self, jobs, json_path=None):
        """Time the batched form over `jobs` for each launch shape (and BVH
        width, when the jobs' scenes can be rebuilt); keep the fastest.
        Returns the chosen tuning."""
        arr = (F.lx_job * len(jobs))(*jobs)
        t = F.lx_tuning()
        check(self.lib.lx_autotune(self.h, arr, len(jobs), json_path.encode() if json_path else None, C.byref(t)))
        return {k: getattr(t, k) for k, _ in F.lx_tuning._fields_}

    def load_tuning(self, json_path):
        check(self.lib.lx_tuning_load(self.h, json_path.encode()))

    def set_tuning(self, **kw):
        t = F.lx_tuning()
        check(self.lib.lx_tuning_get(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        check(self.lib.lx_tuning_set(self.h, C.byref(t)))


def build_opts(bvh_width=0, leaf_size=4, builder=2, sah_bins=16, cdf_precision=F.LX_CDF_FP64):
    o = F.lx_build_opts()
    o.bvh_width, o.leaf_size, o.builder, o.sah_bins, o.cdf_precision = (bvh_width, leaf_size, builder, sah_bins,
                                                                        cdf_precision)
    return o


class Scene:
    """A device-resident scene (meshes, instances, materials, textures, HDRI,
    emitter / HDRI CDFs, two-level BVH)."""

    def __init__(self, ctx, data=None, build=True, **build_kw):
        self.ctx = ctx
        self.lib = ctx.lib
        self.h = C.c_void_p()
        self.data = data
        if data is not None:
            desc, keep = data.to_desc()
            check(self.lib.lx_scene_create(ctx.h, C.byref(desc), C.byref(self.h)))
            del keep
            if build:
                self.build(**build_kw)

    @classmethod
    def load_glb(cls, ctx, path, envmap=None, build=True, **build_kw):
        s = cls(ctx)
        check(s.lib.lx_scene_load_glb(ctx.h, path.encode(), envmap.encode() if envmap else None, C.byref(s.h)))
        if build:
            s.build(**build_kw)
        return s

    def build(self, **kw):
        o = build_opts(**kw)
        check(self.lib.lx_scene_build_accel(self.h, C.byref(o)))

    def info(self):
        i = F.lx_scene_info()
        check(self.lib.lx_scene_get_info(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in F.lx_scene_info._fields_}

    def camera(self):
        c = F.lx_camera_desc()
        d = C.c_int32()
        check(self.lib.lx_scene_get_camera(self.h, C.byref(c), C.byref(d)))
        return dict(width=c.width, height=c.height, position=np.array(c.position[:], np.float32),
                    view=np.array(c.view[:], np.float32), up=np.array(c.up[:], np.float32),
                    right=np.array(c.right[:], np.float32),
                    pixel_length=np.array(c.pixel_length[:], np.float32)), d.value

    def env_cdfs(self):
        i = self.info()
        m = np.empty(i["env_height"], np.float32)
        c = np.empty(i["env_height"] * i["env_width"], np.float32)
        check(self.lib.lx_scene_read_env_cdfs(self.h, m.ctypes.data, c.ctypes.data))
        return m, c

    def emitter_cdfs(self):
        n = C.c_uint32()
        check(self.lib.lx_scene_read_emitter_cdfs(self.h, None, None, C.byref(n)))
        inst = np.empty(n.value, np.uint32)
        pre = np.empty(n.value, np.float32)
        check(self.lib.lx_scene_read_emitter_cdfs(self.h, inst.ctypes.data, pre.ctypes.data, C.byref(n)))
        tris = []
        for k in range(n.value):
            m = C.c_uint32()
            check(self.lib.lx_scene_read_instance_tri_cdf(self.h, k, None, C.byref(m)))
            t = np.empty(m.value, np.float32)
            check(self.lib.lx_scene_read_instance_tri_cdf(self.h, k, t.ctypes.data, C.byref(m)))
            tris.append(t)
        return inst, pre, tris

    def close(self):
        if self.h:
            self.lib.lx_scene_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rays_array(origins, dirs, tmax=1e16):
    n = len(origins)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = origins
    r[:, 4:7] = dirs
    r[:, 7] = tmax
    return np.ascontiguousarray(r)


def trace_closest(scene, rays, alpha=True, iteration=1):
    """rays: (n, 8) float32 [ox oy oz tmin dx dy dz tmax]; returns a
    structured array (t, instance, prim, u, v)."""
    rays = np.ascontiguousarray(rays, np.float32)
    out = np.zeros(len(rays), dtype=[("t", "<f4"), ("instance", "<u4"), ("prim", "<u4"), ("u", "<f4"),
                                     ("v", "<f4")])
    check(scene.lib.lx_trace_closest(scene.h, rays.ctypes.data, len(rays), int(alpha), iteration, out.ctypes.data))
    return out


def trace_occluded(scene, rays, alpha=True, iteration=1):
    rays = np.ascontiguousarray(rays, np.float32)
    out = np.zeros(len(rays), np.uint8)
    check(scene.lib.lx_trace_occluded(scene.h, rays.ctypes.data, len(rays), int(alpha), iteration, out.ctypes.data))
    return out


def trace_emitting(scene, rays, iteration=1):
    """The direct-light query of the tail kernel: trace_closest's hit (alpha
    rules on) where its instance's material emits, a miss elsewhere."""
    rays = np.ascontiguousarray(rays, np.float32)
    out = np.zeros(len(rays), dtype=[("t", "<f4"), ("instance", "<u4"), ("prim", "<u4"), ("u", "<f4"),
                                     ("v", "<f4")])
    check(scene.lib.lx_trace_emitting(scene.h, rays.ctypes.data, len(rays), iteration, out.ctypes.data))
    return out


class Render:
    """pathtraceInit equivalent: per-render device buffers (path state, hit
    records, accumulation)."""

    def __init__(self, scene, max_depth=0, dmis=False, emis=False, envmap_intensity=1.0, seed=0, color_mode=2,
                 camera=None, material_debug_mode=0, shading_model=0):
        """material_debug_mode: LX_DEBUG_* view (one bounce, no tonemap);
        shading_model: LX_SHADING_UBER (default) or LX_SHADING_LEGACY."""
        self.scene = scene
        self.lib = scene.lib
        self.h = C.c_void_p()
        o = F.lx_render_opts()
        o.max_depth, o.dmis, o.emis = max_depth, int(dmis), int(emis)
        o.envmap_intensity, o.seed, o.color_mode = envmap_intensity, seed, color_mode
        o.material_debug_mode, o.shading_model = material_debug_mode, shading_model
        self._cam = scenes.camera_desc(camera) if camera is not None else None
        o.camera = C.pointer(self._cam) if self._cam is not None else None
        check(self.lib.lx_render_create(scene.h, C.byref(o), C.byref(self.h)))
        cam, depth = scene.camera()
        self.width = camera["width"] if camera else cam["width"]
        self.height = camera["height"] if camera else cam["height"]
        self.depth = 1 if material_debug_mode else (max_depth if max_depth > 0 else depth)
        self.next_iter = 1

    def iterate(self, first_iter, n):
        check(self.lib.lx_render_iterate(self.h, first_iter, n))
        self.next_iter = first_iter + n

    def pathtrace(self, iteration=None):
        """One sample per pixel for the given 1-based iteration (reference
        pathtrace(pbo, frame, iter))."""
        it = self.next_iter if iteration is None else iteration
        self.iterate(it, 1)

    def sync(self):
        check(self.lib.lx_render_sync(self.h))

    def accum(self):
        out = np.empty((self.height, self.width, 3), np.float32)
        spp = C.c_uint32()
        check(self.lib.lx_render_read_accum(self.h, out.ctypes.data, C.byref(spp)))
        return out, spp.value

    def copy_image_to_host(self):
        return self.accum()[0]

    def ldr(self, tonemap=F.LX_TONEMAP_ACES):
        out = np.empty((self.height, self.width, 4), np.uint8)
        check(self.lib.lx_render_read_ldr(self.h, out.ctypes.data, tonemap))
        return out

    def reset(self):
        check(self.lib.lx_render_reset(self.h))
        self.next_iter = 1

    def save_png(self, path):
        """Averaged, ACES-tonemapped, gamma 0.45, horizontally mirrored 8-bit
        PNG, as the reference's saveImage writes it."""
        check(self.lib.lx_render_save_png(self.h, path.encode()))

    def set_timing(self, events=True, count_work=False):
        flags = (F.LX_TIMING_EVENTS if events else 0) | (F.LX_TIMING_COUNT_WORK if count_work else 0)
        check(self.lib.lx_render_set_timing(self.h, flags))

    def stats(self):
        s = F.lx_render_stats()
        check(self.lib.lx_render_stats_get(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in F.lx_render_stats._fields_}

    def debug_bounce(self, iteration):
        N = self.width * self.height
        D = self.depth
        surv = np.zeros(D * N, np.uint32)
        sc = np.zeros(D, np.uint32)
        rec = np.zeros(D * N * 8, np.float32)
        rc = np.zeros(D, np.uint32)
        check(self.lib.lx_debug_bounce(self.h, iteration, surv.ctypes.data, sc.ctypes.data, rec.ctypes.data,
                                       rc.ctypes.data))
        survivors = [surv[d * N:d * N + sc[d]].copy() for d in range(D)]
        records = [rec[d * N * 8:(d * N + rc[d]) * 8].reshape(-1, 8).copy() for d in range(D)]
        return survivors, records

    def close(self):
        if self.h:
            self.lib.lx_render_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- glTF input
def parse_glb(path, envmap=None, env_head=0):
    """What lx_scene_load_glb builds from a .glb (+ EXR), read back on the
    host without a device: meshes, instances, materials, textures, camera."""
    lib = load()
    d = F.lx_scene_desc()
    head = np.zeros(max(1, env_head), np.float32)
    check(lib.lx_debug_parse_glb(path.encode(), envmap.encode() if envmap else None, C.byref(d),
                                 head.ctypes.data, env_head))

    def arr(ptr, n, dt):
        if not ptr or n == 0:
            return None
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(dt)), (n,)).copy()

    meshes = []
    for k in range(d.num_meshes):
        m = d.meshes[k]
        meshes.append((arr(m.positions, 3 * m.num_verts, C.c_float).reshape(-1, 3),
                       None if not m.normals else arr(m.normals, 3 * m.num_verts, C.c_float).reshape(-1, 3),
                       None if not m.uvs else arr(m.uvs, 2 * m.num_verts, C.c_float).reshape(-1, 2),
                       arr(m.indices, 3 * m.num_tris, C.c_uint32)))
    insts = [(d.instances[k].mesh, d.instances[k].material, np.array(d.instances[k].transform, np.float32).reshape(3, 4))
             for k in range(d.num_instances)]
    mats = []
    for k in range(d.num_materials):
        m = d.materials[k]
        e = {f: getattr(m, f) for f, _ in F.lx_material_desc._fields_}
        for f in ("base_color", "emission_color"):
            e[f] = tuple(e[f])
        for f in ("albedo_xf", "normal_xf", "emissive_xf", "metallic_roughness_xf"):
            x = e[f]
            e[f] = dict(offset=tuple(x.offset), scale=tuple(x.scale), rotation=x.rotation)
        mats.append(e)
    texs = []
    for k in range(d.num_textures):
        t = d.textures[k]
        if t.format == F.LX_TEX_RGBA16:
            a = arr(t.data, 4 * t.width * t.height, C.c_uint16)
        else:
            a = arr(t.data, 4 * t.width * t.height, C.c_uint8)
        texs.append((a.reshape(t.height, t.width, 4), 16 if t.format == F.LX_TEX_RGBA16 else 8))
    c = d.camera
    cam = dict(width=c.width, height=c.height, position=np.array(c.position), view=np.array(c.view),
               up=np.array(c.up), right=np.array(c.right), pixel_length=tuple(c.pixel_length))
    env = None
    if d.envmap:
        env = (d.envmap[0].width, d.envmap[0].height, head[:env_head].copy())
    return dict(meshes=meshes, instances=insts, materials=mats, textures=texs, camera=cam,
                trace_depth=d.trace_depth, env=env)


# ------------------------------------------------------------ batched form
def make_job(scene, camera, spp, max_depth=0, dmis=False, emis=False, envmap_intensity=1.0, seed=0, first_iter=1,
             deadline_ms=0.0, material_debug_mode=0, shading_model=0):
    j = F.lx_job()
    j.scene = scene.h.value
    j.camera = scenes.camera_desc(camera)
    j.spp, j.max_depth, j.dmis, j.emis = spp, max_depth, int(dmis), int(emis)
    j.envmap_intensity, j.seed, j.first_iter, j.deadline_ms = envmap_intensity, seed, first_iter, deadline_ms
    j.material_debug_mode, j.shading_model = material_debug_mode, shading_model
    return j


def batch_run(ctx, jobs):
    """Render every job to completion, all jobs sharing each bounce launch.
    Returns one dict per job: image (H, W, 3 radiance sums), spp, status,
    timestamps (context clock, ms)."""
    n = len(jobs)
    arr = (F.lx_job * n)(*jobs)
    res = (F.lx_job_result * n)()
    imgs = []
    for k, j in enumerate(jobs):
        img = np.zeros((j.camera.height, j.camera.width, 3), np.float32)
        imgs.append(img)
        res[k].rgb_sum = img.ctypes.data
    check(ctx.lib.lx_batch_run(ctx.h, arr, n, res))
    return [dict(image=imgs[k], spp=res[k].spp_done, status=res[k].status, submit_ms=res[k].submit_ms,
                 start_ms=res[k].start_ms, finish_ms=res[k].finish_ms, rays=res[k].rays) for k in range(n)]


# ------------------------------------------------------------- served form
class Server:
    """Continuous-batching render server (one request thread in the
    library).  Results are copied into per-job numpy images owned here."""

    def __init__(self, ctx, hbm_budget_bytes=0, max_jobs=64, max_paths=0, max_depth=16):
        self.ctx = ctx
        self.lib = ctx.lib
        self.h = C.c_void_p()
        o = F.lx_server_opts()
        o.hbm_budget_bytes, o.max_jobs_in_flight, o.spp_per_round = hbm_budget_bytes, max_jobs, 1
        o.max_paths_per_round, o.max_depth = max_paths, max_depth
        check(self.lib.lx_server_start(ctx.h, C.byref(o), C.byref(self.h)))
        self._jobs = {}

    def submit(self, job):
        img = np.zeros((job.camera.height, job.camera.width, 3), np.float32)
        res = F.lx_job_result()
        res.rgb_sum = img.ctypes.data
        jid = C.c_uint64()
        check(self.lib.lx_server_submit(self.h, C.byref(job), C.byref(res), C.byref(jid)))
        self._jobs[jid.value] = (job, res, img)
        return jid.value

    def poll(self, jid, timeout_ms=-1.0):
        d = C.c_int32()
        check(self.lib.lx_server_poll(self.h, jid, timeout_ms, C.byref(d)))
        return bool(d.value)

    def result(self, jid):
        job, res, img = self._jobs[jid]
        return dict(image=img, spp=res.spp_done, status=res.status, submit_ms=res.submit_ms, start_ms=res.start_ms,
                    finish_ms=res.finish_ms)

    def cancel(self, jid):
        check(self.lib.lx_server_cancel(self.h, jid))

    def inject_budget_failure(self, n=1):
        check(self.lib.lx_server_inject_budget_failure(self.h, n))

    def stats(self):
        s = F.lx_server_stats()
        check(self.lib.lx_server_stats_get(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in F.lx_server_stats._fields_}

    def stop(self):
        if self.h:
            check(self.lib.lx_server_stop(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
