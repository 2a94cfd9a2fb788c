"""MI355X-native HairMSNN hair path tracer: host-side Python API.

The reference's entry points are the executables render_path_tracing,
render_hair_msnn and render_nrc driven by one scene JSON
(render_path_tracing.cu:817-841, render_hair_msnn.cu:1146-1175). The same
operations are available here over the C ABI of libhmsnn.so:

    img = render_path_tracing("scene.json")            # mean image, H x W x 4
    img = render_hair_msnn("scene.json", BETA=1)

Images follow the reference's framebuffer order: row 0 is the bottom row.
There is no CPU fallback; without the HIP library or a GPU every call raises
HmsnnError.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

from ._lib import (BsdfQuery, HairParams, HmsnnError, Info, MsnnStats, NrcStats, check, declared_symbols, load,
                   LIB_PATH)

__all__ = [
    "HmsnnError", "DeviceBuffer", "Scene", "parse_scene", "render_path_tracing", "render_hair_msnn",
    "hair_bsdf_eval", "hair_bsdf_sample", "device_info", "LIB_PATH", "declared_symbols",
]


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def device_info() -> tuple[int, str]:
    lib = load()
    n = C.c_int(0)
    name = C.create_string_buffer(256)
    check(lib.hmsnn_device_count(C.byref(n), name, 256))
    return n.value, name.value.decode()


class DeviceBuffer:
    """A device allocation owned from Python (hmsnn_dev_alloc / hmsnn_dev_free)."""

    def __init__(self, nbytes: int):
        self.lib = load()
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(self.lib.hmsnn_dev_alloc(max(self.nbytes, 1), C.byref(p)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(b.lib.hmsnn_memcpy_h2d(b.ptr, _ptr(a), a.nbytes))
        return b

    def to_numpy(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested view is larger than the allocation")
        if out.nbytes:
            check(self.lib.hmsnn_memcpy_d2h(_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            check(self.lib.hmsnn_dev_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                self.lib.hmsnn_dev_free(self.ptr)
        except Exception:
            pass


def _hair_params(beta_m, beta_n, alpha, gains=(1.0, 1.0, 1.0, 1.0)) -> HairParams:
    p = HairParams()
    p.beta_m, p.beta_n, p.alpha = beta_m, beta_n, alpha
    for i in range(4):
        p.gain[i] = gains[i]
    return p


def hair_bsdf_eval(queries: np.ndarray, beta_m=0.3, beta_n=0.3, alpha=0.0, gains=(1, 1, 1, 1)):
    """f*|cos| (n x 3) and pdf (n) for n x 10 float32 queries (wo, wi, h, sigma_a)."""
    q = np.ascontiguousarray(queries, dtype=np.float32)
    n = q.shape[0]
    lib = load()
    dq, df, dp = DeviceBuffer.from_numpy(q), DeviceBuffer(12 * n), DeviceBuffer(4 * n)
    prm = _hair_params(beta_m, beta_n, alpha, gains)
    check(lib.hmsnn_hair_bsdf_eval_dev(C.byref(prm), dq.ptr, n, df.ptr, dp.ptr, None))
    check(lib.hmsnn_dev_synchronize())
    return df.to_numpy(np.float32, (n, 3)), dp.to_numpy(np.float32, (n,))


def hair_bsdf_sample(queries: np.ndarray, u4: np.ndarray, beta_m=0.3, beta_n=0.3, alpha=0.0, gains=(1, 1, 1, 1)):
    """Samples wi from four uniforms per query; returns wi, f*|cos|, pdf, lobe."""
    q = np.ascontiguousarray(queries, dtype=np.float32)
    u = np.ascontiguousarray(u4, dtype=np.float32)
    n = q.shape[0]
    lib = load()
    dq, du = DeviceBuffer.from_numpy(q), DeviceBuffer.from_numpy(u)
    dwi, df, dp, dl = DeviceBuffer(12 * n), DeviceBuffer(12 * n), DeviceBuffer(4 * n), DeviceBuffer(4 * n)
    prm = _hair_params(beta_m, beta_n, alpha, gains)
    check(lib.hmsnn_hair_bsdf_sample_dev(C.byref(prm), dq.ptr, du.ptr, n, dwi.ptr, df.ptr, dp.ptr, dl.ptr, None))
    check(lib.hmsnn_dev_synchronize())
    return (dwi.to_numpy(np.float32, (n, 3)), df.to_numpy(np.float32, (n, 3)), dp.to_numpy(np.float32, (n,)),
            dl.to_numpy(np.int32, (n,)))


def _info_dict(info: Info) -> dict:
    d = {}
    for name, _ in Info._fields_:
        v = getattr(info, name)
        if hasattr(v, "__len__"):
            v = np.array([list(x) if hasattr(x, "__len__") else x for x in v], dtype=np.float64)
        d[name] = v
    return d


def parse_scene(json_text: str) -> dict:
    """Host-only parse with the reference's defaults and failure rules."""
    info = Info()
    check(load().hmsnn_parse_scene(json_text.encode(), C.byref(info)))
    return _info_dict(info)


def generate_host(json_text: str | dict) -> tuple[np.ndarray, np.ndarray | None]:
    """Procedural segments (n x 8) and environment map (H x W x 4) on the host."""
    if isinstance(json_text, dict):
        json_text = json.dumps(json_text)
    lib = load()
    ns, w, h = C.c_int(), C.c_int(), C.c_int()
    t = json_text.encode()
    check(lib.hmsnn_generate_host(t, None, 0, C.byref(ns), None, 0, C.byref(w), C.byref(h)))
    seg = np.empty((ns.value, 8), np.float32)
    env = np.empty((h.value, w.value, 4), np.float32) if w.value else None
    check(lib.hmsnn_generate_host(t, _ptr(seg) if seg.size else None, seg.size, C.byref(ns),
                                  _ptr(env) if env is not None else None, 0 if env is None else env.size,
                                  C.byref(w), C.byref(h)))
    return seg, env


def generate_curves_host(json_text: str | dict) -> np.ndarray:
    """A .hair scene's round Catmull-Rom segments (n x 20: four control
    points x, y, z, radius; strand; end flags; 0; 0) on the host."""
    if isinstance(json_text, dict):
        json_text = json.dumps(json_text)
    lib = load()
    n = C.c_int()
    t = json_text.encode()
    check(lib.hmsnn_generate_curves_host(t, None, 0, C.byref(n)))
    crv = np.empty((n.value, 20), np.float32)
    if n.value:
        check(lib.hmsnn_generate_curves_host(t, _ptr(crv), crv.size, C.byref(n)))
    return crv


def load_obj(path: str) -> dict:
    """An OBJ head mesh as the library reads it (host only): vertices and
    normals (T x 3 x 3) and diffuse colours (T x 3) per triangle."""
    lib = load()
    n = C.c_int()
    check(lib.hmsnn_obj_load(str(path).encode(), C.byref(n), None, None, None, 0))
    v, nn, kd = (np.empty((n.value, 3, 3), np.float32), np.empty((n.value, 3, 3), np.float32),
                 np.empty((n.value, 3), np.float32))
    check(lib.hmsnn_obj_load(str(path).encode(), C.byref(n), _ptr(v), _ptr(nn), _ptr(kd), n.value))
    return {"v": v, "n": nn, "kd": kd}


class Scene:
    """A scene context: geometry, BVH, environment tables and render state on the GPU."""

    def __init__(self, path: str | None = None, json_text: str | dict | None = None, base_dir: str = "."):
        self.lib = load()
        h = C.c_void_p()
        if path is not None:
            check(self.lib.hmsnn_create(str(path).encode(), C.byref(h)))
        else:
            if isinstance(json_text, dict):
                json_text = json.dumps(json_text)
            check(self.lib.hmsnn_create_from_json(json_text.encode(), base_dir.encode(), C.byref(h)))
        self.h = h
        self.info = self._info()

    def _info(self) -> dict:
        info = Info()
        check(self.lib.hmsnn_get_info(self.h, C.byref(info)))
        return _info_dict(info)

    @property
    def width(self) -> int:
        return int(self.info["width"])

    @property
    def height(self) -> int:
        return int(self.info["height"])

    def close(self) -> None:
        if self.h:
            self.lib.hmsnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- generated content ----
    def segments(self) -> np.ndarray:
        out = np.empty((int(self.info["num_segments"]), 8), dtype=np.float32)
        check(self.lib.hmsnn_export_segments(self.h, _ptr(out)))
        return out

    def curves(self) -> np.ndarray | None:
        """Round Catmull-Rom segments (n x 20), None when the hair is linear."""
        n = int(self.info["num_curves"])
        if n == 0:
            return None
        out = np.empty((n, 20), dtype=np.float32)
        check(self.lib.hmsnn_export_curves(self.h, _ptr(out)))
        return out

    def mesh(self) -> dict | None:
        """The OBJ head mesh: vertices (T x 3 x 3), normals (T x 3 x 3), diffuse (T x 3)."""
        n = int(self.info["num_triangles"])
        if n == 0:
            return None
        v, nn, kd = np.empty((n, 3, 3), np.float32), np.empty((n, 3, 3), np.float32), np.empty((n, 3), np.float32)
        check(self.lib.hmsnn_export_mesh(self.h, _ptr(v), _ptr(nn), _ptr(kd)))
        return {"v": v, "n": nn, "kd": kd}

    def env(self) -> dict:
        W, H = int(self.info["env_width"]), int(self.info["env_height"])
        rgba = np.empty((H, W, 4), np.float32)
        cp, cc = np.empty((H, W + 1), np.float32), np.empty((H, W + 1), np.float32)
        mp, mc = np.empty(H + 1, np.float32), np.empty(H + 1, np.float32)
        check(self.lib.hmsnn_export_env(self.h, _ptr(rgba), _ptr(cp), _ptr(cc), _ptr(mp), _ptr(mc)))
        return {"rgba": rgba, "c_pdf": cp, "c_cdf": cc, "m_pdf": mp, "m_cdf": mc}

    def set_strand_sigma_a(self, sigma: np.ndarray | None) -> None:
        if sigma is None:
            check(self.lib.hmsnn_set_strand_sigma_a(self.h, None, 0))
            return
        s = np.ascontiguousarray(sigma, dtype=np.float32)
        check(self.lib.hmsnn_set_strand_sigma_a(self.h, _ptr(s), s.shape[0]))

    def set_hair_params(self, beta_m: float, beta_n: float, alpha: float) -> None:
        """Global hair roughness and scale tilt (alpha in radians)."""
        check(self.lib.hmsnn_set_hair_params(self.h, float(beta_m), float(beta_n), float(alpha)))
        self.info = self._info()

    @property
    def num_strands(self) -> int:
        return int(self.info["num_strands"])

    def adjoint(self, dl_dimage: np.ndarray, spp: int, per_pixel: bool = False):
        """Path-replay gradient of a loss of the spp-pass mean image given
        dL/d(image) (H x W x 3 or 4): returns (d/d sigma_a per strand
        (num_strands x 3), d/d (beta_m, beta_n, alpha), and optionally each
        pixel's share of the latter (H x W x 3))."""
        H, W = self.height, self.width
        g = np.zeros((H, W, 4), np.float32)
        d = np.asarray(dl_dimage, np.float32)
        g[..., : d.shape[-1]] = d[..., :4]
        gs = np.zeros((self.num_strands, 3))
        gg = (C.c_double * 3)()
        gp = np.zeros((H, W, 3), np.float32) if per_pixel else None
        check(self.lib.hmsnn_adjoint_render(self.h, int(spp), _ptr(g), _ptr(gs), gg,
                                            _ptr(gp) if per_pixel else None))
        out = (gs, np.array(list(gg)))
        return out + (gp,) if per_pixel else out

    def set_seed_offset(self, offset: int) -> None:
        """Offsets the pass index behind every pixel's random stream (0 = the reference's)."""
        check(self.lib.hmsnn_set_seed_offset(self.h, int(offset)))

    def set_pass_concurrency(self, passes_in_flight: int) -> None:
        """Path-tracing passes in flight at once (1..4); the image does not depend on it."""
        check(self.lib.hmsnn_set_pass_concurrency(self.h, int(passes_in_flight)))

    def set_pt_schedule(self, schedule: str) -> None:
        """"stream" (default: a pixel's next pass starts when its path ends) or
        "passes" (one wavefront per pass); the image does not depend on it."""
        check(self.lib.hmsnn_set_pt_schedule(self.h, {"stream": 0, "passes": 1}[schedule]))

    # ---- path tracing ----
    def pt_reset(self) -> None:
        check(self.lib.hmsnn_pt_reset(self.h))

    def pt_passes(self, n: int, sync: bool = True) -> None:
        check(self.lib.hmsnn_pt_passes(self.h, int(n)))
        if sync:
            self.synchronize()

    def synchronize(self) -> None:
        check(self.lib.hmsnn_ctx_synchronize(self.h))

    def image(self, which: int = 0) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), np.float32)
        check(self.lib.hmsnn_get_image(self.h, which, _ptr(out)))
        return out

    def last_pass(self) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), np.float32)
        check(self.lib.hmsnn_get_last_pass(self.h, _ptr(out)))
        return out

    def render_path_tracing(self, spp: int | None = None) -> np.ndarray:
        spp = int(self.info["spp"]) if spp is None else int(spp)
        out = np.empty((self.height, self.width, 4), np.float32)
        check(self.lib.hmsnn_render_path_tracing(self.h, spp, _ptr(out)))
        return out

    def trace_stats(self, passes: int = 1) -> dict:
        o = (C.c_double * 6)()
        check(self.lib.hmsnn_pt_trace_stats(self.h, int(passes), o))
        keys = ["closest_rays", "closest_nodes", "closest_segs", "any_rays", "any_nodes", "any_segs"]
        return dict(zip(keys, list(o)))

    def time_passes(self, passes: int) -> dict:
        o = (C.c_double * 5)()
        check(self.lib.hmsnn_pt_time_passes(self.h, int(passes), o))
        keys = ["total_ms", "extend_ms", "shade_ms", "shadow_ms", "extend_launches"]
        return dict(zip(keys, list(o)))

    def trace_rays(self, rays: np.ndarray, any_hit: bool = False):
        """rays: n x 8 float32 (origin, pad, direction, pad). Returns prim, t, u."""
        r = np.ascontiguousarray(rays, dtype=np.float32)
        n = r.shape[0]
        dr = DeviceBuffer.from_numpy(r)
        dp, dt, du = DeviceBuffer(4 * n), DeviceBuffer(4 * n), DeviceBuffer(4 * n)
        check(self.lib.hmsnn_trace_rays_dev(self.h, dr.ptr, n, int(bool(any_hit)), dp.ptr, dt.ptr, du.ptr))
        prim = dp.to_numpy(np.int32, (n,))
        if any_hit:
            return prim
        return prim, dt.to_numpy(np.float32, (n,)), du.to_numpy(np.float32, (n,))

    # ---- HairMSNN ----
    def msnn_begin(self, BETA: int = 1) -> None:
        check(self.lib.hmsnn_msnn_begin(self.h, int(BETA)))

    def msnn_frames(self, frames: int) -> dict:
        st = MsnnStats()
        check(self.lib.hmsnn_msnn_frames(self.h, int(frames), C.byref(st)))
        return {k: getattr(st, k) for k, _ in MsnnStats._fields_}

    def msnn_image(self, which: int = 0) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), np.float32)
        check(self.lib.hmsnn_msnn_get_image(self.h, which, _ptr(out)))
        return out

    def msnn_save_weights(self, path: str) -> None:
        """Writes the HairMSNN radiance cache to a weight file for tcnn.init_weights."""
        check(self.lib.hmsnn_msnn_save_weights(self.h, str(path).encode()))

    def render_hair_msnn(self, BETA: int = 1, spp: int | None = None) -> np.ndarray:
        spp = int(self.info["spp"]) if spp is None else int(spp)
        out = np.empty((self.height, self.width, 4), np.float32)
        st = MsnnStats()
        check(self.lib.hmsnn_render_hair_msnn(self.h, int(BETA), spp, _ptr(out), C.byref(st)))
        return out

    # ---- neural radiance caching ----
    _NRC_BUFFERS = {
        # name: (id, dtype, shape from (n_tpx, every_nth, n_query, n_rec, w*h))
        "query": (0, np.float32, lambda t, e, q, r, p: (q, 12)),
        "cache": (1, np.float32, lambda t, e, q, r, p: (q, 3)),
        "gb_radiance": (2, np.float32, lambda t, e, q, r, p: (p, 4)),
        "gb_beta": (3, np.float32, lambda t, e, q, r, p: (p, 4)),
        "gb_hit": (4, np.uint8, lambda t, e, q, r, p: (p,)),
        "train_input": (5, np.float32, lambda t, e, q, r, p: (r, 12)),
        "train_target": (6, np.float32, lambda t, e, q, r, p: (r, 3)),
        "train_idx": (7, np.int32, lambda t, e, q, r, p: (t,)),
        "train_bounces": (8, np.int32, lambda t, e, q, r, p: (t,)),
        "train_hit": (9, np.uint8, lambda t, e, q, r, p: (t,)),
    }

    def nrc_begin(self) -> None:
        check(self.lib.hmsnn_nrc_begin(self.h))

    def nrc_set_params(self, c: float, all_unbiased: bool = False) -> None:
        check(self.lib.hmsnn_nrc_set_params(self.h, float(c), int(bool(all_unbiased))))

    def msnn_sizes(self) -> dict:
        v = np.zeros(4, np.int32)
        check(self.lib.hmsnn_msnn_get_sizes(self.h, _ptr(v)))
        return {"train_records": int(v[0]), "every_nth": int(v[1]), "frames": int(v[2]), "train_steps": int(v[3])}

    # name -> (which, dtype, shape from (train records, pixels))
    _MSNN_BUFFERS = {
        "train_input": (0, np.float32, lambda t, p: (t, 12)),
        "train_target": (1, np.float32, lambda t, p: (t, 3)),
        "train_pixel": (2, np.int32, lambda t, p: (t,)),
        "gb_input": (3, np.float32, lambda t, p: (p, 12)),
        "gb_hit": (4, np.uint8, lambda t, p: (p,)),
        "gb_color": (5, np.float32, lambda t, p: (p, 4)),
        "nn_output": (6, np.float32, lambda t, p: (p, 3)),
    }

    def msnn_buffer(self, name: str) -> np.ndarray:
        """One of the last HairMSNN frame's buffers (see _MSNN_BUFFERS)."""
        which, dtype, shape_of = self._MSNN_BUFFERS[name]
        out = np.empty(shape_of(self.msnn_sizes()["train_records"], self.width * self.height), dtype)
        check(self.lib.hmsnn_msnn_get_buffer(self.h, which, _ptr(out), out.nbytes))
        return out

    def nrc_frames(self, frames: int) -> dict:
        st = NrcStats()
        check(self.lib.hmsnn_nrc_frames(self.h, int(frames), C.byref(st)))
        return {k: getattr(st, k) for k, _ in NrcStats._fields_}

    def nrc_image(self) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), np.float32)
        check(self.lib.hmsnn_nrc_get_image(self.h, _ptr(out)))
        return out

    def nrc_sizes(self) -> dict:
        v = np.zeros(4, np.int32)
        check(self.lib.hmsnn_nrc_get_sizes(self.h, _ptr(v)))
        return {"train_pixels": int(v[0]), "every_nth": int(v[1]), "queries": int(v[2]), "train_records": int(v[3])}

    def nrc_buffer(self, name: str) -> np.ndarray:
        which, dtype, shape_of = self._NRC_BUFFERS[name]
        z = self.nrc_sizes()
        shape = shape_of(z["train_pixels"], z["every_nth"], z["queries"], z["train_records"], self.width * self.height)
        out = np.empty(shape, dtype)
        check(self.lib.hmsnn_nrc_get_buffer(self.h, which, _ptr(out), out.nbytes))
        return out

    def nrc_weights(self) -> np.ndarray:
        w = np.empty(20672, np.float32)
        check(self.lib.hmsnn_nrc_get_mlp_weights(self.h, _ptr(w), w.size))
        return w

    def nrc_save_weights(self, path: str) -> None:
        """Writes the NRC cache network to a weight file for tcnn.init_weights."""
        check(self.lib.hmsnn_nrc_save_weights(self.h, str(path).encode()))

    def nrc_set_weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.float32)
        check(self.lib.hmsnn_nrc_set_mlp_weights(self.h, _ptr(w), w.size))

    def render_nrc(self, spp: int | None = None) -> np.ndarray:
        spp = int(self.info["spp"]) if spp is None else int(spp)
        out = np.empty((self.height, self.width, 4), np.float32)
        st = NrcStats()
        check(self.lib.hmsnn_render_nrc(self.h, spp, _ptr(out), C.byref(st)))
        return out


def render_path_tracing(scene_json: str, spp: int | None = None) -> np.ndarray:
    """Python counterpart of the render_path_tracing executable."""
    with Scene(scene_json) as s:
        return s.render_path_tracing(spp)


def render_hair_msnn(scene_json: str, BETA: int = 1, spp: int | None = None) -> np.ndarray:
    """Python counterpart of the render_hair_msnn executable (kernel beta = BETA - 1)."""
    with Scene(scene_json) as s:
        return s.render_hair_msnn(BETA, spp)


def render_nrc(scene_json: str, spp: int | None = None) -> np.ndarray:
    """Python counterpart of the render_nrc executable."""
    with Scene(scene_json) as s:
        return s.render_nrc(spp)
