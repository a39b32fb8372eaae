"""ctypes binding of libtomogpi_mi355x.so and a thin mirror of the
reference's operator classes (Projector::doProjection,
BackProjector::doBackProjection, Projector::updateVolume).

Device arrays are torch tensors on a ROCm device (torch supplies memory and
streams only); host arrays are numpy arrays and go through the library's
TOMO_FLAG_HOST_PTR staging.  Every call raises TomoError on a nonzero status.
The library must be built (``python -m tomogpi_amd.build``); nothing here
falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, asdict, fields

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libtomogpi_mi355x.so")

TOMO_OK = 0
F32, F16, BF16, F64 = 0, 1, 2, 3
PROJ_RS, PROJ_SIDDON, BP_VI = 0, 1, 0
FLAG_RS_FOV_CUBE, FLAG_SIDDON_COMPAT, FLAG_HOST_PTR, FLAG_FDK, FLAG_INDEX_READS, FLAG_STAGED = 0x1, 0x2, 0x4, 0x8, 0x10, 0x20
PAIR_RSVI, PAIR_SIDDONVI = 0, 1
SCHEDULE_DIRECT, SCHEDULE_RECURRENT = 0, 1
ALGO_GD, ALGO_SIRT = 0, 1
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3

# Every symbol include/tomogpi.h declares.
EXPORTS = [
    "tomo_abi_version", "tomo_last_error", "tomo_status_string", "tomo_ctx_create", "tomo_ctx_destroy",
    "tomo_ctx_mem_info", "tomo_alloc", "tomo_free", "tomo_copy_to_device", "tomo_copy_to_host",
    "tomo_synchronize", "tomo_geom_create", "tomo_geom_desc_from_config", "tomo_geom_destroy",
    "tomo_geom_get_info", "tomo_geom_get_phi", "tomo_project", "tomo_backproject", "tomo_project_adjoint",
    "tomo_backproject_adjoint", "tomo_project_work", "tomo_convert", "tomo_diff", "tomo_scale", "tomo_axpy",
    "tomo_sq_norm", "tomo_sq_norm_noalloc", "tomo_dot", "tomo_laplacian", "tomo_gd_run", "tomo_wgd_run", "tomo_cg_run", "tomo_sirt_run", "tomo_phantom_shepp_logan",
    "tomo_phantom_ellipsoids", "tomo_file_read", "tomo_file_write", "tomo_serve_create", "tomo_serve_submit",
    "tomo_serve_poll", "tomo_serve_wait", "tomo_serve_iter_stats", "tomo_serve_get_stats", "tomo_serve_destroy",
]


class TomoError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"status {status}: {message}")
        self.status = status


class GeomDesc(C.Structure):
    _fields_ = [
        ("fod_mm", C.c_double), ("fdd_mm", C.c_double), ("projection_nb", C.c_int32),
        ("num_start_angle", C.c_double), ("den_start_angle", C.c_double),
        ("num_stop_angle", C.c_double), ("den_stop_angle", C.c_double),
        ("u_detector_size_mm", C.c_float), ("v_detector_size_mm", C.c_float),
        ("u_pixel_nb", C.c_int32), ("v_pixel_nb", C.c_int32),
        ("x_volume_size_mm", C.c_float), ("y_volume_size_mm", C.c_float), ("z_volume_size_mm", C.c_float),
        ("x_pixel_nb", C.c_int32), ("y_pixel_nb", C.c_int32), ("z_pixel_nb", C.c_int32),
    ]


class GeomInfo(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("nu", C.c_int32), ("nv", C.c_int32),
        ("nphi", C.c_int32), ("fod", C.c_double), ("fdd", C.c_double),
        ("du", C.c_float), ("dv", C.c_float), ("cu", C.c_float), ("cv", C.c_float),
        ("delta", C.c_float), ("cx", C.c_float), ("cy", C.c_float), ("cz", C.c_float),
        ("step_angle", C.c_double), ("start_angle", C.c_double),
    ]


class GdParams(C.Structure):
    _fields_ = [
        ("pair", C.c_int32), ("flags", C.c_int32), ("gradient_iteration_nb", C.c_int32),
        ("optimal_step_iteration_nb", C.c_int32), ("lam", C.c_double), ("positivity", C.c_int32),
        ("alpha_min_init", C.c_double), ("storage", C.c_int32),
    ]


class GdIterStat(C.Structure):
    _fields_ = [("residual_sq", C.c_double), ("norm_dj", C.c_double), ("norm_dj_proj", C.c_double),
                ("alpha", C.c_double), ("j_reg", C.c_double), ("beta", C.c_double)]


class ServeParams(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_active", C.c_int32), ("use_graph", C.c_int32),
                ("schedule", C.c_int32), ("sub_rounds", C.c_int32)]


class SirtParams(C.Structure):
    _fields_ = [("pair", C.c_int32), ("flags", C.c_int32), ("iterations", C.c_int32), ("positivity", C.c_int32),
                ("relaxation", C.c_double)]


class ServeRequest(C.Structure):
    _fields_ = [("geom", GeomDesc), ("gd", GdParams), ("sinogram", C.c_void_p), ("volume", C.c_void_p),
                ("algorithm", C.c_int32), ("storage", C.c_int32), ("relaxation", C.c_double)]


class ServeResult(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("submit_s", C.c_double), ("admit_s", C.c_double),
                ("done_s", C.c_double), ("latency_s", C.c_double), ("residual_sq", C.c_double),
                ("device_bytes", C.c_uint64)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("rounds", C.c_uint64), ("active", C.c_int32), ("queued", C.c_int32), ("peak_active", C.c_int32),
                ("reserved", C.c_int32), ("bytes_in_use", C.c_uint64), ("peak_bytes", C.c_uint64),
                ("budget_bytes", C.c_uint64), ("latency_p50_s", C.c_double), ("latency_p99_s", C.c_double),
                ("latency_mean_s", C.c_double), ("latency_max_s", C.c_double), ("first_submit_s", C.c_double),
                ("last_done_s", C.c_double), ("throughput_rps", C.c_double)]


_lib = None


def load() -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m tomogpi_amd.build`")
    lib = C.CDLL(LIB_PATH)
    vp, i, f, d, sz = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_size_t
    sig = {
        "tomo_abi_version": ([], i), "tomo_last_error": ([], C.c_char_p), "tomo_status_string": ([i], C.c_char_p),
        "tomo_ctx_create": ([i, C.POINTER(vp)], i), "tomo_ctx_destroy": ([vp], i),
        "tomo_ctx_mem_info": ([vp, C.POINTER(sz), C.POINTER(sz)], i),
        "tomo_alloc": ([vp, sz, C.POINTER(vp)], i), "tomo_free": ([vp, vp], i),
        "tomo_copy_to_device": ([vp, vp, vp, sz], i), "tomo_copy_to_host": ([vp, vp, vp, sz], i),
        "tomo_synchronize": ([vp, vp], i),
        "tomo_geom_create": ([vp, C.POINTER(GeomDesc), C.POINTER(vp)], i),
        "tomo_geom_desc_from_config": ([C.c_char_p, C.POINTER(GeomDesc)], i),
        "tomo_geom_destroy": ([vp], i), "tomo_geom_get_info": ([vp, C.POINTER(GeomInfo)], i),
        "tomo_geom_get_phi": ([vp, C.POINTER(d)], i),
        "tomo_project": ([vp, vp, i, i, vp, i, f, vp, i, f, vp], i),
        "tomo_backproject": ([vp, vp, i, i, vp, i, f, vp, i, f, vp], i),
        "tomo_project_adjoint": ([vp, vp, i, i, vp, i, f, vp, i, f, vp], i),
        "tomo_backproject_adjoint": ([vp, vp, i, i, vp, i, f, vp, i, f, vp], i),
        "tomo_project_work": ([vp, vp, i, i, C.POINTER(C.c_uint64), vp], i),
        "tomo_convert": ([vp, vp, i, f, vp, i, f, sz, vp], i),
        "tomo_diff": ([vp, vp, vp, vp, sz, vp], i), "tomo_scale": ([vp, vp, f, sz, vp], i),
        "tomo_axpy": ([vp, vp, vp, d, i, sz, vp], i),
        "tomo_sq_norm": ([vp, vp, i, f, sz, C.POINTER(d), vp], i),
        "tomo_sq_norm_noalloc": ([vp, vp, i, f, sz, C.POINTER(d), vp], i),
        "tomo_dot": ([vp, vp, vp, sz, C.POINTER(d), vp], i),
        "tomo_laplacian": ([vp, vp, vp, vp, vp], i),
        "tomo_gd_run": ([vp, vp, C.POINTER(GdParams), vp, vp, C.POINTER(GdIterStat), vp], i),
        "tomo_cg_run": ([vp, vp, C.POINTER(GdParams), vp, vp, C.POINTER(GdIterStat), vp], i),
        "tomo_wgd_run": ([vp, vp, C.POINTER(GdParams), vp, vp, vp, C.POINTER(GdIterStat), vp], i),
        "tomo_sirt_run": ([vp, vp, C.POINTER(SirtParams), vp, vp, C.POINTER(d), vp], i),
        "tomo_phantom_shepp_logan": ([vp, vp, vp, vp], i),
        "tomo_phantom_ellipsoids": ([vp, vp, i, C.c_uint64, d, d, vp, vp], i),
        "tomo_file_read": ([C.c_char_p, vp, i, sz], i), "tomo_file_write": ([C.c_char_p, vp, i, sz], i),
        "tomo_serve_create": ([vp, C.POINTER(ServeParams), C.POINTER(vp)], i),
        "tomo_serve_submit": ([vp, C.POINTER(ServeRequest), C.POINTER(C.c_uint64)], i),
        "tomo_serve_poll": ([vp, C.c_uint64, C.POINTER(ServeResult)], i),
        "tomo_serve_wait": ([vp, C.c_uint64, d, C.POINTER(ServeResult)], i),
        "tomo_serve_iter_stats": ([vp, C.c_uint64, C.POINTER(GdIterStat), i], i),
        "tomo_serve_get_stats": ([vp, C.POINTER(ServeStats)], i),
        "tomo_serve_destroy": ([vp], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(status: int):
    if status != TOMO_OK:
        msg = load().tomo_last_error().decode(errors="replace")
        raise TomoError(status, msg)


@dataclass
class AcquisitionDesc:
    """configCT fields (tomo_geom_desc)."""
    fod_mm: float
    fdd_mm: float
    projection_nb: int
    u_detector_size_mm: float
    v_detector_size_mm: float
    u_pixel_nb: int
    v_pixel_nb: int
    x_volume_size_mm: float
    y_volume_size_mm: float
    z_volume_size_mm: float
    x_pixel_nb: int
    y_pixel_nb: int
    z_pixel_nb: int
    num_start_angle: float = 0.0
    den_start_angle: float = 1.0
    num_stop_angle: float = 2.0
    den_stop_angle: float = 1.0

    def to_c(self) -> GeomDesc:
        return GeomDesc(**asdict(self))

    @classmethod
    def from_c(cls, d: GeomDesc) -> "AcquisitionDesc":
        return cls(**{f.name: getattr(d, f.name) for f in fields(cls)})

    @classmethod
    def from_config(cls, workdir: str) -> "AcquisitionDesc":
        d = GeomDesc()
        check(load().tomo_geom_desc_from_config(workdir.encode(), C.byref(d)))
        return cls.from_c(d)

    @classmethod
    def scaled(cls, n: int, nphi: int, nu: int | None = None, nv: int | None = None, pixel_mm: float = 2.0,
               voxel_mm: float = 1.0) -> "AcquisitionDesc":
        """FOD 500 / FDD 1000 mm at 256 voxels scaled by n/256; 1 mm voxels, 2 mm pixels."""
        nu = n if nu is None else nu
        nv = n if nv is None else nv
        k = n / 256.0
        return cls(fod_mm=500.0 * k, fdd_mm=1000.0 * k, projection_nb=nphi, u_detector_size_mm=nu * pixel_mm,
                   v_detector_size_mm=nv * pixel_mm, u_pixel_nb=nu, v_pixel_nb=nv, x_volume_size_mm=n * voxel_mm,
                   y_volume_size_mm=n * voxel_mm, z_volume_size_mm=n * voxel_mm, x_pixel_nb=n, y_pixel_nb=n,
                   z_pixel_nb=n)


def _ptr(a):
    """(pointer, is_host) of a torch tensor or numpy array."""
    if isinstance(a, np.ndarray):
        assert a.flags.c_contiguous
        return a.ctypes.data, True
    assert a.is_contiguous()
    return a.data_ptr(), not a.is_cuda


_DT_NP = {np.dtype(np.float32): F32, np.dtype(np.float16): F16, np.dtype(np.float64): F64}


def dtype_code(a) -> int:
    if isinstance(a, np.ndarray):
        return _DT_NP[a.dtype]
    import torch
    return {torch.float32: F32, torch.float16: F16, torch.bfloat16: BF16}[a.dtype]


def _stream(stream) -> int | None:
    if stream is None:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.current_stream().cuda_stream
        return None
    return stream if isinstance(stream, int) else stream.cuda_stream


class Context:
    """One device (tomo_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.tomo_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.lib.tomo_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mem_info(self):
        f, t = C.c_size_t(), C.c_size_t()
        check(self.lib.tomo_ctx_mem_info(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value


class Geometry:
    """Acquisition geometry uploaded to one context (tomo_geom)."""

    def __init__(self, ctx: Context, desc: AcquisitionDesc):
        self.ctx, self.desc, self.lib = ctx, desc, ctx.lib
        h = C.c_void_p()
        cd = desc.to_c()
        check(self.lib.tomo_geom_create(ctx.h, C.byref(cd), C.byref(h)))
        self.h = h
        info = GeomInfo()
        check(self.lib.tomo_geom_get_info(h, C.byref(info)))
        self.info = {f[0]: getattr(info, f[0]) for f in GeomInfo._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.lib.tomo_geom_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def vol_shape(self):
        i = self.info
        return (i["nz"], i["ny"], i["nx"])

    @property
    def sino_shape(self):
        i = self.info
        return (i["nphi"], i["nv"], i["nu"])

    def phi(self):
        out = np.zeros(self.info["nphi"], dtype=np.float64)
        check(self.lib.tomo_geom_get_phi(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    # -- operators ---------------------------------------------------------
    def _call(self, fn, op, flags, src, src_scale, dst, dst_scale, stream):
        ps, hs = _ptr(src)
        pd, hd = _ptr(dst)
        if hs != hd:
            raise ValueError("input and output must both be host or both be device arrays")
        if hs:
            flags |= FLAG_HOST_PTR
        check(fn(self.ctx.h, self.h, op, flags, C.c_void_p(ps), dtype_code(src), src_scale, C.c_void_p(pd),
                 dtype_code(dst), dst_scale, C.c_void_p(_stream(stream)) if not hs else None))

    def project(self, vol, sino, projector=PROJ_RS, flags=0, vol_scale=1.0, sino_scale=1.0, stream=None):
        self._call(self.lib.tomo_project, projector, flags, vol, vol_scale, sino, sino_scale, stream)
        return sino

    def backproject(self, sino, vol, flags=0, sino_scale=1.0, vol_scale=1.0, stream=None):
        self._call(self.lib.tomo_backproject, BP_VI, flags, sino, sino_scale, vol, vol_scale, stream)
        return vol

    def project_adjoint(self, sino, vol, projector=PROJ_RS, flags=0, sino_scale=1.0, vol_scale=1.0, stream=None):
        self._call(self.lib.tomo_project_adjoint, projector, flags, sino, sino_scale, vol, vol_scale, stream)
        return vol

    def backproject_adjoint(self, vol, sino, flags=0, vol_scale=1.0, sino_scale=1.0, stream=None):
        self._call(self.lib.tomo_backproject_adjoint, BP_VI, flags, vol, vol_scale, sino, sino_scale, stream)
        return sino

    def project_work(self, projector=PROJ_RS, flags=0, stream=None) -> int:
        out = C.c_uint64()
        check(self.lib.tomo_project_work(self.ctx.h, self.h, projector, flags, C.byref(out),
                                         C.c_void_p(_stream(stream))))
        return out.value

    def laplacian(self, x, y, stream=None):
        check(self.lib.tomo_laplacian(self.ctx.h, self.h, C.c_void_p(_ptr(x)[0]), C.c_void_p(_ptr(y)[0]),
                                      C.c_void_p(_stream(stream))))
        return y

    def shepp_logan(self, vol, stream=None):
        check(self.lib.tomo_phantom_shepp_logan(self.ctx.h, self.h, C.c_void_p(_ptr(vol)[0]),
                                                C.c_void_p(_stream(stream))))
        return vol

    def ellipsoids(self, vol, count=64, seed=0, amin=5.0, amax=120.0, stream=None):
        check(self.lib.tomo_phantom_ellipsoids(self.ctx.h, self.h, count, seed, amin, amax,
                                               C.c_void_p(_ptr(vol)[0]), C.c_void_p(_stream(stream))))
        return vol

    def gd_run(self, sino, vol, pair=PAIR_RSVI, flags=0, iters=10, kopt=10, lam=0.0, positivity=0,
               alpha_min_init=-0.0347711, stream=None, conjugate=False, weight=None, storage=F32):
        """Simple-gradient loop (tomo_gd_run), conjugate gradient (tomo_cg_run),
        or with a weight sinogram the weighted gradient (tomo_wgd_run);
        storage F16 / BF16 gives the mixed-precision form."""
        p = GdParams(pair=pair, flags=flags, gradient_iteration_nb=iters, optimal_step_iteration_nb=kopt, lam=lam,
                     positivity=positivity, alpha_min_init=alpha_min_init, storage=storage)
        stats = (GdIterStat * max(iters, 1))()
        ps, hs = _ptr(sino)
        pv, hv = _ptr(vol)
        if hs:
            p.flags |= FLAG_HOST_PTR
        st = None if hs else C.c_void_p(_stream(stream))
        if weight is not None:
            if conjugate:
                raise ValueError("the weighted form is a simple gradient")
            pw, hw = _ptr(weight)
            if hw != hs:
                raise ValueError("weight and sinogram must both be host or both be device arrays")
            check(self.lib.tomo_wgd_run(self.ctx.h, self.h, C.byref(p), C.c_void_p(ps), C.c_void_p(pw),
                                        C.c_void_p(pv), stats, st))
        else:
            fn = self.lib.tomo_cg_run if conjugate else self.lib.tomo_gd_run
            check(fn(self.ctx.h, self.h, C.byref(p), C.c_void_p(ps), C.c_void_p(pv), stats, st))
        return [dict(residual_sq=s.residual_sq, norm_dj=s.norm_dj, norm_dj_proj=s.norm_dj_proj, alpha=s.alpha,
                     j_reg=s.j_reg, beta=s.beta)
                for s in stats[:iters]]

    def sirt_run(self, sino, vol, pair=PAIR_RSVI, flags=0, iters=20, relaxation=1.0, positivity=0, stream=None):
        """SIRT from x0 = 0 (tomo_sirt_run); returns the residuals sum (y - P x)^2."""
        p = SirtParams(pair=pair, flags=flags, iterations=iters, positivity=positivity, relaxation=relaxation)
        res = (C.c_double * max(iters, 1))()
        ps, hs = _ptr(sino)
        pv, hv = _ptr(vol)
        if hs:
            p.flags |= FLAG_HOST_PTR
        check(self.lib.tomo_sirt_run(self.ctx.h, self.h, C.byref(p), C.c_void_p(ps), C.c_void_p(pv), res,
                                     None if hs else C.c_void_p(_stream(stream))))
        return list(res[:iters])


def sq_norm(ctx: Context, x, scale=1.0, stream=None) -> float:
    out = C.c_double()
    check(ctx.lib.tomo_sq_norm(ctx.h, C.c_void_p(_ptr(x)[0]), dtype_code(x), scale, x.numel(), C.byref(out),
                               C.c_void_p(_stream(stream))))
    return out.value


def sq_norm_noalloc(ctx: Context, x, scale=1.0, stream=None) -> float:
    """tomo_sq_norm_noalloc: tomo_sq_norm with its partial sums in the
    library's per-GPU scratch array (no device allocation)."""
    out = C.c_double()
    check(ctx.lib.tomo_sq_norm_noalloc(ctx.h, C.c_void_p(_ptr(x)[0]), dtype_code(x), scale, x.numel(),
                                       C.byref(out), C.c_void_p(_stream(stream))))
    return out.value


def dot(ctx: Context, a, b, stream=None) -> float:
    out = C.c_double()
    check(ctx.lib.tomo_dot(ctx.h, C.c_void_p(_ptr(a)[0]), C.c_void_p(_ptr(b)[0]), a.numel(), C.byref(out),
                           C.c_void_p(_stream(stream))))
    return out.value


def _struct_dict(st) -> dict:
    return {f[0]: getattr(st, f[0]) for f in st._fields_ if f[0] != "reserved"}


class Server:
    """Serving loop (tomo_server): continuous batching of simple-gradient
    reconstructions of uneven sizes on one device."""

    def __init__(self, ctx: Context, budget_bytes: int = 0, max_active: int = 16, use_graph: bool = True,
                 schedule: int = SCHEDULE_DIRECT, sub_rounds: int = 0):
        self.ctx, self.lib = ctx, ctx.lib
        p = ServeParams(hbm_budget_bytes=budget_bytes, max_active=max_active, use_graph=int(use_graph),
                        schedule=schedule, sub_rounds=sub_rounds)
        h = C.c_void_p()
        check(self.lib.tomo_serve_create(ctx.h, C.byref(p), C.byref(h)))
        self.h = h
        self._keep = {}  # request id -> arrays that must outlive the request

    def submit(self, desc: AcquisitionDesc, sino, vol, pair=PAIR_RSVI, flags=0, iters=20, kopt=21, lam=0.01,
               positivity=0, alpha_min_init=-0.0347711, algorithm=ALGO_GD, relaxation=1.0, storage=F32) -> int:
        """Queue one reconstruction; storage F32 or BF16 sets the element type
        of its operator copies (the state and the result stay FP32)."""
        ps, hs = _ptr(sino)
        pv, hv = _ptr(vol)
        if hs != hv:
            raise ValueError("sinogram and volume must both be host or both be device arrays")
        gd = GdParams(pair=pair, flags=flags | (FLAG_HOST_PTR if hs else 0), gradient_iteration_nb=iters,
                      optimal_step_iteration_nb=kopt, lam=lam, positivity=positivity, alpha_min_init=alpha_min_init)
        r = ServeRequest(geom=desc.to_c(), gd=gd, sinogram=ps, volume=pv, algorithm=algorithm, relaxation=relaxation,
                         storage=storage)
        rid = C.c_uint64()
        check(self.lib.tomo_serve_submit(self.h, C.byref(r), C.byref(rid)))
        self._keep[rid.value] = (sino, vol)
        return rid.value

    def poll(self, rid: int) -> dict:
        r = ServeResult()
        check(self.lib.tomo_serve_poll(self.h, rid, C.byref(r)))
        return _struct_dict(r)

    def wait(self, rid: int, timeout: float = -1.0) -> dict:
        r = ServeResult()
        check(self.lib.tomo_serve_wait(self.h, rid, timeout, C.byref(r)))
        self._keep.pop(rid, None)
        return _struct_dict(r)

    def iter_stats(self, rid: int, n: int) -> list:
        st = (GdIterStat * max(n, 1))()
        check(self.lib.tomo_serve_iter_stats(self.h, rid, st, n))
        return [dict(residual_sq=x.residual_sq, norm_dj=x.norm_dj, norm_dj_proj=x.norm_dj_proj, alpha=x.alpha)
                for x in st[:n]]

    def stats(self) -> dict:
        st = ServeStats()
        check(self.lib.tomo_serve_get_stats(self.h, C.byref(st)))
        return _struct_dict(st)

    def close(self):
        if getattr(self, "h", None):
            self.lib.tomo_serve_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- mirror of the reference operator classes ------------------------------
class RegularSamplingProjector:
    """Projector<V,S,T> with doProjection(sino, vol) (src/TomoGPI_lib/inc/Projector.cuh:118)."""

    def __init__(self, geom: Geometry, cube_fov: bool = False):
        self.geom, self.flags = geom, FLAG_RS_FOV_CUBE if cube_fov else 0

    def doProjection(self, estimatedSinogram, volume, **kw):
        return self.geom.project(volume, estimatedSinogram, PROJ_RS, self.flags, **kw)

    @staticmethod
    def updateVolume(ctx: Context, volume_out, volume_in, lam: float, positive: int = 0, stream=None):
        """volume_out += (float)lam * volume_in, clamped at 0 if positive (Projector.cu:109-124)."""
        check(ctx.lib.tomo_axpy(ctx.h, C.c_void_p(_ptr(volume_out)[0]), C.c_void_p(_ptr(volume_in)[0]), lam,
                                positive, volume_out.numel(), C.c_void_p(_stream(stream))))


class SiddonProjector(RegularSamplingProjector):
    """SiddonProjector_compute_C_mem_CPU::doProjection; compat=True keeps the C path's quirks."""

    def __init__(self, geom: Geometry, compat: bool = False):
        self.geom, self.flags = geom, FLAG_SIDDON_COMPAT if compat else 0

    def doProjection(self, estimatedSinogram, volume, **kw):
        return self.geom.project(volume, estimatedSinogram, PROJ_SIDDON, self.flags, **kw)


class VIBackProjector:
    """BackProjector<V,S,T> with doBackProjection(vol, sino) (src/TomoGPI_lib/inc/BackProjector.cuh:144)."""

    def __init__(self, geom: Geometry):
        self.geom = geom

    def doBackProjection(self, estimatedVolume, sinogram, **kw):
        return self.geom.backproject(sinogram, estimatedVolume, **kw)
