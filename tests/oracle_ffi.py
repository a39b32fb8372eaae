"""ctypes bindings of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg load it, and only as the checker or as the timed
CPU baseline.  Build it with ``make -C oracle`` (``__graft_entry__.build()``
does this).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, asdict

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")


class OgDesc(C.Structure):
    _fields_ = [
        ("fod", C.c_double), ("fdd", C.c_double), ("nphi", C.c_int32),
        ("num_start", C.c_double), ("den_start", C.c_double),
        ("num_stop", C.c_double), ("den_stop", C.c_double),
        ("su", C.c_float), ("sv", C.c_float), ("nu", C.c_int32), ("nv", C.c_int32),
        ("sx", C.c_float), ("sy", C.c_float), ("sz", C.c_float),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
    ]


class OgGeom(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("nu", C.c_int32), ("nv", C.c_int32), ("nphi", C.c_int32),
        ("fod", C.c_double), ("fdd", C.c_double),
        ("du", C.c_float), ("dv", C.c_float), ("cu", C.c_float), ("cv", C.c_float),
        ("delta", C.c_float), ("cx", C.c_float), ("cy", C.c_float), ("cz", C.c_float),
        ("start", C.c_double), ("stop", C.c_double), ("step", C.c_double),
        ("phi", C.POINTER(C.c_double)), ("cphi", C.POINTER(C.c_double)),
        ("sphi", C.POINTER(C.c_double)),
        ("alpha_f", C.POINTER(C.c_float)), ("beta_f", C.POINTER(C.c_float)),
    ]


class OgGdParams(C.Structure):
    _fields_ = [
        ("pair", C.c_int32), ("cube", C.c_int32), ("compat", C.c_int32),
        ("iters", C.c_int32), ("kopt", C.c_int32), ("lam", C.c_double),
        ("positivity", C.c_int32), ("alpha_min_init", C.c_double), ("matched", C.c_int32),
    ]


class OgSirtParams(C.Structure):
    _fields_ = [("pair", C.c_int32), ("cube", C.c_int32), ("iters", C.c_int32), ("relax", C.c_double),
                ("positivity", C.c_int32)]


class OgGdStat(C.Structure):
    _fields_ = [("residual_sq", C.c_double), ("norm_dj", C.c_double),
                ("norm_dj_proj", C.c_double), ("alpha", C.c_double), ("j_reg", C.c_double), ("beta", C.c_double)]


@dataclass
class Geometry:
    """configCT values of one acquisition (same fields as tomo_geom_desc)."""
    fod: float
    fdd: float
    nphi: int
    su: float
    sv: float
    nu: int
    nv: int
    sx: float
    sy: float
    sz: float
    nx: int
    ny: int
    nz: int
    num_start: float = 0.0
    den_start: float = 1.0
    num_stop: float = 2.0
    den_stop: float = 1.0

    @property
    def vol_shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def sino_shape(self):
        return (self.nphi, self.nv, self.nu)

    def as_dict(self):
        return asdict(self)


def scaled_geometry(n: int, nphi: int, nu: int | None = None, nv: int | None = None,
                    pixel_mm: float = 2.0, voxel_mm: float = 1.0) -> Geometry:
    """The benchmark acquisition family: FOD 500 mm and FDD 1000 mm at 256
    voxels, scaled by n/256, 1 mm voxels, 2 mm detector pixels."""
    nu = n if nu is None else nu
    nv = n if nv is None else nv
    k = n / 256.0
    return Geometry(fod=500.0 * k, fdd=1000.0 * k, nphi=nphi, su=nu * pixel_mm, sv=nv * pixel_mm,
                    nu=nu, nv=nv, sx=n * voxel_mm, sy=n * voxel_mm, sz=n * voxel_mm, nx=n, ny=n, nz=n)


def _f32p(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f64p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
            lib = C.CDLL(LIB_PATH)
            G = C.POINTER(OgGeom)
            lib.og_geom_init.argtypes = [C.POINTER(OgDesc), G]
            lib.og_geom_free.argtypes = [G]
            fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
            for name, a in [("og_vi_f32", [G, fp, fp]), ("og_vi_f64", [G, dp, dp]),
                            ("og_vi_fdk_f32", [G, fp, fp]), ("og_vi_fdk_f64", [G, dp, dp]),
                            ("og_vi_T_f64", [G, dp, dp]),
                            ("og_rs_f32", [G, C.c_int, fp, fp]), ("og_rs_f64", [G, C.c_int, dp, dp]),
                            ("og_rs_T_f64", [G, C.c_int, dp, dp]),
                            ("og_siddon_f32", [G, fp, fp]), ("og_siddon_f64", [G, dp, dp]),
                            ("og_siddon_T_f64", [G, dp, dp])]:
                getattr(lib, name).argtypes = a
                getattr(lib, name).restype = None
            lib.og_rs_count.argtypes = [G, C.c_int]
            lib.og_rs_count.restype = C.c_uint64
            lib.og_siddon_count.argtypes = [G]
            lib.og_siddon_count.restype = C.c_uint64
            lib.og_siddon_compat.argtypes = [G, fp, fp, C.c_int]
            lib.og_siddon_compat.restype = C.c_int
            lib.og_laplacian_f32.argtypes = [C.c_int] * 3 + [fp, fp]
            lib.og_laplacian_f64.argtypes = [C.c_int] * 3 + [dp, dp]
            lib.og_shepp_logan.argtypes = [C.c_int] * 3 + [fp]
            lib.og_ellipsoids.argtypes = [C.c_int] * 4 + [C.c_uint64, C.c_double, C.c_double, fp]
            lib.og_ellipsoid_table.argtypes = [C.c_int] * 4 + [C.c_uint64, C.c_double, C.c_double, dp, dp]
            lib.og_gd_run.argtypes = [G, C.POINTER(OgGdParams), fp, fp, C.POINTER(OgGdStat)]
            lib.og_cg_run.argtypes = [G, C.POINTER(OgGdParams), fp, fp, C.POINTER(OgGdStat)]
            lib.og_wgd_run.argtypes = [G, C.POINTER(OgGdParams), fp, fp, fp, C.POINTER(OgGdStat)]
            lib.og_wgd_run.restype = C.c_int
            lib.og_cg_run.restype = C.c_int
            lib.og_sirt_weights.argtypes = [G, C.POINTER(OgSirtParams), fp, fp]
            lib.og_sirt_weights.restype = C.c_int
            lib.og_sirt_run.argtypes = [G, C.POINTER(OgSirtParams), fp, fp, dp]
            lib.og_sirt_run.restype = C.c_int
            lib.og_gd_run.restype = C.c_int
            lib.og_set_threads.argtypes = [C.c_int]
            cls._lib = lib
        return cls._lib

    def __init__(self, geom: Geometry):
        self.geom = geom
        d = OgDesc(fod=geom.fod, fdd=geom.fdd, nphi=geom.nphi, num_start=geom.num_start,
                   den_start=geom.den_start, num_stop=geom.num_stop, den_stop=geom.den_stop,
                   su=geom.su, sv=geom.sv, nu=geom.nu, nv=geom.nv, sx=geom.sx, sy=geom.sy, sz=geom.sz,
                   nx=geom.nx, ny=geom.ny, nz=geom.nz)
        self._g = OgGeom()
        if self.lib().og_geom_init(C.byref(d), C.byref(self._g)) != 0:
            raise ValueError("invalid geometry")

    def __del__(self):
        if getattr(self, "_g", None) is not None and Oracle._lib is not None:
            Oracle._lib.og_geom_free(C.byref(self._g))
            self._g = None

    # derived geometry --------------------------------------------------
    @property
    def info(self):
        g = self._g
        return dict(du=g.du, dv=g.dv, cu=g.cu, cv=g.cv, delta=g.delta, cx=g.cx, cy=g.cy, cz=g.cz,
                    step=g.step, start=g.start)

    @property
    def phi(self):
        return np.ctypeslib.as_array(self._g.phi, shape=(self.geom.nphi,)).copy()

    # operators ---------------------------------------------------------
    def _vol(self, dt):
        return np.zeros(self.geom.vol_shape, dtype=dt)

    def _sino(self, dt):
        return np.zeros(self.geom.sino_shape, dtype=dt)

    def vi(self, sino, f64=False):
        if f64:
            s = np.ascontiguousarray(sino, dtype=np.float64)
            out = self._vol(np.float64)
            self.lib().og_vi_f64(C.byref(self._g), _f64p(s), _f64p(out))
        else:
            s = np.ascontiguousarray(sino, dtype=np.float32)
            out = self._vol(np.float32)
            self.lib().og_vi_f32(C.byref(self._g), _f32p(s), _f32p(out))
        return out

    def vi_fdk(self, sino, f64=False):
        if f64:
            s = np.ascontiguousarray(sino, dtype=np.float64)
            out = self._vol(np.float64)
            self.lib().og_vi_fdk_f64(C.byref(self._g), _f64p(s), _f64p(out))
        else:
            s = np.ascontiguousarray(sino, dtype=np.float32)
            out = self._vol(np.float32)
            self.lib().og_vi_fdk_f32(C.byref(self._g), _f32p(s), _f32p(out))
        return out

    def vi_T(self, vol):
        v = np.ascontiguousarray(vol, dtype=np.float64)
        out = self._sino(np.float64)
        self.lib().og_vi_T_f64(C.byref(self._g), _f64p(v), _f64p(out))
        return out

    def rs(self, vol, cube=False, f64=False):
        if f64:
            v = np.ascontiguousarray(vol, dtype=np.float64)
            out = self._sino(np.float64)
            self.lib().og_rs_f64(C.byref(self._g), int(cube), _f64p(v), _f64p(out))
        else:
            v = np.ascontiguousarray(vol, dtype=np.float32)
            out = self._sino(np.float32)
            self.lib().og_rs_f32(C.byref(self._g), int(cube), _f32p(v), _f32p(out))
        return out

    def rs_T(self, sino, cube=False):
        s = np.ascontiguousarray(sino, dtype=np.float64)
        out = self._vol(np.float64)
        self.lib().og_rs_T_f64(C.byref(self._g), int(cube), _f64p(s), _f64p(out))
        return out

    def rs_count(self, cube=False):
        return int(self.lib().og_rs_count(C.byref(self._g), int(cube)))

    def siddon(self, vol, f64=False):
        if f64:
            v = np.ascontiguousarray(vol, dtype=np.float64)
            out = self._sino(np.float64)
            self.lib().og_siddon_f64(C.byref(self._g), _f64p(v), _f64p(out))
        else:
            v = np.ascontiguousarray(vol, dtype=np.float32)
            out = self._sino(np.float32)
            self.lib().og_siddon_f32(C.byref(self._g), _f32p(v), _f32p(out))
        return out

    def siddon_T(self, sino):
        s = np.ascontiguousarray(sino, dtype=np.float64)
        out = self._vol(np.float64)
        self.lib().og_siddon_T_f64(C.byref(self._g), _f64p(s), _f64p(out))
        return out

    def siddon_count(self):
        return int(self.lib().og_siddon_count(C.byref(self._g)))

    def siddon_compat(self, vol, threads=0):
        v = np.ascontiguousarray(vol, dtype=np.float32)
        out = self._sino(np.float32)
        rc = self.lib().og_siddon_compat(C.byref(self._g), _f32p(v), _f32p(out), int(threads))
        if rc != 0:
            raise ValueError("compat Siddon needs a 256^3 volume and at most 256x256 detector cells")
        return out

    def gd(self, sino, vol0=None, pair=0, cube=False, compat=False, iters=3, kopt=10, lam=0.0,
           positivity=0, alpha_min_init=-0.0347711, matched=False, weight=None):
        s = np.ascontiguousarray(sino, dtype=np.float32)
        v = self._vol(np.float32) if vol0 is None else np.array(vol0, dtype=np.float32, order="C")
        p = OgGdParams(pair=pair, cube=int(cube), compat=int(compat), iters=iters, kopt=kopt, lam=lam,
                       positivity=positivity, alpha_min_init=alpha_min_init, matched=int(matched))
        stats = (OgGdStat * max(iters, 1))()
        if weight is None:
            rc = self.lib().og_gd_run(C.byref(self._g), C.byref(p), _f32p(s), _f32p(v), stats)
        else:
            w = np.ascontiguousarray(weight, dtype=np.float32)
            rc = self.lib().og_wgd_run(C.byref(self._g), C.byref(p), _f32p(s), _f32p(w), _f32p(v), stats)
        assert rc == 0
        return v, [dict(residual_sq=x.residual_sq, norm_dj=x.norm_dj, norm_dj_proj=x.norm_dj_proj,
                        alpha=x.alpha, j_reg=x.j_reg, beta=x.beta) for x in stats[:iters]]

    def cg(self, sino, vol0=None, pair=0, cube=False, iters=3, lam=0.0, positivity=0, matched=False):
        s = np.ascontiguousarray(sino, dtype=np.float32)
        v = self._vol(np.float32) if vol0 is None else np.array(vol0, dtype=np.float32, order="C")
        p = OgGdParams(pair=pair, cube=int(cube), compat=0, iters=iters, kopt=0, lam=lam, positivity=positivity,
                       alpha_min_init=0.0, matched=int(matched))
        stats = (OgGdStat * max(iters, 1))()
        assert self.lib().og_cg_run(C.byref(self._g), C.byref(p), _f32p(s), _f32p(v), stats) == 0
        return v, [dict(residual_sq=x.residual_sq, norm_dj=x.norm_dj, norm_dj_proj=x.norm_dj_proj,
                        alpha=x.alpha, j_reg=x.j_reg, beta=x.beta) for x in stats[:iters]]

    def sirt_weights(self, pair=0, cube=False, relax=1.0):
        p = OgSirtParams(pair=pair, cube=int(cube), iters=0, relax=relax, positivity=0)
        rw, cw = self._sino(np.float32), self._vol(np.float32)
        assert self.lib().og_sirt_weights(C.byref(self._g), C.byref(p), _f32p(rw), _f32p(cw)) == 0
        return rw, cw

    def sirt(self, sino, pair=0, cube=False, iters=3, relax=1.0, positivity=0):
        s = np.ascontiguousarray(sino, dtype=np.float32)
        v = self._vol(np.float32)
        res = np.zeros(max(iters, 1), dtype=np.float64)
        p = OgSirtParams(pair=pair, cube=int(cube), iters=iters, relax=relax, positivity=positivity)
        assert self.lib().og_sirt_run(C.byref(self._g), C.byref(p), _f32p(s), _f32p(v), _f64p(res)) == 0
        return v, list(res[:iters])


def laplacian(x):
    x = np.ascontiguousarray(x)
    nz, ny, nx = x.shape
    out = np.empty_like(x)
    if x.dtype == np.float32:
        Oracle.lib().og_laplacian_f32(nx, ny, nz, _f32p(x), _f32p(out))
    else:
        x = x.astype(np.float64)
        out = np.empty_like(x)
        Oracle.lib().og_laplacian_f64(nx, ny, nz, _f64p(x), _f64p(out))
    return out


def shepp_logan(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    out = np.zeros((nz, ny, nx), dtype=np.float32)
    Oracle.lib().og_shepp_logan(nx, ny, nz, _f32p(out))
    return out


def ellipsoids(nx, ny=None, nz=None, count=64, seed=0, amin=5.0, amax=120.0):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    out = np.zeros((nz, ny, nx), dtype=np.float32)
    Oracle.lib().og_ellipsoids(nx, ny, nz, count, seed, amin, amax, _f32p(out))
    return out


def ellipsoid_table(nx, ny, nz, count, seed, amin, amax):
    p7 = np.zeros((count, 7))
    r9 = np.zeros((count, 9))
    Oracle.lib().og_ellipsoid_table(nx, ny, nz, count, seed, amin, amax, _f64p(p7), _f64p(r9))
    return p7, r9.reshape(count, 3, 3)


def set_threads(n: int):
    Oracle.lib().og_set_threads(int(n))
