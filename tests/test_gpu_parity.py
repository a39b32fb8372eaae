"""GPU parity: the HIP operators against the CPU oracle on the same inputs.

The FP32 kernels follow the oracle's FP32 recipe operation for operation
(FP64 per-ray / per-column geometry, explicit fmaf, sequential sums), so the
forward operators are compared bit for bit (tolerance 0), with a relative
L2/Linf fallback stated per test where summation order legitimately
differs.  Transposes use FP32 atomics (order not fixed) and are compared
with the FP64 oracle at rel-L2 <= 1e-5.
"""
import numpy as np
import pytest
import torch

from oracle_ffi import Geometry, Oracle, ellipsoids, shepp_logan, scaled_geometry, laplacian as og_laplacian
from tomogpi_amd import lib

pytestmark = pytest.mark.gpu


def to_desc(g: Geometry) -> lib.AcquisitionDesc:
    return lib.AcquisitionDesc(
        fod_mm=g.fod, fdd_mm=g.fdd, projection_nb=g.nphi, u_detector_size_mm=g.su, v_detector_size_mm=g.sv,
        u_pixel_nb=g.nu, v_pixel_nb=g.nv, x_volume_size_mm=g.sx, y_volume_size_mm=g.sy, z_volume_size_mm=g.sz,
        x_pixel_nb=g.nx, y_pixel_nb=g.ny, z_pixel_nb=g.nz, num_start_angle=g.num_start,
        den_start_angle=g.den_start, num_stop_angle=g.num_stop, den_stop_angle=g.den_stop)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_linf(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# Geometries: a cube case and a ragged one (odd sizes, non-square detector,
# partial angular range, non-zero start angle).
GEOMS = {
    "cube48": scaled_geometry(48, 24),
    # full waves of 64 detector columns, views on both sides of 45 degrees
    "cube128": scaled_geometry(128, 60),
    "ragged": Geometry(fod=230.0, fdd=470.0, nphi=37, su=61 * 1.7, sv=29 * 2.3, nu=61, nv=29, sx=41 * 0.9,
                       sy=35 * 0.9, sz=23 * 0.9, nx=41, ny=35, nz=23, num_start=1.0, den_start=3.0,
                       num_stop=5.0, den_stop=4.0),
}


@pytest.fixture(scope="module", params=list(GEOMS))
def setup(request, ctx):
    g = GEOMS[request.param]
    print(f"[case] geometry={request.param}", flush=True)
    return g, Oracle(g), lib.Geometry(ctx, to_desc(g))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_geometry_info_matches_oracle(setup):
    g, o, G = setup
    oi = o.info
    for k in ("du", "dv", "cu", "cv", "delta", "cx", "cy", "cz"):
        assert G.info[k] == pytest.approx(oi[k], rel=0, abs=0), k
    assert np.array_equal(G.phi(), o.phi)


def test_vi_bit_exact(setup):
    g, o, G = setup
    rng = np.random.default_rng(11)
    sino = rng.random(g.sino_shape, dtype=np.float32)
    ref = o.vi(sino)
    out = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
    G.backproject(dev(sino), out)
    got = out.cpu().numpy()
    assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))


def test_vi_fdk_bit_exact(setup):
    """FDK-weighted VI (TOMO_FLAG_FDK, the reference's back_fdk)."""
    g, o, G = setup
    rng = np.random.default_rng(12)
    sino = rng.random(g.sino_shape, dtype=np.float32)
    ref = o.vi_fdk(sino)
    out = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
    G.backproject(dev(sino), out, flags=lib.FLAG_FDK)
    got = out.cpu().numpy()
    assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))


@pytest.mark.parametrize("staged", [False, True])
@pytest.mark.parametrize("cube", [False, True])
def test_rs_bit_exact(setup, cube, staged):
    """RS with per-lane gathers from the z-pair copies (the default) and with
    LDS-staged taps (TOMO_FLAG_STAGED), both bit-exact."""
    g, o, G = setup
    vol = ellipsoids(g.nx, g.ny, g.nz, count=12, seed=3, amin=2.0, amax=12.0)
    ref = o.rs(vol, cube=cube)
    out = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
    flags = (lib.FLAG_RS_FOV_CUBE if cube else 0) | (lib.FLAG_STAGED if staged else 0)
    G.project(dev(vol), out, lib.PROJ_RS, flags)
    got = out.cpu().numpy()
    assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))


# Detector cells of 4 mm over 1 mm voxels at magnification 2: a workgroup of
# 32 x 8 cells spans up to 64 voxel columns, so the LDS-staged projectors
# shrink their windows to 6, 4, 2 and 1 planes and read many windows from the
# volume directly (their boxes do not fit a buffer even at one plane), beside
# full 8-plane windows.
COARSE = Geometry(fod=100.0, fdd=200.0, nphi=12, su=64 * 4.0, sv=24 * 4.0, nu=64, nv=24, sx=96.0, sy=96.0,
                  sz=40.0, nx=96, ny=96, nz=40)


@pytest.mark.parametrize("op", ["rs", "rs_cube", "siddon"])
def test_staged_window_fallbacks_bit_exact(ctx, op):
    """LDS-staged RS and Siddon on COARSE, where every window size and the
    direct reads occur, bit-exact against the oracle."""
    g = COARSE
    print(f"[case] geometry=coarse op={op}", flush=True)
    o = Oracle(g)
    G = lib.Geometry(ctx, to_desc(g))
    try:
        vol = ellipsoids(g.nx, g.ny, g.nz, count=12, seed=8, amin=3.0, amax=24.0)
        out = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
        if op == "siddon":
            ref = o.siddon(vol)
            G.project(dev(vol), out, lib.PROJ_SIDDON, lib.FLAG_STAGED)
        else:
            cube = op == "rs_cube"
            ref = o.rs(vol, cube=cube)
            G.project(dev(vol), out, lib.PROJ_RS, lib.FLAG_STAGED | (lib.FLAG_RS_FOV_CUBE if cube else 0))
        got = out.cpu().numpy()
        assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))
    finally:
        G.close()


@pytest.mark.parametrize("staged", [False, True])
def test_siddon_bit_exact(setup, staged):
    """Clean Siddon with per-lane gathers from the z-pair copies (the
    default) and with LDS-staged quads (TOMO_FLAG_STAGED)."""
    g, o, G = setup
    rng = np.random.default_rng(5)
    vol = rng.random(g.vol_shape, dtype=np.float32)
    ref = o.siddon(vol)
    out = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
    G.project(dev(vol), out, lib.PROJ_SIDDON, lib.FLAG_STAGED if staged else 0)
    got = out.cpu().numpy()
    assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))


@pytest.mark.parametrize("dt", ["f16", "bf16"])
@pytest.mark.parametrize("op", ["rs", "rs_cube", "siddon"])
def test_16bit_volume_bit_exact(setup, dt, op):
    """16-bit volumes at storage scale 1 projected into FP32 sinograms: the
    sums are the FP32 oracle's on the rounded voxel values, bit for bit (both
    projectors read 16-bit z-pair copies)."""
    g, o, G = setup
    rng = np.random.default_rng(9)
    vol = ellipsoids(g.nx, g.ny, g.nz, count=12, seed=4, amin=2.0, amax=12.0)
    vol = vol + rng.random(g.vol_shape, dtype=np.float32)
    v16 = dev(vol).to(torch.float16 if dt == "f16" else torch.bfloat16)
    rounded = v16.float().cpu().numpy()
    ref = o.siddon(rounded) if op == "siddon" else o.rs(rounded, cube=op == "rs_cube")
    out = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
    G.project(v16, out, lib.PROJ_SIDDON if op == "siddon" else lib.PROJ_RS,
              lib.FLAG_RS_FOV_CUBE if op == "rs_cube" else 0)
    got = out.cpu().numpy()
    assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))


# Detector rows of a quarter voxel at magnification 2: the row step A of a
# column is about 4, so VI's detector boxes exceed the LDS staging buffer and
# every view reads its taps directly.
FINE_ROWS = Geometry(fod=200.0, fdd=400.0, nphi=12, su=40 * 0.5, sv=40 * 0.5, nu=40, nv=40, sx=24.0, sy=24.0,
                     sz=24.0, nx=24, ny=24, nz=24, num_start=0.0, den_start=1.0, num_stop=2.0, den_stop=1.0)


# An FP16 sinogram is interpolated from u-pair words (two FP16 taps per
# word) with FP16 u-weights rounded to steps of 1/2048, so each bilinear
# sample differs from the FP32 form's by at most 2^-12 of the u-difference
# of its taps; against the FP32 oracle on the same (widened) taps the
# backprojection stays within these bounds.
F16_VI_REL_L2 = 1e-4
F16_VI_REL_LINF = 1e-3


def _vi16_case(g, o, G, dt, scale=1.0):
    """16-bit sinogram stored as value * scale (a power of two) into an FP32
    volume.  BF16: VI widens the taps to FP32 (stored / scale, exact) and
    equals the FP32 oracle on the widened values, bit for bit.  FP16: the
    u-pair form, within F16_VI_REL_L2 / F16_VI_REL_LINF of the same."""
    x = o.rs(ellipsoids(g.nx, g.ny, g.nz, count=12, seed=6, amin=2.0, amax=12.0))
    s16 = (dev(x) * scale).to(torch.float16 if dt == "f16" else torch.bfloat16)
    widened = (s16.double() / scale).float().cpu().numpy()
    ref = o.vi(widened)
    out = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
    G.backproject(s16, out, sino_scale=scale)
    got = out.cpu().numpy()
    if dt == "bf16":
        assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))
    else:
        e2, ei = rel_l2(got, ref), rel_linf(got, ref)
        print(f"[f16 vi] rel-L2 {e2:.2e} rel-Linf {ei:.2e} against the FP32 oracle on the same taps", flush=True)
        assert e2 <= F16_VI_REL_L2 and ei <= F16_VI_REL_LINF, (e2, ei)


@pytest.mark.parametrize("dt,scale", [("f16", 1.0), ("bf16", 1.0), ("f16", 2.0 ** -6)])
def test_16bit_vi(setup, dt, scale):
    g, o, G = setup
    _vi16_case(g, o, G, dt, scale)


@pytest.mark.parametrize("dt", ["f16", "bf16"])
def test_16bit_vi_fine_rows(ctx, dt):
    """Every view read directly (boxes over the staging buffer)."""
    G = lib.Geometry(ctx, to_desc(FINE_ROWS))
    try:
        _vi16_case(FINE_ROWS, Oracle(FINE_ROWS), G, dt)
    finally:
        G.close()


# Detectors wider than the volume's shadow (every u-tap of every voxel-view on
# the detector), one with rows of about one voxel (views staged in LDS), one
# with rows of about 2.5 voxels near the source (views staged and read directly).
WIDE_DET = {
    "staged": Geometry(fod=100.0, fdd=200.0, nphi=24, su=180.0, sv=100.0, nu=90, nv=50, sx=32.0, sy=32.0,
                       sz=32.0, nx=32, ny=32, nz=32),
    "direct": Geometry(fod=60.0, fdd=160.0, nphi=24, su=300.0, sv=300.0, nu=300, nv=300, sx=32.0, sy=32.0,
                       sz=32.0, nx=32, ny=32, nz=32),
}


@pytest.mark.parametrize("name", list(WIDE_DET))
@pytest.mark.parametrize("fdk", [False, True])
@pytest.mark.parametrize("scale", [1.0, 2.0 ** -6])
def test_f16_vi_constant_sinogram_exact(ctx, name, fdk, scale):
    """FP16 u-pair VI of a constant sinogram equals the FP32 oracle bit for
    bit: the FP16 u-weights sum to exactly 1, so each row interpolates to the
    constant and the v-interpolation, FDK weight and sums are the FP32 form's."""
    g = WIDE_DET[name]
    print(f"[case] geometry={name} fdk={fdk} scale={scale}", flush=True)
    G = lib.Geometry(ctx, to_desc(g))
    try:
        o = Oracle(g)
        sino = np.full(g.sino_shape, 0.75, dtype=np.float32)
        ref = o.vi_fdk(sino) if fdk else o.vi(sino)
        s16 = (dev(sino) * scale).to(torch.float16)
        out = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
        G.backproject(s16, out, flags=lib.FLAG_FDK if fdk else 0, sino_scale=scale)
        got = out.cpu().numpy()
        assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))
    finally:
        G.close()


# Short source distance: the row step A runs from 1.3 to 4.6 over the volume,
# so within one launch some workgroup-views stage their taps in LDS and the
# others read them directly; on FINE_ROWS every view reads them directly.
NEAR_SOURCE = Geometry(fod=80.0, fdd=160.0, nphi=16, su=160.0, sv=160.0, nu=160, nv=160, sx=64.0, sy=64.0,
                       sz=64.0, nx=64, ny=64, nz=64)


@pytest.mark.parametrize("name", ["near_source", "fine_rows"])
@pytest.mark.parametrize("fdk", [False, True])
def test_vi_bit_exact_unstaged_views(ctx, name, fdk):
    """FP32 VI where the taps of some or all views do not fit the staging buffer."""
    g = NEAR_SOURCE if name == "near_source" else FINE_ROWS
    print(f"[case] geometry={name} fdk={fdk}", flush=True)
    G = lib.Geometry(ctx, to_desc(g))
    try:
        o = Oracle(g)
        rng = np.random.default_rng(13)
        sino = rng.random(g.sino_shape, dtype=np.float32)
        ref = o.vi_fdk(sino) if fdk else o.vi(sino)
        out = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
        G.backproject(dev(sino), out, flags=lib.FLAG_FDK if fdk else 0)
        got = out.cpu().numpy()
        assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))
    finally:
        G.close()


# Degenerate shapes: one voxel, one view and one detector cell; a one-slice
# volume under a two-row detector; a three-column volume 40 voxels tall
# (three z-tiles, the last partial) under a 4 x 50 detector.
EDGE_GEOMS = {
    "single": Geometry(fod=10.0, fdd=20.0, nphi=1, su=2.0, sv=2.0, nu=1, nv=1, sx=1.0, sy=1.0, sz=1.0, nx=1, ny=1,
                       nz=1),
    "thin": Geometry(fod=30.0, fdd=60.0, nphi=3, su=18.0, sv=4.0, nu=9, nv=2, sx=7.0, sy=5.0, sz=1.0, nx=7, ny=5,
                     nz=1, num_start=1.0, den_start=4.0),
    "tall": Geometry(fod=40.0, fdd=80.0, nphi=5, su=8.0, sv=100.0, nu=4, nv=50, sx=3.0, sy=3.0, sz=40.0, nx=3, ny=3,
                     nz=40),
}


@pytest.mark.parametrize("name", list(EDGE_GEOMS))
def test_degenerate_shapes_bit_exact(ctx, name):
    g = EDGE_GEOMS[name]
    print(f"[case] geometry={name}", flush=True)
    G = lib.Geometry(ctx, to_desc(g))
    try:
        o = Oracle(g)
        rng = np.random.default_rng(21)
        sino = rng.random(g.sino_shape, dtype=np.float32)
        vol = rng.random(g.vol_shape, dtype=np.float32)
        outv = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
        outs = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
        cases = [("vi", o.vi(sino), lambda: G.backproject(dev(sino), outv), outv),
                 ("vi_fdk", o.vi_fdk(sino), lambda: G.backproject(dev(sino), outv, flags=lib.FLAG_FDK), outv),
                 ("rs", o.rs(vol), lambda: G.project(dev(vol), outs, lib.PROJ_RS), outs),
                 ("rs_cube", o.rs(vol, cube=True), lambda: G.project(dev(vol), outs, lib.PROJ_RS,
                                                                       lib.FLAG_RS_FOV_CUBE), outs),
                 ("siddon", o.siddon(vol), lambda: G.project(dev(vol), outs, lib.PROJ_SIDDON), outs)]
        for op, ref, run, out in cases:
            run()
            got = out.cpu().numpy()
            assert np.array_equal(got, ref), (op, rel_l2(got, ref), rel_linf(got, ref))
    finally:
        G.close()


@pytest.mark.parametrize("dt", ["f16", "bf16"])
@pytest.mark.parametrize("name", list(EDGE_GEOMS))
def test_degenerate_shapes_16bit_bit_exact(ctx, name, dt):
    """The degenerate shapes with 16-bit volumes: RS reads A-pair copies (one
    plane of groups more than the volume has along the slab axis, zero planes
    at both ends), Siddon 16-bit z-pair copies, both through 32-bit buffer
    offsets and through 64-bit element indices; every sum equals the FP32
    oracle's on the rounded voxel values, bit for bit."""
    g = EDGE_GEOMS[name]
    print(f"[case] geometry={name} volume={dt}", flush=True)
    G = lib.Geometry(ctx, to_desc(g))
    try:
        o = Oracle(g)
        rng = np.random.default_rng(22)
        v16 = dev(rng.random(g.vol_shape, dtype=np.float32)).to(torch.float16 if dt == "f16" else torch.bfloat16)
        rounded = v16.float().cpu().numpy()
        refs = {"rs": o.rs(rounded), "rs_cube": o.rs(rounded, cube=True), "siddon": o.siddon(rounded)}
        out = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
        for op, ref in refs.items():
            for index_reads in (False, True):
                flags = (lib.FLAG_RS_FOV_CUBE if op == "rs_cube" else 0) | (lib.FLAG_INDEX_READS if index_reads else 0)
                G.project(v16, out, lib.PROJ_SIDDON if op == "siddon" else lib.PROJ_RS, flags)
                got = out.cpu().numpy()
                assert np.array_equal(got, ref), (op, index_reads, rel_l2(got, ref), rel_linf(got, ref))
    finally:
        G.close()


@pytest.mark.parametrize("op", ["rs", "rs_cube", "siddon", "vi"])
def test_adjoint_dot(setup, ctx, op):
    """<A x, y> = <x, A^T y> with the GPU operator and its GPU transpose (rel 1e-5)."""
    g, o, G = setup
    rng = np.random.default_rng(1)
    x = dev(rng.random(g.vol_shape, dtype=np.float32))
    y = dev(rng.random(g.sino_shape, dtype=np.float32))
    if op == "vi":
        bx = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
        G.backproject(y, bx)
        bty = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
        G.backproject_adjoint(x, bty)
        lhs, rhs = lib.dot(ctx, bx, x), lib.dot(ctx, y, bty)
    else:
        proj = lib.PROJ_SIDDON if op == "siddon" else lib.PROJ_RS
        flags = lib.FLAG_RS_FOV_CUBE if op == "rs_cube" else 0
        ax = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
        G.project(x, ax, proj, flags)
        aty = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
        G.project_adjoint(y, aty, proj, flags)
        lhs, rhs = lib.dot(ctx, ax, y), lib.dot(ctx, x, aty)
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) <= 1e-5, (lhs, rhs)


@pytest.mark.parametrize("op", ["rs", "siddon", "vi"])
def test_transpose_vs_oracle(setup, op):
    g, o, G = setup
    rng = np.random.default_rng(2)
    if op == "vi":
        x = rng.random(g.vol_shape, dtype=np.float32)
        ref = o.vi_T(x)
        out = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
        G.backproject_adjoint(dev(x), out)
    else:
        y = rng.random(g.sino_shape, dtype=np.float32)
        ref = o.rs_T(y) if op == "rs" else o.siddon_T(y)
        out = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
        G.project_adjoint(dev(y), out, lib.PROJ_RS if op == "rs" else lib.PROJ_SIDDON)
    assert rel_l2(out.cpu().numpy(), ref) <= 1e-5


def test_work_counters(setup):
    g, o, G = setup
    assert G.project_work(lib.PROJ_RS) == o.rs_count()
    assert G.project_work(lib.PROJ_RS, lib.FLAG_RS_FOV_CUBE) == o.rs_count(cube=True)
    assert G.project_work(lib.PROJ_SIDDON) == o.siddon_count()


def test_phantoms_bit_exact(setup):
    g, o, G = setup
    out = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
    G.shepp_logan(out)
    assert np.array_equal(out.cpu().numpy(), shepp_logan(g.nx, g.ny, g.nz))
    G.ellipsoids(out, count=9, seed=17, amin=2.0, amax=9.0)
    assert np.array_equal(out.cpu().numpy(), ellipsoids(g.nx, g.ny, g.nz, count=9, seed=17, amin=2.0, amax=9.0))


def test_laplacian_bit_exact(setup):
    g, o, G = setup
    x = np.random.default_rng(4).random(g.vol_shape, dtype=np.float32)
    y = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
    G.laplacian(dev(x), y)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), og_laplacian(x))


def test_host_pointer_path(setup):
    """TOMO_FLAG_HOST_PTR (the reference's mem_CPU storage) gives the device result."""
    g, o, G = setup
    sino = np.random.default_rng(9).random(g.sino_shape, dtype=np.float32)
    out = np.zeros(g.vol_shape, dtype=np.float32)
    G.backproject(sino, out)
    assert np.array_equal(out, o.vi(sino))


# ---- the reference's C-path Siddon (TOMO_FLAG_SIDDON_COMPAT) -------------
# SiddonProjector_compute_C_mem_CPU.cpp:20-276 runs only at 256^3 (it indexes
# % 256); the compat kernel restates its float arithmetic and quirks and must
# equal the oracle's restatement bit for bit.

def test_siddon_compat_bit_exact_coarse_detector(ctx):
    """256^3, 32 x 32 detector cells of 16 mm, 8 views."""
    g = scaled_geometry(256, 8, nu=32, nv=32, pixel_mm=16.0)
    G = lib.Geometry(ctx, to_desc(g))
    try:
        vol = np.random.default_rng(31).random(g.vol_shape, dtype=np.float32)
        ref = Oracle(g).siddon_compat(vol)
        out = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
        G.project(dev(vol), out, lib.PROJ_SIDDON, lib.FLAG_SIDDON_COMPAT)
        got = out.cpu().numpy()
        assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))
    finally:
        G.close()


# ---- index readers (the form of copies of 4 GiB or more) ----------------
@pytest.mark.parametrize("dt", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("op", ["rs", "rs_cube", "siddon"])
def test_index_reads_bit_exact(setup, op, dt):
    """TOMO_FLAG_INDEX_READS makes the projectors read their copies (z-pair;
    A-pair for RS on 16-bit volumes) through 64-bit element indices, as every
    1024^3 volume does; the results equal the oracle's (and the buffer
    readers') bit for bit."""
    g, o, G = setup
    rng = np.random.default_rng(8)
    vol = ellipsoids(g.nx, g.ny, g.nz, count=12, seed=5, amin=2.0, amax=12.0)
    vol = vol + rng.random(g.vol_shape, dtype=np.float32)
    dv = dev(vol)
    if dt != "f32":
        dv = dv.to(torch.bfloat16 if dt == "bf16" else torch.float16)
        vol = dv.float().cpu().numpy()
    ref = o.siddon(vol) if op == "siddon" else o.rs(vol, cube=op == "rs_cube")
    flags = lib.FLAG_INDEX_READS | (lib.FLAG_RS_FOV_CUBE if op == "rs_cube" else 0)
    out = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
    G.project(dv, out, lib.PROJ_SIDDON if op == "siddon" else lib.PROJ_RS, flags)
    got = out.cpu().numpy()
    assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))


# ---- the 256^3 Shepp-Logan workload at its own size ------------------------
# 256^3 Shepp-Logan (densities 0..2), 256 views over 360 deg, 256 x 256
# detector of 2 mm, FOD 500 mm, FDD 1000 mm: one projection of each kind and
# one backprojection against the oracle, bit for bit, and the adjoint test.

@pytest.fixture(scope="module")
def c0(ctx):
    g = scaled_geometry(256, 256)
    print("[case] 256^3 Shepp-Logan, 256 views", flush=True)
    G = lib.Geometry(ctx, to_desc(g))
    yield g, Oracle(g), G, shepp_logan(256)
    G.close()


@pytest.mark.parametrize("op", ["rs", "rs_cube", "siddon", "siddon_compat"])
def test_c0_projection_bit_exact(c0, op):
    g, o, G, vol = c0
    print(f"[case] c0 projection {op}", flush=True)
    if op == "siddon_compat":
        ref = o.siddon_compat(vol)
    elif op == "siddon":
        ref = o.siddon(vol)
    else:
        ref = o.rs(vol, cube=op == "rs_cube")
    flags = {"rs": 0, "rs_cube": lib.FLAG_RS_FOV_CUBE, "siddon": 0, "siddon_compat": lib.FLAG_SIDDON_COMPAT}[op]
    out = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
    G.project(dev(vol), out, lib.PROJ_RS if op.startswith("rs") else lib.PROJ_SIDDON, flags)
    got = out.cpu().numpy()
    assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))


def test_c0_rsvi_backprojection_bit_exact(c0):
    """VI of the RS sinogram of the phantom (the RSVI pair), bit for bit."""
    g, o, G, vol = c0
    sino = o.rs(vol)
    ref = o.vi(sino)
    out = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
    G.backproject(dev(sino), out)
    got = out.cpu().numpy()
    assert np.array_equal(got, ref), (rel_l2(got, ref), rel_linf(got, ref))


@pytest.mark.parametrize("op", ["rs", "rs_cube", "siddon", "vi"])
def test_c0_adjoint_dot(c0, ctx, op):
    """<A x, y> = <x, A^T y> at 256^3 / 256 views, x, y ~ U[0,1) (seeds 1, 2)."""
    g, o, G, _ = c0
    x = dev(np.random.default_rng(1).random(g.vol_shape, dtype=np.float32))
    y = dev(np.random.default_rng(2).random(g.sino_shape, dtype=np.float32))
    if op == "vi":
        bx = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
        G.backproject(y, bx)
        bty = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
        G.backproject_adjoint(x, bty)
        lhs, rhs = lib.dot(ctx, bx, x), lib.dot(ctx, y, bty)
    else:
        proj = lib.PROJ_SIDDON if op == "siddon" else lib.PROJ_RS
        flags = lib.FLAG_RS_FOV_CUBE if op == "rs_cube" else 0
        ax = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
        G.project(x, ax, proj, flags)
        aty = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
        G.project_adjoint(y, aty, proj, flags)
        lhs, rhs = lib.dot(ctx, ax, y), lib.dot(ctx, x, aty)
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) <= 1e-5, (lhs, rhs)
