"""The 512^3 / 512-view workload at its own size.

Acquisition: 512^3 volume of 1 mm voxels holding 64 random ellipsoids (seed
0, semi-axes 5..120 mm), 512 views over 360 deg, 512 x 512 detector of 2 mm
pixels, FOD 1000 mm, FDD 2000 mm (the benchmark's sweep).  The oracle cannot
run the whole acquisition inside a test, so the checks are:

* views 0..3 of the full 512-view Siddon and RS projections against the
  oracle on a 4-view geometry with the same angular step, bit for bit.  The
  per-view tables are the same numbers: the step is pi (stop - start) / Nphi,
  pi * 2 / 512 for the full acquisition and pi * 2^-6 / 4 for the subset
  (stop = 8 / 512), both pi * 2^-8 exactly;
* VI on that 4-view geometry (full volume and detector) of the oracle's
  4-view Siddon sinogram, bit for bit;
* the adjoint identity <A x, y> = <x, A^T y> for Siddon, RS and VI on a
  16-view acquisition of the full volume and detector (rel 1e-5).

The 1024^3 acquisition of the mixed-precision workload (720 views over 360
deg, 768 x 512 detector of 4 mm pixels, FOD 2000 mm, FDD 4000 mm) is checked
on its first 4 views: its FP32 z-pair copies hold 2.15e9 elements (8.6 GB
each), so the projectors read them through 64-bit element indices, the
readers no smaller volume takes on its own.  The GPU runs the 4-view
geometry itself (pi * 2 / 720 is not a power-of-two multiple of the
subset's step), with the full volume and detector.
"""
import numpy as np
import pytest
import torch

from oracle_ffi import Oracle, scaled_geometry
from tomogpi_amd import lib

pytestmark = pytest.mark.gpu

N, NPHI, SUB = 512, 512, 4


def to_desc(g) -> lib.AcquisitionDesc:
    return lib.AcquisitionDesc(
        fod_mm=g.fod, fdd_mm=g.fdd, projection_nb=g.nphi, u_detector_size_mm=g.su, v_detector_size_mm=g.sv,
        u_pixel_nb=g.nu, v_pixel_nb=g.nv, x_volume_size_mm=g.sx, y_volume_size_mm=g.sy, z_volume_size_mm=g.sz,
        x_pixel_nb=g.nx, y_pixel_nb=g.ny, z_pixel_nb=g.nz, num_start_angle=g.num_start,
        den_start_angle=g.den_start, num_stop_angle=g.num_stop, den_stop_angle=g.den_stop)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def workload(ctx):
    g = scaled_geometry(N, NPHI)
    assert (g.fod, g.fdd, g.nu, g.nv, g.nphi) == (1000.0, 2000.0, 512, 512, 512)
    print("[case] 512^3 / 512 views: phantom and oracle on views 0..3", flush=True)
    G = lib.Geometry(ctx, to_desc(g))
    vol = torch.empty(G.vol_shape, dtype=torch.float32, device="cuda")
    G.ellipsoids(vol, count=64, seed=0, amin=5.0, amax=120.0)
    torch.cuda.synchronize()
    host = vol.cpu().numpy()
    gsub = scaled_geometry(N, SUB)
    gsub.num_stop, gsub.den_stop = 2.0 * SUB, float(NPHI)
    o = Oracle(gsub)
    ref = {"siddon": o.siddon(host), "rs": o.rs(host)}
    ref["vi"] = o.vi(ref["siddon"])
    del host
    yield G, gsub, vol, ref
    G.close()


@pytest.mark.parametrize("op", ["siddon", "rs"])
def test_full_acquisition_views_bit_exact(workload, op):
    G, _, vol, ref = workload
    print(f"[case] 512^3 / 512 views: {op} projection", flush=True)
    out = torch.empty(G.sino_shape, dtype=torch.float32, device="cuda")
    G.project(vol, out, lib.PROJ_SIDDON if op == "siddon" else lib.PROJ_RS)
    got = out[:SUB].cpu().numpy()
    assert got.shape == ref[op].shape
    assert np.array_equal(got, ref[op]), rel_l2(got, ref[op])
    # the rest of the acquisition is finite and non-negative (non-negative volume)
    tail = out[SUB:]
    assert bool(torch.isfinite(tail).all()) and float(tail.min()) >= 0.0


def test_full_size_vi_bit_exact(workload, ctx):
    G, gsub, _, ref = workload
    print("[case] 512^3 / 4-view subset: VI backprojection", flush=True)
    Gs = lib.Geometry(ctx, to_desc(gsub))
    try:
        sino = torch.from_numpy(np.ascontiguousarray(ref["siddon"])).cuda()
        out = torch.empty(Gs.vol_shape, dtype=torch.float32, device="cuda")
        Gs.backproject(sino, out)
        got = out.cpu().numpy()
    finally:
        Gs.close()
    assert np.array_equal(got, ref["vi"]), rel_l2(got, ref["vi"])


@pytest.mark.parametrize("op", ["siddon", "rs", "vi"])
def test_full_size_adjoint_dot(ctx, op):
    g = scaled_geometry(N, 16)
    print(f"[case] 512^3 / 16 views: adjoint {op}", flush=True)
    G = lib.Geometry(ctx, to_desc(g))
    try:
        gen = torch.Generator(device="cuda")
        gen.manual_seed(1)
        x = torch.rand(G.vol_shape, generator=gen, device="cuda")
        gen.manual_seed(2)
        y = torch.rand(G.sino_shape, generator=gen, device="cuda")
        if op == "vi":
            bx = torch.empty(G.vol_shape, dtype=torch.float32, device="cuda")
            G.backproject(y, bx)
            bty = torch.empty(G.sino_shape, dtype=torch.float32, device="cuda")
            G.backproject_adjoint(x, bty)
            lhs, rhs = lib.dot(ctx, bx, x), lib.dot(ctx, y, bty)
        else:
            proj = lib.PROJ_SIDDON if op == "siddon" else lib.PROJ_RS
            ax = torch.empty(G.sino_shape, dtype=torch.float32, device="cuda")
            G.project(x, ax, proj)
            aty = torch.empty(G.vol_shape, dtype=torch.float32, device="cuda")
            G.project_adjoint(y, aty, proj)
            lhs, rhs = lib.dot(ctx, ax, y), lib.dot(ctx, x, aty)
    finally:
        G.close()
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) <= 1e-5, (lhs, rhs)


@pytest.mark.parametrize("op", ["siddon", "rs", "vi"])
def test_1024_index_readers_bit_exact(ctx, op):
    n = 1024
    g = scaled_geometry(n, SUB, nu=768, nv=512, pixel_mm=2.0 * n / 512.0)
    g.num_stop, g.den_stop = 2.0 * SUB, 720.0
    print(f"[case] 1024^3 / 4 of 720 views, 768x512 detector: {op}", flush=True)
    G = lib.Geometry(ctx, to_desc(g))
    try:
        vol = torch.empty(G.vol_shape, dtype=torch.float32, device="cuda")
        G.ellipsoids(vol, count=64, seed=7, amin=10.0, amax=240.0)
        torch.cuda.synchronize()
        host = vol.cpu().numpy()
        o = Oracle(g)
        if op == "vi":
            sino_ref = o.siddon(host)
            del host, vol
            ref = o.vi(sino_ref)
            out = torch.empty(G.vol_shape, dtype=torch.float32, device="cuda")
            G.backproject(torch.from_numpy(sino_ref).cuda(), out)
        else:
            ref = o.siddon(host) if op == "siddon" else o.rs(host)
            del host
            out = torch.empty(G.sino_shape, dtype=torch.float32, device="cuda")
            G.project(vol, out, lib.PROJ_SIDDON if op == "siddon" else lib.PROJ_RS)
        got = out.cpu().numpy()
        del out
    finally:
        G.close()
        torch.cuda.empty_cache()
    assert np.array_equal(got, ref), rel_l2(got, ref)
