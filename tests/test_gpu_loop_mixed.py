"""GPU: the simple-gradient loop against the oracle's loop, and the
mixed-precision storage forms against the FP64 oracle.

Loop tolerance: the operators, vector ops and Laplacian are bit-identical to
the oracle; the FP64 norms are summed in a different order (block partials
vs OpenMP), so alpha can differ in its last bits and the volumes are compared
at rel-L2 <= 1e-6, the step sizes at rel 1e-9.

Mixed forms: FP16 (scaled to put max|x| at 2^14) and BF16 storage of inputs
and outputs, FP32 arithmetic, FP64 geometry.  Error is stated per result as
relative L2 against the FP64 oracle on the FP32 input; the bound asserted is
the mixed form's 1e-3 for FP16 and 4e-3 for BF16 (BF16 keeps 8 significant
bits: rounding a stored result alone costs about 1e-3 relative).

The weighted gradient (tomo_wgd_run) is compared with the oracle's weighted
loop under the same loop tolerance.
"""
import numpy as np
import pytest
import torch

from oracle_ffi import Oracle, ellipsoids, scaled_geometry
from test_gpu_parity import to_desc, rel_l2
from tomogpi_amd import lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(ctx):
    g = scaled_geometry(40, 20)
    return g, Oracle(g), lib.Geometry(ctx, to_desc(g))


@pytest.mark.parametrize("pair,flags,lam,kopt,positivity", [
    (lib.PAIR_RSVI, 0, 0.01, 10, 0),
    (lib.PAIR_RSVI, lib.FLAG_RS_FOV_CUBE, 0.0, 2, 1),
    (lib.PAIR_SIDDONVI, 0, 0.01, 3, 0),
])
def test_gd_loop_matches_oracle(small, pair, flags, lam, kopt, positivity):
    g, o, G = small
    print(f"[case] gd pair={pair} flags={flags} lam={lam} kopt={kopt}", flush=True)
    truth = ellipsoids(g.nx, g.ny, g.nz, count=6, seed=4, amin=2.0, amax=9.0)
    sino = o.rs(truth) if pair == lib.PAIR_RSVI else o.siddon(truth)
    ref_vol, ref_stats = o.gd(sino, pair=pair, cube=bool(flags & lib.FLAG_RS_FOV_CUBE), iters=4, kopt=kopt,
                              lam=lam, positivity=positivity)
    vol = torch.zeros(g.vol_shape, dtype=torch.float32, device="cuda")
    stats = G.gd_run(torch.from_numpy(sino).cuda(), vol, pair=pair, flags=flags, iters=4, kopt=kopt, lam=lam,
                     positivity=positivity)
    torch.cuda.synchronize()
    for it, (a, b) in enumerate(zip(stats, ref_stats)):
        for k in ("alpha", "residual_sq", "norm_dj", "norm_dj_proj", "j_reg", "beta"):
            assert a[k] == pytest.approx(b[k], rel=1e-9, abs=1e-12), (k, a, b)
        # JReg = sum (L x)^2 of the iterate before the update: 0 for the zero start
        assert (a["j_reg"] > 0) == (lam != 0 and it > 0)
    assert rel_l2(vol.cpu().numpy(), ref_vol) <= 1e-6
    assert stats[-1]["residual_sq"] < stats[0]["residual_sq"]


@pytest.mark.parametrize("pair", [lib.PAIR_RSVI, lib.PAIR_SIDDONVI])
def test_weighted_gd_loop_matches_oracle(small, pair):
    """tomo_wgd_run (the reference's doWeightedGradient) against the oracle's
    weighted loop: the same operators, FP64 norms summed in another order."""
    g, o, G = small
    rng = np.random.default_rng(21)
    sino = o.rs(ellipsoids(g.nx, count=6, seed=5, amin=2.0, amax=8.0))
    w = (0.25 + rng.random(g.sino_shape)).astype(np.float32)
    kw = dict(iters=3, kopt=10, lam=0.01, alpha_min_init=-999999999999.0)
    vol = torch.zeros(g.vol_shape, dtype=torch.float32, device="cuda")
    st = G.gd_run(torch.from_numpy(sino).cuda(), vol, pair=pair, weight=torch.from_numpy(w).cuda(), **kw)
    ref_vol, ref_st = o.gd(sino, pair=0 if pair == lib.PAIR_RSVI else 1, weight=w, **kw)
    assert rel_l2(vol.cpu().numpy(), ref_vol) <= 1e-6
    for a, b in zip(st, ref_st):
        assert a["alpha"] == pytest.approx(b["alpha"], rel=1e-6)
    # host pointers give the same volume
    vh = np.zeros(g.vol_shape, dtype=np.float32)
    G.gd_run(sino, vh, pair=pair, weight=w, **kw)
    assert np.array_equal(vh, vol.cpu().numpy())


def test_gd_loop_host_pointers(small):
    g, o, G = small
    truth = ellipsoids(g.nx, g.ny, g.nz, count=6, seed=4, amin=2.0, amax=9.0)
    sino = o.rs(truth)
    vol = np.zeros(g.vol_shape, dtype=np.float32)
    G.gd_run(sino, vol, iters=2, kopt=10, lam=0.01)
    ref_vol, _ = o.gd(sino, iters=2, kopt=10, lam=0.01)
    assert rel_l2(vol, ref_vol) <= 1e-6


def _to16(x, dtype, scale):
    t = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()
    return (t * scale).to(dtype)


def _from16(t, scale):
    return (t.float() / scale).cpu().numpy().astype(np.float64)


@pytest.fixture(scope="module")
def mid(ctx):
    g = scaled_geometry(96, 48)
    return g, Oracle(g), lib.Geometry(ctx, to_desc(g))


@pytest.mark.parametrize("dt", ["f16", "bf16"])
@pytest.mark.parametrize("op", ["vi", "rs", "siddon"])
def test_mixed_precision_error_vs_fp64(mid, dt, op):
    g, o, G = mid
    rng = np.random.default_rng(12)
    tdt = torch.float16 if dt == "f16" else torch.bfloat16
    code = lib.F16 if dt == "f16" else lib.BF16
    bound = 1e-3 if dt == "f16" else 4e-3
    if op == "vi":
        x = o.rs(ellipsoids(g.nx, count=10, seed=8, amin=4.0, amax=20.0))
        ref = o.vi(x.astype(np.float64), f64=True)
        sx = 2.0 ** 14 / np.abs(x).max() if dt == "f16" else 1.0
        sy = 2.0 ** 14 / np.abs(ref).max() if dt == "f16" else 1.0
        xin = _to16(x, tdt, sx)
        out = torch.empty(g.vol_shape, dtype=tdt, device="cuda")
        G.backproject(xin, out, sino_scale=sx, vol_scale=sy)
    else:
        x = ellipsoids(g.nx, count=10, seed=8, amin=4.0, amax=20.0) + rng.random(g.vol_shape).astype(np.float32) * 0.1
        ref = o.rs(x.astype(np.float64), f64=True) if op == "rs" else o.siddon(x.astype(np.float64), f64=True)
        sx = 2.0 ** 14 / np.abs(x).max() if dt == "f16" else 1.0
        sy = 2.0 ** 14 / np.abs(ref).max() if dt == "f16" else 1.0
        xin = _to16(x, tdt, sx)
        out = torch.empty(g.sino_shape, dtype=tdt, device="cuda")
        G.project(xin, out, lib.PROJ_RS if op == "rs" else lib.PROJ_SIDDON, vol_scale=sx, sino_scale=sy)
    torch.cuda.synchronize()
    err = rel_l2(_from16(out, sy), ref)
    print(f"[mixed] op={op} storage={dt} rel_l2_vs_fp64={err:.3e}", flush=True)
    assert err <= bound


def test_mixed_inputs_fp32_outputs_close_to_fp32(mid):
    """BF16-stored inputs with FP32 outputs: the only rounding is the input's."""
    g, o, G = mid
    x = ellipsoids(g.nx, count=10, seed=8, amin=4.0, amax=20.0)
    xin = _to16(x, torch.bfloat16, 1.0)
    out = torch.empty(g.sino_shape, dtype=torch.float32, device="cuda")
    G.project(xin, out, lib.PROJ_SIDDON)
    ref = o.siddon(x.astype(np.float64), f64=True)
    assert rel_l2(out.cpu().numpy(), ref) <= 2e-3


@pytest.mark.parametrize("pair,flags", [(lib.PAIR_RSVI, 0), (lib.PAIR_SIDDONVI, 0), (lib.PAIR_RSVI, lib.FLAG_RS_FOV_CUBE)])
def test_sirt_matches_oracle(small, pair, flags):
    """SIRT (x += C B R (y - P x)): operators and element-wise steps are the
    oracle's arithmetic, so the volume is bit-identical; the FP64 residuals
    are summed in a different order (rel 1e-9)."""
    g, o, G = small
    print(f"[case] sirt pair={pair} flags={flags}", flush=True)
    truth = ellipsoids(g.nx, g.ny, g.nz, count=6, seed=4, amin=2.0, amax=9.0)
    sino = o.rs(truth) if pair == lib.PAIR_RSVI else o.siddon(truth)
    ref, ref_res = o.sirt(sino, pair=pair, cube=bool(flags & lib.FLAG_RS_FOV_CUBE), iters=5, relax=0.9)
    vol = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
    res = G.sirt_run(torch.from_numpy(sino).cuda(), vol, pair=pair, flags=flags, iters=5, relaxation=0.9)
    torch.cuda.synchronize()
    assert np.array_equal(vol.cpu().numpy(), ref)
    assert res == pytest.approx(ref_res, rel=1e-9)
    assert res[-1] < res[0]


@pytest.mark.parametrize("pair,lam", [(lib.PAIR_RSVI, 0.01), (lib.PAIR_SIDDONVI, 0.0)])
def test_cg_loop_matches_oracle(small, pair, lam):
    """Conjugate gradient (tomo_cg_run) against the oracle's loop: same
    operators; FP64 norms and the g.d product summed in a different order."""
    g, o, G = small
    truth = ellipsoids(g.nx, g.ny, g.nz, count=6, seed=4, amin=2.0, amax=9.0)
    sino = o.rs(truth) if pair == lib.PAIR_RSVI else o.siddon(truth)
    ref, ref_stats = o.cg(sino, pair=pair, iters=5, lam=lam)
    vol = torch.zeros(g.vol_shape, dtype=torch.float32, device="cuda")
    stats = G.gd_run(torch.from_numpy(sino).cuda(), vol, pair=pair, iters=5, lam=lam, conjugate=True)
    torch.cuda.synchronize()
    for a, b in zip(stats, ref_stats):
        for k in ("alpha", "residual_sq", "norm_dj", "norm_dj_proj", "j_reg", "beta"):
            assert a[k] == pytest.approx(b[k], rel=1e-8, abs=1e-12), (k, a, b)
    assert rel_l2(vol.cpu().numpy(), ref) <= 1e-6
    assert stats[-1]["residual_sq"] < stats[0]["residual_sq"]


def test_gd_two_threads_two_streams_one_context(small, ctx):
    """Two threads, each on its own stream, run tomo_gd_run on one context at
    once (RSVI and SIDDONVI); each result equals the same call run alone, bit
    for bit (the loops' reductions take their scratch per call)."""
    import threading
    g, o, G = small
    truth = ellipsoids(g.nx, g.ny, g.nz, count=6, seed=7, amin=2.0, amax=9.0)
    jobs = [(lib.PAIR_RSVI, torch.from_numpy(o.rs(truth)).cuda()),
            (lib.PAIR_SIDDONVI, torch.from_numpy(o.siddon(truth)).cuda())]
    kw = dict(iters=6, kopt=7, lam=0.01)
    alone = []
    for pair, y in jobs:
        v = torch.zeros(g.vol_shape, dtype=torch.float32, device="cuda")
        st = G.gd_run(y, v, pair=pair, **kw)
        alone.append((v, st))
    torch.cuda.synchronize()
    for rep in range(3):
        outs = [torch.zeros(g.vol_shape, dtype=torch.float32, device="cuda") for _ in jobs]
        streams = [torch.cuda.Stream() for _ in jobs]
        res, errs = [None, None], []
        torch.cuda.synchronize()

        def run(i):
            try:
                res[i] = G.gd_run(jobs[i][1], outs[i], pair=jobs[i][0], stream=streams[i], **kw)
            except Exception as e:  # reported below
                errs.append(e)

        th = [threading.Thread(target=run, args=(i,)) for i in range(len(jobs))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        torch.cuda.synchronize()
        assert not errs, errs
        for i in range(len(jobs)):
            assert torch.equal(outs[i], alone[i][0]), (rep, i)
            assert res[i] == alone[i][1], (rep, i)


# Mixed-precision loop (tomo_gd_params.storage): every operator reads a 16-bit
# copy of its FP32 input, FP32 accumulation, FP32 state.  Its per-iteration
# residual sum (y - P x)^2 stays within these relative bounds of the FP32
# loop's, and so does the final volume (rel-L2).  A CPU emulation with the
# oracle's operators on rounded inputs measured 1.5e-5..4.2e-5 (FP16) and
# 2.1e-4..4.8e-4 (BF16) for the residuals at 40^3 and 64^3, 10 iterations.
MIXED_LOOP_TOL = {"f16": 1e-3, "bf16": 4e-3}


@pytest.mark.parametrize("dt", ["f16", "bf16"])
@pytest.mark.parametrize("pair", [lib.PAIR_RSVI, lib.PAIR_SIDDONVI])
@pytest.mark.parametrize("conjugate", [False, True])
def test_mixed_loop_residual_tracks_fp32(ctx, dt, pair, conjugate):
    g = scaled_geometry(64, 48)
    print(f"[case] mixed loop {dt} pair={pair} cg={conjugate}", flush=True)
    o = Oracle(g)
    G = lib.Geometry(ctx, to_desc(g))
    try:
        truth = ellipsoids(64, count=8, seed=4, amin=2.0, amax=19.0)
        y = torch.from_numpy(o.rs(truth) if pair == lib.PAIR_RSVI else o.siddon(truth)).cuda()
        kw = dict(pair=pair, iters=10, kopt=11, lam=0.01, conjugate=conjugate)
        x32 = torch.zeros(g.vol_shape, dtype=torch.float32, device="cuda")
        s32 = G.gd_run(y, x32, **kw)
        x16 = torch.zeros(g.vol_shape, dtype=torch.float32, device="cuda")
        s16 = G.gd_run(y, x16, storage=lib.F16 if dt == "f16" else lib.BF16, **kw)
        torch.cuda.synchronize()
        r32 = np.array([s["residual_sq"] for s in s32])
        r16 = np.array([s["residual_sq"] for s in s16])
        rel = np.abs(r16 - r32) / r32
        vol = rel_l2(x16.cpu().numpy(), x32.cpu().numpy())
        print(f"[mixed loop] {dt} pair={pair} cg={conjugate}: residual rel diff per iteration "
              f"{' '.join(f'{v:.1e}' for v in rel)}; final volume rel-L2 {vol:.2e}", flush=True)
        assert rel.max() <= MIXED_LOOP_TOL[dt], rel
        assert vol <= MIXED_LOOP_TOL[dt]
        assert r16[-1] < 0.05 * r16[0]
    finally:
        G.close()
