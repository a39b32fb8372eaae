"""GPU: the serving loop (tomo_serve_*).

* Every request's result is the same alone and in a mix of uneven sizes
  (bit-identical: a problem's arithmetic does not depend on what shares its
  launches), with and without graph replay, with one or several iterations
  per round, and under a memory budget that admits only part of the stream
  at a time.
* Direct schedule vs tomo_gd_run (the plain loop): the operators and vector
  ops are the same code; only the FP64 norms are summed in a different order
  (per-tile partials vs grid-stride partials), so volumes agree to rel-L2
  1e-6 and step sizes to rel 1e-9.
* Recurrent schedule (P x carried by linearity) vs the direct one: FP32
  rounding of the carried sinogram only, rel-L2 <= 1e-4 over 6 iterations.
"""
import numpy as np
import pytest
import torch

from oracle_ffi import Oracle, ellipsoids, scaled_geometry
from test_gpu_parity import rel_l2, to_desc
from tomogpi_amd import lib

pytestmark = pytest.mark.gpu

# (n, nphi, pair, lam, kopt, positivity)
CASES = [
    (24, 18, lib.PAIR_SIDDONVI, 0.01, 10, 0),
    (40, 30, lib.PAIR_RSVI, 0.01, 3, 0),
    (32, 24, lib.PAIR_SIDDONVI, 0.0, 10, 0),
    (48, 20, lib.PAIR_RSVI, 0.02, 10, 1),
    (28, 36, lib.PAIR_RSVI, 0.01, 10, 0),
]
ITERS = 6


def _problem(i, n, nphi, pair):
    g = scaled_geometry(n, nphi)
    o = Oracle(g)
    truth = ellipsoids(n, count=5, seed=100 + i, amin=2.0, amax=n / 3)
    sino = o.rs(truth) if pair == lib.PAIR_RSVI else o.siddon(truth)
    return g, sino


@pytest.fixture(scope="module")
def problems():
    return [_problem(i, *c[:3]) for i, c in enumerate(CASES)]


def _submit(srv, problems, idx, outs):
    ys = {}
    for i in idx:
        g, sino = problems[i]
        outs[i] = torch.zeros(g.vol_shape, dtype=torch.float32, device="cuda")
        ys[i] = torch.from_numpy(sino).cuda()
    torch.cuda.synchronize()  # inputs and outputs ready before the server's stream uses them
    ids = []
    for i in idx:
        g, sino = problems[i]
        n, nphi, pair, lam, kopt, pos = CASES[i]
        ids.append(srv.submit(to_desc(g), ys[i], outs[i], pair=pair, iters=ITERS, kopt=kopt, lam=lam,
                              positivity=pos))
    return ids


def _run(ctx, problems, idx, **kw):
    print(f"[case] serve {idx} {kw}", flush=True)
    srv = lib.Server(ctx, **kw)
    outs = {}
    ids = _submit(srv, problems, idx, outs)
    res = {i: srv.wait(rid, timeout=120.0) for i, rid in zip(idx, ids)}
    hist = {i: srv.iter_stats(rid, ITERS) for i, rid in zip(idx, ids)}
    st = srv.stats()
    srv.close()
    torch.cuda.synchronize()
    return {i: outs[i].cpu().numpy() for i in idx}, res, hist, st


@pytest.fixture(scope="module")
def alone(ctx, problems):
    out = {}
    for i in range(len(CASES)):
        vols, res, hist, _ = _run(ctx, problems, [i], max_active=1)
        assert res[i]["state"] == lib.REQ_DONE
        out[i] = (vols[i], hist[i])
    return out


@pytest.mark.parametrize("use_graph,sub_rounds", [(True, 0), (False, 0), (True, 1)])
def test_mix_equals_alone(ctx, problems, alone, use_graph, sub_rounds):
    idx = list(range(len(CASES)))
    vols, res, hist, st = _run(ctx, problems, idx, max_active=8, use_graph=use_graph, sub_rounds=sub_rounds)
    assert st["completed"] == len(idx) and st["failed"] == 0
    assert st["peak_active"] >= 2  # requests shared rounds (cheap ones may finish before all arrive)
    for i in idx:
        assert np.array_equal(vols[i], alone[i][0]), i
        assert hist[i] == alone[i][1], i


def test_budget_admits_part_of_the_stream(ctx, problems, alone):
    idx = list(range(len(CASES)))
    sizes = []
    srv = lib.Server(ctx, max_active=1)
    outs = {}
    for rid in _submit(srv, problems, idx, outs):
        sizes.append(srv.wait(rid, timeout=120.0)["device_bytes"])
    srv.close()
    budget = sorted(sizes)[-1] + sorted(sizes)[-2]  # the two largest fit together, not all
    assert budget < sum(sizes)
    vols, res, _, st = _run(ctx, problems, idx, max_active=8, budget_bytes=budget)
    assert st["completed"] == len(idx)
    assert 1 <= st["peak_active"] < len(idx)
    assert st["peak_bytes"] <= budget
    for i in idx:
        assert np.array_equal(vols[i], alone[i][0]), i
        assert res[i]["admit_s"] >= res[i]["submit_s"] and res[i]["done_s"] >= res[i]["admit_s"]
    assert st["latency_p99_s"] >= st["latency_p50_s"] > 0


def test_direct_schedule_matches_plain_loop(ctx, problems, alone):
    for i, (n, nphi, pair, lam, kopt, pos) in enumerate(CASES):
        g, sino = problems[i]
        G = lib.Geometry(ctx, to_desc(g))
        vol = torch.zeros(g.vol_shape, dtype=torch.float32, device="cuda")
        stats = G.gd_run(torch.from_numpy(sino).cuda(), vol, pair=pair, iters=ITERS, kopt=kopt, lam=lam,
                         positivity=pos)
        torch.cuda.synchronize()
        assert rel_l2(alone[i][0], vol.cpu().numpy()) <= 1e-6, i
        for a, b in zip(alone[i][1], stats):
            for k in ("residual_sq", "norm_dj", "norm_dj_proj", "alpha"):
                assert a[k] == pytest.approx(b[k], rel=1e-9, abs=1e-12), (i, k, a, b)


def test_recurrent_schedule_close_to_direct(ctx, problems, alone):
    idx = list(range(len(CASES)))
    vols, res, hist, st = _run(ctx, problems, idx, max_active=8, schedule=lib.SCHEDULE_RECURRENT)
    assert st["completed"] == len(idx)
    for i in idx:
        assert rel_l2(vols[i], alone[i][0]) <= 1e-4, i
        assert hist[i][-1]["residual_sq"] < hist[i][0]["residual_sq"]


def test_oracle_loop(ctx, problems, alone):
    """The served result against the CPU oracle's loop (same tolerance as the
    plain loop's test)."""
    i = 0
    n, nphi, pair, lam, kopt, pos = CASES[i]
    g, sino = problems[i]
    ref, _ = Oracle(g).gd(sino, pair=pair, iters=ITERS, kopt=kopt, lam=lam, positivity=pos)
    assert rel_l2(alone[i][0], ref) <= 1e-6


def test_host_pointers(ctx, problems, alone):
    i = 1
    n, nphi, pair, lam, kopt, pos = CASES[i]
    g, sino = problems[i]
    srv = lib.Server(ctx, max_active=2)
    vol = np.zeros(g.vol_shape, dtype=np.float32)
    rid = srv.submit(to_desc(g), np.ascontiguousarray(sino), vol, pair=pair, iters=ITERS,
                     kopt=kopt, lam=lam, positivity=pos)
    srv.wait(rid, timeout=120.0)
    srv.close()
    assert np.array_equal(vol, alone[i][0])


def test_oversized_request_fails_alone(ctx, problems, alone):
    srv = lib.Server(ctx, max_active=4, budget_bytes=64 << 20)
    g_big = scaled_geometry(160, 90)
    big_sino = torch.zeros(g_big.sino_shape, dtype=torch.float32, device="cuda")
    big_vol = torch.zeros(g_big.vol_shape, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rid_big = srv.submit(to_desc(g_big), big_sino, big_vol, iters=2)
    outs = {}
    (rid_ok,) = _submit(srv, problems, [2], outs)
    with pytest.raises(lib.TomoError) as e:
        srv.wait(rid_big, timeout=120.0)
    assert e.value.status == 4  # TOMO_ERR_OUT_OF_MEMORY
    assert srv.wait(rid_ok, timeout=120.0)["state"] == lib.REQ_DONE
    srv.close()
    torch.cuda.synchronize()
    assert np.array_equal(outs[2].cpu().numpy(), alone[2][0])


def test_sirt_served_in_a_mix_equals_plain_sirt(ctx, problems):
    """SIRT requests served beside gradient requests give the plain SIRT loop's
    volume bit for bit (no norm enters SIRT's update)."""
    srv = lib.Server(ctx, max_active=8)
    outs, ids, plain = {}, {}, {}
    gd_ids = _submit(srv, problems, [0, 1], outs)
    for i in (2, 4):
        g, sino = problems[i]
        n, nphi, pair, lam, kopt, pos = CASES[i]
        y = torch.from_numpy(sino).cuda()
        ref = torch.empty(g.vol_shape, dtype=torch.float32, device="cuda")
        lib.Geometry(ctx, to_desc(g)).sirt_run(y, ref, pair=pair, iters=ITERS, relaxation=0.8)
        plain[i] = ref
        outs[i] = torch.zeros(g.vol_shape, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        ids[i] = srv.submit(to_desc(g), y, outs[i], pair=pair, iters=ITERS, algorithm=lib.ALGO_SIRT,
                            relaxation=0.8)
    for rid in list(ids.values()) + gd_ids:
        assert srv.wait(rid, timeout=120.0)["state"] == lib.REQ_DONE
    hist = srv.iter_stats(ids[2], ITERS)
    srv.close()
    torch.cuda.synchronize()
    for i in (2, 4):
        assert np.array_equal(outs[i].cpu().numpy(), plain[i].cpu().numpy()), i
    assert hist[-1]["residual_sq"] < hist[0]["residual_sq"]


def test_bf16_copies_alone_equals_mix_and_close_to_fp32(ctx):
    """BF16 operator copies (storage=BF16): the projectors read BF16 z-pair
    copies rounded from the FP32 state each time they are made; VI, sums,
    state and result stay FP32.  Each result is the same alone and in the
    mix (bit-identical), and within rel-L2 2e-2 of the same request with
    FP32 copies (6 iterations).  The mix has both pairs, GD and SIRT, and a
    geometry whose row steps are near 4 (VI's detector boxes exceed the
    staging buffer there and it reads the taps directly)."""
    from test_gpu_parity import FINE_ROWS
    specs = [  # (geometry, pair, algorithm, lam, kopt)
        (scaled_geometry(32, 24), lib.PAIR_SIDDONVI, lib.ALGO_GD, 0.01, 10),
        (scaled_geometry(40, 30), lib.PAIR_RSVI, lib.ALGO_GD, 0.01, 3),
        (scaled_geometry(36, 20), lib.PAIR_SIDDONVI, lib.ALGO_SIRT, 0.0, 0),
        (FINE_ROWS, lib.PAIR_RSVI, lib.ALGO_GD, 0.01, 10),
    ]
    ys = []
    for i, (g, pair, algo, lam, kopt) in enumerate(specs):
        o = Oracle(g)
        truth = ellipsoids(g.nx, g.ny, g.nz, count=5, seed=300 + i, amin=2.0, amax=g.nx / 3)
        ys.append(torch.from_numpy(o.rs(truth) if pair == lib.PAIR_RSVI else o.siddon(truth)).cuda())

    def run(idx, storage, **kw):
        print(f"[case] serve bf16 {idx} storage={storage} {kw}", flush=True)
        srv = lib.Server(ctx, **kw)
        outs = {i: torch.zeros(specs[i][0].vol_shape, dtype=torch.float32, device="cuda") for i in idx}
        torch.cuda.synchronize()
        ids = {i: srv.submit(to_desc(specs[i][0]), ys[i], outs[i], pair=specs[i][1], iters=ITERS, kopt=specs[i][4],
                             lam=specs[i][3], algorithm=specs[i][2], storage=storage) for i in idx}
        for i in idx:
            assert srv.wait(ids[i], timeout=120.0)["state"] == lib.REQ_DONE, i
        srv.close()
        torch.cuda.synchronize()
        return {i: outs[i].cpu().numpy() for i in idx}

    idx = list(range(len(specs)))
    alone16 = {i: run([i], lib.BF16, max_active=1)[i] for i in idx}
    alone32 = {i: run([i], lib.F32, max_active=1)[i] for i in idx}
    mix16 = run(idx, lib.BF16, max_active=8)
    for i in idx:
        assert np.array_equal(mix16[i], alone16[i]), i
        err = rel_l2(alone16[i], alone32[i])
        print(f"[bf16] request {i}: rel_l2 vs FP32 copies {err:.3e}", flush=True)
        assert 0 < err <= 2e-2, (i, err)
