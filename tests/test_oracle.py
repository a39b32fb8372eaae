"""Pins the CPU oracle (CPU only).

Each check is chosen so that a plausible mistake (wrong sign or axis, off by
a half voxel, a dropped term, a transposed operand) fails it, and none of
them retypes the oracle's own formula:

* VI      - constant-sinogram and single-cell closed forms against an
            independent vector construction of the cone-beam projection
            (tests/geom_ref.py: line source->voxel cut by the detector plane);
            90-degree symmetry; explicit-matrix transpose.
* RS      - uniform-volume closed form (g = delta*L_r exactly for rays whose
            samples stay inside the voxel grid) with L_r from an independent
            vector construction; single-voxel support; explicit-matrix
            transpose; agreement with clean Siddon on a smooth phantom.
* Siddon  - uniform volume gives the exact chord length through cylinder and
            slab; an independent merge-all-crossings Siddon (Jacobs 1998);
            explicit-matrix transpose.
* compat  - the reference C path restated with its quirks; pinned by the
            mirror identity Siddon_ref(phi, s, t) = Siddon_clean(phi+pi, -s, t)
            on a phantom that vanishes near the FOV boundary (where the quirks
            act) and by the telescoping bound on a constant volume.
* FP32 recipe vs FP64 - rel-L2 <= 1e-6.
* Laplacian - constant and quadratic closed forms, symmetry.
* loop    - with the exact transpose as backprojector the step is the exact
            line search of J = |y-Px|^2 + lambda |Lx|^2: J drops by
            normdJ^2 / (2 normdJProj) each iteration; alpha_min rule.
* phantoms - Shepp-Logan centre value and shell volume; ellipsoid table
            invariants and voxel membership.
* FDK VI  - constant-sinogram closed form with the FDK weight evaluated on the
            vector construction (source distance, detector offsets).
* CG      - with the exact transpose, every step is an exact line search on
            the quadratic criterion; the first step is the gradient step; CG
            ends below the gradient loop.
* SIRT    - first iteration composed from the pinned operators (weights,
            relaxation, sign); exact homogeneity; convergence on consistent
            data.
* weighted gradient - W = 1 is the simple gradient; W = 2 doubles the first
            iterate; a binary view mask acts on the residual before the
            backprojection.
"""
import numpy as np
import pytest

import geom_ref as gr
from oracle_ffi import (Geometry, Oracle, ellipsoid_table, ellipsoids, laplacian, scaled_geometry,
                        shepp_logan)


def small_geom(n=16, nphi=8, nu=None, nv=None, **kw):
    g = scaled_geometry(n, nphi, nu, nv)
    for k, v in kw.items():
        setattr(g, k, v)
    return g


# ----------------------------------------------------------------- geometry
def test_geometry_derivation_matches_reference_types():
    g = Geometry(fod=98.0, fdd=230.0, nphi=256, su=14.2, sv=14.2, nu=256, nv=256, sx=6.05, sy=6.05, sz=5.86,
                 nx=256, ny=256, nz=256)
    info = Oracle(g).info
    f = np.float32
    assert info["du"] == f(f(14.2) / f(256))
    assert info["delta"] == f(f(6.05) / f(256))       # dz := dx (Volume.cu:21), zVolumeSize unused
    assert info["cu"] == 127.5 and info["cz"] == 127.5
    assert info["step"] == pytest.approx(2 * np.pi / 256, rel=1e-15)
    phi = Oracle(g).phi
    assert phi[1] - phi[0] == pytest.approx(2 * np.pi / 256, rel=1e-12)


def test_start_angle_is_added_to_index():
    # phi_p = (p + start) * step (Acquisition.cpp:95), not p*step + start*pi
    g = small_geom(8, 4, num_start=1.0, den_start=2.0, num_stop=5.0, den_stop=2.0)
    phi = Oracle(g).phi
    step = np.pi * (2.5 - 0.5) / 4
    assert np.allclose(phi, (np.arange(4) + 0.5) * step, rtol=0, atol=1e-15)


# ----------------------------------------------------------------------- VI
def _projections(g):
    X = gr.voxel_centres(g)
    return [gr.project_point(g, p, X) for p in range(g.nphi)]


def _support_weight(c, n):
    c0 = np.floor(c)
    f = c - c0
    return (1 - f) * ((c0 >= 0) & (c0 < n)) + f * ((c0 + 1 >= 0) & (c0 + 1 < n))


@pytest.mark.parametrize("f64", [True, False])
def test_vi_constant_sinogram_closed_form(f64):
    """g = 1: f = delta * sum_p (valid-tap weight in u) * (valid-tap weight in v)."""
    g = small_geom(20, 12, nu=26, nv=14)
    o = Oracle(g)
    f = o.vi(np.ones(g.sino_shape), f64=f64)
    exp = np.zeros(g.vol_shape)
    for u, v in _projections(g):
        exp += _support_weight(u, g.nu) * _support_weight(v, g.nv)
    exp *= gr.derived(g)["delta"]
    assert np.abs(f - exp).max() <= (1e-9 if f64 else 1e-5) * np.abs(exp).max()
    assert exp.max() > 0 and (exp == exp.max()).sum() > 10  # interior voxels see every view


@pytest.mark.parametrize("f64", [True, False])
def test_vi_fdk_constant_sinogram_closed_form(f64):
    """FDK weighting (FDK3D_kernel_16reg.cuh:71-88) on g = 1: f = sum_p W_p x
    (valid-tap weights), W_p = alphaC (FDD/du) / w^2 / sqrt(betaC + u'^2 +
    v'^2) with w the source distance along the central ray and (u', v') the
    detector offsets of the voxel's projection, from the vector construction."""
    g = small_geom(20, 12, nu=26, nv=14)
    o = Oracle(g)
    f = o.vi_fdk(np.ones(g.sino_shape), f64=f64)
    d = gr.derived(g)
    X = gr.voxel_centres(g)
    ku = g.fdd / d["du"]
    alpha_c = float(np.float32(np.pi * g.fod * g.fdd / g.nphi))
    exp = np.zeros(g.vol_shape)
    for p, phi in enumerate(gr.angles(g)):
        er, _, _ = gr.frame(phi)
        w = X @ er + g.fod
        u, v = gr.project_point(g, p, X)
        W = alpha_c * ku / w ** 2 / np.sqrt(ku * ku + (u - d["cu"]) ** 2 + (v - d["cv"]) ** 2)
        exp += W * _support_weight(u, g.nu) * _support_weight(v, g.nv)
    assert np.abs(f - exp).max() <= (1e-9 if f64 else 2e-5) * np.abs(exp).max()
    assert exp.max() > 0


def test_vi_single_cell_cone():
    """A single detector cell backprojects onto its bilinear cone only."""
    g = small_geom(16, 6, nu=20, nv=18)
    o = Oracle(g)
    p0, m0, n0 = 2, 9, 7
    s = np.zeros(g.sino_shape)
    s[p0, n0, m0] = 1.0
    f = o.vi(s, f64=True)
    u, v = gr.project_point(g, p0, gr.voxel_centres(g))
    exp = np.clip(1 - np.abs(u - m0), 0, None) * np.clip(1 - np.abs(v - n0), 0, None) * gr.derived(g)["delta"]
    assert np.abs(f - exp).max() <= 1e-9
    assert (exp > 0).sum() > 5


def test_vi_quarter_turn_symmetry():
    """Same projection in every view of a 4k-view orbit -> volume invariant
    under a 90-degree rotation about z (square volume and detector)."""
    g = small_geom(12, 8, nu=16, nv=16)
    rng = np.random.default_rng(0)
    one = rng.random((g.nv, g.nu))
    s = np.broadcast_to(one, g.sino_shape).copy()
    f = Oracle(g).vi(s, f64=True)
    # rotation by +90 deg maps voxel (i, j) to (N-1-j, i)
    rot = np.transpose(f, (0, 2, 1))[:, :, ::-1]
    assert np.abs(f - rot).max() <= 1e-9 * np.abs(f).max()


# ----------------------------------------------------------------------- RS
def _rs_expected_lr(g, cube, p, m, n):
    """Independent L_r: vector construction of the ray in voxel units, circle
    intersection, entry-z test and exit-z clip as defined by the reference."""
    d = gr.derived(g)
    S, D = gr.source_detector(g, p, m, n)
    dl = d["delta"]
    S = S / dl
    u = (D / dl - S) / np.linalg.norm(D / dl - S)
    R2 = (g.nx / 2) ** 2 * (2 if cube else 1)
    a = u[0] ** 2 + u[1] ** 2
    b = S[0] * u[0] + S[1] * u[1]
    c = S[0] ** 2 + S[1] ** 2 - R2
    disc = b * b - a * c
    if disc < 0:
        return 0, None, None
    l1, l2 = (-b - np.sqrt(disc)) / a, (-b + np.sqrt(disc)) / a
    cen = np.array([d["cx"], d["cy"], d["cz"]])
    if cube:
        for ax, nn in ((0, g.nx), (1, g.ny)):
            if u[ax] != 0:
                ta, tb = (-cen[ax] - S[ax]) / u[ax], (nn - 1 - cen[ax] - S[ax]) / u[ax]
                l1, l2 = max(l1, min(ta, tb)), min(l2, max(ta, tb))
        if l1 > l2:
            return 0, None, None
    ze = S[2] + l1 * u[2] + cen[2]
    if not (0 <= ze <= g.nz - 1 if cube else 0 <= ze < g.nz):
        return 0, None, None
    zx = S[2] + l2 * u[2] + cen[2]
    if zx > g.nz - 1:
        l2 = (g.nz - 1 - cen[2] - S[2]) / u[2]
    elif zx < 0:
        l2 = (-cen[2] - S[2]) / u[2]
    lr = int(np.floor(abs(l2 - l1)))
    return lr, S + l1 * u + cen, u


@pytest.mark.parametrize("cube", [False, True])
def test_rs_uniform_volume_closed_form(cube):
    g = small_geom(18, 6, nu=24, nv=22)
    o = Oracle(g)
    out = o.rs(np.ones(g.vol_shape), cube=cube, f64=True)
    delta = gr.derived(g)["delta"]
    checked = 0
    for p in range(g.nphi):
        for n in range(g.nv):
            for m in range(g.nu):
                lr, e, u = _rs_expected_lr(g, cube, p, m, n)
                val = out[p, n, m]
                if lr == 0:
                    assert val == 0.0
                    continue
                last = e + (lr - 1) * u
                lo, hi = np.minimum(e, last), np.maximum(e, last)
                if (lo >= 0).all() and hi[0] <= g.nx - 1 and hi[1] <= g.ny - 1 and hi[2] <= g.nz - 1:
                    assert val == pytest.approx(delta * lr, rel=1e-12)
                    checked += 1
                else:
                    assert val <= delta * lr + 1e-9
    assert checked > 100


def test_rs_single_voxel_support():
    g = small_geom(14, 5, nu=18, nv=16)
    o = Oracle(g)
    vol = np.zeros(g.vol_shape)
    i0, j0, k0 = 8, 5, 7
    vol[k0, j0, i0] = 1.0
    s = o.rs(vol, f64=True)
    d = gr.derived(g)
    X0 = np.array([(i0 - d["cx"]) * d["delta"], (j0 - d["cy"]) * d["delta"], (k0 - d["cz"]) * d["delta"]])
    for p in range(g.nphi):
        for n in range(g.nv):
            for m in range(g.nu):
                S, D = gr.source_detector(g, p, m, n)
                u = (D - S) / np.linalg.norm(D - S)
                dist = np.linalg.norm(np.cross(X0 - S, u)) / d["delta"]
                if dist > np.sqrt(3.0):
                    assert s[p, n, m] == 0.0
    assert (s > 0).sum() > 4


# ------------------------------------------------------------------- Siddon
@pytest.mark.parametrize("geom", ["cube", "ragged"])
def test_siddon_uniform_volume_is_chord_length(geom):
    if geom == "cube":
        g = small_geom(20, 7, nu=30, nv=26)
    else:
        g = Geometry(fod=140.0, fdd=330.0, nphi=5, su=40 * 1.3, sv=21 * 1.9, nu=40, nv=21, sx=19 * 1.1,
                     sy=19 * 1.1, sz=13 * 1.1, nx=19, ny=19, nz=13, num_start=1, den_start=5)
    o = Oracle(g)
    s64 = o.siddon(np.ones(g.vol_shape), f64=True)
    s32 = o.siddon(np.ones(g.vol_shape, dtype=np.float32))
    d = gr.derived(g)
    R, zh = 0.5 * g.nx * d["delta"], 0.5 * g.nz * d["delta"]
    exp = np.zeros(g.sino_shape)
    for p in range(g.nphi):
        for n in range(g.nv):
            for m in range(g.nu):
                S, D = gr.source_detector(g, p, m, n)
                exp[p, n, m] = gr.cylinder_slab_chord(S, D, R, zh)
    assert np.abs(s64 - exp).max() <= 1e-9 * exp.max()
    assert np.abs(s32 - exp).max() <= 1e-5 * exp.max()
    assert (exp > 0).mean() > 0.3


def test_siddon_matches_merge_all_crossings():
    g = small_geom(10, 3, nu=9, nv=8)
    rng = np.random.default_rng(3)
    vol = rng.random(g.vol_shape)
    o = Oracle(g)
    out = o.siddon(vol, f64=True)
    for p in range(g.nphi):
        for n in range(g.nv):
            for m in range(g.nu):
                S, D = gr.source_detector(g, p, m, n)
                ref = gr.siddon_merge(g, vol, S, D)
                assert out[p, n, m] == pytest.approx(ref, rel=1e-10, abs=1e-10)


@pytest.mark.parametrize("n,nphi,nu,nv,start", [(10, 5, 13, 11, (1.0, 7.0)), (13, 4, 16, 15, (2.0, 9.0)),
                                                 (16, 3, 19, 17, (1.0, 11.0))])
def test_siddon_segment_count_from_merged_crossings(n, nphi, nu, nv, start):
    """N_seg (og_siddon_count, the numerator of the Siddon roofline: chords
    of positive FP32 length) against an independent merge of every plane
    crossing (geom_ref.siddon_merge_segments), ray by ray.  Segments longer
    than tau = 1e-4 voxel must each be one chord; crossings of different
    planes that lie within tau of each other (corner passages, where FP32
    rounding may or may not leave a sliver) may add at most one chord each.
    So N_robust <= N_seg <= N_robust + N_ambiguous, with N_ambiguous a small
    fraction: a dropped last chord, a counted zero chord or a missed axis
    would move N_seg by about one per ray."""
    g = small_geom(n, nphi, nu=nu, nv=nv, num_start=start[0], den_start=start[1])
    o = Oracle(g)
    dl = gr.derived(g)["delta"]
    tau = 1e-4 * dl
    robust = ambiguous = rays = 0
    for p in range(g.nphi):
        for v in range(g.nv):
            for m in range(g.nu):
                S, D = gr.source_detector(g, p, m, v)
                segs, cuts = gr.siddon_merge_segments(g, S, D, raw=True)
                robust += sum(1 for length, idx in segs if idx is not None and length > tau)
                ambiguous += int((np.diff(cuts) <= tau).sum())
                rays += bool(segs)
    n_seg = o.siddon_count()
    print(f"N_seg {n_seg}, robust {robust}, ambiguous {ambiguous}, rays {rays}")
    assert rays > 100 and ambiguous < rays / 4
    assert robust <= n_seg <= robust + ambiguous


def test_siddon_compat_mirror_identity():
    """Reference C Siddon at (phi, s, t) == clean Siddon at (phi+pi, -s, t) on a
    ball that vanishes near the FOV boundary, where the compat quirks act."""
    g = scaled_geometry(256, 8, nu=32, nv=32, pixel_mm=16.0)
    k, j, i = np.ogrid[:256, :256, :256]
    ball = (((i - 127.5) ** 2 + (j - 131.0) ** 2 + (k - 120.0) ** 2) <= 60.0 ** 2).astype(np.float32)
    o = Oracle(g)
    comp = o.siddon_compat(ball)
    clean = o.siddon(ball)
    mirrored = clean[(np.arange(8) + 4) % 8][:, :, ::-1]
    assert np.abs(comp - mirrored).max() <= 1e-4 * np.abs(mirrored).max()
    assert (mirrored > 0).sum() > 100


def test_siddon_compat_constant_volume_telescopes():
    """Chords telescope: a constant volume gives c*(lambda_last - lambda_min),
    between the clean chord length of the cylinder without z clip and that plus
    one (unclipped) last step, < sqrt(3) delta L / |D_a-S_a| <= a few delta."""
    g = scaled_geometry(256, 4, nu=16, nv=16, pixel_mm=32.0)
    o = Oracle(g)
    comp = o.siddon_compat(np.ones(g.vol_shape, dtype=np.float32))
    d = gr.derived(g)
    R = 0.5 * g.nx * d["delta"]
    for p in range(g.nphi):
        for n in range(g.nv):
            for m in range(g.nu):
                S, D = gr.source_detector(g, p, m, n, mirrored=True)
                chord = gr.cylinder_slab_chord(S, D, R, 1e30)
                assert chord - 1e-2 <= comp[p, n, m] <= chord + 2.0 * d["delta"] + 1e-2


def test_siddon_compat_rejects_other_sizes():
    with pytest.raises(ValueError):
        Oracle(small_geom(16, 2)).siddon_compat(np.zeros((16, 16, 16), dtype=np.float32))


# --------------------------------------------------------------- transposes
def _matrix(fn, n_in, in_shape):
    cols = []
    for c in range(n_in):
        e = np.zeros(n_in)
        e[c] = 1.0
        cols.append(np.asarray(fn(e.reshape(in_shape)), dtype=np.float64).ravel())
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("op", ["vi", "rs", "rs_cube", "siddon"])
def test_transpose_is_exact_matrix_transpose(op):
    g = small_geom(6, 3, nu=7, nv=5)
    o = Oracle(g)
    nvox, nsino = int(np.prod(g.vol_shape)), int(np.prod(g.sino_shape))
    if op == "vi":
        A = _matrix(lambda s: o.vi(s, f64=True), nsino, g.sino_shape)
        AT = _matrix(o.vi_T, nvox, g.vol_shape)
    elif op == "siddon":
        A = _matrix(lambda v: o.siddon(v, f64=True), nvox, g.vol_shape)
        AT = _matrix(o.siddon_T, nsino, g.sino_shape)
    else:
        cube = op == "rs_cube"
        A = _matrix(lambda v: o.rs(v, cube=cube, f64=True), nvox, g.vol_shape)
        AT = _matrix(lambda s: o.rs_T(s, cube=cube), nsino, g.sino_shape)
    assert np.abs(A).max() > 0
    assert np.abs(A.T - AT).max() <= 1e-12 * np.abs(A).max()


def test_rs_and_siddon_agree_on_smooth_phantom():
    """Two independent discretisations of the same line integrals."""
    g = scaled_geometry(64, 16)
    vol = ellipsoids(64, count=8, seed=2, amin=10.0, amax=25.0)
    o = Oracle(g)
    rs, sid = o.rs(vol), o.siddon(vol)
    assert np.linalg.norm(rs - sid) / np.linalg.norm(sid) < 0.05


@pytest.mark.parametrize("op", ["vi", "rs", "rs_cube", "siddon"])
def test_fp32_recipe_close_to_fp64(op):
    g = scaled_geometry(40, 12)
    o = Oracle(g)
    rng = np.random.default_rng(7)
    if op == "vi":
        x = rng.random(g.sino_shape).astype(np.float32)
        a, b = o.vi(x), o.vi(x.astype(np.float64), f64=True)
    else:
        x = rng.random(g.vol_shape).astype(np.float32)
        if op == "siddon":
            a, b = o.siddon(x), o.siddon(x.astype(np.float64), f64=True)
        else:
            a, b = o.rs(x, cube=op == "rs_cube"), o.rs(x.astype(np.float64), cube=op == "rs_cube", f64=True)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-6


# ---------------------------------------------------------------- Laplacian
def test_laplacian_closed_forms():
    n = (7, 6, 5)
    one = np.ones(n[::-1], dtype=np.float64)
    L1 = laplacian(one)
    k, j, i = np.meshgrid(range(5), range(6), range(7), indexing="ij")
    missing = ((i == 0) | (i == 6)).astype(int) + ((j == 0) | (j == 5)) + ((k == 0) | (k == 4))
    assert np.array_equal(L1, missing.astype(np.float64))   # zero padding
    q = (i.astype(np.float64) ** 2)
    Lq = laplacian(q)
    assert np.allclose(Lq[1:-1, 1:-1, 1:-1], -2.0)            # -(second difference of i^2)
    rng = np.random.default_rng(0)
    x, y = rng.random(n[::-1]), rng.random(n[::-1])
    assert np.vdot(laplacian(x), y) == pytest.approx(np.vdot(x, laplacian(y)), rel=1e-13)
    assert np.array_equal(laplacian(x.astype(np.float32)), laplacian(x.astype(np.float32)))


# --------------------------------------------------------------------- loop
def _J(o, g, x, y, lam):
    r = y.astype(np.float64) - o.rs(x.astype(np.float64), f64=True)
    return float(np.sum(r * r) + lam * np.sum(laplacian(x.astype(np.float64)) ** 2))


def test_loop_is_exact_line_search_with_matched_pair():
    g = scaled_geometry(20, 10)
    o = Oracle(g)
    truth = ellipsoids(20, count=5, seed=1, amin=3.0, amax=7.0)
    y = o.rs(truth)
    lam = 0.05
    x = np.zeros(g.vol_shape, dtype=np.float32)
    Js = [_J(o, g, x, y, lam)]
    stats_all = []
    for _ in range(4):
        x, st = o.gd(y, vol0=x, pair=0, iters=1, kopt=10, lam=lam, matched=True)
        Js.append(_J(o, g, x, y, lam))
        stats_all.append(st[0])
    for k, st in enumerate(stats_all):
        predicted = st["norm_dj"] ** 2 / (2.0 * st["norm_dj_proj"])
        assert Js[k] - Js[k + 1] == pytest.approx(predicted, rel=2e-3)
    assert all(a > b for a, b in zip(Js, Js[1:]))


def test_weighted_gradient_pins():
    """og_wgd_run (doWeightedGradient, C/Iter3D_Algo_Gradient.cu:227-371):
    W = 1 reproduces the simple gradient bit for bit (a product by 1 is
    exact); W = 2 from x0 = 0 doubles the first iterate exactly (every
    operator, norm and the step scale by powers of two, alpha is unchanged);
    a binary view mask W acts on the residual before the backprojection,
    so y and W*y give the same first iterate, unlike the unweighted loop."""
    g = scaled_geometry(16, 8)
    o = Oracle(g)
    y = o.rs(ellipsoids(16, count=4, seed=2, amin=2.0, amax=6.0))
    kw = dict(pair=0, iters=3, kopt=10, lam=0.02, alpha_min_init=-999999999999.0)
    x_ref, st_ref = o.gd(y, **kw)
    x_w1, st_w1 = o.gd(y, weight=np.ones_like(y), **kw)
    assert np.array_equal(x_w1, x_ref)
    for a, b in zip(st_w1, st_ref):  # FP64 norms: OpenMP combines partial sums in any order
        assert all(a[k] == pytest.approx(b[k], rel=1e-12) for k in a)
    one = dict(kw, iters=1)
    x1, st1 = o.gd(y, **one)
    x2, st2 = o.gd(y, weight=np.full_like(y, 2.0), **one)
    assert np.array_equal(x2, 2.0 * x1) and st2[0]["alpha"] == pytest.approx(st1[0]["alpha"], rel=1e-12)
    mask = np.zeros_like(y)
    mask[::2] = 1.0  # even views only
    xa, _ = o.gd(y, weight=mask, **one)
    xb, _ = o.gd(y * mask, weight=mask, **one)
    assert np.array_equal(xa, xb)
    assert not np.array_equal(xa, x1)


def test_cg_exact_line_search_and_beats_gradient_with_matched_pair():
    """Conjugate gradient (Fletcher-Reeves) on J = |y-Px|^2 + lambda |Lx|^2 with
    the exact transpose: each step lowers J by |g|^4 / (2 d'Hd) (exact line
    search with g'd = |g|^2), the first step is the gradient step, and after
    a few iterations J is below the gradient loop's."""
    g = scaled_geometry(20, 10)
    o = Oracle(g)
    y = o.rs(ellipsoids(20, count=5, seed=1, amin=3.0, amax=7.0))
    lam = 0.05
    Js = [_J(o, g, np.zeros(g.vol_shape, dtype=np.float32), y, lam)]
    for k in range(1, 6):
        x, st = o.cg(y, pair=0, iters=k, lam=lam, matched=True)
        Js.append(_J(o, g, x, y, lam))
    for k in range(5):
        predicted = st[k]["norm_dj"] ** 2 / (2.0 * st[k]["norm_dj_proj"])
        assert Js[k] - Js[k + 1] == pytest.approx(predicted, rel=2e-3)
    x1_cg, s1 = o.cg(y, pair=0, iters=1, lam=lam, matched=True)
    x1_gd, g1 = o.gd(y, pair=0, iters=1, kopt=10, lam=lam, matched=True)
    assert s1[0]["alpha"] == pytest.approx(g1[0]["alpha"], rel=1e-12)
    assert np.abs(x1_cg - x1_gd).max() <= 1e-6 * np.abs(x1_gd).max()
    xg, _ = o.gd(y, pair=0, iters=5, kopt=10, lam=lam, matched=True)
    assert Js[5] < _J(o, g, xg, y, lam)


def test_loop_alpha_min_rule():
    g = scaled_geometry(16, 8)
    o = Oracle(g)
    y = o.rs(ellipsoids(16, count=4, seed=5, amin=2.0, amax=6.0))
    _, st = o.gd(y, pair=0, iters=4, kopt=2, lam=0.0)
    a1 = st[0]["alpha"]
    assert a1 < 0
    for s in st[1:]:
        assert s["alpha"] == max(a1, -0.0347711)
        assert s["norm_dj"] == 0.0


def test_loop_positivity_clamps():
    g = scaled_geometry(16, 8)
    o = Oracle(g)
    y = -o.rs(ellipsoids(16, count=4, seed=5, amin=2.0, amax=6.0))
    x, _ = o.gd(y, pair=0, iters=2, kopt=10, lam=0.0, positivity=1)
    assert x.min() >= 0.0


# -------------------------------------------------------------------- SIRT
# SIRT is defined by this build: x += C B (R (y - P x)), R = 1/(P 1),
# C = relax/(B 1), with the pair's (independently pinned) operators.
@pytest.mark.parametrize("pair", [0, 1])
def test_sirt_first_iteration_from_pinned_operators(pair):
    g = scaled_geometry(20, 12)
    o = Oracle(g)
    y = o.rs(ellipsoids(20, count=4, seed=6, amin=2.0, amax=7.0))
    proj = (lambda v: o.rs(v)) if pair == 0 else (lambda v: o.siddon(v))
    p1 = proj(np.ones(g.vol_shape, dtype=np.float32))
    b1 = o.vi(np.ones(g.sino_shape, dtype=np.float32))
    rw = np.where(p1 > 0, np.float32(1.0) / np.where(p1 > 0, p1, 1), 0).astype(np.float32)
    cw = np.where(b1 > 0, np.float32(0.75) / np.where(b1 > 0, b1, 1), 0).astype(np.float32)
    ow, oc = o.sirt_weights(pair=pair, relax=0.75)
    assert np.array_equal(ow, rw) and np.array_equal(oc, cw)
    x1, res = o.sirt(y, pair=pair, iters=1, relax=0.75)
    expect = (np.float32(0.0) + cw * o.vi((y - np.float32(0.0)) * rw)).astype(np.float32)
    assert np.array_equal(x1, expect)
    assert res[0] == pytest.approx(float((y.astype(np.float64) ** 2).sum()), rel=1e-12)


def test_sirt_is_homogeneous_and_converges():
    g = scaled_geometry(20, 16)
    o = Oracle(g)
    truth = ellipsoids(20, count=5, seed=7, amin=2.0, amax=7.0)
    y = o.rs(truth)
    x, res = o.sirt(y, pair=0, iters=8)
    x2, _ = o.sirt(2.0 * y, pair=0, iters=8)
    assert np.array_equal(x2, 2.0 * x)          # scaling y by 2 is exact throughout
    assert all(b < a for a, b in zip(res, res[1:]))
    assert res[-1] < 0.2 * res[0]
    x1, _ = o.sirt(y, pair=0, iters=1)
    assert np.linalg.norm(x - truth) < np.linalg.norm(x1 - truth)


# ----------------------------------------------------------------- phantoms
def test_shepp_logan_values():
    n = 65
    v = shepp_logan(n)
    assert v[32, 32, 32] == np.float32(2.0 - 0.98)
    assert v.min() >= -1e-6 and v.max() <= 2.0 + 1e-6
    outer = (v != 0).sum()
    h = (n - 1) / 2
    expected = 4.0 / 3.0 * np.pi * 0.69 * 0.92 * 0.81 * h ** 3
    assert abs(outer - expected) / expected < 0.03


def test_ellipsoid_table_invariants_and_membership():
    n, cnt = 24, 6
    p7, R = ellipsoid_table(n, n, n, cnt, 11, 2.0, 8.0)
    for r in R:
        assert np.allclose(r @ r.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(r) == pytest.approx(1.0, abs=1e-12)
    assert np.all(np.abs(p7[:, :3]) <= 0.4 * n)
    semi = 1.0 / np.sqrt(p7[:, 3:6])
    assert np.all((semi >= 2.0 - 1e-12) & (semi <= 8.0 + 1e-12))
    assert np.all((p7[:, 6] >= 0) & (p7[:, 6] < 1))
    vol = ellipsoids(n, count=cnt, seed=11, amin=2.0, amax=8.0)
    k, j, i = np.meshgrid(*(np.arange(n) - (n - 1) / 2,) * 3, indexing="ij")
    off = np.stack([i, j, k], -1)
    exp = np.zeros((n, n, n))
    for e in range(cnt):
        l = (off - p7[e, :3]) @ R[e].T
        exp += np.where((l ** 2 * p7[e, 3:6]).sum(-1) <= 1.0, p7[e, 6], 0.0)
    assert np.abs(vol - exp).max() <= 1e-6
