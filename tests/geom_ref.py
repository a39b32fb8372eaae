"""Independent geometric constructions used to pin the oracle.

These derive source, detector and voxel positions as vectors in a rotating
frame (e_r = (cos phi, sin phi, 0), e_u = (-sin phi, cos phi, 0), e_z) and
intersect lines with planes, cylinders and ellipsoids directly, instead of
using the closed per-view coefficients of the operators.
"""
import numpy as np


def frame(phi):
    er = np.array([np.cos(phi), np.sin(phi), 0.0])
    eu = np.array([-np.sin(phi), np.cos(phi), 0.0])
    ez = np.array([0.0, 0.0, 1.0])
    return er, eu, ez


def derived(g):
    """Reference-typed derived geometry (float32 sizes, centres N/2-0.5)."""
    f = np.float32
    du = f(f(g.su) / f(g.nu))
    dv = f(f(g.sv) / f(g.nv))
    delta = f(f(g.sx) / f(g.nx))
    return dict(du=float(du), dv=float(dv), cu=g.nu / 2 - 0.5, cv=g.nv / 2 - 0.5, delta=float(delta),
                cx=g.nx / 2 - 0.5, cy=g.ny / 2 - 0.5, cz=g.nz / 2 - 0.5)


def angles(g):
    start = g.num_start / g.den_start
    stop = g.num_stop / g.den_stop
    step = np.pi * (stop - start) / g.nphi
    return (np.arange(g.nphi) + start) * step


def source_detector(g, p, m, n, mirrored=False):
    """Source and detector-cell centre (mm, volume-centred) of ray (p, m, n).
    RS/VI convention: source at -FOD e_r, detector plane at (FDD-FOD) e_r.
    mirrored=True: the reference Siddon C path (+FOD e_r side)."""
    d = derived(g)
    phi = angles(g)[p]
    er, eu, ez = frame(phi)
    s = (m - d["cu"]) * d["du"]
    t = (n - d["cv"]) * d["dv"]
    sign = -1.0 if mirrored else 1.0
    S = -sign * g.fod * er
    D = sign * (g.fdd - g.fod) * er + s * eu + t * ez
    return S, D


def project_point(g, p, X):
    """Detector index coordinates (u, v) of the point X (mm, centred) in view p:
    the line source->X cut by the detector plane."""
    d = derived(g)
    phi = angles(g)[p]
    er, eu, ez = frame(phi)
    S = -g.fod * er
    D0 = (g.fdd - g.fod) * er
    X = np.asarray(X, dtype=np.float64)
    lam = ((D0 - S) @ er) / ((X - S) @ er)
    H = S + lam[..., None] * (X - S)
    u = ((H - D0) @ eu) / d["du"] + d["cu"]
    v = ((H - D0) @ ez) / d["dv"] + d["cv"]
    return u, v


def voxel_centres(g):
    d = derived(g)
    k, j, i = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
    return np.stack([(i - d["cx"]) * d["delta"], (j - d["cy"]) * d["delta"], (k - d["cz"]) * d["delta"]], axis=-1)


def cylinder_slab_chord(S, D, R, zh):
    """Length (mm) of the segment of line S->D inside x^2+y^2<=R^2, |z|<=zh."""
    d = (D - S) / np.linalg.norm(D - S)
    a = d[0] ** 2 + d[1] ** 2
    b = S[0] * d[0] + S[1] * d[1]
    c = S[0] ** 2 + S[1] ** 2 - R * R
    disc = b * b - a * c
    if a == 0 or disc <= 0:
        return 0.0
    l0, l1 = (-b - np.sqrt(disc)) / a, (-b + np.sqrt(disc)) / a
    if d[2] != 0:
        ta, tb = (-zh - S[2]) / d[2], (zh - S[2]) / d[2]
        l0, l1 = max(l0, min(ta, tb)), min(l1, max(ta, tb))
    return max(0.0, l1 - l0)


def siddon_merge_segments(g, S, D, raw=False):
    """The segments of siddon_merge: (length mm, voxel index (i, j, k) or None
    outside the grid) for every pair of consecutive distinct crossings.  With
    raw=True also the sorted crossing list before duplicates are merged."""
    dd = derived(g)
    dl = dd["delta"]
    L = np.linalg.norm(D - S)
    d = (D - S) / L
    R = 0.5 * g.nx * dl
    zh = 0.5 * g.nz * dl
    a = d[0] ** 2 + d[1] ** 2
    b = S[0] * d[0] + S[1] * d[1]
    c = S[0] ** 2 + S[1] ** 2 - R * R
    disc = b * b - a * c
    if a == 0 or disc <= 0:
        return ([], np.zeros(0)) if raw else []
    l0, l1 = (-b - np.sqrt(disc)) / a, (-b + np.sqrt(disc)) / a
    if d[2] != 0:
        ta, tb = (-zh - S[2]) / d[2], (zh - S[2]) / d[2]
        l0, l1 = max(l0, min(ta, tb)), min(l1, max(ta, tb))
    if l1 <= l0:
        return ([], np.zeros(0)) if raw else []
    cen = np.array([dd["cx"], dd["cy"], dd["cz"]])
    nn = [g.nx, g.ny, g.nz]
    cuts = [l0, l1]
    for ax in range(3):
        if d[ax] == 0:
            continue
        planes = (np.arange(nn[ax] + 1) - 0.5 - cen[ax]) * dl
        lam = (planes - S[ax]) / d[ax]
        cuts.extend(lam[(lam > l0) & (lam < l1)].tolist())
    all_cuts = np.sort(np.array(cuts))
    cuts = np.unique(all_cuts)
    out = []
    for lo, hi in zip(cuts[:-1], cuts[1:]):
        mid = S + 0.5 * (lo + hi) * d
        idx = np.floor(mid / dl + cen + 0.5).astype(int)
        inside = all(0 <= idx[a_] < nn[a_] for a_ in range(3))
        out.append((hi - lo, tuple(int(v) for v in idx) if inside else None))
    return (out, all_cuts) if raw else out


def siddon_merge(g, vol, S, D):
    """Siddon by merging every plane crossing (Jacobs et al. 1998 form):
    sort all crossings of x, y, z voxel planes inside the FOV, give each
    segment to the voxel containing its midpoint."""
    dd = derived(g)
    dl = dd["delta"]
    L = np.linalg.norm(D - S)
    d = (D - S) / L
    R = 0.5 * g.nx * dl
    zh = 0.5 * g.nz * dl
    a = d[0] ** 2 + d[1] ** 2
    b = S[0] * d[0] + S[1] * d[1]
    c = S[0] ** 2 + S[1] ** 2 - R * R
    disc = b * b - a * c
    if a == 0 or disc <= 0:
        return 0.0
    l0, l1 = (-b - np.sqrt(disc)) / a, (-b + np.sqrt(disc)) / a
    if d[2] != 0:
        ta, tb = (-zh - S[2]) / d[2], (zh - S[2]) / d[2]
        l0, l1 = max(l0, min(ta, tb)), min(l1, max(ta, tb))
    if l1 <= l0:
        return 0.0
    cen = np.array([dd["cx"], dd["cy"], dd["cz"]])
    nn = [g.nx, g.ny, g.nz]
    cuts = [l0, l1]
    for ax in range(3):
        if d[ax] == 0:
            continue
        planes = (np.arange(nn[ax] + 1) - 0.5 - cen[ax]) * dl
        lam = (planes - S[ax]) / d[ax]
        cuts.extend(lam[(lam > l0) & (lam < l1)].tolist())
    cuts = np.unique(np.array(cuts))
    total = 0.0
    for lo, hi in zip(cuts[:-1], cuts[1:]):
        mid = S + 0.5 * (lo + hi) * d
        idx = np.floor(mid / dl + cen + 0.5).astype(int)
        if all(0 <= idx[a_] < nn[a_] for a_ in range(3)):
            total += (hi - lo) * float(vol[idx[2], idx[1], idx[0]])
    return total
