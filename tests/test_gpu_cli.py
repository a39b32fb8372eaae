"""GPU: the tomogpi command-line driver (TomoGPI_exe surface) on a small
Projections_N work directory, run as a child process; outputs compared
with the oracle.  Config files follow the reference's format
(tests/golden/Projections_256 are the shipped ones)."""
import os
import subprocess

import numpy as np
import pytest

from oracle_ffi import Geometry, Oracle, ellipsoids, shepp_logan

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "tomogpi_amd", "bin", "tomogpi")
N, NPHI = 32, 16


def write_workdir(tmp, compute_line="compute COMPUTE_CUDA", iters=3, kopt=10, lam=0.01, init="none",
                  real="none", gradient="Simple"):
    wd = tmp / f"Projections_{N}"
    wd.mkdir()
    (wd / "configCT").write_text(
        f"[Acquisition]\nFOD 62.5 mm\nFDD 125 mm\nprojectionNb {NPHI}\nnumStartAngle 0\ndenStartAngle 1\n"
        f"numStopAngle 2\ndenStopAngle 1\nacquisitionScheme Uniform (Uniform, NonUniform)\n\n[Detector]\n"
        f"horizontalDetectorSize 64 mm\nverticalDetectorSize 64 mm\nhorizontalDetectorPixelNb {N}\n"
        f"verticalDetectorPixelNb {N}\n\n[Volume]\nxVolumeSize 32 mm\nyVolumeSize 32 mm\nzVolumeSize 32 mm\n"
        f"xVolumePixelNb {N}\nyVolumePixelNb {N}\nzVolumePixelNb {N}\n\n[Initialization]\n"
        f"buildRealSinogramFromTiff 0\ninitVolumeName {init}\nrealVolumeName {real}\n")
    (wd / "configIteration").write_text(
        f"[Iteration]\nglobalIterationNb 1\ngradientIterationNb {iters}\noptimalStepIterationNb {kopt}\n"
        f"[PenaltyFunction]\nregularizationParameter {lam}\npenaltyFunction GGaussian (Huber, GGaussian)\n"
        f"gradient {gradient} (Simple,Conjugate)\nhuberThreshold 0.0001\npositivity 0\nbeta 2.0\n")
    (wd / "configComputeArchitecture").write_text(
        "[precision]\nprecision float (half,float,double)\nstorage mem_CPU (mem_CPU,mem_GPU)\n"
        f"{compute_line}\narchitecture ARCHITECTURE_GPU\npair PAIR_RSVI (PAIR_SFTR,PAIR_RSVI,PAIR_SIDDONVI)\n")
    (wd / "configTiff").write_text("[Tiff]\ntiffName P_ER_GPU\n")
    return wd


def geom():
    return Geometry(fod=62.5, fdd=125.0, nphi=NPHI, su=64.0, sv=64.0, nu=N, nv=N, sx=32.0, sy=32.0, sz=32.0,
                    nx=N, ny=N, nz=N)


def run(tmp, *args):
    print(f"[case] tomogpi {N} {' '.join(args)}", flush=True)
    r = subprocess.run([CLI, str(N), *args], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Time computation :" in r.stdout
    return r.stdout


@pytest.mark.parametrize("storage,cube", [("mem_GPU", False), ("mem_CPU", True)])
def test_cli_proj_back(tmp_path, storage, cube):
    vol = shepp_logan(N)
    vpath = tmp_path / "phantom.v"
    vol.tofile(vpath)
    write_workdir(tmp_path, real=str(vpath))
    run(tmp_path, "--op=proj", f"--storage={storage}")
    o = Oracle(geom())
    ref = o.rs(vol, cube=cube)
    got = np.fromfile(tmp_path / "proj.s", dtype=np.float32).reshape(ref.shape)
    assert np.array_equal(got, ref)
    ref.tofile(tmp_path / f"Projections_{N}" / "P_ER_GPU.s")
    run(tmp_path, "--op=back", f"--storage={storage}")
    back = np.fromfile(tmp_path / "back.v", dtype=np.float32).reshape(vol.shape)
    assert np.array_equal(back, o.vi(ref))


def test_cli_double_precision_files(tmp_path):
    """--precision=double reads and writes FP64 .v/.s files (the reference's
    double containers write sizeof(double) * count) and runs the FP32
    operators: the FP64 outputs hold the FP32 results exactly."""
    vol = shepp_logan(N).astype(np.float64) + 1e-9  # the extra digits round away in FP32
    vpath = tmp_path / "phantom64.v"
    vol.tofile(vpath)
    write_workdir(tmp_path, real=str(vpath))
    run(tmp_path, "--op=proj", "--storage=mem_GPU", "--precision=double")
    o = Oracle(geom())
    ref = o.rs(vol.astype(np.float32))
    assert os.path.getsize(tmp_path / "proj.s") == ref.size * 8
    got = np.fromfile(tmp_path / "proj.s", dtype=np.float64).reshape(ref.shape)
    assert np.array_equal(got, ref.astype(np.float64))
    ref.astype(np.float64).tofile(tmp_path / f"Projections_{N}" / "P_ER_GPU.s")
    run(tmp_path, "--op=back", "--storage=mem_GPU", "--precision=double")
    back = np.fromfile(tmp_path / "back.v", dtype=np.float64).reshape(vol.shape)
    assert np.array_equal(back, o.vi(ref).astype(np.float64))


def test_cli_back_fdk(tmp_path):
    o = Oracle(geom())
    sino = o.rs(shepp_logan(N))
    write_workdir(tmp_path)
    sino.tofile(tmp_path / f"Projections_{N}" / "P_ER_GPU.s")
    run(tmp_path, "--op=back_fdk", "--storage=mem_GPU")
    got = np.fromfile(tmp_path / "back_fdk.v", dtype=np.float32).reshape(o.geom.vol_shape)
    assert np.array_equal(got, o.vi_fdk(sino))


def test_cli_siddon_pair_readme_spelling(tmp_path):
    vol = ellipsoids(N, count=5, seed=2, amin=2.0, amax=8.0)
    vol.tofile(tmp_path / "phantom.v")
    write_workdir(tmp_path, real=str(tmp_path / "phantom.v"))
    out = run(tmp_path, "--op=proj", "--pair=SIDDONVI", "--storage=mem_GPU")
    assert "Siddon" in out
    got = np.fromfile(tmp_path / "proj.s", dtype=np.float32)
    assert np.array_equal(got, Oracle(geom()).siddon(vol).ravel())


def test_cli_iterative_reconstruction(tmp_path):
    o = Oracle(geom())
    truth = ellipsoids(N, count=5, seed=2, amin=2.0, amax=8.0)
    sino = o.rs(truth, cube=True)
    wd = write_workdir(tmp_path, iters=3, kopt=10, lam=0.01)
    sino.tofile(wd / "P_ER_GPU.s")
    out = run(tmp_path)
    res = list((wd / "TomoIter" / "Results").iterdir())
    assert len(res) == 1
    final = np.fromfile(res[0] / "final_Volume.v", dtype=np.float32).reshape(truth.shape)
    ref, stats = o.gd(sino, pair=0, cube=True, iters=3, kopt=10, lam=0.01)
    assert np.linalg.norm(final - ref) / np.linalg.norm(ref) <= 1e-6
    assert (res[0] / "configCT").exists() and (res[0] / "Iteration.dat").exists()
    assert "Gradient iteration 3" in out
    # finalVolume.v (the loop's own save, C/Iter3D_Algo_Gradient.cu:214) holds the same volume
    assert np.array_equal(np.fromfile(res[0] / "finalVolume.v", dtype=np.float32).reshape(truth.shape), final)
    # Iteration.dat in the reference's layout (C/Config/ConfigIteration.cpp:268-303): seven
    # setw(20) left-aligned header names, then IterNb J J_MC JReg lambda |alpha| beta eam
    lines = (res[0] / "Iteration.dat").read_text().splitlines()
    names = ["IterNb", "J", "J_MC", "JReg", "lambda", "alpha", "beta"]
    assert lines[0] == "".join(f"{n:<20}" for n in names)
    assert len(lines) == 4
    for k, (line, st) in enumerate(zip(lines[1:], stats), start=1):
        assert len(line) == 8 * 20 and all(line[20 * c] != " " for c in range(8)), line
        cols = [float(line[20 * c:20 * c + 20]) for c in range(8)]
        assert cols[0] == k and cols[4] == 0.01 and cols[6] == 0.0 and cols[7] == 0.0
        assert cols[2] == pytest.approx(st["residual_sq"], rel=1e-5)       # 6 significant digits
        assert cols[3] == pytest.approx(st["j_reg"], rel=1e-5)
        assert (cols[3] > 0) == (k > 1)   # sum (L x)^2 before the update: 0 for the zero start
        assert cols[1] == pytest.approx(cols[2] + 0.01 * cols[3], rel=1e-5)
        assert cols[5] == pytest.approx(abs(st["alpha"]), rel=1e-5)
    # the timings.log line, appended in the working directory (C/Iter3D_Algo_Gradient.cu:207-210)
    tl = (tmp_path / "timings.log").read_text().splitlines()
    assert len(tl) == 1 and tl[0].startswith("ngpu 1 nstreams 1 storage 0 precision 1 temps gradient ")
    assert tl[0].endswith("ms")


def test_cli_conjugate_gradient(tmp_path):
    o = Oracle(geom())
    sino = o.rs(ellipsoids(N, count=5, seed=2, amin=2.0, amax=8.0), cube=True)
    wd = write_workdir(tmp_path, iters=3, lam=0.01, gradient="Conjugate")
    sino.tofile(wd / "P_ER_GPU.s")
    run(tmp_path)
    res = list((wd / "TomoIter" / "Results").iterdir())
    final = np.fromfile(res[0] / "final_Volume.v", dtype=np.float32).reshape(o.geom.vol_shape)
    ref, _ = o.cg(sino, pair=0, cube=True, iters=3, lam=0.01)
    assert np.linalg.norm(final - ref) / np.linalg.norm(ref) <= 1e-6


def test_cli_weighted_gradient(tmp_path):
    """configIteration `gradient Weighted` reads <workdir>/weightedSinogram.s
    and runs tomo_wgd_run (alpha_min from -999999999999, as the reference's
    doWeightedGradient); the result matches the oracle's weighted loop."""
    wd = write_workdir(tmp_path, iters=3, kopt=10, lam=0.01, gradient="Weighted")
    o = Oracle(geom())
    y = o.rs(ellipsoids(N, count=4, seed=9, amin=3.0, amax=8.0), cube=True)
    w = (0.5 + np.random.default_rng(3).random(y.shape)).astype(np.float32)
    y.tofile(wd / "P_ER_GPU.s")
    w.tofile(wd / "weightedSinogram.s")
    run(tmp_path)
    res = sorted((wd / "TomoIter" / "Results").iterdir())[-1]
    got = np.fromfile(res / "final_Volume.v", dtype=np.float32).reshape(o.geom.vol_shape)
    ref, _ = o.gd(y, pair=0, cube=True, iters=3, kopt=10, lam=0.01, weight=w, alpha_min_init=-999999999999.0)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-6


def test_cli_missing_sinogram_is_an_error(tmp_path):
    write_workdir(tmp_path)
    r = subprocess.run([CLI, str(N), "--op=back"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 5 and "P_ER_GPU.s" in r.stderr
