"""Host-side parts of the C ABI that run without a GPU: the library loads and
exports every symbol include/tomogpi.h declares; configCT parsing of the
reference's shipped config files (tests/golden/Projections_256, copied data);
error codes instead of exit(); raw .v/.s file I/O."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tomogpi_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "Projections_256")


def declared_functions():
    text = open(os.path.join(ROOT, "include", "tomogpi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tomo_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = lib.load()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(lib.EXPORTS) == names
    assert L.tomo_abi_version() == 4


def test_config_ct_of_shipped_dataset():
    d = lib.AcquisitionDesc.from_config(GOLD)
    assert d.fod_mm == 98.0 and d.fdd_mm == 230.0
    assert d.projection_nb == 256
    assert (d.num_start_angle, d.den_start_angle, d.num_stop_angle, d.den_stop_angle) == (0, 1, 2, 1)
    assert d.u_detector_size_mm == np.float32(14.2) and d.v_detector_size_mm == np.float32(14.2)
    assert (d.u_pixel_nb, d.v_pixel_nb) == (256, 256)
    assert d.x_volume_size_mm == np.float32(6.05) and d.z_volume_size_mm == np.float32(5.86)
    assert (d.x_pixel_nb, d.y_pixel_nb, d.z_pixel_nb) == (256, 256, 256)


def test_missing_key_and_file_are_errors_not_exit(tmp_path):
    (tmp_path / "configCT").write_text("FOD 10 mm\nFDD 20 mm\n")
    with pytest.raises(lib.TomoError) as e:
        lib.AcquisitionDesc.from_config(str(tmp_path))
    assert e.value.status == 6 and "projectionNb" in str(e.value)
    with pytest.raises(lib.TomoError) as e:
        lib.AcquisitionDesc.from_config(str(tmp_path / "nowhere"))
    assert e.value.status == 6


def test_token_scan_first_match_and_next_token(tmp_path):
    # the value is the token after the first occurrence of the key, anywhere on a line
    (tmp_path / "configCT").write_text(
        "junk FOD 12.5 mm FOD 99\nFOD 77\nFDD 40\nprojectionNb 8 (comment)\nnumStartAngle 0\ndenStartAngle 1\n"
        "numStopAngle 2\ndenStopAngle 1\nhorizontalDetectorSize 10\nverticalDetectorSize 10\n"
        "horizontalDetectorPixelNb 4\nverticalDetectorPixelNb 4\nxVolumeSize 4\nyVolumeSize 4\nzVolumeSize 4\n"
        "xVolumePixelNb 4\nyVolumePixelNb 4\nzVolumePixelNb\n")
    d = lib.AcquisitionDesc.from_config(str(tmp_path))
    assert d.fod_mm == 12.5 and d.fdd_mm == 40.0 and d.projection_nb == 8
    assert d.z_pixel_nb == 0   # key with no value: atoi("") as in the reference


def test_raw_file_io_roundtrip_and_short_file(tmp_path):
    L = lib.load()
    a = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    p = str(tmp_path / "vol.v").encode()
    lib.check(L.tomo_file_write(p, a.ctypes.data, lib.F32, a.size))
    assert os.path.getsize(p) == a.nbytes
    b = np.zeros_like(a)
    lib.check(L.tomo_file_read(p, b.ctypes.data, lib.F32, b.size))
    assert np.array_equal(a, b)
    big = np.zeros(30, dtype=np.float32)
    rc = L.tomo_file_read(p, big.ctypes.data, lib.F32, big.size)
    assert rc == 5 and b"short file" in L.tomo_last_error()
    assert L.tomo_file_read(str(tmp_path / "none.s").encode(), big.ctypes.data, lib.F32, 1) == 5


def test_status_strings():
    L = lib.load()
    assert L.tomo_status_string(0) == b"ok"
    assert L.tomo_status_string(6) == b"configuration error"


def test_null_arguments_are_errors_without_a_device():
    """Argument checks come before any device call: null handles and bad
    enums return TOMO_ERR_INVALID_ARG (1) on a machine without a GPU."""
    L = lib.load()
    st = lib.GdIterStat()
    assert L.tomo_gd_run(None, None, None, None, None, C.byref(st), None) == 1
    assert L.tomo_cg_run(None, None, None, None, None, C.byref(st), None) == 1
    assert L.tomo_sirt_run(None, None, None, None, None, None, None) == 1
    srv = C.c_void_p()
    p = lib.ServeParams(max_active=4)
    assert L.tomo_serve_create(None, C.byref(p), C.byref(srv)) == 1
    rid = C.c_uint64()
    req = lib.ServeRequest()
    assert L.tomo_serve_submit(None, C.byref(req), C.byref(rid)) == 1
    res = lib.ServeResult()
    assert L.tomo_serve_poll(None, 1, C.byref(res)) == 1
    assert L.tomo_serve_wait(None, 1, 0.0, C.byref(res)) == 1
    assert L.tomo_serve_destroy(None) == 0
    assert b"null" in L.tomo_last_error()
