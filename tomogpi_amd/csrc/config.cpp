// config.cpp - configuration files, geometry derivation and raw file I/O.
//
// Config files are the reference's whitespace key/value files
// (Data3D_0256/Projections_256/config*).  Lookup follows
// Config::getConfigFileField (C/Config/Config.cpp:34-180): read line by line,
// split on whitespace, the first token equal to the key selects the line and
// the next token on that line is the value (empty if none); numbers convert
// with atoi/atof.  Unlike the reference, a missing file or key is reported
// as TOMO_ERR_CONFIG instead of exit(EXIT_FAILURE).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "config.hpp"
#include "tomo_internal.hpp"

namespace tomo {

int config_lookup(const std::string& path, const std::string& key, std::string* value) {
    std::ifstream f(path);
    if (!f.is_open()) return fail(TOMO_ERR_CONFIG, "Unable to open file " + path);
    std::string row, tok;
    while (std::getline(f, row)) {
        std::istringstream ss(row);
        while (ss >> tok) {
            if (tok == key) {
                std::string v;
                ss >> v;
                *value = v;
                return TOMO_OK;
            }
        }
    }
    return fail(TOMO_ERR_CONFIG, "Unable to find field " + key + " in " + path);
}

int config_int(const std::string& path, const std::string& key, int* out) {
    std::string v;
    int rc = config_lookup(path, key, &v);
    if (rc) return rc;
    *out = std::atoi(v.c_str());
    return TOMO_OK;
}

int config_double(const std::string& path, const std::string& key, double* out) {
    std::string v;
    int rc = config_lookup(path, key, &v);
    if (rc) return rc;
    *out = std::atof(v.c_str());
    return TOMO_OK;
}

std::string join_path(const std::string& dir, const std::string& file) {
    if (dir.empty()) return file;
    if (dir.back() == '/') return dir + file;
    return dir + "/" + file;
}

// ConfigCT (C/Config/ConfigCT.cpp:110-164): FOD/FDD and the angles as
// double, detector and volume sizes as float, pixel counts through atoi.
int read_config_ct(const std::string& workdir, tomo_geom_desc* d) {
    const std::string p = join_path(workdir, "configCT");
    double v = 0;
    int iv = 0;
    int rc;
#define RD(key, field)                              \
    if ((rc = config_double(p, key, &v))) return rc; \
    d->field = v;
#define RF(key, field)                              \
    if ((rc = config_double(p, key, &v))) return rc; \
    d->field = (float)v;
#define RI(key, field)                           \
    if ((rc = config_int(p, key, &iv))) return rc; \
    d->field = iv;
    RD("FOD", fod_mm)
    RD("FDD", fdd_mm)
    RI("projectionNb", projection_nb)
    RD("numStartAngle", num_start_angle)
    RD("denStartAngle", den_start_angle)
    RD("numStopAngle", num_stop_angle)
    RD("denStopAngle", den_stop_angle)
    RF("horizontalDetectorSize", u_detector_size_mm)
    RF("verticalDetectorSize", v_detector_size_mm)
    RI("horizontalDetectorPixelNb", u_pixel_nb)
    RI("verticalDetectorPixelNb", v_pixel_nb)
    RF("xVolumeSize", x_volume_size_mm)
    RF("yVolumeSize", y_volume_size_mm)
    RF("zVolumeSize", z_volume_size_mm)
    RI("xVolumePixelNb", x_pixel_nb)
    RI("yVolumePixelNb", y_pixel_nb)
    RI("zVolumePixelNb", z_pixel_nb)
#undef RD
#undef RF
#undef RI
    return TOMO_OK;
}

// Derived geometry (C/Config/ConfigCT.cpp:119, C/Acquisition.cpp:90-97,
// C/Detector.cpp:9-14, C/Volume.cu:17-25, C/Projector.cu:15-38) and the
// per-view FP64 table of the kernels.
int derive_geometry(const tomo_geom_desc& d, tomo_geom* g) {
    if (d.x_pixel_nb <= 0 || d.y_pixel_nb <= 0 || d.z_pixel_nb <= 0 || d.u_pixel_nb <= 0 || d.v_pixel_nb <= 0 ||
        d.projection_nb <= 0)
        return fail(TOMO_ERR_INVALID_ARG, "pixel and projection counts must be positive");
    if (!(d.fod_mm > 0.0) || !(d.fdd_mm > d.fod_mm))
        return fail(TOMO_ERR_INVALID_ARG, "geometry needs 0 < FOD < FDD");
    if (!(d.u_detector_size_mm > 0.0f) || !(d.v_detector_size_mm > 0.0f) || !(d.x_volume_size_mm > 0.0f))
        return fail(TOMO_ERR_INVALID_ARG, "detector and volume sizes must be positive");
    if (d.den_start_angle == 0.0 || d.den_stop_angle == 0.0)
        return fail(TOMO_ERR_INVALID_ARG, "angle denominators must be nonzero");
    const size_t nvox = (size_t)d.x_pixel_nb * d.y_pixel_nb * d.z_pixel_nb;
    const size_t nsino = (size_t)d.u_pixel_nb * d.v_pixel_nb * d.projection_nb;
    if (nvox >= (1ull << 32) || nsino >= (1ull << 32))
        return fail(TOMO_ERR_INVALID_ARG, "volume and sinogram must each hold fewer than 2^32 elements");
    g->desc = d;
    tomo_geom_info& in = g->info;
    in.nx = d.x_pixel_nb; in.ny = d.y_pixel_nb; in.nz = d.z_pixel_nb;
    in.nu = d.u_pixel_nb; in.nv = d.v_pixel_nb; in.nphi = d.projection_nb;
    in.fod = d.fod_mm; in.fdd = d.fdd_mm;
    in.du = d.u_detector_size_mm / (float)d.u_pixel_nb;
    in.dv = d.v_detector_size_mm / (float)d.v_pixel_nb;
    in.cu = (float)(d.u_pixel_nb / 2.0 - 0.5);
    in.cv = (float)(d.v_pixel_nb / 2.0 - 0.5);
    in.delta = d.x_volume_size_mm / (float)d.x_pixel_nb;
    in.cx = (float)(d.x_pixel_nb / 2.0 - 0.5);
    in.cy = (float)(d.y_pixel_nb / 2.0 - 0.5);
    in.cz = (float)(d.z_pixel_nb / 2.0 - 0.5);
    const double start = d.num_start_angle / d.den_start_angle;
    const double stop = d.num_stop_angle / d.den_stop_angle;
    in.start_angle = start;
    in.step_angle = M_PI * (stop - start) / (double)d.projection_nb;
    const int np = d.projection_nb;
    g->phi.resize(np);
    g->view.assign((size_t)np * kViewStride, 0.0);
    g->view_f.resize((size_t)np * 2);
    const double dl = (double)in.delta, ku = in.fdd / (double)in.du, kv = in.fdd / (double)in.dv;
    for (int p = 0; p < np; ++p) {
        const double phi = ((double)p + start) * in.step_angle;
        const double c = std::cos(phi), s = std::sin(phi);
        g->phi[p] = phi;
        double* v = &g->view[(size_t)p * kViewStride];
        v[0] = c;
        v[1] = s;
        v[2] = dl * c;
        v[3] = dl * s;
        v[4] = -(dl * s) * ku;
        v[5] = (dl * c) * ku;
        v[6] = dl * kv;
        g->view_f[2 * p] = (float)c;
        g->view_f[2 * p + 1] = (float)s;
    }
    DevGeom& dv = g->dev;
    dv.nx = in.nx; dv.ny = in.ny; dv.nz = in.nz; dv.nu = in.nu; dv.nv = in.nv; dv.nphi = in.nphi;
    dv.fod = in.fod; dv.fdd = in.fdd;
    dv.du = in.du; dv.dv = in.dv; dv.cu = in.cu; dv.cv = in.cv;
    dv.delta = in.delta; dv.cx = in.cx; dv.cy = in.cy; dv.cz = in.cz;
    return TOMO_OK;
}

size_t dtype_size(int dtype) {
    switch (dtype) {
        case TOMO_F32: return 4;
        case TOMO_F16: return 2;
        case TOMO_BF16: return 2;
        case TOMO_F64: return 8;
        default: return 0;
    }
}

}  // namespace tomo

extern "C" {

int tomo_geom_desc_from_config(const char* workdir, tomo_geom_desc* out) {
    if (!workdir || !out) return tomo::fail(TOMO_ERR_INVALID_ARG, "null argument");
    tomo_geom_desc d{};
    int rc = tomo::read_config_ct(workdir, &d);
    if (rc) return rc;
    *out = d;
    return TOMO_OK;
}

// Headerless native-endian dumps (C/Image3D.cu:140-233, C/Sinogram3D.cu:
// 138-184); a short or missing file is TOMO_ERR_IO instead of exit().
int tomo_file_read(const char* path, void* dst, int dtype, size_t count) {
    const size_t es = tomo::dtype_size(dtype);
    if (!path || !dst || es == 0) return tomo::fail(TOMO_ERR_INVALID_ARG, "bad file read arguments");
    FILE* f = std::fopen(path, "rb");
    if (!f) return tomo::fail(TOMO_ERR_IO, std::string("Unable to open file ") + path);
    const size_t got = std::fread(dst, es, count, f);
    std::fclose(f);
    if (got != count)
        return tomo::fail(TOMO_ERR_IO, std::string("short file ") + path + ": " + std::to_string(got) + " of " +
                                           std::to_string(count) + " elements");
    return TOMO_OK;
}

int tomo_file_write(const char* path, const void* src, int dtype, size_t count) {
    const size_t es = tomo::dtype_size(dtype);
    if (!path || !src || es == 0) return tomo::fail(TOMO_ERR_INVALID_ARG, "bad file write arguments");
    FILE* f = std::fopen(path, "wb");
    if (!f) return tomo::fail(TOMO_ERR_IO, std::string("Unable to create file ") + path);
    const size_t put = std::fwrite(src, es, count, f);
    const int cl = std::fclose(f);
    if (put != count || cl != 0) return tomo::fail(TOMO_ERR_IO, std::string("write failed: ") + path);
    return TOMO_OK;
}

}  // extern "C"
