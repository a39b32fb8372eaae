// tomogpi_main.cpp - command-line driver with the surface of TomoGPI_exe
// (src/TomoGPI_exe/TomoGPI_exe.cu:152-377 and C/Iter3D_main.cu:14-421):
//
//   tomogpi N [--op=proj|back|back_fdk] [--pair=RSVI|SiddonVI] [--compute=CUDA|CPU|C|OCL]
//             [--storage=mem_CPU|mem_GPU] [--precision=float|half|double]
//             [--no_load] [--no_save] [--ngpu=K] [--nstreams=K]
//
// Work directory Projections_N/ with configCT, configIteration,
// configComputeArchitecture, configTiff.  CLI flags override
// configComputeArchitecture.  Operator selection (FP32):
//   mem_GPU                    RS (inscribed cylinder) / clean Siddon + VI, device-resident
//   mem_CPU, RSVI              RS (circumscribed circle + square clip, the reference's default kernel) + VI
//   mem_CPU, SiddonVI, CPU/C   reference C-path Siddon (compat form, 256^3) + VI
//   mem_CPU, SiddonVI, other   clean Siddon + VI
// --precision=half stores volumes and sinograms in FP16 (scaled per array)
// with FP32 accumulation (the loops: every operator reads an FP16 copy of
// its FP32 input, tomo_gd_params.storage); its files stay FP32, as the reference's half
// containers convert on load and save (C/Image3D_CPU.cu:659-735,
// C/Sinogram3D_CPU.cu:556-606).  --precision=double runs the FP32 operators
// and reads and writes FP64 files, the layout of the reference's double
// containers (sizeof(T) * count, C/Image3D.cu:140-233, C/Sinogram3D.cu:
// 138-184).
//   --op=proj   loads configCT:realVolumeName, writes proj.s
//   --op=back   loads <workdir>/<configTiff:tiffName>.s, writes back.v
//   --op=back_fdk  same input, FDK-weighted backprojection, writes back_fdk.v
//   (no --op)   simple-gradient (configIteration gradient Simple), conjugate
//               gradient (gradient Conjugate) or weighted-gradient (gradient
//               Weighted, weights in <workdir>/weightedSinogram.s; a spelling
//               the reference's dispatcher, C/Iter3D.cu:78-90, does not
//               offer) reconstruction from the same sinogram, starting
//               from configCT:initVolumeName (or zero), writes
//               <workdir>/TomoIter/Results/Result_<date>/final_Volume.v
//               (unless --no_save) and finalVolume.v, Iteration.dat and the
//               copied config files there, and appends a line to
//               ./timings.log
// Every failure prints a message and returns a nonzero status.
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "tomogpi.h"
#include "config.hpp"

namespace {

struct Options {
    bool have_op = false;
    std::string op;
    std::string pair, compute, storage, precision;
    bool load = true, save = true;
};

// --key=value or -key=value; a bare --key sets an empty value.
bool flag(int argc, char** argv, const char* key, std::string* value) {
    for (int i = 2; i < argc; ++i) {
        const char* a = argv[i];
        if (a[0] != '-') continue;
        a += (a[1] == '-') ? 2 : 1;
        const char* eq = std::strchr(a, '=');
        const size_t len = eq ? (size_t)(eq - a) : std::strlen(a);
        if (std::strlen(key) == len && std::strncmp(a, key, len) == 0) {
            if (value) *value = eq ? std::string(eq + 1) : std::string();
            return true;
        }
    }
    return false;
}

int die(int rc, const std::string& what) {
    std::fprintf(stderr, "tomogpi: %s: %s (%s)\n", what.c_str(), tomo_status_string(rc), tomo_last_error());
    return rc ? rc : 1;
}

#define CHECK(call, what)                   \
    do {                                    \
        int rc_ = (call);                   \
        if (rc_ != TOMO_OK) return die(rc_, what); \
    } while (0)

std::string cfg(const std::string& dir, const std::string& file, const std::string& key, const std::string& dflt) {
    std::string v;
    if (tomo::config_lookup(tomo::join_path(dir, file), key, &v) != TOMO_OK) return dflt;
    return v;
}

// Both the short tokens the reference parses (ConfigComputeArchitecture.cpp:
// 12-45) and the enum spellings its shipped file uses.
std::string norm_pair(const std::string& s) {
    if (s == "SiddonVI" || s == "SIDDONVI" || s == "PAIR_SIDDONVI") return "SiddonVI";
    return "RSVI";
}
std::string norm_compute(const std::string& s) {
    if (s == "CPU" || s == "C" || s == "COMPUTE_C") return "C";
    if (s == "OCL" || s == "COMPUTE_OCL") return "OCL";
    return "CUDA";
}
std::string norm_storage(const std::string& s) {
    if (s == "mem_GPU" || s == "MEM_GPU") return "mem_GPU";
    return "mem_CPU";
}
std::string norm_precision(const std::string& s) {
    if (s == "half") return "half";
    if (s == "double") return "double";
    return "float";
}

float fp16_scale(const std::vector<float>& x) {
    float m = 0.0f;
    for (float v : x) m = std::fmax(m, std::fabs(v));
    return m > 0.0f ? 16384.0f / m : 1.0f;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::printf("usage: tomogpi N [--op=proj|back|back_fdk] [--pair=RSVI|SiddonVI] [--compute=...] [--storage=...] "
                    "[--precision=...] [--no_load] [--no_save]\n");
        return 1;
    }
    char wd[256];
    std::snprintf(wd, sizeof(wd), "Projections_%u/", (unsigned)std::strtoul(argv[1], nullptr, 10));
    const std::string work = wd;
    std::printf("%s\n", work.c_str());

    Options o;
    o.pair = norm_pair(cfg(work, "configComputeArchitecture", "pair", "RSVI"));
    o.compute = norm_compute(cfg(work, "configComputeArchitecture", "compute", "CUDA"));
    o.storage = norm_storage(cfg(work, "configComputeArchitecture", "storage", "mem_CPU"));
    o.precision = norm_precision(cfg(work, "configComputeArchitecture", "precision", "float"));
    std::string v;
    if (flag(argc, argv, "pair", &v)) o.pair = norm_pair(v);
    if (flag(argc, argv, "compute", &v)) o.compute = norm_compute(v);
    if (flag(argc, argv, "storage", &v)) o.storage = norm_storage(v);
    if (flag(argc, argv, "precision", &v)) o.precision = norm_precision(v);
    if (flag(argc, argv, "no_load", nullptr)) o.load = false;
    if (flag(argc, argv, "no_save", nullptr)) o.save = false;
    o.have_op = flag(argc, argv, "op", &o.op);
    // Devices and streams of the projection (configComputeArchitecture
    // gpuNbP / streamsNbP, overridden by --ngpu / --nstreams): recorded in
    // timings.log; the work runs on one GPU and one stream whatever they say.
    int gpu_nb = std::atoi(cfg(work, "configComputeArchitecture", "gpuNbP", "1").c_str());
    int streams_nb = std::atoi(cfg(work, "configComputeArchitecture", "streamsNbP", "1").c_str());
    if (flag(argc, argv, "ngpu", &v)) gpu_nb = std::atoi(v.c_str());
    if (flag(argc, argv, "nstreams", &v)) streams_nb = std::atoi(v.c_str());
    if (gpu_nb != 1 || streams_nb != 1)
        std::printf("note: %d GPU(s) / %d stream(s) requested; the work runs on one GPU and one stream\n", gpu_nb,
                    streams_nb);

    const bool device = o.storage == "mem_GPU";
    int proj = TOMO_PROJ_RS, pflags = 0;
    if (o.pair == "SiddonVI") {
        proj = TOMO_PROJ_SIDDON;
        if (!device && o.compute == "C") pflags |= TOMO_FLAG_SIDDON_COMPAT;
    } else if (!device) {
        pflags |= TOMO_FLAG_RS_FOV_CUBE;
    }
    const bool half = o.precision == "half";
    std::printf("--pair %s --compute %s --storage %s --precision %s -> %s%s%s\n", o.pair.c_str(), o.compute.c_str(),
                o.storage.c_str(), o.precision.c_str(), proj == TOMO_PROJ_RS ? "RS" : "Siddon",
                (pflags & TOMO_FLAG_RS_FOV_CUBE) ? " (cube FOV)" : (pflags & TOMO_FLAG_SIDDON_COMPAT) ? " (C path)" : "",
                " + VI");

    tomo_geom_desc desc{};
    CHECK(tomo_geom_desc_from_config(work.c_str(), &desc), "configCT");
    tomo_ctx* ctx = nullptr;
    CHECK(tomo_ctx_create(0, &ctx), "device");
    tomo_geom* geom = nullptr;
    CHECK(tomo_geom_create(ctx, &desc, &geom), "geometry");
    tomo_geom_info gi{};
    tomo_geom_get_info(geom, &gi);
    const size_t nvox = (size_t)gi.nx * gi.ny * gi.nz, nsino = (size_t)gi.nu * gi.nv * gi.nphi;
    const std::string tiff = cfg(work, "configTiff", "tiffName", "sinogram");
    const std::string sino_path = tomo::join_path(work, tiff + ".s");

    std::vector<float> vol(nvox, 0.0f), sino(nsino, 0.0f);
    using clk = std::chrono::high_resolution_clock;
    clk::time_point t0, t1;

    // One operator call on host arrays: device staging of in/out, FP16 storage for --precision=half.
    int bflags = 0;  // backprojector flags (TOMO_FLAG_FDK for --op=back_fdk)
    auto run_op = [&](bool forward, const std::vector<float>& in, std::vector<float>& out) -> int {
        const size_t nin = in.size(), nout = out.size();
        if (!half) {
            if (forward)
                return tomo_project(ctx, geom, proj, pflags | TOMO_FLAG_HOST_PTR, in.data(), TOMO_F32, 1.0f,
                                    out.data(), TOMO_F32, 1.0f, nullptr);
            return tomo_backproject(ctx, geom, TOMO_BP_VI, bflags | TOMO_FLAG_HOST_PTR, in.data(), TOMO_F32, 1.0f,
                                    out.data(), TOMO_F32, 1.0f, nullptr);
        }
        void *d32 = nullptr, *d16 = nullptr, *o16 = nullptr, *o32 = nullptr;
        int rc = TOMO_OK;
        const float sin_ = fp16_scale(in);
        const float sout = 1.0f;  // outputs are rescaled below from their FP32 copy
        if ((rc = tomo_alloc(ctx, nin * 4, &d32)) || (rc = tomo_alloc(ctx, nin * 2, &d16)) ||
            (rc = tomo_alloc(ctx, nout * 4, &o32)) || (rc = tomo_alloc(ctx, nout * 2, &o16)))
            goto done;
        if ((rc = tomo_copy_to_device(ctx, d32, in.data(), nin * 4))) goto done;
        if ((rc = tomo_convert(ctx, d32, TOMO_F32, 1.0f, d16, TOMO_F16, sin_, nin, nullptr))) goto done;
        // FP16 in, FP32 accumulate, FP32 out (a per-array output scale needs the output's range)
        rc = forward ? tomo_project(ctx, geom, proj, pflags, d16, TOMO_F16, sin_, o32, TOMO_F32, sout, nullptr)
                     : tomo_backproject(ctx, geom, TOMO_BP_VI, bflags, d16, TOMO_F16, sin_, o32, TOMO_F32, sout, nullptr);
        if (rc) goto done;
        if ((rc = tomo_synchronize(ctx, nullptr))) goto done;
        rc = tomo_copy_to_host(ctx, out.data(), o32, nout * 4);
    done:
        if (d32) tomo_free(ctx, d32);
        if (d16) tomo_free(ctx, d16);
        if (o16) tomo_free(ctx, o16);
        if (o32) tomo_free(ctx, o32);
        return rc;
    };

    // Volume / sinogram files: FP32, or FP64 with --precision=double.
    const int fdt = o.precision == "double" ? TOMO_F64 : TOMO_F32;
    auto read_file = [&](const std::string& path, std::vector<float>& dst, size_t n) -> int {
        if (fdt == TOMO_F32) return tomo_file_read(path.c_str(), dst.data(), TOMO_F32, n);
        std::vector<double> tmp(n);
        const int rc = tomo_file_read(path.c_str(), tmp.data(), TOMO_F64, n);
        if (rc == TOMO_OK)
            for (size_t i = 0; i < n; ++i) dst[i] = (float)tmp[i];
        return rc;
    };
    auto write_file = [&](const std::string& path, const std::vector<float>& src, size_t n) -> int {
        if (fdt == TOMO_F32) return tomo_file_write(path.c_str(), src.data(), TOMO_F32, n);
        const std::vector<double> tmp(src.begin(), src.begin() + (ptrdiff_t)n);
        return tomo_file_write(path.c_str(), tmp.data(), TOMO_F64, n);
    };

    if (o.have_op && o.op == "proj") {
        const std::string real = cfg(work, "configCT", "realVolumeName", "none");
        if (real != "none" && o.load) CHECK(read_file(real, vol, nvox), "realVolumeName");
        t0 = clk::now();
        CHECK(run_op(true, vol, sino), "projection");
        t1 = clk::now();
        if (o.save) CHECK(write_file("proj.s", sino, nsino), "proj.s");
    } else if (o.have_op && (o.op == "back" || o.op == "back_fdk")) {
        // back_fdk: FDK-weighted VI (C/Iter3D_main.cu:296-316), written to back_fdk.v
        const bool fdk = o.op == "back_fdk";
        bflags = fdk ? TOMO_FLAG_FDK : 0;
        if (o.load) CHECK(read_file(sino_path, sino, nsino), sino_path);
        t0 = clk::now();
        CHECK(run_op(false, sino, vol), "backprojection");
        t1 = clk::now();
        const char* name = fdk ? "back_fdk.v" : "back.v";
        if (o.save) CHECK(write_file(name, vol, nvox), name);
    } else if (o.have_op) {
        std::fprintf(stderr, "tomogpi: --op=%s is not supported (proj, back, back_fdk or none)\n", o.op.c_str());
        return TOMO_ERR_UNSUPPORTED;
    } else {
        int iters = 0, kopt = 0, pos = 0;
        double lambda = 0.0, beta = 2.0;
        const std::string ci = tomo::join_path(work, "configIteration");
        CHECK(tomo::config_int(ci, "gradientIterationNb", &iters), "configIteration");
        CHECK(tomo::config_int(ci, "optimalStepIterationNb", &kopt), "configIteration");
        CHECK(tomo::config_double(ci, "regularizationParameter", &lambda), "configIteration");
        CHECK(tomo::config_int(ci, "positivity", &pos), "configIteration");
        const std::string grad = cfg(work, "configIteration", "gradient", "Simple");
        const std::string pen = cfg(work, "configIteration", "penaltyFunction", "GGaussian");
        tomo::config_double(ci, "beta", &beta);
        if (grad != "Simple" && grad != "Conjugate" && grad != "Weighted") {
            std::fprintf(stderr, "tomogpi: gradient %s is not supported (Simple, Conjugate, Weighted)\n", grad.c_str());
            return TOMO_ERR_UNSUPPORTED;
        }
        if (lambda != 0.0 && (pen != "GGaussian" || beta != 2.0)) {
            std::fprintf(stderr, "tomogpi: penalty %s beta %g is not supported (GGaussian, beta 2)\n", pen.c_str(), beta);
            return TOMO_ERR_UNSUPPORTED;
        }
        if (o.load) CHECK(read_file(sino_path, sino, nsino), sino_path);
        const std::string init = cfg(work, "configCT", "initVolumeName", "none");
        if (init != "none" && o.load) CHECK(read_file(init, vol, nvox), "initVolumeName");
        // results directory <workdir>/TomoIter/Results/Result_<date>/ with the config files copied
        char stamp[64];
        const time_t now = time(nullptr);
        strftime(stamp, sizeof(stamp), "Result_%Y_%m_%d_%Hh%M/", localtime(&now));
        const std::string res = work + "TomoIter/Results/" + stamp;
        mkdir((work + "TomoIter").c_str(), 0775);
        mkdir((work + "TomoIter/Results").c_str(), 0775);
        mkdir(res.c_str(), 0775);
        for (const char* f : {"configCT", "configIteration", "configComputeArchitecture", "configTiff"}) {
            std::ifstream src(work + f, std::ios::binary);
            std::ofstream dst(res + f, std::ios::binary);
            dst << src.rdbuf();
        }
        tomo_gd_params gp{};
        gp.pair = o.pair == "SiddonVI" ? TOMO_PAIR_SIDDONVI : TOMO_PAIR_RSVI;
        gp.flags = pflags | TOMO_FLAG_HOST_PTR;
        gp.gradient_iteration_nb = iters;
        gp.optimal_step_iteration_nb = kopt;
        gp.lambda = lambda;
        gp.positivity = pos;
        gp.alpha_min_init = -0.0347711;
        gp.storage = half ? TOMO_F16 : TOMO_F32;  // the half loop (C/Iter3D_half.cu:545-700): FP16 operator inputs
        std::vector<float> weight;
        if (grad == "Weighted") {
            // doWeightedGradient: weights from <workdir>/weightedSinogram.s,
            // alpha_min from -999999999999 (C/Iter3D_Algo_Gradient.cu:237-250)
            weight.assign(nsino, 0.0f);
            CHECK(read_file(work + "weightedSinogram.s", weight, nsino), "weightedSinogram.s");
            gp.alpha_min_init = -999999999999.0;
        }
        std::vector<tomo_gd_iter_stat> st(iters > 0 ? iters : 1);
        t0 = clk::now();
        if (grad == "Conjugate")
            CHECK(tomo_cg_run(ctx, geom, &gp, sino.data(), vol.data(), st.data(), nullptr), "conjugate gradient");
        else if (grad == "Weighted")
            CHECK(tomo_wgd_run(ctx, geom, &gp, sino.data(), weight.data(), vol.data(), st.data(), nullptr),
                  "weighted gradient");
        else
            CHECK(tomo_gd_run(ctx, geom, &gp, sino.data(), vol.data(), st.data(), nullptr), "gradient descent");
        t1 = clk::now();
        // Iteration.dat as ConfigIteration::createIterationLogFile / updateLogFile
        // write it (C/Config/ConfigIteration.cpp:268-303): a header of seven
        // setw(20) left-aligned names, then per iteration IterNb J J_MC JReg
        // lambda |alpha| beta eam in default ostream formatting.  J_MC is the
        // data term sum (y - P x)^2 (the reference leaves its own at 0: the
        // assignment at C/Iter3D_Algo_Gradient.cu:88 is commented out), JReg
        // = sum (L x)^2, J = J_MC + lambda JReg, beta the conjugate
        // gradient's, eam 0 (no reference volume is loaded).
        std::ofstream log((res + "Iteration.dat").c_str(), std::ios::out);
        if (log.is_open()) {
            log << std::setw(20) << std::left << "IterNb" << std::setw(20) << std::left << "J" << std::setw(20)
                << std::left << "J_MC" << std::setw(20) << std::left << "JReg" << std::setw(20) << std::left
                << "lambda" << std::setw(20) << std::left << "alpha" << std::setw(20) << std::left << "beta"
                << std::endl;
        }
        for (int k = 0; k < iters; ++k) {
            std::printf("Gradient iteration %d: |y-Px|^2 = %.9g  alpha = %.9g\n", k + 1, st[k].residual_sq,
                        st[k].alpha);
            if (!log.is_open()) continue;
            const double j_mc = st[k].residual_sq, j_reg = st[k].j_reg;
            const double j = j_mc + lambda * j_reg;
            log << std::setw(20) << std::left << (unsigned short)(k + 1) << std::setw(20) << std::left << j
                << std::setw(20) << std::left << j_mc << std::setw(20) << std::left << j_reg << std::setw(20)
                << std::left << lambda << std::setw(20) << std::left << std::fabs(st[k].alpha) << std::setw(20)
                << std::left << st[k].beta << std::setw(20) << std::left << 0.0 << std::endl;
        }
        log.close();
        // The loop's own outputs (C/Iter3D_Algo_Gradient.cu:202-214): the
        // "timings.log" line appended in the working directory, and
        // finalVolume.v in the results directory whatever --no_save says.
        const long gms = (long)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
        std::printf("Gradient computation :%ldms\n", gms);
        {
            std::ofstream tl("timings.log", std::ios::app);
            tl << "ngpu " << gpu_nb << " nstreams " << streams_nb << " storage " << (device ? 1 : 0) << " precision "
               << (o.precision == "half" ? 0 : o.precision == "double" ? 2 : 1) << " temps gradient " << gms
               << "ms\n";
        }
        CHECK(write_file(res + "finalVolume.v", vol, nvox), "finalVolume.v");
        // doMainIter's own save of the result (C/Iter3D_main.cu:410), skipped with --no_save
        if (o.save) CHECK(write_file(res + "final_Volume.v", vol, nvox), "final_Volume.v");
        std::printf("results in %s\n", res.c_str());
    }
    const long ms = (long)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
    std::printf("Time computation :%ldms\n", ms);
    tomo_geom_destroy(geom);
    tomo_ctx_destroy(ctx);
    return 0;
}
