/*
 * tomogpi.h - C ABI of the MI355X cone-beam projector / backprojector library
 * (libtomogpi_mi355x.so).
 *
 * The library replaces the operator layer of TomoGPI (reference paths are
 * relative to the reference repository root; C = src/TomoGPI_lib/src/CPP,
 * I = src/TomoGPI_lib/inc):
 *
 *   Projector<V,S,T>::doProjection(S<T>* sino, V<T>* vol)      I/Projector.cuh:118
 *   BackProjector<V,S,T>::doBackProjection(V<T>* vol, S<T>* s)  I/BackProjector.cuh:144
 *   Projector<V,S,T>::updateVolume(out, in, lambda, positive)   C/Projector.cu:109-124
 *   Iter3D::doSimpleGradient(...)                               C/Iter3D_Algo_Gradient.cu:25-223
 *
 * Conventions (identical to the reference containers):
 *   volume   f[x + Nx*(y + Ny*z)]          x fastest   (C/Image3D.cu:140-233)
 *   sinogram g[u + Nu*(v + Nv*p)]          u fastest   (C/Sinogram3D.cu:138-184)
 *   Every operator OVERWRITES its output (reference kernels store, never
 *   accumulate: K/BackProjector/backprojection3D_kernel_16reg_UM.cuh:129-145,
 *   K/Projector/projection_ER_kernel_UM.cuh:143).
 *
 * Every entry point returns TOMO_OK (0) or a TOMO_ERR_* code and never calls
 * exit(); tomo_last_error() returns a thread-local message for the last
 * failure on the calling thread.  Pointers are device pointers unless the
 * TOMO_HOST_PTR flag is given; `stream` is a hipStream_t (NULL = the null
 * stream).  Device-pointer calls are stream-ordered and asynchronous;
 * TOMO_HOST_PTR calls are synchronous.  A context may be used from several
 * threads; calls that share a geometry and a stream are ordered by the stream.
 */
#ifndef TOMOGPI_H
#define TOMOGPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TOMOGPI_ABI_VERSION 4

/* ---- status codes ------------------------------------------------------ */
#define TOMO_OK                0
#define TOMO_ERR_INVALID_ARG   1   /* bad pointer, size, enum or geometry   */
#define TOMO_ERR_HIP           2   /* a HIP runtime call failed             */
#define TOMO_ERR_UNSUPPORTED   3   /* combination not implemented           */
#define TOMO_ERR_OUT_OF_MEMORY 4   /* device allocation refused             */
#define TOMO_ERR_IO            5   /* file missing, short or unwritable     */
#define TOMO_ERR_CONFIG        6   /* config file or key missing            */
#define TOMO_ERR_BUSY          7   /* serving queue full / not admitted     */

/* ---- element types of volumes and sinograms ---------------------------- */
#define TOMO_F32   0
#define TOMO_F16   1   /* stored = value * scale, decoded as stored / scale */
#define TOMO_BF16  2
#define TOMO_F64   3   /* host-side I/O only */

/* ---- operators ---------------------------------------------------------- */
#define TOMO_PROJ_RS      0   /* regular-sampling ray-driven projector      */
#define TOMO_PROJ_SIDDON  1   /* Siddon exact-chord projector                */
#define TOMO_BP_VI        0   /* voxel-driven bilinear backprojector         */

/* ---- flags --------------------------------------------------------------- */
#define TOMO_FLAG_RS_FOV_CUBE   0x1  /* RS: circumscribed circle + cube clip
                                        (projection_ER_kernel.cuh:56-312, the
                                        reference's mem_CPU default); default
                                        is the inscribed cylinder
                                        (projection_ER_kernel_UM.cuh:25-153) */
#define TOMO_FLAG_SIDDON_COMPAT 0x2  /* Siddon: reference C-path arithmetic and
                                        quirks (SiddonProjector_compute_C_mem_CPU
                                        .cpp:20-276): float traversal, first
                                        voxel half a voxel behind, last chord
                                        unclipped, indices wrapped modulo N, no
                                        z clip. Default is the clean Siddon. */
#define TOMO_FLAG_HOST_PTR      0x4  /* data pointers are host memory: stage,
                                        compute, copy back, synchronise       */
#define TOMO_FLAG_FDK           0x8  /* VI: FDK-weighted backprojection
                                        (back_fdk; K/BackProjector/FDK3D_kernel
                                        _16reg.cuh:71-88, C/BackProjector.cu:
                                        34-35): f = sum_p W_p bilinear(g_p),
                                        W_p = alphaC (FDD/du) / w^2 /
                                        sqrt(betaC + u'^2 + v'^2), no delta   */
#define TOMO_FLAG_INDEX_READS   0x10 /* RS, clean Siddon: read the z-pair copies
                                        through 64-bit element indices, the
                                        form every copy of 4 GiB or more
                                        takes (1024^3 in FP32 or 16-bit),
                                        also where 32-bit buffer offsets
                                        apply; same results.  Lets that form
                                        be checked at small sizes.           */
#define TOMO_FLAG_STAGED        0x20 /* RS, clean Siddon on FP32 volumes:
                                        the workgroup stages the voxels its
                                        rays cross, a window of slabs at a
                                        time, in LDS and reads its taps there
                                        instead of gathering them per lane
                                        from the z-pair copies; same results */

typedef struct tomo_ctx  tomo_ctx;
typedef struct tomo_geom tomo_geom;

/* Raw acquisition description, field for field the keys of the reference's
 * configCT file (C/Config/ConfigCT.cpp:110-164, Data3D_0256/Projections_256/
 * configCT).  Derived quantities follow the reference constructors:
 *   step = pi*(stop-start)/Nphi, phi_p = (p+start)*step   (ConfigCT.cpp:119,
 *                                                           Acquisition.cpp:90-97)
 *   du = Su/Nu, cu = Nu/2-0.5 (float)                      (Detector.cpp:9-14)
 *   delta = Sx/Nx, delta_z := delta, c = N/2-0.5 (float)   (Volume.cu:17-25)  */
typedef struct tomo_geom_desc {
    double  fod_mm;                 /* FOD: focus-object distance            */
    double  fdd_mm;                 /* FDD: focus-detector distance          */
    int32_t projection_nb;          /* projectionNb                          */
    double  num_start_angle, den_start_angle;   /* start = num/den (x pi)    */
    double  num_stop_angle,  den_stop_angle;    /* stop  = num/den (x pi)    */
    float   u_detector_size_mm;     /* horizontalDetectorSize                */
    float   v_detector_size_mm;     /* verticalDetectorSize                  */
    int32_t u_pixel_nb;             /* horizontalDetectorPixelNb             */
    int32_t v_pixel_nb;             /* verticalDetectorPixelNb               */
    float   x_volume_size_mm, y_volume_size_mm, z_volume_size_mm;
    int32_t x_pixel_nb, y_pixel_nb, z_pixel_nb;
} tomo_geom_desc;

/* Derived geometry as the kernels see it (read-only view). */
typedef struct tomo_geom_info {
    int32_t nx, ny, nz, nu, nv, nphi;
    double  fod, fdd;
    float   du, dv, cu, cv;         /* detector pixel size and centre pixel  */
    float   delta, cx, cy, cz;      /* voxel size (all axes) and centre      */
    double  step_angle, start_angle;
} tomo_geom_info;

/* ---- library / context --------------------------------------------------- */
int         tomo_abi_version(void);
const char* tomo_last_error(void);
const char* tomo_status_string(int status);

/* Binds device `device` (hipSetDevice) for every later call on this context. */
int tomo_ctx_create(int device, tomo_ctx** out);
int tomo_ctx_destroy(tomo_ctx* ctx);
/* Free / total HBM of the context's device (bytes). */
int tomo_ctx_mem_info(tomo_ctx* ctx, size_t* free_bytes, size_t* total_bytes);

/* Device memory helpers for callers without their own allocator. */
int tomo_alloc(tomo_ctx* ctx, size_t bytes, void** dptr);
int tomo_free(tomo_ctx* ctx, void* dptr);
int tomo_copy_to_device(tomo_ctx* ctx, void* dst, const void* src, size_t bytes);
int tomo_copy_to_host(tomo_ctx* ctx, void* dst, const void* src, size_t bytes);
int tomo_synchronize(tomo_ctx* ctx, void* stream);

/* ---- geometry ------------------------------------------------------------ */
/* Builds the per-view FP64 tables and uploads them to the context's device.  */
int tomo_geom_create(tomo_ctx* ctx, const tomo_geom_desc* desc, tomo_geom** out);
/* Reads <workdir>/configCT with the reference's token-scan semantics
 * (C/Config/Config.cpp:34-180: first token equal to the key, value = next
 * token; numbers through atoi/atof).  A missing file or key returns
 * TOMO_ERR_CONFIG instead of exiting.                                        */
int tomo_geom_desc_from_config(const char* workdir, tomo_geom_desc* out);
int tomo_geom_destroy(tomo_geom* geom);
int tomo_geom_get_info(const tomo_geom* geom, tomo_geom_info* out);
/* phi_p in radians, FP64, p = 0..nphi-1 (Acquisition.cpp:90-97). */
int tomo_geom_get_phi(const tomo_geom* geom, double* phi_out);

/* ---- operators (replace doProjection / doBackProjection) ------------------ */
/* sinogram <- P volume.  projector = TOMO_PROJ_RS | TOMO_PROJ_SIDDON.
 * Replaces RegularSamplingProjector_compute_CUDA_mem_{GPU,CPU}::doProjection
 * (C/RegularSamplingProjector_compute_CUDA_mem_GPU.cu:35-92, ..._mem_CPU.cu:26)
 * and SiddonProjector_compute_C_mem_CPU::doProjection
 * (C/SiddonProjector_compute_C_mem_CPU.cpp:20-276).                          */
int tomo_project(tomo_ctx* ctx, const tomo_geom* geom, int projector, int flags,
                 const void* volume, int volume_dtype, float volume_scale,
                 void* sinogram, int sinogram_dtype, float sinogram_scale,
                 void* stream);

/* volume <- B sinogram.  backprojector = TOMO_BP_VI.
 * Replaces VIBackProjector_compute_CUDA_mem_{GPU,CPU}::doBackProjection
 * (C/VIBackProjector_compute_CUDA_mem_GPU.cu:26-111, ..._mem_CPU.cu:19-690).
 * An F32 or BF16 sinogram is interpolated in FP32 (BF16 taps widened first);
 * an F16 sinogram with FP16 u-weights in steps of 1/2048 (two taps per
 * v_dot2_f32_f16) and FP32 sums.                                           */
int tomo_backproject(tomo_ctx* ctx, const tomo_geom* geom, int backprojector, int flags,
                     const void* sinogram, int sinogram_dtype, float sinogram_scale,
                     void* volume, int volume_dtype, float volume_scale,
                     void* stream);

/* Exact transposes, new (the reference has none; needed for <Ax,y>=<x,A^T y>).
 * volume <- P^T sinogram and sinogram <- B^T volume.  Scatter with FP32
 * atomics; output is zeroed first.                                           */
int tomo_project_adjoint(tomo_ctx* ctx, const tomo_geom* geom, int projector, int flags,
                         const void* sinogram, int sinogram_dtype, float sinogram_scale,
                         void* volume, int volume_dtype, float volume_scale,
                         void* stream);
int tomo_backproject_adjoint(tomo_ctx* ctx, const tomo_geom* geom, int backprojector, int flags,
                             const void* volume, int volume_dtype, float volume_scale,
                             void* sinogram, int sinogram_dtype, float sinogram_scale,
                             void* stream);

/* Work units of one projection, for roofline accounting: the number of
 * trilinear samples sum L_r (RS) or of positive-length chords (clean
 * Siddon) over all rays.  Synchronous.                                       */
int tomo_project_work(tomo_ctx* ctx, const tomo_geom* geom, int projector, int flags,
                      uint64_t* units, void* stream);

/* out <- in, element-type conversion with the storage scales.               */
int tomo_convert(tomo_ctx* ctx, const void* in, int in_dtype, float in_scale,
                 void* out, int out_dtype, float out_scale, size_t n, void* stream);

/* ---- loop glue (Image3D/Sinogram3D vector ops, C/Image3D_CPU.cu:108-390,
 *      C/Sinogram3D_CPU.cu:311-375) ------------------------------------------ */
/* out = a - b  (diffSinogram: this = sinogram1 - sinogram2)                  */
int tomo_diff(tomo_ctx* ctx, const float* a, const float* b, float* out, size_t n, void* stream);
/* x *= s  (scalarVolume)                                                     */
int tomo_scale(tomo_ctx* ctx, float* x, float s, size_t n, void* stream);
/* y += lambda * x, then y = max(y,0) if positive (updateVolume)             */
int tomo_axpy(tomo_ctx* ctx, float* y, const float* x, double lambda, int positive,
              size_t n, void* stream);
/* sum_i double(x_i*x_i)  ("L2Norm" of the reference: no square root).
 * Synchronous: the result is on the host when the call returns.             */
int tomo_sq_norm(tomo_ctx* ctx, const void* x, int dtype, float scale, size_t n,
                 double* out, void* stream);
/* tomo_sq_norm without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels, summation order and result as
 * tomo_sq_norm; the partial sums live in a 16 KiB device array of the
 * library (2048 partials and the final sum), one per GPU, held until the
 * process ends.  Every n fits it (the partials are capped at 2048 blocks);
 * TOMO_ERR_INVALID_ARG if the context's device has no array set up.  Calls
 * that overlap on one GPU (other threads or streams) race on that array.
 * Synchronous.                                                             */
int tomo_sq_norm_noalloc(tomo_ctx* ctx, const void* x, int dtype, float scale, size_t n,
                         double* out, void* stream);
/* <a,b> in FP64 (adjoint tests).  Synchronous.                               */
int tomo_dot(tomo_ctx* ctx, const float* a, const float* b, size_t n, double* out, void* stream);
/* y = L x, L = 7-point Laplacian 6x - sum of 6 neighbours, zero padding
 * (separable {-1,2,-1} passes accumulated: C/Convolution3D_CPU_GPU.cu:146-148,
 *  C/GeneralizedGaussianRegularizer_CPU.cu:158-226).                         */
int tomo_laplacian(tomo_ctx* ctx, const tomo_geom* geom, const float* x, float* y, void* stream);

/* ---- simple-gradient reconstruction (Iter3D::doSimpleGradient) ----------- */
#define TOMO_PAIR_RSVI      0
#define TOMO_PAIR_SIDDONVI  1

typedef struct tomo_gd_params {
    int32_t pair;                    /* TOMO_PAIR_*                          */
    int32_t flags;                   /* operator flags (FOV mode, compat)    */
    int32_t gradient_iteration_nb;   /* gradientIterationNb                   */
    int32_t optimal_step_iteration_nb; /* optimalStepIterationNb             */
    double  lambda;                  /* regularizationParameter (GG, beta=2) */
    int32_t positivity;              /* positivity                           */
    double  alpha_min_init;          /* -0.0347711 in the reference          */
    int32_t storage;                 /* element type the operators read
                                        (the reference's --precision half,
                                        C/Iter3D_half.cu:545-700): TOMO_F32
                                        (0), or TOMO_F16 / TOMO_BF16: every
                                        projection and backprojection reads a
                                        16-bit copy of its FP32 input (FP16
                                        scaled by the power of two that puts
                                        max|x| in [2^13, 2^14), BF16
                                        unscaled) and accumulates in FP32;
                                        the iterate, gradient, residual and
                                        norms stay FP32 / FP64.  Served
                                        requests take tomo_serve_request.
                                        storage instead (this field must be
                                        0 there).                            */
} tomo_gd_params;

typedef struct tomo_gd_iter_stat {
    double residual_sq;              /* sum (y - P x)^2 before the update
                                        (J_MC of Iteration.dat)              */
    double norm_dj;                  /* sum dJ^2 (0 when the step is fixed)   */
    double norm_dj_proj;             /* 2 sum (P dJ)^2 + lambda normdJProjReg */
    double alpha;                    /* step used                            */
    double j_reg;                    /* sum (L x)^2 before the update when
                                        lambda != 0, else 0 (JReg of the GG
                                        beta=2 regulariser, C/Generalized
                                        GaussianRegularizer_CPU.cu:178); 0 in
                                        served requests                      */
    double beta;                     /* conjugate gradient: |g|^2/|g_prev|^2
                                        (0 at the first iteration and in the
                                        gradient loops)                      */
} tomo_gd_iter_stat;

/* Runs the loop on device-resident data (or host data with TOMO_HOST_PTR in
 * params->flags).  `volume` holds the initial estimate on entry and the
 * result on exit.  stats (may be NULL) receives gradient_iteration_nb
 * records.  Synchronous (the step rule needs the norms on the host).        */
int tomo_gd_run(tomo_ctx* ctx, const tomo_geom* geom, const tomo_gd_params* params,
                const float* sinogram, float* volume, tomo_gd_iter_stat* stats,
                void* stream);

/* Weighted gradient (Iter3D::doWeightedGradient, C/Iter3D_Algo_Gradient.cu:
 * 227-371; the reference's dispatcher, C/Iter3D.cu:78-90, never selects it):
 * the loop of tomo_gd_run with the residual multiplied by the weight
 * sinogram (same layout as the sinogram; host memory with TOMO_FLAG_HOST_PTR)
 * before the backprojection.  residual_sq and the step's |P dJ|^2 stay
 * unweighted, as in the reference, whose alpha_min starts at -999999999999
 * here (pass it in alpha_min_init).  Synchronous.                          */
int tomo_wgd_run(tomo_ctx* ctx, const tomo_geom* geom, const tomo_gd_params* params,
                 const float* sinogram, const float* weight, float* volume,
                 tomo_gd_iter_stat* stats, void* stream);

/* Conjugate gradient on the same criterion and step rule (the configIteration
 * key "gradient Conjugate"; reference: Iter3D::doConjugateGradient,
 * C/Iter3D_Algo_Gradient.cu:375-559, whose direction starts at 0 and whose
 * beta divides by that zero norm at iteration 2, so the method is defined
 * here):  g = the loop's dJ;  d = g, then d = beta d + g with
 * beta = |g|^2/|g_prev|^2 (Fletcher-Reeves);  alpha = -(g.d) /
 * (2|P d|^2 + lambda 2|L d|^2);  x += alpha d.  optimal_step_iteration_nb and
 * alpha_min_init are not used; stats hold |g|^2 in norm_dj.  Synchronous.    */
int tomo_cg_run(tomo_ctx* ctx, const tomo_geom* geom, const tomo_gd_params* params,
                const float* sinogram, float* volume, tomo_gd_iter_stat* stats,
                void* stream);

/* ---- SIRT ----------------------------------------------------------------- */
/* No counterpart in the reference; defined here for batched benchmarking:
 *   R = 1 / (P 1) on the sinogram, C = relaxation / (B 1) on the volume (0
 *   where the sum is not positive), x0 = 0, and per iteration
 *   r = (y - P x) * R;  x = x + C * (B r)   (clamped at 0 with positivity)
 * with P, B the projector and backprojector of the pair; FP32 element-wise
 * steps without fused multiply-adds.                                        */
typedef struct tomo_sirt_params {
    int32_t pair;                    /* TOMO_PAIR_*                          */
    int32_t flags;                   /* operator flags; TOMO_FLAG_HOST_PTR   */
    int32_t iterations;
    int32_t positivity;
    double  relaxation;              /* 1.0 for plain SIRT                    */
} tomo_sirt_params;

/* volume receives the result (its input contents are ignored); residual_sq
 * (may be NULL) receives sum (y - P x)^2 before each update.  Synchronous.  */
int tomo_sirt_run(tomo_ctx* ctx, const tomo_geom* geom, const tomo_sirt_params* params,
                  const float* sinogram, float* volume, double* residual_sq, void* stream);

/* ---- serving loop: continuous batching of reconstructions ---------------- */
/* No counterpart in the reference (it runs one reconstruction per process).
 * A server owns a worker thread and a stream on the context's device.  Each
 * request is one simple-gradient reconstruction (the loop of tomo_gd_run,
 * zero initial volume) of its own geometry and size.  Requests are admitted
 * in arrival order while the device memory they need fits (free memory or a
 * fixed budget) and a slot is free; admitted requests keep volumes, sinograms, geometry and loop
 * scalars resident on the device.  The server advances every admitted request
 * by one or more iterations per round (more for cheaper requests), each
 * operator of an iteration being one shared persistent launch over all of
 * them (optionally replayed from captured hipGraphs); requests join and leave
 * between rounds.  A request's arithmetic
 * does not depend on what shares its launches, so its result is the same
 * alone and in any mix.  A request is a simple-gradient reconstruction or a
 * SIRT reconstruction (tomo_sirt_run).  Its state is FP32; the copies its
 * operators read are FP32 or BF16 (tomo_serve_request.storage).              */
typedef struct tomo_server tomo_server;

#define TOMO_SCHEDULE_DIRECT     0 /* the operations of tomo_gd_run          */
#define TOMO_SCHEDULE_RECURRENT  1 /* P x and L(L x) carried by linearity:
                                      P x_{k+1} = P x_k + alpha P dJ_k, one
                                      projection per optimal-step iteration
                                      fewer (direct form where positivity is
                                      on or the previous step was fixed)      */

typedef struct tomo_serve_params {
    uint64_t hbm_budget_bytes;  /* device bytes admitted requests may hold;
                                   0 = live: a request is admitted while its
                                   footprint fits the device's free memory
                                   (hipMemGetInfo at admission, plus memory
                                   the server's pool holds unused) less 1/16
                                   of the device                            */
    int32_t  max_active;        /* requests resident at once, 1..64         */
    int32_t  use_graph;         /* replay each round from a captured graph  */
    int32_t  schedule;          /* TOMO_SCHEDULE_*                          */
    int32_t  sub_rounds;        /* iterations per round of the cheapest
                                   requests, 1..8 (0 = 8): each round gives a
                                   request about (work of the largest admitted
                                   request) / (its work) iterations, so small
                                   requests finish early; 1 = one iteration
                                   per request per round                     */
} tomo_serve_params;

#define TOMO_ALGO_GD    0
#define TOMO_ALGO_SIRT  1

typedef struct tomo_serve_request {
    tomo_geom_desc geom;
    tomo_gd_params gd;          /* pair, flags, iterations, Kopt, lambda,
                                   positivity, alpha_min_init; the flag
                                   TOMO_FLAG_HOST_PTR makes both pointers host
                                   pointers.  SIRT uses pair, flags,
                                   gradient_iteration_nb and positivity.     */
    const float* sinogram;      /* measured projections; read at admission  */
    float*       volume;        /* result, written when the request is done */
    int32_t      algorithm;     /* TOMO_ALGO_*                              */
    int32_t      storage;       /* element type of the operator copies (the
                                   z-pair copies the projectors read):
                                   TOMO_F32 (0) or TOMO_BF16, rounded from
                                   the FP32 state each time they are made;
                                   VI reads the FP32 residual, and the
                                   state, the sums and the result stay FP32 */
    double       relaxation;    /* SIRT relaxation (1.0 for plain SIRT)     */
} tomo_serve_request;

#define TOMO_REQ_QUEUED   0
#define TOMO_REQ_RUNNING  1
#define TOMO_REQ_DONE     2
#define TOMO_REQ_FAILED   3

typedef struct tomo_serve_result {
    int32_t  state;             /* TOMO_REQ_*                               */
    int32_t  status;            /* TOMO_OK or the failure's status          */
    double   submit_s;          /* server clock (s since creation)          */
    double   admit_s;
    double   done_s;
    double   latency_s;         /* done - submit                            */
    double   residual_sq;       /* last iteration's sum (y - P x)^2         */
    uint64_t device_bytes;      /* device memory held while admitted        */
} tomo_serve_result;

typedef struct tomo_serve_stats {
    uint64_t submitted, completed, failed;
    uint64_t rounds;            /* rounds run                               */
    int32_t  active, queued, peak_active, reserved;
    uint64_t bytes_in_use, peak_bytes, budget_bytes;
    double   latency_p50_s, latency_p99_s, latency_mean_s, latency_max_s;
    double   first_submit_s, last_done_s;
    double   throughput_rps;    /* completed / (last_done - first_submit)   */
} tomo_serve_stats;

/* Creation captures the server's launch sequences as graphs on its stream;
 * other threads must not synchronise the whole device during this call.     */
int tomo_serve_create(tomo_ctx* ctx, const tomo_serve_params* params, tomo_server** out);
/* Queues a request; the pointers must stay valid until it is done.  The
 * server remembers the latest 65536 requests for poll/wait/iter_stats; older
 * finished ones are forgotten (TOMO_ERR_INVALID_ARG "unknown request id").  */
int tomo_serve_submit(tomo_server* srv, const tomo_serve_request* req, uint64_t* id);
int tomo_serve_poll(tomo_server* srv, uint64_t id, tomo_serve_result* res);
/* Waits until the request is done or failed (TOMO_ERR_BUSY after timeout_s
 * seconds; a negative timeout waits without limit).                          */
int tomo_serve_wait(tomo_server* srv, uint64_t id, double timeout_s, tomo_serve_result* res);
/* Per-iteration records of a finished request (up to n).                    */
int tomo_serve_iter_stats(tomo_server* srv, uint64_t id, tomo_gd_iter_stat* stats, int n);
int tomo_serve_get_stats(tomo_server* srv, tomo_serve_stats* st);
/* Finishes admitted and queued requests, then stops the worker.             */
int tomo_serve_destroy(tomo_server* srv);

/* ---- phantoms (test and benchmark inputs) -------------------------------- */
/* 3D Shepp-Logan, modified-SL geometry with densities 2,-.98,-.02,-.02,.01x6
 * (Data3D_0256/phantom3d.m:175-200), voxel centres mapped to [-1,1].         */
int tomo_phantom_shepp_logan(tomo_ctx* ctx, const tomo_geom* geom, float* volume, void* stream);
/* Sum of `count` random ellipsoids (seeded): centres uniform in the inner 80%
 * of the volume, semi-axes uniform in [amin,amax] voxels, uniformly random
 * rotations, densities U[0,1).                                               */
int tomo_phantom_ellipsoids(tomo_ctx* ctx, const tomo_geom* geom, int count, uint64_t seed,
                            double amin, double amax, float* volume, void* stream);

/* ---- raw file I/O (C/Image3D.cu:140-233, C/Sinogram3D.cu:138-184) ---------- */
/* Headerless native-endian dumps of `count` elements of `dtype`.             */
int tomo_file_read(const char* path, void* host_dst, int dtype, size_t count);
int tomo_file_write(const char* path, const void* host_src, int dtype, size_t count);

#ifdef __cplusplus
}
#endif
#endif /* TOMOGPI_H */
