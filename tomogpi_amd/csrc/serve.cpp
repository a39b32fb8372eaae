// serve.cpp - the serving loop: continuous batching of simple-gradient
// reconstructions of uneven sizes on one device.
//
// A worker thread owns one stream.  Requests wait in arrival order; the head
// request is admitted when its device footprint fits the free device memory
// (or a fixed budget) and a slot is free (no reordering, so a large request
// is never overtaken indefinitely).  Admission carves one stream-ordered
// allocation into the request's resident state (view tables, volume x,
// gradient dJ, Laplacian work volumes, z-pair copies, sinograms y / P x /
// residual / P dJ, the VI box table, FP64 partial sums, loop scalars) and
// stages its inputs.
//
// A round advances every admitted request by one or more iterations of the
// loop of tomo_gd_run (C/Iter3D_Algo_Gradient.cu:25-223 with the beta=2
// regulariser, C/GeneralizedGaussianRegularizer_CPU.cu:158-226) or of
// tomo_sirt_run; cheaper requests get more iterations per round.  Each
// operator of an iteration is one shared persistent launch over all requests
// that take part in it; the launch reads its problem table from device
// memory, so the host writes the tables (one copy per round) and replays the
// same launch sequences, captured once as hipGraphs, every round.  Norms and the step size stay on the
// device (FP64 per-tile partials summed in a fixed order, then a step
// kernel), so a round needs one host synchronisation, at its end, where
// finished requests leave and queued ones join.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.hpp"

namespace tomo {
namespace {

constexpr size_t kAlign = 256;
constexpr int kMaxSubRounds = 8;
constexpr size_t kMaxRecords = 65536;  // finished requests remembered for poll/wait

uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Operators of one round, in launch order.  Each has its own table.
enum Op {
    OP_ZPX_X, OP_ZPY_X, OP_SID_X, OP_RS_X,  // P x (direct form)
    OP_LAP_X, OP_LAP_LX,                    // t1 = L x, t2 = L t1 (direct form)
    OP_AXPY_PX, OP_AXPY_LLX,                // recurrent form: px += alpha pd, t2 += alpha t3
    OP_DIFF,                                // r = y - px, partials of r^2
    OP_FIN_RES,
    OP_VI,                                  // dJ = -2 B r
    OP_AXPY_REG,                            // dJ += 2 lambda t2
    OP_LAP_DJ, OP_FIN_REG,                  // t1 = L dJ, sum t1^2
    OP_LAP_LDJ,                             // recurrent form: t3 = L t1
    OP_ZPX_D, OP_ZPY_D, OP_SID_D, OP_RS_D,  // pd = P dJ
    OP_SQ_DJ, OP_SQ_PD, OP_FIN_DJ, OP_FIN_PD,
    OP_STEP, OP_UPDATE,                     // alpha; x += alpha dJ
    OP_COUNT
};

struct Request {
    uint64_t id = 0;
    tomo_serve_request req{};
    tomo_geom geom;
    bool host = false;
    int state = TOMO_REQ_QUEUED;
    int status = TOMO_OK;
    std::string error;
    double submit_s = 0, admit_s = 0, done_s = 0;
    uint64_t bytes = 0;
    // resident state (valid while admitted)
    void* block = nullptr;
    float *x = nullptr, *dj = nullptr, *t1 = nullptr, *t2 = nullptr, *t3 = nullptr;
    float *qx = nullptr, *qy = nullptr, *y = nullptr, *px = nullptr, *r = nullptr, *pd = nullptr, *tq = nullptr;
    double *part_res = nullptr, *part_vol = nullptr, *part_reg = nullptr, *part_pd = nullptr;
    double *slot = nullptr, *hist = nullptr;
    bool bf16 = false;      // BF16 operator copies
    int k = 0;              // iterations done
    bool prev_opt = false;  // the previous iteration computed P dJ (and L(L dJ))
    std::vector<double> hist_host;
    double residual_sq = 0;
};

// Byte layout of a request's block.
struct Layout {
    size_t off[24];
    size_t total;
};

}  // namespace
}  // namespace tomo

struct tomo_server {
    tomo_ctx* ctx = nullptr;
    tomo_serve_params params{};
    hipStream_t stream = nullptr;
    hipMemPool_t pool = nullptr;           // request memory (kept cached between requests)
    tomo::BatchTable* d_tables = nullptr;  // kMaxSubRounds sets of OP_COUNT tables
    tomo::BatchTable* h_tables = nullptr;  // pinned mirror
    hipGraphExec_t graph[8] = {};          // one captured round per table set
    int sub_rounds = 8;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    bool broken = false;
    std::string broken_msg;
    uint64_t next_id = 1;
    std::deque<std::shared_ptr<tomo::Request>> queue;
    int n_active = 0;  // admitted requests (the worker owns the list)
    std::map<uint64_t, std::shared_ptr<tomo::Request>> all;
    std::vector<double> latencies;
    uint64_t submitted = 0, completed = 0, failed = 0, rounds = 0;
    int peak_active = 0;
    uint64_t bytes_in_use = 0, peak_bytes = 0, budget = 0;
    double first_submit = -1, last_done = 0;
    std::chrono::steady_clock::time_point t0;
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace tomo {
namespace {

uint32_t vec_tiles(uint64_t n) { return cdiv(n, kVecTile); }
uint32_t lap_tiles(const DevGeom& d) { return cdiv(d.nx, 64) * cdiv(d.ny, 4) * (uint32_t)d.nz; }

// Whether a request's operator copies are A-pair copies: BF16 copies (ces 2)
// of an RS request whose A-pair copies take the buffer path (one load per
// sample); otherwise z-pair copies (FP32 or BF16).
bool apair_copies(const DevGeom& d, const tomo_gd_params& gd, size_t ces) {
    return ces == 2 && gd.pair != TOMO_PAIR_SIDDONVI && apair_buffer_path(d.nx, d.ny, d.nz);
}

Layout layout_of(const tomo_geom& g, const tomo_gd_params& gd, int algo, size_t ces) {
    const DevGeom& d = g.dev;
    const size_t nv = g.nvox(), ns = g.nsino();
    const bool lam = gd.lambda != 0.0 && algo == TOMO_ALGO_GD;
    const bool sirt = algo == TOMO_ALGO_SIRT;
    size_t sizes[24] = {0};
    int i = 0;
    sizes[i++] = (size_t)d.nphi * kViewStride * sizeof(double);          // 0 view
    sizes[i++] = (size_t)d.nphi * 2 * sizeof(float);                     // 1 view_f
    sizes[i++] = nv * 4;                                                 // 2 x
    sizes[i++] = nv * 4;                                                 // 3 dj
    sizes[i++] = (lam || sirt) ? nv * 4 : 0;                             // 4 t1 (SIRT: volume weight C)
    sizes[i++] = lam ? nv * 4 : 0;                                       // 5 t2
    sizes[i++] = lam ? nv * 4 : 0;                                       // 6 t3
    const bool ap = apair_copies(d, gd, ces);
    sizes[i++] = (ap ? apair_elems_x(d.nx, d.ny, d.nz) : zpair_elems_x(d.nx, d.ny, d.nz)) * ces;  // 7 qx
    sizes[i++] = (ap ? apair_elems_y(d.nx, d.ny, d.nz) : zpair_elems_y(d.nx, d.ny, d.nz)) * ces;  // 8 qy
    sizes[i++] = ns * 4;                                                 // 9 y
    sizes[i++] = ns * 4;                                                 // 10 px
    sizes[i++] = ns * 4;                                                 // 11 r
    sizes[i++] = ns * 4;                                                 // 12 pd (SIRT: sinogram weight R)
    sizes[i++] = vi_box_count(d) * sizeof(ViBox);                        // 13 tq (VI box table, set at admission)
    sizes[i++] = (size_t)vec_tiles(ns) * 8;                              // 14 part_res
    sizes[i++] = (size_t)vec_tiles(nv) * 8;                              // 15 part_vol
    sizes[i++] = lam ? (size_t)lap_tiles(d) * 8 : 0;                     // 16 part_reg
    sizes[i++] = (size_t)vec_tiles(ns) * 8;                              // 17 part_pd
    sizes[i++] = kSlotCount * 8;                                         // 18 slot
    sizes[i++] = (size_t)std::max(gd.gradient_iteration_nb, 1) * 4 * 8;  // 19 hist
    Layout L{};
    size_t o = 0;
    for (int j = 0; j < 24; ++j) {
        L.off[j] = o;
        o += (sizes[j] + kAlign - 1) / kAlign * kAlign;
    }
    L.total = o;
    return L;
}

// Runs on the worker without the lock.  Returns TOMO_OK or a failure (the
// request is then marked failed by the caller).
int admit(tomo_server* srv, Request& q) {
    q.bf16 = q.req.storage == TOMO_BF16;
    const Layout L = layout_of(q.geom, q.req.gd, q.req.algorithm, q.bf16 ? 2 : 4);
    TOMO_HIP(hipMallocFromPoolAsync(&q.block, L.total, srv->pool, srv->stream));
    char* b = (char*)q.block;
    auto at = [&](int j) { return (void*)(b + L.off[j]); };
    q.geom.d_view = (double*)at(0);
    q.geom.d_view_f = (float*)at(1);
    q.geom.dev.view = q.geom.d_view;
    q.geom.dev.view_f = q.geom.d_view_f;
    const bool sirt = q.req.algorithm == TOMO_ALGO_SIRT;
    const bool lam = q.req.gd.lambda != 0.0 && !sirt;
    q.x = (float*)at(2);
    q.dj = (float*)at(3);
    q.t1 = (lam || sirt) ? (float*)at(4) : nullptr;
    q.t2 = lam ? (float*)at(5) : nullptr;
    q.t3 = lam ? (float*)at(6) : nullptr;
    q.qx = (float*)at(7);
    q.qy = (float*)at(8);
    q.y = (float*)at(9);
    q.px = (float*)at(10);
    q.r = (float*)at(11);
    q.pd = (float*)at(12);
    q.tq = (float*)at(13);
    q.part_res = (double*)at(14);
    q.part_vol = (double*)at(15);
    q.part_reg = lam ? (double*)at(16) : nullptr;
    q.part_pd = (double*)at(17);
    q.slot = (double*)at(18);
    q.hist = (double*)at(19);
    const size_t ns = q.geom.nsino(), nv = q.geom.nvox();
    hipStream_t s = srv->stream;
    TOMO_HIP(hipMemcpyAsync(q.geom.d_view, q.geom.view.data(), q.geom.view.size() * sizeof(double),
                            hipMemcpyHostToDevice, s));
    TOMO_HIP(hipMemcpyAsync(q.geom.d_view_f, q.geom.view_f.data(), q.geom.view_f.size() * sizeof(float),
                            hipMemcpyHostToDevice, s));
    TOMO_HIP(hipMemcpyAsync(q.y, q.req.sinogram, ns * 4, q.host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
    int rc;
    // the VI box table of the geometry, read by every round's backprojection
    if ((rc = launch_vi_boxes(q.geom.dev, q.tq, s))) return rc;
    q.geom.d_vi_boxes = q.tq;  // the single-problem VI below uses it too
    if (sirt) {
        // R = 1/(P 1) into pd, C = relaxation/(B 1) into t1, with the
        // single-problem operators (the same arithmetic as tomo_sirt_run).
        const Buf ones_v{q.x, TOMO_F32, 1.0f};
        const MutBuf pd{q.pd, TOMO_F32, 1.0f};
        if ((rc = launch_fill(q.x, 1.0f, nv, s))) return rc;
        if (q.req.gd.pair == TOMO_PAIR_SIDDONVI) rc = launch_siddon(&q.geom, 0, ones_v, pd, s);
        else rc = launch_rs(&q.geom, (q.req.gd.flags & TOMO_FLAG_RS_FOV_CUBE) ? 1 : 0, ones_v, pd, s);
        if (rc) return rc;
        if ((rc = launch_recip_weight(q.pd, 1.0f, q.pd, ns, s))) return rc;
        if ((rc = launch_fill(q.r, 1.0f, ns, s))) return rc;
        if ((rc = launch_vi(&q.geom, Buf{q.r, TOMO_F32, 1.0f}, MutBuf{q.t1, TOMO_F32, 1.0f}, s))) return rc;
        if ((rc = launch_recip_weight(q.t1, (float)q.req.relaxation, q.t1, nv, s))) return rc;
    }
    if ((rc = launch_fill(q.x, 0.0f, nv, s))) return rc;
    double slot0[kSlotCount] = {0};
    slot0[kSlotAlphaMin] = q.req.gd.alpha_min_init;
    TOMO_HIP(hipMemcpyAsync(q.slot, slot0, sizeof(slot0), hipMemcpyHostToDevice, s));
    // The host slot values are read by the copy before it returns for
    // pageable memory; synchronise so the stack buffer can go.
    TOMO_HIP(hipStreamSynchronize(s));
    q.bytes = L.total;
    q.k = 0;
    q.prev_opt = false;
    return TOMO_OK;
}

struct TableBuilder {
    BatchTable* t;
    void reset() {
        t->count = 0;
        t->total = 0;
        for (uint32_t& c : t->next) c = 0;
    }
    BatchItem& add(const DevGeom& g, uint32_t gx, uint32_t gy, uint32_t gz) {
        BatchItem& it = t->item[t->count++];
        std::memset(&it, 0, sizeof(it));
        it.g = g;
        it.begin = t->total;
        it.gx = gx;
        it.gy = gy;
        it.tiles = gx * gy * gz;
        t->total += it.tiles;
        return it;
    }
    BatchItem& add_vec(const DevGeom& g, uint64_t n) {
        BatchItem& it = add(g, vec_tiles(n), 1, 1);
        it.n = n;
        return it;
    }
    BatchItem& add_one(const DevGeom& g) { return add(g, 1, 1, 1); }
};

// Tables of table set `set` for one iteration of every request in `running`
// (advances each request's iteration count).
void build_tables(tomo_server* srv, int set, const std::vector<Request*>& running) {
    TableBuilder tb[OP_COUNT];
    for (int o = 0; o < OP_COUNT; ++o) {
        tb[o].t = srv->h_tables + (size_t)set * OP_COUNT + o;
        tb[o].reset();
    }
    const bool recurrent_server = srv->params.schedule == TOMO_SCHEDULE_RECURRENT;
    for (Request* qp : running) {
        Request& q = *qp;
        ++q.k;  // this set runs iteration q.k (1-based) below
        const DevGeom& d = q.geom.dev;
        const tomo_gd_params& gd = q.req.gd;
        const size_t nv = q.geom.nvox(), ns = q.geom.nsino();
        const int k = q.k;  // 1-based iteration of this set
        const uint32_t pgx = cdiv(d.nu, kProjTileU), pgz = cdiv(d.nv, kProjTileV);
        const bool sid = gd.pair == TOMO_PAIR_SIDDONVI;
        const int cube = (gd.flags & TOMO_FLAG_RS_FOV_CUBE) ? 1 : 0;
        const bool ap = apair_copies(d, gd, q.bf16 ? 2 : 4);
        const int32_t cflag = (q.bf16 ? kItemBF16 : 0) | (ap ? kItemApair : 0);  // operator copies
        auto add_projection = [&](const float* vol, float* out, int opx, int opy, int ops, int opr) {
            BatchItem& zx = ap ? tb[opx].add(d, cdiv(d.nx + 1, 32), cdiv(d.ny + 2, 32), d.nz + 1)
                               : tb[opx].add(d, cdiv(d.nx, 32), cdiv(d.ny + 2, 32), d.nz + 1);
            zx.in = vol;
            zx.out = q.qx;
            zx.flags = cflag;
            BatchItem& zy = tb[opy].add(d, 1, ap ? d.ny + 1 : d.ny, d.nz + 1);  // whole rows
            zy.in = vol;
            zy.out = q.qy;
            zy.flags = cflag;
            BatchItem& p = tb[sid ? ops : opr].add(d, pgx, d.nphi, pgz);
            p.in = q.qx;
            p.in2 = q.qy;
            p.out = out;
            p.flags = cube | cflag;
        };
        if (q.req.algorithm == TOMO_ALGO_SIRT) {
            add_projection(q.x, q.px, OP_ZPX_X, OP_ZPY_X, OP_SID_X, OP_RS_X);
            BatchItem& df = tb[OP_DIFF].add_vec(d, ns);
            df.in = q.y;
            df.in2 = q.px;
            df.in3 = q.pd;  // R
            df.out = q.r;
            df.aux = q.part_res;
            BatchItem& fr = tb[OP_FIN_RES].add_one(d);
            fr.aux = q.part_res;
            fr.n = vec_tiles(ns);
            fr.slot = q.slot;
            fr.flags = kSlotResidual;
            BatchItem& vi = tb[OP_VI].add(d, cdiv(d.nx, kViStageX), cdiv(d.ny, kViStageY), cdiv(d.nz, kViTileZ));
            vi.in = q.r;
            vi.aux = q.tq;
            vi.out = q.dj;
            vi.a = 1.0f;
            BatchItem& st = tb[OP_STEP].add_one(d);
            st.slot = q.slot;
            st.flags = 2;
            st.aux = q.hist;
            st.n = (uint64_t)(k - 1);
            BatchItem& up = tb[OP_UPDATE].add_vec(d, nv);
            up.in = q.dj;
            up.in3 = q.t1;  // C
            up.out = q.x;
            up.flags = gd.positivity ? 2 : 0;
            continue;
        }
        const bool opt = k < gd.optimal_step_iteration_nb;
        const bool lam = gd.lambda != 0.0;
        // Recurrent form: P x (and L(L x)) from the previous iteration's
        // P dJ (and L(L dJ)); at k = 1 x = 0, so P x = 0 and L(L x) = 0.
        const bool rec = recurrent_server && !gd.positivity && (k == 1 || q.prev_opt);
        const bool rec_zero = rec && k == 1;
        if (!rec) {
            add_projection(q.x, q.px, OP_ZPX_X, OP_ZPY_X, OP_SID_X, OP_RS_X);
            if (lam) {
                BatchItem& l1 = tb[OP_LAP_X].add(d, cdiv(d.nx, 64), cdiv(d.ny, 4), d.nz);
                l1.in = q.x;
                l1.out = q.t1;
                BatchItem& l2 = tb[OP_LAP_LX].add(d, cdiv(d.nx, 64), cdiv(d.ny, 4), d.nz);
                l2.in = q.t1;
                l2.out = q.t2;
            }
        } else if (rec_zero) {
            // px = 0 and t2 = 0: axpy with a zero factor on themselves would
            // read uninitialised memory; fill instead.
            BatchItem& f = tb[OP_AXPY_PX].add_vec(d, ns);
            f.out = q.px;
            f.in = nullptr;
            f.flags = 4;  // fill with a
            f.a = 0.0f;
            if (lam) {
                BatchItem& f2 = tb[OP_AXPY_LLX].add_vec(d, nv);
                f2.out = q.t2;
                f2.flags = 4;
                f2.a = 0.0f;
            }
        } else {
            BatchItem& a1 = tb[OP_AXPY_PX].add_vec(d, ns);
            a1.in = q.pd;
            a1.out = q.px;
            a1.slot = q.slot;
            a1.flags = 1;
            if (lam) {
                BatchItem& a2 = tb[OP_AXPY_LLX].add_vec(d, nv);
                a2.in = q.t3;
                a2.out = q.t2;
                a2.slot = q.slot;
                a2.flags = 1;
            }
        }
        BatchItem& df = tb[OP_DIFF].add_vec(d, ns);
        df.in = q.y;
        df.in2 = q.px;
        df.out = q.r;
        df.aux = q.part_res;
        BatchItem& fr = tb[OP_FIN_RES].add_one(d);
        fr.aux = q.part_res;
        fr.n = vec_tiles(ns);
        fr.slot = q.slot;
        fr.flags = kSlotResidual;
        BatchItem& vi = tb[OP_VI].add(d, cdiv(d.nx, kViStageX), cdiv(d.ny, kViStageY), cdiv(d.nz, kViTileZ));
        vi.in = q.r;
        vi.aux = q.tq;
        vi.out = q.dj;
        vi.a = -2.0f;
        if (lam) {
            BatchItem& ar = tb[OP_AXPY_REG].add_vec(d, nv);
            ar.in = q.t2;
            ar.out = q.dj;
            ar.a = (float)(2.0 * (double)(float)gd.lambda);
            if (opt) {
                BatchItem& ld = tb[OP_LAP_DJ].add(d, cdiv(d.nx, 64), cdiv(d.ny, 4), d.nz);
                ld.in = q.dj;
                ld.out = q.t1;
                ld.aux = q.part_reg;
                BatchItem& fg = tb[OP_FIN_REG].add_one(d);
                fg.aux = q.part_reg;
                fg.n = lap_tiles(d);
                fg.slot = q.slot;
                fg.flags = kSlotRegSq;
                if (recurrent_server && !gd.positivity) {
                    BatchItem& l3 = tb[OP_LAP_LDJ].add(d, cdiv(d.nx, 64), cdiv(d.ny, 4), d.nz);
                    l3.in = q.t1;
                    l3.out = q.t3;
                }
            }
        }
        if (opt) {
            add_projection(q.dj, q.pd, OP_ZPX_D, OP_ZPY_D, OP_SID_D, OP_RS_D);
            BatchItem& s1 = tb[OP_SQ_DJ].add_vec(d, nv);
            s1.in = q.dj;
            s1.aux = q.part_vol;
            BatchItem& s2 = tb[OP_SQ_PD].add_vec(d, ns);
            s2.in = q.pd;
            s2.aux = q.part_pd;
            BatchItem& f1 = tb[OP_FIN_DJ].add_one(d);
            f1.aux = q.part_vol;
            f1.n = vec_tiles(nv);
            f1.slot = q.slot;
            f1.flags = kSlotNormDj;
            BatchItem& f2 = tb[OP_FIN_PD].add_one(d);
            f2.aux = q.part_pd;
            f2.n = vec_tiles(ns);
            f2.slot = q.slot;
            f2.flags = kSlotProjSq;
        }
        BatchItem& st = tb[OP_STEP].add_one(d);
        st.slot = q.slot;
        st.d = lam ? gd.lambda : 0.0;
        st.flags = opt ? 1 : 0;
        st.aux = q.hist;
        st.n = (uint64_t)(k - 1);
        BatchItem& up = tb[OP_UPDATE].add_vec(d, nv);
        up.in = q.dj;
        up.out = q.x;
        up.slot = q.slot;
        up.flags = 1 | (gd.positivity ? 2 : 0);
        q.prev_opt = opt;
    }
}

// The fixed launch sequence of one iteration (every table, empty or not).
int launch_round(tomo_server* srv, int set) {
    BatchTable* T = srv->d_tables + (size_t)set * OP_COUNT;
    hipStream_t s = srv->stream;
    hipError_t e = hipSuccess;
#define TOMO_STEP(call) \
    if ((e = (call)) != hipSuccess) return hip_fail(e, #call)
    TOMO_STEP(launch_batch_zpair_x(T + OP_ZPX_X, s));
    TOMO_STEP(launch_batch_zpair_y(T + OP_ZPY_X, s));
    TOMO_STEP(launch_batch_siddon(T + OP_SID_X, s));
    TOMO_STEP(launch_batch_rs(T + OP_RS_X, s));
    TOMO_STEP(launch_batch_laplacian(T + OP_LAP_X, s));
    TOMO_STEP(launch_batch_laplacian(T + OP_LAP_LX, s));
    TOMO_STEP(launch_batch_axpy(T + OP_AXPY_PX, s));
    TOMO_STEP(launch_batch_axpy(T + OP_AXPY_LLX, s));
    TOMO_STEP(launch_batch_diff_sq(T + OP_DIFF, s));
    TOMO_STEP(launch_batch_finish(T + OP_FIN_RES, s));
    TOMO_STEP(launch_batch_vi(T + OP_VI, s));
    TOMO_STEP(launch_batch_axpy(T + OP_AXPY_REG, s));
    TOMO_STEP(launch_batch_laplacian(T + OP_LAP_DJ, s));
    TOMO_STEP(launch_batch_finish(T + OP_FIN_REG, s));
    TOMO_STEP(launch_batch_laplacian(T + OP_LAP_LDJ, s));
    TOMO_STEP(launch_batch_zpair_x(T + OP_ZPX_D, s));
    TOMO_STEP(launch_batch_zpair_y(T + OP_ZPY_D, s));
    TOMO_STEP(launch_batch_siddon(T + OP_SID_D, s));
    TOMO_STEP(launch_batch_rs(T + OP_RS_D, s));
    TOMO_STEP(launch_batch_sq(T + OP_SQ_DJ, s));
    TOMO_STEP(launch_batch_sq(T + OP_SQ_PD, s));
    TOMO_STEP(launch_batch_finish(T + OP_FIN_DJ, s));
    TOMO_STEP(launch_batch_finish(T + OP_FIN_PD, s));
    TOMO_STEP(launch_batch_step(T + OP_STEP, s));
    TOMO_STEP(launch_batch_axpy(T + OP_UPDATE, s));
#undef TOMO_STEP
    return TOMO_OK;
}

// One round: request q runs m_q iterations, m_q = W_max / W_q clamped to
// [1, sub_rounds] and to its remaining iterations, with W = n^3 Nphi (the
// operators' work) and W_max over the admitted requests; iteration j of the
// round is table set j, whose launches include the requests with m_q > j.
// Cheap requests thus finish in few rounds beside expensive ones, and every
// request's arithmetic is unchanged.
int run_round(tomo_server* srv, const std::vector<std::shared_ptr<Request>>& running) {
    auto work = [](const Request& q) { return (double)q.geom.nvox() * (double)q.geom.info.nphi; };
    double wmax = 0.0;
    for (auto& q : running)
        if (q->k < q->req.gd.gradient_iteration_nb) wmax = std::max(wmax, work(*q));
    std::vector<int> m(running.size(), 0);
    int sets = 0;
    for (size_t i = 0; i < running.size(); ++i) {
        const Request& q = *running[i];
        const int left = q.req.gd.gradient_iteration_nb - q.k;
        if (left <= 0) continue;
        const double ratio = wmax / work(q);
        const int want = ratio >= (double)srv->sub_rounds ? srv->sub_rounds : std::max(1, (int)ratio);
        m[i] = std::min(want, left);
        sets = std::max(sets, m[i]);
    }
    std::vector<Request*> part;
    for (int j = 0; j < sets; ++j) {
        part.clear();
        for (size_t i = 0; i < running.size(); ++i)
            if (m[i] > j) part.push_back(running[i].get());
        build_tables(srv, j, part);
    }
    hipStream_t s = srv->stream;
    TOMO_HIP(hipMemcpyAsync(srv->d_tables, srv->h_tables, sizeof(BatchTable) * OP_COUNT * (size_t)sets,
                            hipMemcpyHostToDevice, s));
    for (int j = 0; j < sets; ++j) {
        if (srv->params.use_graph) {
            TOMO_HIP(hipGraphLaunch(srv->graph[j], s));
        } else {
            const int rc = launch_round(srv, j);
            if (rc) return rc;
        }
    }
    TOMO_HIP(hipStreamSynchronize(s));
    return TOMO_OK;
}

// Captures the launch sequence of every table set as a graph.  Done once, by
// the creating thread before the worker starts: a capture in progress makes
// device-wide synchronisation by other threads fail, so none happens while
// requests are served.
int capture_graphs(tomo_server* srv) {
    hipStream_t s = srv->stream;
    for (int j = 0; j < kMaxSubRounds; ++j) {
        TOMO_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = launch_round(srv, j);
        hipGraph_t captured = nullptr;
        const hipError_t ee = hipStreamEndCapture(s, &captured);
        if (rc) {
            if (captured) (void)hipGraphDestroy(captured);
            return rc;
        }
        TOMO_HIP(ee);
        const hipError_t ie = hipGraphInstantiate(&srv->graph[j], captured, nullptr, nullptr, 0);
        (void)hipGraphDestroy(captured);
        TOMO_HIP(ie);
    }
    return TOMO_OK;
}

// Copies the result out and releases the request's memory (worker, no lock).
// The block is released whatever the copies return.
int retire(tomo_server* srv, Request& q) {
    hipStream_t s = srv->stream;
    const size_t nv = q.geom.nvox();
    hipError_t e = hipMemcpyAsync(q.req.volume, q.x, nv * 4, q.host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s);
    q.hist_host.assign((size_t)q.req.gd.gradient_iteration_nb * 4, 0.0);
    if (e == hipSuccess && !q.hist_host.empty())
        e = hipMemcpyAsync(q.hist_host.data(), q.hist, q.hist_host.size() * 8, hipMemcpyDeviceToHost, s);
    const hipError_t f = hipFreeAsync(q.block, s);
    const hipError_t y = hipStreamSynchronize(s);
    q.block = nullptr;
    TOMO_HIP(e);
    TOMO_HIP(f);
    TOMO_HIP(y);
    if (!q.hist_host.empty()) q.residual_sq = q.hist_host[q.hist_host.size() - 4];
    return TOMO_OK;
}

// Releases a request's block without copying anything out (failed rounds).
void drop_block(tomo_server* srv, Request& q) {
    if (!q.block) return;
    (void)hipFreeAsync(q.block, srv->stream);
    (void)hipStreamSynchronize(srv->stream);
    q.block = nullptr;
}

// Caller holds srv->mu.
void finish_request(tomo_server* srv, Request& q, int status, const std::string& msg) {
    q.done_s = srv->now();
    q.status = status;
    q.error = msg;
    q.state = status == TOMO_OK ? TOMO_REQ_DONE : TOMO_REQ_FAILED;
    if (status == TOMO_OK) {
        ++srv->completed;
        srv->latencies.push_back(q.done_s - q.submit_s);
        srv->last_done = q.done_s;
    } else {
        ++srv->failed;
    }
    if (q.bytes) {
        srv->bytes_in_use -= q.bytes;
        q.bytes = 0;
    }
}

// Bytes further requests may take now.  With a fixed budget: what the budget
// leaves.  Otherwise (live admission) the device's free memory, plus what the
// server's pool holds without handing it out (blocks of finished requests
// stay cached there), less a reserve of 1/16 of the device; the budget shown
// in the statistics follows it.  Caller holds srv->mu; the worker's device is
// current.
uint64_t admission_room(tomo_server* srv) {
    if (srv->params.hbm_budget_bytes)
        return srv->budget > srv->bytes_in_use ? srv->budget - srv->bytes_in_use : 0;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
    uint64_t reserved = 0, used = 0;
    if (hipMemPoolGetAttribute(srv->pool, hipMemPoolAttrReservedMemCurrent, &reserved) != hipSuccess ||
        hipMemPoolGetAttribute(srv->pool, hipMemPoolAttrUsedMemCurrent, &used) != hipSuccess)
        reserved = used = 0;
    const uint64_t avail = (uint64_t)free_b + (reserved > used ? reserved - used : 0);
    const uint64_t keep = (uint64_t)total_b / 16;
    const uint64_t room = avail > keep ? avail - keep : 0;
    srv->budget = srv->bytes_in_use + room;
    return room;
}

void worker_main(tomo_server* srv) {
    (void)hipSetDevice(srv->ctx->device);
    std::vector<std::shared_ptr<Request>> running;  // admitted, owned by this thread
    for (;;) {
        std::vector<std::shared_ptr<Request>> admitted;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            for (;;) {
                // Admission: the head of the queue while it fits the room
                // measured at the start of this pass, less what the pass has
                // admitted already (allocated after the lock is released).
                uint64_t room = srv->queue.empty() ? 0 : admission_room(srv);
                while (!srv->queue.empty() && !srv->broken) {
                    auto q = srv->queue.front();
                    const uint64_t need =
                        layout_of(q->geom, q->req.gd, q->req.algorithm, q->req.storage == TOMO_BF16 ? 2 : 4).total;
                    if ((int)(running.size() + admitted.size()) >= srv->params.max_active) break;
                    if (need > room) {
                        if (!running.empty() || !admitted.empty()) break;
                        // Larger than all the memory it could ever be given.
                        srv->queue.pop_front();
                        finish_request(srv, *q, TOMO_ERR_OUT_OF_MEMORY,
                                       srv->params.hbm_budget_bytes ? "request exceeds the server's memory budget"
                                                                    : "request exceeds the free device memory");
                        srv->cv_done.notify_all();
                        continue;
                    }
                    room -= need;
                    srv->queue.pop_front();
                    q->bytes = need;
                    srv->bytes_in_use += need;
                    srv->peak_bytes = std::max(srv->peak_bytes, srv->bytes_in_use);
                    admitted.push_back(q);
                }
                if (!running.empty() || !admitted.empty()) break;
                if (srv->stop && srv->queue.empty()) return;
                srv->cv_work.wait(lk);
            }
            srv->n_active = (int)(running.size() + admitted.size());
            srv->peak_active = std::max(srv->peak_active, srv->n_active);
        }
        // Device work runs without the lock.
        for (auto& q : admitted) {
            const int rc = admit(srv, *q);
            if (rc && q->block) {
                (void)hipFreeAsync(q->block, srv->stream);
                (void)hipStreamSynchronize(srv->stream);
                q->block = nullptr;
            }
            std::lock_guard<std::mutex> lk(srv->mu);
            if (rc) {
                finish_request(srv, *q, rc, "admission failed");
                --srv->n_active;
                srv->cv_done.notify_all();
            } else {
                q->state = TOMO_REQ_RUNNING;
                q->admit_s = srv->now();
                running.push_back(q);
            }
        }
        bool any = false;
        for (auto& q : running) any = any || q->k < q->req.gd.gradient_iteration_nb;
        if (any) {
            const int rc = run_round(srv, running);
            std::lock_guard<std::mutex> lk(srv->mu);
            ++srv->rounds;
            if (rc) {
                srv->broken = true;
                for (auto& q : running) {
                    drop_block(srv, *q);
                    finish_request(srv, *q, rc, "round failed");
                }
                running.clear();
                for (auto& q : srv->queue) finish_request(srv, *q, rc, "server failed");
                srv->queue.clear();
                srv->n_active = 0;
                srv->cv_done.notify_all();
                continue;
            }
        }
        // Finished requests leave: result copied out, memory released.
        for (size_t i = 0; i < running.size();) {
            Request& q = *running[i];
            if (q.k < q.req.gd.gradient_iteration_nb) {
                ++i;
                continue;
            }
            const int rc = retire(srv, q);
            std::lock_guard<std::mutex> lk(srv->mu);
            finish_request(srv, q, rc, rc ? "result copy failed" : "");
            --srv->n_active;
            running.erase(running.begin() + i);
            srv->cv_done.notify_all();
        }
    }
}

}  // namespace
}  // namespace tomo

using namespace tomo;

extern "C" int tomo_serve_create(tomo_ctx* ctx, const tomo_serve_params* params, tomo_server** out) {
    if (!ctx || !params || !out) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    if (params->max_active < 1 || params->max_active > kBatchMax)
        return fail(TOMO_ERR_INVALID_ARG, "max_active must be in 1..64");
    if (params->schedule != TOMO_SCHEDULE_DIRECT && params->schedule != TOMO_SCHEDULE_RECURRENT)
        return fail(TOMO_ERR_INVALID_ARG, "unknown schedule");
    if (params->sub_rounds < 0 || params->sub_rounds > kMaxSubRounds)
        return fail(TOMO_ERR_INVALID_ARG, "sub_rounds must be in 0..8");
    TOMO_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<tomo_server> srv(new tomo_server());
    srv->ctx = ctx;
    srv->params = *params;
    srv->sub_rounds = params->sub_rounds ? params->sub_rounds : kMaxSubRounds;
    srv->t0 = std::chrono::steady_clock::now();
    srv->budget = params->hbm_budget_bytes;  // 0: live admission (admission_room)
    // A pool of its own, whose memory stays cached between requests
    // (stream-ordered reuse without returning it to the device).
    hipMemPoolProps pp;
    std::memset(&pp, 0, sizeof(pp));
    pp.allocType = hipMemAllocationTypePinned;
    pp.handleTypes = hipMemHandleTypeNone;
    pp.location.type = hipMemLocationTypeDevice;
    pp.location.id = ctx->device;
    TOMO_HIP(hipMemPoolCreate(&srv->pool, &pp));
    uint64_t thr = UINT64_MAX;
    hipError_t e = hipMemPoolSetAttribute(srv->pool, hipMemPoolAttrReleaseThreshold, &thr);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&srv->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        (void)hipMemPoolDestroy(srv->pool);
        return hip_fail(e, "server pool / stream");
    }
    const size_t tbytes = sizeof(BatchTable) * OP_COUNT * kMaxSubRounds;
    e = hipMalloc((void**)&srv->d_tables, tbytes);
    if (e == hipSuccess) e = hipHostMalloc((void**)&srv->h_tables, tbytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        if (srv->d_tables) (void)hipFree(srv->d_tables);
        (void)hipStreamDestroy(srv->stream);
        (void)hipMemPoolDestroy(srv->pool);
        return hip_fail(e, "server tables");
    }
    std::memset(srv->h_tables, 0, tbytes);
    // Empty tables on the device, then the captured rounds.
    e = hipMemcpyAsync(srv->d_tables, srv->h_tables, tbytes, hipMemcpyHostToDevice, srv->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(srv->stream);
    int rc = e == hipSuccess ? TOMO_OK : hip_fail(e, "server tables");
    if (rc == TOMO_OK && params->use_graph) rc = capture_graphs(srv.get());
    if (rc != TOMO_OK) {
        for (auto& gx : srv->graph)
            if (gx) (void)hipGraphExecDestroy(gx);
        (void)hipFree(srv->d_tables);
        (void)hipHostFree(srv->h_tables);
        (void)hipStreamDestroy(srv->stream);
        (void)hipMemPoolDestroy(srv->pool);
        return rc;
    }
    tomo_server* p = srv.release();
    p->worker = std::thread(worker_main, p);
    *out = p;
    return TOMO_OK;
}

extern "C" int tomo_serve_submit(tomo_server* srv, const tomo_serve_request* req, uint64_t* id) {
    if (!srv || !req || !id) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    if (!req->sinogram || !req->volume) return fail(TOMO_ERR_INVALID_ARG, "null data pointer");
    const tomo_gd_params& gd = req->gd;
    if (gd.pair != TOMO_PAIR_RSVI && gd.pair != TOMO_PAIR_SIDDONVI) return fail(TOMO_ERR_INVALID_ARG, "unknown pair");
    if (gd.flags & TOMO_FLAG_SIDDON_COMPAT)
        return fail(TOMO_ERR_UNSUPPORTED, "the serving loop runs the clean Siddon form only");
    if (gd.gradient_iteration_nb < 0) return fail(TOMO_ERR_INVALID_ARG, "negative iteration count");
    if (req->algorithm != TOMO_ALGO_GD && req->algorithm != TOMO_ALGO_SIRT)
        return fail(TOMO_ERR_INVALID_ARG, "unknown algorithm");
    if (req->algorithm == TOMO_ALGO_SIRT && !(req->relaxation > 0.0))
        return fail(TOMO_ERR_INVALID_ARG, "SIRT relaxation must be positive");
    if (req->storage != TOMO_F32 && req->storage != TOMO_BF16)
        return fail(TOMO_ERR_UNSUPPORTED, "operator copies are FP32 or BF16");
    if (gd.storage != TOMO_F32)
        return fail(TOMO_ERR_INVALID_ARG, "served requests take tomo_serve_request.storage; gd.storage must be 0");
    auto q = std::make_shared<Request>();
    q->req = *req;
    q->host = (gd.flags & TOMO_FLAG_HOST_PTR) != 0;
    int rc = derive_geometry(req->geom, &q->geom);
    if (rc) return rc;
    const tomo_geom_info& gi = q->geom.info;
    if (zpair_elems_x(gi.nx, gi.ny, gi.nz) >= (1ull << 32) || zpair_elems_y(gi.nx, gi.ny, gi.nz) >= (1ull << 32))
        return fail(TOMO_ERR_UNSUPPORTED, "volume too large for the Siddon traversal index");
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->broken) return fail(TOMO_ERR_HIP, "server stopped after a device failure");
    if (srv->stop) return fail(TOMO_ERR_BUSY, "server is shutting down");
    q->id = srv->next_id++;
    q->submit_s = srv->now();
    if (srv->first_submit < 0) srv->first_submit = q->submit_s;
    ++srv->submitted;
    srv->queue.push_back(q);
    srv->all[q->id] = q;
    // Bounded history: past kMaxRecords, the oldest finished records go
    // (ids are increasing, so the map's order is submission order).
    for (auto it = srv->all.begin(); srv->all.size() > kMaxRecords && it != srv->all.end();) {
        if (it->second->state >= TOMO_REQ_DONE) it = srv->all.erase(it);
        else ++it;
    }
    *id = q->id;
    srv->cv_work.notify_all();
    return TOMO_OK;
}

namespace {
void fill_result(const Request& q, tomo_serve_result* res) {
    res->state = q.state;
    res->status = q.status;
    res->submit_s = q.submit_s;
    res->admit_s = q.admit_s;
    res->done_s = q.done_s;
    res->latency_s = q.state >= TOMO_REQ_DONE ? q.done_s - q.submit_s : 0.0;
    res->residual_sq = q.residual_sq;
    res->device_bytes = layout_of(q.geom, q.req.gd, q.req.algorithm, q.req.storage == TOMO_BF16 ? 2 : 4).total;
}
}  // namespace

extern "C" int tomo_serve_poll(tomo_server* srv, uint64_t id, tomo_serve_result* res) {
    if (!srv || !res) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) return fail(TOMO_ERR_INVALID_ARG, "unknown request id");
    fill_result(*it->second, res);
    return TOMO_OK;
}

extern "C" int tomo_serve_wait(tomo_server* srv, uint64_t id, double timeout_s, tomo_serve_result* res) {
    if (!srv || !res) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) return fail(TOMO_ERR_INVALID_ARG, "unknown request id");
    std::shared_ptr<Request> q = it->second;
    auto done = [&] { return q->state >= TOMO_REQ_DONE; };
    if (timeout_s < 0) {
        srv->cv_done.wait(lk, done);
    } else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        fill_result(*q, res);
        return fail(TOMO_ERR_BUSY, "request not finished within the timeout");
    }
    fill_result(*q, res);
    if (q->status != TOMO_OK) return fail(q->status, "request failed: " + q->error);
    return TOMO_OK;
}

extern "C" int tomo_serve_iter_stats(tomo_server* srv, uint64_t id, tomo_gd_iter_stat* stats, int n) {
    if (!srv || (!stats && n > 0)) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) return fail(TOMO_ERR_INVALID_ARG, "unknown request id");
    const Request& q = *it->second;
    if (q.state != TOMO_REQ_DONE) return fail(TOMO_ERR_BUSY, "request not done");
    const int m = std::min(n, (int)(q.hist_host.size() / 4));
    for (int i = 0; i < m; ++i) {
        stats[i].residual_sq = q.hist_host[4 * i];
        stats[i].norm_dj = q.hist_host[4 * i + 1];
        stats[i].norm_dj_proj = q.hist_host[4 * i + 2];
        stats[i].alpha = q.hist_host[4 * i + 3];
        stats[i].j_reg = 0.0;
        stats[i].beta = 0.0;
    }
    return TOMO_OK;
}

extern "C" int tomo_serve_get_stats(tomo_server* srv, tomo_serve_stats* st) {
    if (!srv || !st) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = srv->submitted;
    st->completed = srv->completed;
    st->failed = srv->failed;
    st->rounds = srv->rounds;
    st->active = srv->n_active;
    st->queued = (int)srv->queue.size();
    st->peak_active = srv->peak_active;
    st->bytes_in_use = srv->bytes_in_use;
    st->peak_bytes = srv->peak_bytes;
    st->budget_bytes = srv->budget;
    st->first_submit_s = srv->first_submit < 0 ? 0.0 : srv->first_submit;
    st->last_done_s = srv->last_done;
    if (!srv->latencies.empty()) {
        std::vector<double> l = srv->latencies;
        std::sort(l.begin(), l.end());
        // nearest-rank percentiles
        auto pct = [&](double p) {
            size_t r = (size_t)std::ceil(p * (double)l.size());
            if (r < 1) r = 1;
            return l[r - 1];
        };
        st->latency_p50_s = pct(0.50);
        st->latency_p99_s = pct(0.99);
        double s = 0;
        for (double v : l) s += v;
        st->latency_mean_s = s / (double)l.size();
        st->latency_max_s = l.back();
        const double span = srv->last_done - st->first_submit_s;
        st->throughput_rps = span > 0 ? (double)srv->completed / span : 0.0;
    }
    return TOMO_OK;
}

extern "C" int tomo_serve_destroy(tomo_server* srv) {
    if (!srv) return TOMO_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        srv->cv_work.notify_all();
    }
    if (srv->worker.joinable()) srv->worker.join();
    (void)hipSetDevice(srv->ctx->device);
    for (auto& gx : srv->graph)
        if (gx) (void)hipGraphExecDestroy(gx);
    if (srv->d_tables) (void)hipFree(srv->d_tables);
    if (srv->h_tables) (void)hipHostFree(srv->h_tables);
    if (srv->stream) {
        (void)hipStreamSynchronize(srv->stream);
        (void)hipStreamDestroy(srv->stream);
    }
    if (srv->pool) (void)hipMemPoolDestroy(srv->pool);
    delete srv;
    return TOMO_OK;
}
