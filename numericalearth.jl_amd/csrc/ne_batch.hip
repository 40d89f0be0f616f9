// ne_batch.hip — ensembles of models stepped together.
//
// A batch binds existing models (each created by ne_model_create, so each owns its
// descriptors and RK stage buffers) into one device table of member records. A step is
// six launches for the whole ensemble, whatever its size: three RK stages, the member
// clocks, the surface solve and the surface finalize, each a persistent grid walking a
// work list of (member, tile) or (member, column block) items. The six launches can be
// captured once into a HIP graph and replayed per step; every quantity that changes from
// step to step (member clocks, time-interpolation indices) lives in device memory, so
// the graph stays valid for a fixed Δt. The tile shape and the grid sizes come from an
// autotuner that times the candidates on the batch itself.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "ne_internal.h"

using namespace ne;

namespace {

constexpr int NXCD = 8;
constexpr int TILE_ROWS[3] = {4, 8, 16};
constexpr int SURFACE_BLOCKS[3] = {64, 128, 256};

inline int tr_index(int tr) { return tr == 4 ? 0 : tr == 8 ? 1 : tr == 16 ? 2 : -1; }
inline int sb_index(int b) { return b == 64 ? 0 : b == 128 ? 1 : b == 256 ? 2 : -1; }
inline int round8(int n) { return ((n + NXCD - 1) / NXCD) * NXCD; }

template <typename Real>
__global__ void batch_clock_kernel(BatchRec<Real>* recs, int n, double dt) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    BatchRec<Real>& r = recs[m];
    double t = r.time + dt;
    r.time = t;
    r.a.time = time_indices(r.atimes, r.nat, t);
    r.r.time = time_indices(r.rtimes, r.nrt, t);
}

}  // namespace

struct ne_batch {
    ne_ctx* ctx = nullptr;
    int precision = NE_FP64;
    std::vector<ne_model*> members;
    void* recs = nullptr;                 // device BatchRec<Real>[n]
    double* times = nullptr;              // device snapshot times of every member
    WorkItem* stage_work[3] = {nullptr, nullptr, nullptr};    // per tile-row choice
    int stage_nwork[3] = {0, 0, 0};
    // per column-block choice: every member (finalize), and the members of each flux path (solve)
    WorkItem* surface_work[3] = {nullptr, nullptr, nullptr};
    int surface_nwork[3] = {0, 0, 0};
    WorkItem* solve_work[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};  // [fast][choice]
    int solve_nwork[2][3] = {{0, 0, 0}, {0, 0, 0}};
    ne_batch_tuning tuning;
    int cus = 256;
    bool use_graph = false;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double graph_dt = 0.0;
    int64_t cells = 0;
};

namespace {

void drop_graph(ne_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

template <typename Real>
hipError_t enqueue_step(ne_batch* b, double dt, hipStream_t s, ne_counters* counters) {
    const BatchRec<Real>* recs = static_cast<const BatchRec<Real>*>(b->recs);
    const ne_batch_tuning& t = b->tuning;
    int ti = tr_index(t.tile_rows), si = sb_index(t.surface_block);
    if (ti < 0 || si < 0) return hipErrorInvalidValue;
    const double beta[3] = {1.0 / 3.0, 0.5, 1.0};
    hipError_t e;
    for (int st = 0; st < 3; st++) {
        e = launch_batch_stage<Real>(recs, b->stage_work[ti], b->stage_nwork[ti], t.tile_rows, t.stage_blocks, st + 1,
                                     beta[st] * dt, s);
        if (e != hipSuccess) return e;
    }
    int n = (int)b->members.size();
    hipLaunchKernelGGL(batch_clock_kernel<Real>, dim3((n + 63) / 64), dim3(64), 0, s,
                       static_cast<BatchRec<Real>*>(b->recs), n, dt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int fast = 1; fast >= 0; fast--) {
        int nw = b->solve_nwork[fast][si];
        e = launch_batch_surface_solve<Real>(recs, b->solve_work[fast][si], nw, t.surface_block,
                                             std::min(t.surface_blocks, std::max(nw, 1)), fast == 1, counters, s);
        if (e != hipSuccess) return e;
    }
    return launch_batch_surface_finalize<Real>(recs, b->surface_work[si], b->surface_nwork[si], t.surface_block,
                                               t.surface_blocks, dt, s);
}

hipError_t enqueue(ne_batch* b, double dt, hipStream_t s, ne_counters* counters) {
    return b->precision == NE_FP64 ? enqueue_step<double>(b, dt, s, counters)
                                   : enqueue_step<float>(b, dt, s, counters);
}

template <typename Real>
BatchRec<Real> member_record(const ne_model* m, const double* atimes_dev, const double* rtimes_dev) {
    const ne_model_desc& d = m->d;
    BatchRec<Real> r;
    memset(&r, 0, sizeof(r));
    r.g = grid_dev(d.grid);
    r.p = d.params;
    r.a = atmos_dev<Real>(d.atmosphere);
    r.a.time = time_indices(m->atimes.data(), (int)m->atimes.size(), m->time);
    if (d.params.radiation) {
        r.r = rad_dev<Real>(d.radiation);
        r.r.time = time_indices(m->rtimes.data(), (int)m->rtimes.size(), m->time);
    }
    r.afi = d.atmos_fi;
    r.afj = d.atmos_fj;
    r.rfi = d.radiation_fi;
    r.rfj = d.radiation_fj;
    r.T = view<Real>(d.ocean.T);
    r.S = view<Real>(d.ocean.S);
    r.T1 = view<Real>(m->vT1);
    r.S1 = view<Real>(m->vS1);
    r.T2 = view<Real>(m->vT2);
    r.S2 = view<Real>(m->vS2);
    r.u = view<const Real>(d.ocean.u);
    r.v = view<const Real>(d.ocean.v);
    r.w = view<const Real>(d.ocean.w);
    r.f = flux_out(d.fluxes);
    r.n = net_out(d.net);
    r.ro = rad_out(d.radiative);
    r.frazil = view<double>(d.frazil_heat);
    r.atimes = atimes_dev;
    r.rtimes = rtimes_dev;
    r.nat = (int)m->atimes.size();
    r.nrt = (int)m->rtimes.size();
    r.fast = edson_default_constants(d.params) ? 1 : 0;
    r.time = m->time;
    return r;
}

template <typename Real>
int upload_records(ne_batch* b, hipStream_t s) {
    // snapshot times of every member, packed
    std::vector<double> times;
    std::vector<size_t> aoff, roff;
    for (ne_model* m : b->members) {
        aoff.push_back(times.size());
        times.insert(times.end(), m->atimes.begin(), m->atimes.end());
        roff.push_back(times.size());
        times.insert(times.end(), m->rtimes.begin(), m->rtimes.end());
    }
    if (!times.empty()) {
        if (hipMalloc((void**)&b->times, times.size() * sizeof(double)) != hipSuccess) return NE_ERR_DEVICE;
        if (hipMemcpy(b->times, times.data(), times.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return NE_ERR_DEVICE;
    }
    std::vector<BatchRec<Real>> recs;
    for (size_t k = 0; k < b->members.size(); k++) {
        const ne_model* m = b->members[k];
        const double* at = m->atimes.empty() ? nullptr : b->times + aoff[k];
        const double* rt = m->rtimes.empty() ? nullptr : b->times + roff[k];
        recs.push_back(member_record<Real>(m, at, rt));
    }
    size_t bytes = recs.size() * sizeof(BatchRec<Real>);
    if (hipMalloc(&b->recs, bytes) != hipSuccess) return NE_ERR_DEVICE;
    if (hipMemcpyAsync(b->recs, recs.data(), bytes, hipMemcpyHostToDevice, s) != hipSuccess) return NE_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return NE_ERR_DEVICE;
    return NE_OK;
}

int upload_work(WorkItem** dst, int* count, const std::vector<WorkItem>& w) {
    *count = (int)w.size();
    if (w.empty()) return NE_OK;
    if (hipMalloc((void**)dst, w.size() * sizeof(WorkItem)) != hipSuccess) return NE_ERR_DEVICE;
    if (hipMemcpy(*dst, w.data(), w.size() * sizeof(WorkItem), hipMemcpyHostToDevice) != hipSuccess)
        return NE_ERR_DEVICE;
    return NE_OK;
}

int build_work_lists(ne_batch* b) {
    for (int c = 0; c < 3; c++) {
        std::vector<WorkItem> w;
        for (size_t k = 0; k < b->members.size(); k++) {
            const ne_grid& g = b->members[k]->d.grid;
            int nt = batch_tiles_x(g.Nx) * batch_tiles_y(g.Ny, TILE_ROWS[c]);
            for (int t = 0; t < nt; t++) w.push_back(WorkItem{(int)k, t});
        }
        int r = upload_work(&b->stage_work[c], &b->stage_nwork[c], w);
        if (r) return r;
    }
    for (int c = 0; c < 3; c++) {
        std::vector<WorkItem> w, wf[2];
        for (size_t k = 0; k < b->members.size(); k++) {
            const ne_grid& g = b->members[k]->d.grid;
            int fast = edson_default_constants(b->members[k]->d.params) ? 1 : 0;
            int cols = g.Nx * g.Ny;
            for (int s = 0; s < cols; s += SURFACE_BLOCKS[c]) {
                w.push_back(WorkItem{(int)k, s});
                wf[fast].push_back(WorkItem{(int)k, s});
            }
        }
        int r = upload_work(&b->surface_work[c], &b->surface_nwork[c], w);
        for (int f = 0; f < 2 && !r; f++) r = upload_work(&b->solve_work[f][c], &b->solve_nwork[f][c], wf[f]);
        if (r) return r;
    }
    return NE_OK;
}

void default_tuning(ne_batch* b) {
    ne_batch_tuning& t = b->tuning;
    memset(&t, 0, sizeof(t));
    t.tile_rows = 8;
    t.stage_blocks = std::max(NXCD, std::min(round8(b->stage_nwork[1]), 2 * b->cus));
    t.surface_block = 256;
    t.surface_blocks = std::max(1, std::min(b->surface_nwork[2], 4 * b->cus));
}

int check_tuning(const ne_batch* b, const ne_batch_tuning& t) {
    if (tr_index(t.tile_rows) < 0 || sb_index(t.surface_block) < 0) return NE_ERR_UNSUPPORTED;
    if (t.stage_blocks <= 0 || t.stage_blocks % NXCD != 0 || t.surface_blocks <= 0) return NE_ERR_SHAPE;
    if (t.stage_blocks > (1 << 20) || t.surface_blocks > (1 << 20)) return NE_ERR_CAPACITY;
    return NE_OK;
}

// Mean milliseconds of `reps` launches of f after one untimed launch.
template <typename F>
int time_launches(hipStream_t s, int reps, F f, double* ms) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return NE_ERR_DEVICE;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return NE_ERR_DEVICE;
    }
    int r = NE_OK;
    if (f() != hipSuccess || hipEventRecord(e0, s) != hipSuccess) r = NE_ERR_DEVICE;
    for (int k = 0; k < reps && !r; k++)
        if (f() != hipSuccess) r = NE_ERR_DEVICE;
    float t = 0.0f;
    if (!r && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
               hipEventElapsedTime(&t, e0, e1) != hipSuccess))
        r = NE_ERR_DEVICE;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / reps;
    return r;
}

template <typename Real>
int autotune(ne_batch* b, hipStream_t s, ne_batch_tuning* out) {
    const BatchRec<Real>* recs = static_cast<const BatchRec<Real>*>(b->recs);
    ne_batch_tuning best = b->tuning;
    const int reps = 5;
    // RK stage: stage 1 writes only the stage buffers T¹, S¹ from T, S, so timing it leaves
    // the state untouched.
    double best_stage = INFINITY;
    for (int c = 0; c < 3; c++) {
        int nwork = b->stage_nwork[c];
        if (nwork == 0) continue;
        const int per_cu[4] = {1, 2, 4, 0};   // 0: one block per work item
        for (int pc : per_cu) {
            int blocks = pc ? std::min(round8(nwork), pc * b->cus) : round8(nwork);
            blocks = std::max(blocks, NXCD);
            double ms;
            auto f = [&]() {
                return launch_batch_stage<Real>(recs, b->stage_work[c], nwork, TILE_ROWS[c], blocks, 1,
                                                1.0 / 3.0 * 1.0, s);
            };
            int r = time_launches(s, reps, f, &ms);
            if (r) return r;
            if (ms < best_stage) {
                best_stage = ms;
                best.tile_rows = TILE_ROWS[c];
                best.stage_blocks = blocks;
            }
        }
    }
    // Surface solve: recomputes the current fluxes from the current state (idempotent);
    // the solver counters are not accumulated while tuning.
    double best_surface = INFINITY;
    for (int c = 0; c < 3; c++) {
        int nwork = b->surface_nwork[c];
        if (nwork == 0) continue;
        const int per_cu[4] = {2, 4, 8, 0};
        for (int pc : per_cu) {
            int blocks = pc ? std::min(nwork, pc * b->cus) : nwork;
            double ms;
            auto f = [&]() {
                hipError_t e = hipSuccess;
                for (int fast = 1; fast >= 0 && e == hipSuccess; fast--) {
                    int nw = b->solve_nwork[fast][c];
                    e = launch_batch_surface_solve<Real>(recs, b->solve_work[fast][c], nw, SURFACE_BLOCKS[c],
                                                         std::min(blocks, std::max(nw, 1)), fast == 1, nullptr, s);
                }
                return e;
            };
            int r = time_launches(s, reps, f, &ms);
            if (r) return r;
            if (ms < best_surface) {
                best_surface = ms;
                best.surface_block = SURFACE_BLOCKS[c];
                best.surface_blocks = blocks;
            }
        }
    }
    best.stage_ms = best_stage;
    best.surface_ms = best_surface;
    b->tuning = best;
    drop_graph(b);
    if (out) *out = best;
    return NE_OK;
}

int step_host_clocks(ne_batch* b, double dt) {
    for (ne_model* m : b->members) {
        m->time += dt;
        m->iteration += 1;
        m->last_dt = dt;
    }
    return NE_OK;
}

}  // namespace

extern "C" {

int ne_batch_create(ne_ctx* c, int n, ne_model* const* models, ne_batch** out, void* stream) {
    if (!c || !models || !out) return NE_ERR_NULL;
    *out = nullptr;
    if (n <= 0 || n > (1 << 20)) return NE_ERR_SHAPE;
    int precision = -1;
    for (int k = 0; k < n; k++) {
        ne_model* m = models[k];
        if (!m) return NE_ERR_NULL;
        if (m->ctx != c) return NE_ERR_STATE;
        if (m->batch) return NE_ERR_STATE;
        if (m->timing) return NE_ERR_STATE;
        if (precision < 0) precision = m->d.precision;
        if (m->d.precision != precision) return NE_ERR_UNSUPPORTED;
        for (int j = 0; j < k; j++)
            if (models[j] == m) return NE_ERR_STATE;
    }
    ne_batch* b = new (std::nothrow) ne_batch();
    if (!b) return NE_ERR_CAPACITY;
    b->ctx = c;
    b->precision = precision;
    b->members.assign(models, models + n);
    for (ne_model* m : b->members) {
        const ne_grid& g = m->d.grid;
        b->cells += (int64_t)g.Nx * g.Ny * g.Nz;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && cus > 0)
        b->cus = cus;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int r = precision == NE_FP64 ? upload_records<double>(b, s) : upload_records<float>(b, s);
    if (!r) r = build_work_lists(b);
    if (!r && hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking) != hipSuccess) r = NE_ERR_DEVICE;
    if (r) {
        ne_batch_destroy(b);
        return r;
    }
    default_tuning(b);
    for (ne_model* m : b->members) m->batch = b;
    *out = b;
    return NE_OK;
}

int ne_batch_destroy(ne_batch* b) {
    if (!b) return NE_ERR_NULL;
    int r = NE_OK;
    drop_graph(b);
    for (ne_model* m : b->members)
        if (m->batch == b) m->batch = nullptr;
    void* bufs[14] = {b->recs, b->times, b->stage_work[0], b->stage_work[1], b->stage_work[2],
                      b->surface_work[0], b->surface_work[1], b->surface_work[2],
                      b->solve_work[0][0], b->solve_work[0][1], b->solve_work[0][2],
                      b->solve_work[1][0], b->solve_work[1][1], b->solve_work[1][2]};
    for (void* p : bufs)
        if (p && hipFree(p) != hipSuccess) r = NE_ERR_DEVICE;
    if (b->capture_stream && hipStreamDestroy(b->capture_stream) != hipSuccess) r = NE_ERR_DEVICE;
    delete b;
    return r;
}

int ne_batch_time_step(ne_batch* b, double dt, void* stream) {
    if (!b) return NE_ERR_NULL;
    if (!(dt > 0) || !isfinite(dt)) return NE_ERR_SHAPE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (b->use_graph) {
        if (!b->exec || b->graph_dt != dt) {
            drop_graph(b);
            if (hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return NE_ERR_DEVICE;
            hipError_t e = enqueue(b, dt, b->capture_stream, b->ctx->counters);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(b->capture_stream, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return NE_ERR_DEVICE;
            }
            b->graph = g;
            if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
                drop_graph(b);
                return NE_ERR_DEVICE;
            }
            b->graph_dt = dt;
        }
        if (hipGraphLaunch(b->exec, s) != hipSuccess) return NE_ERR_DEVICE;
    } else {
        if (enqueue(b, dt, s, b->ctx->counters) != hipSuccess) return NE_ERR_DEVICE;
    }
    return step_host_clocks(b, dt);
}

int ne_batch_set_graph(ne_batch* b, int enabled) {
    if (!b) return NE_ERR_NULL;
    b->use_graph = enabled != 0;
    if (!b->use_graph) drop_graph(b);
    return NE_OK;
}

int ne_batch_autotune(ne_batch* b, ne_batch_tuning* out, void* stream) {
    if (!b) return NE_ERR_NULL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    return b->precision == NE_FP64 ? autotune<double>(b, s, out) : autotune<float>(b, s, out);
}

int ne_batch_get_tuning(const ne_batch* b, ne_batch_tuning* out) {
    if (!b || !out) return NE_ERR_NULL;
    *out = b->tuning;
    return NE_OK;
}

int ne_batch_set_tuning(ne_batch* b, const ne_batch_tuning* t) {
    if (!b || !t) return NE_ERR_NULL;
    int r = check_tuning(b, *t);
    if (r) return r;
    b->tuning = *t;
    drop_graph(b);
    return NE_OK;
}

int ne_batch_info(const ne_batch* b, int* members, int64_t* cells, int* stage_items, int* surface_items) {
    if (!b) return NE_ERR_NULL;
    if (members) *members = (int)b->members.size();
    if (cells) *cells = b->cells;
    int ti = tr_index(b->tuning.tile_rows), si = sb_index(b->tuning.surface_block);
    if (stage_items) *stage_items = ti >= 0 ? b->stage_nwork[ti] : 0;
    if (surface_items) *surface_items = si >= 0 ? b->surface_nwork[si] : 0;
    return NE_OK;
}

}  // extern "C"
