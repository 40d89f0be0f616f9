// ne_batch.hip — ensembles of models stepped together.
//
// A batch binds existing models (each created by ne_model_create, so each owns its
// descriptors and RK stage buffers) into one device table of member records. A step is
// six or seven launches for the whole ensemble, whatever its size: three RK stages, the
// member clocks, the surface solve (one launch per flux path in use) and the surface
// finalize, each a persistent grid walking a work list of (member, tile) or
// (member, column block) items. The launches can be captured once into a HIP graph and
// replayed per step; every quantity that changes from step to step (member clocks,
// time-interpolation indices) lives in device memory, so the graph stays valid for a
// fixed Δt and a fixed membership.
//
// Membership may change between steps (ne_batch_add / ne_batch_remove, for serving).
// The member table and the work lists live in one device arena and are rebuilt on the
// host into one of two pinned images, then uploaded with a single stream-ordered copy
// before the next step; the arena grows geometrically and retired arenas are released
// with the batch, so no step synchronises the device (a rebuild waits at most for the
// upload made two rebuilds earlier).
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
inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

template <typename Real>
__global__ void batch_clock_kernel(BatchRec<Real>* recs, int n, double dt) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    BatchRec<Real>& r = recs[m];
    double t = r.time + dt;
    r.time = t;
    r.a.time = time_indices(r.atimes, r.nat, t, r.time_indexing, r.time_period);
    r.r.time = time_indices(r.rtimes, r.nrt, t, r.time_indexing, r.time_period);
}

}  // namespace

struct ne_batch {
    ne_ctx* ctx = nullptr;
    int precision = NE_FP64;
    std::vector<ne_model*> members;
    bool dirty = true;                    // membership changed since the last upload

    // device arena (member records, snapshot times, work lists) and two pinned host images,
    // used in turn so that a rebuild waits only for the upload before the previous one
    char* dev = nullptr;
    size_t dev_cap = 0;
    char* host[2] = {nullptr, nullptr};
    size_t host_cap[2] = {0, 0};
    int image = 0;
    std::vector<void*> retired_dev, retired_host;
    hipEvent_t uploaded[2] = {nullptr, nullptr};   // recorded after the last upload from host[i]
    bool upload_pending[2] = {false, false};

    // views into the arena
    void* recs = nullptr;
    // per forcing (without / with the two-colour shortwave) and tile-row choice
    WorkItem* stage_work[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    int stage_nwork[2][3] = {{0, 0, 0}, {0, 0, 0}};
    // per column-block choice: every member (finalize), and the members of each flux path (solve)
    WorkItem* surface_work[3] = {nullptr, nullptr, nullptr};
    int surface_nwork[3] = {0, 0, 0};
    WorkItem* solve_work[3][3] = {};   // [flux path][choice]
    int solve_nwork[3][3] = {};

    ne_batch_tuning tuning;
    int cus = 256;
    bool use_graph = false;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double graph_dt = 0.0;
};

namespace {

void drop_graph(ne_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

inline int blocks_for(int requested, int nwork) { return requested > 0 ? requested : std::max(nwork, 1); }

template <typename Real>
hipError_t enqueue_step(ne_batch* b, double dt, hipStream_t s, ne_counters* counters) {
    int n = (int)b->members.size();
    if (n == 0) return hipSuccess;
    const BatchRec<Real>* recs = static_cast<const BatchRec<Real>*>(b->recs);
    const ne_batch_tuning& t = b->tuning;
    int ti = tr_index(t.tile_rows), si = sb_index(t.surface_block);
    if (ti < 0 || si < 0) return hipErrorInvalidValue;
    const double beta[3] = {1.0 / 3.0, 0.5, 1.0};
    hipError_t e;
    for (int st = 0; st < 3; st++)
        for (int tc = 0; tc < 2; tc++) {
            const int nst = b->stage_nwork[tc][ti];
            if (nst == 0) continue;
            int sblocks = t.stage_blocks > 0 ? std::min(t.stage_blocks, round8(nst)) : round8(nst);
            sblocks = std::max(sblocks, NXCD);
            e = launch_batch_stage<Real>(recs, b->stage_work[tc][ti], nst, t.tile_rows, sblocks, st + 1,
                                         beta[st] * dt, tc == 1, s);
            if (e != hipSuccess) return e;
        }
    hipLaunchKernelGGL(batch_clock_kernel<Real>, dim3((n + 63) / 64), dim3(64), 0, s,
                       static_cast<BatchRec<Real>*>(b->recs), n, dt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int path = 2; path >= 0; path--) {
        int nw = b->solve_nwork[path][si];
        e = launch_batch_surface_solve<Real>(recs, b->solve_work[path][si], nw, t.surface_block,
                                             std::min(blocks_for(t.surface_blocks, nw), std::max(nw, 1)), path,
                                             true, dt, counters, s);
        if (e != hipSuccess) return e;
    }
    int nsu = b->surface_nwork[si];
    return launch_batch_surface_tau<Real>(recs, b->surface_work[si], nsu, t.surface_block,
                                          std::min(blocks_for(t.surface_blocks, nsu), std::max(nsu, 1)), s);
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
    r.a.time = time_indices(m->atimes.data(), (int)m->atimes.size(), m->time, d.time_indexing, d.time_period);
    if (d.params.radiation) {
        r.r = rad_dev<Real>(d.radiation);
        r.r.time = time_indices(m->rtimes.data(), (int)m->rtimes.size(), m->time, d.time_indexing, d.time_period);
    }
    r.time_indexing = d.time_indexing;
    r.time_period = d.time_period;
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
    r.freeze_flags = m->freeze_flags;
    r.residual = m->residual;
    r.beer = m->beer;
    r.atimes = atimes_dev;
    r.rtimes = rtimes_dev;
    r.nat = (int)m->atimes.size();
    r.nrt = (int)m->rtimes.size();
    r.fast = flux_path(d.params);
    r.time = m->time;
    return r;
}

// Host-side layout of the arena: records | times | 2×3 stage lists | 3 surface lists | 3×3 solve lists.
struct Layout {
    size_t recs = 0, times = 0, stage[2][3] = {{0, 0, 0}, {0, 0, 0}}, surface[3] = {0, 0, 0},
           solve[3][3] = {};
    size_t total = 0;
};

template <typename Real>
int rebuild(ne_batch* b, hipStream_t s) {
    const int n = (int)b->members.size();
    // work lists and snapshot times on the host
    std::vector<WorkItem> stage[2][3], surface[3], solve[3][3];
    std::vector<double> times;
    std::vector<size_t> aoff(n), roff(n);
    for (int k = 0; k < n; k++) {
        const ne_model* m = b->members[k];
        const ne_grid& g = m->d.grid;
        const int tc = m->d.params.radiative_forcing == NE_RADIATIVE_FORCING_TWO_COLOR ? 1 : 0;
        for (int c = 0; c < 3; c++) {
            int nt = batch_tiles_x(g.Nx) * batch_tiles_y(g.Ny, TILE_ROWS[c]);
            for (int t = 0; t < nt; t++) stage[tc][c].push_back(WorkItem{k, t});
        }
        int path = flux_path(m->d.params);
        int cols = g.Nx * g.Ny;
        for (int c = 0; c < 3; c++)
            for (int s0 = 0; s0 < cols; s0 += SURFACE_BLOCKS[c]) {
                surface[c].push_back(WorkItem{k, s0});
                solve[path][c].push_back(WorkItem{k, s0});
            }
        aoff[k] = times.size();
        times.insert(times.end(), m->atimes.begin(), m->atimes.end());
        roff[k] = times.size();
        times.insert(times.end(), m->rtimes.begin(), m->rtimes.end());
    }
    Layout L;
    size_t off = 0;
    L.recs = off;
    off = align256(off + (size_t)n * sizeof(BatchRec<Real>));
    L.times = off;
    off = align256(off + times.size() * sizeof(double));
    for (int c = 0; c < 3; c++) {
        for (int f = 0; f < 2; f++) {
            L.stage[f][c] = off;
            off = align256(off + stage[f][c].size() * sizeof(WorkItem));
        }
        L.surface[c] = off;
        off = align256(off + surface[c].size() * sizeof(WorkItem));
        for (int f = 0; f < 3; f++) {
            L.solve[f][c] = off;
            off = align256(off + solve[f][c].size() * sizeof(WorkItem));
        }
    }
    L.total = std::max<size_t>(off, 256);

    // this pinned image may still be read by the upload made from it two rebuilds ago
    const int im = b->image;
    b->image ^= 1;
    if (b->upload_pending[im]) {
        if (hipEventSynchronize(b->uploaded[im]) != hipSuccess) return NE_ERR_DEVICE;
        b->upload_pending[im] = false;
    }
    if (b->host_cap[im] < L.total) {
        size_t cap = std::max(L.total, 2 * b->host_cap[im]);
        void* h = nullptr;
        if (hipHostMalloc(&h, cap, hipHostMallocDefault) != hipSuccess) return NE_ERR_DEVICE;
        if (b->host[im]) b->retired_host.push_back(b->host[im]);
        b->host[im] = static_cast<char*>(h);
        b->host_cap[im] = cap;
    }
    if (b->dev_cap < L.total) {
        // earlier steps may still read the old arena: it is released with the batch
        size_t cap = std::max(L.total, 2 * b->dev_cap);
        void* d = nullptr;
        if (hipMalloc(&d, cap) != hipSuccess) return NE_ERR_DEVICE;
        if (b->dev) b->retired_dev.push_back(b->dev);
        b->dev = static_cast<char*>(d);
        b->dev_cap = cap;
    }
    char* H = b->host[im];
    char* D = b->dev;
    const double* tdev = reinterpret_cast<const double*>(D + L.times);
    BatchRec<Real>* recs = reinterpret_cast<BatchRec<Real>*>(H + L.recs);
    for (int k = 0; k < n; k++) {
        const ne_model* m = b->members[k];
        recs[k] = member_record<Real>(m, m->atimes.empty() ? nullptr : tdev + aoff[k],
                                      m->rtimes.empty() ? nullptr : tdev + roff[k]);
    }
    if (!times.empty()) memcpy(H + L.times, times.data(), times.size() * sizeof(double));
    auto put = [&](size_t o, const std::vector<WorkItem>& w, WorkItem** view, int* count) {
        if (!w.empty()) memcpy(H + o, w.data(), w.size() * sizeof(WorkItem));
        *view = reinterpret_cast<WorkItem*>(D + o);
        *count = (int)w.size();
    };
    for (int c = 0; c < 3; c++) {
        for (int f = 0; f < 2; f++) put(L.stage[f][c], stage[f][c], &b->stage_work[f][c], &b->stage_nwork[f][c]);
        put(L.surface[c], surface[c], &b->surface_work[c], &b->surface_nwork[c]);
        for (int f = 0; f < 3; f++) put(L.solve[f][c], solve[f][c], &b->solve_work[f][c], &b->solve_nwork[f][c]);
    }
    b->recs = D + L.recs;
    if (hipMemcpyAsync(D, H, L.total, hipMemcpyHostToDevice, s) != hipSuccess) return NE_ERR_DEVICE;
    if (hipEventRecord(b->uploaded[im], s) != hipSuccess) return NE_ERR_DEVICE;
    b->upload_pending[im] = true;
    b->dirty = false;
    drop_graph(b);
    return NE_OK;
}

int ensure_uploaded(ne_batch* b, hipStream_t s) {
    if (!b->dirty) return NE_OK;
    return b->precision == NE_FP64 ? rebuild<double>(b, s) : rebuild<float>(b, s);
}

void default_tuning(ne_batch* b) {
    // one block per work item and the shortest tiles: the autotuner's usual choice for
    // ensembles of uneven members
    ne_batch_tuning& t = b->tuning;
    memset(&t, 0, sizeof(t));
    t.tile_rows = 4;
    t.stage_blocks = 0;
    t.surface_block = 256;
    t.surface_blocks = 0;
}

int check_tuning(const ne_batch_tuning& t) {
    if (tr_index(t.tile_rows) < 0 || sb_index(t.surface_block) < 0) return NE_ERR_UNSUPPORTED;
    if (t.stage_blocks < 0 || t.stage_blocks % NXCD != 0 || t.surface_blocks < 0) return NE_ERR_SHAPE;
    if (t.stage_blocks > (1 << 24) || t.surface_blocks > (1 << 24)) return NE_ERR_CAPACITY;
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
    // the state untouched. Grids: 1, 2 or 4 persistent blocks per CU, or one block per item (0).
    double best_stage = INFINITY;
    for (int c = 0; c < 3; c++) {
        const int nwork = b->stage_nwork[0][c] + b->stage_nwork[1][c];
        if (nwork == 0) continue;
        const int per_cu[4] = {1, 2, 4, 0};
        for (int pc : per_cu) {
            const int blocks = pc ? std::max(NXCD, std::min(round8(nwork), pc * b->cus)) : round8(nwork);
            double ms;
            auto f = [&]() {
                // both forcing lists, each with the grid the step would give it
                for (int tc = 0; tc < 2; tc++) {
                    const int nw = b->stage_nwork[tc][c];
                    if (nw == 0) continue;
                    const int bl = std::max(NXCD, pc ? std::min(blocks, round8(nw)) : round8(nw));
                    hipError_t e = launch_batch_stage<Real>(recs, b->stage_work[tc][c], nw, TILE_ROWS[c], bl, 1,
                                                            1.0 / 3.0 * 1200.0, tc == 1, s);
                    if (e != hipSuccess) return e;
                }
                return hipSuccess;
            };
            int r = time_launches(s, reps, f, &ms);
            if (r) return r;
            if (ms < best_stage) {
                best_stage = ms;
                best.tile_rows = TILE_ROWS[c];
                best.stage_blocks = pc && blocks < round8(nwork) ? blocks : 0;
            }
        }
    }
    // Surface solve: recomputes the current fluxes from the current state (idempotent);
    // the solver counters are not accumulated while tuning.
    double best_surface = INFINITY;
    for (int c = 0; c < 3; c++) {
        if (b->surface_nwork[c] == 0) continue;
        const int per_cu[4] = {2, 4, 8, 0};
        for (int pc : per_cu) {
            int req = pc ? pc * b->cus : 0;
            double ms;
            auto f = [&]() {
                hipError_t e = hipSuccess;
                for (int path = 2; path >= 0 && e == hipSuccess; path--) {
                    int nw = b->solve_nwork[path][c];
                    e = launch_batch_surface_solve<Real>(recs, b->solve_work[path][c], nw, SURFACE_BLOCKS[c],
                                                         std::min(blocks_for(req, nw), std::max(nw, 1)), path,
                                                         false, 0.0, nullptr, s);
                }
                return e;
            };
            int r = time_launches(s, reps, f, &ms);
            if (r) return r;
            if (ms < best_surface) {
                best_surface = ms;
                best.surface_block = SURFACE_BLOCKS[c];
                best.surface_blocks = req < b->surface_nwork[c] ? req : 0;
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

void step_host_clocks(ne_batch* b, double dt) {
    for (ne_model* m : b->members) {
        m->time += dt;
        m->iteration += 1;
        m->last_dt = dt;
    }
}

int check_member(const ne_batch* b, const ne_model* m) {
    if (!m) return NE_ERR_NULL;
    if (m->ctx != b->ctx || m->batch || m->timing) return NE_ERR_STATE;
    if (m->d.precision != b->precision) return NE_ERR_UNSUPPORTED;
    // the device clock indexes whole series: windowed snapshot views stay with single models
    if (m->d.atmosphere_slots > 0 || m->d.radiation_slots > 0) return NE_ERR_UNSUPPORTED;
    // the batched step is the split RK3 one
    if (m->d.timestepper != NE_TIMESTEPPER_SPLIT_RK3) return NE_ERR_UNSUPPORTED;
    return NE_OK;
}

}  // namespace

extern "C" {

int ne_batch_create(ne_ctx* c, int32_t precision, int n, ne_model* const* models, ne_batch** out, void* stream) {
    if (!c || !out || (n > 0 && !models)) return NE_ERR_NULL;
    *out = nullptr;
    if (precision != NE_FP64 && precision != NE_MIXED_FP32) return NE_ERR_UNSUPPORTED;
    if (n < 0 || n > (1 << 20)) return NE_ERR_SHAPE;
    ne_batch* b = new (std::nothrow) ne_batch();
    if (!b) return NE_ERR_CAPACITY;
    b->ctx = c;
    b->precision = precision;
    for (int k = 0; k < n; k++) {
        int r = check_member(b, models[k]);
        if (!r && std::find(b->members.begin(), b->members.end(), models[k]) != b->members.end()) r = NE_ERR_STATE;
        if (r) {
            delete b;
            return r;
        }
        b->members.push_back(models[k]);
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && cus > 0)
        b->cus = cus;
    default_tuning(b);
    int r = NE_OK;
    for (int i = 0; i < 2 && !r; i++)
        if (hipEventCreateWithFlags(&b->uploaded[i], hipEventDisableTiming) != hipSuccess) r = NE_ERR_DEVICE;
    if (!r && hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking) != hipSuccess) r = NE_ERR_DEVICE;
    if (!r) r = ensure_uploaded(b, static_cast<hipStream_t>(stream));
    if (r) {
        ne_batch_destroy(b);
        return r;
    }
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
    // the arena may still be read by enqueued steps: wait for the device before releasing it
    if ((b->dev || !b->retired_dev.empty()) && hipDeviceSynchronize() != hipSuccess) r = NE_ERR_DEVICE;
    if (b->dev) b->retired_dev.push_back(b->dev);
    for (int i = 0; i < 2; i++)
        if (b->host[i]) b->retired_host.push_back(b->host[i]);
    for (void* p : b->retired_dev)
        if (hipFree(p) != hipSuccess) r = NE_ERR_DEVICE;
    for (void* p : b->retired_host)
        if (hipHostFree(p) != hipSuccess) r = NE_ERR_DEVICE;
    for (int i = 0; i < 2; i++)
        if (b->uploaded[i] && hipEventDestroy(b->uploaded[i]) != hipSuccess) r = NE_ERR_DEVICE;
    if (b->capture_stream && hipStreamDestroy(b->capture_stream) != hipSuccess) r = NE_ERR_DEVICE;
    delete b;
    return r;
}

int ne_batch_add(ne_batch* b, ne_model* m) {
    if (!b || !m) return NE_ERR_NULL;
    int r = check_member(b, m);
    if (r) return r;
    if (b->members.size() >= (1u << 20)) return NE_ERR_CAPACITY;
    b->members.push_back(m);
    m->batch = b;
    b->dirty = true;
    return NE_OK;
}

int ne_batch_remove(ne_batch* b, ne_model* m) {
    if (!b || !m) return NE_ERR_NULL;
    auto it = std::find(b->members.begin(), b->members.end(), m);
    if (it == b->members.end() || m->batch != b) return NE_ERR_STATE;
    b->members.erase(it);
    m->batch = nullptr;
    b->dirty = true;
    return NE_OK;
}

int ne_batch_time_step(ne_batch* b, double dt, void* stream) {
    if (!b) return NE_ERR_NULL;
    if (!(dt > 0) || !isfinite(dt)) return NE_ERR_SHAPE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int r = ensure_uploaded(b, s);
    if (r) return r;
    if (b->members.empty()) return NE_OK;
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
    step_host_clocks(b, dt);
    return NE_OK;
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
    int r = ensure_uploaded(b, s);
    if (r) return r;
    if (b->members.empty()) return NE_ERR_STATE;
    return b->precision == NE_FP64 ? autotune<double>(b, s, out) : autotune<float>(b, s, out);
}

int ne_batch_get_tuning(const ne_batch* b, ne_batch_tuning* out) {
    if (!b || !out) return NE_ERR_NULL;
    *out = b->tuning;
    return NE_OK;
}

int ne_batch_set_tuning(ne_batch* b, const ne_batch_tuning* t) {
    if (!b || !t) return NE_ERR_NULL;
    int r = check_tuning(*t);
    if (r) return r;
    b->tuning = *t;
    drop_graph(b);
    return NE_OK;
}

int ne_batch_info(const ne_batch* b, int* members, int64_t* cells, int* stage_items, int* surface_items) {
    if (!b) return NE_ERR_NULL;
    if (members) *members = (int)b->members.size();
    if (cells) {
        int64_t n = 0;
        for (const ne_model* m : b->members) n += (int64_t)m->d.grid.Nx * m->d.grid.Ny * m->d.grid.Nz;
        *cells = n;
    }
    // counts of the current upload (the next step rebuilds the lists if membership changed)
    int ti = tr_index(b->tuning.tile_rows), si = sb_index(b->tuning.surface_block);
    if (stage_items) *stage_items = ti >= 0 ? b->stage_nwork[0][ti] + b->stage_nwork[1][ti] : 0;
    if (surface_items) *surface_items = si >= 0 ? b->surface_nwork[si] : 0;
    return NE_OK;
}

}  // extern "C"
