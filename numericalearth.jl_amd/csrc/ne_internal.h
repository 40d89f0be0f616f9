// ne_internal.h — host-side declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "ne_device.h"

struct ne_ctx {
    int device;
    ne_counters* counters;      // device
    double* scratch;            // device workspace (budget partials), grown on demand
    size_t scratch_bytes;
};

struct ne_model {
    ne_ctx* ctx;
    ne_model_desc d;
    std::vector<double> atimes, rtimes;
    void* T1 = nullptr;
    void* S1 = nullptr;
    void* T2 = nullptr;
    void* S2 = nullptr;
    ne_view vT1, vS1, vT2, vS2;
    // one byte per column, written by the step's last RK stage: 1 where a cell of the column is
    // below its freezing point (or not known), so the freeze clamp that follows visits only those
    uint8_t* freeze_flags = nullptr;
    // rounding residual of the FP32 state (residual_bytes): per tile, the running compensated
    // totals Σ V·(x − stored) of T and S (sum, compensation, sum, compensation)
    double* residual = nullptr;
    // columns the FP32 flux solve leaves to its second kernel (count, then one int per column)
    int* deferred = nullptr;
    // Beer's-law factors of TwoColorRadiation at every face (TracerForcing::beer), filled at creation
    double* beer = nullptr;
    double time = 0.0;
    double last_dt = 0.0;
    int ab2_parity = 0;                 // Quasi-Adams-Bashforth 2: which stage-buffer pair takes G
    int64_t iteration = 0;
    bool timing = false;
    int max_steps = 0;
    int recorded = 0;
    std::vector<hipEvent_t> events;
    bool owns_buffers = true;           // stage buffers allocated by the model (no workspace given)
    struct ne_batch* batch = nullptr;   // the batch this model belongs to, if any
    std::vector<int64_t> aslot, rslot;   // stream position held by each snapshot slot (windowed series)
};

namespace ne {

// Atmosphere source fields on the device, storage type Real.
template <typename Real>
struct AtmosDev {
    int Nx, Ny, tx, ty;
    View3<const Real> u, v, T, q, p, rain, snow;
    ne_time_interp time;
};

template <typename Real>
struct RadDev {
    int Nx, Ny, tx, ty;
    View3<const Real> sw, lw;
    ne_time_interp time;
};

struct FluxOut {
    View3<double> Qv, Qt, Jv, rtx, rty, us, ts, qs, Ts;
    View3<int32_t> iters;
};

struct NetOut {
    View3<double> tau_x, tau_y, J_T, J_S, J_w, J_H;
    View3<double> J_sw;     // TwoColorRadiation surface flux J₀ (written when two-colour)
};

// Tracer forcing of T: TwoColorRadiation of the surface flux J0 (radiative_forcing.jl:52-71).
struct TracerForcing {
    int two_color;
    double first_fraction, kappa1, kappa2;
    const double* J0;
    // freeze watch of the step's last stage (nullptr: none): flag[i + j·Nx] = 1 when a stored
    // cell of the column is below T_m = melt0 − liquidus_slope·S (the freeze clamp's test)
    uint8_t* freeze_flags;
    double melt0, liquidus_slope;
    // rounding residual of an FP32 state's last stage (nullptr: none): the stage forms its
    // update conservatively in FP64 and adds Σ V·(x − stored) of T and S over the tile's cells
    // to the tile's slot (the kernel offsets the pointer to it)
    double* residual;
    // Beer's-law factors of the two colours at every face (nullptr: computed in place):
    // beer[2k] = exp(κ₁(z_f[k] − η)), beer[2k+1] = exp(κ₂(z_f[k] − η)), k = 0 … Nz
    const double* beer;
};

inline TracerForcing tracer_forcing(const ne_params* p, const ne_net_ocean_fluxes& n) {
    TracerForcing f = {0, 0.0, 0.0, 0.0, nullptr, nullptr, 0.0, 0.0, nullptr, nullptr};
    if (p && p->radiative_forcing == NE_RADIATIVE_FORCING_TWO_COLOR) {
        f.two_color = 1;
        f.first_fraction = p->first_color_fraction;
        f.kappa1 = p->first_absorption_coefficient;
        f.kappa2 = p->second_absorption_coefficient;
        f.J0 = static_cast<const double*>(n.shortwave_surface_flux.data);
    }
    return f;
}

struct RadOut {
    View3<double> up_lw, down_lw, down_sw;
};

template <typename Real>
AtmosDev<Real> atmos_dev(const ne_prescribed_atmosphere& a);
template <typename Real>
RadDev<Real> rad_dev(const ne_prescribed_radiation& r);
FluxOut flux_out(const ne_interface_fluxes& f);
NetOut net_out(const ne_net_ocean_fluxes& n);
RadOut rad_out(const ne_radiative_fluxes& r);

// Surface launchers (ne_surface.hip). Real is the storage type of the ocean state
// and prescribed fields; fluxes are always double.
hipError_t launch_immersed_levels(const GridDev& g, View3<const double> bottom, int32_t* kmin, hipStream_t s);
hipError_t launch_log(const double* x, double* y, int64_t n, hipStream_t s);
hipError_t launch_atan_ge1(const double* x, double* y, int64_t n, hipStream_t s);
hipError_t launch_cbrt_ge1(const double* x, double* y, int64_t n, hipStream_t s);
hipError_t launch_sqrt(const double* x, double* y, int64_t n, hipStream_t s);
hipError_t launch_exp(const double* x, double* y, int64_t n, hipStream_t s);
hipError_t launch_fractional_indices_nodes(const GridDev& g, const double* xn, int nx, int tx, const double* yn,
                                          int ny, int ty, double* fi, double* fj, hipStream_t s);
hipError_t launch_fractional_indices(const GridDev& g, const ne_source_grid& src, double* fi, double* fj,
                                     hipStream_t s);
template <typename Real>
hipError_t launch_interp_atmos(const GridDev& g, const AtmosDev<Real>& a, const double* fi, const double* fj,
                               const ne_exchange_state& out, hipStream_t s);
template <typename Real>
hipError_t launch_interp_radiation(const GridDev& g, const RadDev<Real>& r, const double* fi, const double* fj,
                                   View3<double> sw, View3<double> lw, hipStream_t s);
hipError_t launch_flux_solve_exchange(const GridDev& g, const ne_params& p, const ne_exchange_state& a,
                                      const ne_ocean_state& o, const FluxOut& out, ne_counters* c,
                                      hipStream_t s);
hipError_t launch_freeze_clamp(const GridDev& g, const ne_params& p, View3<double> T, View3<const double> S,
                               View3<double> frazil, double dt, hipStream_t s);
hipError_t launch_assemble(const GridDev& g, const ne_params& p, const ne_interface_fluxes& f,
                           const ne_exchange_state& a, View3<const double> sst, const NetOut& n, hipStream_t s);
hipError_t launch_radiation(const GridDev& g, const ne_params& p, View3<const double> sw, View3<const double> lw,
                            View3<const double> Ts, const NetOut& n, const RadOut& r, hipStream_t s);

// Fused update_state: K1 interpolates, solves, clamps, assembles and adds radiation per
// column; the τ kernel then averages the momentum fluxes onto the faces.
template <typename Real>
hipError_t launch_surface_update(const GridDev& g, const ne_params& p, const AtmosDev<Real>& a, const double* fi,
                                 const double* fj, const RadDev<Real>& r, const double* rfi, const double* rfj,
                                 View3<Real> T, View3<const Real> S, View3<const Real> u, View3<const Real> v,
                                 const FluxOut& out, const NetOut& n, const RadOut& ro, View3<double> frazil,
                                 double dt_frazil, const uint8_t* freeze_flags, ne_counters* c, int* deferred,
                                 hipStream_t s);
// deferred (FP32 storage; nullptr: none): count + one int per column for the columns K1 leaves to
// its second kernel; the τ kernel, which runs after both, clears the count.
hipError_t launch_surface_tau(const GridDev& g, const ne_params& p, const FluxOut& f, const NetOut& n,
                              int* deferred, hipStream_t s);

// Tracer launchers (ne_tracer.hip).
// stochastic: round the new state stochastically (FP32 storage only; the stage that writes
// the model state), seeded by the model time at the start of the step.
template <typename Real>
hipError_t launch_rk_stage(const GridDev& g, View3<const Real> u, View3<const Real> v, View3<const Real> w,
                           View3<const Real> T0, View3<const Real> S0, View3<const Real> Tin,
                           View3<const Real> Sin, View3<Real> Tout, View3<Real> Sout, View3<const double> JT,
                           View3<const double> JS, View3<const double> Jw, TracerForcing forcing, double beta_dt,
                           bool stochastic, double time, hipStream_t s);

// Seed of a step's stochastic rounding: a hash of the bits of the model time.
__host__ __device__ inline uint32_t step_seed(double time) {
    union {
        double d;
        uint64_t u;
    } b;
    b.d = time;
    uint32_t lo = (uint32_t)b.u, hi = (uint32_t)(b.u >> 32);
    uint32_t x = lo * 0x9e3779b1u ^ (hi + 0x7f4a7c15u);
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}
template <typename Real>
hipError_t launch_tendency(const GridDev& g, View3<const Real> u, View3<const Real> v, View3<const Real> w,
                           View3<const Real> T, View3<const Real> S, View3<const double> JT,
                           View3<const double> JS, View3<const double> Jw, TracerForcing forcing, View3<Real> GT,
                           View3<Real> GS, hipStream_t s);
// Quasi-Adams–Bashforth 2 update of an FP64 state from the step's tendency G and the last
// step's G⁻ (not read when euler), with the freeze flags of the step (ne_tracer.hip).
hipError_t launch_ab2_update(const GridDev& g, View3<double> T, View3<double> S, View3<const double> GT,
                             View3<const double> GS, View3<const double> GmT, View3<const double> GmS, double dt,
                             double chi, bool euler, uint8_t* freeze_flags, double melt0, double slope,
                             hipStream_t s);
// Vertical velocity from continuity (all Nz+1 faces of w).
template <typename Real>
hipError_t launch_w_continuity(const GridDev& g, View3<const Real> u, View3<const Real> v, View3<Real> w,
                               hipStream_t s);
template <typename Real>
hipError_t launch_budget(ne_ctx* ctx, const GridDev& g, View3<const Real> c, double weight, double* out,
                         hipStream_t s);
// Allocation-free budget: the address of the library's fixed partials array on `device`
// (looked up once per device by budget_noalloc_init, with that device current).
hipError_t budget_noalloc_init(int device);
double* budget_noalloc_scratch(int device);
int64_t budget_blocks(const GridDev& g);
hipError_t launch_budget_noalloc(double* partial, const GridDev& g, View3<const double> c, double weight,
                                 double* out, hipStream_t s);
// Rounding residual of an FP32 state: one slot per tile (T sum, T compensation, S sum,
// S compensation), into which the block of the step's last stage adds the tile's Σ V·(x −
// stored); the number of slots a model keeps, and their fixed-order sum into out[4].
constexpr int RESIDUAL_SLOT = 4;
// Beer's-law factors of the two colours at the Nz+1 faces (TracerForcing::beer), as the stage
// kernel would compute them in place.
hipError_t launch_beer_factors(const GridDev& g, double kappa1, double kappa2, double* beer, hipStream_t s);
// Halo of a cell-centred field (ne_fill_halo_*): wrapped in periodic directions, the nearest
// interior value in bounded ones and vertically.
template <typename Real>
hipError_t launch_halo_fill(const GridDev& g, View3<Real> c, int Hx, int Hy, int Hz, hipStream_t s);
// Traffic probe (ne_stream_probe_f64): n_in ≤ NE_STREAM_PROBE_MAX_IN arrays read, n_out written.
hipError_t launch_stream_probe(const double* const* in, int n_in, double* const* out, int n_out, int64_t n,
                               hipStream_t s);
int64_t residual_slots(const GridDev& g);
hipError_t launch_residual_reduce(const double* residual, const GridDev& g, double* out, hipStream_t s);

// ---------------------------------------------------------------------------
// Batched ensemble: one device record per member model, read by persistent kernels
// that walk a work list of (member, tile) or (member, first column) items.
// ---------------------------------------------------------------------------

// Byte strides of the views a tracer stage reads and writes. The six tracer views (T⁰, S⁰,
// c_in, c_out of T and S) share one layout, as do the three 2-D surface flux views; every
// view spans less than 4 GiB, so a byte offset fits 32 bits (checked on the host).
struct TracerLayout {
    uint32_t c_sy, c_sz, u_sy, u_sz, v_sy, v_sz, w_sy, w_sz, j_sy;
};

template <typename Real>
struct BatchRec {
    GridDev g;
    ne_params p;
    AtmosDev<Real> a;
    RadDev<Real> r;
    const double *afi, *afj, *rfi, *rfj;
    View3<Real> T, S, T1, S1, T2, S2;
    View3<const Real> u, v, w;
    FluxOut f;
    NetOut n;
    RadOut ro;
    View3<double> frazil;
    uint8_t* freeze_flags;  // the member's freeze flags (written by stage 3, read by K2)
    double* residual;       // the member's rounding residual (running totals per tile)
    const double* beer;     // the member's Beer's-law factors (two-colour forcing)
    const double* atimes;   // device copies of the snapshot times
    const double* rtimes;
    int nat, nrt;
    int time_indexing;      // NE_TIME_CLAMPED or NE_TIME_CYCLICAL
    double time_period;
    int fast;               // flux path (flux_path): 0 general, 1 default Edson constants, 2 also default constants
    double time;           // device clock of the member (advanced by the clock kernel)
};

template <typename Real>
__host__ __device__ inline TracerLayout tracer_layout(const BatchRec<Real>& r) {
    TracerLayout L;
    L.c_sy = (uint32_t)(r.T.sy * (int64_t)sizeof(Real));
    L.c_sz = (uint32_t)(r.T.sz * (int64_t)sizeof(Real));
    L.u_sy = (uint32_t)(r.u.sy * (int64_t)sizeof(Real));
    L.u_sz = (uint32_t)(r.u.sz * (int64_t)sizeof(Real));
    L.v_sy = (uint32_t)(r.v.sy * (int64_t)sizeof(Real));
    L.v_sz = (uint32_t)(r.v.sz * (int64_t)sizeof(Real));
    L.w_sy = (uint32_t)(r.w.sy * (int64_t)sizeof(Real));
    L.w_sz = (uint32_t)(r.w.sz * (int64_t)sizeof(Real));
    L.j_sy = (uint32_t)(r.n.J_T.sy * (int64_t)sizeof(double));
    return L;
}

struct WorkItem {
    int member;
    int index;              // tile index (tracer stages) or first flattened column (surface)
};

// Persistent-grid geometry of one launch: `blocks` work-groups (a multiple of the 8 XCDs)
// deal the work list out in 8 contiguous chunks, one per XCD.
template <typename Real>
hipError_t launch_batch_stage(const BatchRec<Real>* recs, const WorkItem* work, int nwork, int tile_rows,
                              int blocks, int stage, double beta_dt, bool two_color, hipStream_t s);
template <typename Real>
hipError_t launch_batch_surface_solve(const BatchRec<Real>* recs, const WorkItem* work, int nwork, int block,
                                      int blocks, int path, bool fused, double dt_frazil, ne_counters* c,
                                      hipStream_t s);
template <typename Real>
hipError_t launch_batch_surface_tau(const BatchRec<Real>* recs, const WorkItem* work, int nwork, int block,
                                   int blocks, hipStream_t s);
int batch_tiles_x(int Nx);
int batch_tiles_y(int Ny, int tile_rows);

}  // namespace ne
