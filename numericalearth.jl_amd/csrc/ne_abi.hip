// ne_abi.hip — extern "C" entry points of include/numericalearth_hip.h.
#include <math.h>
#include <cmath>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "ne_internal.h"
#include "numericalearth_hip_testing.h"

using namespace ne;

namespace {

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

inline int dev_err(hipError_t e) { return e == hipSuccess ? NE_OK : NE_ERR_DEVICE; }

int check_grid(const ne_grid* g) {
    if (!g) return NE_ERR_NULL;
    if (g->Nx <= 0 || g->Ny <= 0 || g->Nz <= 0) return NE_ERR_SHAPE;
    if (g->Nz > NE_MAX_LEVELS) return NE_ERR_UNSUPPORTED;
    if (g->topology_x != NE_PERIODIC && g->topology_x != NE_BOUNDED) return NE_ERR_UNSUPPORTED;
    if (g->topology_y != NE_PERIODIC && g->topology_y != NE_BOUNDED) return NE_ERR_UNSUPPORTED;
    if (!g->dz || !g->z_center || !g->dy_xface || !g->dx_yface || !g->Az || !g->lambda_c || !g->phi_c || !g->kmin)
        return NE_ERR_NULL;
    return NE_OK;
}

int check_view3(const ne_view& v) {
    if (!v.data) return NE_ERR_NULL;
    if (v.sx != 1) return NE_ERR_SHAPE;
    return NE_OK;
}

int check_view2(const ne_view& v) { return v.data ? NE_OK : NE_ERR_NULL; }

int check_source(const ne_source_grid& s) {
    if (s.Nx <= 0 || s.Ny <= 0) return NE_ERR_SHAPE;
    if (s.topology_x != NE_FLAT && !(s.dlambda > 0)) return NE_ERR_SHAPE;
    if (s.topology_y != NE_FLAT && !(s.dphi > 0)) return NE_ERR_SHAPE;
    return NE_OK;
}

int check_atmos(const ne_prescribed_atmosphere* a) {
    if (!a) return NE_ERR_NULL;
    int r = check_source(a->grid);
    if (r) return r;
    if (!a->u.data || !a->v.data || !a->T.data || !a->q.data || !a->p.data) return NE_ERR_NULL;
    if (a->time.n1 < 0 || a->time.n2 < 0) return NE_ERR_SHAPE;
    return NE_OK;
}

int check_rad(const ne_prescribed_radiation* r) {
    if (!r) return NE_ERR_NULL;
    int s = check_source(r->grid);
    if (s) return s;
    if (!r->downwelling_shortwave.data || !r->downwelling_longwave.data) return NE_ERR_NULL;
    return NE_OK;
}

int check_ocean(const ne_ocean_state* o) {
    if (!o) return NE_ERR_NULL;
    int r;
    if ((r = check_view3(o->u)) || (r = check_view3(o->v)) || (r = check_view3(o->w)) || (r = check_view3(o->T)) ||
        (r = check_view3(o->S)))
        return r;
    return NE_OK;
}

int check_fluxes(const ne_interface_fluxes* f) {
    if (!f) return NE_ERR_NULL;
    const ne_view* v[] = {&f->latent_heat, &f->sensible_heat, &f->water_vapor, &f->x_momentum, &f->y_momentum,
                          &f->friction_velocity, &f->temperature_scale, &f->water_vapor_scale,
                          &f->interface_temperature};
    for (const ne_view* x : v)
        if (!x->data) return NE_ERR_NULL;
    return NE_OK;
}

int check_net(const ne_net_ocean_fluxes* n) {
    if (!n) return NE_ERR_NULL;
    const ne_view* v[] = {&n->tau_x, &n->tau_y, &n->J_T, &n->J_S, &n->J_w, &n->J_H};
    for (const ne_view* x : v)
        if (!x->data) return NE_ERR_NULL;
    return NE_OK;
}

// A two-colour radiative forcing needs its surface-flux view.
int check_forcing(const ne_grid* g, const ne_params* p, const ne_net_ocean_fluxes* net) {
    if (!p || p->radiative_forcing == NE_RADIATIVE_FORCING_NONE) return NE_OK;
    if (p->radiative_forcing != NE_RADIATIVE_FORCING_TWO_COLOR) return NE_ERR_UNSUPPORTED;
    if (!g->z_face || !net->shortwave_surface_flux.data) return NE_ERR_NULL;
    return NE_OK;
}

int check_params(const ne_params* p) {
    if (!p) return NE_ERR_NULL;
    if (p->stop_criteria != NE_STOP_CONVERGENCE && p->stop_criteria != NE_STOP_FIXED_ITERATIONS)
        return NE_ERR_UNSUPPORTED;
    if (p->stability_functions != NE_STABILITY_NONE && p->stability_functions != NE_STABILITY_EDSON)
        return NE_ERR_UNSUPPORTED;
    if (p->maxiter < 0) return NE_ERR_SHAPE;
    if (p->turbulent_fluxes != NE_FLUXES_SIMILARITY_THEORY && p->turbulent_fluxes != NE_FLUXES_COEFFICIENT_BASED)
        return NE_ERR_UNSUPPORTED;
    if (p->transfer_coefficients != NE_TRANSFER_SCALES && p->transfer_coefficients != NE_TRANSFER_LARGE_YEAGER)
        return NE_ERR_UNSUPPORTED;
    for (int n = 0; n < 3; n++)
        if (p->transfer_kind[n] != NE_COEFFICIENT_CONSTANT && p->transfer_kind[n] != NE_COEFFICIENT_POLYNOMIAL)
            return NE_ERR_UNSUPPORTED;
    if (p->interface_temperature != NE_TEMPERATURE_BULK && p->interface_temperature != NE_TEMPERATURE_SKIN)
        return NE_ERR_UNSUPPORTED;
    return NE_OK;
}

ne_time_interp time_indices(const std::vector<double>& times, double t, const ne_model_desc& d) {
    return ne::time_indices(times.data(), (int)times.size(), t, d.time_indexing, d.time_period);
}

constexpr int64_t NO_POSITION = INT64_MIN;

int64_t floor_mod(int64_t a, int64_t n) {
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

// Stream positions of the two snapshots the interpolation at t reads: the snapshot index
// itself, or with cyclical indexing cycle·n + index, so that the wrap from the last snapshot
// to the first one of the next cycle reads two consecutive positions.
struct SeriesAt {
    ne_time_interp ti;
    int64_t p1, p2;
};

SeriesAt series_at(const double* times, int n, double t, int indexing, double period) {
    SeriesAt s;
    s.ti = ne::time_indices(times, n, t, indexing, period);
    s.p1 = s.ti.n1;
    s.p2 = s.ti.n2;
    if (indexing == NE_TIME_CYCLICAL && n > 1) {
        const double x = t - times[0];
        double tau = std::fmod(x, period);   // as in ne::time_indices
        if (tau != 0.0 && ((tau > 0.0) != (period > 0.0))) tau += period;
        const int64_t cycle = std::llround((x - tau) / period);
        s.p1 = cycle * n + s.ti.n1;
        s.p2 = (s.ti.n2 < s.ti.n1 ? cycle + 1 : cycle) * n + s.ti.n2;
    }
    return s;
}

SeriesAt series_at(const std::vector<double>& times, double t, const ne_model_desc& d) {
    return series_at(times.data(), (int)times.size(), t, d.time_indexing, d.time_period);
}

// Positions a freshly created window of `slots` holds: from the first position read at
// time t, N consecutive ones (kept inside the series with clamped indexing), as an in-memory
// window is filled from the snapshot the current time interpolates from.
void seed_window(std::vector<int64_t>& resident, const std::vector<double>& times, double t,
                 const ne_model_desc& d) {
    const int64_t N = (int64_t)resident.size(), nt = (int64_t)times.size();
    if (N == 0 || nt == 0) return;
    int64_t start = series_at(times, t, d).p1;
    int64_t count = N;
    if (d.time_indexing != NE_TIME_CYCLICAL) {
        start = std::max<int64_t>(0, std::min(start, nt - N));
        count = std::min(N, nt - start);
    }
    for (int64_t p = start; p < start + count; p++) resident[(size_t)floor_mod(p, N)] = p;
}

// Interpolation indices of a windowed series mapped to its slots (position p in slot
// p mod slots); NE_ERR_STATE if a snapshot the interpolation reads is not resident.
int to_slots(SeriesAt& s, int slots, const std::vector<int64_t>& resident) {
    if (slots <= 0) return NE_OK;
    const int s1 = (int)floor_mod(s.p1, slots), s2 = (int)floor_mod(s.p2, slots);
    if (resident[s1] != s.p1 || resident[s2] != s.p2) return NE_ERR_STATE;
    s.ti.n1 = s1;
    s.ti.n2 = s2;
    return NE_OK;
}

// The model's snapshot views at model time t (windowed series mapped to their slots); with
// apply == false only whether they are resident.
int model_time_views(ne_model* m, double t, bool apply = true) {
    ne_model_desc& d = m->d;
    SeriesAt a = series_at(m->atimes, t, d), r = series_at(m->rtimes, t, d);
    int e = to_slots(a, d.atmosphere_slots, m->aslot);
    if (!e && d.params.radiation) e = to_slots(r, d.radiation_slots, m->rslot);
    if (e || !apply) return e;
    d.atmosphere.time = a.ti;
    d.radiation.time = r.ti;
    return NE_OK;
}

// Bytes spanned by a view of nx × ny × nz elements (0 if a stride is not positive).
int64_t extent_bytes(const ne_view& v, int64_t nx, int64_t ny, int64_t nz, int64_t esize) {
    if (v.sx <= 0 || v.sy <= 0 || (nz > 1 && v.sz <= 0)) return 0;
    return ((nx - 1) * v.sx + (ny - 1) * v.sy + (nz - 1) * (nz > 1 ? v.sz : 0) + 1) * esize;
}

constexpr int64_t MAX_VIEW_BYTES = (int64_t)1 << 32;

// The views a tracer stage reads and writes: the tracer views (T⁰, S⁰, c_in, c_out) share
// one layout, the 2-D surface fluxes share one, and every view spans < 4 GiB (the stage
// kernel addresses each view with 32-bit byte offsets).
int check_tracer_views(const ne_grid* g, const ne_view* const* tracers, int ntr, const ne_view& u, const ne_view& v,
                       const ne_view& w, const ne_net_ocean_fluxes* net, int64_t esize) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Nz = g->Nz;
    const ne_view& c = *tracers[0];
    for (int n = 0; n < ntr; n++)
        if (tracers[n]->sx != 1 || tracers[n]->sy != c.sy || tracers[n]->sz != c.sz) return NE_ERR_SHAPE;
    const ne_view* J[4] = {&net->J_T, &net->J_S, &net->J_w, &net->shortwave_surface_flux};
    for (const ne_view* j : J)
        if (j->data && (j->sx != 1 || j->sy != J[0]->sy)) return NE_ERR_SHAPE;
    int64_t ext[5] = {extent_bytes(c, Nx, Ny, Nz, esize), extent_bytes(u, Nx + 1, Ny, Nz, esize),
                      extent_bytes(v, Nx, Ny + 1, Nz, esize), extent_bytes(w, Nx, Ny, Nz + 1, esize),
                      extent_bytes(*J[0], Nx, Ny, 1, 8)};
    for (int64_t e : ext) {
        if (e <= 0) return NE_ERR_SHAPE;
        if (e >= MAX_VIEW_BYTES) return NE_ERR_CAPACITY;
    }
    return NE_OK;
}

// One RK stage buffer (T or S) of a model: the layout of the model's T view, rounded up
// to 256 bytes.
size_t stage_buffer_bytes(const ne_model_desc& d) {
    int64_t esize = d.precision == NE_FP64 ? sizeof(double) : sizeof(float);
    int64_t n = extent_bytes(d.ocean.T, d.grid.Nx, d.grid.Ny, d.grid.Nz, esize);
    return ((size_t)(n > 0 ? n : 0) + 255) & ~(size_t)255;
}

// The model's freeze flags: one byte per column, rounded up to 256 bytes.
size_t freeze_flag_bytes(const ne_model_desc& d) {
    return ((size_t)d.grid.Nx * (size_t)d.grid.Ny + 255) & ~(size_t)255;
}

// The model's rounding residual (totals and tile partials), rounded up to 256 bytes.
size_t residual_bytes(const ne_model_desc& d) {
    return ((size_t)residual_slots(grid_dev(d.grid)) * RESIDUAL_SLOT * sizeof(double) + 255) & ~(size_t)255;
}

// The deferred-column list of the FP32 flux solve: a count and one int per column, rounded up to
// 256 bytes.
size_t deferred_bytes(const ne_model_desc& d) {
    return ((1 + (size_t)d.grid.Nx * (size_t)d.grid.Ny) * sizeof(int) + 255) & ~(size_t)255;
}

// The Beer's-law factors of the two-colour forcing at the Nz+1 faces, rounded up to 256 bytes.
size_t beer_bytes(const ne_model_desc& d) {
    return (2 * ((size_t)d.grid.Nz + 1) * sizeof(double) + 255) & ~(size_t)255;
}

// Device workspace of a model: the four RK stage buffers, the freeze flags, the rounding residual,
// the deferred-column list, the Beer's-law factors.
size_t model_workspace_bytes(const ne_model_desc& d) {
    return 4 * stage_buffer_bytes(d) + freeze_flag_bytes(d) + residual_bytes(d) + deferred_bytes(d) +
           beer_bytes(d);
}

}  // namespace

extern "C" {

int ne_version(void) { return NE_ABI_VERSION; }

int ne_params_default(ne_params* p) {
    if (!p) return NE_ERR_NULL;
    memset(p, 0, sizeof(*p));
    set_default_surface_constants(*p);   // roughness, gustiness, stop criterion, thermodynamics
    const double pi = 3.14159265358979323846;
    p->turbulent_prandtl_number = 1.0;
    p->stability_functions = NE_STABILITY_EDSON;
    p->similarity_form = NE_SIMILARITY_LOG;
    const double em[11] = {50.0, 0.35, 0.7, 0.75, 5.0 / 0.35, 15.0, 2.0, pi / 2.0, 10.15, 3.0, pi / sqrt(3.0)};
    const double es[12] = {50.0, 0.35, 2.0 / 3.0, 3.0 / 2.0, 14.28, 8.525, 15.0, 2.0, 0.0, 34.15, 3.0, pi / sqrt(3.0)};
    memcpy(p->edson_momentum, em, sizeof(em));
    memcpy(p->edson_scalar, es, sizeof(es));
    p->ocean_reference_density = 1020.0;
    p->ocean_heat_capacity = 3991.86795711963;
    p->freshwater_density = 1000.0;
    p->radiation = 1;
    p->stefan_boltzmann_constant = 5.670374419e-8;
    p->ocean_albedo = 0.05;
    p->ocean_emissivity = 0.97;
    p->freshwater_melting_temperature = 0.0;
    p->liquidus_slope = 0.054;
    p->radiative_forcing = NE_RADIATIVE_FORCING_NONE;
    p->first_color_fraction = 0.58;
    p->first_absorption_coefficient = 1.0 / 0.35;
    p->second_absorption_coefficient = 1.0 / 23.0;
    // CoefficientBasedFluxes / LargeYeagerTransferCoefficients / SkinTemperature defaults
    // (coefficient_based_turbulent_fluxes.jl:30-37, :92-99, :215-219; interface_states.jl:364)
    p->turbulent_fluxes = NE_FLUXES_SIMILARITY_THEORY;
    p->transfer_coefficients = NE_TRANSFER_SCALES;
    for (int n = 0; n < 3; n++) {
        p->transfer_kind[n] = NE_COEFFICIENT_CONSTANT;
        p->transfer_coefficient[n] = 1e-3;
    }
    p->interface_temperature = NE_TEMPERATURE_BULK;
    p->drag_polynomial[0] = 2.7;
    p->drag_polynomial[1] = 0.142;
    p->drag_polynomial[2] = 1.0 / 13.09;
    p->drag_polynomial[3] = 3.14807e-10;
    p->high_wind_speed_threshold = 33.0;
    p->high_wind_drag_coefficient = 2.34e-3;
    p->minimum_wind_speed = 0.5;
    p->large_yeager_von_karman_constant = 0.4;
    p->large_yeager_reference_height = 10.0;
    p->stable_heat_transfer_coefficient = 18.0;
    p->unstable_heat_transfer_coefficient = 32.7;
    p->moisture_transfer_coefficient = 34.6;
    p->paulson_momentum[0] = 16.0;
    p->paulson_momentum[1] = 3.14159265358979323846 / 2.0;
    p->paulson_scalar = 16.0;
    p->linear_stable_momentum[0] = 5.0;
    p->linear_stable_momentum[1] = 10.0;
    p->linear_stable_scalar[0] = 5.0;
    p->linear_stable_scalar[1] = 10.0;
    p->skin_diffusivity = 0.0;
    p->skin_depth = 0.0;
    p->skin_max_temperature_change = 5.0;
    p->reference_sublimation_enthalpy = 2834400.0;
    p->water_ice_heat_capacity = 2100.0;
    return NE_OK;
}

int ne_ctx_create(int device, ne_ctx** out) {
    if (!out) return NE_ERR_NULL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return NE_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return NE_ERR_DEVICE;
    ne_ctx* c = new (std::nothrow) ne_ctx();
    if (!c) return NE_ERR_CAPACITY;
    c->device = device;
    c->scratch = nullptr;
    c->scratch_bytes = 0;
    if (hipMalloc((void**)&c->counters, sizeof(ne_counters)) != hipSuccess) { delete c; return NE_ERR_DEVICE; }
    if (hipMemset(c->counters, 0, sizeof(ne_counters)) != hipSuccess ||
        budget_noalloc_init(device) != hipSuccess) {   // once per device, with `device` current
        (void)hipFree(c->counters);
        delete c;
        return NE_ERR_DEVICE;
    }
    *out = c;
    return NE_OK;
}

int ne_ctx_destroy(ne_ctx* c) {
    if (!c) return NE_ERR_NULL;
    (void)hipSetDevice(c->device);
    int r = NE_OK;
    if (c->counters && hipFree(c->counters) != hipSuccess) r = NE_ERR_DEVICE;
    if (c->scratch && hipFree(c->scratch) != hipSuccess) r = NE_ERR_DEVICE;
    delete c;
    return r;
}

int ne_ctx_counters(ne_ctx* c, ne_counters* out, int reset, void* stream) {
    if (!c || !out) return NE_ERR_NULL;
    hipStream_t s = as_stream(stream);
    if (hipMemcpyAsync(out, c->counters, sizeof(ne_counters), hipMemcpyDeviceToHost, s) != hipSuccess)
        return NE_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return NE_ERR_DEVICE;
    if (reset) {
        if (hipMemsetAsync(c->counters, 0, sizeof(ne_counters), s) != hipSuccess) return NE_ERR_DEVICE;
        if (hipStreamSynchronize(s) != hipSuccess) return NE_ERR_DEVICE;
    }
    return NE_OK;
}

int ne_grid_immersed_levels(ne_ctx* c, const ne_grid* g, ne_view bottom, int32_t* kmin, void* stream) {
    if (!c || !g || !kmin) return NE_ERR_NULL;
    if (g->Nx <= 0 || g->Ny <= 0 || g->Nz <= 0) return NE_ERR_SHAPE;
    if (g->Nz > NE_MAX_LEVELS) return NE_ERR_UNSUPPORTED;
    if (!g->z_center || !bottom.data) return NE_ERR_NULL;
    GridDev d = grid_dev(*g);
    return dev_err(launch_immersed_levels(d, view<const double>(bottom), kmin, as_stream(stream)));
}

int ne_fractional_indices_f64(ne_ctx* c, const ne_grid* g, const ne_source_grid* src, double* fi, double* fj,
                              void* stream) {
    if (!c || !g || !src || !fi || !fj) return NE_ERR_NULL;
    if (!g->lambda_c || !g->phi_c) return NE_ERR_NULL;
    if (g->Nx <= 0 || g->Ny <= 0) return NE_ERR_SHAPE;
    int r = check_source(*src);
    if (r) return r;
    return dev_err(launch_fractional_indices(grid_dev(*g), *src, fi, fj, as_stream(stream)));
}

int ne_log_f64(ne_ctx* c, const double* x, double* y, int64_t n, void* stream) {
    if (!c || !x || !y) return NE_ERR_NULL;
    if (n < 0 || n > ((int64_t)1 << 38)) return NE_ERR_SHAPE;
    if (n == 0) return NE_OK;
    return dev_err(launch_log(x, y, n, as_stream(stream)));
}

int ne_atan_ge1_f64(ne_ctx* c, const double* x, double* y, int64_t n, void* stream) {
    if (!c || !x || !y) return NE_ERR_NULL;
    if (n < 0 || n > ((int64_t)1 << 38)) return NE_ERR_SHAPE;
    if (n == 0) return NE_OK;
    return dev_err(launch_atan_ge1(x, y, n, as_stream(stream)));
}

int ne_exp_f64(ne_ctx* c, const double* x, double* y, int64_t n, void* stream) {
    if (!c || !x || !y) return NE_ERR_NULL;
    if (n < 0 || n > ((int64_t)1 << 38)) return NE_ERR_SHAPE;
    if (n == 0) return NE_OK;
    return dev_err(launch_exp(x, y, n, as_stream(stream)));
}

int ne_sqrt_f64(ne_ctx* c, const double* x, double* y, int64_t n, void* stream) {
    if (!c || !x || !y) return NE_ERR_NULL;
    if (n < 0 || n > ((int64_t)1 << 38)) return NE_ERR_SHAPE;
    if (n == 0) return NE_OK;
    return dev_err(launch_sqrt(x, y, n, as_stream(stream)));
}

int ne_cbrt_ge1_f64(ne_ctx* c, const double* x, double* y, int64_t n, void* stream) {
    if (!c || !x || !y) return NE_ERR_NULL;
    if (n < 0 || n > ((int64_t)1 << 38)) return NE_ERR_SHAPE;
    if (n == 0) return NE_OK;
    return dev_err(launch_cbrt_ge1(x, y, n, as_stream(stream)));
}

int ne_stream_probe_f64(ne_ctx* c, const double* const* in, int32_t n_in, double* const* out, int32_t n_out,
                        int64_t n, void* stream) {
    if (!c || !in || !out) return NE_ERR_NULL;
    if (n_in < 0 || n_in > NE_STREAM_PROBE_MAX_IN || n_out < 1 || n_out > NE_STREAM_PROBE_MAX_OUT || n < 0)
        return NE_ERR_SHAPE;
    for (int k = 0; k < n_in; k++)
        if (!in[k]) return NE_ERR_NULL;
    for (int j = 0; j < n_out; j++)
        if (!out[j]) return NE_ERR_NULL;
    if (n == 0) return NE_OK;
    return dev_err(launch_stream_probe(in, n_in, out, n_out, n, as_stream(stream)));
}

int ne_fractional_indices_nodes_f64(ne_ctx* c, const ne_grid* g, const double* lambda_nodes, int32_t nx,
                                    int32_t topology_x, const double* phi_nodes, int32_t ny, int32_t topology_y,
                                    double* fi, double* fj, void* stream) {
    if (!c || !g || !lambda_nodes || !phi_nodes || !fi || !fj) return NE_ERR_NULL;
    if (!g->lambda_c || !g->phi_c) return NE_ERR_NULL;
    if (g->Nx <= 0 || g->Ny <= 0 || nx <= 0 || ny <= 0) return NE_ERR_SHAPE;
    for (int32_t t : {topology_x, topology_y})
        if (t != NE_PERIODIC && t != NE_BOUNDED && t != NE_FLAT) return NE_ERR_UNSUPPORTED;
    return dev_err(launch_fractional_indices_nodes(grid_dev(*g), lambda_nodes, nx, topology_x, phi_nodes, ny,
                                                   topology_y, fi, fj, as_stream(stream)));
}

int ne_interp_atmos_f64(ne_ctx* c, const ne_grid* g, const ne_prescribed_atmosphere* a, const double* fi,
                        const double* fj, const ne_exchange_state* out, void* stream) {
    if (!c || !out || !fi || !fj) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_atmos(a))) return r;
    if (!out->u.data || !out->v.data || !out->T.data || !out->p.data || !out->q.data) return NE_ERR_NULL;
    return dev_err(launch_interp_atmos<double>(grid_dev(*g), atmos_dev<double>(*a), fi, fj, *out, as_stream(stream)));
}

int ne_interp_radiation_f64(ne_ctx* c, const ne_grid* g, const ne_prescribed_radiation* rad, const double* fi,
                            const double* fj, ne_view sw, ne_view lw, void* stream) {
    if (!c || !fi || !fj) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_rad(rad))) return r;
    if (!sw.data || !lw.data) return NE_ERR_NULL;
    return dev_err(launch_interp_radiation<double>(grid_dev(*g), rad_dev<double>(*rad), fi, fj, view<double>(sw),
                                                   view<double>(lw), as_stream(stream)));
}

int ne_ao_flux_solve_f64(ne_ctx* c, const ne_grid* g, const ne_params* p, const ne_exchange_state* a,
                         const ne_ocean_state* o, const ne_interface_fluxes* out, void* stream) {
    if (!c || !a) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_params(p)) || (r = check_ocean(o)) || (r = check_fluxes(out))) return r;
    if (!a->u.data || !a->v.data || !a->T.data || !a->p.data || !a->q.data) return NE_ERR_NULL;
    return dev_err(launch_flux_solve_exchange(grid_dev(*g), *p, *a, *o, flux_out(*out), c->counters,
                                              as_stream(stream)));
}

int ne_freeze_clamp_f64(ne_ctx* c, const ne_grid* g, const ne_params* p, ne_view T, ne_view S, ne_view frazil,
                        double dt, void* stream) {
    if (!c) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_params(p)) || (r = check_view3(T)) || (r = check_view3(S)) ||
        (r = check_view2(frazil)))
        return r;
    return dev_err(launch_freeze_clamp(grid_dev(*g), *p, view<double>(T), view<const double>(S),
                                       view<double>(frazil), dt, as_stream(stream)));
}

int ne_assemble_net_ocean_fluxes_f64(ne_ctx* c, const ne_grid* g, const ne_params* p,
                                     const ne_interface_fluxes* f, const ne_exchange_state* a, ne_view sst,
                                     const ne_net_ocean_fluxes* out, void* stream) {
    if (!c || !a) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_params(p)) || (r = check_fluxes(f)) || (r = check_net(out)) ||
        (r = check_view3(sst)))
        return r;
    return dev_err(launch_assemble(grid_dev(*g), *p, *f, *a, view<const double>(sst), net_out(*out),
                                   as_stream(stream)));
}

int ne_apply_air_sea_radiation_f64(ne_ctx* c, const ne_grid* g, const ne_params* p, ne_view sw, ne_view lw,
                                   ne_view Ts, const ne_net_ocean_fluxes* net, const ne_radiative_fluxes* out,
                                   void* stream) {
    if (!c || !out) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_params(p)) || (r = check_net(net)) || (r = check_view2(sw)) ||
        (r = check_view2(lw)) || (r = check_view2(Ts)) || (r = check_forcing(g, p, net)))
        return r;
    return dev_err(launch_radiation(grid_dev(*g), *p, view<const double>(sw), view<const double>(lw),
                                    view<const double>(Ts), net_out(*net), rad_out(*out), as_stream(stream)));
}

int ne_tracer_tendency_weno5_f64(ne_ctx* c, const ne_grid* g, const ne_params* p, const ne_ocean_state* o,
                                 const ne_net_ocean_fluxes* net, ne_view GT, ne_view GS, void* stream) {
    if (!c) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_ocean(o)) || (r = check_net(net)) || (r = check_view3(GT)) ||
        (r = check_view3(GS)) || (r = check_forcing(g, p, net)))
        return r;
    const ne_view* tr[] = {&o->T, &o->S, &GT, &GS};
    if ((r = check_tracer_views(g, tr, 4, o->u, o->v, o->w, net, sizeof(double)))) return r;
    return dev_err(launch_tendency<double>(grid_dev(*g), view<const double>(o->u), view<const double>(o->v),
                                   view<const double>(o->w), view<const double>(o->T), view<const double>(o->S),
                                   view<const double>(net->J_T), view<const double>(net->J_S),
                                   view<const double>(net->J_w), tracer_forcing(p, *net), view<double>(GT),
                                   view<double>(GS), as_stream(stream)));
}

int ne_tracer_ab2_update_f64(ne_ctx* c, const ne_grid* g, ne_view T, ne_view S, ne_view GT, ne_view GS,
                             ne_view GmT, ne_view GmS, double dt, double chi, int32_t euler, void* stream) {
    if (!c) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g))) return r;
    const ne_view* vs[] = {&T, &S, &GT, &GS};
    for (const ne_view* x : vs)
        if ((r = check_view3(*x))) return r;
    if (!euler && ((r = check_view3(GmT)) || (r = check_view3(GmS)))) return r;
    if (!std::isfinite(dt) || !std::isfinite(chi)) return NE_ERR_SHAPE;
    // T and S are updated in place: neither may share storage with a tendency
    for (const void* q : {GT.data, GS.data, GmT.data, GmS.data})
        if (q && (q == T.data || q == S.data)) return NE_ERR_SHAPE;
    if (T.data == S.data) return NE_ERR_SHAPE;
    return dev_err(launch_ab2_update(grid_dev(*g), view<double>(T), view<double>(S), view<const double>(GT),
                                     view<const double>(GS), view<const double>(GmT), view<const double>(GmS), dt, chi,
                                     euler != 0, nullptr, 0.0, 0.0, as_stream(stream)));
}

int ne_tracer_rk_stage_f64(ne_ctx* c, const ne_grid* g, const ne_params* p, ne_view u, ne_view v, ne_view w,
                           ne_view T0, ne_view S0, ne_view Tin, ne_view Sin, ne_view Tout, ne_view Sout,
                           const ne_net_ocean_fluxes* net, double beta_dt, void* stream) {
    if (!c) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_net(net)) || (r = check_forcing(g, p, net))) return r;
    const ne_view* vs[] = {&u, &v, &w, &T0, &S0, &Tin, &Sin, &Tout, &Sout};
    for (const ne_view* x : vs)
        if ((r = check_view3(*x))) return r;
    if (Tin.data == Tout.data || Sin.data == Sout.data) return NE_ERR_SHAPE;
    const ne_view* tr[] = {&T0, &S0, &Tin, &Sin, &Tout, &Sout};
    if ((r = check_tracer_views(g, tr, 6, u, v, w, net, sizeof(double)))) return r;
    return dev_err(launch_rk_stage<double>(grid_dev(*g), view<const double>(u), view<const double>(v),
                                           view<const double>(w), view<const double>(T0), view<const double>(S0),
                                           view<const double>(Tin), view<const double>(Sin), view<double>(Tout),
                                           view<double>(Sout), view<const double>(net->J_T),
                                           view<const double>(net->J_S), view<const double>(net->J_w),
                                           tracer_forcing(p, *net), beta_dt, false, 0.0, as_stream(stream)));
}

// FP32 forms of the per-phase entry points (the mixed form's convention): FP32 prescribed series
// and 3-D fields, FP64 fractional indices, exchange state and 2-D fluxes; arithmetic as in the
// mixed model (FP32 WENO5, FP64-formed stage update rounded once to FP32).
int ne_interp_atmos_f32(ne_ctx* c, const ne_grid* g, const ne_prescribed_atmosphere* a, const double* fi,
                        const double* fj, const ne_exchange_state* out, void* stream) {
    if (!c || !out || !fi || !fj) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_atmos(a))) return r;
    if (!out->u.data || !out->v.data || !out->T.data || !out->p.data || !out->q.data) return NE_ERR_NULL;
    return dev_err(launch_interp_atmos<float>(grid_dev(*g), atmos_dev<float>(*a), fi, fj, *out, as_stream(stream)));
}

int ne_interp_radiation_f32(ne_ctx* c, const ne_grid* g, const ne_prescribed_radiation* rad, const double* fi,
                            const double* fj, ne_view sw, ne_view lw, void* stream) {
    if (!c || !fi || !fj) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_rad(rad))) return r;
    if (!sw.data || !lw.data) return NE_ERR_NULL;
    return dev_err(launch_interp_radiation<float>(grid_dev(*g), rad_dev<float>(*rad), fi, fj, view<double>(sw),
                                                  view<double>(lw), as_stream(stream)));
}

int ne_tracer_tendency_weno5_f32(ne_ctx* c, const ne_grid* g, const ne_params* p, const ne_ocean_state* o,
                                 const ne_net_ocean_fluxes* net, ne_view GT, ne_view GS, void* stream) {
    if (!c) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_ocean(o)) || (r = check_net(net)) || (r = check_view3(GT)) ||
        (r = check_view3(GS)) || (r = check_forcing(g, p, net)))
        return r;
    const ne_view* tr[] = {&o->T, &o->S, &GT, &GS};
    if ((r = check_tracer_views(g, tr, 4, o->u, o->v, o->w, net, sizeof(float)))) return r;
    return dev_err(launch_tendency<float>(grid_dev(*g), view<const float>(o->u), view<const float>(o->v),
                                          view<const float>(o->w), view<const float>(o->T), view<const float>(o->S),
                                          view<const double>(net->J_T), view<const double>(net->J_S),
                                          view<const double>(net->J_w), tracer_forcing(p, *net), view<float>(GT),
                                          view<float>(GS), as_stream(stream)));
}

int ne_tracer_rk_stage_f32(ne_ctx* c, const ne_grid* g, const ne_params* p, ne_view u, ne_view v, ne_view w,
                           ne_view T0, ne_view S0, ne_view Tin, ne_view Sin, ne_view Tout, ne_view Sout,
                           const ne_net_ocean_fluxes* net, double beta_dt, void* stream) {
    if (!c) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_net(net)) || (r = check_forcing(g, p, net))) return r;
    const ne_view* vs[] = {&u, &v, &w, &T0, &S0, &Tin, &Sin, &Tout, &Sout};
    for (const ne_view* x : vs)
        if ((r = check_view3(*x))) return r;
    if (Tin.data == Tout.data || Sin.data == Sout.data) return NE_ERR_SHAPE;
    const ne_view* tr[] = {&T0, &S0, &Tin, &Sin, &Tout, &Sout};
    if ((r = check_tracer_views(g, tr, 6, u, v, w, net, sizeof(float)))) return r;
    return dev_err(launch_rk_stage<float>(grid_dev(*g), view<const float>(u), view<const float>(v),
                                          view<const float>(w), view<const float>(T0), view<const float>(S0),
                                          view<const float>(Tin), view<const float>(Sin), view<float>(Tout),
                                          view<float>(Sout), view<const double>(net->J_T),
                                          view<const double>(net->J_S), view<const double>(net->J_w),
                                          tracer_forcing(p, *net), beta_dt, false, 0.0, as_stream(stream)));
}

int ne_compute_w_from_continuity_f64(ne_ctx* c, const ne_grid* g, ne_view u, ne_view v, ne_view w, void* stream) {
    if (!c) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_view3(u)) || (r = check_view3(v)) || (r = check_view3(w))) return r;
    if (w.data == u.data || w.data == v.data) return NE_ERR_SHAPE;
    return dev_err(launch_w_continuity<double>(grid_dev(*g), view<const double>(u), view<const double>(v),
                                               view<double>(w), as_stream(stream)));
}

}  // extern "C"

template <typename Real>
static int fill_halo(ne_ctx* c, const ne_grid* g, ne_view cv, int32_t Hx, int32_t Hy, int32_t Hz, void* stream) {
    if (!c) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_view3(cv))) return r;
    if (Hx < 0 || Hy < 0 || Hz < 0 || Hx > 64 || Hy > 64 || Hz > 64) return NE_ERR_SHAPE;
    if (Hx + Hy + Hz == 0) return NE_OK;
    return dev_err(launch_halo_fill<Real>(grid_dev(*g), view<Real>(cv), Hx, Hy, Hz, as_stream(stream)));
}

extern "C" {

int ne_fill_halo_f64(ne_ctx* c, const ne_grid* g, ne_view cv, int32_t Hx, int32_t Hy, int32_t Hz, void* stream) {
    return fill_halo<double>(c, g, cv, Hx, Hy, Hz, stream);
}

int ne_fill_halo_f32(ne_ctx* c, const ne_grid* g, ne_view cv, int32_t Hx, int32_t Hy, int32_t Hz, void* stream) {
    return fill_halo<float>(c, g, cv, Hx, Hy, Hz, stream);
}

int ne_budget_f64(ne_ctx* c, const ne_grid* g, ne_view cv, double weight, double* out, void* stream) {
    if (!c || !out) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_view3(cv))) return r;
    return dev_err(launch_budget<double>(c, grid_dev(*g), view<const double>(cv), weight, out, as_stream(stream)));
}

int ne_budget_f32(ne_ctx* c, const ne_grid* g, ne_view cv, double weight, double* out, void* stream) {
    if (!c || !out) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_view3(cv))) return r;
    return dev_err(launch_budget<float>(c, grid_dev(*g), view<const float>(cv), weight, out, as_stream(stream)));
}

int ne_budget_f64_noalloc(ne_ctx* c, const ne_grid* g, ne_view cv, double weight, double* out, void* stream) {
    if (!c || !out) return NE_ERR_NULL;
    int r;
    if ((r = check_grid(g)) || (r = check_view3(cv))) return r;
    const GridDev gd = grid_dev(*g);
    if (budget_blocks(gd) > NE_BUDGET_NOALLOC_MAX_BLOCKS) return NE_ERR_CAPACITY;
    double* partial = budget_noalloc_scratch(c->device);   // looked up by ne_ctx_create
    if (!partial) return NE_ERR_DEVICE;
    return dev_err(launch_budget_noalloc(partial, gd, view<const double>(cv), weight, out, as_stream(stream)));
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Model
// ---------------------------------------------------------------------------

// ev (optional): ev[3] is recorded before K1 by the caller, ev[4] after K1 and ev[5] after K2.
// after_step: the state was written by the last RK stage, whose freeze flags name the columns
// the freeze clamp has to visit; otherwise (construction, a caller's update_state) every column.
static int model_update_state(ne_model* m, double dt_frazil, hipStream_t s, hipEvent_t* ev = nullptr,
                              bool after_step = false) {
    ne_model_desc& d = m->d;
    GridDev g = grid_dev(d.grid);
    int tv = model_time_views(m, m->time);
    if (tv) return tv;
    FluxOut f = flux_out(d.fluxes);
    NetOut n = net_out(d.net);
    RadOut ro = rad_out(d.radiative);
    View3<double> frazil = view<double>(d.frazil_heat);
    const uint8_t* flags = after_step ? m->freeze_flags : nullptr;
    hipError_t e;
    if (d.precision == NE_FP64) {
        AtmosDev<double> a = atmos_dev<double>(d.atmosphere);
        RadDev<double> r = rad_dev<double>(d.radiation);
        e = launch_surface_update<double>(g, d.params, a, d.atmos_fi, d.atmos_fj, r, d.radiation_fi, d.radiation_fj,
                                          view<double>(d.ocean.T), view<const double>(d.ocean.S),
                                          view<const double>(d.ocean.u), view<const double>(d.ocean.v), f, n, ro,
                                          frazil, dt_frazil, flags, m->ctx->counters, nullptr, s);
    } else {
        AtmosDev<float> a = atmos_dev<float>(d.atmosphere);
        RadDev<float> r = rad_dev<float>(d.radiation);
        e = launch_surface_update<float>(g, d.params, a, d.atmos_fi, d.atmos_fj, r, d.radiation_fi, d.radiation_fj,
                                         view<float>(d.ocean.T), view<const float>(d.ocean.S),
                                         view<const float>(d.ocean.u), view<const float>(d.ocean.v), f, n, ro, frazil,
                                         dt_frazil, flags, m->ctx->counters, m->deferred, s);
    }
    if (e != hipSuccess) return NE_ERR_DEVICE;
    if (ev && hipEventRecord(ev[4], s) != hipSuccess) return NE_ERR_DEVICE;
    e = launch_surface_tau(g, d.params, f, n, d.precision == NE_FP64 ? nullptr : m->deferred, s);
    if (e == hipSuccess && ev && hipEventRecord(ev[5], s) != hipSuccess) return NE_ERR_DEVICE;
    return dev_err(e);
}

// One Quasi-Adams-Bashforth 2 step of an FP64 model: the tendency of the state into one pair of
// stage buffers, then the update with the last step's tendency from the other pair (not read on
// a forward-Euler step); the pairs swap roles every step. Events as the RK step's: ev[1] after
// the tendency, ev[2] and ev[3] after the update.
static int model_ab2_step(ne_model* m, double dt, hipStream_t s, hipEvent_t* ev) {
    ne_model_desc& d = m->d;
    GridDev g = grid_dev(d.grid);
    const bool euler = !(dt == m->last_dt);
    ne_view GT = m->ab2_parity ? m->vT2 : m->vT1, GS = m->ab2_parity ? m->vS2 : m->vS1;
    ne_view GmT = m->ab2_parity ? m->vT1 : m->vT2, GmS = m->ab2_parity ? m->vS1 : m->vS2;
    if (ev && hipEventRecord(ev[0], s) != hipSuccess) return NE_ERR_DEVICE;
    TracerForcing F = tracer_forcing(&d.params, d.net);
    hipError_t e = launch_tendency<double>(g, view<const double>(d.ocean.u), view<const double>(d.ocean.v),
                                           view<const double>(d.ocean.w), view<const double>(d.ocean.T),
                                           view<const double>(d.ocean.S), view<const double>(d.net.J_T),
                                           view<const double>(d.net.J_S), view<const double>(d.net.J_w), F,
                                           view<double>(GT), view<double>(GS), s);
    if (e != hipSuccess) return NE_ERR_DEVICE;
    if (ev && hipEventRecord(ev[1], s) != hipSuccess) return NE_ERR_DEVICE;
    e = launch_ab2_update(g, view<double>(d.ocean.T), view<double>(d.ocean.S), view<const double>(GT),
                          view<const double>(GS), view<const double>(GmT), view<const double>(GmS), dt, NE_AB2_CHI,
                          euler, m->freeze_flags, d.params.freshwater_melting_temperature, d.params.liquidus_slope, s);
    if (e != hipSuccess) return NE_ERR_DEVICE;
    if (ev && (hipEventRecord(ev[2], s) != hipSuccess || hipEventRecord(ev[3], s) != hipSuccess)) return NE_ERR_DEVICE;
    m->ab2_parity ^= 1;
    return NE_OK;
}

template <typename Real>
static int model_ocean_step(ne_model* m, double dt, hipStream_t s, hipEvent_t* ev = nullptr) {
    ne_model_desc& d = m->d;
    GridDev g = grid_dev(d.grid);
    View3<const Real> u = view<const Real>(d.ocean.u), v = view<const Real>(d.ocean.v),
                      w = view<const Real>(d.ocean.w);
    View3<const Real> T0 = view<const Real>(d.ocean.T), S0 = view<const Real>(d.ocean.S);
    View3<const double> JT = view<const double>(d.net.J_T), JS = view<const double>(d.net.J_S),
                        Jw = view<const double>(d.net.J_w);
    const double beta[3] = {1.0 / 3.0, 0.5, 1.0};
    hipError_t e;
    if (ev && hipEventRecord(ev[0], s) != hipSuccess) return NE_ERR_DEVICE;
    const bool sr = sizeof(Real) == 4;   // FP32 state: stochastic rounding of the stage that writes it
    TracerForcing F = tracer_forcing(&d.params, d.net);
    if (F.two_color) F.beer = m->beer;   // the factors filled at creation
    e = launch_rk_stage<Real>(g, u, v, w, T0, S0, T0, S0, view<Real>(m->vT1), view<Real>(m->vS1), JT, JS, Jw, F,
                              beta[0] * dt, false, m->time, s);
    if (e != hipSuccess) return NE_ERR_DEVICE;
    if (ev && hipEventRecord(ev[1], s) != hipSuccess) return NE_ERR_DEVICE;
    e = launch_rk_stage<Real>(g, u, v, w, T0, S0, view<const Real>(m->vT1), view<const Real>(m->vS1),
                              view<Real>(m->vT2), view<Real>(m->vS2), JT, JS, Jw, F, beta[1] * dt, false, m->time,
                              s);
    if (e != hipSuccess) return NE_ERR_DEVICE;
    if (ev && hipEventRecord(ev[2], s) != hipSuccess) return NE_ERR_DEVICE;
    TracerForcing F3 = F;   // the last stage flags the columns the freeze clamp visits
    F3.freeze_flags = m->freeze_flags;
    F3.melt0 = d.params.freshwater_melting_temperature;
    F3.liquidus_slope = d.params.liquidus_slope;
    if (sr) F3.residual = m->residual;   // and an FP32 state's rounding residual
    e = launch_rk_stage<Real>(g, u, v, w, T0, S0, view<const Real>(m->vT2), view<const Real>(m->vS2),
                              view<Real>(d.ocean.T), view<Real>(d.ocean.S), JT, JS, Jw, F3, beta[2] * dt, sr, m->time,
                              s);
    if (e != hipSuccess) return NE_ERR_DEVICE;
    if (ev && hipEventRecord(ev[3], s) != hipSuccess) return NE_ERR_DEVICE;
    return NE_OK;
}

extern "C" {

int ne_model_create(ne_ctx* c, const ne_model_desc* desc, ne_model** out, void* stream) {
    if (!c || !desc || !out) return NE_ERR_NULL;
    *out = nullptr;
    int r;
    if ((r = check_grid(&desc->grid)) || (r = check_params(&desc->params)) || (r = check_atmos(&desc->atmosphere)) ||
        (r = check_ocean(&desc->ocean)) || (r = check_fluxes(&desc->fluxes)) || (r = check_net(&desc->net)) ||
        (r = check_forcing(&desc->grid, &desc->params, &desc->net)))
        return r;
    if (desc->params.radiation && (r = check_rad(&desc->radiation))) return r;
    if (desc->precision != NE_FP64 && desc->precision != NE_MIXED_FP32) return NE_ERR_UNSUPPORTED;
    if (desc->fuse_freeze_clamp != 0) return NE_ERR_UNSUPPORTED;
    if (desc->timestepper != NE_TIMESTEPPER_SPLIT_RK3 && desc->timestepper != NE_TIMESTEPPER_QUASI_ADAMS_BASHFORTH2)
        return NE_ERR_UNSUPPORTED;
    if (desc->timestepper == NE_TIMESTEPPER_QUASI_ADAMS_BASHFORTH2 && desc->precision != NE_FP64)
        return NE_ERR_UNSUPPORTED;   // the mixed form's conservative FP32 stages are the RK step's
    if (!desc->atmos_fi || !desc->atmos_fj) return NE_ERR_NULL;
    if (desc->params.radiation && (!desc->radiation_fi || !desc->radiation_fj)) return NE_ERR_NULL;
    if (desc->n_atmosphere_times < 0 || desc->n_radiation_times < 0) return NE_ERR_SHAPE;
    if (desc->n_atmosphere_times > 0 && !desc->atmosphere_times) return NE_ERR_NULL;
    {
        const ne_view* tr[] = {&desc->ocean.T, &desc->ocean.S};
        int64_t esize = desc->precision == NE_FP64 ? sizeof(double) : sizeof(float);
        if ((r = check_tracer_views(&desc->grid, tr, 2, desc->ocean.u, desc->ocean.v, desc->ocean.w, &desc->net,
                                    esize)))
            return r;
    }
    if (desc->n_radiation_times > 0 && !desc->radiation_times) return NE_ERR_NULL;
    if (desc->time_indexing != NE_TIME_CLAMPED && desc->time_indexing != NE_TIME_CYCLICAL) return NE_ERR_UNSUPPORTED;
    if (desc->time_indexing == NE_TIME_CYCLICAL) {
        if (!(desc->time_period > 0.0)) return NE_ERR_SHAPE;
        for (int n = 0; n < 2; n++) {
            const double* t = n ? desc->radiation_times : desc->atmosphere_times;
            const int nt = n ? desc->n_radiation_times : desc->n_atmosphere_times;
            if (nt > 1 && !(desc->time_period >= t[nt - 1] - t[0])) return NE_ERR_SHAPE;
        }
    }
    if (desc->atmosphere_slots < 0 || desc->radiation_slots < 0) return NE_ERR_SHAPE;
    if (desc->atmosphere_slots == 1 || desc->radiation_slots == 1) return NE_ERR_SHAPE;   // two are read at once

    ne_model* m = new (std::nothrow) ne_model();
    if (!m) return NE_ERR_CAPACITY;
    m->ctx = c;
    m->d = *desc;
    m->atimes.assign(desc->atmosphere_times, desc->atmosphere_times + desc->n_atmosphere_times);
    m->rtimes.assign(desc->radiation_times, desc->radiation_times + desc->n_radiation_times);
    m->aslot.assign((size_t)desc->atmosphere_slots, NO_POSITION);
    m->rslot.assign((size_t)desc->radiation_slots, NO_POSITION);
    seed_window(m->aslot, m->atimes, m->time, *desc);
    seed_window(m->rslot, m->rtimes, m->time, *desc);
    const ne_grid& g = desc->grid;
    size_t n = stage_buffer_bytes(*desc);
    void** bufs[4] = {&m->T1, &m->S1, &m->T2, &m->S2};
    ne_view* views[4] = {&m->vT1, &m->vS1, &m->vT2, &m->vS2};
    if (desc->workspace) {
        if ((reinterpret_cast<uintptr_t>(desc->workspace) & 255) != 0) {   // its buffers start 256-aligned
            delete m;
            return NE_ERR_SHAPE;
        }
        if (desc->workspace_bytes < (int64_t)model_workspace_bytes(*desc)) {
            delete m;
            return NE_ERR_CAPACITY;
        }
        m->owns_buffers = false;
        for (int b = 0; b < 4; b++) *bufs[b] = static_cast<char*>(desc->workspace) + b * n;
        m->freeze_flags = reinterpret_cast<uint8_t*>(static_cast<char*>(desc->workspace) + 4 * n);
        m->residual = reinterpret_cast<double*>(static_cast<char*>(desc->workspace) + 4 * n +
                                                freeze_flag_bytes(*desc));
        m->deferred = reinterpret_cast<int*>(static_cast<char*>(desc->workspace) + 4 * n + freeze_flag_bytes(*desc) +
                                             residual_bytes(*desc));
        m->beer = reinterpret_cast<double*>(static_cast<char*>(desc->workspace) + 4 * n + freeze_flag_bytes(*desc) +
                                            residual_bytes(*desc) + deferred_bytes(*desc));
    } else if (hipMalloc(reinterpret_cast<void**>(&m->freeze_flags), freeze_flag_bytes(*desc)) != hipSuccess ||
               hipMalloc(reinterpret_cast<void**>(&m->residual), residual_bytes(*desc)) != hipSuccess ||
               hipMalloc(reinterpret_cast<void**>(&m->deferred), deferred_bytes(*desc)) != hipSuccess ||
               hipMalloc(reinterpret_cast<void**>(&m->beer), beer_bytes(*desc)) != hipSuccess) {
        ne_model_destroy(m);
        return NE_ERR_DEVICE;
    }
    // every column is visited until a step's last stage has flagged them; the residual starts
    // at zero (creation is never inside a captured graph, so the fills may be memsets)
    if (hipMemsetAsync(m->freeze_flags, 1, freeze_flag_bytes(*desc), as_stream(stream)) != hipSuccess ||
        hipMemsetAsync(m->residual, 0, residual_bytes(*desc), as_stream(stream)) != hipSuccess ||
        hipMemsetAsync(m->deferred, 0, sizeof(int), as_stream(stream)) != hipSuccess ||
        launch_beer_factors(grid_dev(desc->grid), desc->params.first_absorption_coefficient,
                            desc->params.second_absorption_coefficient, m->beer, as_stream(stream)) != hipSuccess) {
        ne_model_destroy(m);
        return NE_ERR_DEVICE;
    }
    for (int b = 0; b < 4; b++) {
        if (m->owns_buffers && hipMalloc(bufs[b], n) != hipSuccess) {
            ne_model_destroy(m);
            return NE_ERR_DEVICE;
        }
        views[b]->data = *bufs[b];
        views[b]->sx = 1;                 // the layout of the model's T view
        views[b]->sy = desc->ocean.T.sy;
        views[b]->sz = desc->ocean.T.sz;
    }
    // reconcile_state!: fluxes are computed at construction (earth_system_model.jl:86-90)
    r = model_update_state(m, INFINITY, as_stream(stream));
    if (r) {
        ne_model_destroy(m);
        return r;
    }
    *out = m;
    return NE_OK;
}

int ne_model_workspace_bytes(const ne_model_desc* desc, int64_t* bytes) {
    if (!desc || !bytes) return NE_ERR_NULL;
    int r = check_grid(&desc->grid);
    if (r || (r = check_view3(desc->ocean.T))) return r;
    if (desc->precision != NE_FP64 && desc->precision != NE_MIXED_FP32) return NE_ERR_UNSUPPORTED;
    *bytes = (int64_t)model_workspace_bytes(*desc);
    return NE_OK;
}

int ne_model_destroy(ne_model* m) {
    if (!m) return NE_ERR_NULL;
    if (m->batch) return NE_ERR_STATE;
    int r = NE_OK;
    void* bufs[4] = {m->T1, m->S1, m->T2, m->S2};
    if (m->owns_buffers) {
        for (void* b : bufs)
            if (b && hipFree(b) != hipSuccess) r = NE_ERR_DEVICE;
        if (m->freeze_flags && hipFree(m->freeze_flags) != hipSuccess) r = NE_ERR_DEVICE;
        if (m->residual && hipFree(m->residual) != hipSuccess) r = NE_ERR_DEVICE;
        if (m->deferred && hipFree(m->deferred) != hipSuccess) r = NE_ERR_DEVICE;
        if (m->beer && hipFree(m->beer) != hipSuccess) r = NE_ERR_DEVICE;
    }
    for (hipEvent_t e : m->events) (void)hipEventDestroy(e);
    delete m;
    return r;
}

int ne_model_time_step(ne_model* m, double dt, void* stream) {
    if (!m) return NE_ERR_NULL;
    if (m->batch) return NE_ERR_STATE;
    if (!(dt > 0)) return NE_ERR_SHAPE;
    hipStream_t s = as_stream(stream);
    // the snapshots of the step's end must be resident before anything is enqueued
    int r = model_time_views(m, m->time + dt, false);
    if (r) return r;
    hipEvent_t* ev = nullptr;
    if (m->timing && m->recorded < m->max_steps) ev = &m->events[(size_t)m->recorded * 6];
    if (m->d.timestepper == NE_TIMESTEPPER_QUASI_ADAMS_BASHFORTH2)
        r = model_ab2_step(m, dt, s, ev);
    else
        r = m->d.precision == NE_FP64 ? model_ocean_step<double>(m, dt, s, ev) : model_ocean_step<float>(m, dt, s, ev);
    if (r) return r;
    m->time += dt;
    m->iteration += 1;
    m->last_dt = dt;
    r = model_update_state(m, dt, s, ev, true);
    if (!r && ev) m->recorded += 1;
    return r;
}

int ne_model_timing_begin(ne_model* m, int max_steps) {
    if (!m) return NE_ERR_NULL;
    if (max_steps <= 0 || max_steps > (1 << 20)) return NE_ERR_CAPACITY;
    for (hipEvent_t e : m->events) (void)hipEventDestroy(e);
    m->events.assign((size_t)max_steps * 6, nullptr);
    for (auto& e : m->events)
        if (hipEventCreate(&e) != hipSuccess) return NE_ERR_DEVICE;
    m->max_steps = max_steps;
    m->recorded = 0;
    m->timing = true;
    return NE_OK;
}

int ne_model_timing_end(ne_model* m, double* ms, int* steps) {
    if (!m || !ms) return NE_ERR_NULL;
    if (!m->timing) return NE_ERR_STATE;
    for (int k = 0; k < NE_TIMED_KERNELS; k++) ms[k] = 0.0;
    for (int n = 0; n < m->recorded; n++) {
        hipEvent_t* ev = &m->events[(size_t)n * 6];
        if (hipEventSynchronize(ev[5]) != hipSuccess) return NE_ERR_DEVICE;
        for (int k = 0; k < NE_TIMED_KERNELS; k++) {
            float t = 0.0f;
            if (hipEventElapsedTime(&t, ev[k], ev[k + 1]) != hipSuccess) return NE_ERR_DEVICE;
            ms[k] += t;
        }
    }
    if (steps) *steps = m->recorded;
    m->timing = false;
    return NE_OK;
}

int ne_model_update_state(ne_model* m, void* stream) {
    if (!m) return NE_ERR_NULL;
    if (m->batch) return NE_ERR_STATE;
    return model_update_state(m, m->iteration == 0 ? INFINITY : m->last_dt, as_stream(stream));
}

int ne_model_update_vertical_velocity(ne_model* m, void* stream) {
    if (!m) return NE_ERR_NULL;
    if (m->batch) return NE_ERR_STATE;
    const ne_model_desc& d = m->d;
    const GridDev g = grid_dev(d.grid);
    hipError_t e;
    if (d.precision == NE_FP64)
        e = launch_w_continuity<double>(g, view<const double>(d.ocean.u), view<const double>(d.ocean.v),
                                        view<double>(d.ocean.w), as_stream(stream));
    else
        e = launch_w_continuity<float>(g, view<const float>(d.ocean.u), view<const float>(d.ocean.v),
                                       view<float>(d.ocean.w), as_stream(stream));
    return dev_err(e);
}

int ne_snapshot_time_indices(const double* times, int32_t n_times, int32_t time_indexing, double time_period,
                             double time, ne_time_interp* interp, int64_t positions[2]) {
    if (!interp || !positions || (n_times > 0 && !times)) return NE_ERR_NULL;
    if (n_times <= 0) return NE_ERR_SHAPE;
    if (time_indexing != NE_TIME_CLAMPED && time_indexing != NE_TIME_CYCLICAL) return NE_ERR_UNSUPPORTED;
    if (time_indexing == NE_TIME_CYCLICAL && !(time_period > 0.0 && time_period >= times[n_times - 1] - times[0]))
        return NE_ERR_SHAPE;
    SeriesAt s = series_at(times, n_times, time, time_indexing, time_period);
    *interp = s.ti;
    positions[0] = s.p1;
    positions[1] = s.p2;
    return NE_OK;
}

int ne_model_snapshots_at(const ne_model* m, double time, int64_t out[4]) {
    if (!m || !out) return NE_ERR_NULL;
    const ne_model_desc& d = m->d;
    SeriesAt a = series_at(m->atimes, time, d), r = series_at(m->rtimes, time, d);
    out[0] = m->atimes.empty() ? NO_POSITION : a.p1;
    out[1] = m->atimes.empty() ? NO_POSITION : a.p2;
    out[2] = m->rtimes.empty() ? NO_POSITION : r.p1;
    out[3] = m->rtimes.empty() ? NO_POSITION : r.p2;
    return NE_OK;
}

int ne_model_snapshot_resident(ne_model* m, int32_t which, int64_t position) {
    if (!m) return NE_ERR_NULL;
    if (which != 0 && which != 1) return NE_ERR_SHAPE;
    std::vector<int64_t>& slot = which ? m->rslot : m->aslot;
    const int64_t nt = (int64_t)(which ? m->rtimes.size() : m->atimes.size());
    if (slot.empty()) return NE_ERR_STATE;           // not a windowed series
    if (position == NO_POSITION) return NE_ERR_SHAPE;
    if (m->d.time_indexing != NE_TIME_CYCLICAL && (position < 0 || position >= nt)) return NE_ERR_SHAPE;
    slot[(size_t)floor_mod(position, (int64_t)slot.size())] = position;
    return NE_OK;
}

int ne_model_rounding_residual(const ne_model* m, double* out, void* stream) {
    if (!m || !out) return NE_ERR_NULL;
    return dev_err(launch_residual_reduce(m->residual, grid_dev(m->d.grid), out, as_stream(stream)));
}

int ne_model_clock(const ne_model* m, double* time, int64_t* iteration) {
    if (!m) return NE_ERR_NULL;
    if (time) *time = m->time;
    if (iteration) *iteration = m->iteration;
    return NE_OK;
}

}  // extern "C"
