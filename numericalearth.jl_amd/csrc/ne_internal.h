// ne_internal.h — host-side declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "ne_device.h"

struct ne_ctx {
    int device;
    ne_counters* counters;      // device
    double* scratch;            // device workspace (budget partials), grown on demand
    size_t scratch_bytes;
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
};

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

// Fused update_state: K1 interpolates and solves every column, K2 clamps, assembles
// and adds radiation.
template <typename Real>
hipError_t launch_surface_solve(const GridDev& g, const ne_params& p, const AtmosDev<Real>& a,
                                const double* fi, const double* fj, View3<const Real> T, View3<const Real> S,
                                View3<const Real> u, View3<const Real> v, const FluxOut& out, ne_counters* c,
                                hipStream_t s);
template <typename Real>
hipError_t launch_surface_finalize(const GridDev& g, const ne_params& p, const AtmosDev<Real>& a,
                                   const double* afi, const double* afj, const RadDev<Real>& r,
                                   const double* rfi, const double* rfj, View3<Real> T, View3<const Real> S,
                                   const FluxOut& f, const NetOut& n, const RadOut& ro, View3<double> frazil,
                                   double dt_frazil, hipStream_t s);

// Tracer launchers (ne_tracer.hip).
template <typename Real>
hipError_t launch_rk_stage(const GridDev& g, View3<const Real> u, View3<const Real> v, View3<const Real> w,
                           View3<const Real> T0, View3<const Real> S0, View3<const Real> Tin,
                           View3<const Real> Sin, View3<Real> Tout, View3<Real> Sout, View3<const double> JT,
                           View3<const double> JS, View3<const double> Jw, double beta_dt, hipStream_t s);
hipError_t launch_tendency(const GridDev& g, View3<const double> u, View3<const double> v, View3<const double> w,
                           View3<const double> T, View3<const double> S, View3<const double> JT,
                           View3<const double> JS, View3<const double> Jw, View3<double> GT, View3<double> GS,
                           hipStream_t s);
template <typename Real>
hipError_t launch_budget(ne_ctx* ctx, const GridDev& g, View3<const Real> c, double weight, double* out,
                         hipStream_t s);

}  // namespace ne
