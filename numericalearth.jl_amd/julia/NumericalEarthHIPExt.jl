# NumericalEarthHIPExt — package extension binding NumericalEarth's coupled surface phases
# to libnumericalearth_hip (MI355X, gfx950) through its C ABI (include/numericalearth_hip.h).
#
# Loaded by Julia's package-extension mechanism when AMDGPU is loaded next to
# NumericalEarth (Project.toml: [weakdeps] AMDGPU, [extensions] NumericalEarthHIPExt = "AMDGPU").
# It adds methods for the phase functions of time_step!/update_state!
# (src/EarthSystemModels/time_step_earth_system_model.jl:14-83) on models whose fields
# live in ROCArrays, so OceanOnlyModel and time_step! are used unchanged.
#
# Not executable in the build container (no Julia there); the same ABI is exercised by
# the Python binding and its tests.
module NumericalEarthHIPExt

using AMDGPU
using NumericalEarth
using NumericalEarth.EarthSystemModels: EarthSystemModel
using NumericalEarth.EarthSystemModels.InterfaceComputations
using Oceananigans: interior

const libne = get(ENV, "NUMERICALEARTH_HIP_LIBRARY", "libnumericalearth_hip.so")

# ---------------------------------------------------------------------------
# ABI mirrors (field order and types as in include/numericalearth_hip.h)
# ---------------------------------------------------------------------------

struct NEView
    data :: Ptr{Cvoid}
    sx :: Int64
    sy :: Int64
    sz :: Int64
end

const NULLVIEW = NEView(C_NULL, 0, 0, 0)

"""
    ne_view(field)

Strided view of an Oceananigans Field's data: pointer to the first interior point of
`parent(field)` and the parent's element strides (any halo).
"""
function ne_view(field)
    p = parent(field)
    Hx, Hy, Hz = field.grid.Hx, field.grid.Hy, ndims(p) == 3 && size(p, 3) > 1 ? field.grid.Hz : 0
    s = strides(p)
    offset = Hx * s[1] + Hy * s[2] + (length(s) >= 3 ? Hz * s[3] : 0)
    ptr = Ptr{Cvoid}(pointer(p) + offset * sizeof(eltype(p)))
    return NEView(ptr, s[1], s[2], length(s) >= 3 ? s[3] : 0)
end

struct NEGrid
    Nx :: Int32; Ny :: Int32; Nz :: Int32
    topology_x :: Int32; topology_y :: Int32; advection :: Int32
    dz :: Ptr{Float64}; z_center :: Ptr{Float64}; dy_xface :: Ptr{Float64}; dx_yface :: Ptr{Float64}
    Az :: Ptr{Float64}; lambda_c :: Ptr{Float64}; phi_c :: Ptr{Float64}; kmin :: Ptr{Int32}
    z_face :: Ptr{Float64}
end

struct NEExchangeState
    u::NEView; v::NEView; T::NEView; p::NEView; q::NEView; rain::NEView; snow::NEView
    downwelling_shortwave::NEView; downwelling_longwave::NEView
end
struct NEOceanState; u::NEView; v::NEView; w::NEView; T::NEView; S::NEView; end
struct NEInterfaceFluxes
    latent_heat::NEView; sensible_heat::NEView; water_vapor::NEView; x_momentum::NEView; y_momentum::NEView
    friction_velocity::NEView; temperature_scale::NEView; water_vapor_scale::NEView
    interface_temperature::NEView; iterations::NEView
end
struct NENetOceanFluxes; tau_x::NEView; tau_y::NEView; J_T::NEView; J_S::NEView; J_w::NEView; J_H::NEView
    shortwave_surface_flux::NEView; end   # J₀ of TwoColorRadiation (surface_flux field), or a NULL view
struct NERadiativeFluxes; upwelling_longwave::NEView; downwelling_longwave::NEView; downwelling_shortwave::NEView; end

# ne_params is an 832-byte block in header order (checked against the header by the
# repository's layout test); it is filled by ne_params_default and then overwritten
# field by field from the flux formulation.
const NE_PARAMS_BYTES = 832
struct NEParams
    bytes :: NTuple{NE_PARAMS_BYTES, UInt8}
end

function default_params()
    r = Ref{NEParams}()
    check(ccall((:ne_params_default, libne), Cint, (Ref{NEParams},), r), "ne_params_default")
    return r[]
end

check(status, what) = status == 0 || throw(ArgumentError("$what failed with ne_status $status"))

# One ne_ctx per device, created lazily.
const CONTEXTS = Dict{Int, Ptr{Cvoid}}()
function context(device = AMDGPU.device_id(AMDGPU.device()) - 1)
    get!(CONTEXTS, device) do
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:ne_ctx_create, libne), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h), "ne_ctx_create")
        h[]
    end
end

stream_handle() = AMDGPU.stream().stream   # hipStream_t of the current task-local stream

# ---------------------------------------------------------------------------
# Phase overrides (used when the exchange grid lives on an AMD GPU)
# ---------------------------------------------------------------------------

const HIPModel = EarthSystemModel  # narrowed by the method signatures below to ROC-resident grids

"""
Replaces `compute_atmosphere_ocean_fluxes!` (atmosphere_ocean_fluxes.jl:17-60) by
`ne_ao_flux_solve_f64`.
"""
function InterfaceComputations.compute_atmosphere_ocean_fluxes!(coupled_model::HIPModel)
    ex = coupled_model.interfaces.exchanger
    atmos = ex.atmosphere.state
    ao = coupled_model.interfaces.atmosphere_ocean_interface
    ocean = coupled_model.ocean.model
    grid = ne_grid(coupled_model)
    params = ne_params(coupled_model)
    rad = ex.radiation   # radiation exchanger (atmosphere_ocean_fluxes.jl:12-15); nothing without radiation
    sw, lw = isnothing(rad) ? (NULLVIEW, NULLVIEW) : (ne_view(rad.state.ℐꜜˢʷ), ne_view(rad.state.ℐꜜˡʷ))
    state = NEExchangeState(ne_view(atmos.u), ne_view(atmos.v), ne_view(atmos.T), ne_view(atmos.p),
                            ne_view(atmos.q), ne_view(atmos.Jʳⁿ), ne_view(atmos.Jˢⁿ), sw, lw)
    oc = NEOceanState(ne_view(ocean.velocities.u), ne_view(ocean.velocities.v), ne_view(ocean.velocities.w),
                      ne_view(ocean.tracers.T), ne_view(ocean.tracers.S))
    f = ao.fluxes
    fl = NEInterfaceFluxes(ne_view(f.latent_heat), ne_view(f.sensible_heat), ne_view(f.water_vapor),
                           ne_view(f.x_momentum), ne_view(f.y_momentum), ne_view(f.friction_velocity),
                           ne_view(f.temperature_scale), ne_view(f.water_vapor_scale), ne_view(ao.temperature),
                           NULLVIEW)
    status = ccall((:ne_ao_flux_solve_f64, libne), Cint,
                   (Ptr{Cvoid}, Ref{NEGrid}, Ref{NEParams}, Ref{NEExchangeState}, Ref{NEOceanState},
                    Ref{NEInterfaceFluxes}, Ptr{Cvoid}),
                   context(), grid, params, state, oc, fl, stream_handle())
    check(status, "ne_ao_flux_solve_f64")
    return nothing
end

# The remaining phases follow the same pattern:
#   interpolate_state!(exchanger, grid, ::PrescribedAtmosphere, model) → ne_interp_atmos_f64
#   interpolate_state!(exchanger, grid, ::PrescribedRadiation, model)  → ne_interp_radiation_f64
#   compute_sea_ice_ocean_fluxes!(::FreezingLimitedEarthSystemModel)  → ne_freeze_clamp_f64
#   update_net_fluxes!(model, ocean)                                   → ne_assemble_net_ocean_fluxes_f64
#   apply_air_sea_radiative_fluxes!(model)                             → ne_apply_air_sea_radiation_f64
# and the ocean's tracer step → ne_tracer_rk_stage_f64 (three calls per Δt).
# ne_grid/ne_params below build the descriptor blocks once per model and cache them.

function ne_grid end     # GridFittedBottom → kmin via ne_grid_immersed_levels; LatLon metrics per row
function ne_params end   # SimilarityTheoryFluxes / ComponentInterfaces → NEParams

end # module
