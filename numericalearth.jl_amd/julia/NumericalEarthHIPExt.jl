# NumericalEarthHIPExt — package extension binding NumericalEarth's coupled surface phases and
# the ocean's WENO5 tracer step to libnumericalearth_hip (MI355X, gfx950) through its C ABI
# (include/numericalearth_hip.h).
#
# Loaded by Julia's package-extension mechanism when AMDGPU is loaded next to NumericalEarth
# (Project.toml: [weakdeps] AMDGPU, [extensions] NumericalEarthHIPExt = "AMDGPU"), the way the
# reference adds back ends in ext/.
#
# Use: build the model as before and hand it over once,
#
#     coupled_model = OceanOnlyModel(ocean; atmosphere, radiation)   # fields on ROCArrays
#     coupled_model = NumericalEarthHIPExt.hip(coupled_model)
#     time_step!(coupled_model, Δt)                                  # or run!(Simulation(coupled_model))
#
# `hip` wraps the ocean simulation in `HIPOcean` (no field is copied). Every method below
# dispatches on a model whose ocean is a HIPOcean (`HIPModel`), so CPU and CUDA models, and ROC
# models that were not handed over, keep the reference's methods; nothing here redefines a method
# of the reference. The phases replaced (time_step_earth_system_model.jl:14-83):
#
#   interpolate_state!(exchanger, grid, ::PrescribedAtmosphere, model)  → ne_interp_atmos_{f64,f32}
#   interpolate_state!(exchanger, grid, ::PrescribedRadiation, model)   → ne_interp_radiation_{f64,f32}
#   compute_atmosphere_ocean_fluxes!(model)                            → ne_ao_flux_solve_f64
#   compute_sea_ice_ocean_fluxes!(::FreezingLimited model)             → ne_freeze_clamp_f64
#   update_net_fluxes!(model, ocean)                                   → ne_assemble_net_ocean_fluxes_f64
#   apply_air_sea_radiative_fluxes!(model)                             → ne_apply_air_sea_radiation_f64
#   time_step!(ocean, Δt): T and S by three split-RK3 stages           → ne_tracer_rk_stage_{f64,f32}
#                          (ocean_simulation.jl:347), then fill_halo_regions! of T and S;
#                          with timestepper = :QuasiAdamsBashforth2 (:323) → ne_tracer_tendency_weno5_f64
#                          into the time stepper's Gⁿ, then               → ne_tracer_ab2_update_f64
#
# The ocean step is the tracer step only: the velocities are those of the wrapped model
# (prescribed, or updated by the caller between steps; ne_compute_w_from_continuity_f64 derives
# w), and closures other than `nothing` are refused. The 2-D surface views keep their halos:
# ne_grid.halo = 1 makes the interpolations and the flux solve write the ring i = 0, Nx+1 and
# j = 0, Ny+1 as the reference launches them over 0:N+1 (InterfaceComputations.jl:99-116).
#
# Not executable in the build container (no Julia there). The structs below mirror the header
# field for field; tests/test_julia_binding.py checks them against the ctypes mirror, which is
# itself checked against the header's layout by compiling it.
module NumericalEarthHIPExt

using AMDGPU
using AMDGPU: ROCArray, ROCBackend
using Oceananigans
using Oceananigans: Simulation, Center, Face
using Oceananigans.Architectures: GPU, CPU, architecture, on_architecture
using Oceananigans.Grids: AbstractGrid, Periodic, Bounded, Flat, topology, λnode, φnode, znode
using Oceananigans.ImmersedBoundaries: ImmersedBoundaryGrid, GridFittedBottom
using Oceananigans.Operators: Δzᶜᶜᶜ, Azᶜᶜᶜ, Δyᶠᶜᶜ, Δxᶜᶠᶜ
using Oceananigans.BoundaryConditions: fill_halo_regions!
using Oceananigans.OutputReaders: Cyclical
using Oceananigans.TimeSteppers: tick!, QuasiAdamsBashforth2TimeStepper
using NumericalEarth
using NumericalEarth.EarthSystemModels: EarthSystemModels, EarthSystemModel, NoSeaIceInterface
using NumericalEarth.EarthSystemModels.InterfaceComputations
using NumericalEarth.EarthSystemModels.InterfaceComputations: SimilarityTheoryFluxes, ConvergenceStopCriteria,
    FixedIterations, ConvectiveGustiness, SubgridVelocityCorrection, MomentumRoughnessLength, ScalarRoughnessLength,
    EdsonMomentumStabilityFunction, EdsonScalarStabilityFunction, LogarithmicSimilarityProfile,
    COARELogarithmicSimilarityProfile, ImpureSaturationSpecificHumidity, BulkTemperature, RelativeVelocity,
    WindVelocity
using NumericalEarth.Atmospheres: PrescribedAtmosphere
using NumericalEarth.Radiations: PrescribedRadiation
using NumericalEarth.SeaIces: FreezingLimitedOceanTemperature
using NumericalEarth.Oceans: get_radiative_forcing, TwoColorRadiation

const libne = get(ENV, "NUMERICALEARTH_HIP_LIBRARY", "libnumericalearth_hip.so")

check(status, what) = status == 0 || throw(ArgumentError("$what failed with ne_status $status"))

# ---------------------------------------------------------------------------
# ABI mirrors (field order and types as in include/numericalearth_hip.h; Julia lays out isbits
# structs as C does, padding included)
# ---------------------------------------------------------------------------

struct NEView
    data :: Ptr{Cvoid}
    sx :: Int64
    sy :: Int64
    sz :: Int64
end

const NULLVIEW = NEView(C_NULL, 0, 0, 0)

struct NEGrid
    Nx :: Int32
    Ny :: Int32
    Nz :: Int32
    topology_x :: Int32
    topology_y :: Int32
    advection :: Int32
    dz :: Ptr{Float64}
    z_center :: Ptr{Float64}
    dy_xface :: Ptr{Float64}
    dx_yface :: Ptr{Float64}
    Az :: Ptr{Float64}
    lambda_c :: Ptr{Float64}
    phi_c :: Ptr{Float64}
    kmin :: Ptr{Int32}
    z_face :: Ptr{Float64}
    halo :: Int32
end

struct NESourceGrid
    Nx :: Int32
    Ny :: Int32
    topology_x :: Int32
    topology_y :: Int32
    lambda0 :: Float64
    dlambda :: Float64
    phi0 :: Float64
    dphi :: Float64
end

struct NETimeInterp
    n1 :: Int32
    n2 :: Int32
    tilde_n :: Float64
end

struct NEPrescribedAtmosphere
    grid :: NESourceGrid
    u :: NEView
    v :: NEView
    T :: NEView
    q :: NEView
    p :: NEView
    rain :: NEView
    snow :: NEView
    time :: NETimeInterp
end

struct NEPrescribedRadiation
    grid :: NESourceGrid
    downwelling_shortwave :: NEView
    downwelling_longwave :: NEView
    time :: NETimeInterp
end

struct NEExchangeState
    u :: NEView
    v :: NEView
    T :: NEView
    p :: NEView
    q :: NEView
    rain :: NEView
    snow :: NEView
    downwelling_shortwave :: NEView
    downwelling_longwave :: NEView
end

struct NEOceanState
    u :: NEView
    v :: NEView
    w :: NEView
    T :: NEView
    S :: NEView
end

struct NEInterfaceFluxes
    latent_heat :: NEView
    sensible_heat :: NEView
    water_vapor :: NEView
    x_momentum :: NEView
    y_momentum :: NEView
    friction_velocity :: NEView
    temperature_scale :: NEView
    water_vapor_scale :: NEView
    interface_temperature :: NEView
    iterations :: NEView
end

struct NENetOceanFluxes
    tau_x :: NEView
    tau_y :: NEView
    J_T :: NEView
    J_S :: NEView
    J_w :: NEView
    J_H :: NEView
    shortwave_surface_flux :: NEView
end

struct NERadiativeFluxes
    upwelling_longwave :: NEView
    downwelling_longwave :: NEView
    downwelling_shortwave :: NEView
end

struct NEParams
    von_karman_constant :: Float64
    turbulent_prandtl_number :: Float64
    convective_gustiness :: Int32
    mesoscale_gustiness :: Int32
    gustiness_parameter :: Float64
    minimum_gustiness :: Float64
    mesoscale_velocity :: Float64
    stability_functions :: Int32
    similarity_form :: Int32
    edson_momentum :: NTuple{11, Float64}
    edson_scalar :: NTuple{12, Float64}
    momentum_roughness :: Int32
    temperature_roughness :: Int32
    water_vapor_roughness :: Int32
    reserved0 :: Int32
    momentum_roughness_length :: Float64
    temperature_roughness_length :: Float64
    water_vapor_roughness_length :: Float64
    roughness_gravitational_acceleration :: Float64
    air_kinematic_viscosity :: Float64
    charnock_parameter :: Float64
    smooth_wall_parameter :: Float64
    maximum_momentum_roughness :: Float64
    reynolds_scaling_A :: Float64
    reynolds_scaling_b :: Float64
    maximum_scalar_roughness :: Float64
    zero_plane_displacement :: Float64
    stop_criteria :: Int32
    maxiter :: Int32
    tolerance :: Float64
    specific_humidity :: Int32
    velocity_difference :: Int32
    water_mole_fraction :: Float64
    fixed_specific_humidity :: Float64
    surface_layer_height :: Float64
    boundary_layer_height :: Float64
    gravitational_acceleration :: Float64
    gas_constant :: Float64
    dry_air_molar_mass :: Float64
    water_molar_mass :: Float64
    dry_air_adiabatic_exponent :: Float64
    water_vapor_heat_capacity :: Float64
    liquid_water_heat_capacity :: Float64
    reference_vaporization_enthalpy :: Float64
    reference_temperature :: Float64
    triple_point_temperature :: Float64
    triple_point_pressure :: Float64
    ocean_reference_density :: Float64
    ocean_heat_capacity :: Float64
    freshwater_density :: Float64
    radiation :: Int32
    radiative_forcing :: Int32
    stefan_boltzmann_constant :: Float64
    ocean_albedo :: Float64
    ocean_emissivity :: Float64
    freshwater_melting_temperature :: Float64
    liquidus_slope :: Float64
    first_color_fraction :: Float64
    first_absorption_coefficient :: Float64
    second_absorption_coefficient :: Float64
    turbulent_fluxes :: Int32
    transfer_coefficients :: Int32
    transfer_kind :: NTuple{3, Int32}
    interface_temperature :: Int32
    transfer_coefficient :: NTuple{3, Float64}
    drag_polynomial :: NTuple{4, Float64}
    high_wind_speed_threshold :: Float64
    high_wind_drag_coefficient :: Float64
    minimum_wind_speed :: Float64
    large_yeager_von_karman_constant :: Float64
    large_yeager_reference_height :: Float64
    stable_heat_transfer_coefficient :: Float64
    unstable_heat_transfer_coefficient :: Float64
    moisture_transfer_coefficient :: Float64
    paulson_momentum :: NTuple{2, Float64}
    paulson_scalar :: Float64
    linear_stable_momentum :: NTuple{2, Float64}
    linear_stable_scalar :: NTuple{2, Float64}
    skin_diffusivity :: Float64
    skin_depth :: Float64
    skin_max_temperature_change :: Float64
    reference_sublimation_enthalpy :: Float64
    water_ice_heat_capacity :: Float64
end

const NE_PARAMS_BYTES = 832
@assert sizeof(NEParams) == NE_PARAMS_BYTES

const NE_PERIODIC, NE_BOUNDED, NE_FLAT = Int32(0), Int32(1), Int32(2)
const NE_TIME_CLAMPED, NE_TIME_CYCLICAL = Int32(0), Int32(1)
const NE_STABILITY_NONE, NE_STABILITY_EDSON = Int32(0), Int32(1)
const NE_SIMILARITY_LOG, NE_SIMILARITY_COARE_LOG = Int32(0), Int32(1)
const NE_ROUGHNESS_CONSTANT, NE_ROUGHNESS_EDSON = Int32(0), Int32(1)
const NE_STOP_CONVERGENCE, NE_STOP_FIXED_ITERATIONS = Int32(0), Int32(1)
const NE_HUMIDITY_IMPURE_SATURATION, NE_HUMIDITY_FIXED = Int32(0), Int32(1)
const NE_VELOCITY_RELATIVE, NE_VELOCITY_WIND = Int32(0), Int32(1)
const NE_RADIATIVE_FORCING_NONE, NE_RADIATIVE_FORCING_TWO_COLOR = Int32(0), Int32(1)

function default_params()
    r = Ref{NEParams}()
    check(ccall((:ne_params_default, libne), Cint, (Ref{NEParams},), r), "ne_params_default")
    return r[]
end

"""
    with(p::NEParams; fields...)

A copy of `p` with the named fields replaced (converted to the field's type).
"""
function with(p::NEParams; kw...)
    values = map(fieldnames(NEParams)) do name
        haskey(kw, name) ? convert(fieldtype(NEParams, name), kw[name]) : getfield(p, name)
    end
    return NEParams(values...)
end

# ---------------------------------------------------------------------------
# Views of Fields and FieldTimeSeries
# ---------------------------------------------------------------------------

"""
    ne_view(field)

Strided view of an Oceananigans Field: pointer to the first interior point of `parent(field)`
and the parent's element strides (any halo; a 2-D (Center, Center, Nothing) field has none in z).
"""
function ne_view(field)
    p = parent(field)
    g = field.grid
    s = strides(p)
    Hz = ndims(p) >= 3 && size(p, 3) > 1 ? g.Hz : 0
    offset = g.Hx * s[1] + g.Hy * s[2] + (length(s) >= 3 ? Hz * s[3] : 0)
    ptr = Ptr{Cvoid}(pointer(p) + offset * sizeof(eltype(p)))
    return NEView(ptr, s[1], s[2], length(s) >= 3 ? s[3] : 0)
end

"""
    series_view(fts)

View of a 2-D FieldTimeSeries whose every snapshot is resident: sx, sy as for a field and
sz = the snapshot stride (the header's convention for time series).
"""
function series_view(fts)
    p = parent(fts.data)
    size(p, 4) == length(fts.times) ||
        throw(ArgumentError("the HIP interpolation reads series whose every snapshot is in memory " *
                            "(size $(size(p, 4)) for $(length(fts.times)) times)"))
    g = fts.grid
    s = strides(p)
    offset = g.Hx * s[1] + g.Hy * s[2]
    ptr = Ptr{Cvoid}(pointer(p) + offset * sizeof(eltype(p)))
    return NEView(ptr, s[1], s[2], s[4])
end

series_view(::Nothing) = NULLVIEW

# ---------------------------------------------------------------------------
# Context, stream
# ---------------------------------------------------------------------------

const CONTEXTS = Dict{Int, Ptr{Cvoid}}()

"""One ne_ctx per device (the library's solver counters and budget workspace), created lazily."""
function context(device = AMDGPU.device_id(AMDGPU.device()) - 1)
    get!(CONTEXTS, device) do
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:ne_ctx_create, libne), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h), "ne_ctx_create")
        h[]
    end
end

stream_handle() = AMDGPU.stream().stream   # hipStream_t of the task-local stream the fields use

# ---------------------------------------------------------------------------
# Grid descriptor: per-row metrics of the immersed LatitudeLongitudeGrid on the device
# ---------------------------------------------------------------------------

ne_topology(::Type{<:Periodic}) = NE_PERIODIC
ne_topology(::Type{<:Bounded}) = NE_BOUNDED
ne_topology(::Type{<:Flat}) = NE_FLAT

struct GridDescriptor
    grid :: NEGrid
    arrays :: Vector{Any}     # the device arrays the descriptor points into (kept alive here)
end

const GRIDS = IdDict{Any, GridDescriptor}()

"""
    ne_grid(grid; halo = 1)

The ne_grid of an ImmersedBoundaryGrid(LatitudeLongitudeGrid, GridFittedBottom), built once per
grid: LatLon metrics depend on j only, so Δyᶠᶜᶜ, Δxᶜᶠᶜ and Azᶜᶜᶜ are taken per row, Δzᶜᶜᶜ and the
heights per level, and kmin from the bottom height by ne_grid_immersed_levels (a cell is
immersed when its centre lies below the bottom). Single-column grids (Flat in x and y) carry no
advection (ocean_simulation.jl:367-379).
"""
function ne_grid(grid::ImmersedBoundaryGrid; halo = 1)
    get!(GRIDS, grid) do
        cpu = on_architecture(CPU(), grid)
        Nx, Ny, Nz = size(grid)
        TX, TY, _ = topology(grid)
        c, f = Center(), Face()
        dz = [Δzᶜᶜᶜ(1, 1, k, cpu) for k in 1:Nz]
        zc = [znode(1, 1, k, cpu, c, c, c) for k in 1:Nz]
        zf = [znode(1, 1, k, cpu, c, c, f) for k in 1:Nz+1]
        dyx = [Δyᶠᶜᶜ(1, j, Nz, cpu) for j in 1:Ny]
        dxy = [Δxᶜᶠᶜ(1, j, Nz, cpu) for j in 1:Ny+1]
        Az = [Azᶜᶜᶜ(1, j, Nz, cpu) for j in 1:Ny]
        λc = [λnode(i, 1, 1, cpu, c, c, c) for i in 1:Nx]
        φc = [φnode(1, j, 1, cpu, c, c, c) for j in 1:Ny]
        arrays = Any[ROCArray{Float64}(a) for a in (dz, zc, dyx, dxy, Az, λc, φc, zf)]
        kmin = AMDGPU.zeros(Int32, Nx * Ny)
        push!(arrays, kmin)
        single_column = TX === Flat && TY === Flat
        ptr(a) = pointer(a)
        g = NEGrid(Nx, Ny, Nz, ne_topology(TX), ne_topology(TY), single_column ? 0 : 1,
                   ptr(arrays[1]), ptr(arrays[2]), ptr(arrays[3]), ptr(arrays[4]), ptr(arrays[5]),
                   ptr(arrays[6]), ptr(arrays[7]), ptr(kmin), ptr(arrays[8]), single_column ? 0 : halo)
        bottom = grid.immersed_boundary.bottom_height
        check(ccall((:ne_grid_immersed_levels, libne), Cint,
                    (Ptr{Cvoid}, Ref{NEGrid}, NEView, Ptr{Int32}, Ptr{Cvoid}),
                    context(), g, ne_view(bottom), pointer(kmin), stream_handle()), "ne_grid_immersed_levels")
        GridDescriptor(g, arrays)
    end.grid
end

ne_grid(grid) = throw(ArgumentError("the HIP path runs on an ImmersedBoundaryGrid(LatitudeLongitudeGrid, " *
                                    "GridFittedBottom); got $(summary(grid))"))

# ---------------------------------------------------------------------------
# Parameters: the flux formulation, interface and component properties lowered to ne_params
# ---------------------------------------------------------------------------

stop_criteria(s::ConvergenceStopCriteria) = (stop_criteria = NE_STOP_CONVERGENCE, maxiter = s.maxiter,
                                             tolerance = s.tolerance)
stop_criteria(s::FixedIterations) = (stop_criteria = NE_STOP_FIXED_ITERATIONS, maxiter = s.iterations)

gustiness(g::ConvectiveGustiness) = (convective_gustiness = 1, mesoscale_gustiness = 0,
                                     gustiness_parameter = g.gustiness_parameter,
                                     minimum_gustiness = g.minimum_gustiness)
function gustiness(g::SubgridVelocityCorrection)
    c = g.convective
    base = isnothing(c) ? (convective_gustiness = 0,) :
           (convective_gustiness = 1, gustiness_parameter = c.gustiness_parameter,
            minimum_gustiness = c.minimum_gustiness)
    g.mesoscale isa Number ||
        isnothing(g.mesoscale) || throw(ArgumentError("the HIP path takes a constant mesoscale velocity"))
    meso = isnothing(g.mesoscale) ? (mesoscale_gustiness = 0,) :
           (mesoscale_gustiness = 1, mesoscale_velocity = g.mesoscale)
    return merge(base, meso)
end
gustiness(::Nothing) = (convective_gustiness = 0, mesoscale_gustiness = 0)

edson(f::EdsonMomentumStabilityFunction) = (f.ζmax, f.A⁺, f.B⁺, f.C⁺, f.D⁺, f.A⁻, f.B⁻, f.C⁻, f.D⁻, f.E⁻, f.F⁻)
edson(f::EdsonScalarStabilityFunction) = (f.ζmax, f.A⁺, f.B⁺, f.C⁺, f.D⁺, f.E⁺, f.A⁻, f.B⁻, f.C⁻, f.D⁻, f.E⁻, f.F⁻)

function stability(s)
    m, h = s.momentum, s.temperature
    if m isa EdsonMomentumStabilityFunction && h isa EdsonScalarStabilityFunction
        s.water_vapor == h || throw(ArgumentError("the HIP path shares the scalar stability function of T and q"))
        return (stability_functions = NE_STABILITY_EDSON, edson_momentum = edson(m), edson_scalar = edson(h))
    elseif m isa Returns && h isa Returns        # stability_functions = nothing (:205-208)
        return (stability_functions = NE_STABILITY_NONE,)
    end
    throw(ArgumentError("the HIP path implements the Edson stability functions or none"))
end

similarity(::LogarithmicSimilarityProfile) = NE_SIMILARITY_LOG
similarity(::COARELogarithmicSimilarityProfile) = NE_SIMILARITY_COARE_LOG

function roughness(r)
    m, t, q = r.momentum, r.temperature, r.water_vapor
    out = (;)
    if m isa MomentumRoughnessLength
        m.wave_formulation isa Number || throw(ArgumentError("the HIP path takes a constant Charnock parameter"))
        out = merge(out, (momentum_roughness = NE_ROUGHNESS_EDSON,
                          roughness_gravitational_acceleration = m.gravitational_acceleration,
                          air_kinematic_viscosity = m.air_kinematic_viscosity, charnock_parameter = m.wave_formulation,
                          smooth_wall_parameter = m.smooth_wall_parameter,
                          maximum_momentum_roughness = m.maximum_roughness_length))
    else
        out = merge(out, (momentum_roughness = NE_ROUGHNESS_CONSTANT, momentum_roughness_length = m))
    end
    for (name, s) in ((:temperature, t), (:water_vapor, q))
        if s isa ScalarRoughnessLength
            f = s.reynolds_number_scaling_function
            out = merge(out, NamedTuple{(Symbol(name, :_roughness), :reynolds_scaling_A, :reynolds_scaling_b,
                                         :maximum_scalar_roughness)}((NE_ROUGHNESS_EDSON, f.A, f.b,
                                                                      s.maximum_roughness_length)))
        else
            out = merge(out, NamedTuple{(Symbol(name, :_roughness), Symbol(name, :_roughness_length))}(
                                        (NE_ROUGHNESS_CONSTANT, s)))
        end
    end
    return out
end

function flux_formulation(f::SimilarityTheoryFluxes)
    f.zero_plane_displacement isa Number ||
        throw(ArgumentError("the HIP path takes a constant zero-plane displacement"))
    return merge((von_karman_constant = f.von_karman_constant,
                  turbulent_prandtl_number = f.turbulent_prandtl_number,
                  similarity_form = similarity(f.similarity_form),
                  zero_plane_displacement = f.zero_plane_displacement),
                 gustiness(f.subgrid_velocities), stability(f.stability_functions),
                 roughness(f.roughness_lengths), stop_criteria(f.solver_stop_criteria))
end

flux_formulation(f) = throw(ArgumentError("the Julia binding lowers SimilarityTheoryFluxes; " *
                                          "$(summary(f)) is reachable through the C ABI's ne_params"))

function interface_properties(p)
    q = p.specific_humidity_formulation
    q isa ImpureSaturationSpecificHumidity && q.water_mole_fraction isa Number ||
        throw(ArgumentError("the HIP path takes ImpureSaturationSpecificHumidity with a constant mole fraction"))
    p.temperature_formulation isa BulkTemperature ||
        throw(ArgumentError("the Julia binding lowers BulkTemperature (SkinTemperature through ne_params)"))
    v = p.velocity_formulation isa WindVelocity ? NE_VELOCITY_WIND : NE_VELOCITY_RELATIVE
    return (specific_humidity = NE_HUMIDITY_IMPURE_SATURATION, water_mole_fraction = q.water_mole_fraction,
            velocity_difference = v)
end

function thermodynamics(atp)
    c, h, t = atp.constitutive, atp.heat_capacity, atp.phase_transitions
    return (gas_constant = c.gas_constant, dry_air_molar_mass = c.dry_air_molar_mass,
            water_molar_mass = c.water_molar_mass, dry_air_adiabatic_exponent = h.dry_air_adiabatic_exponent,
            water_vapor_heat_capacity = h.water_vapor_heat_capacity,
            liquid_water_heat_capacity = h.liquid_water_heat_capacity,
            water_ice_heat_capacity = h.water_ice_heat_capacity,
            reference_vaporization_enthalpy = t.reference_vaporization_enthalpy,
            reference_sublimation_enthalpy = t.reference_sublimation_enthalpy,
            reference_temperature = t.reference_temperature, triple_point_temperature = t.triple_point_temperature,
            triple_point_pressure = t.triple_point_pressure)
end

function radiation_properties(radiation::PrescribedRadiation, ocean)
    s = radiation.surface_properties.ocean
    s.albedo isa Number && s.emissivity isa Number ||
        throw(ArgumentError("the HIP path takes a constant ocean albedo and emissivity"))
    forcing = get_radiative_forcing(ocean.model)
    two = forcing isa TwoColorRadiation
    base = (radiation = 1, stefan_boltzmann_constant = radiation.stefan_boltzmann_constant,
            ocean_albedo = s.albedo, ocean_emissivity = s.emissivity,
            radiative_forcing = two ? NE_RADIATIVE_FORCING_TWO_COLOR : NE_RADIATIVE_FORCING_NONE)
    two || return base
    return merge(base, (first_color_fraction = forcing.first_color_fraction,
                        first_absorption_coefficient = forcing.first_absorption_coefficient,
                        second_absorption_coefficient = forcing.second_absorption_coefficient))
end
radiation_properties(::Nothing, ocean) = (radiation = 0, radiative_forcing = NE_RADIATIVE_FORCING_NONE)

liquidus(sea_ice::FreezingLimitedOceanTemperature) =
    (freshwater_melting_temperature = sea_ice.liquidus.freshwater_melting_temperature,
     liquidus_slope = sea_ice.liquidus.slope)
liquidus(sea_ice) = (;)

"""
    ne_params(coupled_model)

The ne_params block of a model: ne_params_default, then the atmosphere-ocean flux formulation
(similarity_theory_turbulent_fluxes.jl:174-214), the interface properties
(interface_states.jl:8-12), the atmosphere's thermodynamics parameters, the ocean and interface
properties (component_interfaces.jl:425-437, :483-484), the prescribed radiation
(prescribed_radiation.jl:12-24) and the liquidus of FreezingLimitedOceanTemperature.
Formulations the binding does not lower raise ArgumentError, never silently.
"""
function ne_params(cm::EarthSystemModel)
    ao = cm.interfaces.atmosphere_ocean_interface
    op = cm.interfaces.ocean_properties
    atmos = cm.atmosphere
    kw = merge(flux_formulation(ao.flux_formulation), interface_properties(ao.properties),
               thermodynamics(cm.interfaces.atmosphere_properties),
               (ocean_reference_density = op.reference_density, ocean_heat_capacity = op.heat_capacity,
                freshwater_density = op.freshwater_density,
                gravitational_acceleration = cm.interfaces.properties.gravitational_acceleration,
                surface_layer_height = cm.interfaces.properties.surface_layer_height,
                boundary_layer_height = EarthSystemModels.boundary_layer_height(atmos)),
               radiation_properties(cm.radiation, cm.ocean), liquidus(cm.sea_ice))
    return with(default_params(); kw...)
end

# ---------------------------------------------------------------------------
# The ocean component the binding steps
# ---------------------------------------------------------------------------

"""
    HIPOcean(ocean::Simulation)

The reference's ocean simulation, stepped by libnumericalearth_hip: `time_step!` advances T and S
by the three split-RK3 stages of ocean_simulation.jl:347 (WENO5 tracer advection, the top flux
boundary condition, TwoColorRadiation when the model has it) with the model's velocities, then
fills the halos of T and S as Oceananigans does after each step. Property access forwards to the
simulation (`ocean.model`, `ocean.Δt`), so code that reads the ocean keeps working; the stage
buffers are allocated once, with the layout of T.
"""
struct HIPOcean{S, B}
    simulation :: S
    stages :: B               # (T¹, S¹, T², S²), parents shaped like T's
end

function HIPOcean(ocean::Simulation)
    model = ocean.model
    grid = model.grid
    architecture(grid) isa GPU{<:ROCBackend} ||
        throw(ArgumentError("HIPOcean needs an ocean on an AMD GPU (GPU(ROCBackend())), not $(architecture(grid))"))
    isnothing(model.closure) || model.closure == () ||
        throw(ArgumentError("HIPOcean steps advection and surface fluxes only; the closure must be nothing"))
    T = model.tracers.T
    stages = ntuple(_ -> similar(parent(T)), 4)
    return HIPOcean(ocean, stages)
end

Base.getproperty(o::HIPOcean, name::Symbol) =
    name in fieldnames(HIPOcean) ? getfield(o, name) : getproperty(getfield(o, :simulation), name)

# the stage buffer b viewed like T (same halos and strides)
function stage_view(o::HIPOcean, b::Int)
    T = o.simulation.model.tracers.T
    v = ne_view(T)
    offset = v.data - Ptr{Cvoid}(pointer(parent(T)))
    return NEView(Ptr{Cvoid}(pointer(o.stages[b])) + offset, v.sx, v.sy, v.sz)
end

const HIPModel = EarthSystemModel{<:Any, <:Any, <:Any, <:Any, <:HIPOcean}

"""
    hip(coupled_model::EarthSystemModel)

The same model — components, interfaces and clock shared, no field copied — with its ocean
wrapped in HIPOcean, so time_step!/update_state! run the phases on libnumericalearth_hip.
"""
function hip(cm::EarthSystemModel)
    ocean = cm.ocean isa HIPOcean ? cm.ocean : HIPOcean(cm.ocean)
    return EarthSystemModel(cm.architecture, cm.clock, cm.radiation, cm.atmosphere, cm.land, cm.sea_ice, ocean,
                            cm.interfaces)
end

# The ocean component interface forwards to the simulation (as NestedModels/nested_component.jl
# forwards a NestedModel's to its child).
Oceananigans.Grids.grid(o::HIPOcean) = o.simulation.model.grid
EarthSystemModels.reference_density(o::HIPOcean) = EarthSystemModels.reference_density(o.simulation)
EarthSystemModels.heat_capacity(o::HIPOcean) = EarthSystemModels.heat_capacity(o.simulation)
EarthSystemModels.is_sea_ice_component(::HIPOcean) = false
EarthSystemModels.ocean_temperature(o::HIPOcean) = EarthSystemModels.ocean_temperature(o.simulation)
EarthSystemModels.ocean_salinity(o::HIPOcean) = EarthSystemModels.ocean_salinity(o.simulation)
EarthSystemModels.ocean_surface_temperature(o::HIPOcean) = EarthSystemModels.ocean_surface_temperature(o.simulation)
EarthSystemModels.ocean_surface_salinity(o::HIPOcean) = EarthSystemModels.ocean_surface_salinity(o.simulation)
EarthSystemModels.ocean_surface_velocities(o::HIPOcean) = EarthSystemModels.ocean_surface_velocities(o.simulation)
EarthSystemModels.interpolate_state!(exchanger, grid, ::HIPOcean, coupled_model) = nothing
InterfaceComputations.ComponentExchanger(o::HIPOcean, grid) = InterfaceComputations.ComponentExchanger(o.simulation, grid)
InterfaceComputations.net_fluxes(o::HIPOcean) = InterfaceComputations.net_fluxes(o.simulation)

# ---------------------------------------------------------------------------
# Views of the coupled model's fields
# ---------------------------------------------------------------------------

ocean_state(o::HIPOcean) = (m = o.simulation.model;
    NEOceanState(ne_view(m.velocities.u), ne_view(m.velocities.v), ne_view(m.velocities.w),
                 ne_view(m.tracers.T), ne_view(m.tracers.S)))

function exchange_state(cm)
    a = cm.interfaces.exchanger.atmosphere.state
    r = cm.interfaces.exchanger.radiation
    sw, lw = isnothing(r) ? (NULLVIEW, NULLVIEW) : (ne_view(r.state.ℐꜜˢʷ), ne_view(r.state.ℐꜜˡʷ))
    return NEExchangeState(ne_view(a.u), ne_view(a.v), ne_view(a.T), ne_view(a.p), ne_view(a.q),
                           ne_view(a.Jʳⁿ), ne_view(a.Jˢⁿ), sw, lw)
end

function interface_fluxes(cm)
    ao = cm.interfaces.atmosphere_ocean_interface
    f = ao.fluxes
    return NEInterfaceFluxes(ne_view(f.latent_heat), ne_view(f.sensible_heat), ne_view(f.water_vapor),
                             ne_view(f.x_momentum), ne_view(f.y_momentum), ne_view(f.friction_velocity),
                             ne_view(f.temperature_scale), ne_view(f.water_vapor_scale), ne_view(ao.temperature),
                             NULLVIEW)
end

# net ocean fluxes (Oceans.jl:174-196: T, S, u, v, η, freshwater_heat_content), and J₀ of
# TwoColorRadiation (radiative_forcing.jl:86-91) when the ocean carries it
function net_ocean_fluxes(net, ocean)
    forcing = get_radiative_forcing(ocean.model)
    J0 = forcing isa TwoColorRadiation ? ne_view(forcing.surface_flux) : NULLVIEW
    return NENetOceanFluxes(ne_view(net.u), ne_view(net.v), ne_view(net.T), ne_view(net.S), ne_view(net.η),
                            ne_view(net.freshwater_heat_content), J0)
end

# ---------------------------------------------------------------------------
# Source grids and time interpolation of prescribed series
# ---------------------------------------------------------------------------

const INDICES = IdDict{Any, Tuple{Any, Any}}()

"""
Fractional indices (fi[Nx], fj[Ny], 0-based) of the exchange grid in a series' source grid,
computed once per (exchange grid, source grid) by ne_fractional_indices_nodes_f64 from the
source's cell-centre coordinates (prescribed_atmosphere_regridder.jl:51-71).
"""
function fractional_indices(grid, source)
    get!(INDICES, (grid, source)) do
        Nx, Ny, _ = size(grid)
        nx, ny = size(source, 1), size(source, 2)
        c = Center()
        cpu = on_architecture(CPU(), source)
        λ = ROCArray{Float64}([λnode(i, 1, 1, cpu, c, c, c) for i in 1:nx])
        φ = ROCArray{Float64}([φnode(1, j, 1, cpu, c, c, c) for j in 1:ny])
        TX, TY, _ = topology(source)
        fi = AMDGPU.zeros(Float64, Nx)
        fj = AMDGPU.zeros(Float64, Ny)
        check(ccall((:ne_fractional_indices_nodes_f64, libne), Cint,
                    (Ptr{Cvoid}, Ref{NEGrid}, Ptr{Float64}, Int32, Int32, Ptr{Float64}, Int32, Int32,
                     Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                    context(), ne_grid(grid), pointer(λ), nx, ne_topology(TX), pointer(φ), ny, ne_topology(TY),
                    pointer(fi), pointer(fj), stream_handle()), "ne_fractional_indices_nodes_f64")
        (fi, fj)
    end
end

"""
Snapshot indices and weight at time t (cpu_interpolating_time_indices,
interpolate_atmospheric_state.jl:57-60), by the library's ne_snapshot_time_indices.
"""
function time_interp(fts, t)
    times = Vector{Float64}(fts.times)
    ti = fts.time_indexing
    kind, period = ti isa Cyclical ? (NE_TIME_CYCLICAL, Float64(ti.period)) : (NE_TIME_CLAMPED, 0.0)
    out = Ref{NETimeInterp}()
    positions = zeros(Int64, 2)
    check(ccall((:ne_snapshot_time_indices, libne), Cint,
                (Ptr{Float64}, Int32, Int32, Float64, Float64, Ref{NETimeInterp}, Ptr{Int64}),
                times, length(times), kind, period, Float64(t), out, positions), "ne_snapshot_time_indices")
    return out[]
end

# the source grid block is unused by the node-coordinate indices: only its sizes are read
source_block(source) = NESourceGrid(size(source, 1), size(source, 2), ne_topology(topology(source)[1]),
                                    ne_topology(topology(source)[2]), 0, 0, 0, 0)

suffix(::Type{Float64}) = "f64"
suffix(::Type{Float32}) = "f32"

# ---------------------------------------------------------------------------
# Phase overrides
# ---------------------------------------------------------------------------

"""Replaces interpolate_state!(…, ::PrescribedAtmosphere, …) (interpolate_atmospheric_state.jl:9-137)."""
function EarthSystemModels.interpolate_state!(exchanger, grid, atmosphere::PrescribedAtmosphere, cm::HIPModel)
    u = atmosphere.velocities.u
    source = u.grid
    fi, fj = fractional_indices(grid, source)
    desc = NEPrescribedAtmosphere(source_block(source), series_view(u), series_view(atmosphere.velocities.v),
                                  series_view(atmosphere.temperature), series_view(atmosphere.specific_humidity),
                                  series_view(atmosphere.pressure),
                                  series_view(get(atmosphere.freshwater_flux, :rain, nothing)),
                                  series_view(get(atmosphere.freshwater_flux, :snow, nothing)),
                                  time_interp(u, cm.clock.time))
    s = exchanger.state
    out = NEExchangeState(ne_view(s.u), ne_view(s.v), ne_view(s.T), ne_view(s.p), ne_view(s.q), ne_view(s.Jʳⁿ),
                          ne_view(s.Jˢⁿ), NULLVIEW, NULLVIEW)
    f = Symbol(:ne_interp_atmos_, suffix(eltype(parent(u.data))))
    status = ccall(Libc.Libdl.dlsym(Libc.Libdl.dlopen(libne), f), Cint,
                   (Ptr{Cvoid}, Ref{NEGrid}, Ref{NEPrescribedAtmosphere}, Ptr{Float64}, Ptr{Float64},
                    Ref{NEExchangeState}, Ptr{Cvoid}),
                   context(), ne_grid(grid), desc, pointer(fi), pointer(fj), out, stream_handle())
    check(status, String(f))
    return nothing
end

"""Replaces interpolate_state!(…, ::PrescribedRadiation, …) (interpolate_radiation_state.jl:4-69)."""
function EarthSystemModels.interpolate_state!(exchanger, grid, radiation::PrescribedRadiation, cm::HIPModel)
    sw = radiation.downwelling_shortwave
    source = sw.grid
    fi, fj = fractional_indices(grid, source)
    desc = NEPrescribedRadiation(source_block(source), series_view(sw),
                                 series_view(radiation.downwelling_longwave), time_interp(sw, cm.clock.time))
    s = exchanger.state
    f = Symbol(:ne_interp_radiation_, suffix(eltype(parent(sw.data))))
    status = ccall(Libc.Libdl.dlsym(Libc.Libdl.dlopen(libne), f), Cint,
                   (Ptr{Cvoid}, Ref{NEGrid}, Ref{NEPrescribedRadiation}, Ptr{Float64}, Ptr{Float64}, NEView,
                    NEView, Ptr{Cvoid}),
                   context(), ne_grid(grid), desc, pointer(fi), pointer(fj), ne_view(s.ℐꜜˢʷ), ne_view(s.ℐꜜˡʷ),
                   stream_handle())
    check(status, String(f))
    return nothing
end

"""Replaces compute_atmosphere_ocean_fluxes! (atmosphere_ocean_fluxes.jl:17-60)."""
function InterfaceComputations.compute_atmosphere_ocean_fluxes!(cm::HIPModel)
    grid = cm.interfaces.exchanger.grid
    status = ccall((:ne_ao_flux_solve_f64, libne), Cint,
                   (Ptr{Cvoid}, Ref{NEGrid}, Ref{NEParams}, Ref{NEExchangeState}, Ref{NEOceanState},
                    Ref{NEInterfaceFluxes}, Ptr{Cvoid}),
                   context(), ne_grid(grid), ne_params(cm), exchange_state(cm), ocean_state(cm.ocean),
                   interface_fluxes(cm), stream_handle())
    check(status, "ne_ao_flux_solve_f64")
    return nothing
end

const HIPFreezingLimitedModel =
    EarthSystemModel{<:Any, <:Any, <:Any, <:FreezingLimitedOceanTemperature, <:HIPOcean, <:NoSeaIceInterface}

"""Replaces compute_sea_ice_ocean_fluxes!(::FreezingLimitedEarthSystemModel) (freezing_limited_ocean_temperature.jl:73-93)."""
function InterfaceComputations.compute_sea_ice_ocean_fluxes!(cm::HIPFreezingLimitedModel)
    ocean = cm.ocean
    model = ocean.model
    Δt = ocean.Δt
    Δt_frazil = model.clock.iteration == 0 ? Inf : Float64(Δt)   # the guard of :87-88
    status = ccall((:ne_freeze_clamp_f64, libne), Cint,
                   (Ptr{Cvoid}, Ref{NEGrid}, Ref{NEParams}, NEView, NEView, NEView, Float64, Ptr{Cvoid}),
                   context(), ne_grid(model.grid), ne_params(cm), ne_view(model.tracers.T), ne_view(model.tracers.S),
                   ne_view(cm.sea_ice.frazil_heat), Δt_frazil, stream_handle())
    check(status, "ne_freeze_clamp_f64")
    return nothing
end

"""Replaces update_net_fluxes!(model, ocean) (assemble_net_ocean_fluxes.jl:16-70)."""
function EarthSystemModels.update_net_fluxes!(cm::HIPModel, ocean::HIPOcean)
    grid = cm.interfaces.exchanger.grid
    net = net_ocean_fluxes(cm.interfaces.net_fluxes.ocean, ocean)
    status = ccall((:ne_assemble_net_ocean_fluxes_f64, libne), Cint,
                   (Ptr{Cvoid}, Ref{NEGrid}, Ref{NEParams}, Ref{NEInterfaceFluxes}, Ref{NEExchangeState}, NEView,
                    Ref{NENetOceanFluxes}, Ptr{Cvoid}),
                   context(), ne_grid(grid), ne_params(cm), interface_fluxes(cm), exchange_state(cm),
                   ne_view(ocean.model.tracers.T), net, stream_handle())
    check(status, "ne_assemble_net_ocean_fluxes_f64")
    return nothing
end

const HIPRadiatedModel = EarthSystemModel{<:PrescribedRadiation, <:Any, <:Any, <:Any, <:HIPOcean}

"""Replaces apply_air_sea_radiative_fluxes! (apply_air_sea_radiative_fluxes.jl:18-60)."""
function EarthSystemModels.apply_air_sea_radiative_fluxes!(cm::HIPRadiatedModel)
    grid = cm.interfaces.exchanger.grid
    r = cm.interfaces.exchanger.radiation.state
    d = cm.radiation.interface_fluxes.ocean
    out = NERadiativeFluxes(ne_view(d.upwelling_longwave), ne_view(d.downwelling_longwave),
                            ne_view(d.downwelling_shortwave))
    net = net_ocean_fluxes(cm.interfaces.net_fluxes.ocean, cm.ocean)
    status = ccall((:ne_apply_air_sea_radiation_f64, libne), Cint,
                   (Ptr{Cvoid}, Ref{NEGrid}, Ref{NEParams}, NEView, NEView, NEView, Ref{NENetOceanFluxes},
                    Ref{NERadiativeFluxes}, Ptr{Cvoid}),
                   context(), ne_grid(grid), ne_params(cm), ne_view(r.ℐꜜˢʷ), ne_view(r.ℐꜜˡʷ),
                   ne_view(cm.interfaces.atmosphere_ocean_interface.temperature), net, out, stream_handle())
    check(status, "ne_apply_air_sea_radiation_f64")
    return nothing
end

# The tracer step's parameters: the surface constants do not enter it, only the radiative forcing
# (TwoColorRadiation adds its shortwave divergence to Gᵀ).
function tracer_params(ocean::HIPOcean)
    p = with(default_params(); radiation_properties(nothing, ocean)...)
    forcing = get_radiative_forcing(ocean.simulation.model)
    if forcing isa TwoColorRadiation
        p = with(p; radiative_forcing = NE_RADIATIVE_FORCING_TWO_COLOR,
                 first_color_fraction = forcing.first_color_fraction,
                 first_absorption_coefficient = forcing.first_absorption_coefficient,
                 second_absorption_coefficient = forcing.second_absorption_coefficient)
    end
    return p
end

"""
    time_step!(ocean::HIPOcean, Δt)

T and S advanced by the split Runge-Kutta 3 stages c⁽ᵐ⁾ = cⁿ + βₘ Δt G(c⁽ᵐ⁻¹⁾), β = (1/3, 1/2, 1)
(ocean_simulation.jl:347), G the WENO5 advection tendency with the top flux boundary condition
(ne_tracer_rk_stage_f64 / _f32 by T's element type), the last stage writing T and S in place;
then their halos are filled and the ocean clock ticks, as Oceananigans' step leaves them. A model
built with timestepper = :QuasiAdamsBashforth2 (ocean_simulation.jl:323) takes ab2_time_step!.
"""
function Oceananigans.TimeSteppers.time_step!(ocean::HIPOcean, Δt)
    model = ocean.simulation.model
    model.timestepper isa QuasiAdamsBashforth2TimeStepper && return ab2_time_step!(ocean, Δt)
    T, S = model.tracers.T, model.tracers.S
    u, v, w = model.velocities.u, model.velocities.v, model.velocities.w
    net = net_ocean_fluxes(InterfaceComputations.net_fluxes(ocean.simulation), ocean)
    g = ne_grid(model.grid)
    p = tracer_params(ocean)
    f = Symbol(:ne_tracer_rk_stage_, suffix(eltype(parent(T))))
    fn = Libc.Libdl.dlsym(Libc.Libdl.dlopen(libne), f)
    T0, S0 = ne_view(T), ne_view(S)
    T1, S1, T2, S2 = (stage_view(ocean, b) for b in 1:4)
    stages = ((T0, S0, T1, S1, 1/3), (T1, S1, T2, S2, 1/2), (T2, S2, T0, S0, 1.0))
    for (Tin, Sin, Tout, Sout, β) in stages
        status = ccall(fn, Cint,
                       (Ptr{Cvoid}, Ref{NEGrid}, Ref{NEParams}, NEView, NEView, NEView, NEView, NEView, NEView,
                        NEView, NEView, NEView, Ref{NENetOceanFluxes}, Float64, Ptr{Cvoid}),
                       context(), g, p, ne_view(u), ne_view(v), ne_view(w), T0, S0, Tin, Sin, Tout, Sout, net,
                       β * Float64(Δt), stream_handle())
        check(status, String(f))
    end
    fill_halo_regions!((T, S))
    tick!(model.clock, Δt)
    return nothing
end

"""
    ab2_time_step!(ocean::HIPOcean, Δt)

Oceananigans' Quasi-Adams-Bashforth 2 step of T and S for timestepper = :QuasiAdamsBashforth2: the
tendency of the state into the time stepper's Gⁿ (ne_tracer_tendency_weno5_f64), then
c ← c + Δt((3/2 + χ)Gⁿ − (1/2 + χ)G⁻) (ne_tracer_ab2_update_f64, χ the time stepper's), forward
Euler when Δt differs from the clock's last Δt (the first step included); Gⁿ is then kept as G⁻,
the halos filled and the clock ticked with its last Δt set. FP64 fields only.
"""
function ab2_time_step!(ocean::HIPOcean, Δt)
    model = ocean.simulation.model
    ts = model.timestepper
    T, S = model.tracers.T, model.tracers.S
    eltype(parent(T)) == Float64 || throw(ArgumentError("the Quasi-Adams-Bashforth 2 step is FP64"))
    net = net_ocean_fluxes(InterfaceComputations.net_fluxes(ocean.simulation), ocean)
    g = ne_grid(model.grid)
    status = ccall((:ne_tracer_tendency_weno5_f64, libne), Cint,
                   (Ptr{Cvoid}, Ref{NEGrid}, Ref{NEParams}, Ref{NEOceanState}, Ref{NENetOceanFluxes}, NEView, NEView,
                    Ptr{Cvoid}),
                   context(), g, tracer_params(ocean), ocean_state(ocean), net, ne_view(ts.Gⁿ.T), ne_view(ts.Gⁿ.S),
                   stream_handle())
    check(status, "ne_tracer_tendency_weno5_f64")
    euler = Δt != model.clock.last_Δt
    status = ccall((:ne_tracer_ab2_update_f64, libne), Cint,
                   (Ptr{Cvoid}, Ref{NEGrid}, NEView, NEView, NEView, NEView, NEView, NEView, Float64, Float64, Int32,
                    Ptr{Cvoid}),
                   context(), g, ne_view(T), ne_view(S), ne_view(ts.Gⁿ.T), ne_view(ts.Gⁿ.S), ne_view(ts.G⁻.T),
                   ne_view(ts.G⁻.S), Float64(Δt), Float64(ts.χ), Int32(euler), stream_handle())
    check(status, "ne_tracer_ab2_update_f64")
    parent(ts.G⁻.T) .= parent(ts.Gⁿ.T)
    parent(ts.G⁻.S) .= parent(ts.Gⁿ.S)
    fill_halo_regions!((T, S))
    tick!(model.clock, Δt)
    model.clock.last_Δt = Δt
    return nothing
end

end # module
