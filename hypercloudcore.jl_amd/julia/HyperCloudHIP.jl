"""
    HyperCloudHIP

Julia binding of libhcc (include/hcc.h): the MI355X time-step path behind HyperCloudCore.jl's operator API.

A `HIPRKTimeStepper` is a `TimeStepper` with the functor signature and contract of `GeneralRKTimeStepper`
(src/TimeStepping/TimeSteppers.jl:531; src/Types.jl:174-183, "updates the grid in-place"): `ts(eq, pg, t, dt)` advances
the state and, by default, writes it back to `pg.rhos` (and, on a moving grid, the moved positions to
`pg.core.positions`) before returning. `resident = true` keeps the state on the device between calls instead (no
PCIe copy per step; read it with `download!`), which `hip_solve_equation` and device-resident drivers use.

The stepper carries a `divergence_interpolator` field, a `HIPDivergence`, for which `get_time_step(pg, eq, d)` is
defined: the reference's driver loop (test/time_integration.jl:96-111) takes `timestepper.divergence_interpolator`
and calls `get_time_step(pg, eq, div_interp)`, then `timestepper(eq, pg, t, dt)`, then saves `pg.rhos`, and so runs
unchanged on this stepper.

Construction from the reference's objects: `HIPRKTimeStepper(eq, pg, interp::MUSCL, mood::MOOD, tableau)` reads the
MUSCL order, limiter and mode, numerical flux and streamed flag, the MOOD strategy and criterion, the tableau (both
parts) and the boundary-condition map and mover of `pg`; `substep_ratio` defaults to the reference's own
`1 / ssp_coefficient(tableau)` (TimeSteppers.jl:218). PDE types map to compiled kernel variants through
`hip_descriptor`; a PDE, limiter, flux, strategy, criterion or boundary condition without a compiled variant raises
(there is no fallback path).

This file is not exercised in this repository (no Julia toolchain); the Python host package exercises the same C ABI
in the tests, and tests/test_julia_binding.py checks this file's descriptor layouts against include/hcc.h.
"""
module HyperCloudHIP

using HyperCloudCore
using StaticArrays

const libhcc = joinpath(@__DIR__, "..", "lib", "libhcc.so")

const HCC_PDE_LINADV = Int32(1)
const HCC_PDE_EULER  = Int32(2)
const HCC_BC_OUTFLOW = Int32(1)
const HCC_BC_DIRICHLET = Int32(2)

struct DomainDesc
    D::Int32
    mins::NTuple{3, Float64}
    maxs::NTuple{3, Float64}
    periodic::NTuple{3, Int32}
    dx::NTuple{3, Float64}
    R::Float64
    alpha::Float64
    jitter::NTuple{3, Float64}
    seed::UInt64
end

struct StepperDesc
    M::Int32
    pde::Int32
    gamma::Float64
    adv_speed::NTuple{3, NTuple{8, Float64}}
    muscl_order::Int32
    limiter::Int32
    limiter_mode::Int32
    mood_strategy::Int32
    mood_criterion::Int32
    mood_delta::Float64
    stages::Int32
    a_t::NTuple{5, NTuple{5, Float64}}
    b_t::NTuple{5, Float64}
    c_t::NTuple{5, Float64}
    substep_ratio::Float64
    n_bc::Int32
    bc_tag::NTuple{16, Int32}
    bc_kind::NTuple{16, Int32}
    mood_plateau::Float64
    muscl_streamed::Int32
    flux::Int32
    a_im::NTuple{5, NTuple{5, Float64}}
    b_im::NTuple{5, Float64}
    c_im::NTuple{5, Float64}
    mover::Int32             # 0 NoGridMover, 1 LagrangianMover, 2 ScaledLagrangianMover (hip_mover_code)
    mover_alpha::Float64
    mood_ratio::Float64      # MOODu2's ratio (its curvature override at MUSCL order 3)
end

"`(mover, mover_alpha)` of the descriptor for `pg.mover`; PrescribedMover (a closure) has no compiled kernel."
hip_mover_code(m) = error("no compiled grid mover for $(typeof(m))")
hip_mover_code(::HyperCloudCore.NoGridMover) = (Int32(0), 1.0)
hip_mover_code(::HyperCloudCore.LagrangianMover) = (Int32(1), 1.0)
hip_mover_code(m::HyperCloudCore.ScaledLagrangianMover) = (Int32(2), Float64(m.alpha))

checkrc(rc, ctx) = rc == 0 || error("libhcc: ", unsafe_string(ccall((:hcc_last_error, libhcc), Cstring, (Ptr{Cvoid},), ctx)))

"PDE descriptor of a physics type: `(kind, gamma, speeds)`; add a method for each PDE the kernels compile."
hip_descriptor(eq) = error("no compiled HIP kernel for $(typeof(eq)): add a `hip_descriptor` method (there is no fallback)")

pad3(v, x) = ntuple(i -> i <= length(v) ? v[i] : x, 3)

"""
    HIPDivergence

The `divergence_interpolator` of a `HIPRKTimeStepper`: `get_time_step(pg, eq, d::HIPDivergence)` is the CFL step
(CFL number 1) of the device state, computed on the device with the cached weights (hcc_time_step).
"""
mutable struct HIPDivergence <: HyperCloudCore.DivergenceInterpolator
    st::Ptr{Cvoid}
    ctx::Ptr{Cvoid}
    owner::Any   # the HIPRKTimeStepper (set after construction)
end

mutable struct HIPRKTimeStepper <: HyperCloudCore.TimeStepper
    ctx::Ptr{Cvoid}
    grid::Ptr{Cvoid}
    st::Ptr{Cvoid}
    N::Int
    M::Int
    nets::Vector{Ptr{Cvoid}}   # indicator networks owned by the stepper (destroyed before its context)
    resident::Bool             # false: every call exchanges the state with pg.rhos (the reference's contract)
    moving::Bool               # a grid mover: the functor also writes the moved positions to pg.core.positions
    fresh::Bool                # the device state was uploaded from pg.rhos since the last step (by get_time_step)
    divergence_interpolator::HIPDivergence
end

"""
    HIPRKTimeStepper(eq, pg, desc::DomainDesc, sdesc::StepperDesc; device = 0, resident = false)

Uploads the nodes of `pg` and builds the device grid (neighbour search on the GPU) and the stepper from explicit
descriptors. `sdesc.substep_ratio` is `ts.substep_ratio` of the equivalent `GeneralRKTimeStepper` (Julia's bisection
result). The constructor from the reference's objects below fills both descriptors.
"""
function HIPRKTimeStepper(eq, pg, desc::DomainDesc, sdesc::StepperDesc; device::Integer = 0, resident::Bool = false)
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    checkrc(ccall((:hcc_ctx_create, libhcc), Cint, (Cint, Ref{Ptr{Cvoid}}), device, ctx), C_NULL)
    N = pg.meta.N
    pos = Vector{Float64}(undef, N * desc.D)
    for i in 1:N, d in 1:desc.D
        pos[(i - 1) * desc.D + d] = pg.core.positions[i][d]
    end
    isb = UInt8.(pg.core.is_boundary[1:N])
    tags = Int32.(pg.core.tags[1:N])
    grid = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve pos isb tags begin
        checkrc(ccall((:hcc_grid_create, libhcc), Cint,
                      (Ptr{Cvoid}, Ref{DomainDesc}, Int64, Ptr{Float64}, Ptr{UInt8}, Ptr{Int32}, Ref{Ptr{Cvoid}}),
                      ctx[], desc, N, pos, isb, tags, grid), ctx[])
    end
    st = Ref{Ptr{Cvoid}}(C_NULL)
    checkrc(ccall((:hcc_stepper_create, libhcc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{StepperDesc}, Ref{Ptr{Cvoid}}),
                  ctx[], grid[], sdesc, st), ctx[])
    div = HIPDivergence(st[], ctx[], nothing)
    ts = HIPRKTimeStepper(ctx[], grid[], st[], N, Int(sdesc.M), Ptr{Cvoid}[], resident, sdesc.mover != 0, true, div)
    div.owner = ts
    upload!(ts, pg)
    finalizer(ts) do t
        ccall((:hcc_stepper_destroy, libhcc), Cint, (Ptr{Cvoid},), t.st)
        foreach(m -> ccall((:hcc_mlp_destroy, libhcc), Cint, (Ptr{Cvoid},), m), t.nets)
        ccall((:hcc_grid_destroy, libhcc), Cint, (Ptr{Cvoid},), t.grid)
        ccall((:hcc_ctx_destroy, libhcc), Cint, (Ptr{Cvoid},), t.ctx)
    end
    return ts
end

# ---- descriptors from the reference's objects ------------------------------------------------------------------
const HC = HyperCloudCore

hip_limiter(l) = error("no compiled limiter for $(typeof(l)) (compiled: NoLimiter, BarthJespersen, Venkatakrishnan)")
hip_limiter(::HC.NoLimiter) = (Int32(0), Int32(0))
hip_limiter(::HC.BarthJespersenLimiter{Mode}) where {Mode} = (Int32(1), Int32(Mode === :hard))
hip_limiter(::HC.VenkatakrishnanLimiter{Mode}) where {Mode} = (Int32(2), Int32(Mode === :hard))

hip_flux(f) = error("no compiled numerical flux for $(typeof(f)) (compiled: RusanovFlux, AlignedRusanovFlux)")
hip_flux(::HC.RusanovFlux) = Int32(0)
hip_flux(::HC.AlignedRusanovFlux) = Int32(1)

hip_strategy(s) = error("no compiled MOOD strategy for $(typeof(s))")
hip_strategy(::HC.NoStrategy) = Int32(0)
hip_strategy(::HC.EPD0) = Int32(1)
hip_strategy(::HC.EPD1) = Int32(2)
hip_strategy(::HC.EPD2) = Int32(3)
hip_strategy(::HC.StrictEPD0) = Int32(4)

"`(criterion, delta, plateau, ratio)` of a MOOD criterion (MOODu2's `ratio` acts at MUSCL order 3)."
hip_criterion(c) = error("no compiled MOOD criterion for $(typeof(c))")
hip_criterion(::HC.NoMOOD) = (Int32(0), 0.0, 0.0, 0.5)
hip_criterion(c::HC.MOODu1) = (Int32(1), Float64(c.d), 0.0, 0.5)
hip_criterion(::HC.MOODPAD) = (Int32(2), 0.0, 0.0, 0.5)
hip_criterion(::HC.OnlyMOOD) = (Int32(3), 0.0, 0.0, 0.5)
hip_criterion(c::HC.MOODu2) = (Int32(4), Float64(c.d), Float64(c.plateau), Float64(c.ratio))

hip_bc(bc) = error("no compiled boundary condition for $(typeof(bc)) (compiled: OutflowBC)")
hip_bc(::HC.OutflowBC) = HCC_BC_OUTFLOW

_muscl_order(::HC.MUSCL{D, M, T, B, MO}) where {D, M, T, B, MO} = Int32(MO)
_streamed(interp::HC.MUSCL) = Int32(interp.streamed isa Val{true})
_mat5(a) = ntuple(i -> ntuple(j -> (i <= size(a, 1) && j <= size(a, 2)) ? Float64(a[i, j]) : 0.0, 5), 5)
_vec5(v) = ntuple(i -> i <= length(v) ? Float64(v[i]) : 0.0, 5)
_pad(v, n, x) = ntuple(i -> i <= length(v) ? v[i] : x, n)

"""
    hip_domain_desc(pg) -> DomainDesc

The grid descriptor of a `ParticleGrid`: the geometry's box and periodicity, the nominal spacing, the weight
function's range and alpha (the ghost canvas and the bins follow from these as in ComputationalDomain). The jitter
is zero and the seed unused: the nodes are `pg.core.positions`, not regenerated.
"""
function hip_domain_desc(pg::HC.ParticleGrid{D}) where {D}
    geo, wf = pg.geometry, pg.neighbor.weight_func
    DomainDesc(Int32(D), pad3(Float64.(geo.mins), 0.0), pad3(Float64.(geo.maxs), 0.0),
               pad3(Int32.(geo.is_periodic), Int32(0)), pad3(Float64.(pg.meta.dx), 0.0), Float64(pg.meta.R),
               Float64(wf.alpha), (0.0, 0.0, 0.0), UInt64(0))
end

"""
    hip_stepper_desc(eq, pg, interp::MUSCL, mood::MOOD, tableau; substep_ratio) -> StepperDesc

The stepper descriptor of `GeneralRKTimeStepper(eq, interp, mood, tableau)` on `pg` (TimeSteppers.jl:195-229):
`hip_descriptor(eq)` gives the PDE kernel `(kind, gamma, speeds)` (speeds: `speeds[d][m]`, the advection velocity of
component m along axis d); the tableau's explicit part `(a_t, b_t, c_t)` and implicit part `(a, b, c)`; the
boundary-condition map of `pg.geometry`; the mover of `pg`.
"""
function hip_stepper_desc(eq::HC.HyperbolicPDE{D, M}, pg, interp::HC.MUSCL, mood::HC.MOOD, tableau;
                          substep_ratio::Real) where {D, M}
    kind, gamma, speeds = hip_descriptor(eq)
    lim, mode = hip_limiter(interp.limiter)
    crit, delta, plateau, ratio = hip_criterion(mood.criterion)
    s = length(tableau.b_t)
    s <= 5 || error("tableaus of up to 5 stages are compiled")
    bcm = pg.geometry.bc_map
    length(bcm) <= 16 || error("up to 16 boundary-condition tags are compiled")
    tags = sort(collect(keys(bcm)))
    mover, malpha = hip_mover_code(pg.mover)
    adv = ntuple(d -> ntuple(m -> (d <= length(speeds) && m <= length(speeds[d])) ? Float64(speeds[d][m]) : 0.0, 8), 3)
    StepperDesc(Int32(M), Int32(kind), Float64(gamma), adv, _muscl_order(interp), lim, mode, hip_strategy(mood.strategy),
                crit, delta, Int32(s), _mat5(tableau.a_t), _vec5(tableau.b_t), _vec5(tableau.c_t), Float64(substep_ratio),
                Int32(length(tags)), _pad(Int32.(tags), 16, Int32(0)), _pad([hip_bc(bcm[t]) for t in tags], 16, Int32(0)),
                plateau, _streamed(interp), hip_flux(interp.flux), _mat5(tableau.a), _vec5(tableau.b), _vec5(tableau.c),
                mover, malpha, ratio)
end

"""
    HIPRKTimeStepper(eq, pg, interp::MUSCL, mood::MOOD, tableau; substep_ratio, mood_substeps = true, device = 0,
                     resident = false)

The drop-in for `GeneralRKTimeStepper(eq, interp, mood, tableau; mood_substeps)` on `pg`: the same arguments, the
descriptors filled from them (`hip_domain_desc`, `hip_stepper_desc`). `substep_ratio` defaults to the reference's own
value, `1 / ssp_coefficient(tableau)` for an SSP tableau with `mood_substeps` (TimeSteppers.jl:218), else 0.
"""
function HIPRKTimeStepper(eq::HC.HyperbolicPDE, pg::HC.ParticleGrid, interp::HC.MUSCL, mood::HC.MOOD, tableau;
                          substep_ratio::Union{Nothing, Real} = nothing, mood_substeps::Bool = true,
                          device::Integer = 0, resident::Bool = false)
    if substep_ratio === nothing
        C = HC.ssp_coefficient(tableau)
        substep_ratio = (mood_substeps && C > 0) ? 1 / C : 0.0
    end
    sdesc = hip_stepper_desc(eq, pg, interp, mood, tableau; substep_ratio = substep_ratio)
    return HIPRKTimeStepper(eq, pg, hip_domain_desc(pg), sdesc; device = device, resident = resident)
end

"The drop-in built from a constructed `GeneralRKTimeStepper` (its interpolator, MOOD, tableau and substep ratio)."
HIPRKTimeStepper(ts::HC.GeneralRKTimeStepper, eq::HC.HyperbolicPDE, pg::HC.ParticleGrid; kwargs...) =
    HIPRKTimeStepper(eq, pg, ts.divergence_interpolator, ts.mood, ts.tableau; substep_ratio = ts.substep_ratio,
                     kwargs...)

# pg.rhos is a Vector{SVector{M,Float64}}: M contiguous doubles per node, the AoS layout of the ABI
function upload!(ts::HIPRKTimeStepper, pg)
    U = pg.rhos
    GC.@preserve U checkrc(ccall((:hcc_state_upload, libhcc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ts.st, pointer(U)), ts.ctx)
end
function download!(ts::HIPRKTimeStepper, pg)
    U = pg.rhos
    GC.@preserve U checkrc(ccall((:hcc_state_download, libhcc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ts.st, pointer(U)), ts.ctx)
end

"get_time_step(pg, eq, interp) on the device state (CFL number 1)."
function hip_time_step(ts::HIPRKTimeStepper)
    dt = Ref{Float64}(0.0)
    checkrc(ccall((:hcc_time_step, libhcc), Cint, (Ptr{Cvoid}, Ref{Float64}), ts.st, dt), ts.ctx)
    return dt[]
end

"""
    get_time_step(pg, eq, d::HIPDivergence; cached_weights = true)

The reference's `get_time_step(pg, eq, main_grad)` (src/Grid/ParticleGrids.jl:386) for the HIP stepper: unless the
stepper is resident, `pg.rhos` is uploaded first (the CFL step is that of the grid's current state), then the CFL
step is computed on the device.
"""
function HyperCloudCore.get_time_step(pg::HyperCloudCore.ParticleGrid, eq::HyperCloudCore.HyperbolicPDE, d::HIPDivergence;
                                      cached_weights::Bool = true)
    ts = d.owner::HIPRKTimeStepper
    if !ts.resident
        upload!(ts, pg)
        ts.fresh = true
    end
    return hip_time_step(ts)
end

"""
    (ts::HIPRKTimeStepper)(eq, pg, t, dt::Float64)

One time step (`GeneralRKTimeStepper`'s functor, TimeSteppers.jl:531-577). Default mode: the state of `pg.rhos` is
uploaded (unless `get_time_step` just did), stepped on the device and written back to `pg.rhos`; with a grid mover
the moved and wrapped positions go to `pg.core.positions` too. Resident mode (`resident = true`): the device state
is stepped in place and `pg` is left alone. Like the reference, a `dt` of another element type is a `MethodError`.
"""
function (ts::HIPRKTimeStepper)(eq, pg, time::Real, dt::Float64)
    if !ts.resident && !ts.fresh
        upload!(ts, pg)
    end
    ts.fresh = false
    checkrc(ccall((:hcc_step, libhcc), Cint, (Ptr{Cvoid}, Float64, Float64), ts.st, time, dt), ts.ctx)
    if !ts.resident
        download!(ts, pg)
        ts.moving && download_positions!(ts, pg)
    end
    return nothing
end

function mood_flags(ts::HIPRKTimeStepper)
    f = zeros(UInt8, ts.N)
    checkrc(ccall((:hcc_get_mood_flags, libhcc), Cint, (Ptr{Cvoid}, Ptr{UInt8}), ts.st, f), ts.ctx)
    return Bool.(f)
end

function particle_orders(ts::HIPRKTimeStepper)
    o = zeros(Int32, ts.N)
    checkrc(ccall((:hcc_get_orders, libhcc), Cint, (Ptr{Cvoid}, Ptr{Int32}), ts.st, o), ts.ctx)
    return Int.(o)
end

function mood_inadmissible(ts::HIPRKTimeStepper)
    n = Ref{Int64}(0)
    checkrc(ccall((:hcc_get_inadmissible, libhcc), Cint, (Ptr{Cvoid}, Ref{Int64}), ts.st, n), ts.ctx)
    return Int(n[])
end

# Steps, MOOD iterations and HIP-event milliseconds (step, gather passes, flux passes) since the last reset, and the
# steps among them whose passes were timed (set_pass_timing!).
function step_totals(ts::HIPRKTimeStepper; reset::Bool = true)
    n, it, nt = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
    a, b, c = Ref{Float64}(0), Ref{Float64}(0), Ref{Float64}(0)
    checkrc(ccall((:hcc_step_timed_passes, libhcc), Cint, (Ptr{Cvoid}, Ref{Int64}), ts.st, nt), ts.ctx)
    checkrc(ccall((:hcc_step_totals, libhcc), Cint,
                  (Ptr{Cvoid}, Cint, Ref{Int64}, Ref{Int64}, Ref{Float64}, Ref{Float64}, Ref{Float64}),
                  ts.st, Cint(reset), n, it, a, b, c), ts.ctx)
    return (steps = Int(n[]), mood_iterations = Int(it[]), ms_step = a[], ms_gather = b[], ms_flux = c[],
            timed_pass_steps = Int(nt[]))
end

# The gather / flux pass events recorded every `every`-th step from now on (1: every step, the default; 0: never).
function set_pass_timing!(ts::HIPRKTimeStepper, every::Integer)
    checkrc(ccall((:hcc_stepper_set_pass_timing, libhcc), Cint, (Ptr{Cvoid}, Int32), ts.st, Int32(every)), ts.ctx)
    return ts
end

"`pg.core.volumes` of the device grid (`update_volumes!(pg, ShepardVolumes(sweeps = sweeps))` first if `sweeps >= 0`)."
function hip_volumes(ts::HIPRKTimeStepper; sweeps::Integer = -1)
    v = zeros(Float64, ts.N)
    checkrc(ccall((:hcc_grid_volumes, libhcc), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}), ts.grid, Int32(sweeps), v), ts.ctx)
    return v
end

"`total_mass(pg)` (or `interior_mass(pg)`) of the device state: `Σ V_i U_i`."
function hip_total_mass(ts::HIPRKTimeStepper; interior::Bool = false)
    m = zeros(Float64, ts.M)
    checkrc(ccall((:hcc_total_mass, libhcc), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}), ts.st, Int32(interior), m), ts.ctx)
    return m
end

"""
    hip_solve_equation(ts, tmax, dt; is_cfl = false, snapshots = 10, remove_ghosts = false)

`solve_equation` (test/time_integration.jl) with the loop on the stepper's device state (after a non-resident step
that is `pg.rhos`; call `upload!(ts, pg)` after editing `pg.rhos`), snapshots downloaded by the library: returns
`(position_history, state_history, time_history, total_steps, elapsed_wall_time, mass_history)`, the histories as
matrices per snapshot (`D × Ns`, `M × Ns`) and `mass_history[c] = (total_mass, interior_mass)`.
"""
function hip_solve_equation(ts::HIPRKTimeStepper, tmax::Real, dt::Real; is_cfl::Bool = false, snapshots::Integer = 10,
                            remove_ghosts::Bool = false, D::Integer = 2, N_interior::Integer = ts.N)
    ns = remove_ghosts ? N_interior : ts.N
    th = zeros(Float64, snapshots + 1)
    xh = zeros(Float64, D, ns, snapshots + 1)
    uh = zeros(Float64, ts.M, ns, snapshots + 1)
    mh = zeros(Float64, ts.M, 2, snapshots + 1)
    steps, el = Ref{Int64}(0), Ref{Float64}(0)
    checkrc(ccall((:hcc_solve, libhcc), Cint,
                  (Ptr{Cvoid}, Float64, Float64, Int32, Int32, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                   Ptr{Float64}, Ref{Int64}, Ref{Float64}),
                  ts.st, tmax, dt, Int32(is_cfl), Int32(snapshots), Int32(remove_ghosts), th, xh, uh, mh, steps, el), ts.ctx)
    xs = [xh[:, :, c] for c in 1:snapshots + 1]
    us = [uh[:, :, c] for c in 1:snapshots + 1]
    ms = [(mh[:, 1, c], mh[:, 2, c]) for c in 1:snapshots + 1]
    return xs, us, th, Int(steps[]), el[], ms
end

"Decay source `S_i(U) = -rates[i] U` (mode 1 explicit, 2 implicit); rates evaluated on the host (any Julia closure)."
function hip_set_source!(ts::HIPRKTimeStepper, mode::Integer, rates::AbstractVector{Float64})
    r = collect(rates)
    checkrc(ccall((:hcc_set_source, libhcc), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}), ts.st, Int32(mode), r), ts.ctx)
    return ts
end

"""
    HIPEnsembleStepper(pgs, descs, sdesc; device = 0)

Advances several ParticleGrids (same D, R, alpha, spacing and wrap lengths) with one device stepper: the instance
grids are built on the device, concatenated (hcc_grid_concat) and stepped together, each with its own CFL step in
`hip_step_cfl!`. `ts.offsets[k] + 1 : ts.offsets[k] + pgs[k].meta.N` are instance k's rows of the ensemble state.
"""
mutable struct HIPEnsembleStepper
    ctx::Ptr{Cvoid}
    parts::Vector{Ptr{Cvoid}}
    grid::Ptr{Cvoid}
    st::Ptr{Cvoid}
    offsets::Vector{Int64}
    M::Int
end

function HIPEnsembleStepper(pgs::AbstractVector, descs::AbstractVector{DomainDesc}, sdesc::StepperDesc; device::Integer = 0)
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    checkrc(ccall((:hcc_ctx_create, libhcc), Cint, (Cint, Ref{Ptr{Cvoid}}), device, ctx), C_NULL)
    parts = Ptr{Cvoid}[]
    for (pg, desc) in zip(pgs, descs)
        N = pg.meta.N
        pos = [pg.core.positions[i][d] for d in 1:desc.D, i in 1:N]
        isb = UInt8.(pg.core.is_boundary[1:N])
        tags = Int32.(pg.core.tags[1:N])
        g = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve pos isb tags checkrc(ccall((:hcc_grid_create, libhcc), Cint,
            (Ptr{Cvoid}, Ref{DomainDesc}, Int64, Ptr{Float64}, Ptr{UInt8}, Ptr{Int32}, Ref{Ptr{Cvoid}}),
            ctx[], desc, N, pos, isb, tags, g), ctx[])
        push!(parts, g[])
    end
    offsets = zeros(Int64, length(parts) + 1)
    grid = Ref{Ptr{Cvoid}}(C_NULL)
    checkrc(ccall((:hcc_grid_concat, libhcc), Cint, (Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}}, Ref{Ptr{Cvoid}}, Ptr{Int64}),
                  ctx[], Int32(length(parts)), parts, grid, offsets), ctx[])
    st = Ref{Ptr{Cvoid}}(C_NULL)
    checkrc(ccall((:hcc_stepper_create, libhcc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{StepperDesc}, Ref{Ptr{Cvoid}}),
                  ctx[], grid[], sdesc, st), ctx[])
    ts = HIPEnsembleStepper(ctx[], parts, grid[], st[], offsets, Int(sdesc.M))
    U = zeros(Float64, ts.M * offsets[end])
    for (k, pg) in enumerate(pgs), i in 1:pg.meta.N, m in 1:ts.M
        U[(offsets[k] + i - 1) * ts.M + m] = pg.rhos[i][m]
    end
    checkrc(ccall((:hcc_state_upload, libhcc), Cint, (Ptr{Cvoid}, Ptr{Float64}), ts.st, U), ts.ctx)
    finalizer(ts) do t
        ccall((:hcc_stepper_destroy, libhcc), Cint, (Ptr{Cvoid},), t.st)
        ccall((:hcc_grid_destroy, libhcc), Cint, (Ptr{Cvoid},), t.grid)
        foreach(p -> ccall((:hcc_grid_destroy, libhcc), Cint, (Ptr{Cvoid},), p), t.parts)
        ccall((:hcc_ctx_destroy, libhcc), Cint, (Ptr{Cvoid},), t.ctx)
    end
    return ts
end

"One step of every instance with dt_k = min(cfl * CFL step of instance k, dt_cap), computed on the device."
hip_step_cfl!(ts::HIPEnsembleStepper, cfl::Real; dt_cap::Real = Inf) =
    checkrc(ccall((:hcc_step_cfl, libhcc), Cint, (Ptr{Cvoid}, Float64, Float64), ts.st, cfl, dt_cap), ts.ctx)

"The CFL step of every instance (CFL number 1)."
function hip_time_steps(ts::HIPEnsembleStepper)
    dts = zeros(Float64, length(ts.parts))
    checkrc(ccall((:hcc_time_steps, libhcc), Cint, (Ptr{Cvoid}, Ptr{Float64}), ts.st, dts), ts.ctx)
    return dts
end

"""
    hip_set_indicator!(ts::HIPRKTimeStepper, params::Vector{Float32}; threshold = 0f0)

Switches on the learned troubled-node indicator (2-D Euler): a network with the given parameters (HCC_MLP_PARAMS
floats, e.g. the shipped `hcc_hip/indicator_params.npy`) lowers predicted troubled particles to first order before each
flux pass; MOOD's check still runs. Returns the network handle (kept alive by `ts`).
"""
function hip_set_indicator!(ts::HIPRKTimeStepper, params::Vector{Float32}; threshold::Real = 0f0)
    length(params) == 5313 || error("the indicator network has 5313 parameters")
    m = Ref{Ptr{Cvoid}}(C_NULL)
    checkrc(ccall((:hcc_mlp_create, libhcc), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), ts.ctx, m), ts.ctx)
    checkrc(ccall((:hcc_mlp_set_params, libhcc), Cint, (Ptr{Cvoid}, Ptr{Float32}), m[], params), ts.ctx)
    push!(ts.nets, m[])   # destroyed by the stepper's finalizer, after the stepper and before the context
    checkrc(ccall((:hcc_set_indicator, libhcc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Float32), ts.st, m[], Float32(threshold)), ts.ctx)
    return m[]
end

"Copies the device grid's positions (moved and wrapped by a moving stepper) into `pg.core.positions`."
function download_positions!(ts::HIPRKTimeStepper, pg; D::Integer = length(pg.core.positions[1]))
    x = zeros(Float64, D * ts.N)
    checkrc(ccall((:hcc_grid_get_positions, libhcc), Cint, (Ptr{Cvoid}, Ptr{Float64}), ts.grid, x), ts.ctx)
    for i in 1:ts.N
        pg.core.positions[i] = typeof(pg.core.positions[i])(ntuple(d -> x[(i - 1) * D + d], D))
    end
    return pg
end

export hip_set_source!, hip_mover_code, download_positions!, HIPEnsembleStepper, hip_step_cfl!, hip_time_steps,
       hip_set_indicator!
export HIPRKTimeStepper, HIPDivergence, hip_time_step, hip_descriptor, upload!, download!, DomainDesc, StepperDesc,
       step_totals, set_pass_timing!, hip_volumes, hip_total_mass, hip_solve_equation, hip_domain_desc, hip_stepper_desc

end # module
