"""
    HyperCloudHIP

Julia binding of libhcc (include/hcc.h): the MI355X time-step path behind HyperCloudCore.jl's operator API.

A `HIPRKTimeStepper` is a `TimeStepper` with the functor signature of `GeneralRKTimeStepper`
(src/TimeStepping/TimeSteppers.jl:531): `ts(eq, pg, t, dt)` advances `pg.rhos` in place. `hip_time_step(ts, pg, eq)`
replaces `get_time_step(pg, eq, interp)` (src/Grid/ParticleGrids.jl:386). PDE types map to compiled kernel variants
through `hip_descriptor`; a PDE without one raises (there is no fallback path).

This file is not exercised in this repository (no Julia toolchain); the Python host package exercises the same
C ABI in the tests.
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

mutable struct HIPRKTimeStepper <: TimeStepper
    ctx::Ptr{Cvoid}
    grid::Ptr{Cvoid}
    st::Ptr{Cvoid}
    N::Int
    M::Int
    nets::Vector{Ptr{Cvoid}}   # indicator networks owned by the stepper (destroyed before its context)
end

"""
    HIPRKTimeStepper(eq, pg, desc::DomainDesc, interp::MUSCL, mood::MOOD, tableau; device = 0, substep_ratio)

Uploads the nodes of `pg` and builds the device grid (neighbour search on the GPU) and the stepper.
`substep_ratio` is `ts.substep_ratio` of the equivalent `GeneralRKTimeStepper` (Julia's bisection result).
"""
function HIPRKTimeStepper(eq, pg, desc::DomainDesc, sdesc::StepperDesc; device::Integer = 0)
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
    ts = HIPRKTimeStepper(ctx[], grid[], st[], N, Int(sdesc.M), Ptr{Cvoid}[])
    upload!(ts, pg)
    finalizer(ts) do t
        ccall((:hcc_stepper_destroy, libhcc), Cint, (Ptr{Cvoid},), t.st)
        foreach(m -> ccall((:hcc_mlp_destroy, libhcc), Cint, (Ptr{Cvoid},), m), t.nets)
        ccall((:hcc_grid_destroy, libhcc), Cint, (Ptr{Cvoid},), t.grid)
        ccall((:hcc_ctx_destroy, libhcc), Cint, (Ptr{Cvoid},), t.ctx)
    end
    return ts
end

# pg.rhos is a Vector{SVector{M,Float64}}: M contiguous doubles per node, the AoS layout of the ABI
function upload!(ts::HIPRKTimeStepper, pg)
    U = pg.rhos
    GC.@preserve U checkrc(ccall((:hcc_state_upload, libhcc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ts.st, pointer(U)), ts.ctx)
end
function download!(ts::HIPRKTimeStepper, pg)
    U = pg.rhos
    GC.@preserve U checkrc(ccall((:hcc_state_download, libhcc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ts.st, pointer(U)), ts.ctx)
end

"get_time_step(pg, eq, interp) on the device (CFL number 1)."
function hip_time_step(ts::HIPRKTimeStepper)
    dt = Ref{Float64}(0.0)
    checkrc(ccall((:hcc_time_step, libhcc), Cint, (Ptr{Cvoid}, Ref{Float64}), ts.st, dt), ts.ctx)
    return dt[]
end

"The stepper functor: the state stays on the device between steps; call `download!` to read `pg.rhos`."
function (ts::HIPRKTimeStepper)(eq, pg, time::Real, dt::Float64)
    checkrc(ccall((:hcc_step, libhcc), Cint, (Ptr{Cvoid}, Float64, Float64), ts.st, time, dt), ts.ctx)
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

# Steps, MOOD iterations and HIP-event milliseconds (step, gather passes, flux passes) since the last reset.
function step_totals(ts::HIPRKTimeStepper; reset::Bool = true)
    n, it = Ref{Int64}(0), Ref{Int64}(0)
    a, b, c = Ref{Float64}(0), Ref{Float64}(0), Ref{Float64}(0)
    checkrc(ccall((:hcc_step_totals, libhcc), Cint,
                  (Ptr{Cvoid}, Cint, Ref{Int64}, Ref{Int64}, Ref{Float64}, Ref{Float64}, Ref{Float64}),
                  ts.st, Cint(reset), n, it, a, b, c), ts.ctx)
    return (steps = Int(n[]), mood_iterations = Int(it[]), ms_step = a[], ms_gather = b[], ms_flux = c[])
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

`solve_equation` (test/time_integration.jl) with the loop on the device state: returns
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
export HIPRKTimeStepper, hip_time_step, hip_descriptor, upload!, download!, DomainDesc, StepperDesc, step_totals,
       hip_volumes, hip_total_mass, hip_solve_equation

end # module
