"""
    GeoDynamoHIP

`ccall` shim that puts libgdhip (hand-written HIP for MI355X, C ABI in
include/gdhip.h) behind GeoDynamo.jl's own transform and timestep entry
points.  Method names and argument order follow
src/shtnskit_field_functions.jl, src/timestep.jl and src/linear_algebra.jl;
arrays are the reference's column-major layouts, passed as host memory
(GD_HOST) so callers need no device arrays.  Julia was not available where
this shim was written: it is not exercised by the test suite, which drives
the same C ABI from Python (geodynamo.jl_amd/gdhip).
"""
module GeoDynamoHIP

const libgdhip = joinpath(@__DIR__, "..", "gdhip", "libgdhip.so")

const GD_HOST = Cint(1)
const GD_FD_TRANSPOSE = Cint(1)
const PREC = Dict(:fp64 => Cint(0), :fp32 => Cint(1), :bf16x3 => Cint(2))

struct GDError <: Exception
    code::Cint
    msg::String
end

function check(rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:gd_last_error, libgdhip), Cstring, ()))
    rc == 4 && error(msg)                      # timestep.jl:646 raises for a missing l
    throw(GDError(rc, msg))
end

const _ctx = Ref{Ptr{Cvoid}}(C_NULL)
function context(device::Integer=0)
    if _ctx[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:gd_ctx_create, libgdhip), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h))
        _ctx[] = h[]
    end
    return _ctx[]
end

# ---------------------------------------------------------------- transforms
mutable struct HIPSHTPlan
    ptr::Ptr{Cvoid}
    lmax::Int; mmax::Int; nlat::Int; nlon::Int; nlm::Int
end

"create_shtnskit_config (shtnskit_transforms.jl:102) -> device plan"
function create_plan(; lmax::Int, mmax::Int=lmax, nlat::Int, nlon::Int, precision::Symbol=:fp64)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:gd_sht_plan_create, libgdhip), Cint,
                (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint, Ref{Ptr{Cvoid}}),
                context(), lmax, mmax, nlat, nlon, PREC[precision], h))
    nlm = ccall((:gd_sht_nlm, libgdhip), Cint, (Ptr{Cvoid},), h[])
    p = HIPSHTPlan(h[], lmax, mmax, nlat, nlon, nlm)
    finalizer(p) do q
        ccall((:gd_sht_plan_destroy, libgdhip), Cint, (Ptr{Cvoid},), q.ptr)
    end
    return p
end

# spec: data_real/data_imag :: Array{Float64,3} (nlm, 1, nr); phys :: (nlat, nlon, nr)
"shtnskit_spectral_to_physical! (shtnskit_field_functions.jl:11)"
function shtnskit_spectral_to_physical!(p::HIPSHTPlan, re::Array{Float64,3}, im::Array{Float64,3},
                                        phys::Array{Float64,3})
    nr = size(re, 3)
    GC.@preserve re im phys check(ccall((:gd_sht_scalar_synth, libgdhip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Float64}, Cint),
        p.ptr, re, im, nr, phys, GD_HOST))
    return nothing
end

"""
    batch_shtnskit_transforms!(p, specs, physs)

batch_shtnskit_transforms! (shtnskit_field_functions.jl:651-663): the
syntheses of many fields of one plan.  specs[k] = (data_real, data_imag),
each (nlm, 1, nr_k); physs[k] :: (nlat, nlon, nr_k).  The fields are stacked
along r and synthesised in one batched pass (one Legendre launch and one
ring launch), then split back; the results equal the one-by-one calls.
"""
function batch_shtnskit_transforms!(p::HIPSHTPlan, specs::AbstractVector, physs::AbstractVector)
    length(specs) == length(physs) || throw(ArgumentError("batch_shtnskit_transforms!: $(length(specs)) spectral vs $(length(physs)) physical fields"))
    isempty(specs) && return nothing
    re = cat([s[1] for s in specs]...; dims=3)
    im = cat([s[2] for s in specs]...; dims=3)
    out = Array{Float64,3}(undef, size(physs[1], 1), size(physs[1], 2), size(re, 3))
    shtnskit_spectral_to_physical!(p, re, im, out)
    k = 0
    for ph in physs
        n = size(ph, 3)
        ph .= view(out, :, :, k+1:k+n)
        k += n
    end
    return nothing
end

"shtnskit_physical_to_spectral! (shtnskit_field_functions.jl:131)"
function shtnskit_physical_to_spectral!(p::HIPSHTPlan, phys::Array{Float64,3}, re::Array{Float64,3},
                                        im::Array{Float64,3})
    nr = size(re, 3)
    GC.@preserve re im phys check(ccall((:gd_sht_scalar_analysis, libgdhip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}, Ptr{Float64}, Cint),
        p.ptr, phys, nr, re, im, GD_HOST))
    return nothing
end

"shtnskit_vector_synthesis! (shtnskit_field_functions.jl:252): S := pol, T := tor"
function shtnskit_vector_synthesis!(p::HIPSHTPlan, tor_re, tor_im, pol_re, pol_im, vt, vp)
    nr = size(tor_re, 3)
    GC.@preserve tor_re tor_im pol_re pol_im vt vp check(ccall((:gd_sht_vector_synth, libgdhip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Float64},
         Ptr{Float64}, Cint),
        p.ptr, tor_re, tor_im, pol_re, pol_im, nr, vt, vp, GD_HOST))
    return nothing
end

"shtnskit_vector_analysis! (shtnskit_field_functions.jl:291)"
function shtnskit_vector_analysis!(p::HIPSHTPlan, vt, vp, tor_re, tor_im, pol_re, pol_im)
    nr = size(tor_re, 3)
    GC.@preserve vt vp tor_re tor_im pol_re pol_im check(ccall((:gd_sht_vector_analysis, libgdhip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Cint),
        p.ptr, vt, vp, nr, tor_re, tor_im, pol_re, pol_im, GD_HOST))
    return nothing
end

# ---------------------------------------------------------------- banded LU
"factorize_banded (linear_algebra.jl:23): data is the (2bw+1, N) band"
function factorize_banded(data::Matrix{Float64}, bw::Integer)
    N = size(data, 2)
    lu = similar(data)
    GC.@preserve data lu check(ccall((:gd_banded_factor, libgdhip), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Cint),
        context(), N, bw, 1, data, lu, GD_HOST))
    return lu
end

"solve_banded! (linear_algebra.jl:59)"
function solve_banded!(x::Vector{Float64}, lu::Matrix{Float64}, bw::Integer, b::Vector{Float64})
    x === b || copyto!(x, b)
    GC.@preserve lu x check(ccall((:gd_banded_solve, libgdhip), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Ptr{Cint}, Cint, Ptr{Float64}, Cint),
        context(), length(x), bw, 1, lu, C_NULL, 1, x, GD_HOST))
    return x
end

# ---------------------------------------------------------------- CNAB2 operators
mutable struct HIPImplicitMatrices
    ptr::Ptr{Cvoid}
    lmax::Int; nr::Int; dt::Float64; theta::Float64
end

"create_shtns_timestepping_matrices (timestep.jl:498); r is domain.r (N, 7)"
function create_shtns_timestepping_matrices(lmax::Integer, r::Matrix{Float64}, diffusivity::Float64,
                                            dt::Float64; theta::Float64=0.5, bw::Integer=3,
                                            bc_rows::Bool=false)
    nr = size(r, 1)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve r check(ccall((:gd_cnab2_mats_create, libgdhip), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Cdouble, Cdouble, Cdouble, Cint, Cint, Ref{Ptr{Cvoid}}),
        context(), lmax, nr, bw, r, diffusivity, dt, theta, GD_FD_TRANSPOSE, bc_rows, h))
    M = HIPImplicitMatrices(h[], lmax, nr, dt, theta)
    finalizer(M) do q
        ccall((:gd_cnab2_mats_destroy, libgdhip), Cint, (Ptr{Cvoid},), q.ptr)
    end
    return M
end

"build_rhs_cnab2! (timestep.jl:610)"
function build_rhs_cnab2!(rhs_re, rhs_im, u_re, u_im, nl_re, nl_im, np_re, np_im, dt::Float64,
                          M::HIPImplicitMatrices; lmax::Integer, mmax::Integer=lmax)
    dt == M.dt || error("dt differs from the implicit matrices' dt")
    nr = size(u_re, 3)
    GC.@preserve rhs_re rhs_im u_re u_im nl_re nl_im np_re np_im check(ccall((:gd_build_rhs_cnab2, libgdhip), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint),
        M.ptr, lmax, mmax, nr, u_re, u_im, nl_re, nl_im, np_re, np_im, rhs_re, rhs_im, GD_HOST))
    return rhs_re, rhs_im
end

"solve_implicit_step! (timestep.jl:692)"
function solve_implicit_step!(sol_re, sol_im, rhs_re, rhs_im, M::HIPImplicitMatrices; lmax::Integer,
                              mmax::Integer=lmax)
    nr = size(rhs_re, 3)
    GC.@preserve sol_re sol_im rhs_re rhs_im check(ccall((:gd_solve_implicit_step, libgdhip), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint),
        M.ptr, lmax, mmax, nr, rhs_re, rhs_im, sol_re, sol_im, GD_HOST))
    return sol_re, sol_im
end

# ---------------------------------------------------------------- exponential integrators
mutable struct HIPETDCache
    ptr::Ptr{Cvoid}
    lmax::Int; nr::Int; dt::Float64
end

"create_etd_cache (timestep.jl:37); r is domain.r (N, 7)"
function create_etd_cache(lmax::Integer, r::Matrix{Float64}, diffusivity::Float64, dt::Float64; bw::Integer=3)
    nr = size(r, 1)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve r check(ccall((:gd_etd_cache_create, libgdhip), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Cdouble, Cdouble, Cint, Ref{Ptr{Cvoid}}),
        context(), lmax, nr, bw, r, diffusivity, dt, GD_FD_TRANSPOSE, h))
    c = HIPETDCache(h[], lmax, nr, dt)
    finalizer(q -> ccall((:gd_etd_cache_destroy, libgdhip), Cint, (Ptr{Cvoid},), q.ptr), c)
    return c
end

"eab2_update! (timestep.jl:429): u = E u + dt phi1 (3/2 nl - 1/2 nl_prev), in place"
function eab2_update!(u_re, u_im, nl_re, nl_im, np_re, np_im, etd::HIPETDCache, dt::Float64;
                      lmax::Integer, mmax::Integer=lmax)
    nr = size(u_re, 3)
    GC.@preserve u_re u_im nl_re nl_im np_re np_im check(ccall((:gd_eab2_update, libgdhip), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Cdouble, Cint),
        etd.ptr, lmax, mmax, nr, u_re, u_im, nl_re, nl_im, np_re, np_im, dt, GD_HOST))
    return u_re, u_im
end

mutable struct HIPERK2Cache
    ptr::Ptr{Cvoid}
    nr::Int; dt::Float64
end

"create_erk2_cache (timestep.jl:896), dense matrix functions"
function create_erk2_cache(lmax::Integer, r::Matrix{Float64}, diffusivity::Float64, dt::Float64; bw::Integer=3)
    nr = size(r, 1)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve r check(ccall((:gd_erk2_cache_create, libgdhip), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Cdouble, Cdouble, Cint, Ref{Ptr{Cvoid}}),
        context(), lmax, nr, bw, r, diffusivity, dt, GD_FD_TRANSPOSE, h))
    c = HIPERK2Cache(h[], nr, dt)
    finalizer(q -> ccall((:gd_erk2_cache_destroy, libgdhip), Cint, (Ptr{Cvoid},), q.ptr), c)
    return c
end

"The ERK2Cache constructor (timestep.jl:830): per-l matrices given by the caller"
function create_erk2_cache(dt::Float64, l_values::Vector{Int}, E_half::Vector{Matrix{Float64}},
                           E_full::Vector{Matrix{Float64}}, phi1_half::Vector{Matrix{Float64}},
                           phi1_full::Vector{Matrix{Float64}}, phi2_full::Vector{Matrix{Float64}})
    nr = size(E_full[1], 1)
    flat(v) = reduce(vcat, vec.(v))   # column-major matrices one after another
    mats = map(flat, (E_half, E_full, phi1_half, phi1_full, phi2_full))
    lv = Cint.(l_values)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve lv mats check(ccall((:gd_erk2_cache_from_matrices, libgdhip), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Cint}, Cint, Cdouble, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ref{Ptr{Cvoid}}),
        context(), length(lv), lv, nr, dt, mats..., h))
    c = HIPERK2Cache(h[], nr, dt)
    finalizer(q -> ccall((:gd_erk2_cache_destroy, libgdhip), Cint, (Ptr{Cvoid},), q.ptr), c)
    return c
end

mutable struct HIPERK2Buffers
    ptr::Ptr{Cvoid}
    cache::HIPERK2Cache   # keeps the cache alive while the buffers exist
end

"ERK2FieldBuffers (timestep.jl:1298)"
function ERK2FieldBuffers(cache::HIPERK2Cache; lmax::Integer, mmax::Integer=lmax, nr::Integer=cache.nr)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:gd_erk2_buffers_create, libgdhip), Cint, (Ptr{Cvoid}, Cint, Cint, Cint, Ref{Ptr{Cvoid}}),
                cache.ptr, lmax, mmax, nr, h))
    b = HIPERK2Buffers(h[], cache)
    finalizer(q -> ccall((:gd_erk2_buffers_destroy, libgdhip), Cint, (Ptr{Cvoid},), q.ptr), b)
    return b
end

"erk2_prepare_field! (timestep.jl:1316)"
erk2_prepare_field!(b::HIPERK2Buffers, u_re, u_im, nl_re, nl_im, dt::Float64) =
    GC.@preserve u_re u_im nl_re nl_im check(ccall((:gd_erk2_prepare_field, libgdhip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cint),
        b.ptr, u_re, u_im, nl_re, nl_im, dt, GD_HOST))

"erk2_apply_stage! (timestep.jl:1440)"
erk2_apply_stage!(b::HIPERK2Buffers, u_re, u_im) = GC.@preserve u_re u_im check(ccall(
    (:gd_erk2_apply_stage, libgdhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint), b.ptr, u_re, u_im, GD_HOST))

"erk2_store_stage_nonlinear! (timestep.jl:1446)"
erk2_store_stage_nonlinear!(b::HIPERK2Buffers, nl_re, nl_im) = GC.@preserve nl_re nl_im check(ccall(
    (:gd_erk2_store_stage_nonlinear, libgdhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint),
    b.ptr, nl_re, nl_im, GD_HOST))

"erk2_finalize_field! (timestep.jl:1449)"
erk2_finalize_field!(b::HIPERK2Buffers, u_re, u_im, dt::Float64) = GC.@preserve u_re u_im check(ccall(
    (:gd_erk2_finalize_field, libgdhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cint),
    b.ptr, u_re, u_im, dt, GD_HOST))

# ---------------------------------------------------------------- device-resident step
struct GDStepDesc
    E::Cdouble; Pm::Cdouble; Pr::Cdouble; Ra::Cdouble; rratio::Cdouble; dt::Cdouble; theta::Cdouble
    lmax::Cint; mmax::Cint; nlat::Cint; nlon::Cint; nr::Cint; bw::Cint
    prec::Cint; fd_mode::Cint; radials::Cint; coriolis_index::Cint; bc_rows::Cint; full_products::Cint
    radial_force::Cint
end

mutable struct HIPStep
    ptr::Ptr{Cvoid}
end

"The bench step of scripts/bench_timestepping.jl:6-18 on the GPU."
function create_step(; lmax, nlat, nlon, nr, E=1e-3, Pm=5.0, Pr=1.0, Ra=100.0, rratio=0.35, dt=1e-5,
                     theta=0.5, bw=3, precision::Symbol=:fp32, radial_force::Bool=false)
    d = GDStepDesc(E, Pm, Pr, Ra, rratio, dt, theta, lmax, lmax, nlat, nlon, nr, bw, PREC[precision],
                   GD_FD_TRANSPOSE, 0, 0, 1, 0, radial_force ? 1 : 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:gd_step_create, libgdhip), Cint, (Ptr{Cvoid}, Ref{GDStepDesc}, Ref{Ptr{Cvoid}}),
                context(), d, h))
    s = HIPStep(h[])
    finalizer(q -> ccall((:gd_step_destroy, libgdhip), Cint, (Ptr{Cvoid},), q.ptr), s)
    return s
end

set_state!(s::HIPStep, re::Array{Float64}, im::Array{Float64}) = GC.@preserve re im check(
    ccall((:gd_step_set_state, libgdhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint), s.ptr, re, im, GD_HOST))
get_state!(s::HIPStep, re::Array{Float64}, im::Array{Float64}) = GC.@preserve re im check(
    ccall((:gd_step_get_state, libgdhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint), s.ptr, re, im, GD_HOST))
step_once!(s::HIPStep, n::Integer=1) = check(ccall((:gd_step_run_graph, libgdhip), Cint, (Ptr{Cvoid}, Cint), s.ptr, n))

function energies(s::HIPStep)
    out = zeros(3)
    check(ccall((:gd_step_energies, libgdhip), Cint, (Ptr{Cvoid}, Ptr{Float64}), s.ptr, out))
    return out
end

# ---- scalar boundary conditions (thermal.jl:262-313, compositional.jl:318-370,
# scalar_field_common.jl:866-929, 1166-1199; method :tau)
mutable struct HIPScalarBC
    ptr::Ptr{Cvoid}
    nlm::Int
    nr::Int
end

function create_scalar_bc(lmax::Integer, r::Matrix{Float64}; mmax::Integer=lmax, bw::Integer=3)
    nr = size(r, 1)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve r check(ccall((:gd_scalar_bc_create, libgdhip), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Ptr{Float64}, Cint, Ref{Ptr{Cvoid}}),
        context(), lmax, mmax, nr, bw, r, GD_FD_TRANSPOSE, h))
    nlm = sum(min(l, mmax) + 1 for l in 0:lmax)
    b = HIPScalarBC(h[], nlm, nr)
    finalizer(q -> ccall((:gd_scalar_bc_destroy, libgdhip), Cint, (Ptr{Cvoid},), q.ptr), b)
    return b
end

function _scalar_bc!(b::HIPScalarBC, re, im, bc_type_inner::Vector{Int}, bc_type_outer::Vector{Int},
                     boundary_values::Matrix{Float64}, dirichlet::Integer)
    ti, to = Cint.(bc_type_inner), Cint.(bc_type_outer)
    GC.@preserve ti to boundary_values check(ccall((:gd_scalar_bc_set, libgdhip), Cint,
        (Ptr{Cvoid}, Ptr{Cint}, Ptr{Cint}, Ptr{Float64}), b.ptr, ti, to, boundary_values))
    GC.@preserve re im check(ccall((:gd_scalar_bc_apply, libgdhip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Cint), b.ptr, re, im, dirichlet, GD_HOST))
end

apply_temperature_boundary_conditions_spectral!(b::HIPScalarBC, re, im, ti, to, vals) =
    _scalar_bc!(b, re, im, ti, to, vals, 1)
apply_composition_boundary_conditions_spectral!(b::HIPScalarBC, re, im, ti, to, vals) =
    _scalar_bc!(b, re, im, ti, to, vals, 1)
apply_scalar_flux_bc_spectral!(b::HIPScalarBC, re, im, ti, to, vals) = _scalar_bc!(b, re, im, ti, to, vals, 0)

# ---- restart files (outputs_writer.jl:1221-1370): NetCDF classic, the
# reference's variables plus the AB2 history
struct GDRestartMeta
    time::Float64
    last_output_time::Float64
    output_count::Cint
    restart_count::Cint
    grid_file_written::Cint
    mpi_rank::Cint
    mpi_total_ranks::Cint
    geometry::NTuple{32,UInt8}
end

function GDRestartMeta(time, last_output_time, output_count, restart_count, grid_file_written, rank, nranks,
                       geometry::String="shell")
    g = zeros(UInt8, 32)
    copyto!(g, 1, codeunits(geometry), 1, min(31, ncodeunits(geometry)))
    GDRestartMeta(time, last_output_time, output_count, restart_count, grid_file_written, rank, nranks, Tuple(g))
end

"""write_restart!(s, filename, tracker_fields...; precision=:fp64): the step's
state and AB2 history in the file the reference's write_restart! would name."""
function write_restart!(s::HIPStep, filename::String; time::Float64, last_output_time::Float64=0.0,
                        output_count::Integer=0, restart_count::Integer=0, grid_file_written::Bool=false,
                        rank::Integer=0, nranks::Integer=1, geometry::String="shell", precision::Symbol=:fp64)
    meta = GDRestartMeta(time, last_output_time, output_count, restart_count, Cint(grid_file_written), rank, nranks,
                         geometry)
    check(ccall((:gd_restart_write, libgdhip), Cint, (Ptr{Cvoid}, Cstring, Ref{GDRestartMeta}, Cint),
                s.ptr, filename, meta, PREC[precision]))
end

"""read_restart!(s, filename) -> metadata NamedTuple; restores state, AB2
history and the step counter."""
function read_restart!(s::HIPStep, filename::String)
    meta = Ref{GDRestartMeta}()
    check(ccall((:gd_restart_read, libgdhip), Cint, (Ptr{Cvoid}, Cstring, Ref{GDRestartMeta}),
                s.ptr, filename, meta))
    m = meta[]
    g = collect(m.geometry)
    geometry = String(g[1:something(findfirst(==(0x00), g), 33) - 1])
    return (current_time=m.time, last_output_time=m.last_output_time, output_count=Int(m.output_count),
            restart_count=Int(m.restart_count), grid_file_written=m.grid_file_written != 0, geometry=geometry,
            current_step=Int(ccall((:gd_step_count, libgdhip), Cint, (Ptr{Cvoid},), s.ptr)))
end

# ---- reverse mode: the step's adjoint behind ChainRulesCore rrules (the hook
# the reference's transforms take through SHTnsKit's AD extension,
# Manifest.toml:657-669).  ChainRulesCore is a dependency of this module.
import ChainRulesCore
using ChainRulesCore: NoTangent, unthunk

mutable struct HIPAdjoint
    ptr::Ptr{Cvoid}
    step::HIPStep
end

"The checkpointed adjoint of `nsteps` steps of `s` (checkpoints every `every` steps)."
function create_adjoint(s::HIPStep, nsteps::Integer; every::Integer=5)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:gd_adjoint_create, libgdhip), Cint, (Ptr{Cvoid}, Cint, Cint, Ref{Ptr{Cvoid}}),
                s.ptr, nsteps, every, h))
    a = HIPAdjoint(h[], s)
    finalizer(q -> ccall((:gd_adjoint_destroy, libgdhip), Cint, (Ptr{Cvoid},), q.ptr), a)
    return a
end

adjoint_forward!(a::HIPAdjoint) = check(ccall((:gd_adjoint_forward, libgdhip), Cint, (Ptr{Cvoid},), a.ptr))
adjoint_seed!(a::HIPAdjoint, re, im) = GC.@preserve re im check(ccall((:gd_adjoint_seed, libgdhip), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint), a.ptr, re, im, GD_HOST))
function adjoint_backward!(a::HIPAdjoint, re, im)
    GC.@preserve re im check(ccall((:gd_adjoint_backward, libgdhip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint), a.ptr, re, im, GD_HOST))
    return re, im
end

const MISFIT_KIND = Dict(:B_r => Cint(0), :u_r => Cint(1), :T => Cint(2))

"""run_steps(s, re, im, nsteps) -> (re, im) of the state after `nsteps` CNAB2
steps from the state (re, im) (5 fields, l-major (nlm, nr) planes each)."""
function run_steps(s::HIPStep, re::Array{Float64}, im::Array{Float64}, nsteps::Integer)
    set_state!(s, re, im)
    check(ccall((:gd_step_run, libgdhip), Cint, (Ptr{Cvoid}, Cint), s.ptr, nsteps))
    ore, oim = similar(re), similar(im)
    get_state!(s, ore, oim)
    return ore, oim
end

function ChainRulesCore.rrule(::typeof(run_steps), s::HIPStep, re::Array{Float64}, im::Array{Float64},
                              nsteps::Integer; every::Integer=5)
    a = create_adjoint(s, nsteps; every=every)
    set_state!(s, re, im)
    adjoint_forward!(a)
    ore, oim = similar(re), similar(im)
    get_state!(s, ore, oim)
    function run_steps_pullback(dy)
        dre, dim = unthunk(dy)
        adjoint_seed!(a, collect(Float64, unthunk(dre)), collect(Float64, unthunk(dim)))
        gre, gim = adjoint_backward!(a, similar(re), similar(im))
        return (NoTangent(), NoTangent(), gre, gim, NoTangent())
    end
    return (ore, oim), run_steps_pullback
end

"""surface_misfit(s, re, im, nsteps, target; kind=:B_r, level) = 1/2 sum (X - target)^2
over the (nlat, nlon) grid, X the field `kind` (B_r = l(l+1) B_pol / r^2,
magnetic.jl:257-259; u_r; T) at radial level `level` (0-based) after `nsteps`
steps from (re, im)."""
function surface_misfit(s::HIPStep, re::Array{Float64}, im::Array{Float64}, nsteps::Integer,
                        target::Matrix{Float64}; kind::Symbol=:B_r, level::Integer, every::Integer=5)
    J, _ = _misfit_forward(s, re, im, nsteps, target, kind, level, every)
    return J
end

function _misfit_forward(s, re, im, nsteps, target, kind, level, every)
    a = create_adjoint(s, nsteps; every=every)
    set_state!(s, re, im)
    adjoint_forward!(a)
    J = Ref{Cdouble}(0)
    GC.@preserve target check(ccall((:gd_adjoint_misfit, libgdhip), Cint,
        (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Ref{Cdouble}, Cint), a.ptr, MISFIT_KIND[kind], level, target, J,
        GD_HOST))
    return J[], a
end

function ChainRulesCore.rrule(::typeof(surface_misfit), s::HIPStep, re::Array{Float64}, im::Array{Float64},
                              nsteps::Integer, target::Matrix{Float64}; kind::Symbol=:B_r, level::Integer,
                              every::Integer=5)
    J, a = _misfit_forward(s, re, im, nsteps, target, kind, level, every)
    function surface_misfit_pullback(dJ)
        gre, gim = adjoint_backward!(a, similar(re), similar(im))
        c = unthunk(dJ)
        return (NoTangent(), NoTangent(), c .* gre, c .* gim, NoTangent(), NoTangent())
    end
    return J, surface_misfit_pullback
end

end # module
