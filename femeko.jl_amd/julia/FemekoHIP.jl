#=
    FemekoHIP: GPU drop-ins for Femeko.jl's P1 magnetostatics path.

    include("src/Femeko.jl"); include("femeko.jl_amd/julia/FemekoHIP.jl")
    using .FemekoHIP

    The functions keep Femeko's names and signatures and ccall
    libfemeko_hip.so (C ABI: include/femeko_hip.h). The mesh is uploaded once
    per MESH object and stays on the GPU; later calls only move the
    per-element coefficient and the results. Julia's 1-based Int32
    connectivity is passed as is (index_base = 1), no shifted copy is made
    (compare Femeko_jl_C++/Magnetostatics/run.jl:50-52).

        stiffnessMatrix(mesh, f)                  src/FEM3D.jl:40-56
        localStiffnessMatrix(mesh, f)             src/FEM3D.jl:59-76
        tangentialStiffnessMatrix(mesh, H, dmu)   src/FEM3D.jl:79-119
        lineSearch(u, du, A, RHS, minStep)        src/FEM3D.jl:123-152
        lagrange(mesh)                            src/FEM3D.jl:274-281
        BoundaryIntegral(mesh, F, shell_id)       src/FEM3D.jl:284-297
        solveSaddle(mesh, mu, b)                  [A Lag; Lag' 0] \ [b; 0],
                                                  Magnetostatics/magneticField.jl:99-108
        solveDirichlet(mesh, mu, fixed, g; ...)   Jacobi-PCG with Dirichlet
                                                  elimination
                                                  (Fluids/fluids3D.jl:142-174)
        magneticField(mesh, u)                    magneticField.jl:111-130
        newtonSpline(mesh, spl, spl_dmu, ...)     the mu(|H|) loop of
                                                  NewtonRaphson.jl:148-212
        magnetostaticField(mesh, inside, M, vol)  Micromagnetics/LL.jl:162-209
        scalarPotential(u, p, t, surfaceT, ...)   the C++ bridge of
                                                  Femeko_jl_C++/Magnetostatics/
                                                  magnetostatics.cpp:49-93
        fieldObjective(mesh, rho, target, Hstar, fixed, g)
                                                  field-target objective; its
                                                  ChainRulesCore.rrule returns
                                                  the adjoint gradient
        BatchSolver(meshes, nrhs), batchSolve!    many meshes and fields in
                                                  one set of launches

    Each MESH's device copy is kept in a WeakKeyDict: it is freed (finalizer)
    once the MESH is unreachable; invalidate!(mesh) frees it at once and must
    be called after mesh.p or mesh.t are changed in place.
=#
module FemekoHIP

using SparseArrays
using LinearAlgebra
import ChainRulesCore

export stiffnessMatrix, localStiffnessMatrix, tangentialStiffnessMatrix, lagrange, BoundaryIntegral,
       lineSearch, solveDirichlet, solveSaddle, magneticField, newtonSpline, magnetostaticField,
       scalarPotential, fieldObjective, fieldObjectiveGradient, BatchSolver, batchSolve!, batchPotential,
       invalidate!

const LIB = joinpath(@__DIR__, "..", "lib", "libfemeko_hip.so")

const FK_PREC_F64 = Int32(1)
const FK_PREC_F32_REFINED = Int32(2)
const FK_PREC_F32_RESTART = Int32(3)
const FK_ASM_GATHER = Int32(2)
const FK_ERR_STATE = Int32(9)
const mu0 = pi * 4e-7

struct FemekoHIPError <: Exception
    status::Int32
    msg::String
end

mutable struct Context
    handle::Ptr{Cvoid}
end

mutable struct DeviceMesh
    handle::Ptr{Cvoid}
    ctx::Context
    nv::Int
    nt::Int
    nnz::Int
end

const _ctx = Ref{Union{Nothing, Context}}(nothing)
const _meshes = WeakKeyDict{Any, DeviceMesh}()

function check(ctx::Context, status, where)
    if status != 0
        msg = unsafe_string(ccall((:fk_context_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx.handle))
        throw(FemekoHIPError(Int32(status), string(where, ": ", msg)))
    end
end

function context(device::Integer = 0)
    if _ctx[] === nothing
        h = Ref{Ptr{Cvoid}}(C_NULL)
        s = ccall((:fk_context_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h)
        s == 0 || throw(FemekoHIPError(Int32(s), "fk_context_create"))
        ctx = Context(h[])
        finalizer(c -> ccall((:fk_context_destroy, LIB), Cint, (Ptr{Cvoid},), c.handle), ctx)
        _ctx[] = ctx
    end
    return _ctx[]
end

function release!(dm::DeviceMesh)
    if dm.handle != C_NULL
        ccall((:fk_mesh_destroy, LIB), Cint, (Ptr{Cvoid},), dm.handle)
        dm.handle = C_NULL
    end
    return nothing
end

# Upload once per MESH (p 3 x nv Float64, t 4 x nt Int32, 1-based)
function device_mesh(mesh)
    dm = get(_meshes, mesh, nothing)
    dm !== nothing && dm.handle != C_NULL && return dm
    ctx = context()
    h = Ref{Ptr{Cvoid}}(C_NULL)
    s = ccall((:fk_mesh_create, LIB), Cint,
              (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int32}, Int64, Int32, Ptr{Float64}, Ref{Ptr{Cvoid}}),
              ctx.handle, mesh.p, mesh.nv, mesh.t, mesh.nt, Int32(1), mesh.VE, h)
    check(ctx, s, "fk_mesh_create")
    nnz = Ref{Int64}(0)
    ccall((:fk_mesh_info, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int32}, Ptr{Int64}), h[], nnz, C_NULL, C_NULL)
    dm = DeviceMesh(h[], ctx, Int(mesh.nv), Int(mesh.nt), Int(nnz[]))
    finalizer(release!, dm)
    _meshes[mesh] = dm
    return dm
end

"Free the device copy of `mesh` now (call after changing mesh.p or mesh.t in place)."
function invalidate!(mesh)
    dm = get(_meshes, mesh, nothing)
    if dm !== nothing
        release!(dm)
        delete!(_meshes, mesh)
    end
    return nothing
end

function pattern(dm::DeviceMesh)
    colptr = Vector{Int32}(undef, dm.nv + 1)
    rowval = Vector{Int32}(undef, dm.nnz)
    check(dm.ctx, ccall((:fk_pattern, LIB), Cint, (Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Int32}),
                        dm.handle, Int32(1), colptr, rowval), "fk_pattern")
    return colptr, rowval
end

function values_generation(dm::DeviceMesh)
    g = Ref{Int64}(0)
    check(dm.ctx, ccall((:fk_values_generation, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}), dm.handle, g),
          "fk_values_generation")
    return g[]
end

# the assembled matrix is exactly symmetric: CSR arrays == CSC arrays
tosparse(dm, nzval) = (cp = pattern(dm);
    SparseMatrixCSC{Float64, Int64}(dm.nv, dm.nv, Int64.(cp[1]), Int64.(cp[2]), nzval))

# Matrices whose values are resident on the device: lineSearch runs there as
# long as no later assembly on the same mesh replaced them (the generation
# counter of fk_values_generation is stored with the matrix). Keyed by the
# identity of A.nzval; weak references keep neither A nor the mesh alive.
const _resident = Dict{UInt, Tuple{WeakRef, WeakRef, Int64}}()

function remember_resident!(A::SparseMatrixCSC{Float64, Int64}, dm::DeviceMesh)
    filter!(kv -> kv[2][1].value !== nothing && kv[2][2].value !== nothing, _resident)
    _resident[objectid(A.nzval)] = (WeakRef(A.nzval), WeakRef(dm), values_generation(dm))
    return A
end

function resident_mesh(A::SparseMatrixCSC{Float64, Int64})
    e = get(_resident, objectid(A.nzval), nothing)
    (e === nothing || e[1].value !== A.nzval || e[2].value === nothing) && return nothing, 0
    return e[2].value::DeviceMesh, e[3]
end

function set_coefficient!(dm::DeviceMesh, f::Vector{Float64})
    check(dm.ctx, ccall((:fk_set_coefficient, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, f),
          "fk_set_coefficient")
end

function assemble!(dm::DeviceMesh)
    check(dm.ctx, ccall((:fk_assemble, LIB), Cint, (Ptr{Cvoid}, Int32), dm.handle, FK_ASM_GATHER), "fk_assemble")
end

function stiffnessMatrix(mesh, f::Vector{Float64} = ones(mesh.nt))
    dm = device_mesh(mesh)
    set_coefficient!(dm, f)
    assemble!(dm)
    nzval = Vector{Float64}(undef, dm.nnz)
    check(dm.ctx, ccall((:fk_values, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, nzval), "fk_values")
    return remember_resident!(tosparse(dm, nzval), dm)
end

function localStiffnessMatrix(mesh, f::Vector{Float64} = ones(mesh.nt))
    dm = device_mesh(mesh)
    Ak = Matrix{Float64}(undef, 16, mesh.nt)
    check(dm.ctx, ccall((:fk_local_stiffness, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                        dm.handle, f, Ak), "fk_local_stiffness")
    return Ak
end

function tangentialStiffnessMatrix(mesh, H_vec::Matrix{Float64}, dmu::Vector{Float64})
    dm = device_mesh(mesh)
    nzval = Vector{Float64}(undef, dm.nnz)
    check(dm.ctx, ccall((:fk_tangential_stiffness, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), dm.handle, H_vec, dmu, nzval),
          "fk_tangential_stiffness")
    return tosparse(dm, nzval)
end

# lineSearch(u, du, A, RHS, minStep) on the GPU with the values of A resident
# there (A from stiffnessMatrix); mutates u, returns the last residue.
function lineSearch(u::Vector{Float64}, du::Vector{Float64}, A::SparseMatrixCSC{Float64, Int64},
                    RHS::Vector{Float64}, minStep::Float64 = 1e-4)
    dm, gen = resident_mesh(A)
    dm === nothing && throw(FemekoHIPError(FK_ERR_STATE,
        "lineSearch: A must be a matrix returned by FemekoHIP.stiffnessMatrix"))
    values_generation(dm) == gen || throw(FemekoHIPError(FK_ERR_STATE,
        "lineSearch: a later assembly on this mesh replaced A's values on the device"))
    res = Ref{Float64}(0.0)
    check(dm.ctx, ccall((:fk_line_search, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Ref{Float64}),
                        dm.handle, u, du, RHS, minStep, res), "fk_line_search")
    return res[]
end

function lagrange(mesh)
    dm = device_mesh(mesh)
    C = Vector{Float64}(undef, mesh.nv)
    check(dm.ctx, ccall((:fk_lagrange, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, C), "fk_lagrange")
    return C
end

function BoundaryIntegral(mesh, F::Vector{Float64}, shell_id)
    dm = device_mesh(mesh)
    ids = Int32[s for s in shell_id]
    RHS = Vector{Float64}(undef, mesh.nv)
    check(dm.ctx, ccall((:fk_boundary_integral, LIB), Cint,
                        (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Int32,
                         Int32, Ptr{Float64}),
                        dm.handle, mesh.ns, mesh.surfaceT, mesh.normal, mesh.AE, F, ids, Int32(length(ids)),
                        Int32(1), RHS), "fk_boundary_integral")
    return RHS
end

# precision of the solves inside the composite calls below: :f64 or :f32
# (FP32-refined; fk_set_inner_precision)
set_inner_precision!(dm::DeviceMesh, precision::Symbol) =
    check(dm.ctx, ccall((:fk_set_inner_precision, LIB), Cint, (Ptr{Cvoid}, Int32), dm.handle,
                        precision == :f64 ? FK_PREC_F64 : FK_PREC_F32_REFINED), "fk_set_inner_precision")

# u of [A Lag; Lag' 0] \ [b; 0] with A = stiffnessMatrix(mesh, mu)
# (magneticField.jl:99-108; e.g. b = -BoundaryIntegral(mesh, mu0*Hext, shell_id)).
# Returns (u, multiplier).
function solveSaddle(mesh, mu::Vector{Float64}, b::Vector{Float64}; rtol = 1e-10, maxiter = 1_000_000,
                     precision = :f64)
    dm = device_mesh(mesh)
    set_coefficient!(dm, mu)
    assemble!(dm)
    set_inner_precision!(dm, precision)
    u = Vector{Float64}(undef, mesh.nv)
    mult = Ref{Float64}(0.0); iters = Ref{Int64}(0); relres = Ref{Float64}(0.0)
    check(dm.ctx, ccall((:fk_solve_neumann, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{Float64}, Float64, Int64, Ptr{Float64}, Ref{Float64}, Ref{Int64},
                         Ref{Float64}),
                        dm.handle, b, rtol, maxiter, u, mult, iters, relres), "fk_solve_neumann")
    return u, mult[]
end

function set_dirichlet!(dm::DeviceMesh, fixed::AbstractVector{Bool}, g::Vector{Float64})
    mask = UInt8.(fixed)
    check(dm.ctx, ccall((:fk_set_dirichlet, LIB), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Ptr{Float64}), dm.handle, mask, g),
          "fk_set_dirichlet")
end

function potential(dm::DeviceMesh)
    u = Vector{Float64}(undef, dm.nv)
    check(dm.ctx, ccall((:fk_potential, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, u), "fk_potential")
    return u
end

function solve!(dm::DeviceMesh; rtol = 1e-10, precision = :f64, inner_rtol = 0.95, maxiter = 1_000_000)
    iters = Ref{Int64}(0); outer = Ref{Int32}(0); relres = Ref{Float64}(0)
    prec = precision == :f64 ? FK_PREC_F64 : precision == :f32_restart ? FK_PREC_F32_RESTART : FK_PREC_F32_REFINED
    check(dm.ctx, ccall((:fk_solve, LIB), Cint,
                        (Ptr{Cvoid}, Int32, Float64, Int64, Float64, Int32, Int32, Ref{Int64}, Ref{Int32}, Ref{Float64}),
                        dm.handle, prec, rtol, maxiter, inner_rtol, Int32(50), Int32(0), iters, outer, relres),
          "fk_solve")
    return (iterations = iters[], outer = outer[], relres = relres[])
end

# Assemble A(mu), fix u = g where fixed[i], Jacobi-PCG for the free nodes.
function solveDirichlet(mesh, mu::Vector{Float64}, fixed::AbstractVector{Bool}, g::Vector{Float64};
                        rtol = 1e-10, precision = :f64, inner_rtol = 0.95, maxiter = 1_000_000)
    dm = device_mesh(mesh)
    set_coefficient!(dm, mu)
    assemble!(dm)
    set_dirichlet!(dm, fixed, g)
    info = solve!(dm; rtol = rtol, precision = precision, inner_rtol = inner_rtol, maxiter = maxiter)
    return potential(dm), info
end

# The mu(|H|) loop of Magnetostatics/NewtonRaphson.jl:148-212 on the GPU with
# the reference's own splines (spl = Spline1D(HofM, mu), spl_dmu =
# Spline1D(HofM, dmu), Dierckx.jl: fields t and c are FITPACK's knots and
# coefficients) on the elements listed in mesh.InsideElements, Dirichlet data
# (fixed, g). precision = :f32 runs each step's update solve as FP32 PCG
# with FP64 residual replacement (fk_set_inner_precision), :f64 in FP64.
# Returns (u, mu, (outer, cg_iterations, change)).
function newtonSpline(mesh, spl, spl_dmu, fixed::AbstractVector{Bool}, g::Vector{Float64};
                      rtol = 1e-10, maxDeviation = 1e-4, maxAtt = 100, switch_tol = Inf, precision = :f64)
    dm = device_mesh(mesh)
    region = zeros(UInt8, mesh.nt)
    region[mesh.InsideElements] .= 0x01
    set_dirichlet!(dm, fixed, g)
    set_inner_precision!(dm, precision)
    nc_mu = length(spl.t) - 4
    nc_dmu = length(spl_dmu.t) - 4
    outer = Ref{Int32}(0)
    cg = Ref{Int64}(0)
    change = Ref{Float64}(0.0)
    check(dm.ctx, ccall((:fk_nonlinear_spline, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{UInt8}, Ptr{Float64}, Int32, Ptr{Float64}, Ptr{Float64}, Int32,
                         Ptr{Float64}, Float64, Float64, Int32, Float64, Ref{Int32}, Ref{Int64}, Ref{Float64}),
                        dm.handle, region, spl.t, Int32(length(spl.t)), spl.c[1:nc_mu], spl_dmu.t,
                        Int32(length(spl_dmu.t)), spl_dmu.c[1:nc_dmu], rtol, maxDeviation, Int32(maxAtt),
                        switch_tol, outer, cg, change), "fk_nonlinear_spline")
    mu = Vector{Float64}(undef, mesh.nt)
    check(dm.ctx, ccall((:fk_get_coefficient, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, mu),
          "fk_get_coefficient")
    return potential(dm), mu, (outer = outer[], cg_iterations = cg[], change = change[])
end

# H (3 x nt) = -grad(u) per element
function magneticField(mesh, u::Vector{Float64})
    dm = device_mesh(mesh)
    check(dm.ctx, ccall((:fk_set_potential, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, u),
          "fk_set_potential")
    H = Matrix{Float64}(undef, 3, mesh.nt)
    check(dm.ctx, ccall((:fk_field, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                        dm.handle, H, C_NULL, C_NULL), "fk_field")
    return H
end

# magnetostaticField of Micromagnetics/LL.jl:162-209, once per time step on
# the resident A = stiffnessMatrix(mesh) (assembled on the first call):
# inside (nt Bool), M (3 x nv), Volumes (nv, the driver's nodal sums of
# inside VE). warm = true starts the CG from the previous call's potential.
const _demag_gen = WeakKeyDict{DeviceMesh, Int64}()

function magnetostaticField(mesh, inside::AbstractVector{Bool}, M::Matrix{Float64}, Volumes::Vector{Float64};
                            rtol = sqrt(eps()), maxiter = 100_000, warm = false, precision = :f64)
    dm = device_mesh(mesh)
    set_inner_precision!(dm, precision)
    if get(_demag_gen, dm, -1) != values_generation(dm)
        # A = stiffnessMatrix(mesh) (permalloy.jl:99), assembled again only
        # when another call replaced the resident values
        set_coefficient!(dm, ones(mesh.nt))
        assemble!(dm)
        _demag_gen[dm] = values_generation(dm)
    end
    ins = UInt8.(inside)
    Hd = Matrix{Float64}(undef, 3, mesh.nv)
    iters = Ref{Int64}(0); relres = Ref{Float64}(0.0)
    check(dm.ctx, ccall((:fk_demag_field, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{UInt8}, Ptr{Float64}, Ptr{Float64}, Float64, Int64, Int32, Ptr{Float64},
                         Ref{Int64}, Ref{Float64}),
                        dm.handle, ins, M, Volumes, rtol, maxiter, Int32(warm), Hd, iters, relres),
          "fk_demag_field")
    return Hd
end

# Drop-in for the C++ bridge call of Femeko_jl_C++/Magnetostatics/run.jl:66-79:
# the same arguments (0-based t, surfaceT and shell as run.jl:50-52 makes
# them); writes u in place.
function scalarPotential(u::Vector{Float64}, p::Matrix{Float64}, t::Matrix{Int32}, surfaceT::Matrix{Int32},
                         normal::Matrix{Float64}, nv::Int32, nt::Int32, ns::Int32, VE::Vector{Float64},
                         mu::Vector{Float64}, Hext::Vector{Float64}, shell::Int32)
    s = @ccall LIB.fk_scalar_potential(u::Ptr{Float64}, p::Ptr{Float64}, t::Ptr{Int32}, surfaceT::Ptr{Int32},
                                       normal::Ptr{Float64}, nv::Int32, nt::Int32, ns::Int32, VE::Ptr{Float64},
                                       mu::Ptr{Float64}, Hext::Ptr{Float64}, shell::Int32)::Cint
    if s != 0
        msg = unsafe_string(@ccall LIB.fk_scalar_potential_last_error()::Cstring)
        throw(FemekoHIPError(Int32(s), msg))
    end
    return u
end

# ---- field-target objective and its reverse rule ---------------------------
# J(rho) = sum_{k: target[k]} VE_k |H_k - Hstar|^2 with mu_k = mu0 mur(rho_k)
# (default mur = 1 + 999 rho^3) and H from the Dirichlet problem u = g on
# fixed. The gradient comes from the adjoint kernel fk_adjoint_field_target
# (dJ/dmu) times mu0 mur'(rho). precision = :f64 (FP64 PCG) or :f32
# (FP32-refined state and adjoint solves, fk_set_inner_precision).
simp(r) = 1.0 + 999.0 * r^3
dsimp(r) = 3.0 * 999.0 * r^2

function _adjoint_state(mesh, rho, target, Hstar, fixed, g, mur, rtol, precision)
    dm = device_mesh(mesh)
    set_coefficient!(dm, mu0 .* mur.(rho))
    assemble!(dm)
    set_dirichlet!(dm, fixed, g)
    solve!(dm; rtol = rtol, precision = precision)
    set_inner_precision!(dm, precision)
    J = Ref{Float64}(0.0); iters = Ref{Int64}(0)
    dJdmu = Vector{Float64}(undef, mesh.nt)
    check(dm.ctx, ccall((:fk_adjoint_field_target, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{UInt8}, Ptr{Float64}, Float64, Ref{Float64}, Ptr{Float64}, Ref{Int64}),
                        dm.handle, UInt8.(target), Vector{Float64}(Hstar), rtol, J, dJdmu, iters),
          "fk_adjoint_field_target")
    return J[], dJdmu
end

function fieldObjective(mesh, rho::Vector{Float64}, target::AbstractVector{Bool}, Hstar, fixed::AbstractVector{Bool},
                        g::Vector{Float64}; mur = simp, rtol = 1e-12, precision = :f64)
    J, _ = _adjoint_state(mesh, rho, target, Hstar, fixed, g, mur, rtol, precision)
    return J
end

"(J, dJ/drho) of fieldObjective from one state solve, one adjoint solve and the gradient kernel."
function fieldObjectiveGradient(mesh, rho::Vector{Float64}, target::AbstractVector{Bool}, Hstar,
                                fixed::AbstractVector{Bool}, g::Vector{Float64}; mur = simp, dmur = dsimp,
                                rtol = 1e-12, precision = :f64)
    J, dJdmu = _adjoint_state(mesh, rho, target, Hstar, fixed, g, mur, rtol, precision)
    return J, dJdmu .* mu0 .* dmur.(rho)
end

function ChainRulesCore.rrule(::typeof(fieldObjective), mesh, rho::Vector{Float64}, target::AbstractVector{Bool},
                              Hstar, fixed::AbstractVector{Bool}, g::Vector{Float64}; mur = simp, dmur = dsimp,
                              rtol = 1e-12, precision = :f64)
    J, grad = fieldObjectiveGradient(mesh, rho, target, Hstar, fixed, g; mur = mur, dmur = dmur, rtol = rtol,
                                     precision = precision)
    function fieldObjective_pullback(dJ)
        NT = ChainRulesCore.NoTangent()
        return (NT, NT, ChainRulesCore.unthunk(dJ) .* grad, NT, NT, NT, NT)
    end
    return J, fieldObjective_pullback
end

# ---- batched solves ----------------------------------------------------------
# nmesh meshes with nrhs (1..4) Dirichlet right-hand sides each, solved
# together (FP32 PCG with FP64 reliable updates); no reference counterpart.
mutable struct BatchSolver
    handle::Ptr{Cvoid}
    ctx::Context
    nv::Vector{Int64}
    nrhs::Int
end

function BatchSolver(meshes::AbstractVector, nrhs::Integer)
    ctx = context()
    nv = Int64[m.nv for m in meshes]
    nt = Int64[m.nt for m in meshes]
    ps = [m.p for m in meshes]
    ts = [m.t for m in meshes]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve ps ts begin
        pp = Ptr{Float64}[pointer(p) for p in ps]
        tp = Ptr{Int32}[pointer(t) for t in ts]
        s = ccall((:fk_batch_create, LIB), Cint,
                  (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int64}, Ptr{Ptr{Float64}}, Ptr{Ptr{Int32}}, Int32, Int32,
                   Ref{Ptr{Cvoid}}),
                  ctx.handle, Int32(length(meshes)), nv, nt, pp, tp, Int32(1), Int32(nrhs), h)
        check(ctx, s, "fk_batch_create")
    end
    b = BatchSolver(h[], ctx, nv, Int(nrhs))
    finalizer(x -> ccall((:fk_batch_destroy, LIB), Cint, (Ptr{Cvoid},), x.handle), b)
    return b
end

# mus[m]: nt coefficients; fixed[m]: nv Bools; gs[m]: nv x nrhs Dirichlet values
function batchSolve!(b::BatchSolver, mus::AbstractVector, fixed::AbstractVector, gs::AbstractVector;
                     rtol = 1e-10, inner_rtol = 0.95, maxiter = 1_000_000)
    nsys = length(b.nv) * b.nrhs
    masks = [UInt8.(f) for f in fixed]
    gv = [Vector{Float64}(vec(g)) for g in gs]
    GC.@preserve mus masks gv begin
        check(b.ctx, ccall((:fk_batch_set_coefficient, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{Float64}}),
                           b.handle, Ptr{Float64}[pointer(m) for m in mus]), "fk_batch_set_coefficient")
        check(b.ctx, ccall((:fk_batch_assemble, LIB), Cint, (Ptr{Cvoid},), b.handle), "fk_batch_assemble")
        check(b.ctx, ccall((:fk_batch_set_dirichlet, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{UInt8}}, Ptr{Ptr{Float64}}),
                           b.handle, Ptr{UInt8}[pointer(m) for m in masks], Ptr{Float64}[pointer(g) for g in gv]),
              "fk_batch_set_dirichlet")
    end
    iters = Vector{Int64}(undef, nsys); outer = Vector{Int32}(undef, nsys); relres = Vector{Float64}(undef, nsys)
    check(b.ctx, ccall((:fk_batch_solve, LIB), Cint,
                       (Ptr{Cvoid}, Float64, Float64, Int32, Int64, Ptr{Int64}, Ptr{Int32}, Ptr{Float64}),
                       b.handle, rtol, inner_rtol, Int32(50), maxiter, iters, outer, relres), "fk_batch_solve")
    return (iterations = iters, outer = outer, relres = relres)
end

"Potentials of mesh m (1-based) of a batch: nv x nrhs."
function batchPotential(b::BatchSolver, m::Integer)
    phi = Matrix{Float64}(undef, b.nv[m], b.nrhs)
    check(b.ctx, ccall((:fk_batch_potential, LIB), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                       b.handle, Int32(m - 1), phi), "fk_batch_potential")
    return phi
end

end # module
