#=
    FemekoHIP: GPU drop-ins for Femeko.jl's P1 magnetostatics path.

    include("src/Femeko.jl"); include("femeko.jl_amd/julia/FemekoHIP.jl")
    using .FemekoHIP

    The functions keep Femeko's names and signatures (src/FEM3D.jl) and
    ccall libfemeko_hip.so (C ABI: include/femeko_hip.h). The mesh is uploaded
    once per MESH object and stays on the GPU; later calls only move the
    per-element coefficient and the results. Julia's 1-based Int32
    connectivity is passed as is (index_base = 1), no shifted copy is made
    (compare Femeko_jl_C++/Magnetostatics/run.jl:50-52).

        stiffnessMatrix(mesh, f)                  src/FEM3D.jl:40-56
        localStiffnessMatrix(mesh, f)             src/FEM3D.jl:59-76
        tangentialStiffnessMatrix(mesh, H, dmu)   src/FEM3D.jl:79-119
        lagrange(mesh)                            src/FEM3D.jl:274-281
        solveDirichlet(mesh, mu, fixed, g; ...)   Jacobi-PCG with Dirichlet
                                                  elimination
                                                  (Fluids/fluids3D.jl:142-174)
        magneticField(mesh, u)                    magneticField.jl:111-130
        rrule(::typeof(fieldObjective), ...)      ChainRulesCore pullback of
                                                  the adjoint gradient
=#
module FemekoHIP

using SparseArrays

export stiffnessMatrix, localStiffnessMatrix, tangentialStiffnessMatrix, lagrange,
       solveDirichlet, magneticField, fieldObjective, newtonSpline

const LIB = joinpath(@__DIR__, "..", "lib", "libfemeko_hip.so")

const FK_PREC_F64 = Int32(1)
const FK_PREC_F32_REFINED = Int32(2)

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
const _meshes = IdDict{Any, DeviceMesh}()

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

# Upload once per MESH (p 3 x nv Float64, t 4 x nt Int32, 1-based)
function device_mesh(mesh)
    get!(_meshes, mesh) do
        ctx = context()
        h = Ref{Ptr{Cvoid}}(C_NULL)
        s = ccall((:fk_mesh_create, LIB), Cint,
                  (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int32}, Int64, Int32, Ptr{Float64}, Ref{Ptr{Cvoid}}),
                  ctx.handle, mesh.p, mesh.nv, mesh.t, mesh.nt, Int32(1), mesh.VE, h)
        check(ctx, s, "fk_mesh_create")
        nnz = Ref{Int64}(0)
        ccall((:fk_mesh_info, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int32}, Ptr{Int64}), h[], nnz, C_NULL, C_NULL)
        dm = DeviceMesh(h[], ctx, Int(mesh.nv), Int(mesh.nt), Int(nnz[]))
        finalizer(d -> ccall((:fk_mesh_destroy, LIB), Cint, (Ptr{Cvoid},), d.handle), dm)
        dm
    end
end

function pattern(dm::DeviceMesh)
    colptr = Vector{Int32}(undef, dm.nv + 1)
    rowval = Vector{Int32}(undef, dm.nnz)
    check(dm.ctx, ccall((:fk_pattern, LIB), Cint, (Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Int32}),
                        dm.handle, Int32(1), colptr, rowval), "fk_pattern")
    return colptr, rowval
end

# the assembled matrix is exactly symmetric: CSR arrays == CSC arrays
tosparse(dm, nzval) = (cp = pattern(dm);
    SparseMatrixCSC{Float64, Int64}(dm.nv, dm.nv, Int64.(cp[1]), Int64.(cp[2]), nzval))

function stiffnessMatrix(mesh, f::Vector{Float64} = ones(mesh.nt))
    dm = device_mesh(mesh)
    check(dm.ctx, ccall((:fk_set_coefficient, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, f),
          "fk_set_coefficient")
    check(dm.ctx, ccall((:fk_assemble, LIB), Cint, (Ptr{Cvoid}, Int32), dm.handle, Int32(2)), "fk_assemble")
    nzval = Vector{Float64}(undef, dm.nnz)
    check(dm.ctx, ccall((:fk_values, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, nzval), "fk_values")
    return tosparse(dm, nzval)
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

function lagrange(mesh)
    dm = device_mesh(mesh)
    C = Vector{Float64}(undef, mesh.nv)
    check(dm.ctx, ccall((:fk_lagrange, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, C), "fk_lagrange")
    return C
end

# Assemble A(mu), fix u = g where fixed[i], Jacobi-PCG for the free nodes.
function solveDirichlet(mesh, mu::Vector{Float64}, fixed::AbstractVector{Bool}, g::Vector{Float64};
                        rtol = 1e-10, precision = :f64, inner_rtol = 0.95, maxiter = 1_000_000)
    dm = device_mesh(mesh)
    check(dm.ctx, ccall((:fk_set_coefficient, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, mu),
          "fk_set_coefficient")
    check(dm.ctx, ccall((:fk_assemble, LIB), Cint, (Ptr{Cvoid}, Int32), dm.handle, Int32(2)), "fk_assemble")
    mask = UInt8.(fixed)
    check(dm.ctx, ccall((:fk_set_dirichlet, LIB), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Ptr{Float64}), dm.handle, mask, g),
          "fk_set_dirichlet")
    iters = Ref{Int64}(0); outer = Ref{Int32}(0); relres = Ref{Float64}(0)
    prec = precision == :f64 ? FK_PREC_F64 : FK_PREC_F32_REFINED
    check(dm.ctx, ccall((:fk_solve, LIB), Cint,
                        (Ptr{Cvoid}, Int32, Float64, Int64, Float64, Int32, Int32, Ref{Int64}, Ref{Int32}, Ref{Float64}),
                        dm.handle, prec, rtol, maxiter, inner_rtol, Int32(50), Int32(0), iters, outer, relres),
          "fk_solve")
    u = Vector{Float64}(undef, mesh.nv)
    check(dm.ctx, ccall((:fk_potential, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, u), "fk_potential")
    return u, (iterations = iters[], outer = outer[], relres = relres[])
end

# The mu(|H|) loop of Magnetostatics/NewtonRaphson.jl:148-212 on the GPU with
# the reference's own splines (spl = Spline1D(HofM, mu), spl_dmu =
# Spline1D(HofM, dmu), Dierckx.jl: fields t and c are FITPACK's knots and
# coefficients) on the elements listed in mesh.InsideElements, Dirichlet data
# (fixed, g). Returns (u, mu, (outer, cg_iterations, change)).
function newtonSpline(mesh, spl, spl_dmu, fixed::AbstractVector{Bool}, g::Vector{Float64};
                      rtol = 1e-10, maxDeviation = 1e-4, maxAtt = 100, switch_tol = Inf)
    dm = device_mesh(mesh)
    region = zeros(UInt8, mesh.nt)
    region[mesh.InsideElements] .= 0x01
    mask = UInt8.(fixed)
    check(dm.ctx, ccall((:fk_set_dirichlet, LIB), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Ptr{Float64}), dm.handle, mask, g),
          "fk_set_dirichlet")
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
    u = Vector{Float64}(undef, mesh.nv)
    check(dm.ctx, ccall((:fk_potential, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, u), "fk_potential")
    mu = Vector{Float64}(undef, mesh.nt)
    check(dm.ctx, ccall((:fk_get_coefficient, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), dm.handle, mu),
          "fk_get_coefficient")
    return u, mu, (outer = outer[], cg_iterations = cg[], change = change[])
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

end # module
