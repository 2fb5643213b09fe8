"""
    GridapMHDHIP

Julia side of the libmhd drop-in: adds a `:hip` solver to GridapMHD through
the reference's own extension points (no existing method changes):

  * `default_solver_params(::Val{:hip})`   (src/parameters.jl:219-287)
  * `uses_petsc(::Val{:hip})`              (src/parameters.jl:289-295)
  * `space_uses_multigrid(::Val{:hip},_)`  (src/parameters.jl:297-303)
  * `_multi_field_style(::Val{:hip})`      (src/fespaces.jl:4-9)
  * `_solver(::Val{:hip}, op, params)`     (src/main.jl:181-190)
  * `Gridap.Algebra.solve!` for the returned solver (src/main.jl:150, 275)

`GridapMHD.main(params)` stays unchanged: with `params[:solver] = :hip` the
Newton / FGMRES / block-preconditioner hot path runs in libmhd.so on the GPU.

The glue hands libmhd plain arrays: per-cell vertex coordinates, cell dof ids
(in libmhd's local order, see `local_order_*`), thin-wall faces, Dirichlet
values and the fluid parameters.  libmhd's RT basis is nodal (normal flux at
the 2x2 Gauss points of each face); Gridap's is the moment-dual basis.  Both
span the same space, so solutions agree; dof values are converted with the
cell-local matrix `rt_basis_change` (evaluate Gridap's dof functionals on
libmhd's shape functions).
"""
module GridapMHDHIP

using Gridap
using Gridap.Algebra
using Gridap.FESpaces
using Gridap.MultiField
using GridapMHD

const libmhd = joinpath(@__DIR__, "..", "libmhd.so")

# ---------------------------------------------------------------- C structs
struct MhdMesh
  ncells::Int32; nfaces_tw::Int32; cell_xyz::Ptr{Float64}; cell_verts::Ptr{Int32}; nverts::Int32; pad::Int32
  tw_face_cell::Ptr{Int32}; tw_face_lface::Ptr{Int8}; tw_tau::Ptr{Float64}; tw_cw::Ptr{Float64}; tw_jw::Ptr{Float64}
end
struct MhdSpaces
  nfree::NTuple{4,Int32}; ndir::Int32; block_style::Int32
  cell_dofs::Ptr{Int32}; rt_sign::Ptr{Int8}; dir_vals::Ptr{Float64}
end
struct MhdParams
  alpha::Float64; beta::Float64; gamma::Float64; sigma::Float64
  B::NTuple{3,Float64}; f0::NTuple{3,Float64}; fx::NTuple{9,Float64}; g0::NTuple{3,Float64}; gx::NTuple{9,Float64}
  zeta_u::Float64; zeta_j::Float64; convection::Int32; pad::Int32
end
mutable struct MhdSolverOpts
  max_newton::Int32; restart::Int32; max_krylov::Int32; precision::Int32
  rtol::Float64; atol::Float64; ls_rtol::Float64; inner_iters::Int32; verbose::Int32
end
mutable struct MhdSolveInfo
  newton_iters::Int32; krylov_iters::Int32; converged::Int32; flags::Int32
  residual0::Float64; residual::Float64; t_residual::Float64; t_jacobian::Float64; t_setup::Float64; t_solve::Float64
  newton_residuals::NTuple{64,Float64}; krylov_per_step::NTuple{64,Int32}
end
MhdSolveInfo() = MhdSolveInfo(0,0,0,0,0,0,0,0,0,0,ntuple(_->0.0,64),ntuple(_->Int32(0),64))

last_error(ctx) = unsafe_string(ccall((:mhd_last_error, libmhd), Cstring, (Ptr{Cvoid},), ctx))
check(rc, ctx, what) = rc < 0 ? error("$what failed ($rc): $(last_error(ctx))") : rc

# ---------------------------------------------------------------- extension points
function GridapMHD.default_solver_params(::Val{:hip})
  Dict(
    :solver => :hip,
    :matrix_type => SparseMatrixCSC{Float64,Int64},
    :vector_type => Vector{Float64},
    :solver_postpro => ((cache, info) -> hip_postpro(cache, info)),
    :niter => 20, :niter_ls => 400, :rtol => 1e-6, :atol => 1e-8,
    :precision => :fp64,           # :fp64 | :mixed | :mixed_bf16
    :inner_iters => 3, :restart => 200,
    :levels => :auto,              # [u,j] V-cycle levels (x/y semi-coarsening), see attach_levels!
    :coarse_iters => 0,            # 0: exact coarsest solve
    :smoother_weight => 0.175,
    :coarse_step => :minres,       # :minres | :full
  )
end
GridapMHD.uses_petsc(::Val{:hip}) = false
GridapMHD.space_uses_multigrid(::Val{:hip}, solver) = fill(false, 4)
GridapMHD._multi_field_style(::Val{:hip}) = BlockMultiFieldStyle(3, (2, 1, 1), (1, 3, 2, 4))  # ([u,j], p, phi)

struct HIPNewtonSolver <: Gridap.Algebra.NonlinearSolver
  ctx::Base.RefValue{Ptr{Cvoid}}
  opts::MhdSolverOpts
  to_mhd::Function      # Gridap free values -> libmhd free values
  from_mhd::Function    # libmhd free values -> Gridap free values
  keep::Vector{Any}
end

function GridapMHD._solver(::Val{:hip}, op, params)
  U = get_trial(op)
  data = extract_problem(U, params)             # arrays below, built once
  ctx = Ref{Ptr{Cvoid}}(C_NULL)
  rc = ccall((:mhd_create, libmhd), Cint,
             (Ptr{Ptr{Cvoid}}, Ref{MhdMesh}, Ref{MhdSpaces}, Ref{MhdParams}, Int32),
             ctx, data.mesh, data.spaces, data.params, Int32(0))
  check(rc, ctx[], "mhd_create")
  s = params[:solver]
  prec = Dict(:fp64 => 0, :mixed => 1, :mixed_bf16 => 2)[s[:precision]]
  opts = MhdSolverOpts(s[:niter], s[:restart], s[:niter_ls], prec, s[:rtol], s[:atol], s[:rtol] / 10,
                       s[:inner_iters], 0)
  solver = HIPNewtonSolver(ctx, opts, data.to_mhd, data.from_mhd, data.keep)
  finalizer(x -> ccall((:mhd_destroy, libmhd), Cvoid, (Ptr{Cvoid},), x.ctx[]), solver)
  solver
end

function Gridap.Algebra.solve!(x::AbstractVector, ns::HIPNewtonSolver, op::Gridap.Algebra.NonlinearOperator,
                               cache::Nothing)
  xm = ns.to_mhd(x)
  info = MhdSolveInfo()
  rc = ccall((:mhd_newton_solve, libmhd), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ref{MhdSolverOpts}, Ref{MhdSolveInfo}),
             ns.ctx[], xm, ns.opts, info)
  check(rc, ns.ctx[], "mhd_newton_solve")
  copyto!(x, ns.from_mhd(xm))
  (info = info, rc = rc)
end

"""
    attach_levels!(ctx, coarse_ctxs, prolongations, injections; coarse_iters=0, ω=0.175, step=:minres)

Chains coarser contexts (built from `hunt_generate_mesh_hierarchy`-style
x,y-semi-coarsened models) onto `ctx` through `mhd_attach_coarse`; each
prolongation is a `SparseMatrixCSR` fine x coarse on the [u; j] dofs.
"""
function attach_levels!(ctx, coarse_ctxs, prolongations, injections; coarse_iters=0, ω=0.175, step=:minres)
  for lev in length(coarse_ctxs):-1:1
    fine = lev == 1 ? ctx : coarse_ctxs[lev-1]
    P = prolongations[lev]
    rc = ccall((:mhd_attach_coarse, libmhd), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Ptr{Int64}, Ptr{Int32}, Ptr{Float64}, Ptr{Int32}, Int32, Float64,
                Int32),
               fine, coarse_ctxs[lev], size(P, 1), size(P, 2), P.rowptr .- 1, Int32.(P.colval .- 1), P.nzval,
               Int32.(injections[lev] .- 1), coarse_iters, ω, step == :full ? 0 : 1)
    check(rc, fine, "mhd_attach_coarse")
  end
end

function hip_postpro(cache, out)
  out[:nls_iters] = cache.info.newton_iters
  out[:ls_iters] = cache.info.krylov_iters
  out[:nls_residuals] = collect(cache.info.newton_residuals[1:cache.info.newton_iters+1])
end

# ---------------------------------------------------------------- data extraction
"""
    extract_problem(U, params)

Builds the C structs from Gridap objects: cell vertex coordinates
(`get_cell_coordinates`), topological vertex ids (`get_cell_node_ids` of the
periodic topology), cell dof ids of every field (`get_cell_dof_ids`),
renumbered into libmhd's local order and into the [u; j], p, phi block
numbering, Dirichlet values (`get_dirichlet_dof_values`), thin-wall faces of
`params[:bcs][:thin_wall]` (owner cell and local face of each boundary face)
and the fluid parameters of `params[:fluid]`.  `to_mhd` / `from_mhd` apply the
dof permutation and the RT basis change.
"""
function extract_problem end

end # module
