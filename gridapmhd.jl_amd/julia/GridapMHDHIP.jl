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
using Gridap.ReferenceFEs
using Gridap.Geometry
using Gridap.Fields: GenericField
using FillArrays: Fill
using LinearAlgebra: dot
using GridapMHD

const libmhd = joinpath(@__DIR__, "..", "libmhd.so")

# ---------------------------------------------------------------- C structs
struct MhdMesh
  ncells::Int32; nfaces_tw::Int32; cell_xyz::Ptr{Float64}; cell_verts::Ptr{Int32}; nverts::Int32; pad::Int32
  tw_face_cell::Ptr{Int32}; tw_face_lface::Ptr{Int8}; tw_tau::Ptr{Float64}; tw_cw::Ptr{Float64}; tw_jw::Ptr{Float64}
end
struct MhdSkeleton
  nfaces::Int32; pad::Int32; face_cell::Ptr{Int32}; face_lface::Ptr{Int8}; face_perm::Ptr{Int8}; mu::Ptr{Float64}
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
    :precision => :fp64,           # :fp64 | :mixed | :mixed_bf16 | :mixed_fp16
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

mutable struct HIPNewtonSolver <: Gridap.Algebra.NonlinearSolver
  ctx::Base.RefValue{Ptr{Cvoid}}
  opts::MhdSolverOpts
  to_mhd::Function      # Gridap free values -> libmhd free values
  from_mhd::Function    # libmhd free values -> Gridap free values
  covec_to_mhd::Function  # a Gridap covector -> libmhd's (transpose of from_mhd's linear part)
  keep::Vector{Any}
end

function GridapMHD._solver(::Val{:hip}, op, params)
  U = get_trial(op)
  data = extract_problem(U, params)             # arrays below, built once
  ctx = Ref{Ptr{Cvoid}}(C_NULL)
  if data.skeleton.nfaces > 0               # params[:bcs][:stabilization]
    rc = ccall((:mhd_create_stabilized, libmhd), Cint,
               (Ptr{Ptr{Cvoid}}, Ref{MhdMesh}, Ref{MhdSpaces}, Ref{MhdParams}, Ref{MhdSkeleton}, Int32),
               ctx, data.mesh, data.spaces, data.params, data.skeleton, Int32(0))
  else
    rc = ccall((:mhd_create, libmhd), Cint,
               (Ptr{Ptr{Cvoid}}, Ref{MhdMesh}, Ref{MhdSpaces}, Ref{MhdParams}, Int32),
               ctx, data.mesh, data.spaces, data.params, Int32(0))
  end
  check(rc, ctx[], "mhd_create")
  s = params[:solver]
  prec = Dict(:fp64 => 0, :mixed => 1, :mixed_bf16 => 2, :mixed_fp16 => 3)[s[:precision]]
  opts = MhdSolverOpts(s[:niter], s[:restart], s[:niter_ls], prec, s[:rtol], s[:atol], s[:rtol] / 10,
                       s[:inner_iters], 0)
  solver = HIPNewtonSolver(ctx, opts, data.to_mhd, data.from_mhd, data.covec_to_mhd, data.keep)
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
function attach_levels!(ctx, coarse_ctxs, prolongations, interpolations; coarse_iters=0, ω=0.175, step=:minres)
  for lev in length(coarse_ctxs):-1:1
    fine = lev == 1 ? ctx : coarse_ctxs[lev-1]
    P, R = prolongations[lev], interpolations[lev]   # SparseMatrixCSR: fine x coarse [u; j], coarse u x fine u
    rc = ccall((:mhd_attach_coarse, libmhd), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Ptr{Int64}, Ptr{Int32}, Ptr{Float64},
                Ptr{Int64}, Ptr{Int32}, Ptr{Float64}, Int32, Float64, Int32),
               fine, coarse_ctxs[lev], size(P, 1), size(P, 2), Int64.(P.rowptr .- 1), Int32.(P.colval .- 1), P.nzval,
               Int64.(R.rowptr .- 1), Int32.(R.colval .- 1), R.nzval, coarse_iters, ω, step == :full ? 0 : 1)
    check(rc, fine, "mhd_attach_coarse")
  end
end

# ---------------------------------------------------------------- differentiation
"""
    pressure_drop_gradient(ns::HIPNewtonSolver, x, g, Ha, Re) -> (J, dJ/dHa, dJ/dcw)

Adjoint gradient of J = g ⋅ x at the converged state `x` (Gridap free values)
for θ = (Ha, cw): A(x)ᵀ λ = g (`mhd_adjoint_solve`, Badia2024's preconditioner
on the transposed operator), dJ/dθ = -λ ⋅ ∂R/∂θ with ∂R/∂γ and ∂R/∂cw from
`mhd_dresidual_dparam` and ∂γ/∂Ha = 2Ha/Re (the :cfd scaling).  `g` is in
Gridap's free-dof basis; it is moved to libmhd's as a covector (the transpose
of the value conversion) and λ ⋅ ∂R/∂θ is basis-independent.
"""
function pressure_drop_gradient(ns::HIPNewtonSolver, x, g, Ha, Re)
  xm = ns.to_mhd(x)
  # covector conversion: J = g ⋅ x = g ⋅ from_mhd(xm) = (from_mhdᵀ g) ⋅ xm; from_mhd is
  # linear, so its transpose is applied column by column of the cell blocks
  gm = ns.covec_to_mhd(g)
  λ = similar(xm)
  info = MhdSolveInfo()
  rc = ccall((:mhd_adjoint_solve, libmhd), Cint,
             (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{MhdSolverOpts}, Ref{MhdSolveInfo}),
             ns.ctx[], xm, gm, λ, ns.opts, info)
  check(rc, ns.ctx[], "mhd_adjoint_solve")
  rc == 0 || error("mhd_adjoint_solve did not converge (rc = $rc)")
  dRγ, dRcw = similar(xm), similar(xm)
  check(ccall((:mhd_dresidual_dparam, libmhd), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Float64}),
              ns.ctx[], xm, Int32(0), dRγ), ns.ctx[], "mhd_dresidual_dparam")
  check(ccall((:mhd_dresidual_dparam, libmhd), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Float64}),
              ns.ctx[], xm, Int32(1), dRcw), ns.ctx[], "mhd_dresidual_dparam")
  J = dot(g, x)
  (J, -(2Ha / Re) * dot(λ, dRγ), -dot(λ, dRcw))
end

"""
The ChainRules reverse rule of the pressure drop, as a maintainer would add it
(ChainRulesCore is not a dependency of this module):

    function ChainRulesCore.rrule(::typeof(pressure_drop), ns, x, g, Ha, Re)
      J, dHa, dcw = GridapMHDHIP.pressure_drop_gradient(ns, x, g, Ha, Re)
      J, ΔJ -> (NoTangent(), NoTangent(), NoTangent(), NoTangent(), ΔJ * dHa, NoTangent(), ΔJ * dcw)
    end
"""
pressure_drop_pullback(ns, x, g, Ha, Re) = begin
  J, dHa, dcw = pressure_drop_gradient(ns, x, g, Ha, Re)
  J, ΔJ -> (ΔJ * dHa, ΔJ * dcw)
end

function hip_postpro(cache, out)
  out[:nls_iters] = cache.info.newton_iters
  out[:ls_iters] = cache.info.krylov_iters
  out[:nls_residuals] = collect(cache.info.newton_residuals[1:cache.info.newton_iters+1])
end

# ---------------------------------------------------------------- data extraction
#
# libmhd's local dof order of a cell (129): u c*27 + a with the Q2 node
# a = i + 3j + 9k at (i/2, j/2, k/2); p the dP1 basis {1-x-y-z, x, y, z};
# j the nodal RT basis (faces z-, z+, y-, y+, x-, x+, 4 Gauss points each,
# then 12 interior); phi dQ1 at vertex v = i + 2j + 4k.  Gridap's local dofs
# are mapped onto these by the node coordinates of its Lagrangian reference
# FEs (u, phi) and by a cell-local basis change for p and j: M[i, k] = the
# i-th Gridap dof functional applied to libmhd's k-th reference shape function.

const G2 = (0.5 - 0.5 / sqrt(3.0), 0.5 + 0.5 / sqrt(3.0))
lag2(t) = ((2t - 1) * (t - 1), 4t * (1 - t), t * (2t - 1))
glin(t) = ((G2[2] - t) / (G2[2] - G2[1]), (t - G2[1]) / (G2[2] - G2[1]))
tangential(d) = d == 1 ? (2, 3) : (d == 2 ? (1, 3) : (1, 2))

"libmhd's reference RT1 shape function r (1-based, 36) at the reference point ξ."
function mhd_rt_shape(r::Int, ξ)
  if r <= 24
    f, k = divrem(r - 1, 4)
    d = 3 - f ÷ 2          # faces z-, z+, y-, y+, x-, x+
    m = isodd(f) ? 3 : 1
  else
    d, k = divrem(r - 25, 4); d += 1
    m = 2
  end
  t1, t2 = tangential(d)
  v = lag2(ξ[d])[m] * glin(ξ[t1])[k % 2 + 1] * glin(ξ[t2])[k ÷ 2 + 1]
  VectorValue(ntuple(c -> c == d ? v : 0.0, 3))
end
mhd_p_shape(k::Int, ξ) = k == 1 ? 1 - ξ[1] - ξ[2] - ξ[3] : ξ[k - 1]

"M[i, k] = Gridap dof i of `reffe` applied to libmhd's shape function k."
function basis_change(reffe, shape, n)
  dofs = get_dof_basis(reffe)
  M = zeros(n, n)
  for k in 1:n
    M[:, k] = evaluate(dofs, GenericField(ξ -> shape(k, ξ)))
  end
  M
end

"Gridap local dof -> libmhd local index, for a Lagrangian reference FE on HEX."
function lagrangian_local_order(reffe, offset, nodes_per_comp, node_index)
  X = get_node_coordinates(reffe)
  d2n = get_dof_to_node(reffe)
  d2c = get_dof_to_comp(reffe)
  [offset + (d2c[i] - 1) * nodes_per_comp + node_index(X[d2n[i]]) for i in eachindex(d2n)]
end
q2_index(x) = round(Int, 2x[1]) + 3round(Int, 2x[2]) + 9round(Int, 2x[3]) + 1
q1_index(x) = round(Int, x[1]) + 2round(Int, x[2]) + 4round(Int, x[3]) + 1

"""
    extract_problem(U, params)

Builds the C structs from Gridap objects: cell vertex coordinates
(`get_cell_coordinates`), topological vertex ids (the periodic grid
topology), cell dof ids of every field (`get_cell_dof_ids`) renumbered into
libmhd's local order and the [u; j], p, phi block numbering, Dirichlet values
(`get_dirichlet_dof_values`), thin-wall faces of `params[:bcs][:thin_wall]`
and the fluid parameters of `params[:fluid]`.  `to_mhd` / `from_mhd` convert
free-value vectors (the p and j basis changes, RT face signs).

Not exercised: there was no Julia where this was written.  The Gridap calls
are those of Gridap 0.17 / 0.18; the face-orientation signs of the RT space
are read from `get_sign_flip`, which assumes Gridap's own orientation rule.
"""
function extract_problem(U, params)
  model = params[:model]
  Ω = Triangulation(model)
  ncells = num_cells(Ω)
  # geometry: Gridap's HEX vertex order is lexicographic, libmhd's v = i + 2j + 4k
  coords = get_cell_coordinates(Ω)
  cell_xyz = Vector{Float64}(undef, 24 * ncells)
  for c in 1:ncells, v in 1:8, i in 1:3
    cell_xyz[24(c - 1) + 3(v - 1) + i] = coords[c][v][i]
  end
  topo = get_grid_topology(model)
  c2v = get_faces(topo, 3, 0)
  cell_verts = Vector{Int32}(undef, 8 * ncells)
  for c in 1:ncells, v in 1:8
    cell_verts[8(c - 1) + v] = c2v.data[c2v.ptrs[c] + v - 1] - 1
  end
  nverts = num_faces(topo, 0)
  # fields in the original order u, p, j, phi; reference FEs as GridapMHD builds them
  Uu, Up, Uj, Uφ = U[1], U[2], U[3], U[4]
  reffe_u = LagrangianRefFE(VectorValue{3,Float64}, HEX, 2)
  reffe_p = LagrangianRefFE(Float64, HEX, 1; space=:P)
  reffe_j = RaviartThomasRefFE(Float64, HEX, 1)
  reffe_φ = LagrangianRefFE(Float64, HEX, 1)
  loc_u = lagrangian_local_order(reffe_u, 0, 27, q2_index)            # libmhd index (1-based) of Gridap dof i
  loc_φ = lagrangian_local_order(reffe_φ, 121, 8, q1_index)
  Mp = basis_change(reffe_p, mhd_p_shape, 4)
  Mj = basis_change(reffe_j, mhd_rt_shape, 36)
  nfree = map(num_free_dofs, (Uu, Up, Uj, Uφ))
  off = Dict(:u => 0, :j => nfree[1], :p => nfree[1] + nfree[3], :φ => nfree[1] + nfree[3] + nfree[2])
  ndir_u = num_dirichlet_dofs(Uu)
  cell_dofs = fill(Int32(0), 129 * ncells)
  ids_u, ids_p, ids_j, ids_φ = map(get_cell_dof_ids, (Uu, Up, Uj, Uφ))
  gid(g, o, ndo) = g > 0 ? Int32(o + g - 1) : Int32(-(ndo - g - 1) - 1)    # free 0-based / Dirichlet -(k+1)
  # p and j: libmhd's local dofs are combinations of Gridap's; p and the
  # interior j dofs are cell-owned, so the cell numbering of Gridap is kept
  # and the values are converted by to_mhd / from_mhd
  for c in 1:ncells
    base = 129(c - 1)
    for (i, g) in enumerate(ids_u[c]); cell_dofs[base + loc_u[i]] = gid(g, off[:u], 0); end
    for (i, g) in enumerate(ids_p[c]); cell_dofs[base + 81 + i] = gid(g, off[:p], 0); end
    for (i, g) in enumerate(ids_j[c]); cell_dofs[base + 85 + i] = gid(g, off[:j], ndir_u); end
    for (i, g) in enumerate(ids_φ[c]); cell_dofs[base + loc_φ[i]] = gid(g, off[:φ], 0); end
  end
  flip = Gridap.FESpaces.get_sign_flip(model, Fill(reffe_j, ncells))   # cell -> dof signs of Gridap's RT space
  # Dirichlet values: u as they are; j face values converted face by face (a
  # face's dofs depend only on that face's shape functions, a 4 x 4 block of Mj)
  dir_u = get_dirichlet_dof_values(Uu)
  dir_jG = get_dirichlet_dof_values(Uj)
  dir_j = copy(dir_jG)
  for c in 1:ncells, k in 1:6
    fb = 4(k - 1) .+ (1:4)
    g = ids_j[c][fb]
    all(<(0), g) || continue
    s = flip[c][fb[1]]
    dir_j[-g] = s .* (Mj[fb, fb] \ (s .* dir_jG[-g]))
  end
  dir_vals = vcat(dir_u, dir_j)
  # free-vector conversion, cell by cell (shared dofs get the same value from every cell)
  Mpi, Mji = inv(Mp), inv(Mj)
  function convert(x, Mp_, Mj_)
    y = copy(x)
    for c in 1:ncells
      gp = [off[:p] + g for g in ids_p[c]]
      y[gp] = Mp_ * x[gp]
      gj = ids_j[c]
      s = [i <= 24 ? flip[c][i] : 1.0 for i in 1:36]
      loc = [g > 0 ? x[off[:j] + g] : (Mj_ === Mj ? dir_j[-g] : dir_jG[-g]) for g in gj] .* s
      new = (Mj_ * loc) .* s
      for (i, g) in enumerate(gj)
        g > 0 && (y[off[:j] + g] = new[i])
      end
    end
    y
  end
  to_mhd(x) = convert(x, Mpi, Mji)
  from_mhd(x) = convert(x, Mp, Mj)
  # the transpose of from_mhd's linear part: u and phi map one to one, p
  # through Mpᵀ per cell, j through s Mjᵀ s per cell, each Gridap face dof
  # taken from the first cell that produces it (every cell produces the same value)
  function covec_to_mhd(g)
    gm = copy(g)
    for c in 1:ncells
      gp = [off[:p] + k for k in ids_p[c]]
      gm[gp] = Mp' * g[gp]
    end
    jr = off[:j] .+ (1:nfree[3])
    gm[jr] .= 0.0
    seen = falses(nfree[3])
    for c in 1:ncells
      gj = ids_j[c]
      s = [i <= 24 ? flip[c][i] : 1.0 for i in 1:36]
      w = zeros(36)
      for (i, k) in enumerate(gj)
        if k > 0 && !seen[k]
          w[i] = g[off[:j] + k]
          seen[k] = true
        end
      end
      t = s .* (Mj' * (s .* w))
      for (i, k) in enumerate(gj)
        k > 0 && (gm[off[:j] + k] += t[i])
      end
    end
    gm
  end
  # thin walls: boundary faces of the tags, as (cell, local face)
  tw_cell, tw_lf, tw_tau, tw_cw, tw_jw = Int32[], Int8[], Float64[], Float64[], Float64[]
  if haskey(params[:bcs], :thin_wall)
    labels = get_face_labeling(model)
    c2f = get_faces(topo, 3, 2)
    f2c = get_faces(topo, 2, 3)
    lf_of = (0, 1, 2, 3, 4, 5)               # Gridap HEX faces: z-, z+, y-, y+, x-, x+ (lexicographic)
    for tw in params[:bcs][:thin_wall]
      mask = get_face_mask(labels, tw[:domain], 2)
      for c in 1:ncells, k in 1:6
        f = c2f.data[c2f.ptrs[c] + k - 1]
        if mask[f] && length(f2c.ptrs[f]:f2c.ptrs[f+1]-1) == 1
          push!(tw_cell, c - 1); push!(tw_lf, lf_of[k]); push!(tw_tau, tw[:τ]); push!(tw_cw, tw[:cw])
          push!(tw_jw, get(tw, :jw, 0.0))
        end
      end
    end
  end
  # skeleton stabilisation (src/weakforms.jl:168-177, 685-688): every face of two
  # cells, the second cell's 3x3 face Gauss point at each point of the first,
  # matched through the face's corner vertex ids (periodic images identified)
  sk_cell, sk_lf, sk_perm, sk_mu = Int32[], Int8[], Int8[], Float64[]
  stab = get(params[:bcs], :stabilization, [])
  stab isa AbstractDict && (stab = [stab])
  if !isempty(stab)
    μ = sum(Float64(st[:μ]) for st in stab)
    c2f = get_faces(topo, 3, 2)
    f2c = get_faces(topo, 2, 3)
    # corner (i0, i1) of local face lf: local vertex side << d | i0 << ta | i1 << tb
    function corner_vertex(lf, k)
      d = 2 - lf ÷ 2
      ta, tb = d == 0 ? (1, 2) : (d == 1 ? (0, 2) : (0, 1))
      ((lf & 1) << d) | ((k & 1) << ta) | ((k >> 1) << tb)
    end
    localface(c, f) = findfirst(k -> c2f.data[c2f.ptrs[c] + k - 1] == f, 1:6) - 1
    bit(v, i) = (v >> i) & 1
    for f in 1:num_faces(topo, 2)
      cs = f2c.data[f2c.ptrs[f]:f2c.ptrs[f+1]-1]
      length(cs) == 2 || continue
      c0, c1 = cs
      l0, l1 = localface(c0, f), localface(c1, f)
      P = [cell_verts[8(c0 - 1) + corner_vertex(l0, k) + 1] for k in 0:3]
      M = [cell_verts[8(c1 - 1) + corner_vertex(l1, k) + 1] for k in 0:3]
      m = [findfirst(==(P[k]), M) - 1 for k in 1:4]     # minus corner of each plus corner
      push!(sk_cell, c0 - 1, c1 - 1); push!(sk_lf, l0, l1); push!(sk_mu, μ)
      for q in 0:8
        qa, qb = q % 3, q ÷ 3
        s0 = 2bit(m[1], 0) + (bit(m[2], 0) - bit(m[1], 0)) * qa + (bit(m[3], 0) - bit(m[1], 0)) * qb
        s1 = 2bit(m[1], 1) + (bit(m[2], 1) - bit(m[1], 1)) * qa + (bit(m[3], 1) - bit(m[1], 1)) * qb
        push!(sk_perm, s0 + 3s1)
      end
    end
  end
  skeleton = MhdSkeleton(length(sk_mu), 0, pointer(sk_cell), pointer(sk_lf), pointer(sk_perm), pointer(sk_mu))
  f = params[:fluid]
  conv = Dict(:none => 0, :picard => 1, :newton => 2)[f[:convection]]
  B, fv = Tuple(Float64.(f[:B])), Tuple(Float64.(f[:f]))
  mparams = MhdParams(f[:α], f[:β], f[:γ], get(f, :σ, 1.0), B, fv, ntuple(_ -> 0.0, 9), (0.0, 0.0, 0.0),
                      ntuple(_ -> 0.0, 9), get(f, :ζᵤ, 0.0), get(f, :ζⱼ, 0.0), Int32(conv), Int32(0))
  # Gridap's dof values carry the RT orientation; libmhd's local face signs are the same flips
  rt_sign = Int8[flip[c][4(k - 1) + 1] > 0 ? 1 : -1 for k in 1:6 for c in 1:ncells]
  rt_sign = reshape(permutedims(reshape(rt_sign, ncells, 6)), :)
  keep = Any[cell_xyz, cell_verts, cell_dofs, dir_vals, rt_sign, tw_cell, tw_lf, tw_tau, tw_cw, tw_jw,
             sk_cell, sk_lf, sk_perm, sk_mu]
  mesh = MhdMesh(ncells, length(tw_cell), pointer(cell_xyz), pointer(cell_verts), nverts, 0,
                 pointer(tw_cell), pointer(tw_lf), pointer(tw_tau), pointer(tw_cw), pointer(tw_jw))
  spaces = MhdSpaces(Int32.((nfree[1], nfree[2], nfree[3], nfree[4])), Int32(length(dir_vals)), Int32(1),
                     pointer(cell_dofs), pointer(rt_sign), pointer(dir_vals))
  (mesh = mesh, spaces = spaces, params = mparams, skeleton = skeleton, to_mhd = to_mhd, from_mhd = from_mhd,
   covec_to_mhd = covec_to_mhd, keep = keep)
end

end # module
