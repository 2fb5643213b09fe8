"""MI355X-native backend for GridapMHD's H1-HDiv assembly + block-preconditioned Newton-Krylov hot path.

Public surface (mirrors the reference's entry points):
  main(params) -> (xh, full_params, output)        src/main.jl:122-169
  add_default_params(params)                        src/parameters.jl:85-137
  hunt(...), thin_wall_duct(...)                    src/Applications/hunt.jl, FullyDevelopedFlow.jl
  expansion(...)                                    src/Applications/expansion.jl (gmsh hex meshes)
  meshers.hunt_generate_base_mesh, fespaces.setup_fe_spaces
  mesh_io.read_gmsh / write_gmsh / write_vtu        GridapGmsh, writevtk
  _lib.Context                                      ctypes binding of include/mhd.h
"""
from . import meshers, fespaces, _lib, parameters, multigrid, analytic, postprocess, mesh_io  # noqa: F401
from .parameters import add_default_params, default_solver_params  # noqa: F401
from .main import main, MultiFieldFEFunction  # noqa: F401
from .applications import hunt, thin_wall_duct, fully_developed, expansion  # noqa: F401
