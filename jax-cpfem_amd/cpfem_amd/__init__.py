"""cpfem_amd — MI355X (gfx950) crystal-plasticity FE hot path behind the reference's model/solver API.

    from cpfem_amd import CrystalPlasticity, solver, box_mesh, Mesh

All compute runs in libcpfem_hip.so (hand-written HIP kernels, C ABI in include/cpfem.h);
this package is the host-side mirror of the reference's Python surface.
"""
from .generate_mesh import Mesh, box_mesh, get_meshio_cell_type, read_gmsh
from .materials import (CALIB_COPPER, CALIB_STEEL304, CALIB_TANTALUM, COPPER, DP_FERRITE, DP_MARTENSITE, PRESETS,
                        SENS_ROT, STEEL304, TANTALUM, Material)
from .models import CrystalPlasticity
from .problem import BlockEllMatrix, Problem
from .solver import solver
from .adjoint import ad_wrapper, implicit_vjp, simulate, LoadHistory
from .rotations import alpha_rotate_tensor_rank_2, get_rot_mat, rotate_theta_transfer
from .optimize import minimize_lbfgsb, OptimizeOutcome

__all__ = ["CrystalPlasticity", "Problem", "solver", "Mesh", "box_mesh", "get_meshio_cell_type", "read_gmsh", "Material",
           "BlockEllMatrix", "COPPER", "TANTALUM", "STEEL304", "CALIB_COPPER", "CALIB_TANTALUM",
           "CALIB_STEEL304", "SENS_ROT", "DP_FERRITE", "DP_MARTENSITE", "PRESETS", "ad_wrapper", "implicit_vjp", "simulate", "LoadHistory",
           "get_rot_mat", "rotate_theta_transfer", "alpha_rotate_tensor_rank_2", "minimize_lbfgsb", "OptimizeOutcome"]
