"""MI355X-native Galerkin BEM assembly and dense solve for the Maxwell EFIE/MFIE
on Raviart-Thomas spaces: the ``assemble(op, X, Y; quadstrat)`` path of BEAST.jl
(src/operator.jl:100-116) on hand-written HIP kernels for gfx950, behind a C ABI
(include/beast_hip.h).  This package is the host-side mirror of the reference's
API for that path; all arithmetic runs in the native library.
"""
from .mesh import Mesh, readmesh, writemesh, meshsphere, meshellipsoid, meshcube, meshrectangle
from .rt import RTBasis, raviartthomas, assemblydata, mesh_edges
from .quadrules import trgauss, legendre01
from .operators import (Maxwell3D, MWSingleLayer3D, MWDoubleLayer3D, DoubleLayerRotatedMW3D, NCross, Identity,
                        LocalOperator, PlaneWaveMW, curl,
                        LinearCombinationOfOperators,
                        DoubleNumWiltonSauterQStrat, DoubleNumSauterQstrat, DoubleNumQStrat, SingleNumQStrat,
                        defaultquadstrat, gamma_wavenumber_handler, n, cross, x_hat, y_hat, z_hat,
                        TangTraceMW, CrossTraceMW, FarField)
from .assembly import (Context, default_context, assemble, assemble_rhs, classify, local_blocks, element_to_dof,
                       DeviceMatrix, lu)
from .postproc import (MWFarField3D, MWDoubleLayerFarField3D, MWDoubleLayerRotatedFarField3D,
                       MWFarField3DDropConstant, potential)
from ._native import BHError, UnsupportedOnGPU, SingularException, NoConvergence

__all__ = [k for k in dir() if not k.startswith("_")]
