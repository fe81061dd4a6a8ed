"""MI355X-native relativistic HLLC (RHLLC) step for unstructured tetrahedral meshes.

The path is the reference's ``rhllc::Hllc3dStab`` time step
(solvers/rhllc/src/rhllc_3d_stab.cpp:10-65) with its callees, rebuilt as
gfx950 HIP kernels behind a C ABI (include/rhllc_hip.h).
"""
