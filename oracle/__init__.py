"""CPU oracle of the voxel elasticity hot path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference's algorithms in NumPy (and
loads the reference's own C++ oracle from oracle/_ref/ where it was built) so that tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg can check and time against it.  The
product package (matrix-free-voxel-elasticity_amd/) never imports it, and no product code
path falls back to it.

Modules: hexref (element algebra, products, layouts), mgref (multigrid, flexible CG,
Chebyshev), solvers (Jacobi-PCG, mixed IR), simpref (cone filter, OC, sensitivity,
problem generator), cref (ctypes loader of oracle/_ref/reference.so).
"""
