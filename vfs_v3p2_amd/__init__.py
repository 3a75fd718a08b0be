"""MI355X-native pressure-projection engine for the VFS curvilinear IB flow solver.

The engine is the HIP/C++ shared library built from vfs_v3p2_amd/csrc and driven
through the C ABI declared in include/vfs_pp.h.  `engine` holds the host-side
mirror of the reference's PoissonSolver/UpdatePressure/Projection interface,
`cases` the workload generators.
"""
