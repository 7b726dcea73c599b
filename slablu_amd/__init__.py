"""MI355X-native SlabLU factor+solve (HPS discretizations of 2-D elliptic PDEs).

The drop-in modules mirror the reference package layout:
  slablu_amd.src_disc    (pdo, built_in_funcs, hps_subdomain_disc,
                          hps_multidomain_disc, hps_parallel_leaf_ops)
  slablu_amd.src_solver  (slablu_solver, block_tridiag_solver_dense,
                          block_tridiag_solver_structured)
All dense linear algebra runs in libslabhip.so (HIP, gfx950) through
slablu_amd.hip.
"""
__version__ = "0.1.0"
