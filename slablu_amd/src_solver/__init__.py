"""SlabLU solvers (drop-in for the reference's src_solver package)."""
