"""HPS discretization (drop-in for the reference's src_disc package)."""
