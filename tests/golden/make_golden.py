"""Write the golden fixtures under tests/golden/ from the reference.

Run in the development container only (the reference checkout is not present
where the GPU tests run):

    python tests/golden/make_golden.py /root/reference

Only quantities that no closed form or invariant pins are taken from the
reference:
  * Interp_mat (src_disc/hps_subdomain_disc.py:53-80) at the (a, p) pairs of
    the workloads: a least-squares construction whose exact entries depend on
    a and on LAPACK gelsd.
  * End-to-end figures of the reference's own driver path
    (argparse_driver.py:154-275): the known-solution error `trueres_solve`,
    the reduced residual and the leaf residual, for the exact interior solve
    (`solver='superLU'`) and for SlabLU (`solver='slabLU'`, torch RNG seeded),
    plus the superLU-path solution vector at the small configuration.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    sys.path.insert(0, ref_root)
    import torch
    torch.set_default_dtype(torch.double)
    from src_disc import hps_subdomain_disc
    from src_solver.slablu_solver import Domain_Solver
    import src_disc.built_in_funcs as bif
    import src_disc.pdo as pdo
    from src_disc import hps_multidomain_disc
    import scipy.sparse

    # Version skew: the reference indexes scipy CSR matrices with torch index
    # tensors (slablu_solver.py:419, 706, ...), which scipy >= 1.13 rejects.
    # Converting the index objects to numpy leaves every arithmetic step of
    # the reference unchanged.
    def _np_key(k):
        if isinstance(k, tuple):
            return tuple(_np_key(x) for x in k)
        if isinstance(k, torch.Tensor):
            return k.numpy()
        return k

    class _CSR(scipy.sparse.csr_matrix):
        def __getitem__(self, key):
            out = super().__getitem__(_np_key(key))
            return _CSR(out) if scipy.sparse.issparse(out) and out.format == "csr" else out

    _orig_sparse_mat = hps_multidomain_disc.HPS_Multidomain.sparse_mat

    def _sparse_mat(self, device, verbose=False):
        A, t = _orig_sparse_mat(self, device, verbose)
        return _CSR(A), t

    hps_multidomain_disc.HPS_Multidomain.sparse_mat = _sparse_mat

    # ---- Interp_mat fixtures ------------------------------------------------
    interp = {}
    for (n, p) in [(256, 10), (80, 10), (400, 22), (1000, 22), (160, 22), (1848, 16), (476, 16), (112, 16)]:
        a = 1 / (2 * (n / (p - 2)))
        H = hps_subdomain_disc.HPS_Disc(a, p, d=2)
        interp["p%d_n%d" % (p, n)] = np.asarray(H.Interp_mat)
        interp["a_p%d_n%d" % (p, n)] = np.asarray(a)
    np.savez_compressed(os.path.join(HERE, "interp_mat.npz"), **interp)

    # ---- end-to-end figures ---------------------------------------------------
    figures = {}
    solutions = {}

    def run(tag, n, p, ppw, solver, buf_constant=1.0, seed=0, store_sol=False):
        torch.manual_seed(seed)
        nwaves = int(n / ppw)
        kh = (nwaves + 0.03) * 2 * np.pi + 1.8
        op = pdo.PDO_2d(pdo.ones, pdo.ones, c=lambda xx: bif.bfield_constant(xx, kh))
        box_geom = torch.tensor([[0, 1.0], [0, 1.0]])
        a = 1 / (2 * (n / (p - 2)))
        dom = Domain_Solver(box_geom, op, kh, a, p=p, buf_constant=buf_constant)
        dom.build(sparse_assembly="reduced_cpu", solver_type=solver, verbose=False)
        uu_dir = lambda xx: bif.uu_dir_func_greens(xx, kh)
        sol, res, true_res, resloc, toc = dom.solve(uu_dir, None, known_sol=True)
        N = (p - 2) * (p * dom.hps.n[0].item() * dom.hps.n[1].item() + dom.hps.n[0].item() + dom.hps.n[1].item())
        figures[tag] = dict(n=n, p=p, ppw=ppw, solver=solver, buf_constant=buf_constant, kh=kh, N=int(N),
                            res=float(res), trueres=float(true_res), resloc=float(resloc),
                            buf=int(dom.buf_pans), Npan=int(dom.Npan))
        if store_sol:
            solutions[tag] = np.asarray(sol).reshape(-1)
        print(tag, figures[tag], flush=True)

    run("small_superLU", 80, 10, 10, "superLU", store_sol=True)
    run("p16_superLU", 112, 16, 10, "superLU", store_sol=True)
    run("test2d_square_superLU", 400, 22, 40, "superLU")
    run("test2d_square_slabLU", 400, 22, 40, "slabLU")
    with open(os.path.join(HERE, "reference_figures.json"), "w") as fh:
        json.dump(figures, fh, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "reference_solutions.npz"), **solutions)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
