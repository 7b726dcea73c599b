"""Command-line driver with the reference's flags and result pickles
(argparse_driver.py of the reference: flags :17-38, parameter derivation
:86-163, boundary data :195-246, the second of two solves reported
:250-251, pickles :169-187, 253-299).

    python -m slablu_amd.argparse_driver --disc hps --p 22 --n 400 \\
        --pde bfield_constant --domain square --bc free_space --ppw 40 \\
        --solver slabLU --pickle out.pkl

writes the same two pickled dicts in order (build_info, then solve_info)
with the reference's keys, so tests/test_2d.py-style consumers and the
plotting utilities read them unchanged.  Differences, each explicit:

  * HPS only (--disc fd exits with an error); --solver slabLU only;
  * the solver runs on the GPU: --disable_cuda and --sparse_assembly are
    accepted and reported as without effect;
  * --pde: poisson, bfield_constant and bfield_lshaped_waveguide (the
    benchmark's coefficient); the reference's other coefficient fields are
    not provided (exit with an error);
  * added flags: --precision fp64|mixed, --center X Y (Green's-function
    source, the reference's default (-0.1, 0.5)).
"""
import argparse
import pickle

import numpy as np
import torch


def parse(argv=None):
    ap = argparse.ArgumentParser("Call direct solver for 2D domain.")
    ap.add_argument("--disc", type=str, required=True)
    ap.add_argument("--p", type=int, required=False)
    ap.add_argument("--n", type=int, required=True)
    ap.add_argument("--pde", type=str, required=True)
    ap.add_argument("--domain", type=str, required=True)
    ap.add_argument("--box_xlim", type=float, required=False, default=1.0)
    ap.add_argument("--box_ylim", type=float, required=False, default=1.0)
    ap.add_argument("--bc", type=str, required=True)
    ap.add_argument("--ppw", type=int, required=False)
    ap.add_argument("--nwaves", type=float, required=False)
    ap.add_argument("--solver", type=str, required=False)
    ap.add_argument("--sparse_assembly", type=str, required=False, default="reduced_gpu")
    ap.add_argument("--pickle", type=str, required=False)
    ap.add_argument("--store_sol", action="store_true")
    ap.add_argument("--disable_cuda", action="store_true")
    ap.add_argument("--buf_constant", type=float, required=False)
    ap.add_argument("--periodic_bc", action="store_true")
    ap.add_argument("--precision", type=str, default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--center", type=float, nargs=2, default=None)
    return ap.parse_args(argv)


def wavenumber(n, ppw, nwaves):
    """kh = (nwaves + 0.03) 2 pi + 1.8, nwaves = int(n / ppw) or given
    (argparse_driver.py:86-91)."""
    if ppw is None and nwaves is None:
        raise ValueError("oscillatory bfield chosen but ppw and nwaves NOT set")
    if ppw is not None and nwaves is not None:
        raise ValueError("ppw and nwaves both set")
    nw = int(n / ppw) if ppw is not None else nwaves
    return (nw + 0.03) * 2 * np.pi + 1.8


def main(argv=None):
    args = parse(argv)
    from .src_disc import built_in_funcs, pdo
    from .src_solver.slablu_solver import Domain_Solver
    n, disc = args.n, args.disc
    box_geom = torch.tensor([[0, args.box_xlim], [0, args.box_ylim]], dtype=torch.float64)
    print("\n RUNNING PROBLEM WITH %s discretization with %d points. PDE is %s with boundary %s on %s domain with "
          "periodic_bc=%s." % (disc, n, args.pde, args.bc, args.domain, args.periodic_bc))
    if disc != "hps":
        raise SystemExit("only the HPS discretization is provided (--disc hps)")
    if args.solver not in (None, "slabLU"):
        raise SystemExit("only --solver slabLU is provided")
    if args.disable_cuda or args.sparse_assembly != "reduced_gpu":
        print("note: this build factors and solves on the GPU; --disable_cuda / --sparse_assembly have no effect")
    if not torch.cuda.is_available():
        raise SystemExit("no GPU is visible: the SlabLU path runs on the GPU only")
    curved = args.domain != "square"
    geom = {"square": "square", "curved": "sinusoidal", "annulus": "annulus", "curvy_annulus": "curvy_annulus"}
    if args.domain not in geom:
        raise SystemExit("unknown --domain %s" % args.domain)
    if args.pde == "poisson":
        if args.ppw is not None or curved:
            raise SystemExit("poisson runs on the square without --ppw")
        kh = 0
        op, param_map = pdo.PDO_2d(pdo.ones, pdo.ones), (lambda xx: xx)
    elif args.pde in ("bfield_constant", "bfield_lshaped_waveguide"):
        kh = wavenumber(n, args.ppw, args.nwaves)
        print("kh is %5.2f" % kh)
        bfield = getattr(built_in_funcs, args.pde)
        op, param_map, _ = pdo.get_param_map_and_pdo(geom[args.domain], bfield, kh)
    else:
        raise SystemExit("--pde %s: coefficient field not provided (poisson, bfield_constant, "
                         "bfield_lshaped_waveguide)" % args.pde)
    if args.p is None:
        raise SystemExit("HPS selected but p not provided")
    buf_constant = 1.0 if args.buf_constant is None else args.buf_constant
    p = args.p
    a = 1 / (2 * (n / (p - 2)))
    dom = Domain_Solver(box_geom, op, kh, a, p=p, buf_constant=buf_constant, periodic_bc=args.periodic_bc,
                        precision=args.precision)
    N = (p - 2) * (p * dom.hps.n0 * dom.hps.n1 + dom.hps.n0 + dom.hps.n1)

    build_info = dom.build(sparse_assembly=args.sparse_assembly, solver_type="slabLU", verbose=True)
    build_info.update(N=N, n=n, disc=disc, buf=dom.buf_pans * (p - 2), pde=args.pde, bc=args.bc,
                      domain=args.domain, solver=args.solver, sparse_assembly=args.sparse_assembly,
                      box_geom=box_geom, kh=kh, periodic_bc=args.periodic_bc, a=a, p=p, precision=args.precision)

    print("SOLVE RESULTS")
    ff_body, known_sol = None, False
    center = tuple(args.center) if args.center is not None else (-0.1, 0.5)
    if args.bc == "free_space":
        # the Green's function solves the constant-coefficient equation only
        if args.pde != "bfield_constant":
            raise SystemExit("free_space data need --pde bfield_constant")
        uu_dir = lambda xx: built_in_funcs.uu_dir_func_greens(param_map(torch.as_tensor(xx)), kh, center=center)
        known_sol = True
    elif args.bc == "pulse":
        uu_dir = lambda xx: built_in_funcs.uu_dir_pulse(param_map(torch.as_tensor(xx)), kh)
    elif args.bc == "ones":
        uu_dir = lambda xx: torch.ones(torch.as_tensor(xx).shape[0], 1, dtype=torch.float64)
    elif args.bc == "mms":
        if args.pde != "poisson":
            raise SystemExit("mms data are for poisson")
        Lx, Ly = 4 * np.pi, 1
        uu_dir = lambda xx: built_in_funcs.uu_dir_func_mms(xx, Lx, Ly)
        ff_body = lambda xx: built_in_funcs.ff_body_func_mms(xx, Lx, Ly)
        known_sol = True
    elif args.bc == "log_dist":
        if args.pde != "poisson":
            raise SystemExit("log_dist data are for poisson")
        uu_dir = lambda xx: built_in_funcs.uu_dir_func_greens(xx, kh, center=center)
        known_sol = True
    else:
        raise SystemExit("unknown --bc %s" % args.bc)

    # the second of two solves is reported (argparse_driver.py:250-251)
    dom.solve(uu_dir, ff_body, known_sol=known_sol)
    uu_sol, res, true_res, resloc_hps, toc_solve = dom.solve(uu_dir, ff_body, known_sol=known_sol)
    print("\t--Slab solver solved Ax=b residual %5.2e with known solution residual %5.2e and resloc_HPS %5.2e in "
          "time %5.2f s" % (res, true_res, resloc_hps, toc_solve))
    solve_info = dict(res_solve=res, trueres_solve=true_res, resloc_hps_solve=resloc_hps, toc_solve=toc_solve,
                      dof_per_s=N / (build_info["toc_build"] + toc_solve))
    if args.precision == "mixed":
        solve_info.update(refine_converged=dom.refine_converged, refine_history=list(dom.refine_history))
    if args.store_sol:
        print("\t--Storing solution")
        solve_info["xx"] = dom.hps.xx_tot
        solve_info["sol"] = uu_sol
    if args.pickle is not None:
        print("Pickling results to file %s" % args.pickle)
        with open(args.pickle, "wb+") as f:
            pickle.dump(build_info, f)
            pickle.dump(solve_info, f)
    return build_info, solve_info


if __name__ == "__main__":
    main()
