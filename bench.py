"""SlabLU factor+solve throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1000] [--p 22]

Workload (default): Helmholtz -Lap u - kh^2 (1 - b(x)) u = 0 on the unit
square, n=1000, p=22 (50 x 50 leaves, N = 1,102,000 DOFs), 10 points per
wavelength (kh = 630.31), b(x) = Gaussian bumps (amplitude 1, width 0.005)
on a 0.04 lattice with one row and one column removed (L-shaped waveguide),
Dirichlet data from the free-space Green's function.  A step is one full
factor + solve of the mixed-precision path (FP32 leaf/slab stages, FP64
interface LU, FP64-residual refinement of the full discrete system until the face
solution changes by less than 1e-8, relative).  The FP64 path is timed on the
same problem and its solution is the reference for the error of the mixed
solution.

Prints one JSON line (metric, value = DOF/s, ...).  roofline: the leaf
interior solve, run in every refinement step of the mixed path
(k_getrs_wave, HBM-bound: it streams every leaf's LU factors once per
call), timed with HIP events on the stream it runs on against its
algorithmic bytes; traffic is the measured HBM bytes per launch recorded in
profiles/leaf_solve_traffic.json when present.  cpu_baseline: the oracle on
a bounded sample of the same workload on the host cores.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "factor+solve DOF/s at N≈1M p=22: mixed+refined vs FP64 at 1e-8 rel err"
PEAK_HBM_GBS = 8000.0


def leaf_solve_bytes(nboxes, ni, itemsize, nrhs=1):
    """Algorithmic HBM bytes of one leaf interior solve launch: every leaf's
    LU factors (ni x ni) once, the permutation, and the right-hand side read
    and written."""
    return nboxes * (ni * ni * itemsize + ni * 4 + 2 * ni * nrhs * itemsize)


def time_leaf_solve(dom, reps=20):
    """Average duration (ms) of one leaf interior solve launch on the factored
    leaves of dom, with HIP events on the current stream."""
    import torch
    from slablu_amd import hip
    lf = dom.hps.leaf
    M = lf["M"]
    ni = M.shape[1]
    LU = M[:, :, :ni]
    x0 = torch.randn(M.shape[0], ni, 1, dtype=M.dtype, device=M.device)
    x = x0.clone()
    hip.getrs(LU, lf["ipiv"], x, perm=lf["perm"])
    ms = []
    for _ in range(reps):
        x.copy_(x0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        hip.getrs(LU, lf["ipiv"], x, perm=lf["perm"])
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    return sum(ms) / len(ms), M.shape[0], ni, M.element_size()


def measured_traffic():
    """HBM bytes per leaf-solve launch from profiles/leaf_solve_traffic.json
    (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "leaf_solve_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    return d.get("hbm_bytes_per_launch")


def make_problem(n, p, ppw, precision, periodic=False, domain="square", box_x=1.0, pde="waveguide"):
    import torch
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    nwaves = int(n / ppw)
    kh = (nwaves + 0.03) * 2 * math.pi + 1.8
    bfield = bif.bfield_lshaped_waveguide if pde == "waveguide" else bif.bfield_constant
    op, zmap, _ = pdo.get_param_map_and_pdo(domain, bfield, kh)
    a = 1 / (2 * (n / (p - 2)))
    dom = Domain_Solver(torch.tensor([[0, box_x], [0, 1.0]]), op, kh, a, p=p, buf_constant=1.0,
                        precision=precision, periodic_bc=periodic)
    uu_dir = lambda xx: bif.uu_dir_func_greens(zmap(torch.as_tensor(xx)), kh)
    return dom, uu_dir, kh


def run_path(dom, uu_dir, steps, warmup, leaf_events):
    import torch
    ts, tf, tsol = [], [], []
    leaf_ms = []
    stages = {}
    sol = None
    for it in range(warmup + steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if leaf_events is not None and it >= warmup:
            dom.hps.leaf_event_pair = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        else:
            dom.hps.leaf_event_pair = None
        info = dom.build(verbose=False)
        sol, rel, _, _, toc = dom.solve(uu_dir)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if it >= warmup:
            ts.append(t1 - t0)
            tf.append(info["toc_build"])
            tsol.append(toc)
            for k in ("toc_DtN", "toc_Aloc", "toc_Ared", "toc_Tred", "toc_Tform"):
                stages[k] = stages.get(k, 0.0) + 1e3 * info[k] / steps
            if dom.hps.leaf_event_pair is not None:
                e0, e1 = dom.hps.leaf_event_pair
                leaf_ms.append(e0.elapsed_time(e1))
    return ts, tf, tsol, leaf_ms, sol, stages


def cpu_baseline(n, p, ppw):
    """The oracle (numpy/scipy FP64) on a bounded sample of the workload:
    leaf DtNs of 100 leaves, the slab stage of one slab, two interface LU
    steps and the interface solve, scaled to the full problem."""
    import numpy as np
    import scipy.linalg
    from oracle import hps_oracle as O
    cores = len(os.sched_getaffinity(0))
    nwaves = int(n / ppw)
    kh = (nwaves + 0.03) * 2 * math.pi + 1.8
    a = 1 / (2 * (n / (p - 2)))
    import torch
    from slablu_amd.src_disc import built_in_funcs as bif
    bfun = lambda xx, k: bif.bfield_lshaped_waveguide(torch.as_tensor(xx), k).numpy()
    hps = O.HPSOracle(O.helmholtz_pdo(kh, bfun), np.array([[0, 1.0], [0, 1.0]]), a, p)
    part = O.SlabPartition(hps.n0, hps.n1, p, 1.0)
    nleaf = min(100, hps.nboxes)
    t0 = time.perf_counter()
    hps.dtns(np.arange(nleaf))
    t_leaf = (time.perf_counter() - t0) * hps.nboxes / nleaf
    # one slab of the two-level slab stage (elim LU, rem Schur + LU, dense T_slab)
    DtN = np.zeros((hps.nboxes, 4 * (p - 2), 4 * (p - 2)))
    boxes = np.arange(part.buf * hps.n1)
    DtN[boxes] = hps.dtns(boxes)
    A = hps.sparse_mat(DtN)
    sl = O.SlabLUOracle(A, part)
    t0 = time.perf_counter()
    sl.build_slab(0)
    t_slab = (time.perf_counter() - t0) * part.Npan
    L = hps.n1 * (p - 2)
    rng = np.random.default_rng(0)
    S = rng.standard_normal((L, L)) + L * np.eye(L)
    t0 = time.perf_counter()
    for _ in range(2):
        lu = scipy.linalg.lu_factor(S)
        S2 = S - S @ scipy.linalg.lu_solve(lu, S)
    t_tred = (time.perf_counter() - t0) / 2 * (part.Npan - 1)
    t_total = t_leaf + t_slab + t_tred
    N = O.dof_count(p, hps.n0, hps.n1)
    return dict(value=N / t_total, unit="DOF/s", cores=cores, kind="port",
                sample="oracle (numpy/scipy FP64, %d BLAS threads) on %d of %d leaves, 1 of %d slabs, 2 of %d "
                       "interface LU steps; times scaled to the full %d-DOF problem (leaf %.1f s, slab %.1f s, "
                       "interface %.1f s)" % (cores, nleaf, hps.nboxes, part.Npan, part.Npan - 1, N, t_leaf,
                                             t_slab, t_tred))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1000)
    ap.add_argument("--p", type=int, default=22)
    ap.add_argument("--ppw", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--domain", default="square", choices=["square", "sinusoidal", "annulus", "curvy_annulus"],
                    help="reference rectangle mapped to this domain (the c12 cross term for curved ones)")
    ap.add_argument("--box-x", type=float, default=1.0, help="width of the reference rectangle")
    ap.add_argument("--pde", default="waveguide", choices=["waveguide", "constant"],
                    help="coefficient b(x): L-shaped waveguide bumps, or constant (kh^2)")
    ap.add_argument("--periodic", action="store_true",
                    help="x-periodic slabs (cyclic interface ring), Dirichlet data on bottom and top only")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the SlabLU path on a single MI355X "
              "(--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from oracle import hps_oracle as O
    torch.set_num_threads(len(os.sched_getaffinity(0)))
    n, p, ppw = args.n, args.p, args.ppw
    print("bench: mixed path n=%d p=%d" % (n, p), file=sys.stderr, flush=True)
    extra = dict(periodic=args.periodic, domain=args.domain, box_x=args.box_x, pde=args.pde)
    dom, uu_dir, kh = make_problem(n, p, ppw, "mixed", **extra)
    N = O.dof_count(p, dom.hps.n0, dom.hps.n1)
    ts, tf, tsol, leaf_ms, sol_mixed, stages = run_path(dom, uu_dir, args.steps, args.warmup, leaf_events=True)
    refine_hist = list(dom.refine_history)
    nboxes = dom.hps.nboxes
    n0, n1 = dom.hps.n0, dom.hps.n1
    roof_dom = time_leaf_solve(dom)
    del dom
    torch.cuda.empty_cache()
    fp64 = None
    if not args.no_fp64:
        print("bench: fp64 path", file=sys.stderr, flush=True)
        dom64, uu_dir, _ = make_problem(n, p, ppw, "fp64", **extra)
        ts64, tf64, tsol64, leaf64, sol64, stages64 = run_path(dom64, uu_dir, max(1, args.steps), args.warmup,
                                                                leaf_events=True)
        err = (torch.linalg.norm(sol_mixed - sol64) / torch.linalg.norm(sol64)).item()
        fp64 = dict(dof_per_s=N / (sum(ts64) / len(ts64)), ms_per_step=1e3 * sum(ts64) / len(ts64),
                    factor_ms=1e3 * sum(tf64) / len(tf64), solve_ms=1e3 * sum(tsol64) / len(tsol64),
                    leaf_ms=sum(leaf64) / max(1, len(leaf64)), stages_ms=stages64, err_mixed_vs_fp64=err)
        del dom64
        torch.cuda.empty_cache()
    t_step = sum(ts) / len(ts)
    roof = None
    if roof_dom is not None:
        avg_ms, nb, ni, isz = roof_dom
        nbytes = leaf_solve_bytes(nb, ni, isz)
        ach = nbytes / (avg_ms * 1e-3) / 1e9
        roof = dict(bound="hbm", kernel="k_getrs_wave<float,1>: leaf interior solve A_cc^{-1} r over %d leaves "
                                          "(%d x %d FP32 LU factors), run in every refinement step" % (nb, ni, ni),
                    achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s", frac=ach / PEAK_HBM_GBS,
                    traffic=measured_traffic(), algorithmic_bytes=nbytes, avg_ms=avg_ms)
    cpu = None
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(n, p, ppw)
    out = {
        "metric": METRIC, "value": N / t_step, "unit": "DOF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "Helmholtz, %s, %s, n=%d p=%d ppw=%d kh=%.2f, "
                               "%dx%d leaves, N=%d, mixed precision + FP64 refinement%s"
                               % ("L-shaped waveguide bumps" if args.pde == "waveguide" else "constant coefficient",
                                  "unit square" if args.domain == "square" and args.box_x == 1.0
                                  else "%s domain (reference box %g x 1)" % (args.domain, args.box_x),
                                  n, p, ppw, kh, n0, n1, N,
                                  ", x-periodic" if args.periodic else ""),
                   "N": N, "p": p, "leaves": nboxes, "precision": "mixed", "periodic": bool(args.periodic)},
        "factor_ms": 1e3 * sum(tf) / len(tf), "solve_ms": 1e3 * sum(tsol) / len(tsol),
        "refine_corrections": refine_hist, "stages_ms": stages,
        "fp64": fp64,
        "speedup_vs_fp64": (fp64["ms_per_step"] / (1e3 * t_step)) if fp64 else None,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
