"""SlabLU factor+solve throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1000] [--p 22]

Headline workload: Helmholtz -Lap u - kh^2 (1 - b(x)) u = 0 on the unit
square, n=1000, p=22 (50 x 50 leaves, N = 1,102,000 DOFs), 10 points per
wavelength (kh = 630.31), b(x) = Gaussian bumps (amplitude 1, width 0.005)
on a 0.04 lattice with one half row and one half column removed (L-shaped
waveguide), Dirichlet data from the free-space Green's function of a source
at (1.5, 1.5).  A step is one full factor + solve of the mixed-precision
path (FP32 leaf/slab stages, FP64 interface LU, FP64-residual refinement of
the full discrete system to 1e-8); value = N / (factor + solve) in DOF/s.

Beside it, in the same JSON line:
  fp64          the plain FP64 path on the same problem; its solution is the
                reference for the error of the mixed one (the run fails when
                that error exceeds 1e-8 or the refinement did not converge);
  cfg2_fp64     the constant-coefficient problem, FP64, same size;
  adjoint       34 x 34 leaves, p=16 (N = 259,896): forward (factor + solve)
                and gradient (adjoint) times of the per-node coefficient;
  multi_rhs     132 x 132 leaves, p=16 (N = 3,906,672), 64 right-hand sides:
                FP64 factor + graph-replayed 64-column solve (autotuned sweep
                tiles), and the mixed + refined form, error against FP64;
  hbs           the mixed path with the slab Schur complements recovered as
                HBS matrices from random samples (rank by the reference's
                rule) instead of formed exactly: T formation time against
                the exact one, compressed size, refined error against FP64;
  stages        per-stage times with their algorithmic work (fixed formulas)
                and the fraction of the MI355X peak they reach;
  rem_factorization  the rem-chain factorization re-run on this build's
                Schur blocks (HIP events), against its algorithmic flops;
  roofline      the dominant kernel of the mixed step, the FP32 MFMA GEMM
                k_gemm<float, int, true, true>: every launch of it in the
                timed steps bracketed by HIP events on its stream (the
                library's launch probe), algorithmic flops 2 m n k and bytes
                (A, B, C read once, C written once) per product summed over
                the launches; bound by HBM when their flop/byte ratio is below
                the FP32 MFMA ridge point (it is: rank-32 updates dominate);
                traffic (HBM bytes per launch) from
                profiles/gemm_f32_traffic.json when present;
  cpu_baseline  the oracle (numpy/scipy FP64) on a bounded sample of the same
                workload, factor and solve phases, extrapolated.
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
PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}
SRC = (1.5, 1.5)


def algorithmic_work(p, n0, n1, buf, Npan):
    """Per-stage algorithmic flops of the factorization (fixed formulas:
    n_i = (p-2)^2, n_b = 4(p-2), L = n1 (p-2), B = buf (p-2))."""
    f = p - 2
    ni, nb_, L, B = f * f, 4 * f, n1 * f, buf * f
    nboxes = n0 * n1
    return dict(
        leaf=nboxes * ((2 / 3) * ni ** 3 + 2 * ni ** 2 * nb_ + 2 * nb_ ** 2 * (ni + nb_)),
        rem_factor=Npan * (n1 - 1) * (14 / 3) * B ** 3,
        tslab=Npan * (n1 - 1) * 6 * B ** 2 * 2 * L,
        interface=(Npan - 1) * (14 / 3) * L ** 3,
    )


def make_problem(n, p, ppw, precision, periodic=False, domain="square", box_x=1.0, pde="waveguide", center=SRC,
                 hbs_rank=None):
    import torch
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    nwaves = int(n / ppw)
    kh = (nwaves + 0.03) * 2 * math.pi + 1.8
    bfield = bif.bfield_lshaped_waveguide if pde == "waveguide" else bif.bfield_constant
    op, zmap, _ = pdo.get_param_map_and_pdo(domain, bfield, kh)
    a = 1 / (2 * (n / (p - 2)))
    dom = Domain_Solver(torch.tensor([[0, box_x], [0, 1.0]]), op, kh, a, p=p, buf_constant=1.0,
                        precision=precision, periodic_bc=periodic, hbs_rank=hbs_rank)
    uu_dir = lambda xx: bif.uu_dir_func_greens(zmap(torch.as_tensor(xx)), kh, center=center)
    return dom, uu_dir, kh


def run_path(dom, uu_dir, steps, warmup, probe=None):
    """warmup + steps factor+solve steps; with probe (a hip.PROBE_* kernel
    id) the launch probe times that kernel over the timed steps."""
    import torch
    from slablu_amd import hip
    ts, tf, tsol = [], [], []
    stages = {}
    sol = None
    probed = None
    for it in range(warmup + steps):
        if probe is not None and it == warmup:
            torch.cuda.synchronize()
            hip.probe_start(probe, 1 << 16)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
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
    if probe is not None:
        probed = hip.probe_stop()
        probed["steps"] = steps
    return dict(t_step=sum(ts) / len(ts), t_factor=sum(tf) / len(tf), t_solve=sum(tsol) / len(tsol),
                stages=stages, sol=sol, zero_pivots=info["zero_pivots"], probe=probed)


def stage_table(dom, stages_ms, solve_ms, precision):
    """Per-stage times against their algorithmic work and the peak of the
    arithmetic they run in."""
    W = algorithmic_work(dom.hps.p, dom.hps.n0, dom.hps.n1, dom.buf_pans, dom.Npan)
    slab_prec = "fp32" if precision == "mixed" else "fp64"
    rows = []

    def add(name, ms, flops, prec):
        tf = flops / (ms * 1e-3) / 1e12 if ms and ms > 0 else float("nan")
        rows.append(dict(stage=name, ms=ms, algorithmic_gflop=flops / 1e9, achieved_tflops=tf,
                         peak_tflops=PEAK_TFLOPS[prec], arithmetic=prec, frac=tf / PEAK_TFLOPS[prec]))

    add("leaf DtN (batched LU + TRSM + GEMM, all leaves)", stages_ms["toc_DtN"], W["leaf"], slab_prec)
    add("T_slab formation (multi-RHS rem sweeps)", stages_ms["toc_Tform"], W["tslab"], slab_prec)
    add("interface block-tridiagonal factorization", stages_ms["toc_Tred"] - stages_ms["toc_Tform"],
        W["interface"], "fp64")
    rows.append(dict(stage="Aloc + Ared (elim LU, couplings, rem Schur + rem block LU)",
                     ms=stages_ms["toc_Aloc"] + stages_ms["toc_Ared"]))
    lf = dom.hps.leaf
    fac_bytes = (lf["M"].numel() * lf["M"].element_size() if lf["M"] is not None else 0) + \
        lf["DtN"].numel() * lf["DtN"].element_size() + dom.Aloc_solver.nbytes + dom.Ared_solver.nbytes + \
        dom.Tred_solver.nbytes
    rows.append(dict(stage="solve (all refinement steps)" if precision == "mixed" else "solve", ms=solve_ms,
                     stored_factor_bytes=fac_bytes,
                     note="HBM-bound: a solve reads the stored factors about once per sweep"))
    return rows, W


def time_rem_factor(dom, reps=5):
    """Average duration (ms) of the rem-chain factorization of this build
    (all slabs: pieces, spikes and separators of the separated chains, or
    the twisted single chains), with HIP events on the current stream, on
    fresh copies of the Schur blocks."""
    import copy
    import torch
    diag, sub, sup = dom.rem_schur_blocks()
    b, m, bs, _ = diag.shape
    proto = dom.Ared_solver
    ms = []
    for _ in range(reps):
        d, l, u = diag.clone(), sub.clone(), sup.clone()
        S = copy.copy(proto)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        S.build_sweep(d, l, u)
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
        del S
    return sum(ms) / len(ms), (b, m, bs, str(diag.dtype).replace("torch.", ""))


def measured_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("hbm_bytes_per_launch")


def hbs_leg(n, p, ppw, extra, sol_fp64, tform_exact_ms):
    """The mixed path with the slab Schur complements recovered as HBS
    matrices (rank by the reference's rule) instead of formed exactly: one
    warm-up and two timed factor + solve steps, the T formation stage, the
    compressed size, and the refined solution against the exact FP64 one."""
    import torch
    dom, uu, _ = make_problem(n, p, ppw, "mixed", hbs_rank="auto", **extra)
    r = run_path(dom, uu, 2, 1)
    h = dom.Tslab_hbs
    err = (torch.linalg.norm(r["sol"] - sol_fp64) / torch.linalg.norm(sol_fp64)).item() if sol_fp64 is not None \
        else None
    out = dict(rank=dom.hbs_rank_used(), leaf_blocks="%d x %d" % (h.m, h.bs), samples_per_side=h.bs + h.k,
               hbs_matrices=h.b, ms_per_step=1e3 * r["t_step"], factor_ms=1e3 * r["t_factor"],
               solve_ms=1e3 * r["t_solve"], tform_ms=r["stages"].get("toc_Tform"), tform_exact_ms=tform_exact_ms,
               compressed_bytes=h.nbytes(), dense_bytes=h.b * h.n * h.n * 8,
               refine_steps=len(dom.refine_history), refine_converged=bool(dom.refine_converged),
               err_vs_fp64=err)
    del dom
    torch.cuda.empty_cache()
    return out


def adjoint_leg(reps=2):
    """Forward factor + solve and adjoint gradient times of the adjoint
    workload (34 x 34 leaves, p = 16, 10 ppw, b = 1 + 0.1 * seeded Fourier
    field, 1024 observation nodes, leaf factors recomputed in 4 chunks)."""
    import torch
    from slablu_amd.src_disc import built_in_funcs as bif
    from slablu_amd.src_solver.adjoint import HelmholtzAdjoint, observed_loss
    n0, p = 34, 16
    n = n0 * (p - 2)
    kh = (int(n / 10) + 0.03) * 2 * math.pi + 1.8
    uu = lambda xx: bif.uu_dir_func_greens(xx, kh, center=SRC)
    per_leaf = (p - 2) ** 2 * ((p - 2) ** 2 + 4 * (p - 2)) * 8
    prob = HelmholtzAdjoint(torch.tensor([[0, 1.0], [0, 1.0]]), kh, 1 / (2 * n0), p, uu, buf_constant=1.0,
                            leaf_memory_budget=n0 * n0 * per_leaf // 4)
    b = bif.fourier_field(prob.xx).cuda()
    obs = prob.observation_nodes(1024).cuda()
    with torch.no_grad():
        d = prob.solve(b + 0.02 * (bif.fourier_field(prob.xx, seed=1).cuda() - 1.0))[obs, 0]
    tf, tb = [], []
    for _ in range(reps + 1):
        bt = b.clone().requires_grad_(True)
        L = observed_loss(prob.solve(bt), obs, d)
        L.backward()
        tf.append(prob.timings["forward_s"])
        tb.append(prob.timings["backward_s"])
    tf, tb = tf[1:], tb[1:]
    N = (p - 2) * (p * n0 * n0 + 2 * n0)
    return dict(workload="Helmholtz -Lap u - kh^2 b u = 0, 34x34 leaves p=16 10ppw kh=%.2f, N=%d, b = 1 + 0.1 * "
                         "(32 seeded Fourier modes |k|<=8), L = 1/2 ||u(obs) - d||^2 at 1024 interior nodes, leaf "
                         "factors recomputed in 4 chunks" % (kh, N),
                forward_ms=1e3 * sum(tf) / len(tf), gradient_ms=1e3 * sum(tb) / len(tb),
                gradient_over_forward=sum(tb) / sum(tf))


def multirhs_leg(reps=2):
    """The batched multi-right-hand-side workload: 132 x 132 leaves, p = 16
    (N = 3,906,672), waveguide coefficient, 10 ppw, 64 Green's-function
    columns from sources uniform in the annulus 1.2 < |x - (0.5, 0.5)| < 2
    (seed 0).  FP64: factor, sweep tiles autotuned, the 64-column solve
    captured as one HIP graph and replayed; mixed: factor and the refined
    64-column solve.  Times are means over reps after a warm-up."""
    import torch
    from slablu_amd import autotune
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    n0, p, nc = 132, 16, 64
    n = n0 * (p - 2)
    kh = (int(n / 10) + 0.03) * 2 * math.pi + 1.8
    op = pdo.PDO_2d(pdo.ones, pdo.ones, c=lambda xx: bif.bfield_lshaped_waveguide(xx, kh))
    srcs = bif.annulus_sources(nc, seed=0)
    uu = lambda xx: bif.uu_dir_func_greens_multi(xx, kh, srcs)
    N = (p - 2) * (p * n0 * n0 + 2 * n0)
    out = dict(workload="Helmholtz, L-shaped waveguide bumps, 132x132 leaves p=16 10ppw kh=%.2f, N=%d, %d "
                        "right-hand sides (annulus sources, seed 0)" % (kh, N, nc))
    ref = None
    for precision in ("fp64", "mixed"):
        dom = Domain_Solver(torch.tensor([[0, 1.0], [0, 1.0]]), op, kh, 1 / (2 * n0), p=p, buf_constant=1.0,
                            precision=precision)
        dom.build(verbose=False)
        tuning = autotune.tune_sweeps(dom, nc)
        tfac, tsol = [], []
        for it in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dom.build(verbose=False)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            if precision == "fp64":
                dom.solve_graph(uu)       # capture (after two eager warm-ups) once per build
                torch.cuda.synchronize()
                ts = time.perf_counter()
                sol = dom.solve_graph(uu)
            else:
                ts = t1
                sol = dom.solve(uu)[0]
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if it > 0:
                tfac.append(t1 - t0)
                tsol.append(t2 - ts)
        fac_ms, sol_ms = 1e3 * sum(tfac) / reps, 1e3 * sum(tsol) / reps
        leg = dict(factor_ms=fac_ms, solve_ms=sol_ms, dof_rhs_per_s=N * nc / ((fac_ms + sol_ms) * 1e-3),
                   tiles={k: v["nr"] for k, v in tuning.items()},
                   rem_chains="%s, %d x %d blocks of %d" % (type(dom.Ared_solver).__name__, dom.Npan, dom.rem_nblocks,
                                                          dom.rem_bs))
        if precision == "fp64":
            leg["solve"] = "64 columns, one captured HIP graph, replayed"
            ref = sol.cpu()
        else:
            leg["solve"] = "64 columns, FP32 two-level preconditioner, FP64 FGMRES refinement to 1e-8"
            leg["refinement_converged"] = bool(dom.refine_converged)
            leg["corrections"] = len(dom.refine_history)
            err = torch.linalg.norm(sol.cpu() - ref, dim=0) / torch.linalg.norm(ref, dim=0)
            leg["max_column_err_vs_fp64"] = err.max().item()
        out[precision] = leg
        del dom, sol
        torch.cuda.empty_cache()
    return out


def cpu_baseline(n, p, ppw):
    """The oracle (numpy/scipy FP64) on a bounded sample of the headline
    workload, extrapolated to the full problem: leaf DtNs of 100 leaves, the
    slab stage of one slab (elim LU, rem Schur + LU, dense T_slab), two
    interface LU steps of the interface size; solve phase: the leaf recovery
    of 100 leaves (re-factored, as the reference's solve does), one slab's
    two slab solves and the interface sweeps."""
    import numpy as np
    import scipy.linalg
    import torch
    from oracle import hps_oracle as O
    from slablu_amd.src_disc import built_in_funcs as bif
    cores = len(os.sched_getaffinity(0))
    kh = O.kh_from_ppw(n, ppw)
    a = 1 / (2 * (n / (p - 2)))

    def with_coef(h):
        # the coefficient at the nodes from the field's cell-local evaluation
        # (an input here, as the coefficient callables are to the solver)
        h.c_nodes = bif.bfield_lshaped_waveguide(torch.as_tensor(h.xx_tot), kh).numpy().reshape(h.nboxes, -1)
        return h

    t0 = time.perf_counter()
    hps = with_coef(O.HPSOracle(O.helmholtz_pdo(kh), np.array([[0, 1.0], [0, 1.0]]), a, p))
    t_coef = time.perf_counter() - t0
    part = O.SlabPartition(hps.n0, hps.n1, p, 1.0)
    nleaf = min(100, hps.nboxes)
    hps.dtns(np.arange(4))            # untimed: BLAS threads up
    t0 = time.perf_counter()
    hps.dtns(np.arange(nleaf))
    t_leaf = (time.perf_counter() - t0) * hps.nboxes / nleaf
    # one slab of the two-level slab stage
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
    C = rng.standard_normal((L, L))
    t0 = time.perf_counter()
    for _ in range(2):
        lu = scipy.linalg.lu_factor(S)
        S2 = S - C @ scipy.linalg.lu_solve(lu, C)
    t_tred = (time.perf_counter() - t0) / 2 * (part.Npan - 1)
    # solve phase
    nC = len(part.I_C)
    f = rng.standard_normal((nC, 1))
    t0 = time.perf_counter()
    sl.slab_solve(0, f)
    sl.slab_solve(0, f)
    t_ssolve = (time.perf_counter() - t0) * part.Npan
    t0 = time.perf_counter()
    for _ in range(2 * (part.Npan - 1)):
        scipy.linalg.lu_solve(lu, f[:L])
    t_isolve = time.perf_counter() - t0
    sub = with_coef(O.HPSOracle(O.helmholtz_pdo(kh), np.array([[0, 2 * a * 10], [0, 2 * a * 10]]), a, p))
    t0 = time.perf_counter()
    sub.leaf_solve(np.zeros((sub.xx_active.shape[0], 1)))
    t_rec = (time.perf_counter() - t0) * hps.nboxes / sub.nboxes
    t_factor = t_coef + t_leaf + t_slab + t_tred
    t_solve = t_ssolve + t_isolve + t_rec
    N = O.dof_count(p, hps.n0, hps.n1)
    return dict(value=N / (t_factor + t_solve), unit="DOF/s", cores=cores, kind="port",
                sample="oracle (numpy/scipy FP64, %d BLAS threads), extrapolated factor+solve: leaf DtNs of %d of %d "
                       "leaves, the slab stage of 1 of %d slabs, 2 of %d interface LU steps (random interface-size "
                       "blocks), the leaf recovery of %d leaves (re-factored), one slab's two slab solves and the "
                       "interface sweeps; scaled to the full %d-DOF problem (factor %.1f s: leaf %.1f, slab %.1f, "
                       "interface %.1f; solve %.1f s)"
                       % (cores, nleaf, hps.nboxes, part.Npan, part.Npan - 1, sub.nboxes, N, t_factor, t_leaf,
                          t_slab, t_tred, t_solve))


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
    ap.add_argument("--no-extras", action="store_true",
                    help="headline mixed path only (no constant-coefficient FP64 leg, no adjoint leg, no roofline "
                         "re-run): one profiled step")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 256 MiB device copy (calibrates the FETCH_SIZE / WRITE_SIZE counters)")
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
    torch.set_num_threads(len(os.sched_getaffinity(0)))
    n, p, ppw = args.n, args.p, args.ppw
    extra = dict(periodic=args.periodic, domain=args.domain, box_x=args.box_x, pde=args.pde)
    print("bench: mixed path n=%d p=%d" % (n, p), file=sys.stderr, flush=True)
    dom, uu_dir, kh = make_problem(n, p, ppw, "mixed", **extra)
    hps = dom.hps
    N = (p - 2) * (p * hps.n0 * hps.n1 + hps.n0 + hps.n1)
    from slablu_amd import hip
    mx = run_path(dom, uu_dir, args.steps, args.warmup, probe=hip.PROBE_GEMM_F32)
    refine = dict(converged=dom.refine_converged, corrections=list(dom.refine_history),
                  residuals=list(dom.refine_residuals))
    stages, W = stage_table(dom, mx["stages"], 1e3 * mx["t_solve"], "mixed")
    roof = None
    pr = mx["probe"]
    if pr and pr["launches"] > 0:
        nl = pr["launches"]
        avg_ms = pr["total_ms"] / nl
        fl, by = pr["flops"] / nl, pr["bytes"] / nl
        tflops = fl / (avg_ms * 1e-3) / 1e12
        gbs = by / (avg_ms * 1e-3) / 1e9
        # arithmetic intensity of the launches against the ridge point of the
        # FP32 MFMA roof (157.3 TFLOP/s over 8 TB/s = 19.7 flop/byte): most
        # launches are rank-32 updates (about 8 flop/byte), so the kernel is
        # bound by HBM
        ai = fl / by
        ridge = PEAK_TFLOPS["fp32"] * 1e12 / (PEAK_HBM_GBS * 1e9)
        hbm = ai < ridge
        roof = dict(bound="hbm" if hbm else "mfma",
                    kernel="k_gemm<float, int, true, true>: FP32 MFMA GEMM (64 x 64 output tiles, 16-byte rows of "
                           "A, B and C), every launch in the timed steps (leaf LU trailing updates, slab Schur "
                           "products; the leaf DtN product runs in its own small-output kernel)",
                    achieved=gbs if hbm else tflops, peak=PEAK_HBM_GBS if hbm else PEAK_TFLOPS["fp32"],
                    unit="GB/s" if hbm else "TFLOP/s",
                    frac=(gbs / PEAK_HBM_GBS) if hbm else (tflops / PEAK_TFLOPS["fp32"]),
                    traffic=measured_traffic("gemm_f32_traffic.json"),
                    arithmetic_intensity=ai, ridge_point=ridge, achieved_tflops=tflops,
                    attainable_tflops=min(PEAK_TFLOPS["fp32"], ai * PEAK_HBM_GBS / 1e3),
                    algorithmic_flops_per_launch=fl, algorithmic_bytes_per_launch=by, avg_ms_per_launch=avg_ms,
                    launches_per_step=nl / pr["steps"], untimed_launches=pr["dropped"],
                    timing="HIP events around each launch on its stream (the library's launch probe), timed steps")
    remf = None
    if not args.no_extras:
        print("bench: rem factorization timing", file=sys.stderr, flush=True)
        avg_ms, shape = time_rem_factor(dom)
        nb_, m_, bs_, dt_ = shape
        remf = dict(chains="%d slabs x %d blocks of %d x %d (%s)" % (nb_, m_, bs_, bs_, dt_),
                    form=type(dom.Ared_solver).__name__, ms=avg_ms, algorithmic_gflop=W["rem_factor"] / 1e9,
                    achieved_tflops=W["rem_factor"] / (avg_ms * 1e-3) / 1e12)
    del dom
    torch.cuda.empty_cache()
    fp64 = None
    accuracy_ok = bool(refine["converged"])
    if not args.no_fp64:
        print("bench: fp64 path", file=sys.stderr, flush=True)
        dom64, uu_dir, _ = make_problem(n, p, ppw, "fp64", **extra)
        r64 = run_path(dom64, uu_dir, max(1, min(args.steps, 3)), min(args.warmup, 1))
        err = (torch.linalg.norm(mx["sol"] - r64["sol"]) / torch.linalg.norm(r64["sol"])).item()
        st64, _ = stage_table(dom64, r64["stages"], 1e3 * r64["t_solve"], "fp64")
        fp64 = dict(dof_per_s=N / r64["t_step"], ms_per_step=1e3 * r64["t_step"], factor_ms=1e3 * r64["t_factor"],
                    solve_ms=1e3 * r64["t_solve"], stages=st64, err_mixed_vs_fp64=err)
        accuracy_ok = accuracy_ok and err <= 1e-8
        del dom64
        torch.cuda.empty_cache()
    cfg2 = adj = mrhs = hbs = None
    if not args.no_extras and not args.no_fp64:
        print("bench: constant-coefficient fp64", file=sys.stderr, flush=True)
        dom2, uu2, _ = make_problem(n, p, ppw, "fp64", pde="constant")
        r2 = run_path(dom2, uu2, max(1, min(args.steps, 3)), min(args.warmup, 1))
        cfg2 = dict(workload="constant-coefficient Helmholtz, n=%d p=%d ppw=%d, FP64" % (n, p, ppw),
                    dof_per_s=N / r2["t_step"], ms_per_step=1e3 * r2["t_step"], factor_ms=1e3 * r2["t_factor"],
                    solve_ms=1e3 * r2["t_solve"])
        del dom2
        torch.cuda.empty_cache()
        print("bench: hbs-compressed T_slab", file=sys.stderr, flush=True)
        hbs = hbs_leg(n, p, ppw, extra, r64["sol"] if fp64 else None, mx["stages"].get("toc_Tform"))
        print("bench: adjoint", file=sys.stderr, flush=True)
        adj = adjoint_leg()
        print("bench: multi-rhs", file=sys.stderr, flush=True)
        mrhs = multirhs_leg()
    if args.calibrate_copy:
        src = torch.ones(64 * 2 ** 20, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
    cpu = None
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(n, p, ppw)
    t_step = mx["t_step"]
    out = {
        "metric": METRIC, "value": N / t_step, "unit": "DOF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "Helmholtz, %s, %s, n=%d p=%d ppw=%d kh=%.2f, %dx%d leaves, N=%d, source at (1.5, "
                               "1.5), mixed precision + FP64 refinement to 1e-8%s"
                               % ("L-shaped waveguide bumps" if args.pde == "waveguide" else "constant coefficient",
                                  "unit square" if args.domain == "square" and args.box_x == 1.0
                                  else "%s domain (reference box %g x 1)" % (args.domain, args.box_x),
                                  n, p, ppw, kh, hps.n0, hps.n1, N, ", x-periodic" if args.periodic else ""),
                   "N": N, "p": p, "leaves": hps.nboxes, "precision": "mixed", "periodic": bool(args.periodic)},
        "factor_ms": 1e3 * mx["t_factor"], "solve_ms": 1e3 * mx["t_solve"],
        "refinement": refine, "accuracy_ok": accuracy_ok, "stages": stages,
        "fp64": fp64, "speedup_vs_fp64": (fp64["ms_per_step"] / (1e3 * t_step)) if fp64 else None,
        "cfg2_fp64": cfg2, "adjoint": adj, "multi_rhs": mrhs, "hbs": hbs, "rem_factorization": remf,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    sys.stdout.flush()
    if not accuracy_ok:
        print("bench: the mixed solution missed the 1e-8 target against FP64 (or the refinement did not converge)",
              file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
