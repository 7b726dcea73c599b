"""Tile autotuner of the multi-right-hand-side solve.

The chain sweeps (elim chains, rem pieces and separators) run one workgroup
per (chain, tile of NR right-hand-side columns).  Which NR is fastest
depends on the block size, the factor form, the number of columns and how
many chains share the GPU, so for a built Domain_Solver and a column count
the autotuner times every candidate width on the solver's own chains (HIP
events on the current stream, on a random right-hand side), keeps the
fastest through slab_set_sweep_tile and records its choices
(solver.tuning, and a JSON file when a path is given).  A candidate whose
solution differs from the built-in rule's by more than a few units of the
working precision (relative, over all columns) is rejected and recorded as
such; it is never selected.
"""
import json

import torch

from . import hip


def _chain_sets(dom):
    """(label, solver) of every block-tridiagonal solver a solve sweeps (the
    rem solver's pieces and separators share one block size)."""
    return [("elim", dom.Aloc_solver), ("rem", dom.Ared_solver)]


def _sweep_kind(solver):
    """(dtype, block size, inverse factors) of the chains a solver sweeps."""
    inv = getattr(solver, "linv", None) is not None or getattr(solver, "plinv", None) is not None
    dt = solver.Tdata.dtype if hasattr(solver, "Tdata") else solver.dtype
    return dt, solver.bs, inv


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def tune_sweeps(dom, nrhs, candidates=(4, 8, 16, 32), reps=3, path=None):
    """Pick the sweep tile width of every chain kind of the built solver dom
    for nrhs right-hand-side columns; returns {label: {"nr", "ms", "timings"}}."""
    result = {}
    for label, solver in _chain_sets(dom):
        dt, bs, inv = _sweep_kind(solver)
        rows = solver.m * solver.bs
        f0 = torch.randn(solver.b, rows, nrhs, dtype=dt, device=dom.device)
        work = torch.empty_like(f0)

        def run():
            work.copy_(f0)
            solver.lusolve_sweep(work)

        timings, rejected = {}, {}
        tol = 1e-10 if dt == torch.float64 else 1e-4
        prev = hip.set_sweep_tile(dt, bs, inv, 0)
        try:
            timings[0] = _time(run, reps)               # the built-in rule
            ref = work.clone()
            scale = torch.linalg.norm(ref).item()
            for nr in candidates:
                if nr > nrhs:
                    continue
                hip.set_sweep_tile(dt, bs, inv, nr)
                t = _time(run, reps)
                diff = torch.linalg.norm(work - ref).item() / max(scale, 1e-300)
                if not diff <= tol:
                    rejected[str(nr)] = diff
                    continue
                timings[nr] = t
        finally:
            hip.set_sweep_tile(dt, bs, inv, prev)
        best = min(timings, key=timings.get)
        hip.set_sweep_tile(dt, bs, inv, best)
        result[label] = dict(dtype=str(dt).replace("torch.", ""), bs=bs, inverse_factors=inv, nrhs=nrhs,
                             nr=best if best else "built-in", ms=timings[best],
                             timings={str(k if k else "built-in"): v for k, v in timings.items()},
                             rejected=rejected)
    dom.tuning = result
    if path is not None:
        with open(path, "w") as fh:
            json.dump(result, fh, indent=1)
    return result
