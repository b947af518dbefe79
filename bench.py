#!/usr/bin/env python3
"""Benchmark of the MI355X multigrid solve path.

Workload (one GPU): 3D jump-coefficient diffusion -div(kappa grad u) = f on
the unit cube, trilinear Q1 hexes from one hex refined 7 times (129^3 =
2,146,689 unknowns, 27-point stencil), kappa = 10^U[-2,2] per finest
element, f ~ U[-1,1]; ls + lmgc V(2,2) with multicolour Gauss-Seidel,
base level 4 (17^3 vectors, ex band LU applied as its dense inverse), stop
at 1e-10 relative defect.

A "step" is one complete solve to 1e-10 from x = 0 with the hierarchy
resident in HBM.  The headline value is the mixed-precision form (FP32
V-cycles inside FP64 refinement); the plain FP64 form is timed beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-1e-10 defect reduction, mixed vs FP64 GPU; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0


def cpu_info():
    model, cores = "unknown", os.cpu_count()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, cores


def load_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("traffic_bytes")
    except (OSError, ValueError):
        return None


def serving_run(n_req, rate, seed, budget_gb, slots):
    """Solve-serving stream: 2D Q1 Poisson and 3D Q1 jump-coefficient
    (kappa = 10^U[-2,2]) solves with unknowns log-uniform in [1e3, 1e6] (the
    largest uniformly refined level not above the drawn size), relative
    defect 1e-8, FP64 MC-GS V(2,2).  Two streams on one server: Poisson
    arrivals at `rate` (latency), then a burst of as many requests submitted
    at once (arrivals above capacity: saturated throughput), whose shapes
    find their groups resident."""
    import numpy as np
    from ug_amd import problems, ugmg

    rng = np.random.default_rng(seed)

    def draw(k0):
        reqs = []
        for i in range(n_req):
            dim = 2 if rng.random() < 0.5 else 3
            target = 10.0 ** rng.uniform(3.0, 6.0)
            cells = 2 if dim == 2 else 1
            size = lambda r: (cells * 2 ** r + 1) ** dim
            nref = 1
            while size(nref + 1) <= target:
                nref += 1
            while size(nref) < 1000:
                nref += 1
            kappa = problems.KAPPA_ONE if dim == 2 else problems.KAPPA_JUMP
            P = problems.q1(dim, cells, nref, kappa, seed=int(seed * 1000 + k0 + i))
            # base level: the largest level with at most 300 vectors (3D:
            # 5^3, 2D: 17^2).  The 9^3 level (BFS bandwidth 259) would cost
            # ~40 ms of host band LU + inverse per request against ~0.15 ms
            # per cycle for smoothing it instead.
            bl = max(l for l in range(P.top + 1) if P.levels[l].n <= 300)
            prm = ugmg.default_params(nu1=2, nu2=2, gamma=1, baselevel=bl, smoother=ugmg.SMOOTH_GS, damp=1.0,
                                      base_abslimit=1e-300, use_graph=1)
            reqs.append((P, prm))
        return reqs

    def run(srv, reqs, arrivals):
        ids = []
        t0 = time.perf_counter()
        k = 0
        done0 = srv.step()[2]
        while True:
            now = time.perf_counter() - t0
            while k < len(reqs) and arrivals[k] <= now:
                P, prm = reqs[k]
                ids.append(srv.submit(P, prm, red=1e-8, maxiter=500))
                k += 1
            a, q, d = srv.step()
            if k == len(reqs) and d - done0 >= len(reqs):
                break
            if a == 0 and k < len(reqs):
                time.sleep(min(0.001, max(0.0, arrivals[k] - now)))
        lat, gpu, ok, rows = [], [], 0, []
        t_first, t_last = None, 0.0
        for i, rid in enumerate(ids):
            st, res, (ts, ta, td) = srv.status(rid)
            if st == ugmg.REQ_DONE and res.converged:
                ok += 1
            lat.append(td - ts)
            gpu.append(td - ta)
            rows.append((td - ts, ta - ts, td - ta, reqs[i][0].levels[-1].n, res.number_of_linear_iterations))
            t_first = ts if t_first is None else min(t_first, ts)
            t_last = max(t_last, td)
        slow = max(rows)
        return {
            "requests": len(reqs), "converged": ok,
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
            "admitted_to_done_p50_ms": round(float(np.percentile(gpu, 50)) * 1e3, 2),
            "slowest": {"unknowns": int(slow[3]), "submit_to_admit_ms": round(slow[1] * 1e3, 2),
                        "admit_to_done_ms": round(slow[2] * 1e3, 2), "iterations": int(slow[4])},
            "solves_per_s": round(len(reqs) / (t_last - t_first), 2),
        }

    srv = ugmg.Server(0, budget_bytes=budget_gb * 1e9, max_slots=slots)
    # cold start: a burst on an empty server, every shape's first request
    # waits for its group's host layout and device set-up
    cold = draw(0)
    cold_run = run(srv, cold, np.zeros(n_req))
    g0, l0, c0 = srv.stats()
    # Poisson arrivals with the groups resident
    reqs = draw(n_req)
    gaps = rng.exponential(1.0 / rate, n_req)
    lat_run = run(srv, reqs, np.cumsum(gaps) - gaps[0])
    g1, l1, c1 = srv.stats()
    # arrivals above capacity: a burst
    burst = draw(2 * n_req)
    cap_run = run(srv, burst, np.zeros(n_req))
    g2, l2, c2 = srv.stats()
    srv.close()
    sizes = [r[0].levels[-1].n for r in cold + reqs + burst]
    cold_run.update({"arrival": f"all {n_req} submitted at once to an empty server",
                     "resident_groups": g0, "solves_per_launch": round(c0 / max(l0, 1), 2)})
    lat_run.update({"arrival_rate_per_s": rate, "solves_per_launch": round((c1 - c0) / max(l1 - l0, 1), 2)})
    cap_run.update({"arrival": f"all {n_req} submitted at once (above capacity), groups resident",
                    "solves_per_launch": round((c2 - c1) / max(l2 - l1, 1), 2),
                    "resident_groups": g2})
    return {"slots": slots, "hbm_budget_gb": budget_gb,
            "unknowns_min_median_max": [int(min(sizes)), int(np.median(sizes)), int(max(sizes))],
            "poisson_arrivals": lat_run, "saturated": cap_run, "cold_start": cold_run,
            "p50_ms": lat_run["p50_ms"], "p99_ms": lat_run["p99_ms"],
            "capacity_solves_per_s": cap_run["solves_per_s"],
            "note": "latency = completion - submission; each group iteration is one batched launch sequence over "
                    "the group's solves (solves_per_launch); cold_start includes each shape's first host layout "
                    "build and device set-up"}


def assemble_run(problems, ugmg, torch, nref, P, base):
    """Device-side set-up of the finest level from its element list (hexes
    with kappa_e times the reference Q1 matrix, in device memory):
    ugmg_assemble_level alone (flatten to CSR in UG's row layout, with the
    copy of the level to the host), whether it equals the host generator's
    level, and the set-up of the whole hierarchy after it: the assembled
    level's CSR stays on the device for the device layout build (coarse
    levels from the host generator), against the same hierarchy laid out on
    the host."""
    import numpy as np
    top = P.levels[-1]
    el = problems.q1_elements(3, 1, nref, problems.KAPPA_JUMP, seed=0x5EED0002)
    elem = torch.from_numpy(el.elem).cuda()
    kref = torch.from_numpy(el.kref).cuda()
    scale = torch.from_numpy(el.scale).cuda()
    skip = torch.from_numpy(el.skip.view(np.int32)).cuda()
    ctx = ugmg.Context(0)
    ctx.set_nlevels(1, 1)
    torch.cuda.synchronize()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        ctx.assemble_level(0, el.n, elem, kref=kref, scale=scale, skip=skip)
        ctx.sync()
        times.append(time.perf_counter() - t0)
    rp, ci, vals, _ = ctx.get_level(0)
    same = bool(np.array_equal(rp, top.row_ptr) and np.array_equal(ci, top.col_idx)
                and np.array_equal(vals.view(np.int64), top.vals.view(np.int64)))
    ctx.close()
    out = {"elements": int(el.nelem), "nnz": int(rp[-1]), "assemble_s": round(min(times), 3),
           "equals_host_level": same}
    # the hierarchy's set-up: device assembly + device layout, device
    # layout of the uploaded levels, host layout (each includes the base
    # level's band LU and inverse)
    for key, where, assemble in (("assemble_and_device_layout_s", ugmg.SETUP_DEVICE, True),
                                 ("device_layout_s", ugmg.SETUP_DEVICE, False),
                                 ("host_layout_s", ugmg.SETUP_HOST, False)):
        ctx = ugmg.Context(0)
        ctx.set_setup(where)
        ctx.set_hierarchy(P)
        ctx.set_params(**base, precision=ugmg.FP64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if assemble:
            ctx.assemble_level(P.top, el.n, elem, kref=kref, scale=scale, skip=skip)
        ctx.preprocess()
        ctx.sync()
        out[key] = round(time.perf_counter() - t0, 3)
        ctx.close()
    out["note"] = ("element list on the GPU -> CSR in UG row layout (incl. its copy to the host); "
                   "set-up times cover levels 4..7 incl. the 17^3 base level's band LU and inverse")
    return out


def bcgs_run(nref, nrhs, steps, warmup, cpu_sample):
    """BiCGStab preconditioned by one V(1,1) multicolour-GS cycle on the SUPG
    P1 convection-diffusion hierarchy (eps 1e-3, wind angle from the seed),
    nrhs right-hand sides ~U[-1,1] solved together to 1e-8: FP32 cycle with
    FP64 BiCGStab vectors, dots and norms, beside the all-FP64 form and a
    single right-hand side."""
    import numpy as np
    import torch
    from ug_amd import problems, ugmg

    P = problems.p1_cd(nref, 1e-3, float("nan"), seed=0x5EED0003)
    top, n = P.top, P.levels[-1].n
    R = problems.splitmix_uniform(0x5EED0013, nrhs * n, -1.0, 1.0).reshape(nrhs, n)
    R[:, P.levels[-1].skip == 1] = 0.0
    bl = max(l for l, lv in enumerate(P.levels) if lv.n <= 729)
    prm = dict(nu1=1, nu2=1, gamma=1, baselevel=bl, smoother=ugmg.SMOOTH_GS, damp=1.0,
               base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, use_graph=1)
    Rt = torch.from_numpy(R).cuda()
    X = torch.zeros_like(Rt)
    B = torch.empty_like(Rt)
    out = {"workload": f"p1cd_{2 ** (nref + 1) + 1}^2: -1e-3 Lap u + b.grad u = f, SUPG P1, "
                       f"theta={P.theta:.4f}, BiCGStab + V(1,1) MC-GS, base level {bl}, {nrhs} rhs, tol 1e-8",
           "unknowns": n, "nnz_fine": int(P.levels[-1].nnzb), "levels": top + 1, "nrhs": nrhs}

    def run(ctx, k, cols):
        lat, res = [], None
        for _ in range(k):
            X[:cols].zero_()
            B[:cols].copy_(Rt[:cols])
            torch.cuda.synchronize()
            t = time.perf_counter()
            res, _ = ctx.bcgs(top, X[:cols], B[:cols], red=1e-8, maxiter=200, maxdef=1)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t)
        return float(np.mean(lat)), res

    for name, prec in (("mixed", ugmg.MIXED), ("fp64", ugmg.FP64)):
        ctx = ugmg.Context(0)
        ctx.set_hierarchy(P)
        ctx.set_params(**prm, precision=prec, inner_cycles=1)
        ctx.set_batch(nrhs)
        ctx.preprocess()
        run(ctx, warmup, nrhs)
        t_all, res = run(ctx, steps, nrhs)
        # true defect of every column, FP64 on the GPU
        Bt = Rt.clone()
        torch.cuda.synchronize()
        for j in range(nrhs):
            ctx.dmatmul_minus(top, Bt[j], X[j])
        ctx.sync()
        rel = (torch.linalg.norm(Bt, dim=1) / torch.linalg.norm(Rt, dim=1)).max().item()
        run(ctx, warmup, 1)
        t_one, _ = run(ctx, steps, 1)
        its = [r.number_of_linear_iterations for r in res]
        out[name] = {"ms": round(t_all * 1e3, 3), "ms_single_rhs": round(t_one * 1e3, 3),
                     "batch_gain_per_rhs": round(t_one * nrhs / t_all, 2),
                     "half_steps_max": max(its), "half_steps_mean": round(float(np.mean(its)), 1),
                     "all_converged": all(r.converged for r in res), "max_true_rel_defect": rel}
        ctx.close()
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["ms"] / out["mixed"]["ms"], 3)
    if cpu_sample > 0:
        from oracle import oracle as O
        H = O.Hierarchy(P.levels)
        p = O.params(smoother=O.MCGS, damp=1.0, baselevel=bl, nu1=1, nu2=1)
        t = time.perf_counter()
        its = []
        for j in range(cpu_sample):
            r = O.bcgs_solve(H, top, p, R[j], red=1e-8, maxiter=200)
            its.append(int(r["nit"]))
        per = (time.perf_counter() - t) / cpu_sample
        out["cpu_baseline"] = {
            "value": round(per * nrhs * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
            "per_rhs_ms": round(per * 1e3, 1), "half_steps": its,
            "sample": f"{cpu_sample} of the {nrhs} right-hand sides, each solved to 1e-8 by the oracle "
                      f"(the CPU restatement of UG's BCGSSolver + lmgc, FP64, 1 thread; not compiled UG), "
                      f"one at a time as UG solves them; the value is {nrhs} x the measured time per rhs"}
    return out


def elasticity_run(nref, steps, warmup, inner, cpu_cycles):
    """3D P1 linear elasticity (3x3 blocks) on the Kuhn-split unit cube,
    clamped at x = 0: ls + lmgc V(2,2) multicolour block GS to 1e-10 per
    component, FP64 and FP32 cycle in FP64 refinement."""
    import numpy as np
    import torch
    from ug_amd import problems, ugmg

    t0 = time.time()
    P = problems.elasticity(nref, E=1.0, nu=0.3, seed=0x5EED0004)
    gen_s = time.time() - t0
    top, n = P.top, P.levels[-1].n
    bl = max(l for l, lv in enumerate(P.levels) if 3 * lv.n <= 8192)
    prm = dict(nu1=2, nu2=2, gamma=1, baselevel=bl, smoother=ugmg.SMOOTH_GS, damp=1.0,
               base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, use_graph=1)
    rhs = torch.from_numpy(P.rhs).cuda()
    x = torch.zeros_like(rhs)
    b = torch.empty_like(rhs)
    out = {"workload": f"elasticity_p1_{2 ** nref + 1}^3: P1 tets (Kuhn split), E=1 nu=0.3, clamped x=0, "
                       f"3x3 blocks, V(2,2) multicolour block GS, base level {bl}, tol 1e-10 per component",
           "nodes": n, "unknowns": 3 * n, "nnz_blocks_fine": int(P.levels[-1].nnzb), "levels": top + 1,
           "generate_s": round(gen_s, 2)}
    A64 = None
    for name, prec in (("fp64", ugmg.FP64), ("mixed", ugmg.MIXED)):
        ctx = ugmg.Context(0)
        ctx.set_hierarchy(P)
        ctx.set_params(**prm, precision=prec, inner_cycles=inner)
        ctx.preprocess()
        lat, res = [], None
        for k in range(warmup + steps):
            x.zero_()
            b.copy_(rhs)
            torch.cuda.synchronize()
            t = time.perf_counter()
            res, _ = ctx.solve(top, x, b, red=1e-10, abslimit=1e-300, maxiter=500, record=False)
            torch.cuda.synchronize()
            if k >= warmup:
                lat.append(time.perf_counter() - t)
        bb = rhs.clone()
        torch.cuda.synchronize()
        ctx.dmatmul_minus(top, bb, x)
        ctx.sync()
        tr = (bb.view(-1, 3).norm(dim=0) / rhs.view(-1, 3).norm(dim=0)).cpu().numpy()
        out[name] = {"ms": round(float(np.mean(lat)) * 1e3, 3), "iterations": res.number_of_linear_iterations,
                     "converged": bool(res.converged), "true_rel_defect_per_component": [float(v) for v in tr]}
        if name == "fp64":
            A64 = ctx
        else:
            ctx.close()
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["ms"] / out["mixed"]["ms"], 3)
    A64.close()
    if cpu_cycles > 0:
        from oracle import oracle as O
        H = O.Hierarchy(P.levels)
        p = O.params(smoother=O.GS, damp=1.0, baselevel=bl)
        t = time.perf_counter()
        want = O.ls_solve(H, top, p, P.rhs, red=1e-10, maxiter=500)
        sec = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(sec * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
                               "iterations": int(want["nit"]), "converged": bool(want["converged"]),
                               "sample": "the whole solve to 1e-10 per component by the oracle (the CPU "
                                         "restatement of UG's ls + lmgc with lexicographic block GS, FP64, "
                                         "1 thread; not compiled UG) on the same hierarchy"}
    return out


def elasticity_local_run(nuni, depth, steps, warmup, inner, cpu):
    """3D P1 linear elasticity (E = 1, nu = 0.3, clamped
    on x = 0, 3x3 blocks) on the Kuhn-split unit cube refined 5 times
    uniformly, then 3 times toward the corner (0,0,0) with UG's green
    closure of tetrahedra; ls + lmgc V(2,2) block Jacobi (0.8) ON_SURFACE,
    base level 3, nodal body force ~U[-1,1]^3 on the surface vectors, stop at
    1e-8 per component: FP32 cycles in FP64 refinement (the headline form of
    the config) beside the plain FP64 solve."""
    import numpy as np
    import torch
    from ug_amd import problems, ugmg

    t0 = time.time()
    P = problems.elasticity_local(nuni, depth)
    gen_s = time.time() - t0
    bl, top = 3, P.top
    frl = problems.fullrefinelevel(P)
    fs = problems.surface_load(P, bl, 0x5EED0004)
    nsurf = sum(int(problems.surface_mask(P, l).sum()) for l in range(frl, top + 1))
    out = {"workload": f"elasticity_local_{2 ** nuni + 1}^3+{depth}: P1 tets (Kuhn split), {nuni} uniform + "
                       f"{depth} local levels toward (0,0,0) (UG green closure), E=1 nu=0.3, clamped x=0, "
                       f"3x3 blocks, ls + lmgc V(2,2) block Jacobi 0.8 ON_SURFACE from level {frl}, base level "
                       f"{bl}, tol 1e-8 per component",
           "nodes_per_level": [int(lv.n) for lv in P.levels], "surface_nodes": nsurf,
           "surface_unknowns": 3 * nsurf, "fullrefinelevel": frl, "levels": top + 1,
           "generate_s": round(gen_s, 2)}
    fsd = [None if f is None else torch.from_numpy(f).cuda() for f in fs]
    xs = [None if f is None else torch.zeros_like(f) for f in fsd]
    bs = [None if f is None else torch.empty_like(f) for f in fsd]

    def apply(lv, u):
        rows = np.repeat(np.arange(lv.n), np.diff(lv.row_ptr))
        r = np.zeros((lv.n, 3))
        np.add.at(r, rows, np.einsum("kij,kj->ki", lv.vals.reshape(-1, 3, 3), u.reshape(-1, 3)[lv.col_idx]))
        return r.ravel()

    def surface_defect(xh):
        acc = np.zeros(3)
        for l in range(frl, top + 1):
            m = problems.surface_mask(P, l)
            x = np.zeros(P.levels[l].n * 3) if xh is None else xh[l]
            acc += ((fs[l] - apply(P.levels[l], x)).reshape(-1, 3)[m] ** 2).sum(axis=0)
        return np.sqrt(acc)

    first = surface_defect(None)
    sols = {}
    for name, prec in (("fp64", ugmg.FP64), ("mixed", ugmg.MIXED)):
        ctx = ugmg.Context(0)
        ctx.set_hierarchy(P)
        ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=bl, smoother=ugmg.SMOOTH_JAC, damp=0.8,
                       base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, use_graph=1, precision=prec,
                       inner_cycles=inner)
        ctx.preprocess()
        lat, res = [], None
        for k in range(warmup + steps):
            for l in range(bl, top + 1):
                xs[l].zero_()
                bs[l].copy_(fsd[l])
            torch.cuda.synchronize()
            t = time.perf_counter()
            res, _ = ctx.solve_surface(top, xs, bs, red=1e-8, maxiter=300, record=False)
            torch.cuda.synchronize()
            if k >= warmup:
                lat.append(time.perf_counter() - t)
        cyc, by = ctx.solve_stats()
        ctx.close()
        xh = [None if t is None else t.cpu().numpy() for t in xs]
        sols[name] = xh
        tr = surface_defect(xh) / first
        ms = float(np.mean(lat)) * 1e3
        out[name] = {"ms": round(ms, 3), "iterations": res.number_of_linear_iterations, "cycles": int(cyc),
                     "converged": bool(res.converged),
                     "true_rel_surface_defect_per_component": [float(v) for v in tr],
                     "solve_roofline": {"bytes_per_solve": by, "achieved": round(by / ms / 1e6, 1),
                                        "frac": round(by / ms / 1e6 / HBM_PEAK_GBS, 4)}}
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["ms"] / out["mixed"]["ms"], 3)
    if cpu:
        from oracle import oracle as O
        H = O.Hierarchy(P.levels)
        t = time.perf_counter()
        r = O.ls_solve_surface(H, top, O.params(baselevel=bl, smoother=O.JAC, damp=0.8), fs, red=1e-8,
                               maxiter=300)
        sec = time.perf_counter() - t
        err = {k: max(float(np.max(np.abs(v[l] - r["x"][l])) / np.max(np.abs(r["x"][l])))
                      for l in range(bl, top + 1)) for k, v in sols.items()}
        out["error_vs_oracle_fp64_solve"] = err
        out["cpu_baseline"] = {"value": round(sec * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
                               "iterations": int(r["nit"]), "converged": bool(r["converged"]),
                               "sample": "the whole surface solve to 1e-8 per component by the oracle (the CPU "
                                         "restatement of UG's ls + lmgc ON_SURFACE with block Jacobi, FP64, "
                                         "1 thread; not compiled UG) on the same hierarchy"}
    return out


def hbm_probe(torch, reps=20):
    """Achievable HBM bandwidth on this GPU: a device copy of 1 GiB (reads
    and writes counted) and a sum over 1 GiB (reads), best of `reps`, timed
    with events; reported beside the 8 TB/s peak the rooflines divide by."""
    m = 1 << 27
    src = torch.ones(m, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    out = {}
    for name, fn, by in (("copy", lambda: dst.copy_(src), 16.0 * m), ("read", lambda: src.sum(), 8.0 * m)):
        best = None
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            b.synchronize()
            t = a.elapsed_time(b) * 1e-3
            best = t if best is None else min(best, t)
        out[name + "_gbs"] = round(by / best / 1e9, 1)
    del src, dst
    torch.cuda.empty_cache()
    return out


def roofline_entry(kernel, ms, by, traffic=None, **extra):
    ach = by / (ms * 1e-3) / 1e9
    e = {"kernel": kernel, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "algorithmic_bytes": by,
         "avg_launch_us": round(ms * 1e3, 2)}
    e.update(extra)
    return e


def poisson2d_run(steps, warmup, cpu):
    """2D Q1 Poisson on the unit square (2x2 quads refined
    6 times, 129^2 nodes), ls + lmgc V(2,2) damped Jacobi (0.8), exact coarse
    solve on level 0, stop at 1e-10: FP64 and mixed time-to-tolerance on the
    GPU, and the oracle's whole CPU solve."""
    import numpy as np
    import torch
    from ug_amd import problems, ugmg

    P = problems.poisson2d_129()
    top, n = P.top, P.levels[-1].n
    prm = dict(nu1=2, nu2=2, gamma=1, baselevel=0, smoother=ugmg.SMOOTH_JAC, damp=0.8,
               base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, use_graph=1)
    rhs = torch.from_numpy(P.rhs).cuda()
    x = torch.zeros_like(rhs)
    b = torch.empty_like(rhs)
    out = {"workload": "poisson2d_q1_129^2: -Lap u = f, unit square, Q1 (2x2 quads, 6 refinements), "
                       "V(2,2) damped Jacobi 0.8, exact coarse solve on level 0, tol 1e-10",
           "unknowns": n, "levels": top + 1}
    for name, prec in (("fp64", ugmg.FP64), ("mixed", ugmg.MIXED)):
        ctx = ugmg.Context(0)
        ctx.set_hierarchy(P)
        ctx.set_params(**prm, precision=prec, inner_cycles=7)
        ctx.preprocess()
        lat, res = [], None
        for k in range(warmup + steps):
            x.zero_()
            b.copy_(rhs)
            torch.cuda.synchronize()
            t = time.perf_counter()
            res, _ = ctx.solve(top, x, b, red=1e-10, abslimit=1e-300, maxiter=200, record=False)
            torch.cuda.synchronize()
            if k >= warmup:
                lat.append(time.perf_counter() - t)
        cyc, by = ctx.solve_stats()
        bb = rhs.clone()
        torch.cuda.synchronize()
        ctx.dmatmul_minus(top, bb, x)
        ctx.sync()
        out[name] = {"ms": round(float(np.mean(lat)) * 1e3, 3), "iterations": res.number_of_linear_iterations,
                     "cycles": int(cyc), "converged": bool(res.converged),
                     "true_rel_defect": float(torch.linalg.norm(bb) / torch.linalg.norm(rhs))}
        ctx.close()
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["ms"] / out["mixed"]["ms"], 3)
    if cpu:
        from oracle import oracle as O
        H = O.Hierarchy(P.levels)
        t = time.perf_counter()
        want = O.ls_solve(H, top, O.params(smoother=O.JAC, damp=0.8, baselevel=0), P.rhs, red=1e-10, maxiter=200)
        sec = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(sec * 1e3, 2), "unit": "ms", "cores": 1, "kind": "port",
                               "iterations": int(want["nit"]),
                               "sample": "the whole solve to 1e-10 by the oracle (the CPU restatement of UG's "
                                         "ls + lmgc + jac + ex, FP64, 1 thread; not compiled UG)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nref", type=int, default=7, help="refinements of the unit hex (7: 129^3)")
    ap.add_argument("--inner", type=int, default=7,
                    help="FP32 V-cycles per FP64 refinement step (at most: a step runs fewer when the "
                         "rate predicts the tolerance met)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=10.0, help="Poisson arrival rate (1/s)")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--bcgs-nref", type=int, default=8, help="P1 refinements of the BiCGStab leg (8: 513^2)")
    ap.add_argument("--bcgs-rhs", type=int, default=32)
    ap.add_argument("--no-bcgs", action="store_true")
    ap.add_argument("--elasticity-nref", type=int, default=6, help="Kuhn-cube refinements (6: 65^3 nodes)")
    ap.add_argument("--no-elasticity", action="store_true")
    ap.add_argument("--no-poisson2d", action="store_true")
    ap.add_argument("--kernels-only", nargs="?", const="spmv", choices=["spmv", "gs"],
                    help="only the finest level's kernels (for PMC runs): 'spmv' (10 FP64 SpMV "
                         "launches) or 'gs' (10 FP64 MC-GS steps: 8 colour sweeps + defect each), "
                         "then a copy kernel of known size")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU; --gpus {args.gpus} is not supported "
              "(levels would be partitioned by coordinate bisection with halo exchanges; "
              "that sharded form is designed, not built)", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from ug_amd import problems, ugmg

    t0 = time.time()
    P = problems.jump3d(args.nref)
    gen_s = time.time() - t0
    top = P.top
    n = P.levels[-1].n
    nnz = int(P.levels[-1].nnzb)
    base = dict(nu1=2, nu2=2, gamma=1, baselevel=4, smoother=ugmg.SMOOTH_GS, damp=1.0,
                base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, use_graph=1)

    ctx64 = ugmg.Context(0)
    ctx64.set_hierarchy(P)
    ctx64.set_params(**base, precision=ugmg.FP64)
    t0 = time.time()
    ctx64.preprocess()
    prep_s = time.time() - t0
    if args.kernels_only:
        # 10 FP64 SpMV launches (or MC-GS steps) on the finest level, then a
        # copy kernel of known size (torch elementwise: reads 8*m bytes,
        # writes 8*m bytes)
        kind = ugmg.K_SPMV64 if args.kernels_only == "spmv" else ugmg.K_GS64
        ms, by = ctx64.time_kernel(top, kind, 10)
        m = 1 << 26
        src = torch.ones(m, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({f"{args.kernels_only}_fp64_us": ms * 1e3,
                          f"{args.kernels_only}_fp64_algorithmic_bytes": by,
                          "copy_kernel_bytes_read": 8 * m, "copy_kernel_bytes_written": 8 * m}))
        ctx64.close()
        return 0
    print("bench: 129^3 set up", file=sys.stderr, flush=True)
    asm = assemble_run(problems, ugmg, torch, args.nref, P, base)
    ctxmx = ugmg.Context(0)
    ctxmx.set_hierarchy(P)
    ctxmx.set_params(**base, precision=ugmg.MIXED, inner_cycles=args.inner)
    ctxmx.preprocess()

    rhs = torch.from_numpy(P.rhs).cuda()
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(rhs)
    torch.cuda.synchronize()

    def solve(ctx):
        x.zero_()
        b.copy_(rhs)
        torch.cuda.synchronize()
        res, d = ctx.solve(top, x, b, red=1e-10, abslimit=1e-300, maxiter=500, record=True)
        return res, d

    def timed(ctx, k):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = None
        lat = []
        for _ in range(k):
            s = time.perf_counter()
            out = solve(ctx)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - s)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / k, out, lat

    for _ in range(args.warmup):
        solve(ctx64)
        solve(ctxmx)
    t64, (r64, d64), _ = timed(ctx64, args.steps)
    cyc64, by64 = ctx64.solve_stats()
    x64 = x.detach().clone()
    tmx, (rmx, dmx), lat = timed(ctxmx, args.steps)
    cycmx, bymx = ctxmx.solve_stats()
    xmx = x.detach().clone()
    print("bench: 129^3 solves timed", file=sys.stderr, flush=True)

    # equal final accuracy: true defect |f - A x| / |f| of both solutions in FP64
    def true_rel(xs):
        bb = rhs.clone()
        torch.cuda.synchronize()
        ctx64.dmatmul_minus(top, bb, xs)
        ctx64.sync()
        return float(torch.linalg.norm(bb) / torch.linalg.norm(rhs))

    acc64, accmx = true_rel(x64), true_rel(xmx)
    diff = float(torch.linalg.norm(xmx - x64) / torch.linalg.norm(x64))

    # rooflines on the finest level, bytes as the data are laid out
    # (ugmg_time_kernel), durations from hipEvents on the context's stream
    ms_gs, by_gs = ctx64.time_kernel(top, ugmg.K_GS64, 10)
    ms_gsf, by_gsf = ctx64.time_kernel(top, ugmg.K_GSF64, 10)
    ms_spmv, by_spmv = ctx64.time_kernel(top, ugmg.K_SPMV64, 20)
    ms_spmv32, by_spmv32 = ctxmx.time_kernel(top, ugmg.K_SPMV32, 20)
    ms_gs32, by_gs32 = ctxmx.time_kernel(top, ugmg.K_GS32, 10)
    # the damped-Jacobi step (t = w D^-1 b, b -= A t, c += t) on the same
    # level, from a Jacobi hierarchy (list order, no colouring)
    ctxj = ugmg.Context(0)
    ctxj.set_hierarchy(P)
    ctxj.set_params(**dict(base, smoother=ugmg.SMOOTH_JAC, damp=0.8), precision=ugmg.FP64)
    ctxj.preprocess()
    ms_jac, by_jac = ctxj.time_kernel(top, ugmg.K_JAC64, 20)
    ctxj.close()
    # SURVEY-formula bytes of the FP64 SpMV (int32 column index per entry):
    # what a CSR layout would move
    survey_spmv = nnz * 12.0 + 4.0 * (n + 1) + 24.0 * n

    roof = roofline_entry("multicolour GS step FP64 on the 129^3 level: 8 colour sweeps over the L slab "
                          "+ the defect pass over the U slab (b -= A(wv), c += wv)", ms_gs, by_gs,
                          load_traffic("gs_step_fp64_jump3d_129_traffic.json"))
    kernels = {
        "gs_step_fp64_fused": roofline_entry("the GS step as a cycle runs its steps after the first: "
                                             "8 colour sweeps with the previous defect folded in",
                                             ms_gsf, by_gsf),
        "gs_step_fp32": roofline_entry("the FP32 GS step of the mixed form's cycles", ms_gs32, by_gs32),
        "spmv_fp64": roofline_entry("k_spmv<double> (dmatmul_minus, b -= A x)", ms_spmv, by_spmv,
                                    load_traffic("spmv_fp64_jump3d_129_traffic.json"),
                                    survey_formula_bytes=survey_spmv),
        "spmv_fp32": roofline_entry("k_spmv<float> (the FP32 copy of A)", ms_spmv32, by_spmv32),
        "jacobi_step_fp64": roofline_entry("damped-Jacobi step (k_jac + k_spmv mode 2) from a Jacobi "
                                           "hierarchy", ms_jac, by_jac),
    }
    solve_roof = {
        "fp64": {"bytes_per_solve": by64, "cycles": int(cyc64),
                 "achieved": round(by64 / t64 / 1e9, 1), "frac": round(by64 / t64 / 1e9 / HBM_PEAK_GBS, 4)},
        "mixed": {"bytes_per_solve": bymx, "fp32_cycles": int(cycmx),
                  "achieved": round(bymx / tmx / 1e9, 1), "frac": round(bymx / tmx / 1e9 / HBM_PEAK_GBS, 4)},
        "note": "algorithmic bytes of every kernel a solve launches (ugmg_solve_stats, the layout-true "
                "model of ugmg_time_kernel) / ms_per_step",
    }
    fracs = [roof["frac"]] + [k["frac"] for k in kernels.values()] + [solve_roof["fp64"]["frac"],
                                                                     solve_roof["mixed"]["frac"]]

    out = {
        "metric": METRIC,
        "value": round(tmx * 1e3, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tmx * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle in f64 refinement",
        "data": "synthetic (seeded SplitMix64 kappa and f)",
        "config": {
            "workload": f"jump3d_q1_{2 ** args.nref + 1}^3: -div(kappa grad u)=f, unit cube, Q1 hex, "
                        f"kappa=10^U[-2,2] per element, V(2,2) multicolour GS, base level 4 (17^3, dense-inverse ex "
                        f"in FP64), tol 1e-10",
            "unknowns": n,
            "nnz_fine": nnz,
            "levels": top + 1,
            "inner_cycles": args.inner,
        },
        "fp64_ms": round(t64 * 1e3, 3),
        "speedup_mixed_vs_fp64": round(t64 / tmx, 3),
        "fp64_iterations": r64.number_of_linear_iterations,
        "mixed_outer_iterations": rmx.number_of_linear_iterations,
        "mixed_fp32_cycles": int(cycmx),
        "final_true_rel_defect": {"fp64": acc64, "mixed": accmx},
        "mixed_vs_fp64_solution_rel_diff": diff,
        "mixed_latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                             "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "roofline": roof,
        "roofline_kernels": kernels,
        "solve_roofline": solve_roof,
        "roofline_fracs_le_1": bool(all(f <= 1.0 for f in fracs)),
        "hbm_probe": hbm_probe(torch),
        "setup_s": {"generate": round(gen_s, 2), "preprocess": round(prep_s, 2), "device_assembly": asm},
    }
    ctx64.close()
    ctxmx.close()

    if not args.no_poisson2d:
        out["poisson2d"] = poisson2d_run(args.steps, args.warmup, not args.no_cpu)
        print("bench: poisson2d leg done", file=sys.stderr, flush=True)

    if not args.no_bcgs:
        out["bcgs_cd2d"] = bcgs_run(args.bcgs_nref, args.bcgs_rhs, args.steps, args.warmup,
                                    0 if args.no_cpu else 2)
        print("bench: bcgs leg done", file=sys.stderr, flush=True)

    if not args.no_elasticity:
        out["elasticity_3d"] = elasticity_local_run(5, 3, args.steps, args.warmup, args.inner, not args.no_cpu)
        print("bench: locally refined elasticity leg done", file=sys.stderr, flush=True)
        out["elasticity_3d_uniform"] = elasticity_run(args.elasticity_nref, args.steps, args.warmup, args.inner,
                                                      0 if args.no_cpu else 1)
        print("bench: uniform elasticity leg done", file=sys.stderr, flush=True)

    if not args.no_serve:
        out["serving"] = serving_run(args.serve_requests, args.serve_rate, 0x5EED0005, 32.0, args.serve_slots)
        print("bench: serving leg done", file=sys.stderr, flush=True)

    if not args.no_cpu:
        from oracle import oracle as O
        model, cores = cpu_info()
        H = O.Hierarchy(P.levels)
        p = O.params(smoother=O.GS, damp=1.0, baselevel=4)
        t = time.perf_counter()
        want = O.ls_solve(H, top, p, P.rhs, red=1e-10, maxiter=500)
        sec = time.perf_counter() - t
        out["cpu_baseline"] = {
            "value": round(sec * 1e3, 1),
            "unit": "ms",
            "cores": 1,
            "kind": "port",
            "iterations": int(want["nit"]),
            "converged": bool(want["converged"]),
            "sample": f"the whole solve to 1e-10 by the oracle (the CPU restatement of UG's ls + lmgc V(2,2) "
                      f"with lexicographic GS and the ex base solve, FP64; not compiled UG, which this image "
                      f"cannot build) on the same {n}-unknown hierarchy, its own cycle count, 1 thread of "
                      f"{cores} ({model})",
        }
    out["host"] = platform.node()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
