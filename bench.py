#!/usr/bin/env python3
"""Benchmark: RK cell-updates/s of the Navier-Stokes Taylor-Green vortex on a periodic 128^3
hex mesh (2,097,152 cells), FP64 GPU path vs mixed FP32/FP64 GPU path, with the mixed path's
error against FP64, the stage kernel's HBM roofline fraction, and the reference CPU path
timed on this host's cores (first, before the GPU is touched).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--cpu-cores 16]
    python bench.py --workload tet|channel|batch   (the other configurations)

One step = one 3-stage LSERK step (src/simulation/advance.c:47-117) of every cell; the state
stays resident in HBM for the whole timed region (no host transfers inside it).  Each path
first lets the autotuner pick its launch shape; per-kernel durations come from a separate
run with HIP events around every launch.
"""
import argparse
import ctypes as C
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import teal_amd  # noqa: E402
from teal_amd import _abi  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
GAMMA, MACH, RE, PR, CFL = 1.4, 0.1, 1600.0, 0.71, 0.5


def tgv_primitive(center):
    """Taylor-Green vortex, rho = 1, u = sin x cos y cos z, v = -cos x sin y cos z, w = 0,
    p = 1/(gamma M^2) + (cos 2x + cos 2y)(cos 2z + 2)/16."""
    x, y, z = center[:, 0], center[:, 1], center[:, 2]
    P = np.zeros((len(center), 5))
    P[:, 0] = 1.0
    P[:, 1] = np.sin(x) * np.cos(y) * np.cos(z)
    P[:, 2] = -np.cos(x) * np.sin(y) * np.cos(z)
    P[:, 4] = 1.0 / (GAMMA * MACH * MACH) + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2) / 16
    return P


def byte_model(ni, nc, nf, mixed, viscous=True, vk=False):
    """Algorithmic HBM bytes: every operand array read once, every result written once.
    s = state/geometry bytes (8 FP64, 4 mixed), u = 8 (FP64 U0/U), i = 4 (index).
    `per_cell_update` is the path-level model of one RK step (3 residual passes + the
    per-step dt/U0 work) per cell; the per-kernel figures are per launch of this library's
    kernels: gradient+limiter; the stage kernel (face fluxes in LDS, accumulation, RK update);
    and, for the split path, face fluxes (writes F) and accumulation (reads F)."""
    s, u, i = (4 if mixed else 8), 8, 4
    v = 1 if viscous else 0
    k = 1 if vk else 0
    m = 1 if mixed else 0
    nn = 2 * nf - (nc - ni)  # incidences (CSR neighbour entries) of inner cells
    nfv = 10 if viscous else 5  # values per face flux record
    b_grad = 5 * s * nc + (2 * i + 6 * s) * nf + 15 * s * ni
    b_lim = 5 * s * nc + 15 * s * ni + i * (ni + 1 + nn) + 3 * s * nn + k * s * ni + 5 * s * ni
    b_flux = (25 * s * nc + (2 * i + 16 * s + 4 * v * s) * nf + s * ni + 5 * u * ni + 5 * s * ni
              + m * 5 * u * ni)
    b_step = 5 * (m * u + (1 - m) * s) * ni + 4 * s * ni + 5 * u * ni
    faces = (2 * i + 16 * s + 4 * v * s) * nf  # face index pair + area, basis, offsets, correction
    b_stage = 5 * s * nc + 20 * s * ni + faces + i * nf + i * nn + 5 * u * ni + s * ni + 5 * s * ni
    b_face = faces + 5 * s * nc + 20 * s * ni + nfv * s * nf
    b_accum = i * nn + nfv * s * nf + 5 * u * ni + s * ni + 5 * s * ni
    return {"grad_limit": b_grad + b_lim, "flux": b_flux, "step": b_step, "stage": b_stage,
            "face": b_face, "accum": b_accum,
            "per_cell_update": (b_step + 3 * (b_grad + b_lim + b_flux)) / ni}


def measured_traffic(precision, block_rows,
                     path=os.path.join(ROOT, "profiles", "hbm_traffic_tgv128.json")):
    """HBM bytes per launch of the fused stage kernel in the block layout of `block_rows` rows
    (RK stages and last stages weighted by their dispatch counts) from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated against a copy of known
    size), or None."""
    try:
        with open(path) as f:
            data = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    typ = "float" if precision == "mixed" else "double"
    # k_flux_block<type, flux, viscous, order, mode, rows, geometry prefetch>
    rows = [v for k, v in data.items()
            if k.startswith(f"k_flux_block<{typ},")
            and re.search(rf", {block_rows}(, \d+)?>$", k)]
    n = sum(r["dispatches"] for r in rows)
    if not n:
        return None
    return {"bytes": sum(r["traffic_bytes"] * r["dispatches"] for r in rows) / n,
            "average_ns": sum(r["average_ns"] * r["dispatches"] for r in rows) / n,
            "source": os.path.relpath(path, ROOT)}


def picked_shape(ms):
    """The autotuner's pick from its timings: the fastest stage shape / gradient workgroup."""
    nst = len(teal_amd.AUTOTUNE_STAGE_SHAPES)
    names = teal_amd.AUTOTUNE_SHAPES
    return names[int(np.nanargmin(ms[:nst]))] + "/" + names[nst + int(np.nanargmin(ms[nst:]))]


def make_gpu(mesh, precision):
    g = teal_amd.Gpu(mesh, precision)
    g.configure(space_order=2, flux="hllc", viscous=1, limiter="minmod", navierstokes=1,
                gamma=GAMMA, mu=1.0 / RE, prandtl=PR)
    return g


def timed_run(g, P0, steps, warmup, time_kernels, tune_steps=10):
    """Warm up, let the autotuner pick the launch shape, then time `steps` resident steps
    (no per-kernel events: they would rule out graph replay); per-kernel durations come from
    a separate run of up to 100 steps with HIP events around each launch."""
    lib = g.lib
    t = C.c_double(0.0)
    g.upload(P0)
    if warmup:
        teal_amd.check(lib.teal_gpu_run_lserk(g.h, warmup, C.byref(t), CFL, 2, 3), "warmup")
    shapes = teal_amd.AUTOTUNE_SHAPES
    ms = g.autotune(max(2, tune_steps), CFL) if tune_steps else None
    g.upload(P0)
    t = C.c_double(0.0)
    t0 = time.perf_counter()
    rc = lib.teal_gpu_run_lserk(g.h, steps, C.byref(t), CFL, 2, 3)  # ends with a sync
    t1 = time.perf_counter()
    teal_amd.check(rc, "run_lserk")
    out = {"seconds": t1 - t0, "sim_time": t.value}
    if ms is not None:
        out["shape"] = picked_shape(ms)
        out["autotune_ms_per_step"] = dict(zip(shapes, [None if np.isnan(x) else x for x in ms.tolist()]))
    if time_kernels:
        ksteps = min(steps, 100)
        g.upload(P0)
        teal_amd.check(lib.teal_gpu_timing(g.h, 3 * ksteps + 3), "timing")
        t = C.c_double(0.0)
        teal_amd.check(lib.teal_gpu_run_lserk(g.h, ksteps, C.byref(t), CFL, 2, 3), "run_lserk")
        nl = C.c_int(0)
        for k, name in enumerate(("grad", "face", "stage")):
            out[name + "_ms"] = lib.teal_gpu_timing_ms(g.h, k, C.byref(nl))
            out[name + "_launches"] = nl.value
        out["fused"] = out["face_launches"] == 0
        lib.teal_gpu_timing(g.h, 0)
    return out


def conserved(P):
    U = np.empty_like(P)
    U[:, 0] = P[:, 0]
    U[:, 1:4] = P[:, :1] * P[:, 1:4]
    U[:, 4] = P[:, 4] / (GAMMA - 1) + 0.5 * P[:, 0] * np.sum(P[:, 1:4] ** 2, axis=1)
    return U


def accuracy(g64, gmx, P0, steps, volume):
    """Advance both paths for `steps` steps with the FP64 path's step sizes (the mixed path
    is capped at the FP64 step and its own CFL limit is set 5% looser, so both take the same
    dt), then measure the mixed state against the FP64 state.

    Norms, all volume-weighted over the inner cells, d = mixed - FP64:
      l2[v]       = sqrt(sum V d_v^2 / sum V) per conserved variable (rho, rho u, rho v,
                    rho w, E) -- the formula of equations_norm (src/equations/norm.c:8-38)
      l2_relative = l2 divided by the volume-weighted RMS of the FP64 field; the three momentum
                    components are divided by the RMS of the momentum magnitude (rho w starts
                    at zero in this flow, so its own RMS is not a scale)
      linf, linf_relative (same scales)
      l2_primitive = equations_norm's formula on the primitive variables (rho, u, v, w, p)
    The "<= 1e-6" error bound is read as l2_relative_max; `target_met` says whether it holds,
    and `l2_max` (absolute) is reported beside it: the energy E ~ 180 and p ~ 71 of this state
    put the absolute figure two orders of magnitude above the relative one."""
    g64.upload(P0)
    gmx.upload(P0)
    t64 = tmx = 0.0
    for _ in range(steps):
        s0, t_new, _ = g64.lserk(t64, CFL)
        dt = t_new - t64
        t64 = t_new
        _, tmx, _ = gmx.lserk(tmx, CFL * 1.05, max_step=dt)
    ni = len(volume)
    P64 = g64.download()[:ni]
    Pmx = gmx.download()[:ni]
    U64 = conserved(P64)
    Umx = conserved(Pmx)
    d = Umx - U64
    w = volume / volume.sum()
    l2 = np.sqrt((w[:, None] * d * d).sum(axis=0))
    linf = np.abs(d).max(axis=0)
    scale = np.sqrt((w[:, None] * U64 * U64).sum(axis=0))
    scale[1:4] = np.sqrt((w * (U64[:, 1:4] ** 2).sum(axis=1)).sum())
    dp_ = Pmx - P64
    l2p = np.sqrt((w[:, None] * dp_ * dp_).sum(axis=0))
    names = ["rho", "rho_u", "rho_v", "rho_w", "E"]
    pnames = ["rho", "u", "v", "w", "p"]
    rel = l2 / scale
    return {"steps": steps, "time_fp64": t64, "time_mixed": tmx,
            "normalisation": "volume-weighted RMS of the FP64 field per conserved variable; "
                             "momentum components by the RMS momentum magnitude",
            "mixed_arithmetic": "FP32 face fluxes, gradients, limiters and geometry; each FP32 "
                                "face contribution summed into its cell in FP64 (exact, in "
                                "place of compensated FP32 sums); FP64 RK update and U",
            "l2": dict(zip(names, l2.tolist())),
            "l2_relative": dict(zip(names, rel.tolist())),
            "linf": dict(zip(names, linf.tolist())),
            "linf_relative": dict(zip(names, (linf / scale).tolist())),
            "l2_primitive": dict(zip(pnames, l2p.tolist())),
            "l2_max": float(l2.max()), "l2_relative_max": float(rel.max()),
            "target": "l2_relative_max <= 1e-6", "target_met": bool(rel.max() <= 1e-6)}


def _cpu_stepper(case):
    """(kind, mesh arrays, primitive view, step function) of the reference's own lserk
    (oracle/_ref, built from the reference sources) on `case`, or of the oracle restatement
    when that build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol

    om = ol.oracle_mesh(case)
    arrays = _abi.mesh_arrays(om)
    nc = arrays["num_cells"]
    try:
        lib = ol.ref()
        e = ol.ref_eqns(case, om)
        kind, prim = "reference", lib.ref_eqns_primitive
        step = lambda t, r: lib.ref_lserk(e, C.byref(t), r, 1e300, CFL, 2, 3)
    except OSError:
        lib = ol.oracle()
        e = ol.oracle_eqns(case, om)
        kind, prim = "port", lib.oracle_eqns_primitive
        step = lambda t, r: lib.oracle_lserk(e, C.byref(t), r, 1e300, CFL, 2, 3)
    return kind, arrays, np.ctypeslib.as_array(prim(e), shape=(nc * 5,)), step


def _tgv_case(n, nz=None, z0=0.0, z1=2 * np.pi):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol

    return ol.Case((n, n, nz or n), lo=(0, 0, z0), hi=(2 * np.pi, 2 * np.pi, z1),
                   kind="navierstokes", mu=1 / RE, prandtl=PR)


def host_cores():
    """Host CPU topology as this process sees it: logical CPUs it may run on, and one logical
    CPU per distinct physical core among them (from `lscpu -p=cpu,core,socket`), plus the
    machine's physical core count."""
    import subprocess

    allowed = sorted(os.sched_getaffinity(0))
    first = {}  # (socket, core) -> first allowed logical cpu
    total = set()
    try:
        out = subprocess.run(["lscpu", "-p=cpu,core,socket"], capture_output=True, text=True,
                             check=True).stdout
        for line in out.splitlines():
            if line.startswith("#"):
                continue
            cpu, core, sock = (int(x) for x in line.split(","))
            total.add((sock, core))
            if cpu in allowed:
                first.setdefault((sock, core), cpu)
    except (OSError, ValueError, subprocess.CalledProcessError):
        first = {(0, c): c for c in allowed}
        total = set(first)
    return {"logical_allowed": len(allowed), "physical_allowed": sorted(first.values()),
            "physical_host": len(total)}


def _step_for(step, budget_s, max_steps):
    t = C.c_double(0.0)
    r = (C.c_double * 5)()
    done = 0
    t0 = time.perf_counter()
    while True:
        step(t, r)
        done += 1
        if time.perf_counter() - t0 >= budget_s or done >= max_steps:
            break
    return done, time.perf_counter() - t0


def cpu_worker(spec):
    """Child process of the multi-core baseline: slab k of K of the n^3 TGV (z in
    [2 pi k/K, 2 pi (k+1)/K], periodic over the slab in place of the halo exchange -- the
    compute of one MPI rank of a K-rank slab decomposition), pinned to logical CPU `cpu`,
    stepped in `samples` consecutive samples of `budget` seconds from wall-clock `start`;
    prints one JSON line with the steps and seconds of each sample."""
    n, k, K, budget, start, cpu, samples = spec.split(",")
    n, k, K, budget, start = int(n), int(k), int(K), float(budget), float(start)
    cpu, samples = int(cpu), int(samples)
    if cpu >= 0:
        os.sched_setaffinity(0, {cpu})
    h = 2 * np.pi / K
    kind, a, p, step = _cpu_stepper(_tgv_case(n, n // K, k * h, (k + 1) * h))
    nc, ni = a["num_cells"], a["num_inner"]
    P0 = tgv_primitive(a["center"][:nc])
    P0[ni:] = 0
    p[:] = P0.ravel()
    late = max(0.0, time.time() - start)
    time.sleep(max(0.0, start - time.time()))
    runs = []
    for _ in range(samples):
        done, secs = _step_for(step, budget, 1000)
        runs.append((done, secs))
    print(json.dumps({"kind": kind, "cells": int(ni), "runs": runs, "late": late}), flush=True)


def cpu_baseline(n, max_cores=16, budget_s=15.0, samples=3):
    """The reference's own lserk (oracle/_ref, built from the reference sources; the oracle
    restatement when that build is absent) on this host's cores, a bounded sample of the NS
    TGV n^3: K = min(max_cores, physical cores this process may use) concurrent processes,
    each pinned to its own physical core and stepping one slab of a K-rank slab decomposition
    (the reference's MPI layout, with the exchange replaced by periodic copies, so an upper
    bound on what K MPI ranks reach).  All processes run `samples` consecutive samples of
    budget_s / samples seconds side by side; a sample's rate is the sum of the processes'
    rates, and the value is the median over the samples."""
    import subprocess

    topo = host_cores()
    cpus = topo["physical_allowed"][:max(1, max_cores)]
    cores = len(cpus)
    while cores > 1 and n % cores:
        cores -= 1
    cpus = cpus[:cores]
    per = budget_s / samples
    start = time.time() + 10.0  # after every child's setup
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                               f"{n},{k},{cores},{per},{start},{cpu},{samples}"],
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
             for k, cpu in enumerate(cpus)]
    outs = []
    for p in procs:
        o, e = p.communicate(timeout=600)
        if p.returncode:
            raise RuntimeError(f"cpu worker failed: {e[-500:]}")
        outs.append(json.loads(o.strip().splitlines()[-1]))
    rates = [sum(o["cells"] * o["runs"][s][0] / o["runs"][s][1] for o in outs)
             for s in range(samples)]
    steps = [r[0] for o in outs for r in o["runs"]]
    return {"value": float(np.median(rates)), "unit": "cell-updates/s", "cores": cores,
            "kind": outs[0]["kind"],
            "sample": f"{cores} concurrent single-thread processes, each pinned to its own "
                      f"physical core and stepping one {n}x{n}x{n // cores} slab of NS-TGV "
                      f"{n}^3 ({outs[0]['cells']} cells, periodic over the slab in place of the "
                      f"halo exchange); {samples} samples of ~{per:.0f} s ({min(steps)}-"
                      f"{max(steps)} LSERK steps per process per sample); value = median over "
                      "the samples of the sum of the processes' rates",
            "sample_rates": rates,
            "host_physical_cores": topo["physical_host"],
            "cores_allowed": {"logical": topo["logical_allowed"],
                              "physical": len(topo["physical_allowed"])},
            "core_limit": max_cores,
            "max_late_start_s": float(max(o["late"] for o in outs))}


def tet_workload(args, n=96, seed=2024):
    """Euler on n^3 hexes split into 6 tets each, interior nodes jittered by +-0.2 h (seeded),
    uniform Mach 2 flow (rho = 1, p = 1/gamma, u = 2) plus a Gaussian density bump of
    amplitude 0.1 (seeded centre, sigma 0.1), supersonic inflow at x = 0 and supersonic
    outflow elsewhere; LSQ + Barth-Jespersen, LSERK order 2 / 3 stages, CFL 0.5.
    Prints one JSON line with FP64 and mixed cell-updates/s for Roe and AUSMD."""
    lib = teal_amd.lib()
    t0 = time.perf_counter()
    mesh = lib.mesh_create_tets(n, 0.2, seed)
    lib.mesh_generate(mesh)
    a = _abi.mesh_arrays(mesh)
    nc, ni, nf = a["num_cells"], a["num_inner"], a["num_faces"]
    print(f"# tet mesh: {ni} cells, {nf} faces, built in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    free = (1.0, 2.0, 0.0, 0.0, 1.0 / GAMMA)
    x = a["center"][:nc]
    x0 = np.random.default_rng(seed).uniform(0.3, 0.7, 3)
    P0 = np.tile(np.array(free), (nc, 1))
    P0[:, 0] += 0.1 * np.exp(-((x - x0) ** 2).sum(axis=1) / (2 * 0.1 ** 2))
    P0[ni:] = 0
    bcs = {"left": ("supersonic inflow", free)}
    bcs.update({s: ("supersonic outflow", None) for s in ("right", "bottom", "top", "back", "front")})
    out = {}
    for flux in ("roe", "ausmd"):
        for name, prec in (("fp64", teal_amd.FP64), ("mixed", teal_amd.MIXED)):
            g = teal_amd.Gpu(mesh, prec)
            g.configure(space_order=2, flux=flux, viscous=0, limiter="minmod", navierstokes=0,
                        gamma=GAMMA, bcs=bcs)
            r = timed_run(g, P0, args.steps, args.warmup, time_kernels=True)
            g.close()
            ups = ni * args.steps / r["seconds"]
            out[f"{flux}_{name}"] = {"cell_updates_per_s": ups, "ms_per_step": 1e3 * r["seconds"] / args.steps,
                                     "grad_ms": r["grad_ms"], "stage_ms": r["stage_ms"]}
            print(f"# tet {flux} {name}: {1e3 * r['seconds'] / args.steps:.3f} ms/step, {ups:.4g} "
                  f"cell-updates/s; grad {r['grad_ms']:.3f} ms, stage {r['stage_ms']:.3f} ms",
                  file=sys.stderr, flush=True)
    best = out["roe_mixed"]
    print(json.dumps({
        "metric": "RK cell-updates/s, Euler on a jittered tet mesh, mixed and FP64 GPU",
        "value": best["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": best["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 flux/gradient/geometry + f64 accumulation and update (mixed)", "data": "synthetic",
        "config": {"workload": f"Euler Mach-2 bump, {n}^3 hexes x 6 jittered tets ({ni} cells), "
                               "Roe, LSQ + Barth-Jespersen, LSERK 2/3, CFL 0.5",
                   "cells": ni, "faces": nf},
        "runs": out}), flush=True)


def tube_states(rng, count):
    """Seeded left/right primitive states of `count` shock tubes: rho, p in [0.1, 1], u in
    [-1, 1], v = w = 0."""
    lr = np.zeros((count, 2, 5))
    lr[:, :, 0] = rng.uniform(0.1, 1, (count, 2))
    lr[:, :, 1] = rng.uniform(-1, 1, (count, 2))
    lr[:, :, 4] = rng.uniform(0.1, 1, (count, 2))
    return lr


def tube_primitive(a, lr):
    nc, ni = a["num_cells"], a["num_inner"]
    P = np.where((a["center"][:nc, 0] < 0.5)[:, None], lr[0], lr[1])
    P[ni:] = 0
    return P


def batch_workload(args, seed=2026):
    """Many independent Euler Riemann shock tubes: `--batch` instances on [0,1]^3 Cartesian
    meshes of n^3 cells, n drawn log-uniformly from [16, 64] (seeded), supersonic outflow at
    x = 0 and 1, symmetry on the y and z sides, seeded left/right states split at x = 0.5; HLLC,
    LSQ + minmod, LSERK order 2 / 3 stages, CFL 0.5, each instance with its own time step.
    Times the batch (one handle, every launch covering all instances, autotuned launch shape)
    against the same instances advanced one at a time (one handle each), FP64, and checks
    that both give bit-identical states; the mixed batch is timed beside them."""
    lib = teal_amd.lib()
    rng = np.random.default_rng(seed)
    sizes = np.rint(16 * 4 ** rng.uniform(0, 1, args.batch)).astype(int)
    states = tube_states(rng, args.batch)
    t0 = time.perf_counter()
    meshes, arrays, prims = [], [], []
    for i, n in enumerate(sizes):
        m = teal_amd.mesh_box((n, n, n), periodic=(0, 0, 0))
        meshes.append(m)
        arrays.append(_abi.mesh_arrays(m))
        prims.append(tube_primitive(arrays[-1], states[i]))
    cells = int(sum(a["num_inner"] for a in arrays))
    print(f"# batch: {args.batch} tubes, {cells} cells (n = {sizes.min()}..{sizes.max()}), meshes "
          f"built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    bcs = {"left": ("supersonic outflow", None), "right": ("supersonic outflow", None)}
    bcs.update({side: ("symmetry", None) for side in ("bottom", "top", "back", "front")})
    cfg = dict(space_order=2, flux="hllc", viscous=0, limiter="minmod", navierstokes=0,
               gamma=GAMMA, bcs=bcs)
    steps = args.steps

    # one at a time: each instance on its own handle, eager launches
    single_s = 0.0
    single_P = []
    for i, m in enumerate(meshes):
        g = teal_amd.Gpu(m, teal_amd.FP64)
        g.configure(**cfg)
        g.upload(prims[i])
        t = C.c_double(0.0)
        t1 = time.perf_counter()
        teal_amd.check(lib.teal_gpu_run_lserk(g.h, steps, C.byref(t), CFL, 2, 3), "run_lserk")
        single_s += time.perf_counter() - t1
        single_P.append(g.download()[:arrays[i]["num_inner"]])
        g.close()
    print(f"# one at a time (fp64): {single_s:.2f} s, {cells * steps / single_s:.4g} cell-updates/s",
          file=sys.stderr, flush=True)

    runs = {}
    for name, prec in (("fp64", teal_amd.FP64), ("mixed", teal_amd.MIXED)):
        b = teal_amd.Batch(meshes, prec)
        b.configure(**cfg)
        b.upload_batch(prims)
        ms = b.autotune(max(2, args.tune_steps), CFL)
        shapes = teal_amd.AUTOTUNE_SHAPES
        best = picked_shape(ms)
        times = np.zeros(args.batch)
        if args.warmup:
            b.run_lserk_batch(args.warmup, times, CFL)
        b.upload_batch(prims)
        times = np.zeros(args.batch)
        t1 = time.perf_counter()
        times, _ = b.run_lserk_batch(steps, times, CFL)  # ends with a sync
        secs = time.perf_counter() - t1
        r = {"seconds": secs, "cell_updates_per_s": cells * steps / secs,
             "ms_per_step": 1e3 * secs / steps, "autotune_ms_per_step": dict(zip(shapes, ms.tolist())),
             "shape": best, "layout": b.info()}
        if name == "fp64":
            got = b.download_batch()
            r["equal_to_one_at_a_time"] = bool(all(
                np.array_equal(got[i][:a["num_inner"]], single_P[i]) for i, a in enumerate(arrays)))
        # per-kernel times of the stage kernel over a short eager run (events disable graphs)
        b.upload_batch(prims)
        teal_amd.check(lib.teal_gpu_timing(b.h, 3 * args.kernel_steps + 3), "timing")
        b.run_lserk_batch(args.kernel_steps, np.zeros(args.batch), CFL)
        nl = C.c_int(0)
        for k, kern in enumerate(("grad", "face", "stage")):
            r[kern + "_ms"] = lib.teal_gpu_timing_ms(b.h, k, C.byref(nl))
        lib.teal_gpu_timing(b.h, 0)
        runs[name] = r
        b.close()
        print(f"# batch {name}: {r['ms_per_step']:.3f} ms/step, {r['cell_updates_per_s']:.4g} "
              f"cell-updates/s, shape {best}, autotune {np.round(ms, 4).tolist()} ms/step; grad "
              f"{r['grad_ms']:.3f} ms, stage {r['stage_ms']:.3f} ms", file=sys.stderr, flush=True)

    fp = runs["fp64"]
    u = teal_amd.lib().teal_mesh_union((C.c_void_p * len(meshes))(*meshes), len(meshes))
    ua = _abi.mesh_arrays(u)
    model = byte_model(ua["num_inner"], ua["num_cells"], ua["num_faces"], mixed=False, viscous=False)
    lib.mesh_destroy(u)
    fused = fp["layout"]["fused"] == 1
    kern = "stage" if fused else "accum"
    achieved = model[kern] / (fp["stage_ms"] * 1e-3) / 1e9 if fp["stage_ms"] > 0 else None
    print(json.dumps({
        "metric": "RK cell-updates/s, batched Euler shock tubes vs one-at-a-time launches",
        "value": fp["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": fp["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.batch} Euler shock tubes, n^3 cells with n log-uniform in "
                               f"[16, 64] ({cells} cells), x outflow, y/z symmetry, HLLC, "
                               "LSQ + minmod, LSERK 2/3, CFL 0.5",
                   "cells": cells, "instances": args.batch, "seed": seed},
        "one_at_a_time": {"seconds": single_s, "cell_updates_per_s": cells * steps / single_s},
        "speedup_vs_one_at_a_time": fp["cell_updates_per_s"] / (cells * steps / single_s),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None,
                     "kernel": kern},
        "runs": runs}), flush=True)


def channel_mesh(n):
    """The 64^3-style channel [0,2] x [-1,1] x [0,1]: periodic in x and z, node levels in y
    stretched geometrically (ratio 1.05) toward both walls."""
    lib = teal_amd.lib()
    mesh = lib.mesh_create(teal_amd.Vector(0, -1, 0), teal_amd.Vector(2, 1, 1),
                           teal_amd.Triple(n, n, n), teal_amd.Triple(1, 0, 1))
    half = n // 2
    widths = 1.05 ** np.arange(half)
    edges = np.concatenate([[0.0], np.cumsum(widths)]) / widths.sum()
    levels = np.concatenate([-1 + edges, (1 - edges[::-1])[1:]])
    coord = _abi.mesh_arrays(mesh)["coord"]
    coord[:, 1] = levels[np.rint((coord[:, 1] + 1) / (2.0 / n)).astype(int)]
    lib.mesh_generate(mesh)
    return mesh


def channel_workload(args, n=64, seed=5):
    """Navier-Stokes implicit Euler with Jacobian-free Newton-GMRES (NewtonKrylov {0.1, 0.3,
    30}) in the stretched channel of channel_mesh: no-slip (adiabatic) walls in y, parabolic
    profile plus 1% seeded noise, Re = 100 on the half height, M = 0.2, HLLC + central viscous,
    unlimited second order.  Runs each path until the largest residual component falls below
    1e-8 of its first value (the reference's termination test, src/simulation/run.c:80-82,
    taken relative), or `args.steps` steps, or `args.budget` seconds.

    No forcing: the reference's only Navier-Stokes wall is adiabatic
    (src/navierstokes/boundary.c:6-26), so a bulk-velocity forcing source (periodic_hill)
    feeds the gas heat it can never lose and the energy residual never falls; the steady state
    reached here is the decay of the initial flow to rest."""
    mesh = channel_mesh(n)
    a = _abi.mesh_arrays(mesh)
    nc, ni = a["num_cells"], a["num_inner"]
    y = a["center"][:nc, 1]
    P0 = np.zeros((nc, 5))
    P0[:, 0] = 1.0
    P0[:, 1] = (1 - y ** 2) * (1 + 0.01 * np.random.default_rng(seed).standard_normal(nc))
    P0[:, 4] = 1.0 / (GAMMA * 0.2 ** 2)
    P0[ni:] = 0
    out = {}
    for name, prec in (("fp64", teal_amd.FP64), ("mixed", teal_amd.MIXED)):
        g = teal_amd.Gpu(mesh, prec)
        g.configure(space_order=2, flux="hllc", viscous=1, limiter=None, navierstokes=1,
                    gamma=GAMMA, mu=0.01, prandtl=PR,
                    bcs={"bottom": ("wall", None), "top": ("wall", None)})
        g.upload(P0)
        g.implicit_euler(0.0, args.courant, 0.1, 0.3, 30)  # warm-up step (allocations)
        g.upload(P0)
        t, newton, krylov, res0, hist = 0.0, 0, 0, None, []
        t0 = time.perf_counter()
        steps = 0
        while steps < args.steps:
            _, t, res = g.implicit_euler(t, args.courant, 0.1, 0.3, 30)
            steps += 1
            st = g.krylov_stats()
            newton += st["newton"]
            krylov += st["krylov"]
            r = float(max(res))
            res0 = r if res0 is None else res0
            elapsed = time.perf_counter() - t0
            if steps % 250 == 1 or r <= 1e-8 * res0:
                hist.append((steps, t, r / res0))
                print(f"#   {name} step {steps}: t {t:.4g}, residual {r:.3e} (drop {r / res0:.2e}), "
                      f"{newton} Newton / {krylov} GMRES iterations, {elapsed:.1f} s",
                      file=sys.stderr, flush=True)
            if r <= 1e-8 * res0 or elapsed > args.budget:
                break
        secs = time.perf_counter() - t0
        g.close()
        converged = r <= 1e-8 * res0
        out[name] = {"converged": converged, "steps": steps, "seconds": secs,
                     "time_to_solution_s": secs if converged else None,
                     "ms_per_step": 1e3 * secs / steps, "newton_iterations": newton,
                     "gmres_iterations": krylov, "matvecs_per_s": krylov / secs,
                     "residual_first": res0, "residual_last": r, "drop": r / res0,
                     "sim_time": t, "history": hist}
        print(f"# channel {name}: {'converged' if converged else 'stopped'} after {steps} steps, "
              f"{secs:.1f} s, drop {r / res0:.2e}, {newton} Newton / {krylov} GMRES iterations, "
              f"{krylov / secs:.0f} matvecs/s", file=sys.stderr, flush=True)
    mx = out["mixed"]
    print(json.dumps({
        "metric": "implicit Euler (Newton-GMRES) time to a 1e-8 residual drop, NS channel 64^3",
        "value": mx["time_to_solution_s"], "unit": "s", "n_gpus": 1, "steps": mx["steps"],
        "warmup": 1, "ms_per_step": mx["ms_per_step"], "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 matvec residuals + f64 Krylov vectors / Gram-Schmidt (mixed)",
        "data": "synthetic",
        "config": {"workload": f"NS channel {n}^3, walls stretched 1.05, Re 100, M 0.2, unforced "
                               f"decay to rest, implicit Euler CFL {args.courant}, NewtonKrylov "
                               "(0.1, 0.3, 30), unlimited second order",
                   "cells": ni, "termination": "max residual <= 1e-8 x first"},
        "runs": out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 1000; tet and batch 500; channel: at most 30000 "
                         "implicit steps toward steady state)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--accuracy-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-cores", type=int, default=16,
                    help="most physical host cores for the reference CPU baseline (one pinned "
                         "process each); 16 is the host share of one GPU on an 8-GPU MI355X node")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates rocprofv3 FETCH_SIZE/"
                         "WRITE_SIZE against a kernel of known traffic)")
    ap.add_argument("--batch", type=int, default=256, help="batch workload: instances")
    ap.add_argument("--tune-steps", type=int, default=10, help="batch workload: autotuner steps")
    ap.add_argument("--kernel-steps", type=int, default=20,
                    help="batch workload: steps of the per-kernel timing run")
    ap.add_argument("--workload", choices=["tgv", "tet", "channel", "batch"], default="tgv",
                    help="tgv: the NS Taylor-Green vortex (the headline metric); tet: Euler "
                         "Mach-2 bump on the jittered tetrahedral mesh (Roe and AUSMD); channel: "
                         "NS implicit Euler (Newton-GMRES) in a 64^3 stretched-wall channel; "
                         "batch: 256 Euler shock tubes of 16^3..64^3 cells, batched vs one at "
                         "a time")
    ap.add_argument("--courant", type=float, default=200.0, help="channel workload CFL")
    ap.add_argument("--budget", type=float, default=240.0,
                    help="channel workload: wall-clock seconds per path before it stops")
    ap.add_argument("--channel-n", type=int, default=64, help="channel workload cells per side")
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(args.cpu_worker)
        return
    if args.steps is None:  # the step counts of the configurations
        args.steps = {"batch": 500, "tet": 500, "channel": 30000}.get(args.workload, 1000)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this library drives one GPU per process; runs on several GPUs are "
              "not part of it", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        import torch

        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 1 << 30, "copy_probe_bytes_written": 1 << 30,
                          "launches": 3}))
        return

    if args.workload == "tet":
        tet_workload(args)
        return
    if args.workload == "channel":
        channel_workload(args, n=args.channel_n)
        return
    if args.workload == "batch":
        batch_workload(args)
        return
    n = args.n
    mesh = teal_amd.mesh_box((n, n, n), (0.0, 0.0, 0.0), (2 * np.pi,) * 3, (1, 1, 1))
    a = _abi.mesh_arrays(mesh)
    nc, ni, nf = a["num_cells"], a["num_inner"], a["num_faces"]
    P0 = tgv_primitive(a["center"][:nc])
    P0[ni:] = 0
    volume = a["volume"][:ni].copy()

    cpu = None  # first, before this process touches the GPU (it starts child processes)
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(n, max(1, args.cpu_cores))
        except Exception as exc:  # reported, not fatal
            cpu = {"value": None, "error": str(exc)}
        print(f"# cpu baseline: {cpu}", file=sys.stderr, flush=True)

    results = {}
    gpus = {}
    for name, prec in (("fp64", teal_amd.FP64), ("mixed", teal_amd.MIXED)):
        g = make_gpu(mesh, prec)
        gpus[name] = g
        r = timed_run(g, P0, args.steps, args.warmup, time_kernels=True)
        r["layout"] = g.info()
        r["ms_per_step"] = 1e3 * r["seconds"] / args.steps
        r["cell_updates_per_s"] = ni * args.steps / r["seconds"]
        model = byte_model(ni, nc, nf, mixed=(prec == teal_amd.MIXED))
        r["bytes_per_cell_update"] = model["per_cell_update"]
        r["hbm_fraction_step"] = r["cell_updates_per_s"] * model["per_cell_update"] / (HBM_PEAK_GBS * 1e9)
        kernels = (("grad", "grad_limit"), ("stage", "stage" if r["fused"] else "accum"))
        if not r["fused"]:
            kernels += (("face", "face"),)
        for kern, key in kernels:
            r[kern + "_bytes"] = model[key]
            r[kern + "_gbs"] = model[key] / (r[kern + "_ms"] * 1e-3) / 1e9
        results[name] = r
        print(f"# {name}: {r['ms_per_step']:.3f} ms/step, {r['cell_updates_per_s']:.4g} "
              f"cell-updates/s; " + ", ".join(f"{k} {r[k + '_ms']:.3f} ms ({r[k + '_gbs']:.0f} GB/s)"
                                              for k, _ in kernels),
              file=sys.stderr, flush=True)

    # the configuration's 1000 steps, whatever the number of timed steps
    acc_steps = args.accuracy_steps if args.accuracy_steps is not None else 1000
    err = accuracy(gpus["fp64"], gpus["mixed"], P0, acc_steps, volume)
    print(f"# mixed vs fp64 after {acc_steps} steps: L2 {err['l2_max']:.3g} "
          f"(relative {err['l2_relative_max']:.3g})", file=sys.stderr, flush=True)
    for g in gpus.values():
        g.close()

    mx, fp = results["mixed"], results["fp64"]
    KEYS = ["cell_updates_per_s", "ms_per_step", "hbm_fraction_step", "bytes_per_cell_update",
            "fused", "layout", "grad_ms", "grad_gbs", "stage_ms", "stage_gbs", "face_ms", "face_gbs",
            "shape", "autotune_ms_per_step"]
    timed = [k for k in ("grad", "face", "stage") if k + "_gbs" in mx]
    dom = max(timed, key=lambda k: mx[k + "_ms"])  # dominant kernel
    traffic = (measured_traffic("mixed", mx["layout"]["block_rows"])
               if (dom == "stage" and mx["fused"]) else None)
    line = {
        "metric": "RK cell-updates/s, TGV 128^3, mixed vs FP64 GPU at stated L2 error; %HBM",
        "value": mx["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mx["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 flux/gradient/geometry + f64 accumulation and update (mixed)",
        "data": "synthetic",
        "config": {
            "workload": f"Navier-Stokes Taylor-Green vortex, periodic {n}^3 hex mesh on [0,2pi]^3 "
                        f"({ni} cells), rho=1, M=0.1, Re=1600, Pr=0.71, HLLC + central viscous, "
                        "LSQ gradients + Barth-Jespersen, LSERK order 2 / 3 stages, CFL 0.5",
            "cells": ni, "faces": nf,
        },
        "fp64": {k: fp[k] for k in KEYS if k in fp},
        "mixed": {k: mx[k] for k in KEYS if k in mx},
        "mixed_over_fp64": mx["cell_updates_per_s"] / fp["cell_updates_per_s"],
        "mixed_error_vs_fp64": err,
        "roofline": {
            "bound": "hbm",
            "kernel": f"{dom} kernel of the mixed path",
            "achieved": mx[dom + "_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": mx[dom + "_gbs"] / HBM_PEAK_GBS,
            "traffic": traffic["bytes"] if traffic else None,
            "traffic_source": traffic["source"] if traffic else None,
            "traffic_rate_gbs": (traffic["bytes"] / (traffic["average_ns"] * 1e-9) / 1e9
                                 if traffic else None),
            "algorithmic_bytes_per_launch": mx[dom + "_bytes"],
            "mean_launch_ms": mx[dom + "_ms"],
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
