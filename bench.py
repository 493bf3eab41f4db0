"""Benchmark: VPM steps/s on one MI355X, two coaxial leapfrogging rings on the 256^3 grid,
mixed form (FP32 projection, interpolation and sine-transform passes with the inner-box solve's
inverse passes in FP64; FP64 boundary sums, stencils and diagnostics) beside the plain FP64
form, plus a leg of 64 independent rings stepped as one batch (queued and graph-replayed) and
a leg timing the gradient of a probe energy through 10 checkpointed RK2 steps (adjoint).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

A step is the application's step with RK2: two vpm(WhatToDo=2) evaluations (projection,
yaps infinite-domain Poisson solve, curl + stretching, interpolation), the particle
updates, vpm(WhatToDo=0) and remesh_particles_3d(1) -- all on the device, particles
resident in HBM.  Workload: rings R=1 and R=1.2, Gaussian core sigma=0.08, axial
separation 0.5, box [-2,2]^3, h=1/64 (N=273 nodes per axis, 289^3 yaps block, 43^3
coarse grid), particles at cell centres where |w|>1e-6, remeshed once before timing
(test.f90:50-55), dt=0.01.  Both forms start from the same remeshed particles; K steps of
each are timed.  Prints one JSON line; `value` is the mixed form's steps/s.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "VPM steps/s + HBM GB/s, 256³/2M particles; mixed-FP32 vs FP64 @1e-5 L2"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--h", type=float, default=1.0 / 64)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=256,
                    help="boundary-sum target stride of the bounded CPU sample")
    ap.add_argument("--forms", default="fp64,mixed", help="forms to time: fp64, mixed or both")
    ap.add_argument("--only", choices=("batch", "adjoint"), default=None,
                    help="run one side leg alone and print its JSON object (for profiling; not the bench line)")
    ap.add_argument("--adjoint-steps", type=int, default=10, help="steps of the adjoint leg (0: skip)")
    ap.add_argument("--adjoint-ckpt", type=int, default=2, help="checkpoint spacing of the adjoint leg")
    ap.add_argument("--batch-steps", type=int, default=5,
                    help="steps of the 64-ring batch leg, queued and graph-replayed (0: skip the leg)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(h):
    from vpm_free_amd import cases
    return cases.leapfrog_rings(h=h)


def stencil_nodes(xp, g):
    """Distinct grid nodes covered by the particles' 4x4x4 M4' stencils (nodes inode-1..inode+2)."""
    n = g["nn"]
    x0, h = np.array(g["xbound"][:3]), np.array(g["dpm"])
    cell = np.floor((xp - x0) / h).astype(np.int64)  # inode - 1
    occ = np.zeros((n[2], n[1], n[0]), dtype=bool)
    ok = np.all((cell >= 0) & (cell < np.array(n)), axis=1)
    c = cell[ok]
    occ[c[:, 2], c[:, 1], c[:, 0]] = True
    for ax in range(3):
        out = occ.copy()
        for s in (-1, 1, 2):
            out |= np.roll(occ, s, axis=ax)
        occ = out
    return int(occ.sum())


def chosen_variant(tuning, mixed, axis):
    """The tuner's choice for the y or z pass of the 289^3 block, as its report names it
    ("8.8.9 plan, capped, 2 work-groups per CU"), or None."""
    head = "sine transforms %s pass %s L=%d " % (axis, "FP32" if mixed else "FP64", 576 if mixed else 288)
    for line in tuning:
        # (not the zero-BC solve's z pass, which stops after the diagonal)
        if line.startswith(head) and " -> " in line and "(forward and diagonal" not in line:
            return line.split(" -> ", 1)[1].rsplit(" (", 1)[0]
    return None


def measured_traffic(mixed, axis, variant):
    """HBM bytes per launch of the y or z pass in the tuner's chosen variant, from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy), or None when
    that variant was not profiled."""
    path = os.path.join(ROOT, "profiles", "dst_passes_289_traffic.md")
    key = "traffic_bytes_per_launch_" + axis
    want = "float" if mixed else "double"
    try:
        for line in open(path):
            t = line.strip().split(None, 2)
            if len(t) == 3 and t[0] == key and t[1].startswith(want + "=") and t[2] == variant:
                return int(t[1].split("=")[1])
    except (OSError, ValueError):
        return None
    return None


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def cpu_baseline(h, xp, qp, stride):
    """Oracle (C/OpenMP restatement of the reference path) on the host cores: one
    vpm(WhatToDo=2) with the boundary sums evaluated on every stride-th boundary node,
    scaled back to all nodes, plus one remesh; step = 2 x vpm2 + remesh."""
    from oracle import oracle as O
    p = O.Params.make(h)
    g = O.define_sizes(p, xp)
    O.set_boundsum_stride(stride)
    O.timers(reset=True)
    t0 = time.perf_counter()
    O.vpm2(p, g, xp, qp)
    wall_vpm = time.perf_counter() - t0
    t = O.timers(reset=True)
    O.set_boundsum_stride(1)
    vpm_est = wall_vpm - t[2] + t[2] * stride
    t0 = time.perf_counter()
    O.remesh(p, g, xp, qp)
    wall_rm = time.perf_counter() - t0
    step = 2.0 * vpm_est + wall_rm
    return {
        "value": 1.0 / step,
        "unit": "steps/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": (f"oracle vpm(2) on the same 256^3 leapfrog particles with boundary sums on 1/{stride} "
                   f"of the boundary nodes (measured {wall_vpm:.1f}s, boundary part {t[2]:.2f}s scaled x{stride}) "
                   f"+ one full remesh ({wall_rm:.1f}s); step = 2 vpm(2) + remesh = {step:.1f}s"),
    }


def time_form(ctx, precision, xp0, qp0, args, g):
    """Upload the starting particles, run W warmup steps and K timed steps in one form."""
    ctx.set_precision(precision)
    ctx.upload(xp0, qp0)
    circ0, imp0 = ctx.diagnostics()
    ntime = 1
    for _ in range(args.warmup):
        ctx.step_rk2(args.dt, ntime=ntime)
        ntime += 1
    stage_tot, nparts = {}, []
    ctx.kernel_times(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.step_rk2(args.dt, ntime=ntime)
        ntime += 1
        for k, v in ctx.stage_times().items():
            stage_tot[k] = stage_tot.get(k, 0.0) + v
        nparts.append(ctx.count())
    elapsed = time.perf_counter() - t0
    kt = ctx.kernel_times()
    work = ctx.kernel_work()
    circ1, imp1 = ctx.diagnostics()
    xf, _, _, _ = ctx.download()
    npart = float(np.mean(nparts))
    G = g["nn"][0] * g["nn"][1] * g["nn"][2]
    res = {
        "steps_per_s": args.steps / elapsed,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "particles": int(npart),
        "particles_range": [int(nparts[0]), int(nparts[-1])],
        "stage_ms_per_step": {k: round(v / args.steps, 3) for k, v in stage_tot.items()},
        "diagnostics": {"circulation_drift": float(np.max(np.abs(circ1 - circ0))),
                        "impulse_rel_drift": float(np.linalg.norm(imp1 - imp0) / np.linalg.norm(imp0))},
        "kernels": {},
    }
    # per-kernel device time from each dispatch's start/stop events on the library's stream
    for name in ("dst_y_fine", "dst_z_fine"):
        dst_ms, dst_n = kt[name]
        if dst_n:
            b = work[name]  # read + write of every interior point: 2 x points x comps x word
            gbs = b / (dst_ms / dst_n * 1e-3) / 1e9
            res["kernels"][name] = {"ms": round(dst_ms / dst_n, 4), "launches": int(dst_n),
                                    "ms_per_step": round(dst_ms / args.steps, 4),
                                    "algorithmic_bytes": int(b), "GB/s": round(gbs, 1),
                                    "frac": round(gbs / HBM_PEAK_GBS, 4)}
    bs_ms, bs_n = kt["boundsum_fine"]
    if bs_n:
        fl = work["boundsum_fine"]
        res["kernels"]["boundsum_fine"] = {"ms": round(bs_ms / bs_n, 4), "launches": int(bs_n), "flops": fl,
                                           "TFLOP/s": round(fl / (bs_ms / bs_n * 1e-3) / 1e12, 3)}
    p2m_ms, p2m_n = kt["p2m"]
    if p2m_n:
        b = npart * 6 * 8 + G * 3 * 8  # FP64 positions + strengths read, FP64 RHS written
        gbs = b / (p2m_ms / p2m_n * 1e-3) / 1e9
        res["kernels"]["p2m"] = {"ms": round(p2m_ms / p2m_n, 4), "launches": int(p2m_n), "algorithmic_bytes": int(b),
                                 "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    m2p_ms, m2p_n = kt["m2p"]
    if m2p_n:
        T = stencil_nodes(xf, g)
        b = npart * 4 * 8 + npart * 6 * 8 + T * 6 * 8
        gbs = b / (m2p_ms / m2p_n * 1e-3) / 1e9
        res["kernels"]["m2p"] = {"ms": round(m2p_ms / m2p_n, 4), "launches": int(m2p_n), "algorithmic_bytes": int(b),
                                 "stencil_nodes": int(T), "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return res


def batch_leg(V, steps):
    """64 independent rings (radius U[0.5,1.5], sigma U[0.05,0.15], random axis, grids 32^3..128^3
    drawn with probability 1/N^3, seed 3), mixed form: one RK2 step with remesh of every ring,
    queued on the rings' streams vs one captured hipGraph replayed."""
    from vpm_free_amd import cases
    clouds = cases.batch_rings(seed=3, count=64)
    cells = int(sum(round(2 * np.abs(xp).max() / h) ** 3 for h, xp, _ in clouds))
    out = {"clouds": len(clouds), "cells": cells, "particles": int(sum(xp.shape[0] for _, xp, _ in clouds))}
    for graph in (False, True):
        with V.Batch(len(clouds)) as b:
            for i, (h, xp, qp) in enumerate(clouds):
                c = b.cloud(i, V.Params.make(h))
                c.set_precision(1)
                c.upload(xp, qp)
                c.vpm_resident(0, 0)
            b.step_rk2(0.01, graph=graph)  # prepares the batch (and captures the graph)
            dev, host = [], []
            t0 = time.perf_counter()
            for _ in range(steps):
                b.step_rk2(0.01, graph=graph)
                d, e = b.times()
                dev.append(d)
                host.append(e)
            wall = (time.perf_counter() - t0) / steps
        key = "graph" if graph else "queued"
        out[key] = {"ms_per_step": round(1000 * wall, 3), "device_ms": round(float(np.median(dev)), 3),
                    "host_enqueue_ms": round(float(np.median(host)), 3),
                    "clouds_per_s": round(len(clouds) / wall, 1)}
        log(f"batch {key}: {1000 * wall:.2f} ms per step of {len(clouds)} clouds "
            f"(device {np.median(dev):.2f} ms, host queueing {np.median(host):.2f} ms)")
    return out


def adjoint_leg(V, nsteps, ckpt, dt=0.01):
    """Gradient of the probe energy J = 1/2 sum |u(p)|^2 at 4096 random probes (seed 4, inside
    [-1.5,1.5]^3) after `nsteps` RK2 steps with remesh, with respect to the starting particle
    strengths and positions: one ring (R=1, sigma=0.22) on [-2,2]^3 at h=1/32 (128^3 class:
    N=129 global nodes), plain FP64 form, particles kept every `ckpt` steps.  Timed against the
    same steps run forward alone."""
    from vpm_free_amd import cases
    h = 1.0 / 32
    xp, qp = cases.single_ring(h=h, sigma=0.22)
    probes = np.random.default_rng(4).uniform(-1.5, 1.5, size=(4096, 3))
    with V.VPM(V.Params.make(h)) as c:
        c.set_precision(0)
        c.upload(xp, qp)
        c.vpm_resident(0, 0)
        g = V.geom_dict(c.geometry)
        c.step_rk2(dt, ntime=1)  # warm-up: buffers and tuned launches
        c.upload(xp, qp)
        t0 = time.perf_counter()
        for k in range(nsteps):
            c.step_rk2(dt, ntime=k + 1)
        n_end = c.count()
        fwd = time.perf_counter() - t0
        c.adjoint_rollout(xp, qp, probes, min(nsteps, 2), dt, ckpt=ckpt)  # warm-up: plans, records
        t0 = time.perf_counter()
        J, dq, dx = c.adjoint_rollout(xp, qp, probes, nsteps, dt, ckpt=ckpt)
        adj = time.perf_counter() - t0
        # the same gradient with every step's record kept from the forward pass (no recomputation)
        c.adjoint_rollout(xp, qp, probes, nsteps, dt, ckpt=0)  # warm-up: all records allocated
        t0 = time.perf_counter()
        J0, dq0, dx0 = c.adjoint_rollout(xp, qp, probes, nsteps, dt, ckpt=0)
        adj0 = time.perf_counter() - t0
    agree = float(max(np.linalg.norm(dq0 - dq) / np.linalg.norm(dq), np.linalg.norm(dx0 - dx) / np.linalg.norm(dx)))
    out = {"workload": "one ring R=1 sigma=0.22, h=1/32, box [-2,2]^3, %d RK2 steps dt=%g + remesh, 4096 probes"
                       % (nsteps, dt),
           "grid_nodes": g["nn"], "block_nodes": g["nnbl"][0], "particles": [int(xp.shape[0]), int(n_end)],
           "ckpt": ckpt, "forward_ms": round(1000 * fwd, 2), "adjoint_ms": round(1000 * adj, 2),
           "adjoint_over_forward": round(adj / fwd, 3),
           "keep_all": {"adjoint_ms": round(1000 * adj0, 2), "adjoint_over_forward": round(adj0 / fwd, 3),
                        "rel_diff_vs_ckpt": agree},
           "J": J, "grad_norms": {"dq": float(np.linalg.norm(dq)), "dx": float(np.linalg.norm(dx))}}
    log(f"adjoint: {nsteps} steps forward {1000 * fwd:.1f} ms, gradient (forward + backward) {1000 * adj:.1f} ms "
        f"= {adj / fwd:.2f}x with checkpoints every {ckpt} steps, {1000 * adj0:.1f} ms = {adj0 / fwd:.2f}x keeping "
        f"every step; particles {xp.shape[0]} -> {n_end}")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU (the yaps blocks would be the units of a "
                                   "sharded run; it is not launched)", "n_gpus": args.gpus}))
        sys.exit(2)
    from vpm_free_amd import vpm as V

    if args.only == "batch":
        print(json.dumps(batch_leg(V, args.batch_steps)))
        return
    if args.only == "adjoint":
        print(json.dumps(adjoint_leg(V, args.adjoint_steps, args.adjoint_ckpt)))
        return
    t0 = time.perf_counter()
    xp, qp = workload(args.h)
    log(f"workload: {xp.shape[0]} particles generated in {time.perf_counter() - t0:.1f}s")
    ctx = V.VPM(V.Params.make(args.h))
    ctx.upload(xp, qp)
    ctx.vpm_resident(0, 0)  # define_sizes (NTIME=0): box fixed from here on (idefine=1)
    ctx.remesh_resident(1)  # initial remesh, as the reference driver does before its loop
    g = V.geom_dict(ctx.geometry)
    nb = g["nnbl"][0]
    xp0, qp0, _, _ = ctx.download()
    log(f"grid N={g['nn']} block={nb} coarse={g['nn_coarse']} particles after remesh={xp0.shape[0]}")

    # accuracy of the mixed form: one vpm(2) on the same particles in both forms
    ctx.set_precision(0)
    ctx.vpm_resident(2, 1)
    _, _, up64, gp64 = ctx.download()
    ctx.set_precision(1)
    ctx.upload(xp0, qp0)
    ctx.vpm_resident(2, 1)
    _, _, up32, gp32 = ctx.download()
    acc = {"up_rel_l2": rel_l2(up32, up64), "gp_rel_l2": rel_l2(gp32, gp64)}
    log(f"mixed vs FP64 on one vpm(2): UP rel L2 {acc['up_rel_l2']:.3e}, GP rel L2 {acc['gp_rel_l2']:.3e}")

    forms = {}
    for name in args.forms.split(","):
        forms[name] = time_form(ctx, 1 if name == "mixed" else 0, xp0, qp0, args, g)
        forms[name]["tuning"] = ctx.tuning_report().strip().splitlines()
        f = forms[name]
        log(f"{name}: {f['ms_per_step']:.2f} ms/step over {args.steps} steps; particles "
            f"{f['particles_range'][0]}..{f['particles_range'][1]}")
        log("  stage ms/step: " + ", ".join(f"{k}={v:.2f}" for k, v in f["stage_ms_per_step"].items()))
        log("  kernels: " + json.dumps(f["kernels"]))
    ctx.close()
    batch = batch_leg(V, args.batch_steps) if args.batch_steps > 0 else None
    adjoint = adjoint_leg(V, args.adjoint_steps, args.adjoint_ckpt) if args.adjoint_steps > 0 else None

    mixed = "mixed" in forms
    head = forms["mixed"] if mixed else forms["fp64"]
    # dominant kernel: the y or the z pass of the 289^3 block's sine transforms, whichever
    # takes more device time per step (both read and write the block once per launch)
    ks = head["kernels"]
    dom = max(("dst_y_fine", "dst_z_fine"), key=lambda k: ks.get(k, {}).get("ms_per_step", -1.0))
    dk = ks.get(dom, {})
    axis = "y" if dom == "dst_y_fine" else "z"
    variant = chosen_variant(head.get("tuning", []), mixed, axis)
    roofline = {
        "bound": "hbm",
        "achieved": dk.get("GB/s"),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": dk.get("frac"),
        "traffic": measured_traffic(mixed, axis, variant),
        "traffic_source": "profiles/dst_passes_289_traffic.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
        "kernel": "k_dst_wave<DST_%s,%s>: %s pass of the 289^3 block's sine-transform Poisson solve (3 comps%s)"
                  % (axis.upper(), "float" if mixed else "double", axis,
                     "" if axis == "y" else "; forward transform, diagonal and inverse transform -- the second"
                     " fine solve's; the zero-BC solve's z pass stops after the diagonal and is not counted"),
        "variant": variant,
        "ms_per_launch": dk.get("ms"),
        "launches_per_step": round(dk.get("launches", 0) / args.steps, 2),
        "algorithmic_bytes_per_launch": dk.get("algorithmic_bytes"),
    }

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline sample ...")
        cpu = cpu_baseline(args.h, xp0, qp0, args.cpu_stride)
        log(f"cpu baseline: {cpu['value']:.4g} steps/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": round(head["steps_per_s"], 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (mixed form: FP32 projection weights and sums, fine and coarse Poisson passes, the inner-box "
                 "solve's forward passes, interpolation; FP64 inner-box inverse passes, particles, grids in HBM, "
                 "boundary sums, curl/stretching, remesh, diagnostics)"
                 if mixed else "f64",
        "data": "synthetic",
        "config": {
            "workload": "two coaxial leapfrogging vortex rings (R=1, 1.2; sigma=0.08; sep 0.5), box [-2,2]^3, "
                        "h=1/64 (256^3 class: N=273, yaps block 289^3, coarse 43^3), RK2 dt=0.01 + remesh each step",
            "grid_nodes": g["nn"],
            "block_nodes": nb,
            "particles": head["particles"],
            "form": "mixed" if mixed else "fp64",
        },
        "particles_per_s": round(head["particles"] * head["steps_per_s"], 1),
        "mixed_vs_fp64": {
            "speedup": round(forms["mixed"]["steps_per_s"] / forms["fp64"]["steps_per_s"], 3)
            if mixed and "fp64" in forms else None,
            "accuracy_one_vpm2": acc,
        },
        "forms": forms,
        "batched_64_rings": batch,
        "adjoint_ring_128": adjoint,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
