"""Benchmark: 512^3 FP64 G-equation WENO5-RK3 + HCR-2 on one MI355X.

Workload (the 3D roofline workload): 512^3 interior cells on the unit
cube, periodic, G0 = x - 0.3 - eta(y,z) with 8 seeded surface modes, velocity
(1,0,0) + 16 seeded solenoidal modes of amplitude 0.1, S_L = 0.08, CFL 0.5,
HCR-2 (2 iterations) after every 5th RK3 step.  A step is one RK3 step plus
1/5 of an HCR-2 invocation; inputs are resident in HBM before timing starts.

Prints ONE JSON line (rank 0).  Options:
  --steps K --warmup W     timed / untimed steps (K a multiple of 5 keeps the
                           HCR-2 share exact; defaults 20 and 10: the first few
                           steps after start-up run slower while the clocks
                           settle)
  --mode exact|fast        arithmetic mode of the solver
  --n N                    grid edge (default 512)
  --no-cpu-baseline        skip the CPU reference timing
  --no-adjoint             skip the reduced adjoint-vs-FD leg that the default
                           line carries under "adjoint" (16 parameter sets)
  --no-extra               skip the "other_mode" (the other arithmetic mode,
                           10 steps of the same workload), "batch" (the
                           batched ensemble below) and "reference_logs" (the
                           1601^2 Zalesak and 64^3 deformation log workloads in
                           exact mode, --config zalesak) summaries of the
                           default line
  --profile-pass           launch the calibration triad (3 x 1 GiB) and the stage
                           kernels only (for rocprofv3 --pmc runs)
  --summarize-pmc F W      turn two rocprofv3 --pmc CSV directories (FETCH_SIZE,
                           WRITE_SIZE) into profiles/stage_kernel_traffic_512_<mode>.json
  --summarize-sq D...      SQ counters of the stage kernels from rocprofv3 --pmc
                           CSV directories, per wave-plane, as a markdown table
                           (--sq-kernel NAME --sq-cells N for another kernel:
                           per 64 cells of an N-cell dispatch)
  --config batch           the batched-ensemble workload instead: 96 2D
                           flames of seeded sizes in {128,192,256,384,512}^2,
                           S_L ~ U[0.05,0.2], 300 RK3 steps each, in one
                           autotuned, hipGraph-replayed batch vs a serial
                           per-instance loop (prints its own JSON line)
  --config sample          the reference's sample/ scheme instead: its 800x800
                           Bunsen-flame driver loop, 2000 steps, next to the
                           reference's own sample code on 16 CPU threads
  --config adjoint         the adjoint workload instead: flame-inversion
                           gradients (128^3, 8^3 S_L nodes + 16 mode amplitudes,
                           100 steps) for --sets seeded parameter sets (32),
                           batched over --contexts concurrent adjoint contexts
                           (8), vs central finite differences on the forward
                           solver
  --config zalesak         the reference's own logged workloads instead: the 2D
                           Zalesak rotation at 201^2..1601^2 (S_L 0, no reinit),
                           end to end, against the ms/step of its RTX 4090 and
                           32-rank MPI logs (vs_baseline: the 1601^2 4090 run),
                           plus the 3D 64^3 deformation log workload
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s (WENO5-RK3+HCR2 512³ FP64), % HBM peak; adjoint vs FD grads/s"
HBM_PEAK = 8.0e12            # MI355X HBM3E, bytes/s
STAGE_BYTES = {1: 40, 2: 48, 3: 48}   # algorithmic bytes per interior cell per stage
STEP_BYTES = 136 + 83 / 5.0           # RK3 + amortised HCR-2, bytes per cell-update
def traffic_file(mode):
    """Measured HBM bytes of the stage-2 kernel at 512^3, one file per arithmetic mode."""
    return os.path.join(ROOT, "profiles", "stage_kernel_traffic_512_%s.json" % mode)
# the 3D stage-2 kernels (shared-face z-march) the 512^3 workload launches,
# fast and exact arithmetic: variant 4 = full tiles, scalar S_L
FAST_STAGE_KERNEL = "k_stage_zface<2, false, 4>"
EXACT_STAGE_KERNEL = "k_stage_zface<2, true, 4>"


def progress(msg):
    """A progress line on stderr (long phases under a profiler stay visible)."""
    sys.stderr.write("[bench] %s\n" % msg)
    sys.stderr.flush()


def load_ls():
    import __graft_entry__ as ge
    return ge.load_package()


def setup(ls, n, mode):
    import torch
    p = ls.default_params(3, n, s_l=0.08, reinit_interval=5, reinit_iters=2,
                          mode=ls.MODE_FAST if mode == "fast" else ls.MODE_EXACT)
    G, u, v, w = ls.init_case(p, ls.CASE_PLANE_MODES3D, seed=3)
    s = ls.Solver(p)
    s.fill_ghosts(G)
    speed = (u.abs() + v.abs() + w.abs()).max().item()
    dt = 0.5 * min(p.dx, p.dy, p.dz) / (speed + p.s_l)
    torch.cuda.synchronize()
    return p, s, G, u, v, w, dt


def time_kernels(ls, s, G, u, v, w, dt, fused, reps=10):
    """Average device time of one stage-2 kernel, from HIP events on the launch
    stream.  Where ls_rk3_stage also launches a ghost fill (fused=False: a
    linear-extrapolation axis), the fill is timed separately and subtracted;
    the 3D z-march kernels of both modes write their ghosts themselves."""
    import torch
    st = torch.cuda.current_stream()
    a, b = s.empty(), s.empty()
    s.rk3_stage(1, G, G, a, u, v, w, dt=dt)
    torch.cuda.synchronize()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record(st)
    for _ in range(reps):
        s.rk3_stage(2, a, G, b, u, v, w, dt=dt)
    e1.record(st)
    for _ in range(reps):
        s.fill_ghosts(b)
    e2.record(st)
    torch.cuda.synchronize()
    t_stage_fill = e0.elapsed_time(e1) / reps * 1e-3
    t_fill = e1.elapsed_time(e2) / reps * 1e-3
    del a, b
    return (t_stage_fill if fused else t_stage_fill - t_fill), t_fill


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _slab(ls, p, fields, nz):
    """The first nz interior z planes of padded 3D fields, as a periodic slab
    (host arrays, ghosts filled by the oracle)."""
    import numpy as np
    from oracle.pyoracle import Oracle, make_grid
    g = make_grid(3, p.nx, p.ny, nz, bc=(0, 0, 0), s_l=p.s_l)
    g.dx, g.dy, g.dz = p.dx, p.dy, p.dz
    o = Oracle(fma=True)
    out = []
    for t in fields:
        a = np.zeros(g.shape)
        a[3:3 + nz] = t[3:3 + nz].cpu().numpy()
        out.append(o.fill_ghosts(g, a))
    return g, out


def cpu_baseline(ls, n):
    """The CPU baseline: this repository's C restatement of the reference's
    CPU solver (oracle/liblsoracle_fma.so: the reference's flags, -O3 with FMA
    contraction, OpenMP over cells instead of MPI y-slabs) timed on this host's
    cores over the reference's timed region (lso_timed_loop: fixed steps,
    HCR-2 trigger, in-loop L2/area every 10 (2D) or 50 (3D) steps), at P = 1
    and P = all cores, on the seeded inputs the GPU runs (generated on the
    GPU, copied to the host).  Bounded samples: the 2D 256^2 circle in full,
    and z slabs of the 256^3 and 512^3 workloads (6 steps, one HCR-2).  Where
    the reference's own 3D solver was built here from its sources
    (oracle/_ref/ref_bench3d), its mpirun time on a 512x512x32 slab is added as
    a cross-check row."""
    import numpy as np
    import torch
    from oracle.pyoracle import Oracle, make_grid
    cores = max(1, min(16, os.cpu_count() or 1))
    o = Oracle(fma=True)
    rows = []

    def run(name, g, G, u, v, w, dt, steps, ri, di, threads, sample):
        sec, _, _, _ = o.timed_loop(g, G, G, u, v, w, dt=dt, nsteps=steps, reinit_interval=ri,
                                    diag_interval=di, threads=threads)
        cells = g.nx * g.ny * (g.nz if g.dim == 3 else 1)
        rows.append({"config": name, "threads": threads, "sample": sample, "steps": steps,
                     "seconds": sec, "cell_updates_per_s": cells * steps / sec})

    # C1: 2D 256^2 circle, S_L 0.1, CFL 0.5, zero-gradient edges, HCR-2 every 10, 200 steps
    p1 = ls.default_params(2, 256, s_l=0.1, reinit_interval=10, bc=(1, 1, 0))
    G, u, v, _ = ls.init_case(p1, ls.CASE_CIRCLE2D)
    g1 = make_grid(2, 256, 256, bc=(1, 1, 0), s_l=0.1)
    G1 = o.fill_ghosts(g1, G.cpu().numpy())
    u1, v1 = u.cpu().numpy(), v.cpu().numpy()
    dt1 = ls.cfl_dt(p1, 0.5, 0.5, 0.0, 0.0)
    for th in (1, cores):
        run("C1 2D 256^2 circle", g1, G1, u1, v1, None, dt1, 200, 10, 10, th, "full (200 steps)")
    # C2 / C3: 3D swirl 256^3 and plane+modes 512^3, periodic, HCR-2 every 5
    for name, cid, N, s_l, nz1, nzp in (("C2 3D 256^3 sphere in swirl", ls.CASE_SPHERE_SWIRL3D, 256, 0.05, 32, 256),
                                         ("C3 3D 512^3 plane+modes", ls.CASE_PLANE_MODES3D, 512, 0.08, 8, 64)):
        p = ls.default_params(3, N, s_l=s_l, reinit_interval=5)
        fields = ls.init_case(p, cid, seed=3)
        speed = (fields[1].abs() + fields[2].abs() + fields[3].abs()).max().item()
        dt = 0.5 * p.dx / (speed + s_l)
        for th, nz in ((1, nz1), (cores, nzp)):
            g, (Gs, us, vs, ws) = _slab(ls, p, fields, nz)
            run(name, g, Gs, us, vs, ws, dt, 6, 5, 50, th, "%dx%dx%d z slab, 6 steps" % (N, N, nz))
            del Gs, us, vs, ws
        del fields
        torch.cuda.empty_cache()
    out = {"value": rows[-1]["cell_updates_per_s"], "unit": "cell-updates/s", "cores": cores, "kind": "port",
           "sample": "C restatement of the reference CPU solver (oracle/liblsoracle_fma.so, -O3 FMA, OpenMP, "
                     "%d threads) on a 512x512x64 z slab of the 512^3 plane+modes workload, 6 RK3 steps with one "
                     "HCR-2, the reference's timed region (in-loop diagnostics, no output)" % cores,
           "cpu": "%s, %d cores used" % (cpu_model(), cores), "rows": rows}
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench3d")
    mpirun = shutil.which("mpirun") or ("/opt/conda/bin/mpirun" if os.path.exists("/opt/conda/bin/mpirun") else None)
    if os.path.exists(exe) and mpirun:
        try:
            r = subprocess.run([mpirun, "-np", str(cores), exe, "512", "512", "32", "10", "5"],
                               stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
            d = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
            out["reference_crosscheck"] = {
                "value": d["cell_updates_per_s"], "unit": "cell-updates/s", "cores": cores, "kind": "reference",
                "sample": "the reference's level-set_MPI_3D solver built here from its sources (-O3 FMA) under "
                          "mpirun -np %d, max-over-ranks MPI_Wtime, 512x512x32 periodic slab, its own fixed "
                          "4-mode plane input, 10 RK3 steps, HCR-2 every 5, %.2f s" % (cores, d["seconds"])}
        except Exception as e:  # pragma: no cover - reported, the port rows stand
            sys.stderr.write("reference cross-check failed: %s\n" % e)
    return out


def ensemble(ls, count=96, seed=4, mode="fast"):
    """The batched workload (the package's ensemble builder)."""
    return ls.ensemble(count, seed, ls.MODE_FAST if mode == "fast" else ls.MODE_EXACT)


def bench_batch(ls, steps, mode, emit=True):
    import torch
    members, params, fields = ensemble(ls, mode=mode)
    progress("batch: %d members built" % len(members))
    b = ls.Batch(members, mode=ls.MODE_FAST if mode == "fast" else ls.MODE_EXACT)
    G, U, V = b.empty(), b.empty(), b.empty()
    for q, (g, u, v) in enumerate(fields):
        b.view(G, q).copy_(g)
        b.view(U, q).copy_(u)
        b.view(V, q).copy_(v)
    b.fill_ghosts(G)
    G0 = G.clone()
    shape, probe_ms = b.autotune(G, U, V, probe_steps=10)
    b.advance(G, U, V, steps, graph=True)  # capture + first replay (warm-up)
    torch.cuda.synchronize()
    G.copy_(G0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.advance(G, U, V, steps, graph=True)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    progress("batch: timed run done")
    cells = sum(m["nx"] * m["ny"] for m in members)
    # serial per-instance loop over the single-grid solver
    solvers = [ls.Solver(p) for p in params]
    singles = []
    for q, (g, u, v) in enumerate(fields):
        gg = g.clone()
        solvers[q].fill_ghosts(gg)
        singles.append(gg)
    solvers[0].advance(singles[0].clone(), fields[0][1], fields[0][2], dt=members[0]["dt"], t_final=1e30,
                       max_steps=2)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for q in range(len(members)):
        solvers[q].advance(singles[q], fields[q][1], fields[q][2], dt=members[q]["dt"], t_final=1e30,
                           max_steps=steps)
    torch.cuda.synchronize()
    t_serial = time.perf_counter() - t0
    same = all(torch.equal(b.view(G, q), singles[q]) for q in range(len(members)))
    out = {"metric": "batched ensemble cell-updates/s (96 uneven 2D flames, 300 RK3 steps)",
           "value": cells * steps / t_batch, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "96 2D flames, sizes in {128,192,256,384,512}^2 (seed 4), S_L~U[0.05,0.2], "
                                  "circle/plane G0, uniform+8-mode velocity, zero-gradient edges",
                      "members": len(members), "cells": cells, "mode": mode, "tile_shape": shape,
                      "autotune_probe_ms_per_step": probe_ms},
           "batched_s": t_batch, "serial_loop_s": t_serial, "speedup_vs_serial": t_serial / t_batch,
           "serial_cell_updates_per_s": cells * steps / t_serial, "bitwise_equal_to_serial": bool(same)}
    if emit:
        print(json.dumps(out), flush=True)
    return out


def bench_adjoint(ls, mode, nsets, fd_sample=16, emit=True, nctx=8, serial_sets=4):
    """The adjoint workload: parameter gradients of the flame-inversion
    loss for `nsets` seeded parameter sets, batched over `nctx` concurrent
    adjoint contexts (FlameBatch: ls_adjoint_run_sets), vs central finite
    differences on the plain forward solver (2 forward solves per parameter
    per gradient).  The first `serial_sets` sets also run one by one, for the
    batching speedup and a bitwise check of the batch against them."""
    import numpy as np
    import torch
    m = ls.MODE_FAST if mode == "fast" else ls.MODE_EXACT
    prob = ls.FlameInversion(n=128, nc=8, nmodes=16, nsteps=100, reinit_interval=5, ckpt=10, mode=m)
    sets = [prob.draw(100 + b) for b in range(nsets)]
    nctx = max(1, min(nctx, nsets))
    fb = ls.FlameBatch(prob, nctx)
    fb.loss_and_grad(sets[:nctx])  # warm-up of every context
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grads = fb.loss_and_grad(sets)
    torch.cuda.synchronize()
    t_adj = (time.perf_counter() - t0) / nsets
    # the same sets one by one through the single context
    ns = max(1, min(serial_sets, nsets))
    prob.loss_and_grad(*sets[0])  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    serial = [prob.loss_and_grad(*sets[q]) for q in range(ns)]
    torch.cuda.synchronize()
    t_serial = (time.perf_counter() - t0) / ns
    same = all(serial[q][0] == grads[q][0] and torch.equal(serial[q][1], grads[q][1]) and
               torch.equal(serial[q][2], grads[q][2]) for q in range(ns))
    fb.close()
    # one plain forward solve + loss: the unit of finite differences
    reps = 8
    prob.loss_plain(*sets[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(reps):
        prob.loss_plain(*sets[r % nsets])
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / reps
    npar = prob.nc ** 3 + prob.nmodes
    t_fd = 2 * npar * t_fwd
    # agreement on a seeded sample of parameters of set 0 (h = 1e-6)
    rng = np.random.default_rng(0)
    n_s = fd_sample - 4
    idx_s = sorted(rng.choice(prob.nc ** 3, n_s, replace=False).tolist())
    idx_a = [0, 5, 10, 15]
    sl_c, amps = sets[0]
    h = 1e-6
    fd = []
    for q in idx_s:
        sp, sm = sl_c.clone(), sl_c.clone()
        sp[q] += h
        sm[q] -= h
        fd.append((prob.loss_plain(sp, amps) - prob.loss_plain(sm, amps)) / (2 * h))
    for q in idx_a:
        ap, am = amps.clone(), amps.clone()
        ap[q] += h
        am[q] -= h
        fd.append((prob.loss_plain(sl_c, ap) - prob.loss_plain(sl_c, am)) / (2 * h))
    fd = np.array(fd)
    adj = np.concatenate([grads[0][1].cpu().numpy()[idx_s], grads[0][2].cpu().numpy()[idx_a]])
    err = float(np.abs(adj - fd).max() / np.abs(fd).max())
    out = {"metric": "adjoint vs FD grads/s (128³ sphere flame, 8³ S_L + 16 velocity modes, 100 RK3 steps)",
           "value": 1.0 / t_adj, "unit": "gradients/s", "n_gpus": 1, "steps": prob.nsteps,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "3D 128^3 sphere flame r=0.2 in swirl + 16 seeded solenoidal modes, S_L on "
                                  "8^3 coarse nodes (trilinear), loss sum (H_e(G_T)-H_e(G_target))^2, "
                                  "e=1.5dx, HCR-2 every 5 steps, checkpoints every 10 steps",
                      "parameter_sets": nsets, "parameters": npar, "mode": mode, "dt": prob.dt,
                      "concurrent_contexts": nctx},
           "adjoint_s_per_gradient": t_adj, "forward_solve_s": t_fwd,
           "serial_s_per_gradient": t_serial, "batch_speedup_vs_serial": t_serial / t_adj,
           "batch_bitwise_equal_to_serial": bool(same), "serial_sets_checked": ns,
           "fd_s_per_gradient": t_fd, "fd_gradients_per_s": 1.0 / t_fd, "speedup_vs_fd": t_fd / t_adj,
           "fd_sample": {"parameters": len(fd), "h": h, "max_rel_err": err},
           "adjoint_workspace_bytes": prob.adj.workspace_bytes(),
           "losses": [g[0] for g in grads[:4]]}
    prob.close()
    if emit:
        print(json.dumps(out), flush=True)
    return out


# The reference's published run logs for the 2D Zalesak rotation (S_L = 0, no
# reinitialisation, one full rotation): steps, final L2 error and ms per step
# on an RTX 4090 (Local_code/level-set_GPU_2D/log/*.log) and on 32 MPI ranks
# (Local_code/level-set_MPI_2D/log/*.log).
ZALESAK_LOGS = {201: (6284, "8.576633e-02", 0.295, 0.389), 401: (12567, "8.373019e-02", 0.722, 1.263),
                801: (25133, "8.250949e-02", 2.381, 4.631), 1601: (50266, "8.243448e-02", 9.510, 18.166)}


def bench_zalesak(ls, mode, sizes=None, emit=True):
    """The reference's own logged workloads, timed end to end against the
    ms/step its logs publish: the 2D Zalesak rotation (one full rotation, CFL
    0.2) and the 3D 64^3 deformation, with L2 and area every 50 steps."""
    import torch
    T = 6.283185
    rows = []
    for N, (steps_ref, l2_ref, ms_4090, ms_mpi) in sorted(ZALESAK_LOGS.items()):
        if sizes is not None and N not in sizes:
            continue
        p = ls.default_params(2, N, reinit_interval=0, mode=ls.MODE_EXACT if mode == "exact" else ls.MODE_FAST)
        dt = ls.cfl_dt(p, 0.2, 1.0, 0.0, 0.0)
        G, u, v, _ = ls.init_case(p, ls.CASE_ZALESAK2D, period=T)
        s = ls.Solver(p)
        s.fill_ghosts(G)
        G0 = G.clone()
        s.advance(G.clone(), u, v, dt=dt, t_final=50 * dt, max_steps=50)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t, n = 0.0, 0
        while n < steps_ref + 5:
            k, t, _ = s.advance(G, u, v, dt=dt, t_final=T, step0=n, max_steps=50, t0=t)
            n += k
            l2, area = s.diag(G, G0)
            if k < 50 or t >= T:
                break
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = wall / n * 1e3
        rows.append({"N": N, "steps": n, "log_steps": steps_ref, "ms_per_step": ms,
                     "cell_updates_per_s": N * N / (ms * 1e-3), "final_l2": "%.6e" % l2, "log_l2": l2_ref,
                     "reference_rtx4090_ms_per_step": ms_4090, "speedup_vs_rtx4090": ms_4090 / ms,
                     "reference_mpi32_ms_per_step": ms_mpi, "speedup_vs_mpi32": ms_mpi / ms})
        s.close()
    # the 3D log workload: sphere in the time-reversed deformation field, 64^3,
    # dt 0.001 to T = 1.5 (1501 steps, a sliver last step), velocity recomputed
    # every step (Local_code/level-set_MPI_3D/log/simulation.log: 5.990 ms/step
    # on 8 MPI ranks; level-set_GPU_3D: 10.653; level-set_MPI_GPU_3D: 14.973)
    p = ls.default_params(3, 64, reinit_interval=0, velocity=ls.VEL_DEFORM3D, vel_period=1.5,
                          mode=ls.MODE_EXACT if mode == "exact" else ls.MODE_FAST)
    G, u, v, w = ls.init_case(p, ls.CASE_SPHERE_DEFORM3D, period=1.5)
    s = ls.Solver(p)
    s.fill_ghosts(G)
    G0 = G.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, n = 0.0, 0
    while True:
        k, t, _ = s.advance(G, u, v, w, dt=0.001, t_final=1.5, step0=n, max_steps=50, t0=t)
        n += k
        l2, vol = s.diag(G, G0)
        if k < 50 or t >= 1.5:
            break
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / n * 1e3
    deform = {"N": 64, "steps": n, "log_steps": 1501, "ms_per_step": ms, "final_l2": "%.6e" % l2,
              "log_l2": "9.158634e-04", "reference_mpi8_ms_per_step": 5.990, "speedup_vs_mpi8": 5.990 / ms,
              "reference_gpu_ms_per_step": 10.653, "speedup_vs_gpu": 10.653 / ms}
    s.close()
    big = rows[-1]
    out = {"metric": "2D Zalesak rotation cell-updates/s (WENO5-RK3, S_L 0, no reinit, FP64) vs the reference's "
                     "RTX 4090 log",
           "value": big["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1, "steps": big["steps"],
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "vs_baseline": big["cell_updates_per_s"] / (1601 * 1601 / (9.510e-3)),
           "config": {"workload": "Zalesak slotted disk, one rotation (T = 6.283185), CFL 0.2, periodic x / "
                                  "zero-gradient y, L2 and area every 50 steps", "grid": [1601, 1601], "mode": mode},
           "rows": rows, "deformation3d": deform}
    if emit:
        print(json.dumps(out))
    return out


def bench_sample(ls, steps, n=800):
    """The sample/ scheme: the reference's Bunsen-flame driver loop (800 x 800,
    dt 2.5e-6, S_L 0.4, forced velocity at 10 Hz) on the GPU, with the
    reference's own sample code (OpenMP, 16 threads) timed beside it."""
    import torch
    s = ls.Sample(n, n)
    progress("sample: %dx%d set up" % (n, n))
    G0 = s.bunsen()
    G = G0.clone()
    u, v = s.empty(), s.empty()
    t = s.advance(G, u, v, G0, 20, freq=10)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t = s.advance(G, u, v, G0, steps, freq=10, t0=t)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    progress("sample: timed loop done")
    # RK3 kernels alone (3 stage launches), HIP events on the launch stream
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 50
    e0.record(st)
    for _ in range(reps):
        s.rk3_step(G, u, v, 2.5e-6, 0.4, G0)
    e1.record(st)
    torch.cuda.synchronize()
    t_step = e0.elapsed_time(e1) * 1e-3 / reps
    cells = n * n
    bytes_step = 112  # stage 1: G,u,v -> G1 32 B; stages 2, 3: 40 B each (algorithmic, per cell)
    ach = bytes_step * cells / t_step
    out = {"metric": "sample/ scheme cell-updates/s (LLF WENO5-RK3, 800² Bunsen flame, FP64)",
           "value": cells * steps / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "sample/main.cpp Bunsen flame %dx%d on 0.02 m, dt 2.5e-6, S_L 0.4, "
                                  "forced velocity 10 Hz, anchored+periodic boundary" % (n, n), "grid": [n, n]},
           "ms_per_step": wall / steps * 1e3,
           "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": ach / HBM_PEAK, "traffic": None,
                        "kernel": "k_sample_ymarch<1..3> (3 launches per RK3 step)",
                        "rk3_kernels_ms": t_step * 1e3, "algorithmic_bytes_per_cell_step": bytes_step}}
    exe = os.path.join(ROOT, "oracle", "_ref", "sample_bench")
    if os.path.exists(exe):
        env = dict(os.environ, OMP_NUM_THREADS="16")
        cpu_steps = 40
        r = subprocess.run([exe, str(n), str(cpu_steps)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, env=env, timeout=600)
        if r.returncode == 0:
            j = json.loads(r.stdout.strip().splitlines()[-1])
            out["cpu_baseline"] = {"value": j["cell_updates_per_s"], "unit": "cell-updates/s",
                                   "cores": j["threads"], "kind": "reference",
                                   "sample": "the reference's sample/ RK3_step loop (its own sources and flags, "
                                             "OpenMP) on the same %dx%d case, %d steps" % (n, n, cpu_steps)}
    print(json.dumps(out), flush=True)
    s.close()


def read_traffic(n, mode):
    path = traffic_file(mode)
    if n != 512 or not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        if d.get("mode") != mode:
            return None
        return d.get("hbm_bytes_per_launch")
    except Exception:
        return None


def calib_lib():
    """liblscalib.so: the triad / copy calibration kernels (built by build())."""
    import ctypes as C
    L = C.CDLL(os.path.join(ROOT, "level-set-method-code_amd", "liblscalib.so"))
    L.lsc_triad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int64, C.c_void_p]
    L.lsc_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    return L


CALIB_N = 1 << 27  # doubles per array: 1 GiB each


def triad_bandwidth(reps=10):
    """Achievable HBM rate: the calibration triad c = a + s*b over three 1 GiB
    arrays (2 GiB read, 1 GiB written per launch), best of `reps` launches timed
    with HIP events on the launch stream.  Returns bytes/s."""
    import ctypes as C
    import torch
    L = calib_lib()
    a = torch.ones(CALIB_N, dtype=torch.float64, device="cuda")
    b = torch.full_like(a, 0.5)
    c = torch.empty_like(a)
    st = torch.cuda.current_stream()
    args = (C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), C.c_void_p(c.data_ptr()), 2.0, CALIB_N,
            C.c_void_p(st.cuda_stream))
    for _ in range(3):
        assert L.lsc_triad(*args) == 0
    torch.cuda.synchronize()
    best = 1e30
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        assert L.lsc_triad(*args) == 0
        e1.record(st)
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    assert float(c[12345]) == 2.0
    del a, b, c
    torch.cuda.empty_cache()
    return 24.0 * CALIB_N / best


def profile_pass(ls, n, mode):
    """Kernels for the --pmc runs: the calibration triad over 1 GiB arrays (the
    kernel of known traffic: 2 GiB read + 1 GiB written), then one stage-1 and
    three stage-2 launches and one HCR-2 invocation (the two-iteration pair)."""
    import ctypes as C
    import torch
    p, s, G, u, v, w, dt = setup(ls, n, mode)
    L = calib_lib()
    a = torch.ones(CALIB_N, dtype=torch.float64, device="cuda")
    b = torch.ones_like(a)
    c = torch.empty_like(a)
    torch.cuda.synchronize()
    assert L.lsc_triad(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), C.c_void_p(c.data_ptr()), 2.0,
                       CALIB_N, C.c_void_p(torch.cuda.current_stream().cuda_stream)) == 0
    torch.cuda.synchronize()
    del a, b, c
    x, y = s.empty(), s.empty()
    s.rk3_stage(1, G, G, x, u, v, w, dt=dt)
    for _ in range(3):
        s.rk3_stage(2, x, G, y, u, v, w, dt=dt)
    s.reinit(y)  # the HCR-2 pair kernel
    torch.cuda.synchronize()
    print("profile pass done", flush=True)


def profile_pass_2d(ls, config, mode):
    """Kernels for --pmc runs of the 2D workloads, a few launches each (every
    dispatch is serialised and slow under counter collection): one RK3 step of
    the sample/ scheme on its 800^2 Bunsen case, or two RK3 steps of the
    96-member ensemble on its 8-row y-march chunks (schedule shape 3)."""
    import torch
    if config == "sample":
        s = ls.Sample(800, 800)
        G0 = s.bunsen()
        G = G0.clone()
        u, v = s.empty(), s.empty()
        s.forced_velocity(0.0, 10, u, v)
        s.rk3_step(G, u, v, 2.5e-6, 0.4, G0)
    else:
        members, params, fields = ensemble(ls, mode=mode)
        b = ls.Batch(members, mode=ls.MODE_FAST if mode == "fast" else ls.MODE_EXACT)
        G, U, V = b.empty(), b.empty(), b.empty()
        for q, (g, u, v) in enumerate(fields):
            b.view(G, q).copy_(g)
            b.view(U, q).copy_(u)
            b.view(V, q).copy_(v)
        b.fill_ghosts(G)
        b.set_tile(3)
        b.advance(G, U, V, 2, graph=False)
    torch.cuda.synchronize()
    print("profile pass done", flush=True)


def _pmc_rows(d):
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def summarize_pmc(fetch_dir, write_dir, n, mode, out_path=None):
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) for the stage-2 kernel, each
    scaled by what the same counter reads for the calibration triad (2 GiB
    read, 1 GiB written)."""
    def per_kernel(rows, counter):
        out = {}
        for r in rows:
            if r.get("Counter_Name") != counter:
                continue
            out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return out
    fr = per_kernel(_pmc_rows(fetch_dir), "FETCH_SIZE")
    wr = per_kernel(_pmc_rows(write_dir), "WRITE_SIZE")
    copy_k = [k for k in fr if "k_triad" in k]
    stage_k = [k for k in fr if ((mode == "fast" and FAST_STAGE_KERNEL in k) or
                                 (mode == "exact" and EXACT_STAGE_KERNEL in k))]
    if not copy_k or not stage_k:
        raise SystemExit("kernels not found in pmc output: %s" % list(fr))
    ck, sk = copy_k[0], stage_k[0]
    gib = float(1 << 30)
    f_scale = 2.0 * gib / (fr[ck][0] * 1024.0)
    w_scale = gib / (wr[ck][0] * 1024.0)
    fetch = sum(fr[sk]) / len(fr[sk]) * 1024.0 * f_scale
    write = sum(wr[sk]) / len(wr[sk]) * 1024.0 * w_scale
    cells = n ** 3
    d = {"kernel": sk, "mode": mode, "grid": [n, n, n],
         "fetch_kib_raw": sum(fr[sk]) / len(fr[sk]), "write_kib_raw": sum(wr[sk]) / len(wr[sk]),
         "calib_kernel": ck, "calib_fetch_kib_raw": fr[ck][0], "calib_write_kib_raw": wr[ck][0],
         "fetch_scale": f_scale, "write_scale": w_scale,
         "hbm_bytes_per_launch": fetch + write,
         "hbm_bytes_per_cell": (fetch + write) / cells,
         "algorithmic_bytes_per_cell": STAGE_BYTES[2]}
    out_path = out_path or traffic_file(mode)
    os.makedirs(os.path.dirname(os.path.abspath(out_path)), exist_ok=True)
    with open(out_path, "w") as fh:
        json.dump(d, fh, indent=1)
    print(json.dumps(d))


def summarize_sq(dirs, n, kernel="k_stage_zface", cells=None):
    """SQ counters (rocprofv3 --pmc CSV directories) of the kernels whose name
    contains `kernel`, as a markdown table: each counter averaged over the
    kernel's dispatches and divided by the 64-cell units of one dispatch
    (`cells` / 64; by default the n^3 cells of the 3D stage launch, i.e. one
    64-wide wavefront row per plane)."""
    vals = {}
    for d in dirs:
        for r in _pmc_rows(d):
            k = r["Kernel_Name"].replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").strip()
            if kernel not in k:
                continue
            vals.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    wp = (cells if cells else n ** 3) / 64.0
    names = sorted({c for v in vals.values() for c in v})
    print("| counter, per 64 cells | " + " | ".join(sorted(vals)) + " |")
    print("|---|" + "---|" * len(vals))
    for c in names:
        row = []
        for k in sorted(vals):
            x = vals[k].get(c)
            row.append("%.2f" % (sum(x) / len(x) / wp) if x else "")
        print("| %s | %s |" % (c, " | ".join(row)))


def timed_steps(ls, n, mode, steps, warmup):
    """Whole-job rate of `steps` solver steps on the n^3 workload, plus the stage kernel's roofline."""
    import torch
    p, s, G, u, v, w, dt = setup(ls, n, mode)
    cells = p.cells
    # warmup (untimed); steps counted from 1 so that HCR-2 runs exactly once per 5 steps
    s.advance(G, u, v, w, dt=dt, t_final=1e30, step0=1, max_steps=warmup)
    torch.cuda.synchronize()
    step0 = 1 + warmup
    st = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(st)
    n_done, _, _ = s.advance(G, u, v, w, dt=dt, t_final=1e30, step0=step0, max_steps=steps)
    ev1.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev = ev0.elapsed_time(ev1) * 1e-3
    assert n_done == steps
    elapsed = max(wall, dev)
    value = cells * steps / elapsed
    ok = s.check_finite(G) == 0

    t_stage, t_fill = time_kernels(ls, s, G, u, v, w, dt, fused=True)  # periodic: ghosts written in-kernel
    ach = STAGE_BYTES[2] * cells / t_stage
    traffic = read_traffic(n, mode)
    roof = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": ach / HBM_PEAK, "traffic": traffic,
            "kernel": (FAST_STAGE_KERNEL if mode == "fast" else EXACT_STAGE_KERNEL) +
                      " (fused WENO5/u_eff RHS + stage-2 combine)",
            "kernel_ms": t_stage * 1e3, "ghost_fill_ms": t_fill * 1e3,
            "algorithmic_bytes_per_cell": STAGE_BYTES[2],
            "step_frac": value * STEP_BYTES / HBM_PEAK}
    del s, G, u, v, w
    torch.cuda.empty_cache()
    return value, elapsed, roof, dt, ok


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mode", choices=["exact", "fast"], default="fast")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-pass", action="store_true")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
    ap.add_argument("--traffic-out", default=None)
    ap.add_argument("--summarize-sq", nargs="+", metavar="PMC_DIR")
    ap.add_argument("--sq-kernel", default="k_stage_zface",
                    help="with --summarize-sq: kernel name substring to tabulate")
    ap.add_argument("--sq-cells", type=int, default=None,
                    help="with --summarize-sq: cells per dispatch (default n^3)")
    ap.add_argument("--config", choices=["plane_modes", "batch", "adjoint", "sample", "zalesak"],
                    default="plane_modes")
    ap.add_argument("--sets", type=int, default=32)
    ap.add_argument("--contexts", type=int, default=8)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other-mode and batched-ensemble summaries on the default line")
    a = ap.parse_args()

    if a.summarize_pmc:
        summarize_pmc(a.summarize_pmc[0], a.summarize_pmc[1], a.n, a.mode, a.traffic_out)
        return
    if a.summarize_sq:
        summarize_sq(a.summarize_sq, a.n, a.sq_kernel, a.sq_cells)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus > 1 or world > 1:
        print("bench.py measures one GPU: the solver has no multi-GPU path (it would shard by "
              "y-slabs with a 3-layer halo exchange per stage)", file=sys.stderr)
        sys.exit(2)

    import torch
    ls = load_ls()
    torch.cuda.set_device(0)
    if a.profile_pass:
        if a.config in ("sample", "batch"):
            profile_pass_2d(ls, a.config, a.mode)
        else:
            profile_pass(ls, a.n, a.mode)
        return
    if a.config == "batch":
        bench_batch(ls, 300 if a.steps == 20 else a.steps, a.mode)
        return
    if a.config == "adjoint":
        bench_adjoint(ls, a.mode, a.sets, nctx=a.contexts)
        return
    if a.config == "zalesak":
        bench_zalesak(ls, a.mode)
        return
    if a.config == "sample":
        bench_sample(ls, 2000 if a.steps == 20 else a.steps)
        return

    value, elapsed, roof, dt, ok = timed_steps(ls, a.n, a.mode, a.steps, a.warmup)
    # the achievable HBM rate on this part: the calibration triad (liblscalib.so)
    bw = triad_bandwidth()
    roof["achievable_peak"] = bw / 1e9
    roof["achievable_kernel"] = "k_triad (c = a + s*b, 3 x 1 GiB FP64, 16-byte accesses)"
    roof["frac_of_achievable"] = roof["achieved"] * 1e9 / bw
    roof["step_achieved"] = value * STEP_BYTES / 1e9
    roof["step_frac_of_achievable"] = value * STEP_BYTES / bw
    out = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": elapsed / a.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "3D plane+modes flame, %d^3 FP64, S_L=0.08, CFL 0.5, WENO5-RK3, "
                                  "HCR-2 every 5 steps (2 iterations), periodic" % a.n,
                      "grid": [a.n, a.n, a.n], "mode": a.mode, "dt": dt,
                      "parallelism": "single GPU", "finite": ok},
           "roofline": roof}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ls, a.n)
    if not a.no_adjoint:
        # second half of the metric: adjoint vs finite-difference gradients (reduced batch)
        adj = bench_adjoint(ls, a.mode, nsets=16, fd_sample=8, emit=False, nctx=8, serial_sets=2)
        out["adjoint"] = {k: adj[k] for k in ("value", "unit", "adjoint_s_per_gradient", "forward_solve_s",
                                               "fd_s_per_gradient", "speedup_vs_fd", "fd_sample",
                                               "serial_s_per_gradient", "batch_speedup_vs_serial",
                                               "batch_bitwise_equal_to_serial")}
        out["adjoint"]["workload"] = adj["config"]["workload"]
        out["adjoint"]["parameter_sets"] = 16
        out["adjoint"]["concurrent_contexts"] = adj["config"]["concurrent_contexts"]
    if not a.no_extra:
        # the other arithmetic mode on the same workload (short run), and the batched 2D ensemble
        other = "exact" if a.mode == "fast" else "fast"
        ov, oe, oroof, _, ook = timed_steps(ls, a.n, other, 10, 5)
        out["other_mode"] = {"mode": other, "value": ov, "unit": "cell-updates/s", "steps": 10,
                             "ms_per_step": oe / 10 * 1e3, "kernel": oroof["kernel"],
                             "kernel_ms": oroof["kernel_ms"], "frac": oroof["frac"],
                             "traffic": oroof["traffic"], "finite": ook}
        bt = bench_batch(ls, 300, a.mode, emit=False)
        out["batch"] = {k: bt[k] for k in ("value", "unit", "steps", "batched_s", "serial_loop_s",
                                           "speedup_vs_serial", "bitwise_equal_to_serial")}
        out["batch"]["workload"] = bt["config"]["workload"]
        out["batch"]["tile_shape"] = bt["config"]["tile_shape"]
        # the reference's own logged workloads (the only published timings), exact mode
        rl = bench_zalesak(ls, "exact", sizes=(1601,), emit=False)
        z, dfm = rl["rows"][-1], rl["deformation3d"]
        out["reference_logs"] = {
            "zalesak2d_1601": {k: z[k] for k in ("steps", "ms_per_step", "final_l2", "log_l2",
                                                 "reference_rtx4090_ms_per_step", "speedup_vs_rtx4090")},
            "deformation3d_64": {k: dfm[k] for k in ("steps", "ms_per_step", "final_l2", "log_l2",
                                                     "reference_gpu_ms_per_step", "speedup_vs_gpu")},
            "mode": "exact"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
