#!/usr/bin/env python3
"""Throughput of the monodomain time step on one MI355X.

Workload: ten Tusscher 2006 on a synthetic truncated-ellipsoid LV shell
(1,029,600 nodes, 5,904,000 P1 tets) driven by a seeded Purkinje tree (4,351
bar nodes, 256 PMJs, Corrias Purkinje cells), endo/M/epi layers (30/40/30 %)
and an apicobasal G_Ks gradient, fibers +60 -> -60 degrees, FP32 ionic + SpMV
with FP64 CG dots, Rush-Larsen ionic integrator, dt = 0.02 ms, CG stopping at
||r|| <= 1e-7 ||b - A V^n|| (the reference's r0, EP_CrankNicholson.cc:104),
each solve started from the last increment (--cg-x0) with the Purkinje rows
of that guess solved locally first, activation times and a 10-electrode
pseudo-ECG recorded every 1 ms.  A step is one EP_CrankNicholson::Solve over
all nodes; the metric is node-steps per second.  The CG iterations per step
follow the activity (about 15 while the tree conducts, 27-29 in the plateau
that holds most of the beat), so the run always covers the 400 ms beat: the
autotuner runs in the first steps, the state is advanced untimed to the
window's start, then W warm-up and exactly K timed steps follow, and the rest
of the beat runs after them.  The window starts at 150 ms (mid-plateau)
unless W + K spans the rest of the beat (the default: 100 + 19,800 steps
from 2 ms).  Beside the K-step value the line reports the whole beat
(`whole_beat`: every step outside the autotuner and the 50 profiled steps,
wall-clocked) and the per-phase table.

Prints ONE JSON line (rank 0).  --gpus N with N > 1 is refused: this build
measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BEAT_MS = 400.0
# Every 200th step's two solves are followed by an FP64 true residual
# ||b - A x||_D / ||b||_D (two launches, ~0.3 us per step on average); the
# largest goes into the JSON line beside the CG tolerance.
TRUE_RESIDUAL_EVERY = 200


# Multiply-adds per node of the learned step as the MFMA tiles run it: layer 0
# 64 x 128 (25 inputs as BF16 his and los, padded), three 128 x 128, 128 x 32.
SURR_MACS = 64 * 128 + 3 * 128 * 128 + 128 * 32


def lv_workload(scale: float = 1.0, purkinje: str = "high"):
    from ucla_cmg_amd import meshgen
    return meshgen.lv_workload(scale, purkinje)


def his_stimulus(mesh, n=3):
    """The Purkinje root and the next nodes of the first branch."""
    from ucla_cmg_amd import meshgen
    return meshgen.his_nodes(mesh, n)


def endo_apex_stimulus(mesh, radius=0.6):
    depth = mesh.meta["depth"]
    z = mesh.xyz[:, 2]
    apex_z = z.min()
    sel = (depth == 0.0) & (z < apex_z + radius)
    return np.nonzero(sel)[0].astype(np.int32)


def electrodes10():
    from ucla_cmg_amd import meshgen
    return meshgen.lv_electrodes()


MODEL_KW = dict(D=(1.2 / 1400.0, 0.2 / 1400.0, 0.2 / 1400.0), dt=0.02, cg_tol=1e-7)


# Operator and CG-vector layout of this build (SELL-64 with entries in groups
# of four per lane, off-diagonal entries only, slices whose rows share one set
# of column offsets storing no per-row columns; CG rows as {r, w, s, p}
# records; each XCD's blocks taking its slice range in turn): a traffic
# measurement taken on another layout or work split does not describe this
# kernel.
LAYOUT = "sell64-groups4-offdiag-sharedoffsets/cg-records4/xcd-window"


def measured_traffic(nnz, stored):
    """HBM bytes per k_cg_iter launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), when it was
    measured on this operator and layout; else None."""
    path = os.path.join(ROOT, "profiles", "k_cg_iter_lv1m_purkinje_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except OSError:
        return None
    if t.get("nnz") != nnz or t.get("sell_stored_entries") != stored or t.get("layout") != LAYOUT:
        return None
    return t["hbm_read_bytes"] + t["hbm_write_bytes"]


def host_cpu():
    """nproc, the CPUs this process may run on, and the CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpus_available": len(os.sched_getaffinity(0)),
            "model": model}


def oracle_for_baseline():
    """The oracle package, on its AVX2/FMA build when this host has those
    instructions (same sources, -march=x86-64-v3); returns (module, build)."""
    fast = os.path.join(ROOT, "oracle", "liboracle_avx2.so")
    try:
        with open("/proc/cpuinfo") as f:
            flags = next((ln for ln in f if ln.startswith("flags")), "").split()
    except OSError:
        flags = []
    build = "-O3"
    if os.path.exists(fast) and "avx2" in flags and "fma" in flags and "oracle" not in sys.modules:
        os.environ["ORACLE_LIB"] = fast
        build = "-O3 -march=x86-64-v3"
    import oracle
    return oracle, build


def cpu_baseline(mesh, scales, pidx, V, ragged, k_state, seconds_all=12.0, seconds_one=5.0):
    """The oracle -- the reference's FP64 integrators (ten Tusscher backward
    Euler + Newton, Purkinje forward Euler with adaptive sub-steps) and
    Jacobi-PCG, OpenMP C -- on the same mesh, started from the GPU run's
    state at step k_state (mid-plateau): first on every CPU this process may
    use, then on one core."""
    oracle, build = oracle_for_baseline()
    host = host_cpu()
    n = mesh.n_nodes
    pk = (mesh.tissue == 11) | (mesh.tissue == 12) if mesh.tissue is not None else np.zeros(n, bool)
    states = np.tile(oracle.tt_init(), (n, 1))
    for i in np.nonzero(~pk)[0]:
        states[i] = ragged[i]
    tt_const = oracle.tt_constants_scaled(scales)

    def timed(threads, seconds, min_steps):
        sim = oracle.Simulation(mesh, oracle.MODEL_TT, dt=0.02, D=MODEL_KW["D"], cg_tol=1e-7,
                                tt_const=tt_const, param_idx=pidx, nthreads=threads,
                                states=states, V0=V)
        if sim.pk_states is not None:
            for i in np.nonzero(pk)[0]:
                sim.pk_states[i] = ragged[i]
        steps, t0 = 0, time.perf_counter()
        while True:
            sim.step()
            steps += 1
            el = time.perf_counter() - t0
            if el >= seconds and steps >= min_steps:
                break
        its = sim.cg_iters_total / max(1, 2 * steps)
        return n * steps / el, steps, el, its

    threads = host["cpus_available"]
    v_all, s_all, e_all, its = timed(threads, seconds_all, 2)
    v_one, s_one, e_one, _ = timed(1, seconds_one, 1)
    return {
        "value": v_all,
        "unit": "node-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{s_all} steps ({e_all:.1f} s) of the full {n}-node LV workload from the GPU "
                  f"run's state at {k_state * 0.02:g} ms, on {threads} threads; FP64 "
                  f"backward-Euler/Newton tT, Jacobi-PCG tol 1e-7 ({its:.1f} its/solve); oracle "
                  f"built {build}",
        "one_core": {"value": v_one, "steps": s_one, "seconds": e_one},
        "host": host,
    }


# --value-streaming: how the CG reads the operator's values (-1: the library's
# choice by operator size; 0 plain loads; 1 nontemporal loads).
VALUE_STREAMING = -1


def _streaming(solver):
    if VALUE_STREAMING != -1:
        solver.setValueStreaming(VALUE_STREAMING)
    return solver


def sweep_bench(vg, torch, steps, tune):
    """The batched sweep: 96 ten Tusscher slabs, 6.0k-236k nodes each
    (log-uniform, 6.62M in all), each with its own G_Na/G_Kr/G_CaL scaling
    and anisotropy ratio, in one handle (per-instance CG); mixed precision,
    Rush-Larsen, dt 0.02 ms; each slab paced at its corner column for 2 ms.
    Aggregate node-steps/s over ``steps`` steps after an autotuned warm-up."""
    from ucla_cmg_amd import autotune, meshgen
    t0 = time.perf_counter()
    meshes, scales, ratios = meshgen.slab_sweep()
    mesh, inst, offs = meshgen.concat(meshes)
    dl = MODEL_KW["D"][0]
    solver = vg.CardiacSolver(mesh, cell_model=vg.MODEL_TT, precision=vg.PREC_MIXED,
                              integrator=vg.INTEG_RUSH_LARSEN, dt=0.02, D=(dl, dl, dl),
                              cg_tol=1e-7, param_scales=scales, param_index=inst.astype(np.uint16),
                              instance=inst, cg_x0=vg.X0_LAST)
    _streaming(solver)
    stim = np.concatenate([meshgen.slab_corner_nodes(*m.meta["shape"], 3) + offs[k]
                           for k, m in enumerate(meshes)]).astype(np.int32)
    solver.enableActivation(-40.0)
    solver.setStimulusCurrent(stim, 50.0 * 2000.0)
    setup = time.perf_counter() - t0
    done = [0]

    def step(m):
        for _ in range(m):
            if done[0] == 100:
                solver.clearStimulus()
            solver.Solve(1)
            done[0] += 1

    step(10)
    tuned = autotune.tune(solver, steps_each=6, blocks=(0, 2048, 4096), step_fn=step) if tune else None
    step(10)
    solver.synchronize()
    st0 = solver.stats()
    t1 = time.perf_counter()
    step(steps)
    solver.synchronize()
    wall = time.perf_counter() - t1
    st1 = solver.stats()
    n = mesh.n_nodes
    out = {
        "metric": "aggregate node-steps/s, 96 batched tT slabs (per-instance CG)",
        "value": n * steps / wall,
        "unit": "node-steps/s",
        "ms_per_step": wall * 1e3 / steps,
        "steps": steps,
        "instances": len(meshes),
        "nodes": n,
        "value_streaming": solver.valueStreaming(),
        "instance_nodes_min_max": [int(min(m.n_nodes for m in meshes)),
                                   int(max(m.n_nodes for m in meshes))],
        "cg_iters_per_instance_solve": (st1["cg_iterations"] - st0["cg_iterations"]) /
                                       max(1, st1["cg_solves"] - st0["cg_solves"]),
        "cg_not_converged": st1["cg_not_converged"] - st0["cg_not_converged"],
        "cg_blocks": solver.cgBlocks(),
        "graph": tuned["graph"] if tuned else False,
        "autotune": tuned["table"] if tuned else None,
        "setup_s": setup,
    }
    solver.close()
    return out


def sweep_full(args, vg, torch):
    """The batched sweep over its configured horizon (96 ten Tusscher slabs,
    6.62M nodes, dt 0.02 ms, --sweep-ms, 500 ms by default): the product form
    (mixed precision, Rush-Larsen, CG to 1e-7, autotuned launch shape) timed
    over the whole run, then the same 96 instances with the reference's own
    integrators in FP64 (backward-Euler/Newton ten Tusscher, CG to 1e-10) --
    the oracle's method, which the batched FP64 path matches per instance to
    1e-6 mV in the GPU tests -- on every --sweep-check-every-th instance (that
    run costs ~5.5x the mixed one per node), and per checked instance: the
    largest activation-time difference, activation mismatches and the largest
    V RMS difference over frames every 50 ms.  One JSON line."""
    from ucla_cmg_amd import autotune, meshgen
    t_setup = time.perf_counter()
    meshes, scales, ratios = meshgen.slab_sweep()
    mesh, inst, offs = meshgen.concat(meshes)
    dl = MODEL_KW["D"][0]
    dt = 0.02
    T = int(round(args.sweep_ms / dt))
    stim_off = int(round(2.0 / dt))
    frame_every = int(round(50.0 / dt))
    n = mesh.n_nodes
    setup_s = time.perf_counter() - t_setup

    def log(msg):
        print(f"[sweep] {msg}", file=sys.stderr, flush=True)

    def run(name, prec, integ, tol, tune, parts):
        sub, sinst, soffs = (mesh, inst, offs) if len(parts) == len(meshes) else \
            meshgen.concat([meshes[k] for k in parts])
        solver = vg.CardiacSolver(sub, cell_model=vg.MODEL_TT, precision=prec, integrator=integ,
                                  dt=dt, D=(dl, dl, dl), cg_tol=tol, param_scales=scales[parts],
                                  param_index=sinst.astype(np.uint16), instance=sinst,
                                  cg_x0=vg.X0_LAST)
        solver.enableActivation(-40.0)
        solver.setStimulusCurrent(np.concatenate(
            [meshgen.slab_corner_nodes(*meshes[k].meta["shape"], 3) + soffs[q]
             for q, k in enumerate(parts)]).astype(np.int32), 50.0 * 2000.0)
        state = {"k": 0, "solve_s": 0.0, "frames": []}

        def step(m):
            # Advances m steps in chunks that end at the stimulus switch and at
            # frame steps; only the solves are timed.
            end = state["k"] + m
            while state["k"] < end:
                k = state["k"]
                if k == stim_off:
                    solver.clearStimulus()
                nxt = min(end, (k // frame_every + 1) * frame_every,
                          stim_off if k < stim_off else end)
                t0 = time.perf_counter()
                solver.Solve(nxt - k)
                solver.synchronize()
                state["solve_s"] += time.perf_counter() - t0
                state["k"] = nxt
                if nxt % frame_every == 0:
                    state["frames"].append(solver.getVoltage())
                    log(f"{name}: {nxt * dt:.0f} ms, {state['solve_s']:.1f} s")

        tuned = None
        if tune:
            step(10)
            tuned = autotune.tune(solver, steps_each=6, blocks=(0, 2048, 4096), step_fn=step)
        step(T - state["k"])
        st = solver.stats()
        res = {"at": solver.activationTimes(), "frames": np.array(state["frames"]),
               "solve_s": state["solve_s"], "stats": st, "tuned": tuned, "offs": soffs,
               "nodes": sub.n_nodes}
        solver.close()
        return res

    every = list(range(len(meshes)))
    checked = every[::args.sweep_check_every]
    mixed = run("mixed", vg.PREC_MIXED, vg.INTEG_RUSH_LARSEN, 1e-7, True, every)
    ref = run("fp64", vg.PREC_FP64, vg.INTEG_REFERENCE, 1e-10, False, checked)
    dat, mism, rms, act = [], [], [], []
    for q, k in enumerate(checked):
        a, b = offs[k], offs[k + 1]
        ra, rb = ref["offs"][q], ref["offs"][q + 1]
        ag, ar = mixed["at"][a:b], ref["at"][ra:rb]
        both = (ag >= 0) & (ar >= 0)
        dat.append(float(np.abs(ag[both] - ar[both]).max()) if both.any() else 0.0)
        mism.append(int(((ag >= 0) != (ar >= 0)).sum()))
        rms.append(float(np.sqrt(((mixed["frames"][:, a:b] - ref["frames"][:, ra:rb]) ** 2)
                                 .mean(axis=1)).max()))
        act.append(float(both.mean()))
    st = mixed["stats"]
    out = {
        "metric": "aggregate node-steps/s, 96 batched tT slabs (per-instance CG), whole run",
        "value": n * T / mixed["solve_s"], "unit": "node-steps/s", "n_gpus": 1, "steps": T,
        "ms_per_step": mixed["solve_s"] * 1e3 / T, "higher_is_better": True,
        "dtype": "fp32 (ionic, SpMV, vectors) + fp64 (CG dots)",
        "data": "synthetic slabs (seeded sizes, scalings, anisotropy), model initial states",
        "config": {"workload": f"96 ten Tusscher slabs, {n} nodes "
                               f"({min(m.n_nodes for m in meshes)}-{max(m.n_nodes for m in meshes)} "
                               f"each), G_Na/G_Kr/G_CaL ~ U[0.5, 1.5], anisotropy ~ U[2, 10], "
                               f"corner stimulus 2 ms, dt 0.02 ms, {args.sweep_ms:g} ms"},
        "cg_iters_per_instance_solve": st["cg_iterations"] / max(1, st["cg_solves"]),
        "cg_not_converged": st["cg_not_converged"], "cg_truncated": st["cg_truncated"],
        "autotune": mixed["tuned"]["table"] if mixed["tuned"] else None,
        "launch": {"cg_blocks": mixed["tuned"]["cg_blocks"], "graph": mixed["tuned"]["graph"]}
                  if mixed["tuned"] else None,
        "fp64_reference": {"ms_per_step": ref["solve_s"] * 1e3 / T, "nodes": ref["nodes"],
                           "cg_not_converged": ref["stats"]["cg_not_converged"]},
        "per_instance_vs_fp64": {
            "instances": checked,
            "nodes": [int(meshes[k].n_nodes) for k in checked],
            "at_max_err_ms": [round(v, 4) for v in dat],
            "activation_mismatch": mism,
            "v_rms_max_mv": [round(v, 4) for v in rms],
            "activated_frac": [round(v, 4) for v in act],
            "worst_at_err_ms": max(dat), "median_at_err_ms": float(np.median(dat)),
            "worst_v_rms_mv": max(rms), "median_v_rms_mv": float(np.median(rms)),
            "frames_ms": [50.0 * (j + 1) for j in range(mixed["frames"].shape[0])],
        },
        "setup_s": setup_s,
    }
    print(json.dumps(out))


def phase_singularities(Vt, Vlag, vstar=-40.0):
    from ucla_cmg_amd import analysis
    return analysis.phase_singularities(Vt, Vlag, vstar)


def mahajan_bench(args, vg):
    """Mahajan rabbit ventricle on a 6 x 6 x 1 cm slab at h = 0.02 cm
    (301 x 301 x 51 = 4,620,651 nodes), isotropic sigma = 1 mS/cm: an S1
    plane wave from x = 0, then an S2 cross-field stimulus over y <= 3 cm at
    a seeded coupling interval in [180, 230] ms; dt = 0.01 ms.  Reports
    node-steps/s, the phase-singularity count on the z = 0 surface and a
    10-lead pseudo-ECG."""
    from ucla_cmg_amd import meshgen
    nx = ny = 301
    nz = 51
    h_ = 0.02
    t0 = time.perf_counter()
    mesh = meshgen.slab(nx, ny, nz, h_)
    dt = 0.01
    steps = int(round(args.mj_ms / dt))
    rng = np.random.default_rng(args.mj_seed)
    s2 = float(rng.uniform(180.0, 230.0))
    amp = 50.0 * 6007.8  # 50 uA/cm^2 through the model's own chi
    sigma = 1.0 / 1400.0
    s = vg.CardiacSolver(mesh, cell_model=vg.MODEL_MJ, precision=vg.PREC_MIXED,
                         integrator=vg.INTEG_REFERENCE, D=(sigma, sigma, sigma), dt=dt,
                         dt_ionic=dt, cg_tol=1e-7, cg_x0=vg.X0_LAST, graph=args.mj_graph,
                         row_order=args.row_order, cg_blocks=args.cg_blocks)
    _streaming(s)
    setup_s = time.perf_counter() - t0
    ring = meshgen.electrodes_around((3.0, 3.0, 0.5), 5.0, 8)
    el = np.vstack([ring, [[3.0, 3.0, 3.0], [3.0, 3.0, -2.0]]])
    s.setupECG(el)
    s.enableActivation(-40.0)
    s.recordECG(100, steps // 100 + 2)
    frame_every = int(round(5.0 / dt))
    events = [(0, "s1_on"), (int(round(2.0 / dt)), "s1_off"), (int(round(s2 / dt)), "s2_on"),
              (int(round((s2 + 2.0) / dt)), "s2_off")]
    surf = []
    ps = []
    k = 0
    t_run = time.perf_counter()
    while k < steps:
        for kk, what in events:
            if kk == k:
                if what == "s1_on":
                    s.setStimulusCurrent(0, 0.06, amp)
                elif what == "s2_on":
                    s.setStimulusCurrent(1, 3.0, amp)
                else:
                    s.clearStimulus()
        nxt = min([steps] + [kk for kk, _ in events if kk > k] + [(k // frame_every + 1) * frame_every])
        s.Solve(nxt - k)
        k = nxt
        if k % frame_every == 0:
            V = s.getVoltage()[: nx * ny].reshape(ny, nx)
            surf.append(V)
            if len(surf) >= 3:
                ps.append((k * dt, *phase_singularities(surf[-1], surf[-3])))
                surf.pop(0)
            if k % (frame_every * 10) == 0:
                print(f"[mahajan] t = {k * dt:.0f} ms, {time.perf_counter() - t_run:.1f} s, "
                      f"PS {ps[-1][1] if ps else 0}", file=sys.stderr, flush=True)
    s.synchronize()
    wall = time.perf_counter() - t_run
    st = s.stats()
    at = s.activationTimes()
    ecg = s.ecgFrames()
    n = mesh.n_nodes
    late = [p for p in ps if p[0] >= s2 + 100.0]
    out = {
        "metric": "node-steps/s, 4.6M-node Mahajan slab, S1-S2 cross-field, 1 GPU",
        "value": n * steps / wall, "unit": "node-steps/s", "n_gpus": 1, "steps": steps,
        "ms_per_step": wall * 1e3 / steps, "higher_is_better": True,
        "dtype": "fp32 gates/voltage/SpMV + fp64 Ca-cycling and Na_i accumulators, CG dots",
        "config": {"workload": f"Mahajan 2008 slab 6x6x1 cm, h 0.02 cm, {n} nodes, sigma 1 mS/cm "
                               f"isotropic, S1 x<=0.06 cm at 0 ms, S2 y<=3 cm at {s2:.1f} ms, "
                               f"dt 0.01 ms, {args.mj_ms:g} ms",
                   "nodes": n, "tets": mesh.n_elems, "nnz": st["nnz"]},
        "cg_iters_per_step": st["cg_iterations"] / max(1, st["steps"]),
        "cg_not_converged": st["cg_not_converged"], "graph": bool(args.mj_graph),
        "row_order": {1: "mesh", 2: "geometric"}.get(st["row_order"], st["row_order"]),
        "sell_wide_slices": st["sell_wide_slices"], "cg_blocks_used": s.cgBlocks(),
        "value_streaming": s.valueStreaming(),
        "activated_frac": float((at >= 0).mean()),
        "phase_singularities": {"series_ms_count_charge": ps[::4],
                                "max": max((p[1] for p in ps), default=0),
                                "mean_after_s2_plus_100ms": float(np.mean([p[1] for p in late])) if late else None},
        "ecg_frames": int(ecg.shape[0]), "ecg_peak_abs": float(np.abs(ecg).max()) if ecg.size else None,
        "setup_s": setup_s,
    }
    s.close()
    print(json.dumps(out))


def surrogate_kernel_timing(args, vg):
    """Cost of the learned step's kernel alone on the LV: an untrained
    network whose envelopes admit every node (increments scaled to ~1e-12 so
    the run stays physical), against the plain ionic kernel."""
    mesh, scales, pidx = lv_workload(args.scale, args.purkinje)
    h = vg.CardiacSolver(mesh, cell_model=vg.MODEL_TT, precision=vg.PREC_MIXED,
                         integrator=vg.INTEG_RUSH_LARSEN, param_scales=scales, param_index=pidx,
                         cg_x0=vg.X0_LAST, **MODEL_KW)
    s = vg.Surrogate(args.surr_seed)
    s.reserve(1 << 20)
    s.collect(h, 1, every=1, node_stride=50)
    s.fit_scaling(0.0)
    sc = s.get_scaling()
    sc["lo"][:], sc["hi"][:] = -1e30, 1e30
    sc["t_lo"][:], sc["t_hi"][:] = -8.0, 8.0
    sc["mu_y"][:], sc["sd_y"][:], sc["c_y"][:] = 0.0, 1.0, 1e-12
    s.set_scaling(sc)

    def ionic_us(steps=100):
        h.profile(True)
        h.Solve(steps)
        h.synchronize()
        ms, k = h.profileGet(vg.KC_IONIC)
        h.profile(False)
        return ms * 1e3 / k

    h.Solve(20)
    plain = ionic_us()
    s.attach(h)
    h.Solve(20)
    surr = ionic_us()
    used, fb = vg.Surrogate.usage(h)
    n = mesh.n_nodes
    print(json.dumps({"metric": "learned ionic step kernel time, 1.03M-node LV", "unit": "us",
                      "plain_ionic_us": plain, "surrogate_ionic_us": surr,
                      "network_rows": used, "fallback_rows": fb,
                      "surrogate_tflops": 2 * SURR_MACS * int(mesh.meta.get("n_myocardium", n))
                      / (surr * 1e-6) / 1e12}))


def surrogate_bench(args, vg, torch):
    """The learned ten Tusscher step on the LV: record the plain model's
    ionic updates from plain mixed-precision runs -- by default of 24 slabs
    of the batched-sweep generator (its sizes, anisotropies and G_Na/G_Kr/
    G_CaL draws), each carrying one of the LV's 24 G_Ks/G_to parameter sets,
    paced at a corner; with --surr-train lv, of the LV beat itself -- train
    the network on the GPU, then run the plain FP64 (the judge), plain mixed
    and surrogate forms over the LV beat, all host-driven, and compare
    activation times, ECG and speed (ionic update alone and whole step)."""
    import math
    mesh, scales, pidx = lv_workload(args.scale, args.purkinje)
    if args.purkinje == "none":
        stim, amp = endo_apex_stimulus(mesh), 50.0 * 2000.0
    else:
        stim, amp = his_stimulus(mesh), 2.0 * 50.0 * 2500.0
    dt = MODEL_KW["dt"]
    T = int(round(args.surr_ms / dt))
    stim_off = int(round(2.0 / dt))
    el = electrodes10()
    kw = dict(cell_model=vg.MODEL_TT, integrator=vg.INTEG_RUSH_LARSEN, param_scales=scales,
              param_index=pidx, cg_x0=vg.X0_LAST, **MODEL_KW)

    def make(prec, graph):
        h = vg.CardiacSolver(mesh, precision=prec, graph=graph, **kw)
        h.enableActivation(-40.0)
        h.setupECG(el)
        h.recordECG(50, T // 50 + 2)
        h.setStimulusCurrent(stim, amp)
        return h

    def result(h, wall=None):
        h.synchronize()
        return dict(at=h.activationTimes(), ecg=h.ecgFrames(), V=h.getVoltage(), wall=wall)

    def log(msg):
        print(f"[surrogate] {msg}", file=sys.stderr, flush=True)

    def timed_run(h, name):
        h.Solve(stim_off)
        h.clearStimulus()
        h.synchronize()
        t0 = time.perf_counter()
        k = stim_off
        while k < T:  # chunks, each ending in a progress line
            m = min(2500, T - k)
            h.Solve(m)
            h.synchronize()
            k += m
            log(f"{name}: step {k}/{T}, {time.perf_counter() - t0:.1f} s")
        return time.perf_counter() - t0

    def ionic_us(h, steps=100):
        h.setLaunch(graph=False)
        h.profile(True)
        h.Solve(steps)
        h.synchronize()
        ms, n = h.profileGet(vg.KC_IONIC)
        h.profile(False)
        return ms * 1e3 / max(1, n)

    n = mesh.n_nodes
    out = {"metric": "node-steps/s, 1M-node tenTusscher heart, 1 GPU, learned ionic step",
           "unit": "node-steps/s", "n_gpus": 1, "higher_is_better": True, "steps": T,
           "data": "synthetic LV shell mesh; network trained on this run's plain-model updates"}
    # 1. Samples from plain mixed runs.
    t0 = time.perf_counter()
    sur = vg.Surrogate(args.surr_seed)
    if args.surr_train == "slabs":
        from ucla_cmg_amd import meshgen
        meshes, sweep_scales, _ = meshgen.slab_sweep()
        pick = list(range(0, len(meshes), 4))[:24]
        rows = np.array([[sweep_scales[k][0], sweep_scales[k][1], scales[q][2],
                          sweep_scales[k][3], scales[q][4]] for q, k in enumerate(pick)])
        tmesh, tinst, toffs = meshgen.concat([meshes[k] for k in pick])
        dl = MODEL_KW["D"][0]
        h = vg.CardiacSolver(tmesh, cell_model=vg.MODEL_TT, precision=vg.PREC_MIXED,
                             integrator=vg.INTEG_RUSH_LARSEN, dt=dt, D=(dl, dl, dl), cg_tol=1e-7,
                             param_scales=rows, param_index=tinst.astype(np.uint16),
                             instance=tinst, cg_x0=vg.X0_LAST)
        h.setStimulusCurrent(np.concatenate(
            [meshgen.slab_corner_nodes(*meshes[k].meta["shape"], 3) + toffs[q]
             for q, k in enumerate(pick)]).astype(np.int32), 50.0 * 2000.0)
        n_rec = tmesh.n_nodes
        out["data"] = (f"synthetic LV shell mesh; network trained on the plain-model updates of "
                       f"{len(pick)} slabs of the sweep generator ({n_rec} nodes, corner pacing, "
                       f"each with one of the LV's G_Ks/G_to sets), not on the LV")
    else:
        h = make(vg.PREC_MIXED, 0)
        n_rec = int(mesh.meta.get("n_myocardium", n))
    records = -(-(T - stim_off) // args.surr_every)
    # Node stride: --surr-stride, or more to keep to about 2e8 samples.
    stride = max(args.surr_stride, -(-n_rec * records // 200_000_000))
    per = -(-n_rec // stride)
    sur.reserve(per * records)
    h.Solve(stim_off)
    h.clearStimulus()
    k = stim_off
    while k < T:
        m = min(2500, T - k)
        sur.collect(h, m, every=args.surr_every, node_stride=stride)
        k += m
        log(f"collect: step {k}/{T}, {sur.num_samples()} samples")
    h.close()
    out["samples"] = sur.num_samples()
    out["collect_s"] = time.perf_counter() - t0
    # 2. Training (cosine learning-rate decay).
    sur.fit_scaling(0.02)
    if args.surr_v_weight != 1.0:
        # V's transformed increment spread by k: its squared error weighs k^2 in the
        # loss (the step's inverse transform uses the same spread; its output
        # envelope, in the same units, scales with it).
        sc = sur.get_scaling()
        k = float(args.surr_v_weight)
        sc["sd_y"][0] /= k
        sc["t_lo"][0] *= k
        sc["t_hi"][0] *= k
        sur.set_scaling(sc)
    # Fallback settings judged beside the network alone: "up:V_min:h_min"
    # (upstroke) and/or "band:V_lo:V_hi", joined by "+".
    inf = float("inf")
    forms = [("fp64", vg.PREC_FP64, None), ("plain", vg.PREC_MIXED, None),
             ("surrogate", vg.PREC_MIXED, {"up": (inf, inf), "band": (inf, -inf)})]
    for spec in [u for u in args.surr_fallback.split(",") if u]:
        fb = {"up": (inf, inf), "band": (inf, -inf)}
        for part in spec.split("+"):
            kind, a, b = part.split(":")
            fb[kind] = (float(a), float(b))
        forms.append((f"surrogate_{spec}", vg.PREC_MIXED, fb))
    head_fb = forms[3][2] if len(forms) > 3 else forms[2][2]

    # Checkpoint selection (--surr-select): a sweep-generator slab that neither
    # training nor the judge sees, run coupled with the headline form against
    # the plain mixed model; the training checkpoint with the lowest ECG error
    # there is the one judged.
    val = None
    if args.surr_select:
        from ucla_cmg_amd import meshgen
        vmeshes, vscales, _ = meshgen.slab_sweep()
        vk = 30  # 38,025 nodes, not among the training slabs (every 4th)
        vmesh = vmeshes[vk]
        vrow = np.array([[vscales[vk][0], vscales[vk][1], scales[12][2], vscales[vk][3],
                          scales[12][4]]])
        lo_, hi_ = vmesh.xyz.min(axis=0), vmesh.xyz.max(axis=0)
        ctr = 0.5 * (lo_ + hi_)
        vel = np.vstack([meshgen.electrodes_around(ctr, 2.0 * float((hi_ - lo_)[:2].max()), 8),
                         [ctr + [0, 0, 2.0], ctr - [0, 0, 2.0]]])
        vT = int(round(args.surr_val_ms / dt))
        dl = MODEL_KW["D"][0]

        def val_run(net):
            g = vg.CardiacSolver(vmesh, cell_model=vg.MODEL_TT, precision=vg.PREC_MIXED,
                                 integrator=vg.INTEG_RUSH_LARSEN, dt=dt, D=(dl, dl, dl),
                                 cg_tol=1e-7, param_scales=vrow,
                                 param_index=np.zeros(vmesh.n_nodes, np.uint16), cg_x0=vg.X0_LAST)
            g.enableActivation(-40.0)
            g.setupECG(vel)
            g.recordECG(50, vT // 50 + 2)
            if net:
                sur.set_upstroke_fallback(*head_fb["up"])
                sur.set_band_fallback(*head_fb["band"])
                sur.attach(g)
            g.setStimulusCurrent(meshgen.slab_corner_nodes(*vmesh.meta["shape"], 3), 50.0 * 2000.0)
            g.Solve(stim_off)
            g.clearStimulus()
            g.Solve(vT - stim_off)
            r = dict(at=g.activationTimes(), ecg=g.ecgFrames())
            g.close()
            return r

        vref = val_run(False)
        val = {"slab": vk, "nodes": vmesh.n_nodes, "ms": args.surr_val_ms, "checks": []}
        best = None

    t0 = time.perf_counter()
    its, done, curve = args.surr_iters, 0, []
    while done < its:
        m = min(2000, its - done)
        lr = 2e-5 + 0.5 * (args.surr_lr - 2e-5) * (1 + math.cos(math.pi * done / its))
        loss = sur.train(m, batch=args.surr_batch, lr=lr)
        done += m
        if done % 20000 == 0 or done == its:
            curve.append((done, loss))
            log(f"train: {done}/{its} loss {loss:.3e}")
            if val is not None and done >= its // 4:
                r = val_run(True)
                both = (r["at"] >= 0) & (vref["at"] >= 0)
                e = float(np.linalg.norm(r["ecg"] - vref["ecg"]) / np.linalg.norm(vref["ecg"]))
                a = float(np.abs(r["at"][both] - vref["at"][both]).max()) if both.any() else 1e9
                val["checks"].append({"iterations": done, "loss": loss, "ecg_rel_l2": e,
                                      "at_max_err_ms": a})
                log(f"validation at {done}: ECG {e:.4f}, max dAT {a:.3f} ms")
                if best is None or e < best[0]:
                    best = (e, done, sur.get_params())
    train_s = time.perf_counter() - t0
    out["train"] = {"iterations": its, "batch": args.surr_batch, "seconds": train_s, "lr": args.surr_lr,
                    "loss_curve": curve,
                    "tflops": 3 * 2 * SURR_MACS * its * args.surr_batch / train_s / 1e12}
    if val is not None:
        sur.set_params(best[2])
        val["chosen_iterations"] = best[1]
        out["train"]["selection"] = val
    # 3. The three forms over the same beat, host-driven (every solve to its
    # tolerance, checked below).
    runs = {}
    for name, prec, use in forms:
        h = make(prec, 0)
        if use:
            sur.set_upstroke_fallback(*use["up"])
            sur.set_band_fallback(*use["band"])
            sur.attach(h)
        wall = timed_run(h, name)
        runs[name] = result(h, wall)
        st = h.stats()
        runs[name]["cg"] = {"not_converged": st["cg_not_converged"],
                            "truncated": st["cg_truncated"],
                            "its_per_step": st["cg_iterations"] / max(1, st["steps"])}
        if use:
            runs[name]["usage"] = vg.Surrogate.usage(h)
        if name != "fp64":
            runs[name]["ionic_us"] = ionic_us(h)
        h.close()
    ref = runs["fp64"]
    for name in [f[0] for f in forms[1:]]:
        r = runs[name]
        both = (r["at"] >= 0) & (ref["at"] >= 0)
        dat = np.abs(r["at"][both] - ref["at"][both])
        out[name] = {
            "node_steps_per_s": n * (T - stim_off) / r["wall"],
            "ms_per_step": r["wall"] * 1e3 / (T - stim_off),
            "ionic_kernel_us": r["ionic_us"],
            "at_max_err_ms": float(dat.max()), "at_mean_err_ms": float(dat.mean()),
            "at_p99_err_ms": float(np.percentile(dat, 99)),
            "at_frac_over_1ms": float((dat > 1.0).mean()),
            "activated_frac": float(both.mean()),
            "activated_mismatch": int(((r["at"] >= 0) != (ref["at"] >= 0)).sum()),
            "ecg_rel_l2": float(np.linalg.norm(r["ecg"] - ref["ecg"]) / np.linalg.norm(ref["ecg"])),
            "v_rms_mv_end": float(np.sqrt(np.mean((r["V"] - ref["V"]) ** 2))),
            "cg": r["cg"],
        }
    out["fp64_cg"] = ref["cg"]
    for name, _, use in forms[2:]:
        u = runs[name]["usage"]
        out[name]["network_frac"] = u[0] / max(1, u[0] + u[1])
        out[name]["fallback"] = {k: list(v) for k, v in use.items() if v[0] != inf}
        out[name]["gates"] = {"at_le_1ms": out[name]["at_max_err_ms"] <= 1.0,
                              "ecg_le_3pct": out[name]["ecg_rel_l2"] <= 0.03,
                              "speedup_ge_2": out["plain"]["ms_per_step"]
                              / out[name]["ms_per_step"] >= 2.0}
    # The headline form: the first fallback setting given (the default is the
    # upstroke fallback), else the network alone.
    main_form = forms[3][0] if len(forms) > 3 else "surrogate"
    out["form"] = main_form
    out["value"] = out[main_form]["node_steps_per_s"]
    out["ms_per_step"] = out[main_form]["ms_per_step"]
    out["speedup_vs_plain"] = out["plain"]["ms_per_step"] / out[main_form]["ms_per_step"]
    out["ionic_speedup_vs_plain"] = out["plain"]["ionic_kernel_us"] / out[main_form]["ionic_kernel_us"]
    out["gates"] = out[main_form]["gates"]
    out["train"]["v_weight"] = args.surr_v_weight
    out["config"] = {"workload": f"tT2006 LV shell + Purkinje tree ({args.purkinje}), {n} nodes, "
                                 f"{args.surr_ms:g} ms beat, dt {dt} ms, network 25-128x4-19 SiLU BF16",
                     "sample_every_steps": args.surr_every, "sample_node_stride": stride}
    print(json.dumps(out))


def c1_bench(args, vg, torch):
    """The small Luo-Rudy slab (41 x 41 x 9 nodes, 15,129; 100 ms at dt 0.01
    ms, 10,000 steps): the GPU in FP64 (CG to 1e-10, the configuration's
    tolerance) and mixed precision, and the oracle over the same 10,000 steps
    on every available CPU plus a one-core sample.  One JSON line."""
    from ucla_cmg_amd import meshgen
    oracle, build = oracle_for_baseline()
    mesh = meshgen.slab()
    n = mesh.n_nodes
    D = (1.2 / 1400.0, 0.2 / 1400.0, 0.2 / 1400.0)
    dt, ms = 0.01, 100.0
    steps = int(round(ms / dt))
    stim = meshgen.slab_corner_nodes()
    amp = 50.0 * oracle.lr_xi()
    out = {"metric": "node-steps/s, 15,129-node Luo-Rudy slab, 100 ms", "unit": "node-steps/s",
           "config": {"workload": "LR1991 slab 41x41x9 (h 0.025 cm, 64,000 tets), corner "
                                  "stimulus 2 ms, dt 0.01 ms, 10,000 steps"}}
    for name, prec, tol in (("gpu_fp64", vg.PREC_FP64, 1e-10), ("gpu_mixed", vg.PREC_MIXED, 1e-7)):
        s = vg.CardiacSolver(mesh, cell_model=vg.MODEL_LR, precision=prec, D=D, dt=dt, cg_tol=tol,
                             cg_x0=vg.X0_LAST)
        s.enableActivation(-40.0)
        s.setStimulusCurrent(stim, amp)
        s.Solve(int(2.0 / dt))
        s.clearStimulus()
        s.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.Solve(steps - int(2.0 / dt))
        s.synchronize()
        el = time.perf_counter() - t0
        st = s.stats()
        at = s.activationTimes()
        out[name] = {"node_steps_per_s": n * (steps - int(2.0 / dt)) / el,
                     "ms_per_step": el * 1e3 / (steps - int(2.0 / dt)),
                     "cg_iters_per_step": st["cg_iterations"] / st["steps"],
                     "activated": float((at >= 0).mean()), "timed_steps": steps - int(2.0 / dt)}
        s.close()
    host = host_cpu()

    # Both start from the library's initial state (the Luo-Rudy rest state).
    s0 = vg.CardiacSolver(mesh, cell_model=vg.MODEL_LR, precision=vg.PREC_FP64, D=D, dt=dt)
    rest_states, rest_V = s0.getStateVariables(), s0.getVoltage()
    s0.close()

    def run_oracle(threads, nsteps):
        sim = oracle.Simulation(mesh, oracle.MODEL_LR, dt=dt, D=D, cg_tol=1e-10, nthreads=threads,
                                states=rest_states, V0=rest_V)
        sim.istim[stim] = amp
        t0 = time.perf_counter()
        for k in range(nsteps):
            if k == int(2.0 / dt):
                sim.istim[:] = 0.0
            sim.step()
        return n * nsteps / (time.perf_counter() - t0), time.perf_counter() - t0
    v_all, e_all = run_oracle(host["cpus_available"], steps)
    v_one, e_one = run_oracle(1, args.c1_one_core_steps)
    out["cpu_baseline"] = {"value": v_all, "unit": "node-steps/s", "cores": host["cpus_available"],
                           "kind": "port", "sample": f"all {steps} steps, {e_all:.1f} s, oracle "
                                                     f"built {build}",
                           "one_core": {"value": v_one, "steps": args.c1_one_core_steps,
                                        "seconds": e_one},
                           "host": host}
    out["value"] = out["gpu_mixed"]["node_steps_per_s"]
    out["gpu_fp64_over_cpu"] = out["gpu_fp64"]["node_steps_per_s"] / v_all
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=19800,
                    help="timed steps (default: the rest of the 400 ms beat after the warm-up)")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--scale", type=float, default=1.0, help="mesh size factor (1.0 = 1.03M nodes)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU baseline: seconds on all available CPUs (then --cpu-seconds-one "
                         "on one core)")
    ap.add_argument("--cpu-seconds-one", type=float, default=5.0)
    ap.add_argument("--window-ms", type=float, default=None,
                    help="start of the warm-up + timed window in the beat (default 150 ms, "
                         "mid-plateau; 2 ms when the window spans the rest of the beat)")
    ap.add_argument("--no-beat", action="store_true",
                    help="stop after the timed window instead of finishing the 400 ms beat")
    ap.add_argument("--cg-blocks", type=int, default=0,
                    help="workgroups per CG launch (0 = library default)")
    ap.add_argument("--cg-x0", choices=["zero", "previous", "last", "extrapolate"], default="last",
                    help="initial guess of each half-step solve (same stopping bound)")
    ap.add_argument("--graph", type=int, default=0, choices=[0, 1],
                    help="replay each step as a captured hipGraph (no host syncs)")
    ap.add_argument("--value-streaming", type=int, default=-1, choices=[-1, 0, 1],
                    help="CG operator values: -1 library's choice by size, 0 plain, 1 nontemporal")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep --cg-blocks/--graph instead of autotuning them in the warm-up")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the batched 96-slab sweep measurement")
    ap.add_argument("--sweep-steps", type=int, default=100)
    ap.add_argument("--sweep-full", action="store_true",
                    help="the 96-slab sweep over --sweep-ms (whole run timed) with per-instance "
                         "errors against the FP64 reference-integrator run, instead")
    ap.add_argument("--sweep-ms", type=float, default=500.0)
    ap.add_argument("--sweep-check-every", type=int, default=4,
                    help="--sweep-full: instances checked against the FP64 reference run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of this many MiB 20 times (PMC byte "
                         "counter calibration) and exit")
    ap.add_argument("--c1", action="store_true",
                    help="the 15,129-node Luo-Rudy slab for 100 ms: GPU (FP64, mixed) and the "
                         "oracle over the same run on all CPUs, plus a one-core sample")
    ap.add_argument("--c1-one-core-steps", type=int, default=2000)
    ap.add_argument("--mahajan", action="store_true",
                    help="the 4.6M-node Mahajan S1-S2 slab instead (phase singularities, ECG)")
    ap.add_argument("--mj-ms", type=float, default=1000.0)
    ap.add_argument("--mj-seed", type=int, default=7)
    ap.add_argument("--row-order", type=int, default=0, choices=[0, 1, 2],
                    help="myocardium row order: 0 auto (mesh order unless a column is more than "
                         "32767 rows away), 1 mesh, 2 geometric")
    ap.add_argument("--mj-graph", type=int, default=0, choices=[0, 1],
                    help="Mahajan slab: 1 = captured step (hipGraph), 0 = host-driven CG")
    ap.add_argument("--surrogate", action="store_true",
                    help="evaluate the learned ionic step on the LV instead (collect, train, compare)")
    ap.add_argument("--surr-ms", type=float, default=400.0)
    ap.add_argument("--surr-train", choices=["slabs", "lv"], default="slabs",
                    help="training data: 24 sweep-generator slabs (default) or the LV beat")
    ap.add_argument("--surr-every", type=int, default=20)
    ap.add_argument("--surr-stride", type=int, default=5)
    ap.add_argument("--surr-iters", type=int, default=160000)
    ap.add_argument("--surr-batch", type=int, default=16384)
    ap.add_argument("--surr-seed", type=int, default=1)
    ap.add_argument("--surr-lr", type=float, default=1e-3, help="peak Adam step (cosine decay)")
    ap.add_argument("--surr-fallback", default="up:-80:0.05,up:-80:0.05+band:-200:-80",
                    help="comma-separated fallback settings judged beside the network alone, "
                         "each 'up:V_min:h_min' (upstroke) and/or 'band:V_lo:V_hi' joined "
                         "by '+' (empty: none)")
    ap.add_argument("--surr-select", action="store_true",
                    help="judge the training checkpoint with the lowest ECG error on a held-out "
                         "sweep slab (checked every 20k iterations from a quarter of training)")
    ap.add_argument("--surr-val-ms", type=float, default=300.0)
    ap.add_argument("--surr-v-weight", type=float, default=1.0,
                    help="weight k of V's output in the training loss (as k^2)")
    ap.add_argument("--surr-kernel-only", action="store_true",
                    help="with --surrogate: time the learned step's kernel alone")
    ap.add_argument("--purkinje", choices=["high", "low", "none"], default="high",
                    help="Purkinje tree with high/low PMJ resistance, or none")
    args = ap.parse_args()
    global VALUE_STREAMING
    VALUE_STREAMING = args.value_streaming
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch
    from ucla_cmg_amd import voomgpu as vg

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(20):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "launches": 20}))
        return
    if args.c1:
        c1_bench(args, vg, torch)
        return
    if args.sweep_full:
        sweep_full(args, vg, torch)
        return
    if args.mahajan:
        mahajan_bench(args, vg)
        return
    if args.surrogate:
        if args.surr_kernel_only:
            surrogate_kernel_timing(args, vg)
        else:
            surrogate_bench(args, vg, torch)
        return
    t_setup = time.perf_counter()
    mesh, scales, pidx = lv_workload(args.scale, args.purkinje)
    if args.purkinje == "none":
        stim, amp = endo_apex_stimulus(mesh), 50.0 * 2000.0
    else:
        stim, amp = his_stimulus(mesh), 2.0 * 50.0 * 2500.0
    solver = vg.CardiacSolver(mesh, cell_model=vg.MODEL_TT, precision=vg.PREC_MIXED,
                              integrator=vg.INTEG_RUSH_LARSEN, param_scales=scales,
                              param_index=pidx, cg_blocks=args.cg_blocks,
                              cg_x0={"zero": vg.X0_ZERO, "previous": vg.X0_PREVIOUS,
                                     "last": vg.X0_LAST,
                                     "extrapolate": vg.X0_EXTRAPOLATE}[args.cg_x0],
                              graph=args.graph, row_order=args.row_order,
                              true_residual_every=TRUE_RESIDUAL_EVERY, **MODEL_KW)
    _streaming(solver)
    solver.enableActivation(-40.0)
    solver.setupECG(electrodes10())
    solver.recordECG(50, int(BEAT_MS / 1.0) + (args.warmup + args.steps + 200) // 50 + 4)
    solver.setStimulusCurrent(stim, amp)
    setup_s = time.perf_counter() - t_setup
    n = mesh.n_nodes
    stim_off = int(2.0 / 0.02)

    def run(k0, k1):
        k = k0
        while k < k1:
            if k == stim_off:
                solver.clearStimulus()
            nxt = min(k1, stim_off) if k < stim_off else k1
            solver.Solve(nxt - k)
            k = nxt

    class Stepper:
        """Advances the run k steps at a time, switching the stimulus off
        at 2 ms whatever the chunking."""
        k = 0

        def __call__(self, m):
            run(self.k, self.k + m)
            self.k += m

    step = Stepper()
    tuned = None
    beat_steps = int(round(BEAT_MS / 0.02))
    if args.no_tune:
        step(min(args.warmup, 25))
    else:
        from ucla_cmg_amd import autotune
        step(25)
        tuned = autotune.tune(solver, steps_each=8, blocks=(0, 2048), step_fn=step)
    solver.synchronize()
    # Where the warm-up + timed window starts: mid-plateau by default, the
    # phase that holds most of the beat's steps and cost (27-29 CG iterations
    # per step against ~15 while the tree conducts); right after the His
    # stimulus when the window spans the rest of the beat.
    if args.window_ms is not None:
        w0 = int(round(args.window_ms / 0.02))
    elif args.warmup + args.steps >= beat_steps - stim_off:
        w0 = stim_off
    else:
        w0 = int(round(150.0 / 0.02))
    w0 = max(w0, step.k)

    # Every stretch of the beat outside the autotuner and the profiled steps
    # is timed too (wall clock, synchronised at its ends), for the whole-beat
    # figure and the per-phase table.
    phases = []
    beat = {"steps": 0, "s": 0.0, "its": 0}

    def segment(m, record=True):
        tw, sw = time.perf_counter(), solver.stats()["cg_iterations"]
        step(m)
        solver.synchronize()
        dt_s, its = time.perf_counter() - tw, solver.stats()["cg_iterations"] - sw
        if record:
            phases.append({"t_ms": round((step.k - m) * 0.02, 2), "steps": m,
                           "ms_per_step": dt_s * 1e3 / m, "cg_iters_per_step": its / m})
            beat["steps"] += m
            beat["s"] += dt_s
            beat["its"] += its

    def roll(until):
        while step.k < until:
            segment(min(2000, until - step.k))

    roll(w0)
    if args.warmup:
        segment(args.warmup)
    # The CPU baseline starts from the GPU run's state at mid-plateau (150 ms):
    # before the timed window when that starts there, else at the step of the
    # timed window that reaches 150 ms, where the window is split in two
    # stretches (each synchronised at both ends, both timed) around the
    # download of the state.
    snap = None
    k_snap = int(round(150.0 / 0.02))

    def take_snap():
        return (solver.getVoltage(), solver.getStateVariablesRagged()[1], step.k)

    if not args.no_cpu_baseline and not (step.k < k_snap < step.k + args.steps):
        snap = take_snap()
    st0 = solver.stats()
    k0 = step.k
    stream = torch.cuda.ExternalStream(solver.stream())
    parts = ([k_snap - k0, k0 + args.steps - k_snap] if not args.no_cpu_baseline and snap is None
             else [args.steps])
    wall, ev_ms = 0.0, 0.0
    for q, m in enumerate(parts):
        if q:
            snap = take_snap()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record(stream)
        t0 = time.perf_counter()
        step(m)
        ev1.record(stream)
        solver.synchronize()
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
        ev_ms += ev0.elapsed_time(ev1)
    st1 = solver.stats()
    cg_its = st1["cg_iterations"] - st0["cg_iterations"]
    k_per_step = cg_its / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = n * args.steps / wall
    phases.append({"t_ms": round(k0 * 0.02, 2), "steps": args.steps, "ms_per_step": ms_per_step,
                   "cg_iters_per_step": k_per_step, "timed": True})
    beat["steps"] += args.steps
    beat["s"] += wall
    beat["its"] += cg_its

    # Per-kernel times: HIP events around every launch on the solver's stream,
    # over a window of the same run right after the timed one (the events
    # cost a little time of their own, so the timed window runs without).
    prof_steps = 50
    solver.profile(True)
    stp0 = solver.stats()
    tp0 = time.perf_counter()
    step(prof_steps)
    solver.synchronize()
    prof_wall_ms = (time.perf_counter() - tp0) * 1e3
    prof = {name: solver.profileGet(kc) for kc, name in enumerate(vg.KC_NAMES)}
    prof_its = solver.stats()["cg_iterations"] - stp0["cg_iterations"]
    solver.profile(False)
    if not args.no_beat:
        roll(beat_steps)
    beat_stats = solver.stats()

    # Dominant kernel: the fused CG iteration k_cg_iter<float>.  Algorithmic
    # bytes per launch: every stored nonzero once (FP32 value + int32
    # column) and per row the vectors it must read (r, w, s, p, x, diag) and
    # write (r, w, s, p, x), 4 B each; neighbours' r, w, s are the same
    # vectors, counted once.
    it_ms, it_launches = prof["cg_iter"]
    nnz = st1["nnz"]
    it_bytes = 8 * nnz + 44 * n
    # The library times each solve's iteration launches as one interval (an
    # event pair around all of them, so the launches run back to back as in
    # the timed region, launch boundaries included).  Launches after
    # convergence exit at once; their few microseconds are charged to the
    # launches that did an iteration (conservative).
    it_avg_s = (it_ms / 1e3) / max(1, prof_its) if it_launches else float("nan")
    achieved = it_bytes / it_avg_s / 1e9
    traffic = measured_traffic(nnz, st1["sell_padded_nnz"])
    ionic_ms, ionic_launches = prof["ionic"]
    ionic_bytes = n * (18 * 4 * 2 + 4 + 4 + 4 + 4 + 2)
    kern_tot = sum(v[0] for v in prof.values())
    out = {
        "metric": "node-steps/s, 1M-node tenTusscher heart, 1 GPU; surrogate vs plain at <=1 ms dAT",
        "value": value,
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (ionic, SpMV, vectors) + fp64 (CG dots, residual, ECG sums)",
        "data": "synthetic LV shell mesh, model initial states",
        "config": {
            "workload": f"tT2006 LV shell + Purkinje tree ({args.purkinje} PMJ resistance), "
                        f"{n} nodes, {mesh.n_elems} elements, plain path (Rush-Larsen ionic), "
                        "dt 0.02 ms, CG tol 1e-7, AT + 10-lead ECG/ms; timed: steps "
                        f"{k0}-{k0 + args.steps} ({k0 * 0.02:g}-{(k0 + args.steps) * 0.02:g} ms "
                        "of the 400 ms beat, after the state was advanced there untimed)",
            "nodes": n, "tets": mesh.n_elems, "nnz": nnz, "cg_iters_per_step": k_per_step,
            "cg_blocks": args.cg_blocks,
            "cg_x0": args.cg_x0,
            "cg_blocks_used": solver.cgBlocks(),
            "value_streaming": solver.valueStreaming(),
            "graph": tuned["graph"] if tuned else bool(args.graph),
            "row_order": {1: "mesh", 2: "geometric"}.get(solver.stats()["row_order"]),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_cg_iter<float>",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": it_bytes,
            "us_per_iteration": it_avg_s * 1e6,
            "avg_launch_us": (it_ms * 1e3) / max(1, it_launches),
            "launches_per_iteration": it_launches / max(1, prof_its),
        },
        "whole_beat": {
            "steps": beat["steps"],
            "t_end_ms": round(step.k * 0.02, 2),
            "ms_per_step": beat["s"] * 1e3 / max(1, beat["steps"]),
            "node_steps_per_s": n * beat["steps"] / beat["s"] if beat["s"] else None,
            "cg_iters_per_step": beat["its"] / max(1, beat["steps"]),
            "cg_not_converged": beat_stats["cg_not_converged"],
            "cg_truncated": beat_stats["cg_truncated"],
            "true_residual_max": beat_stats["max_true_relres"],
            "true_residual_checks": beat_stats["true_residual_checks"],
            "note": "every step of the run except the autotuner's and the 50 profiled ones, "
                    "each stretch wall-clocked between stream synchronisations",
        },
        "phases": phases,
        "kernels_ms": {k: v[0] for k, v in prof.items()},
        "kernel_launches": {k: v[1] for k, v in prof.items()},
        "ionic_GBps": ionic_bytes / ((ionic_ms / 1e3) / max(1, ionic_launches)) / 1e9 if ionic_launches else None,
        "event_ms_timed_region": ev_ms,
        "profiled_ms_per_step": prof_wall_ms / prof_steps,
        "kernel_busy_frac": kern_tot / prof_wall_ms,
        "setup_s": setup_s,
    }
    if tuned:
        out["autotune"] = tuned["table"]
    solver.close()
    del solver
    if not args.no_sweep:
        try:
            out["sweep96"] = sweep_bench(vg, torch, args.sweep_steps, not args.no_tune)
        except Exception as e:  # reported, never hides the headline line
            out["sweep96"] = {"error": repr(e)}
    if snap is not None:
        out["cpu_baseline"] = cpu_baseline(mesh, scales, pidx, snap[0], snap[1], snap[2],
                                           args.cpu_seconds, args.cpu_seconds_one)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
