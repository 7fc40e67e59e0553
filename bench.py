"""SPH step benchmark: particle-updates/s on the 1,000,000-particle dam break (one MI355X).

Workload: 100^3 jittered lattice (spacing 0.01, +-5e-4 jitter), box 3 x 1.5 x 1.2 (z-up),
h = 0.02, rho0 = 1000, k = 2000, mu = 0.001, m = 0.001, dt = 0.001, reference formulas.

Protocol (reset-and-step): the reference formulas throw almost every particle into the box
corners within one step, after which the neighbour count is ~N/8 and the run stops being a
dam break.  Each timed step therefore restores the initial state on the device (a
device-to-device copy, inside the timed region) and runs one full step: cell keys, sort,
cell table, gather, density, forces, leapfrog, wall clamp.  Inputs are resident in HBM
before timing starts.  The timed steps record no phase events; the phase split (HIP events on
the engine stream, the roofline's kernel time) comes from as many steps right after.

Prints one JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "sph-particle-simulator_amd"))

METRIC = "particle-updates/s @1M dam break: plain GPU SPH vs surrogate at ≤1% ρ error"

# Algorithmic HBM bytes per particle-update (FP32 SoA, state kept cell-sorted):
#   density pass:  read x,y,z 12 + write rho 4                       = 16
#   force pass:    read x,y,z,v,rho 28 + write x,v 24                = 52
#   whole step:    keys 16 + histogram 4 + scan ~0.7 + reorder 60 + 16 + 52 ~= 149
ALG_BYTES = {"density": 16.0, "force": 52.0, "step": 149.0}
# Measured HBM traffic and VALU issue of the force walk on the 1M dam break: rocprofv3 counter
# runs of the build named in the file (FETCH_SIZE and WRITE_SIZE, each scaled by what the same
# counter reads for a device copy of known size).  Reported only for the launch shape the
# counters were taken with.
COUNTER_PROFILE = os.path.join(REPO, "profiles", "force_walk_dam_break_1M_counters.json")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # published; counts packed FP32 (two lanes per VALU slot)
FP32_PEAK_UNPACKED_TFLOPS = 78.6  # one wave64 FP32 instruction per SIMD per 4 cycles
# algorithmic FP32 work per neighbour pair (FMA = 2, sqrt/rsqrt = 1): density 17, force 37
FLOP_PER_PAIR = 17 + 37
BF16_PEAK_TFLOPS = 2500.0  # dense MFMA


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", type=int, default=1, help="scene: 0 (10k), 1 (1M), 2 (9.6M)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default walk launch shapes instead of tuning them first")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-ensemble and graph-mode lines")
    ap.add_argument("--ensemble-scale", type=float, default=1.0,
                    help="particle-count scale of the 64-scene ensemble (1.0 = full size)")
    ap.add_argument("--full-run", type=int, default=1000,
                    help="consecutive steps of the full run (consistent formulation, dt=1e-4) "
                         "with its energy trajectory; 0 skips it")
    ap.add_argument("--no-cpu-legs", action="store_true",
                    help="CPU baseline of the headline scene only (no 10k and ensemble legs)")
    ap.add_argument("--walk-shapes", type=str, default=None,
                    help="density,force launch shapes (indices of sphk.walk_shape_names()); "
                         "skips the autotuner")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device three times "
                         "(reference launches for HBM byte counters)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the headline step from a captured HIP graph")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the learned macro-step surrogate line")
    ap.add_argument("--surrogate-snapshots", type=int, default=2000,
                    help="training snapshots of the surrogate (macro-steps of plain SPH)")
    ap.add_argument("--surrogate-macro", type=int, default=64,
                    help="macro-steps of the timed, judged surrogate run")
    return ap.parse_args()


def _cpu_reset_step(sph_oracle, scene, repeats):
    o = sph_oracle.Oracle(scene.n, scene.oracle_params(sph_oracle), fast=True)
    times = []
    for _ in range(repeats):
        o.set_particles(scene.pos, scene.vel)
        t0 = time.perf_counter()
        o.step(scene.dt)
        times.append(time.perf_counter() - t0)
    cores = int(o.lib.or_num_threads())
    del o
    return statistics.median(times), cores


def cpu_baseline(scene, repeats, sph_scenes, legs=True):
    """Reference step restated in C (oracle, built with the reference's -O3 -mavx2 -ffast-math
    -fopenmp), on this host's cores: reset-and-step of step 0, median of `repeats`.  The other
    legs: the 10k dam break (reset-and-step, and its full 100-step run under the reference
    formulas) and the 64-scene ensemble stepped scene after scene (step 0 of each)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import sph_oracle

    med, cores = _cpu_reset_step(sph_oracle, scene, repeats)
    out = {
        "value": scene.n / med,
        "unit": "particle-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{scene.name} step 0 (reset-and-step), median of {repeats}, "
                  f"{med:.3f} s/step; C restatement of the reference step (hash-map build, "
                  f"125-cell query, per-particle neighbour vectors, OpenMP), reference flags",
    }
    if not legs:
        return out
    c0 = sph_scenes.config0()
    med0, _ = _cpu_reset_step(sph_oracle, c0, repeats)
    o = sph_oracle.Oracle(c0.n, c0.oracle_params(sph_oracle), fast=True)
    o.set_particles(c0.pos, c0.vel)
    t0 = time.perf_counter()
    for _ in range(c0.steps):
        o.step(c0.dt)
    t_full = time.perf_counter() - t0
    del o
    out["dam_break_10k"] = {
        "reset_and_step": {"value": c0.n / med0, "s_per_step": med0},
        "full_run": {"steps": c0.steps, "value": c0.n * c0.steps / t_full,
                     "s_per_step": t_full / c0.steps,
                     "regime": "reference formulas: the block is thrown into the box corners "
                               "by step 1 (pile-up, ~N/8 neighbours)"},
    }
    ens = sph_scenes.config3(seed=0)
    tot, t_ens = 0, 0.0
    for sc in ens:
        o = sph_oracle.Oracle(sc.n, sc.oracle_params(sph_oracle), fast=True)
        o.set_particles(sc.pos, sc.vel)
        t0 = time.perf_counter()
        o.step(sc.dt)
        t_ens += time.perf_counter() - t0
        tot += sc.n
        del o
    out["ensemble_64_sequential"] = {"particles": tot, "value": tot / t_ens,
                                     "s_per_step": t_ens,
                                     "sample": "step 0 of each of the 64 scenes in turn"}
    return out


def _load_counters(shape_name):
    """Counter-measured traffic and VALU issue of the force walk, if taken for this shape."""
    try:
        with open(COUNTER_PROFILE) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None
    if prof.get("force_shape") != shape_name:
        return None
    return prof


def time_reset_steps(e, dt, steps, warmup):
    """Reset-and-step timing of an engine whose initial state is its snapshot."""
    for _ in range(warmup):
        e.snapshot_restore()
        e.step(dt)
    e.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        e.snapshot_restore()
        e.step(dt)
    e.synchronize()
    return (time.perf_counter() - t0) / steps


def pool_line(sphk, sph_scenes, shapes, steps=10):
    """Drop into pool, 9,614,147 particles (400 x 60 x 400 pool + a lattice sphere falling at
    2 m/s), reset-and-step with the 1M scene's tuned walk shapes."""
    scene = sph_scenes.config2()
    e = sphk.Engine(scene.n, scene.params(sphk))
    e.set_walk_shapes(*shapes)
    e.upload(scene.pos, scene.vel)
    e.snapshot_save()
    t = time_reset_steps(e, scene.dt, steps, 2)
    e.reset_stats()
    e.snapshot_restore()
    e.step(scene.dt)
    st = e.stats()
    out = {"particles": scene.n, "steps": steps, "ms_per_step": t * 1e3, "value": scene.n / t,
           "unit": "particle-updates/s",
           "mean_neighbors": st.total_neighbors / max(st.total_neighbor_queries, 1),
           "max_neighbors": st.max_neighbors,
           "walk_shapes": [sphk.walk_shape_names()[i] for i in shapes]}
    del e
    return out


def ensemble_line(sphk, sph_scenes, scale, steps):
    """64 independent dam-break scenes (log-uniform 10k..500k particles) stepped together
    through the segmented cell grid, graph-captured, autotuned."""
    import numpy as np

    import sph_autotune

    ens = sph_scenes.config3(seed=0, n_scenes=64, scale=scale)
    offs = np.cumsum([0] + [sc.n for sc in ens]).astype(np.uint64)
    pos = np.concatenate([sc.pos for sc in ens])
    n = int(offs[-1])
    e = sphk.Engine(n, ens[0].params(sphk))
    e.upload(pos)
    e.set_scenes(offs, np.array([sc.bounds for sc in ens], np.float32))
    tune = sph_autotune.autotune(e, ens[0].dt, steps=2, use_cache=False)
    e.snapshot_save()
    e.set_graph(True)
    t = time_reset_steps(e, ens[0].dt, steps, 2)
    out = {"scenes": len(ens), "particles": n, "min_scene": int(min(sc.n for sc in ens)),
           "max_scene": int(max(sc.n for sc in ens)), "ms_per_step": t * 1e3,
           "value": n / t, "unit": "particle-updates/s", "graph": True,
           "walk_shapes": [tune["names"][tune["density"]], tune["names"][tune["force"]]]}
    del e
    return out


def full_run_line(sphk, scene, steps, dt=1e-4, every=100, damping=None, shapes=None):
    """`steps` consecutive steps of the 1M dam break under the consistent formulation (the
    reference formulas pile the block into the box corners within a step), timed in segments
    of `every` steps.  Between segments (untimed) the energies are read back: kinetic
    (get_total_energy), potential in the gravity field (which acts on y), and the internal
    energy of the equation of state P = k (rho - rho0): U = sum m k (ln(rho/rho0) + rho0/rho - 1),
    from de = P / rho^2 drho with e(rho0) = 0, at the summation density of the state.  Beside
    them, the reference's conservation check (compute_conservation_errors: |sum rho h^3 - N m|
    / N m, energy_error fixed at 0)."""
    import numpy as np

    params = scene.params(sphk, formulation=1)
    if damping is not None:
        params.damping = damping
    e = sphk.Engine(scene.n, params)
    if shapes is not None:
        e.set_walk_shapes(*shapes)
    e.upload(scene.pos, scene.vel)
    p = e.params()
    m = e.masses().astype(np.float64)
    g = -float(p.gravity)
    k, rho0 = float(p.gas_constant), float(p.rest_density)

    def energies():
        y = e.positions()[:, 1].astype(np.float64)
        rho = e.surrogate_densities().astype(np.float64)  # summation density of this state
        ke = float(e.total_energy())
        pe = float(np.sum(m * g * (y - float(p.ymin))))
        u = float(np.sum(m * k * (np.log(rho / rho0) + rho0 / rho - 1.0)))
        return ke, pe, u, rho

    traj = []
    t_steps = 0.0
    done = 0
    while True:
        ke, pe, u, rho = energies()
        row = {"step": done, "ke": ke, "pe": pe, "u": u, "e": ke + pe + u,
               "rho_mean": float(rho.mean()),
               "rho_max_dev_frac": float(np.abs(rho - rho0).max() / rho0)}
        if done:
            me, ee = e.conservation_errors()
            row.update(mass_error=float(me), energy_error=float(ee))
        traj.append(row)
        if done >= steps:
            break
        kk = min(every, steps - done)
        e.synchronize()
        t0 = time.perf_counter()
        for _ in range(kk):
            e.step(dt)
        e.synchronize()
        t_steps += time.perf_counter() - t0
        done += kk
    e0, e1 = traj[0]["e"], traj[-1]["e"]
    out = {
        "workload": f"{scene.name}: {steps} consecutive steps, consistent formulation, dt={dt}, "
                    f"damping {p.damping:g} per step",
        "steps": steps, "value": scene.n * steps / t_steps, "unit": "particle-updates/s",
        "ms_per_step": t_steps / steps * 1e3,
        "energy_drift_frac": (e1 - e0) / e0,
        "mass_exact": bool(np.array_equal(e.masses().astype(np.float64), m)),
        "reference_conservation_check": {"mass_error_final": traj[-1]["mass_error"],
                                         "energy_error_final": traj[-1]["energy_error"],
                                         "note": "the reference's check: |sum rho h^3 - N m| "
                                                 "/ N m, energy_error hard-wired to 0"},
        "trajectory": traj,
    }
    del e
    return out


def graph_line(sphk, sph_scenes, steps):
    """10,000-particle dam break, eager vs graph-captured step (launch-bound size)."""
    s = sph_scenes.config0()
    res = {}
    for graph in (False, True):
        e = sphk.Engine(s.n, s.params(sphk))
        e.upload(s.pos, s.vel)
        e.snapshot_save()
        e.set_graph(graph)
        t = time_reset_steps(e, s.dt, steps, 5)
        res["graph" if graph else "eager"] = {"ms_per_step": t * 1e3, "value": s.n / t}
        del e
    return res


def surrogate_line(sphk, sph_scenes, snapshots, macro_steps, judge_every=4, shapes=None):
    """Learned macro-step surrogate on the 1M dam break (consistent formulation, dt = 1e-4):
    trained on the engine's own trajectory, then timed against plain SPH over the same
    simulated time from the trained-to state.  Density errors are mean |rho - rho_SPH| in
    kg/m^3 (rest density 1000).

    The headline surrogate figure is the main run: judge runs every `judge_every`
    macro-steps measure the network's divergence per macro-step e and bound the run's density
    error against the plain run by (4/3) e (N - 1)^(3/2); the network runs only while that bound
    is within 1 % of the rest density, so the error is held by construction (and measured
    here).  The network alone (no judge, no bound) is timed beside it as a diagnostic, with its
    error against the plain run."""
    import numpy as np

    scene = sph_scenes.config1()
    dt = 1e-4
    e = sphk.Engine(scene.n, scene.params(sphk, formulation=1))
    if shapes is not None:
        e.set_walk_shapes(*shapes)
    e.upload(scene.pos, scene.vel)
    e.surrogate_init(seed=0)
    cfg = e.surrogate_config()
    sub = cfg.substeps
    n = scene.n
    t0 = time.perf_counter()
    curve = e.surrogate_train(snapshots=snapshots, dt=dt, batch_chunks=1024,
                              iters_per_snapshot=1, lr=1e-3, seed=1)
    e.synchronize()
    t_train = time.perf_counter() - t0
    trained = e.surrogate_params()
    e.surrogate_init(seed=0, config=sphk.surrogate_default_config(judge_every=judge_every))
    e.surrogate_set_params(trained)
    cfg = e.surrogate_config()
    e.snapshot_save()
    # warm both paths (first-use allocations), then restore; the plain step also counts the
    # neighbour lists of the trained-to state (edges of the network's layer 1)
    e.reset_stats()
    e.step(dt)
    st = e.stats()
    edges = float(st.total_neighbors - st.total_neighbor_queries)  # self excluded
    e.surrogate_macro_step(dt)
    e.snapshot_restore()
    e.synchronize()
    t0 = time.perf_counter()
    for _ in range(sub * macro_steps):
        e.step(dt)
    e.synchronize()
    t_plain = time.perf_counter() - t0
    rho_plain_run = e.surrogate_densities()
    e.snapshot_restore()
    e.synchronize()
    t0 = time.perf_counter()
    rep = e.surrogate_run(macro_steps, dt)
    e.synchronize()
    t_sur = time.perf_counter() - t0
    rho_sur_run = e.surrogate_densities()
    # the network alone over the same simulated time (no judge): timed, and its error
    e.snapshot_restore()
    e.synchronize()
    t0 = time.perf_counter()
    for _ in range(macro_steps):
        e.surrogate_macro_step(dt)
    e.synchronize()
    t_net = time.perf_counter() - t0
    rho_net_run = e.surrogate_densities()
    # a short judged run (8 macro-steps), for which the bound at the measured divergence stays
    # within the tolerance: the network runs between its judge runs (1 and 5)
    short = 8
    e.snapshot_restore()
    e.synchronize()
    t0 = time.perf_counter()
    for _ in range(sub * short):
        e.step(dt)
    e.synchronize()
    t_plain_s = time.perf_counter() - t0
    rho_plain_s = e.surrogate_densities()
    e.snapshot_restore()
    e.synchronize()
    t0 = time.perf_counter()
    rep_s = e.surrogate_run(short, dt)
    e.synchronize()
    t_sur_s = time.perf_counter() - t0
    d_s = np.abs(e.surrogate_densities() - rho_plain_s)
    drun = np.abs(rho_sur_run - rho_plain_run)
    dnet = np.abs(rho_net_run - rho_plain_run)
    ups_plain = n * sub * macro_steps / t_plain
    ups_sur = n * sub * macro_steps / t_sur
    ups_net = n * sub * macro_steps / t_net
    t_macro = t_net / macro_steps
    out = {
        "workload": f"dam_break_1M, consistent formulation, dt={dt}, macro-step = {sub} steps, "
                    f"judge every {cfg.judge_every} macro-steps, network only while the "
                    f"bound (4/3) e (N-1)^1.5 on the run's density error stays within "
                    f"{cfg.tolerance} kg/m^3",
        "value": ups_sur,
        "unit": "particle-updates/s",
        "plain_value": ups_plain,
        "speedup_vs_plain": ups_sur / ups_plain,
        "rho_error_vs_plain_run": {"mean": float(drun.mean()),
                                   "mean_frac_rest": float(drun.mean()) / 1000.0,
                                   "p99": float(np.quantile(drun, 0.99))},
        "ms_per_plain_step": t_plain / (sub * macro_steps) * 1e3,
        "macro_steps": macro_steps,
        "judged_run": {"network_macro_steps": int(rep.macro_steps),
                       "plain_macro_steps": int(rep.plain_macro_steps),
                       "judges": int(rep.judges), "fell_back": bool(rep.fell_back),
                       "fallback_at": int(rep.fallback_at) if rep.fell_back else None,
                       "error_bound": rep.error_bound,
                       "divergence_allowing_the_network": cfg.tolerance / (
                           4.0 / 3.0 * (macro_steps - 1) ** 1.5),
                       "judge_mean_rho_error_max": rep.max_error,
                       "judge_max_particle_rho_error": rep.max_abs_error},
        "short_run": {"macro_steps": short, "value": n * sub * short / t_sur_s,
                      "speedup_vs_plain": t_plain_s / t_sur_s,
                      "network_macro_steps": int(rep_s.macro_steps),
                      "judges": int(rep_s.judges), "fell_back": bool(rep_s.fell_back),
                      "error_bound": rep_s.error_bound,
                      "rho_error_vs_plain_run": {"mean": float(d_s.mean()),
                                                 "mean_frac_rest": float(d_s.mean()) / 1000.0}},
        "network_only": {"value": ups_net, "speedup_vs_plain": ups_net / ups_plain,
                         "ms_per_macro_step": t_macro * 1e3,
                         "rho_error_vs_plain_run": {
                             "mean": float(dnet.mean()),
                             "mean_frac_rest": float(dnet.mean()) / 1000.0,
                             "p99": float(np.quantile(dnet, 0.99))},
                         "note": "no judge and no bound: a diagnostic, not the surrogate figure"},
        "training": {"snapshots": snapshots, "batch_particles": 1024 * 64, "seconds": t_train,
                     "loss_first": float(curve[0]), "loss_last": float(curve[-20:].mean())},
        "dtype": "bf16 MFMA operands, f32 accumulation",
        # useful MFMA work of one network evaluation: layer 1 over every edge (K = 10 of the
        # 16-wide operand: 9 features + bias) and the node MLP per particle
        "mfma": {"useful_flop_per_macro_step": 2.0 * edges * 10 * 128 +
                 2.0 * n * (128 * 128 + 132 * 128 + 3 * 128),
                 "achieved_tflops_over_macro_step": (2.0 * edges * 10 * 128 + 2.0 * n * (
                     128 * 128 + 132 * 128 + 3 * 128)) / t_macro / 1e12,
                 "peak_tflops": BF16_PEAK_TFLOPS},
    }
    out["mfma"]["frac"] = out["mfma"]["achieved_tflops_over_macro_step"] / BF16_PEAK_TFLOPS
    del e
    return out


def hip_copy_calibration(mib):
    """Three device-to-device copies of `mib` MiB through the HIP runtime libsphk uses: the
    reference launches whose FETCH_SIZE / WRITE_SIZE counters scale the walks' counters to
    bytes."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so.7")
    nbytes = C.c_size_t(mib << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), nbytes) == 0 and hip.hipMalloc(C.byref(dst), nbytes) == 0
    assert hip.hipMemset(src, 0, nbytes) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, nbytes, 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design only)", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import sph_scenes
    import sphk

    scene = sph_scenes.CONFIGS[args.config]()
    e = sphk.Engine(scene.n, scene.params(sphk))
    e.upload(scene.pos, scene.vel)
    tune = None
    if args.walk_shapes:
        d, f = (int(x) for x in args.walk_shapes.split(","))
        e.set_walk_shapes(d, f)
    elif not args.no_autotune:
        import sph_autotune

        tune = sph_autotune.autotune(e, scene.dt, steps=3, use_cache=False)
    e.snapshot_save()
    if args.graph:
        e.set_graph(True)

    def one_step():
        e.snapshot_restore()
        e.step(scene.dt)

    for _ in range(args.warmup):
        one_step()
    e.synchronize()
    e.reset_stats()
    # the timed steps run as a user runs them, without phase events: each event record idles
    # the GPU ~4.5 us (four per step, ~1.5 % at 1M).  The phase split and the roofline's kernel
    # time come from as many steps right after with the engine's HIP events around each phase
    # (eagerly: a graph replay records none), and that pass's wall time is reported beside.
    e.set_timing(False)
    e.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_step()
    e.synchronize()
    elapsed = time.perf_counter() - t0
    e.set_graph(False)
    e.set_timing(True)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        one_step()
    e.synchronize()
    elapsed_events = time.perf_counter() - t1
    st = e.stats()
    e.set_timing(False)

    k = max(st.timed_steps, 1)
    t_force = st.force_computation_time / k
    t_dens = st.density_computation_time / k
    t_nbr = st.neighbor_search_time / k
    n = scene.n
    achieved = ALG_BYTES["force"] * n / t_force / 1e9
    mean_nbrs = st.total_neighbors / max(st.total_neighbor_queries, 1)
    ups = n * args.steps / elapsed
    # FP32 pair arithmetic of the two walks (the candidate filter is integer work)
    flops_per_update = FLOP_PER_PAIR * mean_nbrs
    force_shape = sphk.walk_shape_names()[e.walk_shapes()[1]]
    counters = _load_counters(force_shape) if args.config == 1 else None
    out = {
        "metric": METRIC,
        "value": ups,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{scene.name}: plain SPH dam break, {n} particles, FP32, reference "
                        f"formulas, dt={scene.dt}, reset-and-step",
            "particles": n,
            "protocol": "reset-and-step (device-side restore of the initial state inside the "
                        "timed region)",
            "mean_neighbors": mean_nbrs,
            "max_neighbors": int(st.max_neighbors),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_force_walk",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None if counters is None else counters["hbm_bytes_per_launch"],
            "traffic_source": None if counters is None else
            {"profile": os.path.relpath(COUNTER_PROFILE, REPO), "build": counters["build"],
             "force_shape": counters["force_shape"]},
            "alg_bytes_per_update": ALG_BYTES["force"],
            "kernel_ms": t_force * 1e3,
            "limiter": "VALU issue and gather latency (HBM traffic is far below its peak)",
            "valu_busy_frac": None if counters is None else counters["valu_busy_frac"],
        },
        "valu": {
            "fp32_flop_per_update": flops_per_update,
            "achieved_tflops": flops_per_update * ups / 1e12,
            "peak_tflops": FP32_PEAK_TFLOPS,
            "frac": flops_per_update * ups / 1e12 / FP32_PEAK_TFLOPS,
            "peak_unpacked_tflops": FP32_PEAK_UNPACKED_TFLOPS,
            "frac_unpacked": flops_per_update * ups / 1e12 / FP32_PEAK_UNPACKED_TFLOPS,
        },
        "phase_ms": {"neighbor_search": t_nbr * 1e3, "density": t_dens * 1e3,
                     "force_integrate": t_force * 1e3, "step_events": st.total_time / k * 1e3,
                     "wall_ms_per_step_with_events": elapsed_events / args.steps * 1e3},
        "step_hbm_frac": ALG_BYTES["step"] * ups / 1e9 / HBM_PEAK_GBS,
        "walk_shapes": {"density": sphk.walk_shape_names()[e.walk_shapes()[0]],
                        "force": sphk.walk_shape_names()[e.walk_shapes()[1]],
                        "tuning_ms": None if tune is None else
                        {"density": tune["density_ms"], "force": tune["force_ms"]}},
    }
    if not args.no_extras:
        if args.config == 1:
            out["drop_into_pool_9.6M"] = pool_line(sphk, sph_scenes, e.walk_shapes())
        out["ensemble"] = ensemble_line(sphk, sph_scenes, args.ensemble_scale, 5)
        out["small_scene_10k"] = graph_line(sphk, sph_scenes, 50)
    if args.full_run:
        out["full_run"] = full_run_line(sphk, scene, args.full_run, shapes=e.walk_shapes())
        undamped = sph_scenes.CONFIGS[args.config]()
        out["full_run_undamped"] = full_run_line(sphk, undamped, args.full_run, damping=1.0,
                                                  shapes=e.walk_shapes())
    if not args.no_surrogate and args.config == 1:
        out["surrogate"] = surrogate_line(sphk, sph_scenes, args.surrogate_snapshots,
                                          args.surrogate_macro, shapes=e.walk_shapes())
    if args.calibrate_copy:
        hip_copy_calibration(args.calibrate_copy)
    if not args.no_cpu_baseline:
        del e
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_repeats, sph_scenes,
                                           legs=not args.no_cpu_legs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
