#!/usr/bin/env python3
"""Throughput of the force engine's hot path on one MI355X.

One step = one DKD leapfrog step of the whole synthetic cone plume
(N = 1,048,576 ions, 10 um long, 20 deg, lognormal radius, 5 % -1e) with FP64
state: the FMM force (FP32 expansions and pairs, compensated sums, order p=10,
near field two leaf cells wide), the pwl background field and the fused
kick/drift update.  Inputs are resident in HBM when the timed region starts.

Also measured, on the same state: the force error of the FMM against the GPU
FP64 direct sum on a seeded 8192-ion subsample (relative L2 as
integration_schemes.cpp:411-455 defines it), one FP64 direct leapfrog step
(the speed-up denominator), per-pass FMM kernel times, and a CPU baseline (the
oracle's FP64 direct sum on a bounded target sample), and the relative drift of
the total energy (kinetic + pair potential + field work) over the timed steps.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "solveit-code_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector peak (dense)
FLOP_PER_PAIR = 20


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE of rocprofv3
    --pmc runs of this bench, scaled by the counters' reading for a device copy
    of known size), from profiles/traffic.json; None when not profiled."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    return None if t is None else t["fetch_bytes"] + t["write_bytes"]


def copy_calibration(nbytes=256 << 20):
    """A plain device copy of nbytes (torch copy kernel), run before the
    workload so that --pmc runs of this bench can scale FETCH_SIZE and
    WRITE_SIZE by what they read for a known number of bytes."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def cpu_baseline(d, n_targets, threads=None):
    """The oracle's FP64 direct sum (the reference's direct method restated,
    one target per loop iteration) on the host cores: n_targets targets of the
    same plume on one thread, then threads x n_targets targets split over
    `threads` threads (ctypes releases the GIL), each scaled to
    particle-steps/s of a full N x N direct step."""
    from concurrent.futures import ThreadPoolExecutor
    import oracle
    n = d["x"].shape[0]
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    oracle.direct_f64(d, i0=0, i1=n_targets)
    dt1 = time.perf_counter() - t0
    rate1 = n_targets * n / dt1
    nt = min(n, threads * n_targets)
    bounds = [(k * nt // threads, (k + 1) * nt // threads) for k in range(threads)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda b: oracle.direct_f64(d, i0=b[0], i1=b[1]), bounds))
    dtm = time.perf_counter() - t0
    ratem = nt * n / dtm
    return {"value": ratem / n, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "value_1core": rate1 / n,
            "sample": f"oracle FP64 direct sum over {n} sources of the same plume: {n_targets} "
                      f"targets on 1 thread ({dt1:.1f} s, {rate1:.3e} pairs/s) and {nt} targets "
                      f"on {threads} threads ({dtm:.1f} s, {ratem:.3e} pairs/s), scaled to one "
                      f"N x N step"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--order", type=int, default=10)
    ap.add_argument("--leaf-size", type=int, default=20)
    ap.add_argument("--dt", type=float, default=1e-12)
    ap.add_argument("--cpu-targets", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 256 MiB device copy first (reference kernel for --pmc byte counts)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import solveit_hip
    from solveit_hip import synth

    if args.copy_calibration:
        copy_calibration()
    n = args.n
    d = synth.cone(n)
    mesh = synth.tip_field_mesh()
    eng = solveit_hip.Engine(0)
    eng.efield_setup(mesh)
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    acc = tuple(torch.zeros(n, dtype=torch.float64, device="cuda:0") for _ in range(3))
    opts = solveit_hip.make_step_opts(method=solveit_hip.SHP_METHOD_FMM, use_efield=True,
                                      fmm_order=args.order, leaf_size=args.leaf_size, fmm_ring=2)
    t = 0.0
    eng.leapfrog_run(st, opts, t, args.dt, args.warmup, acc=acc)
    t += args.warmup * args.dt
    eng.sync()
    torch.cuda.synchronize()
    # Total energy at the start of the timed steps (FP64 pair sum for PE).
    x0, y0, z0 = st.x.clone(), st.y.clone(), st.z.clone()
    ke0, pe0 = eng.energy(st)
    eng.sync()
    torch.cuda.synchronize()
    eng.fmm_timing(reset=True)
    t0 = time.perf_counter()
    eng.leapfrog_run(st, opts, t, args.dt, args.steps, acc=acc)
    eng.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t += args.steps * args.dt
    ms_per_step = wall * 1e3 / args.steps
    value = n * args.steps / wall
    # Per-pass HIP-event times and counters of the FMM calls of the timed steps.
    tm = eng.fmm_timing(reset=True)
    calls = max(1, tm["calls"])
    mean = {k: tm[k] / calls for k in solveit_hip.FMM_STAT_KEYS}

    # Energy drift over the timed steps: KE + PE + work of the background
    # field along the reference's z-then-radial path from the start positions.
    ke1, pe1 = eng.energy(st)
    fe = eng.field_energy(x0, y0, z0, st.x, st.y, st.z, st.q, nsteps=4 * args.steps)
    dfe = float(fe[(st.status == 1) | (st.status == 2)].sum().item())
    e_start = ke0 + pe0
    drift = (ke1 + pe1 + dfe - e_start) / abs(e_start)

    # Accuracy of the FMM force against the GPU FP64 direct sum on this state.
    a_f = torch.stack(eng.fmm_dev(st, order=args.order, leaf_size=args.leaf_size, ring=2), 1)
    eng.sync()
    fp64_step_ms = None
    l2 = l2max = l2_all = None
    if not args.no_fp64:
        a_d = torch.stack(eng.direct_dev(st, precision=solveit_hip.SHP_PREC_FP64), 1)
        eng.sync()
        # Integrator::compareAcceleration's metric (GPU, FP64) on a seeded
        # 8192-ion target subsample and on every ion
        sub = torch.as_tensor(np.sort(np.random.default_rng(8192).choice(n, size=min(8192, n),
                                                                         replace=False)),
                              device="cuda:0")
        m_sub = eng.compare_accel(st.status[sub].contiguous(),
                                  tuple(a_f[sub, k].contiguous() for k in range(3)),
                                  tuple(a_d[sub, k].contiguous() for k in range(3)))
        m_all = eng.compare_accel(st.status, tuple(a_f[:, k].contiguous() for k in range(3)),
                                  tuple(a_d[:, k].contiguous() for k in range(3)))
        l2, l2max, l2_all = m_sub["l2"], m_sub["max"], m_all["l2"]
        opts64 = solveit_hip.make_step_opts(method=solveit_hip.SHP_METHOD_DIRECT,
                                            precision=solveit_hip.SHP_PREC_FP64, use_efield=True)
        eng.sync()
        t1 = time.perf_counter()
        eng.leapfrog_run(st, opts64, t, args.dt, 1, acc=acc)
        eng.sync()
        fp64_step_ms = (time.perf_counter() - t1) * 1e3

    m2l_tf = 2.0 * mean["m2l_madds"] / (mean["m2l_ms"] * 1e-3) / 1e12
    leaf_tf = FLOP_PER_PAIR * mean["p2p_pairs"] / (mean["leaf_ms"] * 1e-3) / 1e12
    # Dominant kernel of the timed steps: the leaf kernel (P2P + L2P;
    # algorithmic flops = 20 per near-field pair) or the first M2L pass (output
    # degrees 0..4; 2 flops per multiply-add of the box pairs that exist),
    # whichever has the longer mean launch time.
    if mean["leaf_ms"] >= mean["m2l_pass0_ms"]:
        kname, kms, kflop = "k_fmm_leaf", mean["leaf_ms"], FLOP_PER_PAIR * mean["p2p_pairs"]
    else:
        kname, kms, kflop = ("k_fmm_m2l_tiled<10,8,7,0,4>", mean["m2l_pass0_ms"],
                             2.0 * mean["m2l_pass0_madds"])
    roof = {"kernel": kname, "achieved": kflop / (kms * 1e-3) / 1e12, "kernel_ms": kms,
            "flop_per_launch": kflop, "bound": "fp32-vector", "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "traffic": profiled_traffic(kname)}
    roof["frac"] = roof["achieved"] / FP32_PEAK_TFLOPS

    line = {
        "metric": "particle-steps/s at N=1M, FP32-comp FMM vs FP64 direct at rel L2 err<=1e-5",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"cone plume N={n} (10 um, 20 deg, lognormal r sigma 0.5 x 100 nm, "
                        "5% -1e), DKD leapfrog step with FMM force (FP32 expansions and pairs, "
                        "compensated sums) + pwl tip field, FP64 state",
            "n": n,
            "method": "fmm",
            "order": args.order,
            "leaf_level": int(round(mean["leaf_level"])),
            "leaf_level_mean": mean["leaf_level"],
            "near_ring": 2,
            "rel_l2_vs_fp64_direct": l2,
            "max_rel_vs_fp64_direct": l2max,
            "rel_l2_all_ions_vs_fp64_direct": l2_all,
            "fp64_direct_step_ms": fp64_step_ms,
            "speedup_vs_fp64_direct": (fp64_step_ms / ms_per_step) if fp64_step_ms else None,
            "pass_ms": {k: mean[k] for k in ("p2m_m2m_ms", "ttable_ms", "m2l_ms", "l2l_ms",
                                             "leaf_ms", "m2l_pass0_ms")},
            "fmm_calls_timed": tm["calls"],
            "energy_drift_rel": drift,
            "energy_start_J": {"kinetic": ke0, "potential": pe0},
            "m2l_pairs": mean["m2l_pairs"],
            "p2p_pairs": mean["p2p_pairs"],
            "m2l_tflops": m2l_tf,
            "p2p_tflops": leaf_tf,
        },
        "roofline": roof,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(d, args.cpu_targets)
    eng.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
