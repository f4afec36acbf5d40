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
(the speed-up denominator), per-pass FMM kernel times, the relative drift of
the total energy (kinetic + pair potential + field work) over the timed steps,
and a CPU baseline: the oracle's FP64 restatement of the same FMM on the host
cores (bounded sample: every 16th leaf box's near field), beside the oracle's
FP64 direct sum on a bounded target sample.
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


def host_cpu_info():
    """nproc, the cores this process may run on, the box's CPU share used for
    worker threads, and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = nproc
    # Worker threads: the cores available to this process, at most the host's
    # CPU share per GPU (16 on these boxes; OMP_NUM_THREADS states it).
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or avail
    return {"nproc": nproc, "available": avail, "threads": max(1, min(avail, share)),
            "model": model}


def cpu_baseline(d, n_targets, fmm_args=None):
    """CPU legs on the host cores, on bounded samples of the timed workload.

    fmm (value, like for like): the oracle's FP64 restatement of the same FMM
    (same order, shell orders, near ring and leaf level) on the plume state at
    the start of the timed steps, OpenMP over the available cores: the tree and
    far-field passes over every box, L2P + P2P on every k-th leaf box, scaled
    to all leaf boxes.  direct: the oracle's FP64 direct sum (the reference's
    direct method restated) on n_targets targets on 1 thread and on
    threads x n_targets targets, scaled to one N x N step."""
    from concurrent.futures import ThreadPoolExecutor
    import oracle
    info = host_cpu_info()
    threads = info["threads"]
    n = d["x"].shape[0]
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
    out = {"unit": "particle-steps/s", "cores": threads, "kind": "port",
           "nproc": info["nproc"], "cores_available": info["available"], "cpu_model": info["model"],
           "direct_value": ratem / n, "direct_value_1core": rate1 / n,
           "direct_sample": f"oracle FP64 direct sum over {n} sources: {n_targets} targets on 1 thread "
                            f"({dt1:.1f} s, {rate1:.3e} pairs/s) and {nt} targets on {threads} threads "
                            f"({dtm:.1f} s, {ratem:.3e} pairs/s), scaled to one N x N step"}
    if fmm_args is None:
        out["value"] = out["direct_value"]
        out["sample"] = out["direct_sample"]
        return out
    order, leaf_level, ring, shells, stride = fmm_args
    oracle.set_threads(threads)
    oracle.fmm_set_shell_orders(shells)
    try:
        _, st = oracle.fmm_f64(d, p=order, L=leaf_level, ring=ring, leaf_stride=stride)
    finally:
        oracle.fmm_set_shell_orders(None)
        oracle.set_threads(0)
    scale = st["leaves"] / max(1, st["leaves_evaluated"])
    t_full = st["far_s"] + st["near_s"] * scale
    out["value"] = float(n / t_full)
    out["sample"] = (f"oracle FP64 FMM (p={order}, shell orders {shells[3:6]}, ring {ring}, leaf level "
                     f"{leaf_level}) on the {n}-ion plume at the start of the timed steps, {threads} "
                     f"OpenMP threads: tree + P2M/M2M/M2L/L2L over all boxes {st['far_s']:.1f} s, "
                     f"L2P + P2P on every {stride}th of {st['leaves']} leaf boxes {st['near_s']:.1f} s "
                     f"({st['p2p_pairs']:.3e} pairs), scaled to all leaf boxes: {t_full:.1f} s per step")
    out["fmm_far_s"] = st["far_s"]
    out["fmm_near_s_sampled"] = st["near_s"]
    return out


def _timed(fn, reps=2):
    """Wall ms of fn() (synchronized), the last of `reps` calls, and its result."""
    import torch
    out = None
    ms = 0.0
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
    return ms, out


def _metric(eng, st, a, ref, sub=None):
    """compareAcceleration's rel L2 / max (GPU, FP64), on all ions or a subsample."""
    import torch
    if sub is None:
        m = eng.compare_accel(st.status, tuple(a), tuple(ref))
    else:
        m = eng.compare_accel(st.status[sub].contiguous(), tuple(x[sub].contiguous() for x in a),
                              tuple(x[sub].contiguous() for x in ref))
    return {"rel_l2": m["l2"], "max_rel": m["max"], "n": m["n"]}


def _energy_total(eng, st, x0, y0, z0, nsteps_field):
    """KE + PE + field work from (x0, y0, z0) (J)."""
    ke, pe = eng.energy(st)
    fe = 0.0
    if nsteps_field:
        f = eng.field_energy(x0, y0, z0, st.x, st.y, st.z, st.q, nsteps=nsteps_field)
        fe = float(f[(st.status == 1) | (st.status == 2)].sum().item())
    return ke + pe + fe


def run_config(k, args):
    """One JSON line for workload k measured on this GPU: 1 the FP64 direct
    leapfrog of a 4096-ion hemisphere with the tip field, 2 FP32C against FP64
    direct on 65,536- and 262,144-ion cones, 3 direct / Barnes-Hut / FMM on
    the 1M-ion cone, 4 FMM with symmetry images at 262,144 ions, 5 the
    batched adjoint of 64 plumes; the comment of each branch says what is
    timed."""
    import numpy as np
    import torch
    import solveit_hip
    from solveit_hip import synth
    eng = solveit_hip.Engine(0)
    P = solveit_hip
    res = {"config": k}

    def log(msg):
        print(f"[config {k}] {msg}", file=sys.stderr, flush=True)

    if k == 1:
        # FP64 direct, N=4096 hemisphere, 1000 DKD steps at dt=1e-12 s with the
        # synthetic tip field: particle-steps/s and relative energy drift.
        d = synth.hemisphere(4096)
        eng.efield_setup(synth.tip_field_mesh())
        st = P.DeviceState.from_numpy(d)
        x0, y0, z0 = st.x.clone(), st.y.clone(), st.z.clone()
        e0 = _energy_total(eng, st, x0, y0, z0, 0)
        opts = P.make_step_opts(method=P.SHP_METHOD_DIRECT, precision=P.SHP_PREC_FP64, use_efield=True)
        nsteps = 1000
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.leapfrog_run(st, opts, 0.0, 1e-12, nsteps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        e1 = _energy_total(eng, st, x0, y0, z0, 64)
        res.update(workload="hemisphere 20 nm N=4096, FP64 direct + pwl tip field, 1000 DKD steps dt=1e-12 s",
                   particle_steps_per_s=4096 * nsteps / wall, ms_per_step=wall * 1e3 / nsteps,
                   energy_drift_rel=(e1 - e0) / abs(e0))
        import oracle
        t0 = time.perf_counter()
        oracle.leapfrog_run_f64(d, 0.0, 1e-12, 20, mesh=synth.tip_field_mesh())
        dtc = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": 4096 * 20 / dtc, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                               "sample": f"oracle FP64 DKD leapfrog with the pwl field, 20 steps ({dtc:.1f} s)"}
    elif k == 2:
        # FP32C direct at N=65536 and 262144 (cone): one force evaluation (kernel
        # time, pair rate, fraction of the FP32 peak, rel L2 vs the GPU FP64
        # path), then 100 DKD steps with FP32C and with FP64 forces from the
        # same state: energy drift of each and the final position difference.
        out = []
        for n in (65536, 262144):
            d = synth.cone(n)
            st = P.DeviceState.from_numpy(d)
            nsrc = int(((st.status == 1) | (st.status == 2)).sum().item())
            nt = int((st.status == 1).sum().item())
            ms32, a32 = _timed(lambda: eng.direct_dev(st, precision=P.SHP_PREC_FP32C))
            k32 = eng.last_kernel_ms()
            ms64, a64 = _timed(lambda: eng.direct_dev(st, precision=P.SHP_PREC_FP64))
            k64 = eng.last_kernel_ms()
            acc = _metric(eng, st, a32, a64)
            pairs = float(nt) * nsrc
            row = {"n": n, "fp32c_kernel_ms": k32, "fp64_kernel_ms": k64,
                   "fp32c_pairs_per_s": pairs / (k32 * 1e-3),
                   "fp32c_tflops_20flop": FLOP_PER_PAIR * pairs / (k32 * 1e-3) / 1e12,
                   "fp32c_frac_of_157": FLOP_PER_PAIR * pairs / (k32 * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                   "fp64_tflops_20flop": FLOP_PER_PAIR * pairs / (k64 * 1e-3) / 1e12,
                   "rel_l2_fp32c_vs_fp64": acc["rel_l2"], "max_rel": acc["max_rel"]}
            drift = {}
            fin = {}
            for name, prec in (("fp32c", P.SHP_PREC_FP32C), ("fp64", P.SHP_PREC_FP64)):
                s2 = P.DeviceState.from_numpy(d)
                x0, y0, z0 = s2.x.clone(), s2.y.clone(), s2.z.clone()
                e0 = _energy_total(eng, s2, x0, y0, z0, 0)
                opts = P.make_step_opts(method=P.SHP_METHOD_DIRECT, precision=prec)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                eng.leapfrog_run(s2, opts, 0.0, 1e-12, 100)
                torch.cuda.synchronize()
                row[f"{name}_100_steps_ms"] = (time.perf_counter() - t0) * 1e3
                drift[name] = (_energy_total(eng, s2, x0, y0, z0, 0) - e0) / abs(e0)
                fin[name] = torch.stack([s2.x, s2.y, s2.z], 1)
            row["energy_drift_rel"] = drift
            import oracle
            t0 = time.perf_counter()
            oracle.direct_f64(d, i0=0, i1=1024)
            dtc = time.perf_counter() - t0
            row["cpu_baseline"] = {"pairs_per_s": 1024 * nsrc / dtc, "cores": 1, "kind": "port",
                                   "particle_steps_per_s": 1024 * nsrc / dtc / n,
                                   "sample": f"oracle FP64 direct sum, 1024 targets ({dtc:.1f} s)"}
            row["final_pos_rel_diff_fp32c_vs_fp64"] = float(
                (fin["fp32c"] - fin["fp64"]).norm() / (fin["fp64"] - torch.stack([
                    torch.as_tensor(d["x"]), torch.as_tensor(d["y"]), torch.as_tensor(d["z"])],
                    1).to(fin["fp64"].device)).norm())
            out.append(row)
            log(json.dumps(row))
        res.update(workload="cone plume, FP32C direct vs GPU FP64 direct", runs=out)
    elif k == 3:
        # N=1M cone: Barnes-Hut theta 0.3/0.5/0.7, the reference FMM's operators
        # and the engine's FMM at p=4/6/8/10 (leaf size 64): ms per force
        # evaluation and rel L2 vs FP64 direct on the seeded 8192-target
        # subsample; the engine's FMM also against the reference operators at
        # the same order.
        n = 1048576
        d = synth.cone(n)
        st = P.DeviceState.from_numpy(d)
        ms64, a64 = _timed(lambda: eng.direct_dev(st, precision=P.SHP_PREC_FP64), reps=1)
        sub = torch.as_tensor(np.sort(np.random.default_rng(8192).choice(n, 8192, replace=False)),
                              device="cuda:0")
        rows = [{"method": "fp64_direct", "ms": ms64}]
        for th in (0.3, 0.5, 0.7):
            ms, a = _timed(lambda: eng.bh_dev(st, theta=th))
            walk = eng.last_kernel_ms()  # k_bh_walk alone
            bs = eng.bh_stats()
            inter = bs["accepted_cells"] + bs["ion_terms"]
            visits = inter + bs["opened_cells"]
            rows.append({"method": "bh", "theta": th, "ms": ms, "walk_ms": walk,
                         "accepted_cells": bs["accepted_cells"], "ion_terms": bs["ion_terms"],
                         "opened_cells": bs["opened_cells"], "max_depth": bs["max_depth"],
                         "interactions_per_s": inter / (walk * 1e-3), "node_visits_per_s": visits / (walk * 1e-3),
                         **_metric(eng, st, a, a64, sub)})
            log(json.dumps(rows[-1]))
        # The reference FMM's own operators (spherical harmonics, rotation-based
        # translations, FP64, M2L on the matrix cores; shp_fmm_sph_host) on the
        # plume's float AoS with the reference's level choice: ms per call with
        # the host arrays in and out, rel L2 against the FP64 direct sum of the
        # same float inputs (the reference's direct contract, shp_direct_host).
        pos4, vel4, acc4 = synth.to_reference_aos(d)
        ms_ref, a_ref = _timed(lambda: eng.direct_host(pos4, vel4, acc4), reps=1)
        subn = sub.cpu().numpy()
        tgt = acc4[subn, 3] == 1.0

        def rel_l2_np(a, b):
            e = np.linalg.norm(a - b, axis=1) ** 2 / np.linalg.norm(b, axis=1) ** 2
            return float(np.sqrt(np.mean(e))), float(np.sqrt(np.max(e)))

        sph = {}
        for p in (4, 6, 8, 10):
            ms, a = _timed(lambda: eng.fmm_sph_host(pos4, vel4, acc4, order=p))
            sph[p] = a
            l2, mx = rel_l2_np(a[subn][tgt], a_ref[subn][tgt])
            rows.append({"method": "fmm_reference_operators", "order": p, "ms_host_in_out": ms,
                         **eng.fmm_sph_stats(), "rel_l2": l2, "max_rel": mx,
                         "vs": "FP64 direct of the same float inputs"})
            log(json.dumps(rows[-1]))
        rows[0]["ref_contract_direct_ms"] = ms_ref
        for p in (4, 6, 8, 10):
            ms, a = _timed(lambda: eng.fmm_dev(st, order=p, leaf_size=64, ring=2))
            at = torch.stack(a, 1)[sub].cpu().numpy()
            rows.append({"method": "fmm", "order": p, "ms": ms, "leaf_level": eng.fmm_leaf_level(),
                         **_metric(eng, st, a, a64, sub),
                         "rel_l2_vs_reference_operators_same_order": rel_l2_np(at[tgt], sph[p][subn][tgt])[0]})
            log(json.dumps(rows[-1]))
        # CPU legs (one thread, bounded): the oracle's Barnes-Hut (reference
        # arithmetic) on a whole 65536-ion cone of the same shape, with the GPU
        # on the same input beside it, and the oracle's FP64 FMM (the engine's
        # algorithm) on a 131072-ion cone.
        import oracle
        dbh = synth.cone(65536)
        t0 = time.perf_counter()
        oracle.bh_ref(*synth.to_reference_aos(dbh), theta=0.5)
        tb = time.perf_counter() - t0
        stb = P.DeviceState.from_numpy(dbh)
        msb, _ = _timed(lambda: eng.bh_dev(stb, theta=0.5))
        dsm = synth.cone(131072)
        oracle.fmm_set_shell_orders([0, 0, 0, 10, 8, 7])
        oracle.set_threads(1)  # one host thread, like the Barnes-Hut leg
        try:
            t0 = time.perf_counter()
            oracle.fmm_f64(dsm, 10, 6, ring=2)
            tf = time.perf_counter() - t0
        finally:
            oracle.set_threads(0)
            oracle.fmm_set_shell_orders(None)
        msf, _ = _timed(lambda: eng.fmm_dev(P.DeviceState.from_numpy(dsm), order=10, leaf_level=6, ring=2))
        res["cpu_baseline"] = {
            "bh_theta0.5_s_65536": tb, "gpu_bh_theta0.5_ms_65536": msb,
            "fmm_p10_s_131072": tf, "gpu_fmm_p10_ms_131072": msf,
            "cores": 1, "kind": "port",
            "sample": "oracle Barnes-Hut (reference arithmetic) on a whole 65536-ion cone; oracle FP64 FMM "
                      "(p 10/8/7, ring 2, leaf level 6) on a 131072-ion cone; the GPU times on the same inputs"}
        res.update(workload="cone plume N=1048576 (one force evaluation each)", runs=rows)
    elif k == 4:
        # N=262144 per cell/sector: translational FMM (3x3x3 shell of 2 um
        # cells, uniform fill + wrapped cone) and rotational FMM (6 sectors,
        # sector-folded cone), vs FP64 direct over the same images.
        n = 262144
        rows = []
        for kind in ("translational", "rotational"):
            if kind == "translational":
                d = d_cell = synth.cell_fill(n)
                sym = sym_cell = P.make_symmetry("translational", lattice=synth.full_lattice(1, (1, 1, 1)),
                                                 domain=(2e-6, 2e-6, 2e-6))
                nimg = 27
            else:
                d = synth.fold_sectors(synth.cone(n), 6)
                sym = P.make_symmetry("rotational", nsec=6)
                nimg = 6
            st = P.DeviceState.from_numpy(d)
            ms64, a64 = _timed(lambda: eng.direct_dev(st, precision=P.SHP_PREC_FP64, symmetry=sym), reps=1)
            ms32, a32 = _timed(lambda: eng.direct_dev(st, precision=P.SHP_PREC_FP32C, symmetry=sym), reps=1)
            msf, af = _timed(lambda: eng.fmm_dev(st, order=10, leaf_size=20, ring=2, symmetry=sym))
            rows.append({"kind": kind, "n": n, "images": nimg, "fmm_ms": msf, "fp64_direct_ms": ms64,
                         "fp32c_direct_ms": ms32, "fmm_speedup_vs_fp64": ms64 / msf,
                         "fmm": _metric(eng, st, af, a64), "fp32c": _metric(eng, st, a32, a64)})
            log(json.dumps(rows[-1]))
        # The reference's rotational FMM with its own operators (4 quadrants,
        # shp_fmm_sph_host) on the cone folded into the first quadrant, against
        # the reference-contract direct sum over the same 4 images, with the
        # engine's FMM on the same ions and images beside it.
        d4 = synth.fold_sectors(synth.cone(n), 4)
        pos4, vel4, acc4 = synth.to_reference_aos(d4)
        sym4 = P.make_symmetry("rotational", nsec=4)
        ms_ref4, a_ref4 = _timed(lambda: eng.direct_host(pos4, vel4, acc4, symmetry=sym4), reps=1)
        tgt4 = acc4[:, 3] == 1.0

        def rel_l2_np(a, b):
            e = np.linalg.norm(a - b, axis=1) ** 2 / np.linalg.norm(b, axis=1) ** 2
            return float(np.sqrt(np.mean(e))), float(np.sqrt(np.max(e)))

        for p in (4, 10):
            ms, a = _timed(lambda: eng.fmm_sph_host(pos4, vel4, acc4, order=p, symmetry=sym4))
            l2, mx = rel_l2_np(a[tgt4], a_ref4[tgt4])
            rows.append({"kind": "rotational_reference_operators", "n": n, "images": 4, "order": p,
                         "ms_host_in_out": ms, **eng.fmm_sph_stats(), "rel_l2": l2, "max_rel": mx,
                         "ref_contract_direct_ms": ms_ref4, "vs": "reference-contract direct sum over the 4 images"})
            log(json.dumps(rows[-1]))
        ms, a = _timed(lambda: eng.fmm_host(pos4, vel4, acc4, order=10, leaf_size=20, ring=2, symmetry=sym4))
        l2, mx = rel_l2_np(a[tgt4], a_ref4[tgt4])
        rows.append({"kind": "rotational_engine_fmm_4_quadrants", "n": n, "images": 4, "order": 10,
                     "ms_host_in_out": ms, "rel_l2": l2, "max_rel": mx,
                     "vs": "reference-contract direct sum over the 4 images"})
        log(json.dumps(rows[-1]))
        # The reference's translational FMM with its own operators (ternary tree
        # of the 2 um cell, lists wrapped periodically) on the cell fill, against
        # the reference-contract direct sum over the 27 images.  A float
        # coordinate that rounds onto the +face of [-L/2, L/2) is the same ion
        # one period lower.
        posc, velc, accc = synth.to_reference_aos(d_cell)
        cell = np.float32(2e-6)
        bmin = -cell / np.float32(2)
        on_face = (posc[:, :3] - bmin) / cell >= np.float32(1)
        posc[:, :3] = np.where(on_face, posc[:, :3] - cell, posc[:, :3])
        ms_refc, a_refc = _timed(lambda: eng.direct_host(posc, velc, accc, symmetry=sym_cell), reps=1)
        tgtc = accc[:, 3] == 1.0
        for p in (4, 10):
            ms, a = _timed(lambda: eng.fmm_sph_host(posc, velc, accc, order=p, symmetry=sym_cell))
            l2, mx = rel_l2_np(a[tgtc], a_refc[tgtc])
            rows.append({"kind": "translational_reference_operators", "n": n, "images": 27, "order": p,
                         "ms_host_in_out": ms, **eng.fmm_sph_stats(), "rel_l2": l2, "max_rel": mx,
                         "ref_contract_direct_ms": ms_refc,
                         "vs": "reference-contract direct sum over the 27 images"})
            log(json.dumps(rows[-1]))
        import oracle
        t0 = time.perf_counter()
        oracle.direct_f64(d, i0=0, i1=128, sym="rotational", nsec=6)
        dtc = time.perf_counter() - t0
        res["cpu_baseline"] = {"rotational_fp64_direct_s_per_eval_est": dtc / 128 * n, "cores": 1, "kind": "port",
                               "sample": f"oracle FP64 direct sum over the 6 sectors, 128 targets ({dtc:.1f} s), "
                                         "scaled to 262144 targets"}
        res.update(workload="FMM with symmetry images, N=262144 per cell/sector", runs=rows)
    elif k == 5:
        res.update(run_config5(eng, args, log))
    eng.close()
    print(json.dumps(res), flush=True)


def run_config5(eng, args, log):
    """64 cone plumes, N lognormal (median 32768, sigma ln 4 / 2) clipped to
    [8192, 131072], in one batch: K=200 graph-captured DKD steps of --adj-dt
    (FP32C forces, FP64 state) with checkpoints every 20; the checkpointed discrete
    adjoint of J = mean sin^2 of the velocity divergence angle with respect to
    charges and initial velocities; central differences of the FP64 rollout
    (h=1e-6) on `--fd-dirs` seeded directions."""
    import numpy as np
    import torch
    import solveit_hip as P
    from solveit_hip import synth
    rng = np.random.default_rng(20261208)
    sizes = np.clip(np.round(np.exp(rng.normal(np.log(32768), np.log(4) / 2, 64))), 8192, 131072).astype(int)
    parts = [synth.cone(int(m), seed=1000 + i) for i, m in enumerate(sizes)]
    keys = ("x", "y", "z", "vx", "vy", "vz", "q", "status", "species", "t_inj")
    allb = {key: np.concatenate([p[key] for p in parts]) for key in keys}
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n = int(off[-1])
    pairs_step = float(np.sum(sizes.astype(float) ** 2))
    log(f"64 instances, N total {n}, min {sizes.min()} max {sizes.max()}, pairs/step {pairs_step:.3e}")
    dev = "cuda:0"
    base = {key: torch.as_tensor(allb[key], device=dev) for key in ("x", "y", "z")}
    status = torch.as_tensor(allb["status"], dtype=torch.int32, device=dev)
    species = torch.as_tensor(allb["species"], dtype=torch.int32, device=dev)
    q0 = torch.as_tensor(allb["q"], device=dev)
    v0 = [torch.as_tensor(allb[key], device=dev) for key in ("vx", "vy", "vz")]
    dt, K, C = args.adj_dt, 200, 20
    plan = eng.batch(off, P.SHP_PREC_FP32C)
    plan64 = eng.batch(off, P.SHP_PREC_FP64)
    opts = P.make_step_opts(method=P.SHP_METHOD_DIRECT, precision=P.SHP_PREC_FP32C)
    opts64 = P.make_step_opts(method=P.SHP_METHOD_DIRECT, precision=P.SHP_PREC_FP64)

    def state(q, v):
        return P.DeviceState(base["x"].clone(), base["y"].clone(), base["z"].clone(), q.clone(), status,
                             species, v[0].clone(), v[1].clone(), v[2].clone(), torch.zeros_like(q))

    def copy(st):
        return P.DeviceState(st.x.clone(), st.y.clone(), st.z.clone(), st.q, st.status, st.species,
                             st.vx.clone(), st.vy.clone(), st.vz.clone(), st.t_inj)

    mask = status == 1

    def objective(st):
        s2 = (st.vx ** 2 + st.vy ** 2) / (st.vx ** 2 + st.vy ** 2 + st.vz ** 2)
        return s2[mask].mean()

    tune_ms, chosen = plan.tune(state(q0, v0))
    log(f"tuner ms {['%.2f' % m for m in tune_ms]} chosen {chosen}")
    # forward with checkpoints
    st = state(q0, v0)
    ckpt = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s0 in range(0, K, C):
        ckpt.append(copy(st))
        plan.leapfrog(st, opts, s0 * dt, dt, C, use_graph=True)
    torch.cuda.synchronize()
    fwd_s = time.perf_counter() - t0
    J = float(objective(st))
    log(f"forward {fwd_s:.2f} s, J {J:.6e}")
    # backward: dJ/dv at the end, then the checkpointed reverse sweep
    vx, vy, vz = st.vx, st.vy, st.vz
    r2 = vx ** 2 + vy ** 2 + vz ** 2
    m = mask.double() / mask.sum()
    gvx = m * (2 * vx / r2 - 2 * vx * (vx ** 2 + vy ** 2) / r2 ** 2)
    gvy = m * (2 * vy / r2 - 2 * vy * (vx ** 2 + vy ** 2) / r2 ** 2)
    gvz = m * (-2 * vz * (vx ** 2 + vy ** 2) / r2 ** 2)
    xb = torch.zeros(3, n, dtype=torch.float64, device=dev)
    vb = torch.stack([gvx, gvy, gvz]).contiguous()
    qb = torch.zeros(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for seg in range(len(ckpt) - 1, -1, -1):
        s = copy(ckpt[seg])
        starts = []
        for j in range(C):
            starts.append(copy(s))
            plan.leapfrog(s, opts, (seg * C + j) * dt, dt, 1, use_graph=False)
        for s in reversed(starts):
            xb, vb = plan.step_vjp(s, dt, xb, vb, qb)
        log(f"backward segment {seg} done")
    torch.cuda.synchronize()
    bwd_s = time.perf_counter() - t0
    grads = [qb, vb[0], vb[1], vb[2]]

    def J64(q, v):
        s = state(q, v)
        for s0 in range(0, K, C):
            plan64.leapfrog(s, opts64, s0 * dt, dt, C, use_graph=True)
        return float(objective(s))

    gen = torch.Generator(device="cpu").manual_seed(2026)
    h = 1e-6
    worst, fd_rows = 0.0, []
    # directions fd_first .. fd_first + fd_dirs - 1 of the seeded sequence (a long
    # check can be split over runs; every run draws the same sequence)
    for _ in range(args.fd_first):
        for _ in range(4):
            torch.randn(n, generator=gen, dtype=torch.float64)
    t0 = time.perf_counter()
    for i in range(args.fd_first, args.fd_first + args.fd_dirs):
        dirs = [torch.randn(n, generator=gen, dtype=torch.float64).to(dev) for _ in range(4)]
        dq, dv = dirs[0] * 0.01, [x * 10.0 for x in dirs[1:]]
        ana = float(sum((g * x).sum() for g, x in zip(grads, [dq] + dv)))
        jp = J64(q0 + h * dq, [a + h * b for a, b in zip(v0, dv)])
        jm = J64(q0 - h * dq, [a - h * b for a, b in zip(v0, dv)])
        fd = (jp - jm) / (2 * h)
        worst = max(worst, abs(ana - fd) / abs(fd))
        fd_rows.append({"adjoint": ana, "fd": fd})
        log(f"direction {i}: adjoint {ana:.8e} fd {fd:.8e}")
    fd_s = time.perf_counter() - t0
    plan.close()
    plan64.close()
    # CPU leg: the oracle's FP64 discrete adjoint (rollout + reverse sweep) of
    # one 4096-ion cone over 20 steps
    import oracle
    d1 = synth.cone(4096, seed=1000)
    t0 = time.perf_counter()
    oracle.rollout_grad_f64(d1, dt, 20, lambda x, v: (np.zeros_like(x), np.ones_like(v)))
    cpu_s = time.perf_counter() - t0
    return {"workload": "64 cone plumes N lognormal in [8192, 131072], K=200 steps, checkpoints every 20",
            "dt": dt,
            "n_total": n, "sizes_min_max": [int(sizes.min()), int(sizes.max())],
            "pairs_per_step": pairs_step, "tuner_ms": tune_ms, "tuner_choice": chosen,
            "forward_s": fwd_s, "forward_particle_steps_per_s": n * K / fwd_s,
            "forward_fp32c_tflops_20flop": FLOP_PER_PAIR * pairs_step * K / fwd_s / 1e12,
            "backward_s": bwd_s, "objective": J, "fd_first": args.fd_first, "fd_dirs": args.fd_dirs, "fd_s": fd_s,
            "fd_worst_rel_diff": worst, "fd": fd_rows,
            "cpu_baseline": {"adjoint_particle_steps_per_s": 4096 * 20 / cpu_s, "cores": 1, "kind": "port",
                             "sample": f"oracle FP64 rollout + discrete adjoint, one 4096-ion cone, 20 steps "
                                       f"({cpu_s:.1f} s)"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--order", type=int, default=10)
    ap.add_argument("--leaf-size", type=int, default=20)
    ap.add_argument("--dt", type=float, default=1e-12)
    ap.add_argument("--cpu-targets", type=int, default=1024)
    ap.add_argument("--cpu-leaf-stride", type=int, default=16,
                    help="CPU FMM leg: evaluate L2P + P2P on every k-th leaf box")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--config", type=int, default=0,
                    help="measure workload 1-5 (run_config) instead of the headline step (one JSON line)")
    ap.add_argument("--fd-dirs", type=int, default=32, help="workload 5: finite-difference directions")
    ap.add_argument("--fd-first", type=int, default=0, help="workload 5: index of the first direction checked")
    ap.add_argument("--adj-dt", type=float, default=1e-14,
                    help="workload 5: time step of the differentiated rollout (s)")
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

    if args.config:
        run_config(args.config, args)
        return
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
    cpu_state = None if args.no_cpu else st.to_numpy()  # plume state at the start of the timed steps
    leaf0 = eng.fmm_leaf_level()
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
        shells = [0, 0, 0] + list(solveit_hip.fmm_shell_orders(args.order, 2))
        line["cpu_baseline"] = cpu_baseline(cpu_state, args.cpu_targets,
                                            (args.order, leaf0, 2, shells, args.cpu_leaf_stride))
    eng.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
