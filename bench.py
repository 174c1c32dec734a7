#!/usr/bin/env python3
"""Benchmark: ns/day of the ~200k-site SPC/E + NaCl Ewald system on one MI355X.

Workload: 64,000 SPC/E + 1,152 Na+/Cl- pairs (194,304 sites), 1.03 g/cm3,
seeded random lattice placement, 12 A cutoff, Ewald alpha/kappa auto-set
by cook's formulas for 0.05 K/A r- and k-space force errors.  A step is
one device-resident MD step of cook's Shake(): force call (binning, pair
kernel, exclusion corrections, Ewald k-space), leap-frog + SHAKE and the
Berendsen thermostat, h = 2 fs.  Inputs are resident in HBM.  The timed
steps run as the engine runs them (unmeasured steps replayed from the
captured step graph); the per-kernel times of the roofline come from a
further short eager run with HIP events on the engine's stream.

Prints one JSON line (rank 0, one GPU).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

PEAK = {"fp32_valu": 157.3e12, "fp64_valu": 78.6e12, "fp32_mfma": 157.3e12,
        "fp16_mfma": 2.5e15, "hbm": 8.0e12}
# stencil w, stencil columns per pair part, kz chunk, k parts, matrix-core pass 2
# (0: FMA chains, 1 or 2: charge tiles of 32 per wave)
PROFILED_PIN = {"mixed": "2,3,32,16,1", "fp64": "2,3,8,64,0"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mode", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--workload", choices=["c2", "c1", "c3", "c4"], default="c2",
                    help="c2: 194,304-site SPC/E+NaCl; c1: 1000 SPC/E; c3: Drude-polarizable "
                         "brine, 97,200 sites + 32,600 Drudes, ASPC k=4 with one SCF iteration")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--form", choices=["lj", "exp6"], default="lj",
                    help="site-site form of the c2 brine (exp6: the exp-6 brine of the same minima)")
    ap.add_argument("--rdf", type=int, default=0,
                    help="RDF grid: histograms on, and the cost of a measured force call with "
                         "and without them reported")
    ap.add_argument("--kstride", type=int, default=0,
                    help="CPU baseline: k-space over every n-th charge (0 = auto)")
    ap.add_argument("--pin", default="",
                    help="W,PCOLS,KZC,KPARTS[,MFMA]: stencil, stencil columns per pair part, "
                         "k-space kz chunk and parts, pass 2 on the matrix cores (mixed) instead "
                         "of the autotuner (c4: W,KZC,KPARTS)")
    ap.add_argument("--kpass1", type=int, default=-1, choices=[-1, 0, 1, 2, 4, 5, 6, 7, 8, 16],
                    help="mixed mode: k-space pass 1 as FMA chains (0), on the FP32 matrix cores "
                         "(1, 2: row units; 16, 8: tiles per wave) or on the FP16 matrix cores with "
                         "split operands (4..7: row-unit shapes) (-1: the engine's default)")
    ap.add_argument("--kprod", type=int, default=-1, choices=[-1, 0, 3, 6],
                    help="mixed mode: pass 2's products of FP32 operands (0) or of split-FP16 "
                         "operands, 3 or 6 products (-1: the engine's default, 6)")
    ap.add_argument("--tune", action="store_true",
                    help="c2: run the autotuner instead of the profiled configuration")
    ap.add_argument("--no-overlap", action="store_true",
                    help="k-space on the same stream as the pair kernel (each kernel's own "
                         "duration under a profiler)")
    ap.add_argument("--no-fp64-ref", action="store_true",
                    help="mixed mode: skip the FP64 path's rate and the mixed-vs-FP64 force error")
    ap.add_argument("--forces-only", type=int, default=0,
                    help="run this many force calls (no MD, no JSON line), for counter passes")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="with --forces-only: end with a 1 GiB device copy, the reference for "
                         "scaling FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def copy_calibration(nbytes=1 << 30):
    """a plain device-to-device copy of nbytes through the HIP runtime the
    engine is linked against (the runtime's copy kernel reads and writes
    nbytes), the reference for scaling FETCH_SIZE / WRITE_SIZE.  No torch
    here: torch brings its own HIP runtime into the process."""
    hip = C.CDLL("libamdhip64.so.7")        # already loaded with libmdx.so
    src, dst = C.c_void_p(), C.c_void_p()
    size = C.c_size_t(nbytes)
    for call in (lambda: hip.hipMalloc(C.byref(src), size), lambda: hip.hipMalloc(C.byref(dst), size),
                 lambda: hip.hipMemset(src, 1, size), lambda: hip.hipMemcpy(dst, src, size, 3),
                 lambda: hip.hipDeviceSynchronize()):
        rc = call()
        if rc:
            raise RuntimeError("copy calibration: HIP error %d" % rc)
    hip.hipFree(src)
    hip.hipFree(dst)


def traffic_lookup(workload, mode, kernel, info):
    """HBM bytes per launch measured with rocprofv3 --pmc for this kernel
    and configuration (profiles/pmc_traffic.json), or None"""
    path = os.path.join(REPO, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        doc = json.load(fh)
    for e in doc["kernels"]:
        if (e["workload"], e["mode"], e["kernel"]) != (workload, mode, kernel):
            continue
        if kernel.startswith("kpass") and (e["kz_chunk"], e["k_parts"]) != (
                info["kchunk"], info.get("kparts_asked", info.get("kparts"))):
            continue
        if kernel == "pair" and (e["stencil_w"], e.get("pair_cols_per_part", 0)) != (
                info["stencil"], info.get("pair_cols", 0)):
            continue
        if kernel == "kpass2" and int(e.get("kspace_pass2_mfma", 0)) != int(info.get("kmfma", 0)):
            continue
        if kernel == "kpass2" and int(e.get("kspace_pass2_products", 0)) != int(info.get("kprod", 0)):
            continue
        if kernel == "kpass1" and int(e.get("kspace_pass1_tiles", 0)) != int(info.get("kpass1_tiles", 0)):
            continue
        return e["bytes"]
    return None


def cpu_model():
    """the host CPU's model name (/proc/cpuinfo)"""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, bs_oracle, kstride, threads):
    """The FP64 CPU restatement on a bounded sample of the same force call:
    full r-space pass, k-space over every kstride-th charge (linear in the
    charges), scaled back to the full system."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle
    lib = oracle.load()
    lib.oracle_time_sample.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_int,
                                       C.POINTER(C.c_double)]
    r = np.ascontiguousarray(wl["r_full"])
    t = (C.c_double * 3)()
    rc = lib.oracle_time_sample(C.byref(bs_oracle.struct), r.ctypes.data_as(C.POINTER(C.c_double)),
                                threads, kstride, t)
    if rc:
        raise RuntimeError("oracle_time_sample failed (%d)" % rc)
    nq = int(np.count_nonzero(bs_oracle.info["charge"]) + np.count_nonzero(bs_oracle.info["chargepol"]))
    t_force = t[0] + t[1] * nq / t[2]
    return t_force, dict(r_space_s=t[0], k_space_sample_s=t[1], k_charges=int(t[2]), nq=nq)


def bench_replicas(a):
    """The replica batch: 256 atomic systems (salt melts with Ewald, LJ
    mixtures) of 300-6000 atoms, one batch step = every system's force call
    and leap-frog + Berendsen step, h = 0.005 LJ time units (5 fs with
    sigma = 1 A, epsilon = 1 K, m = 1 program unit).  value = aggregate
    ns/day (all systems advance h per batch step)."""
    from macsimus_amd import abi, workloads
    from macsimus_amd.engine import Batch

    sy = workloads.replicas_c4()
    bss = [workloads.build_replica(sd) for sd in sy]
    mode = abi.MDX_MIXED if a.mode == "mixed" else abi.MDX_FP64
    b = Batch(bss, mode)
    rs = [sd["r"] for sd in sy]
    if a.pin:
        pin = [int(x) for x in a.pin.split(",")]
        w, kzc, kparts = pin[:3]
        b.set_stencil(w)
        if mode == abi.MDX_MIXED:
            b.set_kspace_mfma(len(pin) > 3 and pin[3] != 0)
        b.set_kspace(kzc, kparts)
        tune_ms = None
    else:
        # the batch autotuner: stencil, then k-space kz chunk and parts
        w, kzc, kparts, _, tune_ms = b.autotune(rs)
    info = b.info()
    if a.forces_only:
        for _ in range(a.forces_only):
            fs, outs = b.forces(rs, measure=0)
        print("forces-only: %d calls, pairs %d, config %s" % (a.forces_only, sum(o.npair for o in outs),
                                                            (w, kzc, kparts)), flush=True)
        if a.copy_calibration:
            copy_calibration()
            print("copy-calibration: 1 GiB device copy done", flush=True)
        return
    h = 0.005
    ps = [abi.Step(h=h, eps=1e-8, maxit=0, thermostat=1, T=sd["T"], tau=0.1,
                   nf=3.0 * sum(sd["counts"]) - 3, measure_every=0) for sd in sy]
    fs, outs0 = b.forces(rs, measure=1)
    b.upload_state(rs, [sd["v"] * h for sd in sy])
    if a.warmup:
        for p in ps:
            p.measure_every = a.warmup
        b.md_steps(ps, a.warmup)
    for p in ps:
        p.measure_every = a.steps
    t0 = time.perf_counter()
    outs = b.md_steps(ps, a.steps)
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / a.steps
    ksteps = max(2, min(a.steps, 10))
    b.set_timing(True)
    for p in ps:
        p.measure_every = ksteps
    b.md_steps(ps, ksteps)
    tm = b.timing()
    b.set_timing(False)
    kern = {k: (v[0] / max(v[1], 1), v[1]) for k, v in tm.items() if v[1]}
    nsys = len(sy)
    ns_day = nsys * 86400.0 / (ms_step * 1e-3) * h * 1e-3
    npair = sum(o.npair for o in outs0)
    nlj = sum(o.nlj for o in outs0)
    qk = sum(int(np.count_nonzero(bs.info["charge"])) * bs.info["nk"] for bs in bss)
    work = {"kpass1": 8.0 * qk, "kpass2": 12.0 * qk, "pair": 24.0 * npair + 7.0 * nlj}
    dom = max(("kpass1", "kpass2", "pair"), key=lambda k: kern.get(k, (0, 0))[0])
    peak = PEAK["fp64_valu"] if mode == abi.MDX_FP64 else PEAK["fp32_valu"]
    if dom == "kpass2" and info["kmfma"]:
        # FP32 matrix cores, or 6 FP16 products per FP32 product
        peak = PEAK["fp32_mfma"] if info["kmfma"] == 1 else PEAK["fp16_mfma"] / 6
    achieved = work[dom] / (kern[dom][0] * 1e-3)
    force_ms = sum(kern[k][0] for k in ("bin", "pair", "kpass1", "kcoef", "kpass2") if k in kern)

    cpu = None
    if not a.no_cpu_baseline:
        # oracle force calls of every 16th system, one after another, scaled
        # to the whole batch (the CPU runs the systems in turn)
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import oracle
        threads = min(16, os.cpu_count() or 1)
        prov = oracle.provider()
        stride = 16
        t_sample = 0.0
        for k in range(0, nsys, stride):
            bso = workloads.build_replica(sy[k], provider=prov)
            tc = time.perf_counter()
            oracle.forces(bso, sy[k]["r"], 0, nthreads=threads)
            t_sample += time.perf_counter() - tc
        t_batch = t_sample * stride
        cpu = {"value": nsys * 86400.0 / t_batch * h * 1e-3, "unit": "ns/day (aggregate)",
               "cores": threads, "kind": "port",
               "sample": "oracle/ FP64 restatement, one force call of each of %d systems (every "
                         "%dth, %.2f s), scaled x%d to the batch; leap-frog cost not included"
                         % (len(range(0, nsys, stride)), stride, t_sample, stride),
               "s_per_batch_force": t_batch}

    line = {
        "metric": "aggregate ns/day of 256 batched LJ + Ewald replicas",
        "value": ns_day, "unit": "ns/day (aggregate)", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32" if mode == abi.MDX_MIXED else "f64",
        "data": "synthetic",
        "config": {"workload": "replicas_256", "systems": nsys, "sites": info["sites"],
                   "padded_sites": info["padded"], "cells": info["cells"],
                   "charges": info["charges"], "k_segments": info["kseg"], "pairs": npair,
                   "timestep_ps": h, "precision": a.mode, "parallelism": "1 GPU",
                   "tuned": {"stencil_w": w, "kz_chunk": kzc, "k_parts": kparts,
                             # 0 FP32 FMA, 1 FP32 matrix cores, 2 split-FP16 matrix cores
                             "kspace_pass2_mfma": int(info["kmfma"]), "force_ms": tune_ms}},
        "us_per_force": force_ms * 1e3,
        "replica_steps_per_s": nsys / (ms_step * 1e-3),
        "atom_steps_per_s": info["sites"] / (ms_step * 1e-3),
        "kernel_ms": {k: v[0] for k, v in kern.items()},
        "roofline": {"bound": "mfma" if (dom == "kpass2" and info["kmfma"]) else "valu",
                     "kernel": dom, "achieved": achieved / 1e12,
                     "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": traffic_lookup("replicas_256", a.mode, dom, info),
                     "work_per_launch": work[dom]},
        "cpu_baseline": cpu,
        "energy_K": {"En_pot_sum": sum(o.frc.En_pot + o.frc.En_el for o in outs),
                     "En_kin_sum": sum(o.En_kin for o in outs)},
    }
    print(json.dumps(line), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (several GPUs "
              "would each own a slab of the box with its halo and split the k-vectors)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "c4":
        return bench_replicas(a)
    from macsimus_amd import abi, workloads
    from macsimus_amd.engine import Engine
    from macsimus_amd.system import build_system

    wl = {"c2": workloads.brine_c2, "c1": workloads.water_c1, "c3": workloads.polar_c3}[a.workload]()
    if a.form == "exp6":
        if a.workload != "c2":
            sys.exit("bench.py: --form exp6 is for the c2 brine")
        wl = workloads.brine_c2_exp6()
    polar = bool(wl.get("polar"))
    # Drude systems: the configuration has the Drude offsets after the sites
    wl["r_full"] = np.concatenate([wl["r"], np.zeros_like(wl["r"])]) if polar else wl["r"]
    L = [wl["L"]] * 3
    bs = build_system(wl["ff"], wl["counts"], L, wl["cutoff"], epsr=0.05, epsk=0.05)
    mode = abi.MDX_MIXED if a.mode == "mixed" else abi.MDX_FP64
    eng = Engine(bs, mode)
    if a.workload == "c2" and not a.pin and not a.tune:
        # the configuration the committed rocprof summaries and HBM traffic
        # counts were taken at (the autotuner's choices time the same within
        # run-to-run noise)
        a.pin = PROFILED_PIN[a.mode]
    if a.pin:
        pin = [int(x) for x in a.pin.split(",")]
        w, pcols, kzc, kparts = pin[:4]
        eng.set_stencil(w, pcols)
        eng.set_kchunk(kzc, kparts)
        if mode == abi.MDX_MIXED:
            eng.set_kspace_mfma(pin[4] if len(pin) > 4 else 0)
        tune = (w, kzc, kparts, pcols)
    else:
        # the engine's autotuner picks the cell stencil, its split over
        # pair-kernel parts and the k-space force kernel's kz chunk and work
        # split on this configuration (HIP events)
        tune = eng.autotune(wl["r_full"], reps=2)[:4]
    if a.kpass1 >= 0 and mode == abi.MDX_MIXED:
        eng.set_kspace_pass1(a.kpass1)
    if a.kprod >= 0 and mode == abi.MDX_MIXED:
        eng.set_kspace_products(a.kprod)
    if a.no_overlap:
        eng.set_overlap(False)
    info = eng.info()
    if a.forces_only:
        for _ in range(a.forces_only):
            f, out = eng.forces(wl["r_full"], measure=0)
        print("forces-only: %d calls, pairs %d, tuned %s" % (a.forces_only, out.npair, tune), flush=True)
        if a.copy_calibration:
            copy_calibration()
            print("copy-calibration: 1 GiB device copy done", flush=True)
        return
    h = 0.002
    nf = workloads.degrees_of_freedom(wl["ff"], wl["counts"])
    p = abi.Step(h=h, eps=1e-7, maxit=0, thermostat=1, T=298.0, tau=0.1, nf=float(nf),
                 measure_every=0, aspc_k=4 if polar else 0, scf_iter=1, scf_omega=0.0)
    r1 = np.zeros_like(wl["r_full"])
    # census of the starting configuration (pairs inside the cutoff)
    f0, out0 = eng.forces(wl["r_full"], measure=1)
    rdf_cost = None
    if a.rdf:
        # a measured force call without and with the RDF pass (host clock
        # around synchronous calls, 3 each after one warm call)
        def timed(n=3):
            eng.forces(wl["r_full"], measure=1)
            t = time.perf_counter()
            for _ in range(n):
                eng.forces(wl["r_full"], measure=1)
            return (time.perf_counter() - t) * 1e3 / n
        t_off = timed()
        eng.set_rdf(a.rdf, wl["cutoff"])
        t_on = timed()
        h_ = eng.rdf()
        rdf_cost = {"grid": a.rdf, "records": int(h_["hist"].shape[0]), "bins": int(h_["hist"].shape[1]),
                    "ms_measured_call": t_off, "ms_measured_call_with_rdf": t_on,
                    "pairs_counted_per_call": int(h_["hist"].sum()) // max(h_["nmeas"], 1)}
    eng.upload_state(wl["r_full"], r1)
    if a.warmup:
        p.measure_every = a.warmup
        eng.md_steps(p, a.warmup)
    # timed region: the steps as the engine runs them (unmeasured steps
    # replayed from the captured step graph), energies on the last step
    p.measure_every = a.steps
    t0 = time.perf_counter()
    out = eng.md_steps(p, a.steps)
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / a.steps
    # per-kernel times: a further short run launched eagerly with HIP events
    # recorded on the engine's stream around each kernel family
    ksteps = max(2, min(a.steps, 10))
    eng.set_timing(True)
    p.measure_every = ksteps
    eng.md_steps(p, ksteps)
    tm = eng.timing()
    eng.set_timing(False)
    ns_day = 86400.0 / (ms_step * 1e-3) * h * 1e-3

    # per-kernel algorithmic work (reference counting rule)
    nq, nk = info["nq"], bs.info["nk"]
    npair, nlj = out0.npair, out0.nlj
    work = {"kpass1": 8.0 * nq * nk, "kpass2": 12.0 * nq * nk,
            "pair": 24.0 * npair + 7.0 * nlj}
    kern = {k: (v[0] / max(v[1], 1), v[1]) for k, v in tm.items() if v[1]}
    dom = max(("kpass1", "kpass2", "pair"), key=lambda k: kern.get(k, (0, 0))[0])
    t_dom = kern[dom][0] * 1e-3
    peak = PEAK["fp32_valu"] if mode == abi.MDX_MIXED else PEAK["fp64_valu"]
    achieved = work[dom] / t_dom
    force_ms = sum(kern[k][0] for k in ("bin", "pair", "excl", "kpass1", "kcoef", "kpass2") if k in kern)

    # every hot kernel against its own roof: FLOP/s for the pair kernel
    # and both k-space passes (matrix cores where they run there; with
    # split-FP16 operands each FP32 product costs kprod FP16 products, so
    # the roof is the FP16 matrix peak / kprod), HBM bytes/s for binning
    # (44 B per site + 8 per cell) and the SHAKE step (96 B per site: r, p,
    # r1 read, r, r1 written)
    ncell = info["ncx"] * info["ncy"] * info["ncz"]
    mfma = {"kpass1": bool(info["kpass1_tiles"]), "kpass2": bool(info["kmfma"])}

    def mfma_peak(k):
        if k == "kpass2" and info["kprod"]:
            return PEAK["fp16_mfma"] / info["kprod"]
        if k == "kpass1" and 4 <= info["kpass1_tiles"] <= 7:
            return PEAK["fp16_mfma"] / (3 if info["kprod"] == 3 else 6)
        return PEAK["fp32_mfma"]
    rooflines = {}
    for k in ("pair", "kpass1", "kpass2"):
        if k not in kern:
            continue
        pk = mfma_peak(k) if (mode == abi.MDX_MIXED and mfma.get(k)) else peak
        ach = work[k] / (kern[k][0] * 1e-3)
        rooflines[k] = {"bound": "mfma" if (mode == abi.MDX_MIXED and mfma.get(k)) else "valu",
                        "ms": kern[k][0], "achieved": ach / 1e12, "peak": pk / 1e12,
                        "unit": "TFLOP/s", "frac": ach / pk, "work_per_launch": work[k]}
    bytes_ = {"bin": 44.0 * bs.struct.ns + 8.0 * ncell, "shake": 96.0 * bs.struct.ns}
    for k, b_ in bytes_.items():
        if k in kern:
            ach = b_ / (kern[k][0] * 1e-3)
            rooflines[k] = {"bound": "hbm", "ms": kern[k][0], "achieved": ach / 1e9,
                            "peak": PEAK["hbm"] / 1e9, "unit": "GB/s", "frac": ach / PEAK["hbm"],
                            "bytes_per_launch": b_}

    # the plain FP64 GPU path on the same system and steps, and the
    # mixed-vs-FP64 error of one force call on the starting configuration
    fp64_ref = None
    if mode == abi.MDX_MIXED and not a.no_fp64_ref:
        eng64 = Engine(bs, abi.MDX_FP64)
        pin64 = [int(x) for x in PROFILED_PIN["fp64"].split(",")] if a.workload == "c2" else None
        if pin64:
            eng64.set_stencil(pin64[0], pin64[1])
            eng64.set_kchunk(pin64[2], pin64[3])
        else:
            eng64.autotune(wl["r_full"], reps=2)
        f64, o64 = eng64.forces(wl["r_full"], measure=1)
        eng64.upload_state(wl["r_full"], r1)
        if a.warmup:
            p.measure_every = a.warmup
            eng64.md_steps(p, a.warmup)
        p.measure_every = a.steps
        t0 = time.perf_counter()
        eng64.md_steps(p, a.steps)
        ms64 = (time.perf_counter() - t0) * 1e3 / a.steps
        eng64.close()
        E0, E64 = out0.En_pot + out0.En_el, o64.En_pot + o64.En_el
        pscale = max(abs(x) for x in o64.Pvir[:3])
        fp64_ref = {"ns_per_day": 86400.0 / (ms64 * 1e-3) * h * 1e-3, "ms_per_step": ms64,
                    "mixed_over_fp64": ms64 / ms_step,
                    "rel_rms_force_err": float(np.sqrt(((f0 - f64) ** 2).sum() / (f64 ** 2).sum())),
                    "rel_energy_err": abs(E0 - E64) / abs(E64),
                    "rel_virial_tensor_err": max(abs(x - y) for x, y in zip(out0.Pvir, o64.Pvir)) / pscale,
                    "same_census": out0.npair == o64.npair}

    cpu = None
    if not a.no_cpu_baseline:
        # the host's CPU share of one GPU on the box (16 threads)
        threads = min(16, os.cpu_count() or 1)
        kstride = a.kstride or max(1, nq // 48000)
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import oracle
        bso = build_system(wl["ff"], wl["counts"], L, wl["cutoff"], epsr=0.05, epsk=0.05,
                           provider=oracle.provider())
        t_force, detail = cpu_baseline(wl, bso, kstride, threads)
        cpu = {"value": 86400.0 / t_force * h * 1e-3, "unit": "ns/day", "cores": threads,
               "kind": "port", "cpu_model": cpu_model(),
               "sample": "oracle/ FP64 restatement, one force call of the same system: full "
                         "r-space pass (%.2f s) + k-space over %d of %d charges (%.2f s) scaled "
                         "to all charges; SHAKE step cost not included" % (
                             detail["r_space_s"], detail["k_charges"], detail["nq"],
                             detail["k_space_sample_s"]),
               "s_per_force": t_force}

    line = {
        "metric": "ns/day, ~200k-atom SPC/E+NaCl Ewald: mixed FP32 vs FP64, rel force err <=1e-5",
        "value": ns_day, "unit": "ns/day", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32" if mode == abi.MDX_MIXED else "f64",
        "data": "synthetic",
        "config": {"workload": wl["name"], "sites": int(bs.struct.ns), "charges": nq,
                   "drude": polar,
                   "box_A": wl["L"], "cutoff_A": wl["cutoff"], "alpha": bs.info["alpha"],
                   "kappa": bs.info["kappa"], "k_vectors": nk, "pairs": npair,
                   "timestep_fs": h * 1e3, "precision": a.mode, "parallelism": "1 GPU",
                   "tuned": {"stencil_w": tune[0], "kz_chunk": tune[1], "k_parts": tune[2],
                             "pair_cols_per_part": tune[3], "kspace_pass2_mfma": bool(info["kmfma"]),
                             "kspace_pass2_products": info["kprod"],
                             "kspace_pass1_tiles": info["kpass1_tiles"]}},
        "us_per_force": force_ms * 1e3,
        "kernel_ms": {k: v[0] for k, v in kern.items()},
        "roofline": {"bound": rooflines[dom]["bound"],
                     "kernel": dom, "achieved": achieved / 1e12, "peak": rooflines[dom]["peak"],
                     "unit": "TFLOP/s", "frac": rooflines[dom]["frac"],
                     "traffic": traffic_lookup(wl["name"], a.mode, dom, info),
                     "work_per_launch": work[dom]},
        "rooflines": rooflines,
        "fp64_path": fp64_ref,
        "cpu_baseline": cpu,
        "energy_K": {"En_pot": out.frc.En_pot + out.frc.En_el, "En_kin": out.En_kin},
    }
    if a.no_overlap:
        line["config"]["kspace_overlap"] = False
    if a.form != "lj":
        line["config"]["site_site_form"] = a.form
    if rdf_cost:
        line["rdf"] = rdf_cost
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
