#!/usr/bin/env python3
"""SU(3) checkerboard Metropolis throughput on one MI355X.

One step = one full checkerboard sweep (mu = 0..3, parity 0/1: 8 sub-pass
launches) of single-generator Metropolis over an SU(3) Wilson-action
lattice, 10 proposals per link (generator h mod 8 for hit h: the reference
sweep's 8 generators, then generators 0 and 1 again), links stored FP32,
accept test and reunitarisation FP64.  Default workload: 32^4, beta = 6.0,
proposal angle 0.35, cold start as the reference's run_chain
(su3_ensemble.rs:412) followed by the warm-up sweeps; the links stay
resident in HBM for the whole timed region.  The same line carries the
32^4 8-proposal sweep (the reference's proposal set) and 24^4 at 10 hits,
each with its own roofline.

The second half of the metric, plaquette ESS/s of the learned proposal
against plain Metropolis, is measured on SU(3) 16^3x32 (learned_vs_plain).

Prints one JSON line with the throughput (link updates/s), the roofline of
the dominant kernel (k_su3_metropolis_ws -- k_su3_metropolis on lattices of
fewer than 256 x 256 links per sub-pass -- one launch per checkerboard sub-pass:
algorithmic bytes per launch / mean launch duration, the device time of the
timed region by one HIP event pair on the kernel's stream over its launch
count), the CPU baseline (the reference-exact restatement of
metropolis_sweep on every host thread of the job, at 16^4 and 12^3x24) and
the GPU's own algorithm on the host (the threaded FP64 device-order replay
at 24^4 and 32^4).
"""
import argparse
import ctypes as C
import json

import numpy as np
import os
import sys
import threading
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "yang-mills_amd"))

METRIC = "SU(3) link updates/s (% HBM roofline) + plaquette ESS/s learned vs plain"
HBM_PEAK_GBS = 8000.0
# Compulsory traffic per link update of one checkerboard sub-pass: the updated
# link is read and written, the opposite-parity mu links and the three other
# directions (both parities) are read once: 4V reads + V/2 writes per V/2
# updates = 9 link transfers of 18 FP32 = 72 bytes.
BYTES_PER_UPDATE_FP32 = 9 * 72


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def job_threads():
    """Host threads this job may use: the CPUs it is scheduled on, capped by
    OMP_NUM_THREADS when the environment sets it (the GPU pool grants each
    one-GPU job a share of the host's cores that way)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        avail = min(avail, int(env))
    return max(1, avail)


def cpu_baseline():
    """Reference-exact metropolis_sweep (the oracle's restatement of
    su3_ensemble.rs:342-391: lexicographic order, SplitMix64, checked FP64
    products, Taylor exp) on every host thread of the job, one independent
    chain per thread, at the reference's largest admissible SU(3) sizes:
    half the threads on 16^4 (2^18 links, su3_lattice.rs:17), half on
    12^3x24, one sweep each, run at the same time."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as O

    L = O.lib()
    cores = job_threads()
    geoms = [(16, 16, 16, 16), (12, 12, 12, 24)]
    plan = [geoms[i % 2] for i in range(cores)]
    secs = [0.0] * cores

    def work(i):
        n = plan[i]
        links = O.su3_identity(n)
        rng = C.c_uint64(O.lib().ymo_derive_su3_chain_seed(0x2C02198105010002, i))
        p, a = C.c_uint64(), C.c_uint64()
        t = time.perf_counter()
        assert L.ymo_su3_metropolis_sweep(O.ext(n), O.dptr(links), 6.0, 0.35, C.byref(rng), C.byref(p),
                                          C.byref(a)) == 0
        secs[i] = time.perf_counter() - t

    t0 = time.perf_counter()
    th = [threading.Thread(target=work, args=(i,)) for i in range(cores)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    dt = time.perf_counter() - t0
    total = sum(4 * int(np.prod(n)) for n in plan)
    per_geom = {}
    for g in geoms:
        idx = [i for i in range(cores) if plan[i] == g]
        if idx:
            per_geom["x".join(str(v) for v in g)] = {
                "threads": len(idx), "per_core": float(np.mean([4 * np.prod(g) / secs[i] for i in idx]))}
    return {
        "value": total / dt,
        "unit": "link updates/s",
        "cores": cores,
        "per_core": total / dt / cores,
        "kind": "port",
        "host_cpus": os.cpu_count(),
        "by_lattice": per_geom,
        "sample": (f"{cores} threads (all the job's host threads; host reports {os.cpu_count()} CPUs), one "
                   f"reference-order sweep each of SU(3) 16^4 or 12^3x24 at beta=6.0 (8 checked-product Metropolis "
                   f"proposals per link, SplitMix64), {total} link updates in {dt:.1f} s"),
    }


def cpu_device_order_fp64(extents=(24, 32), hits=10):
    """The GPU's algorithm on the host: the device-order replay (checkerboard
    sub-passes, Philox draws, FP64 link storage, FP64 accept and
    reunitarisation, 10 hits) with the links of each sub-pass on OpenMP
    threads, one sweep of a hot-start lattice per extent."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as O

    L = O.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS") or job_threads())
    rows = []
    for e in extents:
        n = (e,) * 4
        vh = L.ymd_slots_per_parity(O.ext(n))
        buf = np.zeros(8 * 18 * vh, dtype=np.float64)
        key = O.lib().ymo_derive_su3_chain_seed(0x2C02198105010002, 0)
        assert L.ymd_su3_hot_start(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, key) == 0
        prm = O.DevSu3Params(6.0, 0.35, hits, 1, key, 0)
        p, a = C.c_uint64(), C.c_uint64()
        t = time.perf_counter()
        assert L.ymd_su3_metropolis_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, C.byref(prm), C.byref(p),
                                          C.byref(a)) == 0
        dt = time.perf_counter() - t
        rows.append({"extent": e, "seconds_per_sweep": dt, "link_updates_per_s": 4 * e ** 4 / dt})
        del buf
    return {"kind": "device-order replay, FP64 storage", "cores": threads, "unit": "link updates/s",
            "hits_per_link": hits, "results": rows,
            "sample": f"one hot-start sweep of SU(3) {' and '.join(f'{e}^4' for e in extents)} at beta=6.0, "
                      f"{hits} hits, links of each checkerboard sub-pass on {threads} OpenMP threads"}


def sweep_record(ym_amd, L, hits, beta, steps, warmup, hot_start, repeats=4):
    """Autotune, warm up and time `steps` sweeps of SU(3) L^4 (FP32 storage):
    updates/s by the host clock, the mean sub-pass launch time from one HIP
    event pair over the region (on the lattice stream), the median of
    `repeats` more regions, and the roofline of the sub-pass kernel."""
    n = (L,) * 4
    V = L ** 4
    angle = 0.35
    key = ym_amd.derive_su3_chain_seed(0x2C02198105010002, 0)
    lat = ym_amd.Su3Lattice4(n, 32)
    if hot_start:
        lat.hot_start(key)
    best, cand = lat.autotune(beta, angle, hits)
    log(f"SU(3) {L}^4 hits={hits}: autotune {best} "
        f"({', '.join(f'{m}/{w}: {t:.1f} us' for (m, w), t in cand.items())})")
    lat.metropolis_sweeps(beta, angle, key, 0, warmup, hits)
    # One HIP event pair on the lattice stream around the whole timed region
    # (8 sub-pass launches per sweep, nothing else on the stream): the mean
    # launch duration is the region's device time over the launch count, so
    # it includes the gaps between back-to-back launches.  Events between
    # the launches are left out because each one delays the next launch.
    t0 = time.perf_counter()
    prop, acc, total_ms = lat.metropolis_sweeps_timed(beta, angle, key, warmup, steps, hits)
    wall = time.perf_counter() - t0
    rep = [total_ms / steps]
    for r in range(repeats):
        rep.append(lat.metropolis_sweeps_timed(beta, angle, key, warmup + (r + 1) * steps, steps, hits)[2] / steps)
    avg_launch_s = total_ms / (8 * steps) / 1e3
    bytes_per_launch = (V // 2) * BYTES_PER_UPDATE_FP32
    achieved = bytes_per_launch / avg_launch_s / 1e9
    per_link = best[0].endswith("/per-link") or (V // 2 + 255) // 256 < 256
    med = sorted(rep)[len(rep) // 2]
    return lat, {
        "workload": f"SU(3) Wilson action {L}^4, beta={beta}, single-generator Metropolis, {hits} proposals/link/sweep",
        "value": 4 * V * steps / wall, "unit": "link updates/s", "steps": steps, "warmup": warmup,
        "ms_per_step": wall * 1e3 / steps, "acceptance": acc / prop,
        "device_time_ms_per_step": total_ms / steps, "repeat_device_ms_per_step": rep,
        "median_repeat_link_updates_per_s": 4 * V / (med * 1e-3),
        "autotuned_tile_mapping": {"mode": best[0], "strip": best[1]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(L, hits),
                     "kernel": "k_su3_metropolis" if per_link else "k_su3_metropolis_ws",
                     "bytes_per_launch": bytes_per_launch, "mean_launch_us": avg_launch_s * 1e6,
                     "note": "achieved = 648 B per update (9 link transfers of 18 FP32 reals) / launch time; "
                             "the staple reads rows 0 and 1 of each neighbour link (48 of 72 B) and rebuilds "
                             "the third, so the measured traffic is below the algorithmic bytes; what limits "
                             "the kernel is the VALU issue of its FP32 staple and FP64 hits (rocprofv3 VALUBusy "
                             "64% at 32^4, 10 hits; with its loads served from the caches it is only 2-3.5 us "
                             "faster)"},
    }


def learned_vs_plain(chains=8, sweeps=2500, discard=500, therm=2000, equil=1000, configs=20, epochs=10):
    """Plaquette ESS/s of the learned proposal (one MH-corrected independence
    proposal per link per sweep, BF16 MFMA inference, no Metropolis hits)
    against plain 8-hit Metropolis on SU(3) 16^3x32 at beta = 6.0.  Chain c
    is thermalised on its own (cold start, `therm` plain sweeps, its own key);
    the network is trained on pairs from chain 0 after that; then every chain
    runs `equil` learned sweeps, because plain Metropolis from a cold start is
    still drifting after 10^4 sweeps at this volume (its plaquette approaches
    the equilibrium value from above for that long) while the learned chain
    equilibrates within ~10^3.  The plain and the learned chain c both
    continue from that equilibrated configuration for `sweeps` sweeps,
    measuring the plaquette after every sweep.  ESS from
    the reference diagnostics (positive-window tau_int, su2_ensemble.rs:
    375-463) summed over chains (su3_confirmation.rs:214).  Two clocks:
    `seconds` is the wall time of the sweeps and the per-sweep measurements
    (host calls included), `device_seconds` the HIP-event time of the sweep
    kernels alone.  The plain chains use the autotuned tile mapping.
    Training (pair collection from one thermalised chain + Adam) is reported
    separately as setup cost."""
    import numpy as np

    import ym_amd

    n = (16, 16, 16, 32)
    keys = [ym_amd.derive_su3_chain_seed(0x2C02198105010002, 1 + c) for c in range(chains)]
    starts = []
    for c in range(chains):
        lat = ym_amd.Su3Lattice4(n, 32)
        lat.metropolis_sweeps(6.0, 0.35, keys[c], 0, therm, 8)
        starts.append(lat)
    t0 = time.perf_counter()
    mlp = ym_amd.learned.Mlp(seed=11, dataset_capacity=configs * 4 * int(np.prod(n)))
    src = starts[0].clone()
    for k in range(configs):
        src.metropolis_sweeps(6.0, 0.35, keys[0], therm + 5 * k, 5, 8)
        mlp.collect_pairs(src)
    pairs = mlp.dataset_size()[0]
    losses, train_ms = mlp.train(epochs=epochs, batch=16384, lr=2e-3)
    setup_s = time.perf_counter() - t0
    for c in range(chains):
        ym_amd.learned.learned_sweeps(starts[c], mlp, 6.0, keys[c], 50000, equil, metro_hits=0, bf16=True)
    out = {"workload": "SU(3) Wilson action 16^3x32, beta=6.0, plaquette after every sweep",
           "training_pairs": pairs, "training_epochs": epochs, "final_training_loss": float(losses[-1]),
           "setup_seconds": setup_s, "chains": chains, "sweeps_per_chain": sweeps, "discarded": discard,
           "thermalisation_sweeps_per_chain": f"{therm} plain + {equil} learned"}
    for kind in ("plain", "learned"):
        rows, acc, wall, dev_ms = [], [], 0.0, 0.0
        for c in range(chains):
            l = starts[c].clone()
            if kind == "plain":
                best, cand = l.autotune(6.0, 0.35, 8)
                if c == 0:
                    out["plain_sub_pass_mapping"] = {"best": f"{best[0]}/{best[1]}",
                                                     "us_per_sub_pass": {f"{m}/{w}": t for (m, w), t in cand.items()}}
            vals = []
            t1 = time.perf_counter()
            for s in range(sweeps):
                sw = 100000 + s
                if kind == "plain":
                    dev_ms += l.metropolis_sweeps_timed(6.0, 0.35, keys[c], sw, 1, 8)[2]
                else:
                    r = ym_amd.learned.learned_sweeps(l, mlp, 6.0, keys[c], sw, 1, metro_hits=0, bf16=True)
                    acc.append(r.learned_accepts / max(1, r.learned_proposals))
                    dev_ms += r.device_ms
                vals.append(l.average_plaquette())
            wall += time.perf_counter() - t1
            rows.append(np.array(vals[discard:]))
        rows = np.array(rows)
        lag = min(100, rows.shape[1] // 2)
        pool = ym_amd.stats.pool_chains(rows, lag)
        taus = [ym_amd.diagnostics(r, lag)[0].tau_int for r in rows]
        means = rows.mean(axis=1)
        out[kind] = {"plaquette": pool.pooled_mean, "plaquette_se": pool.pooled_se,
                     "between_chain_se": float(means.std(ddof=1) / np.sqrt(chains)),
                     "ess": pool.aggregate_ess, "tau_int": taus, "mean_tau_int": float(np.mean(taus)),
                     "seconds": wall, "device_seconds": dev_ms * 1e-3,
                     "device_us_per_sweep": dev_ms * 1e3 / (chains * sweeps),
                     "ess_per_s": pool.aggregate_ess / wall,
                     "ess_per_device_s": pool.aggregate_ess / (dev_ms * 1e-3)}
        if acc:
            out[kind]["acceptance"] = float(np.mean(acc))
    out["ess_per_s_ratio"] = out["learned"]["ess_per_s"] / out["plain"]["ess_per_s"]
    out["ess_per_device_s_ratio"] = out["learned"]["ess_per_device_s"] / out["plain"]["ess_per_device_s"]

    # The same chains run concurrently: one host thread per chain, each
    # lattice on its own HIP stream, the learned chains sharing the one
    # network.  At 16^3x32 a single chain leaves most of the GPU idle (the
    # learned sub-pass is one wave per SIMD), so this is how several chains
    # are run in practice; ESS per wall second of the whole set.
    for kind in ("plain", "learned"):
        lats = [starts[c].clone() for c in range(chains)]
        if kind == "plain":
            for l in lats:
                l.autotune(6.0, 0.35, 8)
        vals = [[] for _ in range(chains)]
        errors = []

        def run(c, kind=kind, lats=lats, vals=vals, errors=errors):
            try:
                l = lats[c]
                for s in range(sweeps):
                    sw = 200000 + s
                    if kind == "plain":
                        l.metropolis_sweeps(6.0, 0.35, keys[c], sw, 1, 8)
                    else:
                        ym_amd.learned.learned_sweeps(l, mlp, 6.0, keys[c], sw, 1, metro_hits=0, bf16=True)
                    vals[c].append(l.average_plaquette())
            except Exception as e:  # reported after the join
                errors.append(e)

        th = [threading.Thread(target=run, args=(c,)) for c in range(chains)]
        t1 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        wall = time.perf_counter() - t1
        if errors:
            raise errors[0]
        rows = np.array([v[discard:] for v in vals])
        lag = min(100, rows.shape[1] // 2)
        pool = ym_amd.stats.pool_chains(rows, lag)
        out[kind]["concurrent"] = {"seconds": wall, "plaquette": pool.pooled_mean, "plaquette_se": pool.pooled_se,
                                   "ess": pool.aggregate_ess, "ess_per_s": pool.aggregate_ess / wall,
                                   "sweeps_per_s": chains * sweeps / wall}
    out["concurrent_ess_per_s_ratio"] = out["learned"]["concurrent"]["ess_per_s"] / \
        out["plain"]["concurrent"]["ess_per_s"]
    d = out["learned"]["plaquette"] - out["plain"]["plaquette"]
    se = max(out["learned"]["plaquette_se"], out["learned"]["between_chain_se"]) ** 2 + \
        max(out["plain"]["plaquette_se"], out["plain"]["between_chain_se"]) ** 2
    out["plaquette_z"] = abs(d) / se ** 0.5
    return out


def side_workload(name):
    """Workloads beside the headline sweep, each printed as its own JSON line:
    FP32-vs-FP64 observable errors for SU(2) (32^4, 48^4 x beta 2.3, 2.5)
    and SU(3) (24^4, 32^4 with 10 stout levels), and the batched 256-chain
    glueball production; SU(2) FP32 heat-bath and over-relaxation sweep
    throughput at 32^4 and 48^4 (su2-sweeps); average_plaquette calls on
    SU(3) and SU(2) FP32 lattices (plaquette)."""
    import ym_amd

    if name == "plaquette":
        rows = []
        for group, n in ((3, (32,) * 4), (3, (16, 16, 16, 32)), (2, (32,) * 4)):
            lat = (ym_amd.Su3Lattice4 if group == 3 else ym_amd.Lattice4)(n, 32)
            lat.hot_start(5)
            for _ in range(3):
                lat.average_plaquette()
            reps = 20
            lat.timer_begin()
            for _ in range(reps):
                lat.average_plaquette()
            us = lat.timer_end() / reps * 1e3
            sites = n[0] * n[1] * n[2] * n[3]
            nbytes = sites * 4 * (72 if group == 3 else 16)
            log(f"SU({group}) {n}: {us:.1f} us per average_plaquette call")
            rows.append({"group": group, "extents": list(n), "us_per_call": us, "link_bytes": nbytes,
                         "algorithmic_gb_per_s_per_call": nbytes / (us * 1e-6) / 1e9})
        return {"workload": "average_plaquette on FP32 hot-start lattices (kernel + partial sums + host copy "
                            "per call; the kernel reads every link once)", "results": rows}
    if name == "su2-sweeps":
        rows = []
        for ext in (32, 48):
            n = (ext,) * 4
            lat = ym_amd.Lattice4(n, 32)
            key = ym_amd.derive_chain_seed(0x2B00201211160101, 0)
            lat.hot_start(key)
            for s in range(5):
                lat.heatbath_sweeps(2.3, key, s, 1)
                lat.overrelax_sweeps(3)
            reps = 10
            # one call of `reps` sweeps: no host synchronisation between them
            lat.timer_begin()
            lat.heatbath_sweeps(2.3, key, 5, reps)
            hb = lat.timer_end() / reps
            lat.timer_begin()
            lat.overrelax_sweeps(3 * reps)
            orx = lat.timer_end() / (3 * reps)
            upd = 4 * ext ** 4
            log(f"SU(2) {ext}^4: heat-bath sweep {hb * 1e3:.1f} us, over-relaxation sweep {orx * 1e3:.1f} us")
            rows.append({"extent": ext, "beta": 2.3, "heatbath_us_per_sweep": hb * 1e3,
                         "overrelax_us_per_sweep": orx * 1e3,
                         "heatbath_link_updates_per_s": upd / (hb * 1e-3),
                         "overrelax_link_updates_per_s": upd / (orx * 1e-3),
                         "overrelax_hbm_fraction": upd * 144 / (orx * 1e-3) / (HBM_PEAK_GBS * 1e9),
                         "heatbath_hbm_fraction": upd * 144 / (hb * 1e-3) / (HBM_PEAK_GBS * 1e9)})
        return {"workload": "SU(2) FP32 links, hot start, beta=2.3, checkerboard heat-bath / over-relaxation "
                            "(144 B per link update: 9 transfers of 16 B)", "results": rows}
    if name == "stout":
        rows = []
        for ext in (24, 32):
            n = (ext,) * 4
            for prec in (32, 64):
                src = ym_amd.Su3Lattice4(n, prec)
                key = ym_amd.derive_su3_chain_seed(0x2C02198105010002, 3)
                src.hot_start(key)
                src.metropolis_sweeps(6.0, 0.35, key, 0, 20, 10)
                dst = ym_amd.Su3Lattice4(n, prec)
                ym_amd.lattice.lib().ym_su3_stout_spatial(dst.handle, src.handle, 0.1, 10)  # warm-up
                dst.timer_begin()
                st = ym_amd.lattice.lib().ym_su3_stout_spatial(dst.handle, src.handle, 0.1, 10)
                ms = dst.timer_end()
                assert st == 0
                V = ext ** 4
                # per level: the 3V spatial links read and written, the V temporal links copied
                # (read + write), neighbours of the staples from the caches: 8V link transfers
                per_level_bytes = 8 * V * (72 if prec == 32 else 144)
                us = ms * 1e3 / 10
                log(f"stout {ext}^4 FP{prec}: {us:.1f} us per level")
                rows.append({"extent": ext, "precision": prec, "exp": "Cayley-Hamilton" if prec == 32 else "Taylor",
                             "us_per_level": us, "spatial_links_per_s": 3 * V / (us * 1e-6),
                             "algorithmic_bytes_per_level": per_level_bytes,
                             "hbm_fraction": per_level_bytes / (us * 1e-6) / (HBM_PEAK_GBS * 1e9),
                             "plaquette_after_10": dst.average_plaquette()})
        return {"workload": "SU(3) spatial stout smearing, 10 levels at rho=0.1 (config 3), thermalised "
                            "24^4 / 32^4, beta=6.0; device time by one event pair on the output lattice's stream",
                "results": rows}
    if name == "su2-precision":
        rows = []
        for ext in (32, 48):
            for beta in (2.3, 2.5):
                log(f"SU(2) {ext}^4 beta={beta}")
                rows.append(ym_amd.precision.su2_precision(ext, beta))
        return {"workload": "SU(2) FP32 storage vs FP64, hot start, 1 HB + 3 OR per update", "results": rows}
    if name == "su3-precision":
        rows = []
        for ext in (24, 32):
            log(f"SU(3) {ext}^4")
            rows.append(ym_amd.precision.su3_precision(ext))
        return {"workload": "SU(3) FP32 storage vs FP64, beta=6.0, 10 stout levels rho=0.1", "results": rows}
    if name == "ensemble":
        log("batched production: 256 chains, 68 sweeps, raw + 4 stout levels")
        t0 = time.perf_counter()
        rep = ym_amd.ensemble.production(chains_per_geometry=64, stout_levels=4)
        wall = time.perf_counter() - t0
        links = 64 * sum(4 * g[0] * g[1] * g[2] * g[3] for g in rep.geometries)
        corr = {}
        for (g, lev, kind), ev in rep.correlators.items():
            if kind == 0 and lev in (0, 2, 4):
                corr[f"{g[0]}^3x{g[3]} stout{lev} C(1)/C(0)"] = float(ev.full[1] / ev.full[0])
        analysis = {}
        for (g, lev), an in rep.analysis.items():
            if lev not in (0, 4):
                continue
            em, pr, fit = an["effective_mass"], an["projection"], an["fit"]
            analysis[f"{g[0]}^3x{g[3]} stout{lev}"] = {
                "m_eff(1)": (f"{em.full_estimate[0]:.4f} +- {em.standard_error[0]:.4f}"
                             if isinstance(em, ym_amd.stats.EffectiveMassSeries) else f"status {em['status']}"),
                "gevp_lambda(t1=2)": (f"{pr.principal_eigenvalue:.5f}" if isinstance(pr, ym_amd.stats.PrincipalProjection)
                                      else f"status {pr['status']}"),
                "fit": (f"window {fit.window_start}-{fit.window_end} m = {fit.fitted_mass:.3f} +- "
                        f"{fit.mass_standard_error:.3f}" if not isinstance(fit, dict) else f"status {fit['status']}"),
            }
        return {"workload": "256 SU(3) chains (64 each of 4^3x10, 6^3x12, 8^3x16, 12^3x24), beta=5.7, "
                            "68 sweeps from cold, A1++ on raw and stout 1-4, C(t) + jackknife, host analysis",
                "analysis": analysis,
                "device_ms": rep.device_ms, "wall_s": wall,
                "link_updates_per_s": links * rep.sweeps / (rep.device_ms * 1e-3),
                "mean_plaquette_by_geometry": {f"{g[0]}^3x{g[3]}": float(np.mean(rep.plaquette[i * 64:(i + 1) * 64]))
                                               for i, g in enumerate(rep.geometries)},
                "mean_acceptance": float(np.mean(rep.acceptance)), "correlator_ratios": corr}
    raise ValueError(name)


def load_traffic(extent, hits):
    path = os.path.join(REPO, "profiles", f"su3_metropolis_{extent}x4_h{hits}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="su3",
                    choices=["su3", "su2-sweeps", "su2-precision", "su3-precision", "ensemble", "plaquette", "stout"],
                    help="su3 (default): the SU(3) 32^4 sweep metric; the others print their own JSON line")
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--extent", type=int, default=32)
    ap.add_argument("--hits", type=int, default=10, help="proposals per link per sweep (config: 10 hits)")
    ap.add_argument("--beta", type=float, default=6.0)
    ap.add_argument("--no-side-records", action="store_true", help="skip the 32^4 8-hit and 24^4 sweep records")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain ESS/s measurement")
    ap.add_argument("--learned-chains", type=int, default=8)
    ap.add_argument("--learned-sweeps", type=int, default=2500, help="measured sweeps per chain (a fifth, at "
                    "most 500, discarded)")
    ap.add_argument("--learned-therm", type=int, default=2000, help="plain thermalisation sweeps per chain")
    ap.add_argument("--learned-equil", type=int, default=1000, help="learned equilibration sweeps per chain")
    ap.add_argument("--hot-start", action="store_true",
                    help="Haar-random start instead of the reference's cold (identity) start")
    ap.add_argument("--traffic-calibration", action="store_true",
                    help="after the timed region copy the link field once (a plain device copy of "
                         "known size) to calibrate the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(independent chains shard across GPUs with no communication; run one process per GPU)",
              file=sys.stderr)
        sys.exit(2)

    import ym_amd

    if args.workload != "su3":
        print(json.dumps(side_workload(args.workload)), flush=True)
        return

    L = args.extent
    n = (L, L, L, L)
    V = L ** 4
    start = "hot" if args.hot_start else "cold"
    log(f"SU(3) {L}^4 FP32 {start} start, beta={args.beta}, hits={args.hits}: timed {args.steps} sweeps")
    lat, head = sweep_record(ym_amd, L, args.hits, args.beta, args.steps, args.warmup, args.hot_start,
                             repeats=0 if args.traffic_calibration else 4)
    if args.traffic_calibration:
        log(f"calibration: device copy of {lat.device_bytes()} bytes")
        lat.clone()
    plaq = lat.average_plaquette()
    unit_res, det_res = lat.link_residuals()
    del lat
    log(f"{head['value']:.4e} link updates/s, {head['roofline']['kernel']} mean "
        f"{head['roofline']['mean_launch_us']:.1f} us, {head['roofline']['achieved']:.0f} GB/s algorithmic "
        f"({100 * head['roofline']['frac']:.1f}% of peak)")
    others = []
    if not args.no_side_records:
        for (Lx, hx) in ((32, 8), (24, 10)):
            if (Lx, hx) == (L, args.hits):
                continue
            _, rec = sweep_record(ym_amd, Lx, hx, args.beta, args.steps, args.warmup, args.hot_start)
            log(f"  {Lx}^4 hits={hx}: {rec['value']:.4e} link updates/s, {rec['roofline']['mean_launch_us']:.1f} us "
                f"per launch, {100 * rec['roofline']['frac']:.1f}% of peak")
            others.append(rec)

    learned = None
    if not args.no_learned:
        log("learned vs plain plaquette ESS/s (16^3x32)")
        learned = learned_vs_plain(chains=max(2, args.learned_chains), sweeps=args.learned_sweeps,
                                   discard=min(500, args.learned_sweeps // 5), therm=args.learned_therm,
                                   equil=args.learned_equil)
        log(f"ESS/s learned {learned['learned']['ess_per_s']:.1f} vs plain {learned['plain']['ess_per_s']:.1f} "
            f"(x{learned['ess_per_s_ratio']:.2f}; device time x{learned['ess_per_device_s_ratio']:.2f}), "
            f"acceptance {learned['learned'].get('acceptance', 0):.3f}")
        log(f"concurrent chains: ESS/s learned {learned['learned']['concurrent']['ess_per_s']:.1f} vs plain "
            f"{learned['plain']['concurrent']['ess_per_s']:.1f} (x{learned['concurrent_ess_per_s_ratio']:.2f})")

    cpu = cpu_dev = None
    if not args.no_cpu_baseline:
        log("cpu baseline: reference-order sweeps of 16^4 and 12^3x24 on every job thread")
        cpu = cpu_baseline()
        log(f"  {cpu['value']:.4e} link updates/s on {cpu['cores']} threads ({cpu['per_core']:.0f} per thread)")
        log("cpu device-order replay, FP64 storage, 24^4 and 32^4")
        cpu_dev = cpu_device_order_fp64()

    rf = head["roofline"]
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "link updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("synthetic: Haar-random hot start from Philox4x32-10" if args.hot_start else
                 "synthetic: cold (identity) start as the reference's run_chain") +
                ", key derive_su3_chain_seed(0x2c02198105010002, 0)",
        "config": {
            "workload": f"SU(3) Wilson action {L}^4, beta={args.beta}, single-generator Metropolis, "
                        f"{args.hits} proposals/link/sweep (generators h mod 8), checkerboard sweeps",
            "extents": list(n),
            "beta": args.beta,
            "proposal_angle": 0.35,
            "hits_per_link": args.hits,
            "link_storage": "fp32 (18 reals)",
            "accept_test": "fp64",
            "reunitarisation": "fp64 per update",
            "acceptance": head["acceptance"],
            "final_plaquette": plaq,
            "unitarity_residual_fp32_storage": unit_res,
            "determinant_residual_fp32_storage": det_res,
            "device_time_ms_per_step": head["device_time_ms_per_step"],
            "repeat_device_ms_per_step": head["repeat_device_ms_per_step"],
            "median_repeat_link_updates_per_s": head["median_repeat_link_updates_per_s"],
            "autotuned_tile_mapping": head["autotuned_tile_mapping"],
        },
        "roofline": {k: rf[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic", "kernel",
                                        "bytes_per_launch", "mean_launch_us", "note")},
        "other_configs": others,
        "cpu_baseline": cpu,
        "cpu_device_order_fp64": cpu_dev,
        "learned_vs_plain": learned,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
