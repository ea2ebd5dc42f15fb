#!/usr/bin/env python3
"""Electron-moves/s of 128-electron DMC on one MI355X (FP64 and mixed precision).

Workload: the seeded 128-electron molecule ae32 (32 Z=4 all-electron atoms, 64+64 electrons,
416 spline basis functions, Slater-Jastrow), DMC with timestep 0.01 and 8192 walkers resident
in HBM.  A step is one DMC step of the whole population: 128 single-electron moves per walker,
the local energy, the weight update (and branching every fifth step).

Prints one JSON line: the headline value is the fastest mixed-precision rate whose DMC energy
agrees with FP64's (otherwise FP64); every precision is reported, with the roofline of the
dominant kernel (the electron-move sweep), a CPU baseline timed on this host, the energy
agreement (paired with FP64: a short stretch of the 128-electron workload, and equilibrated
DMC of the reference's N2 deck) and the latencies of a stream of serving jobs.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ONCE_PER_SWEEP = 1 << 20  # a recompute interval longer than any spin channel
PEAK = {"f64": 78.6e12, "f32": 157.3e12, "bf16": 2.5e15}
HBM = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flops_per_move(nb, ns, ne, na, nshell, p_acc, n_ee=6, n_ei=4):
    """Algorithmic flops of one electron move, term by term as the workload's cost model
    states it: dense LCAO (5 components), Sherman-Morrison on acceptance, ratio + gradient +
    laplacian, Jastrow pairs (e-e over the Ne-1 partners, e-n over the atoms, n_ee / n_ei
    radial functions each), distances and spline shells."""
    return (10.0 * nb * ns + 4.0 * ns * ns * p_acc + 10.0 * ns + (ne - 1) * (12 + 30 * n_ee)
            + na * (12 + 30 * n_ei) + 10 * na + 60 * nshell)


def bytes_per_move(ns, ne, s_inv, p_acc):
    """Algorithmic HBM bytes of one move with the inverse held on chip for the whole sweep:
    MO-cache row read (gradients) and write on acceptance, position and inverse traffic
    amortised over the spin channel's moves (load + store of the inverse once per sweep)."""
    return 3 * ns * s_inv + 5 * ns * s_inv * p_acc + 2.0 * ns * ns * s_inv / ns + 24


def run_gpu(deck, mol, precision, nw, steps, warmup, tau, refresh=0):
    from mainline_amd import engine as E
    from mainline_amd import synth
    eng = E.Engine(deck, precision, seed=20261204, refresh_interval=refresh)
    pos = synth.random_configs(mol, nw, seed=5)
    eng.set_walkers(pos)
    log("[%s] warmup %d steps" % (precision, warmup))
    eng.run(E.method("DMC", nw, nblock=1, nstep=max(1, warmup), timestep=tau))
    log("[%s] timed %d steps" % (precision, steps))
    t0 = time.perf_counter()
    res = eng.run(E.method("DMC", nw, nblock=1, nstep=steps, timestep=tau))
    dt = time.perf_counter() - t0
    # per-kernel average and launch count over the timed steps only (event times on the
    # engine's stream; the engine restarts its kernel times at every run)
    kt = {k: [a, int(c)] for k, (a, c) in eng.kernel_times().items() if c > 0}
    moves = nw * mol.nelec * steps
    return dict(moves_per_s=moves / dt, seconds=dt, ms_per_step=1e3 * dt / steps, energy=float(res["eblock"][0, 0]),
                acceptance=float(res["acceptance"]), kernels=kt, tuning=eng.tuning())


N2_DECK = os.path.join(ROOT, "tests", "golden", "decks", "n2_pp", "n2.wfeval")


def accuracy_leg_n2(precs, nw=4096, nblock=40, nstep=20, tau=0.02, nequil=200, refresh=32):
    """The precision comparison on a physical system with a stationary DMC energy: the
    reference's N2 deck (tests/n2 basis, orbitals and BFD pseudopotential, with a Jastrow).
    One FP64 equilibration (walkers from the reference's random guess, 50 VMC steps, nequil
    DMC steps) gives the first population.  Then, block by block, FP64 runs nstep DMC steps
    from population P_b and leaves P_(b+1); every other precision runs the same block from the
    same P_b with the same random streams (each block starts a fresh DMC run on both sides:
    trial energy from the population, step counters equal).  The trajectories of a block stay
    close, so the block-paired differences measure what the precision changes with an error
    far below the FP64 energy's own; the FP64 blocks give the energy and its error bar."""
    from mainline_amd import engine as E
    eq = E.Engine(N2_DECK, "fp64", seed=98)
    eq.set_walkers(eq.random_walkers(nw, seed=7))
    eq.run(E.method("VMC", nw, nblock=1, nstep=50, timestep=0.3))
    eq.run(E.method("DMC", nw, nblock=1, nstep=nequil, timestep=tau))
    pos, wts = eq.get_walkers(nw)
    eq.close()
    m = E.method("DMC", nw, nblock=1, nstep=nstep, timestep=tau)
    ref = E.Engine(N2_DECK, "fp64", seed=99)
    others = {p: E.Engine(N2_DECK, p, seed=99, refresh_interval=refresh) for p in precs if p != "fp64"}
    blocks = {p: [] for p in ["fp64"] + list(others)}
    for b in range(nblock):
        for p, eng in others.items():
            eng.set_walkers(pos, wts)
            blocks[p].append(float(eng.run(m)["eblock"][0, 0]))
        ref.set_walkers(pos, wts)
        blocks["fp64"].append(float(ref.run(m)["eblock"][0, 0]))
        pos, wts = ref.get_walkers(nw)
    ref.close()
    for eng in others.values():
        eng.close()
    out = {}
    for p, eb in blocks.items():
        eb = np.array(eb)
        out[p] = dict(mean=float(eb.mean()), err=float(eb.std(ddof=1) / np.sqrt(len(eb))), blocks=eb.tolist())
    r = out["fp64"]
    log("[accuracy N2 fp64] E = %.5f +- %.5f" % (r["mean"], r["err"]))
    for p, v in out.items():
        if p == "fp64":
            continue
        d = np.array(v["blocks"]) - np.array(r["blocks"])
        v["paired_delta"] = float(d.mean())
        v["paired_err"] = float(d.std(ddof=1) / np.sqrt(len(d)))
        v["agrees_2sigma"] = bool(abs(v["paired_delta"]) <= 2 * v["paired_err"])
        v["within_fp64_error_bar"] = bool(abs(v["paired_delta"]) <= r["err"])
        log("[accuracy N2 %s] E = %.5f +- %.5f, paired difference from FP64 %+.2e +- %.2e Ha" % (
            p, v["mean"], v["err"], v["paired_delta"], v["paired_err"]))
    return out


def accuracy_leg(deck, mol, precs, nw, nblock, nstep, tau, nequil, refresh=32):
    """DMC of every precision from the same walkers with the same random streams: an
    equilibration stretch of nequil steps, then nblock blocks of nstep steps.  Per precision:
    block-averaged energies, their mean and standard error, and the block-by-block difference
    from FP64 (the runs share their random numbers, so the paired difference is far more
    precise than either error bar)."""
    from mainline_amd import engine as E
    from mainline_amd import synth
    pos = synth.random_configs(mol, nw, seed=9)
    out = {}
    for p in precs:
        eng = E.Engine(deck, p, seed=99, refresh_interval=0 if p == "fp64" else refresh)
        eng.set_walkers(pos)
        eng.run(E.method("DMC", nw, nblock=1, nstep=nequil, timestep=tau))
        eb = eng.run(E.method("DMC", nw, nblock=nblock, nstep=nstep, timestep=tau))["eblock"][:, 0]
        out[p] = dict(mean=float(eb.mean()), err=float(eb.std(ddof=1) / np.sqrt(len(eb))), blocks=eb.tolist())
        log("[accuracy %s] E = %.5f +- %.5f" % (p, out[p]["mean"], out[p]["err"]))
        eng.close()
    ref = out.get("fp64")
    for p, v in out.items():
        if ref is None:
            continue
        d = np.array(v["blocks"]) - np.array(ref["blocks"])
        v["delta_vs_fp64"] = v["mean"] - ref["mean"]
        v["combined_err"] = float(np.hypot(v["err"], ref["err"]))
        v["agrees_2sigma"] = bool(abs(v["delta_vs_fp64"]) <= 2 * v["combined_err"])
        v["paired_delta"] = float(d.mean())
        v["paired_err"] = float(d.std(ddof=1) / np.sqrt(len(d))) if len(d) > 1 else 0.0
    return out


def fixed_config_errors(deck, mol, precs, K=10000, refresh=32):
    """Errors per quantity of each mixed precision against FP64 on K fixed configurations
    drawn from a seeded FP64 VMC equilibration of the workload (random Gaussian clouds, then
    30 VMC steps at timestep 0.3).  The FP64 engine is the reference: it equals the oracle to
    1e-10 on this molecule (the parity tests), and the oracle is too slow for K of this size.
      fresh:       each configuration evaluated from scratch (FP64 recompute of the inverse
                   from the FP32 MO values): ln|Psi| (absolute), gradient of ln Psi (per
                   configuration max |error| / max |gradient|), local energy (absolute, Ha)
      after_sweep: one VMC step of the mixed engine (a sweep of each spin channel with FP32
                   Sherman-Morrison updates, FP64 recompute every `refresh` moves), then the
                   local energy the method computed against FP64 at the same final positions.
    Reported as max and median over the K configurations."""
    from mainline_amd import engine as E
    from mainline_amd import synth
    ref_eng = E.Engine(deck, "fp64", seed=515)
    ref_eng.set_walkers(synth.random_configs(mol, K, seed=21))
    ref_eng.run(E.method("VMC", K, nblock=1, nstep=30, timestep=0.3, ndecorr=1))
    cfg, _ = ref_eng.get_walkers(K)
    ref = ref_eng.eval(cfg)

    def stats(x):
        return dict(max=float(np.max(x)), median=float(np.median(x)))

    out = {"configurations": K, "source": "seeded FP64 VMC equilibration (30 steps, timestep 0.3)"}
    for p in precs:
        eng = E.Engine(deck, p, seed=515, refresh_interval=refresh)
        a = eng.eval(cfg)
        gerr = np.max(np.abs(a["grad"] - ref["grad"]), axis=(1, 2)) / np.max(np.abs(ref["grad"]), axis=(1, 2))
        el_a, el_r = a["eparts"].sum(axis=1), ref["eparts"].sum(axis=1)
        r = dict(logpsi_abs=stats(np.abs(a["logpsi"] - ref["logpsi"])), grad_rel=stats(gerr),
                 eloc_abs_ha=stats(np.abs(el_a - el_r)), eloc_rel=stats(np.abs(el_a - el_r) / np.abs(el_r)),
                 sign_mismatches=int(np.sum(a["sign"] != ref["sign"])))
        eng.set_walkers(cfg)
        eng.run(E.method("VMC", K, nblock=1, nstep=1, timestep=0.3, ndecorr=1))
        pos_f, _ = eng.get_walkers(K)
        st = eng.walker_state(K)
        rf = ref_eng.eval(pos_f)
        el_r2 = rf["eparts"].sum(axis=1)
        r["after_sweep"] = dict(eloc_abs_ha=stats(np.abs(st["elast"][:, 3] - el_r2)),
                                eloc_rel=stats(np.abs(st["elast"][:, 3] - el_r2) / np.abs(el_r2)),
                                refresh_interval=refresh)
        out[p] = r
        log("[errors %s] ln|Psi| max %.2e, grad max %.2e, E_L max %.2e Ha (median %.2e); after a sweep E_L max "
            "%.2e Ha" % (p, r["logpsi_abs"]["max"], r["grad_rel"]["max"], r["eloc_abs_ha"]["max"],
                         r["eloc_abs_ha"]["median"], r["after_sweep"]["eloc_abs_ha"]["max"]))
        eng.close()
    ref_eng.close()
    return out


def batching_leg(precision, steps, tmp, warmup=90):
    """The batching workload: eight molecules of 8..192 electrons (Z=4 atoms, 13 functions per
    atom, seeded walker counts in 512..4096) as DMC (timestep 0.01) in shared launches: one
    sweep launch per spin channel and 32-move recompute interval covering all eight, one
    recompute, local-energy, branching and copy launch, replayed as a HIP graph, LCAO loads in
    flight tuned per size bucket during the warm-up.  Beside it the same engines run one after
    another with their own launches."""
    from mainline_amd import engine as E
    from mainline_amd import synth
    engs, sizes = [], []
    for mol, nw in synth.batch_molecules():
        deck = mol.write(os.path.join(tmp, mol.name))
        e = E.Engine(deck, precision, seed=77)
        e.set_walkers(synth.nonsingular_configs(mol, nw, lambda p: e.eval(p)["logpsi"], seed=3))
        engs.append(e)
        sizes.append((mol.nelec, nw))
    B = E.Batch(engs)
    log("[batching %s] warm-up %d steps (tuning)" % (precision, warmup))
    B.run(E.method("DMC", 0, nblock=1, nstep=warmup, timestep=0.01))
    r = B.run(E.method("DMC", 0, nblock=1, nstep=steps, timestep=0.01))
    _, buckets = B.kernel_times()
    moves_solo, secs_solo = 0.0, 0.0
    per = []
    for e, (ne, nw) in zip(engs, sizes):
        rr = e.run(E.method("DMC", 0, nblock=1, nstep=steps, timestep=0.01))
        moves_solo += rr["moves"]
        secs_solo += rr["seconds"]
        per.append(dict(electrons=ne, walkers=nw, alone_moves_per_s=rr["moves_per_s"]))
    B.close()
    res = dict(precision=precision, molecules=per, steps=steps, moves_per_s=r["moves_per_s"],
               ms_per_step=1e3 * r["seconds"] / steps, tuned=r["tuned"], lcao_loads_per_bucket=buckets,
               one_after_another_moves_per_s=moves_solo / secs_solo,
               over_one_after_another=r["moves_per_s"] / (moves_solo / secs_solo))
    log("[batching %s] %.3e electron-moves/s in shared launches, %.3e one after another" % (
        precision, res["moves_per_s"], res["one_after_another_moves_per_s"]))
    return res


def small_configs_leg(tmp, c2_walkers=4096, c2_steps=10, c2_warmup=3):
    """The two smaller workloads at their own sizes in FP64: the 4-electron H4 chain VMC (256
    walkers, timestep 0.5, 50 blocks of 20 steps, the whole run timed) and the 64-electron
    pseudopotential molecule DMC (16 Z=4 atoms, 208 functions, 4096 walkers, tau 0.01;
    c2_warmup untimed and c2_steps timed steps) with that step's kernel times: the local
    energy (k_local, which holds the nonlocal pseudopotential's 12-point quadrature ratios)
    beside the sweeps.  The pseudopotential's algorithmic flops per walker-step are the cost
    model's estimate: 12 points x (2 Nb Nsigma + 2 Nsigma + Ne n_ee 12) per electron-atom pair
    inside the cutoff, at 1.5 such pairs per electron."""
    from mainline_amd import engine as E
    from mainline_amd import synth
    out = {}
    h4 = synth.h4()
    eng = E.Engine(h4.write(os.path.join(tmp, "c1")), "fp64", seed=11)
    eng.set_walkers(synth.random_configs(h4, 256, seed=5))
    t0 = time.perf_counter()
    r = eng.run(E.method("VMC", 256, nblock=50, nstep=20, timestep=0.5))
    dt = time.perf_counter() - t0
    out["h4_vmc"] = dict(walkers=256, blocks=50, steps=20, seconds=dt, moves_per_s=r["moves"] / dt,
                         energy=float(r["eblock"][:, 0].mean()), acceptance=float(r["acceptance"]))
    eng.close()
    log("[h4 VMC] %.3e electron-moves/s, whole run %.3f s" % (out["h4_vmc"]["moves_per_s"], dt))
    pp = synth.pp16()
    nw = c2_walkers
    eng = E.Engine(pp.write(os.path.join(tmp, "c2")), "fp64", seed=12)
    eng.set_walkers(synth.nonsingular_configs(pp, nw, lambda q: eng.eval(q)["logpsi"], seed=6))
    eng.run(E.method("DMC", nw, nblock=1, nstep=c2_warmup, timestep=0.01))
    t0 = time.perf_counter()
    r = eng.run(E.method("DMC", nw, nblock=1, nstep=c2_steps, timestep=0.01))
    dt = time.perf_counter() - t0
    kt = {k: (a, int(c)) for k, (a, c) in eng.kernel_times().items() if c > 0}
    eng.close()
    per_step = {k: a * c / c2_steps for k, (a, c) in kt.items()}
    local = sum(v for k, v in per_step.items() if k.startswith("k_local"))
    sweep = sum(v for k, v in per_step.items() if k.startswith("k_sweep"))
    ns, nb, ne = pp.nup, pp.nbasis, pp.nelec
    f_pp = ne * 1.5 * 12 * (2.0 * nb * ns + 2 * ns + ne * 6 * 12)
    f_em = flops_per_move(nb, ns, ne, pp.natoms, 5 * pp.natoms, float(r["acceptance"]), n_ee=6, n_ei=4)
    out["pp16_dmc"] = dict(walkers=nw, steps=c2_steps, seconds=dt, moves_per_s=nw * ne * c2_steps / dt,
                           ms_per_step=1e3 * dt / c2_steps, kernel_ms_per_step=per_step,
                           local_energy_ms_per_step=local, sweep_ms_per_step=sweep,
                           pp_flops_per_walker_step=f_pp, move_flops_per_walker_step=f_em * ne,
                           pp_over_moves_flops=f_pp / (f_em * ne))
    log("[pp16 DMC] %.3e electron-moves/s, %.2f ms/step (sweeps %.2f, local energy + PP %.2f)" % (
        out["pp16_dmc"]["moves_per_s"], out["pp16_dmc"]["ms_per_step"], sweep, local))
    return out


def serving_leg(precision, njobs, rate, seed, tmp):
    """Poisson stream of VMC/DMC requests (mean `rate` jobs/s) over molecules of 8..512
    electrons (log-uniform, up to 1664 basis functions), 1024 walkers and 1000..20000
    walker-steps each, through the serving loop; job latency is arrival to result."""
    from mainline_amd import engine as E
    from mainline_amd import synth
    rng = np.random.default_rng(seed)
    sizes = [8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512]
    decks = {}
    for ne in sizes:
        m = synth.z4_molecule(ne // 4, 20261300 + ne, name="srv%d" % ne)
        decks[ne] = (m.write(tmp), m)
    srv = E.Server(max_active=32, hbm_budget=128 << 30)
    t_arr = np.cumsum(rng.exponential(1.0 / rate, size=njobs))
    jobs = []
    t0 = time.perf_counter()
    for i in range(njobs):
        ne_req = float(np.exp(rng.uniform(np.log(8), np.log(512))))
        ne = min(sizes, key=lambda x: abs(np.log(x) - np.log(ne_req)))
        dmc = rng.uniform() < 0.5
        nsteps = int(np.ceil(rng.uniform(1000, 20000) / 1024))
        m = E.method("DMC" if dmc else "VMC", 1024, nblock=1, nstep=nsteps, timestep=0.01 if dmc else 0.3,
                     ndecorr=1)
        wait = t_arr[i] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        jobs.append((srv.submit(decks[ne][0], m, 1024, precision=precision, seed=1000 + i,
                                equil_steps=3 if dmc else 0), ne))
    lat, moves, t_end, queue, setup, run, batched = [], 0.0, 0.0, [], [], [], 0
    for jid, ne in jobs:
        eb, st = srv.wait(jid)
        lat.append(st[3])
        moves += st[4]
        t_end = max(t_end, st[2])
        queue.append(st[1] - st[0])
        setup.append(st[8])
        run.append(st[2] - st[1] - st[8])
        batched += int(st[9])
    srv.close()
    lat = np.array(lat)
    span = t_end - 0.0
    q50 = lambda x: float(np.percentile(x, 50))  # noqa: E731
    res = dict(jobs=njobs, rate_per_s=rate, p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
               mean_s=float(lat.mean()), moves_per_s=float(moves / max(span, 1e-9)), wall_s=float(span),
               queue_p50_s=q50(queue), setup_p50_s=q50(setup), setup_total_s=float(np.sum(setup)),
               run_p50_s=q50(run), jobs_in_batches=batched)
    log("[serving %s] %d jobs: p50 %.3f s, p99 %.3f s, %.3e electron-moves/s (p50 queue %.3f s, set-up %.3f s, "
        "run %.3f s; set-up total %.1f s; %d jobs in batches)" % (
            precision, njobs, res["p50_s"], res["p99_s"], res["moves_per_s"], res["queue_p50_s"], res["setup_p50_s"],
            res["run_p50_s"], res["setup_total_s"], batched))
    return res


def profile_traffic(precision, walkers):
    """HBM bytes per launch of the sweep kernel (FETCH_SIZE + WRITE_SIZE, scaled by a copy
    kernel's reading) from the counter runs summarised in profiles/traffic.tsv, when a run of
    this precision and walker count is recorded there; None otherwise."""
    path = os.path.join(ROOT, "profiles", "traffic.tsv")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.rstrip("\n").split("\t")
        if len(f) >= 5 and not line.startswith("#") and f[1] == precision and int(f[2]) == walkers:
            return float(f[3]) + float(f[4])
    return None


def copy_probe(nbytes):
    """Three device-to-device copies of nbytes through the HIP runtime (its copy kernel):
    the reference reading for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    log("copied %d bytes x3" % nbytes)


REF_QWALK = os.path.join(ROOT, "oracle", "_ref", "qwalk")


def cpu_baseline_reference(deck, mol, seconds_target):
    """The reference program itself (built from its sources by oracle/Makefile.ref, -O3
    -ffast-math as its own Linux.mk), one serial process per host core as the stand-in for MPI
    ranks, each running DMC of the same workload with 8 walkers and its own seed.  Two rounds
    that differ only in the step count cancel the set-up time (deck reading, spline fits,
    initial walkers): rate = extra moves / extra wall time."""
    import subprocess
    procs = min(16, os.cpu_count() or 1)
    work = os.path.dirname(deck)
    text = open(deck).read().split("\n", 1)  # INCLUDE <sys>, then the trial function
    def round_(nstep, tag):
        paths = []
        for i in range(procs):
            p = os.path.join(work, "ref_%s_%d.in" % (tag, i))
            with open(p, "w") as f:
                f.write("METHOD {\n  DMC\n  NCONFIG 8\n  NBLOCK 1\n  NSTEP %d\n  TIMESTEP 0.01\n}\n" % nstep)
                f.write("RANDOMSEED { %d %d }\n" % (1000 + i, 77 + i))
                f.write("\n".join(text))
            paths.append(p)
        t0 = time.perf_counter()
        ps = [subprocess.Popen([REF_QWALK, os.path.basename(p)], cwd=work, stdout=subprocess.DEVNULL,
                               stderr=subprocess.DEVNULL) for p in paths]
        codes = [q.wait() for q in ps]
        if any(codes):
            raise RuntimeError("reference qwalk exited with %s" % codes)
        return time.perf_counter() - t0
    t_short = round_(4, "a")
    per_step = max(1e-3, (round_(24, "b") - t_short) / 20)
    n_long = 4 + max(20, int(seconds_target / per_step))
    t_long = round_(n_long, "c")
    moves = procs * 8 * mol.nelec * (n_long - 4)
    return dict(value=moves / (t_long - t_short), unit="electron-moves/s", cores=procs, kind="reference",
                sample="reference qwalk DMC, %d processes x 8 walkers x %d steps of the 128-electron workload "
                       "minus a %d-step run (%.1f s - %.1f s)" % (procs, n_long, 4, t_long, t_short))


def cpu_baseline(deck, mol, seconds_target):
    """The oracle's FP64 CPU engine (same algorithm, threads over walkers) on a bounded sample;
    used where the reference program is not built."""
    from oracle import oracle as O
    from mainline_amd import synth
    threads = min(16, os.cpu_count() or 1)
    orc = O.Oracle(deck)
    nw = max(threads, 16)
    pos = synth.random_configs(mol, nw, seed=6)
    t0 = time.perf_counter()
    orc.run(pos, "DMC", nblock=1, nstep=1, timestep=0.01, seed=3, nthreads=threads)
    t1 = time.perf_counter() - t0
    nstep = max(1, int(seconds_target / max(t1, 1e-3)))
    t0 = time.perf_counter()
    orc.run(pos, "DMC", nblock=1, nstep=nstep, timestep=0.01, seed=4, nthreads=threads)
    dt = time.perf_counter() - t0
    moves = nw * mol.nelec * nstep
    return dict(value=moves / dt, unit="electron-moves/s", cores=threads, kind="port",
                sample="oracle FP64 DMC, %d walkers x %d steps of the 128-electron workload (%.1f s)" % (nw, nstep, dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--walkers", type=int, default=8192)
    ap.add_argument("--precision", default="auto", help="fp64, mixed_fp32, mixed_bf16, mixed_bf16x6 or auto")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--accuracy-blocks", type=int, default=10, help="0 disables the energy-agreement leg")
    ap.add_argument("--serving-jobs", type=int, default=100, help="0 disables the serving leg")
    ap.add_argument("--sustainable-rate", type=float, default=2.0,
                    help="arrival rate (jobs/s) of the second serving stream")
    ap.add_argument("--sustainable-jobs", type=int, default=40, help="jobs of the second serving stream (0: none)")
    ap.add_argument("--batch-steps", type=int, default=20, help="timed steps of the batching leg (0 disables it)")
    ap.add_argument("--refresh", type=int, default=32,
                    help="mixed precision: FP64 inverse recompute every this many moves")
    ap.add_argument("--atoms", type=int, default=32,
                    help="Z=4 atoms of the timed molecule (32 = the 128-electron workload)")
    ap.add_argument("--small-configs", type=int, default=1,
                    help="1 = also time the H4 VMC and 64-electron pseudopotential DMC workloads (FP64)")
    ap.add_argument("--error-configs", type=int, default=10000,
                    help="fixed configurations of the per-quantity error report (0 disables it)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many bytes device-to-device (counter calibration run) and exit")
    a = ap.parse_args()
    if a.copy_probe > 0:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the walker engine runs on a single MI355X; sharding the "
              "walkers over several GPUs is not run by this benchmark.", file=sys.stderr)
        sys.exit(2)
    from mainline_amd import synth
    # the 128-electron workload; --atoms N times the same construction at another size (for
    # profiling the 65..256-electrons-per-spin kernels), reported under its own name
    mol = synth.ae32() if a.atoms == 32 else synth.z4_molecule(a.atoms, 20261204 + a.atoms, name="ae%d" % a.atoms)
    tmp = tempfile.mkdtemp(prefix="qwbench_")
    deck = mol.write(tmp)
    tau = 0.01
    results = {}
    precs = ["fp64", "mixed_fp32", "mixed_bf16x6", "mixed_bf16"] if a.precision == "auto" else [a.precision]
    for p in precs:
        # the workload's precision variant: FP64 recompute of the FP32 inverse every 32 moves
        results[p] = run_gpu(deck, mol, p, a.walkers, a.steps, a.warmup, tau, refresh=0 if p == "fp64" else a.refresh)
        log("[%s] %.3e electron-moves/s, %.2f ms/step" % (p, results[p]["moves_per_s"], results[p]["ms_per_step"]))
    variants = {}
    if "mixed_fp32" in results and a.refresh < ONCE_PER_SWEEP:
        # once per sweep of a spin channel (64 FP32 updates between recomputes): reported beside
        variants["mixed_fp32_recompute_once_per_sweep"] = run_gpu(deck, mol, "mixed_fp32", a.walkers, a.steps,
                                                                  a.warmup, tau, refresh=ONCE_PER_SWEEP)
    acc = acc_n2 = errs = None
    mixed = [p for p in results if p != "fp64"]
    if a.accuracy_blocks > 0:
        acc = accuracy_leg(deck, mol, list(results), a.walkers, a.accuracy_blocks, 10, tau, 40, a.refresh)
        acc_n2 = accuracy_leg_n2(list(results), refresh=a.refresh)
    if a.error_configs > 0 and mixed:
        errs = fixed_config_errors(deck, mol, mixed, K=a.error_configs, refresh=a.refresh)
    # headline: the fastest mixed precision whose DMC energy equals FP64's on the N2 deck (a
    # system with a stationary DMC energy): the block-paired difference within two of its
    # standard errors AND within the FP64 run's own error bar; FP64 itself when none does
    head = "fp64" if "fp64" in results else precs[0]
    gate = {}
    for p in mixed:
        v = (acc_n2 or {}).get(p)
        ok = v is not None and abs(v["paired_delta"]) <= 2 * v["paired_err"] and \
            abs(v["paired_delta"]) <= acc_n2["fp64"]["err"]
        gate[p] = bool(ok)
        if ok and (head == "fp64" or results[p]["moves_per_s"] > results[head]["moves_per_s"]):
            head = p
    R = results[head]
    ns, nb, ne, na = mol.nup, mol.nbasis, mol.nelec, mol.natoms
    p_acc = R["acceptance"] if R["acceptance"] > 0 else 0.99
    kname = "k_sweep" if "k_sweep" in R["kernels"] else next(iter(R["kernels"]))
    avg_ms = R["kernels"][kname][0]
    # moves per sweep launch, from the launches counted over the timed steps
    moves_per_launch = a.walkers * ne * a.steps / R["kernels"][kname][1]
    # 5 spline shells per atom (2s, 2p, 1d), n_ee = 6, n_ei = 4
    f_em = flops_per_move(nb, ns, ne, na, 5 * na, p_acc, n_ee=6, n_ei=4)
    dt = "f64" if head == "fp64" else "f32"
    achieved = f_em * moves_per_launch / (avg_ms * 1e-3)
    # The algorithmic flops (dense-LCAO convention) priced against the peak of the arithmetic
    # type; the FP64 and FP32 sweeps run them on the vector ALUs (FP32 non-packed), sparse over
    # the active basis functions, the BF16 modes their LCAO on the matrix cores.
    roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK[dt] / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK[dt], "traffic": profile_traffic(head, a.walkers),
            "executes_on": "matrix cores (LCAO) + VALU" if head in ("mixed_bf16", "mixed_bf16x6") else
            "VALU (%s, non-packed)" % ("FP64" if head == "fp64" else "FP32"),
            "kernel": kname, "avg_ms": avg_ms, "flops_per_move": f_em, "moves_per_launch": moves_per_launch,
            "launch_choice": R.get("tuning", {}).get("sweep")}
    small = small_configs_leg(tmp) if a.small_configs else None
    batching = None
    if a.batch_steps > 0:
        batching = batching_leg("mixed_fp32", a.batch_steps, tmp)
    serving = {}
    if a.serving_jobs > 0:
        for p in dict.fromkeys(["fp64", head]):
            serving[p] = serving_leg(p, a.serving_jobs, 20.0, 4242, tmp)
            if a.sustainable_jobs > 0:
                # a rate the device keeps up with (the 20/s stream queues: its latency is mostly
                # waiting for admission)
                serving[p + "_at_%g_per_s" % a.sustainable_rate] = serving_leg(
                    p, a.sustainable_jobs, a.sustainable_rate, 4343, tmp)
    cpu = None
    if not a.no_cpu:
        cpu = (cpu_baseline_reference if os.access(REF_QWALK, os.X_OK) else cpu_baseline)(deck, mol, a.cpu_seconds)
    line = {
        "metric": "electron-moves/s, 128e DMC, mixed vs FP64 at equal E_L; p50/p99 job latency",
        "value": R["moves_per_s"], "unit": "electron-moves/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": R["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "mixed(%s LCAO, f32 inverse, f64 log-det/weights/sums)" % (
            {"mixed_fp32": "f32", "mixed_bf16": "bf16x3 MFMA", "mixed_bf16x6": "bf16x6 MFMA"}[head]),
        "data": "synthetic (seeded molecule, random-init walkers)",
        "config": {"workload": "%s: %d-electron DMC, %d Z=4 atoms, %d spline basis functions, "
                               "Slater-Jastrow, tau 0.01" % (mol.name, mol.nelec, mol.natoms, mol.nbasis),
                   "walkers": a.walkers, "electrons": ne, "basis": nb, "precision": head,
                   "fp64_recompute_every_moves": a.refresh if head != "fp64" else None},
        "roofline": roof, "cpu_baseline": cpu, "equal_energy_gate": gate,
    }
    if small is not None:
        line["smaller_workloads_fp64"] = small
    for p in results:
        line[p] = {k: results[p][k] for k in ("moves_per_s", "ms_per_step", "energy", "acceptance")}
        if p != "fp64" and "fp64" in results:
            line[p]["over_fp64"] = results[p]["moves_per_s"] / results["fp64"]["moves_per_s"]
            line[p]["fp64_recompute_every_moves"] = a.refresh
    for k, v in variants.items():
        line[k] = {kk: v[kk] for kk in ("moves_per_s", "ms_per_step", "energy", "acceptance")}
        if "fp64" in results:
            line[k]["over_fp64"] = v["moves_per_s"] / results["fp64"]["moves_per_s"]
    if acc is not None:
        line["energy_agreement"] = acc
    if acc_n2 is not None:
        line["energy_agreement_n2"] = acc_n2
    if errs is not None:
        line["errors_vs_fp64"] = errs
    if batching is not None:
        line["batching"] = batching
    if serving:
        line["serving"] = serving
        line["p50_job_latency_s"] = serving[head]["p50_s"]
        line["p99_job_latency_s"] = serving[head]["p99_s"]
    line["kernels"] = R["kernels"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
