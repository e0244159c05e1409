"""Benchmark of the ED impurity solver on one MI355X.

Workload (one "step"): one full zero-temperature ED solve of the
single-orbital Bethe-bath impurity at Ns=14 (Nbath=13), half-filled sector
(7,7) with D = 11,778,624 states, U=2 by default, seeded bath
(eps ~ U[-1,1], V ~ U[0.2,0.6]):
  ground state: 200 Lanczos steps + eigenvector from the stored Krylov basis
  + FP64 Rayleigh quotient; then 4 Green's-function chains (c+/c for both
  spins) of 150 steps each and G(i w_n) on 2048 Matsubara points (beta=100).
Vectors are FP32 with FP64 dot products, alpha/beta, tridiagonal solve and
continued fraction.  The same solve in FP64 vectors is timed for the
FP32-vs-FP64 speed-up.  dE0, max_dG, max_d_occ, max_d_docc are the FP32
solve's errors against the FP64 CPU oracle's solve of the same problem
(tests/golden/ns14_oracle.npz); errors_vs_oracle holds them for both
precisions at U = 1, 2 and 4.

value = ED solves/s over the timed steps (inputs resident on the device).
roofline = the H.v kernel: algorithmic bytes per launch (x read once + y
written once, 2*4*D) / average launch time from HIP events on the solver's
stream.  cpu_baseline = the FP64 C restatement of the reference algorithm
(oracle/, OpenMP) timed on the host: the median of three runs of a few
Lanczos steps at the same size, scaled to one solve, on up to 16 host
threads and (one_core) on one; with --cpu-full also three whole oracle
solves timed end to end.

The same JSON line also carries:
  batched   512 independent single-orbital problems (Ns in {8,10,12}, random
            U, filling and bath) through the batched solver: solves/s;
  served    a Poisson stream of 200 requests/s (Ns in {10,12,14}, four
            chains each) replayed against the wall clock for
            --serve-seconds (60): solves/s, p50/p99 latency (result time -
            arrival time) overall and by Ns, and the queue depth over time
            (backlog growth: whether the GPU keeps up with the offered rate);
  streaming_ceiling  a hand-written copy and three-stream add of
            Ns=14-sized FP32 vectors: what plain streaming reaches here;
  reorth    the Ns=14 FP32 solve with each reorthogonalization mode of
            the Green's-function chains (none, partial, full): Green's-phase
            time, reorthogonalized steps, G's error against the oracle;
  kanamori  the two-orbital Kanamori Ns=14 problem, FP32 vs FP64 vectors:
            solve times, errors against the oracle's solve, H.v time, and
            the oracle's ms per Lanczos step on the host cores;
  batched_kanamori  two-orbital problems through the batched solver (all 8
            chains of a problem in shared launches) against the same
            problems solved one by one: 4 at Ns=14, 48 at Ns=10 (skipped,
            like `batched`, with --no-batched);
  dmft_ns6  the Ns=6 problem through the DMFT loop (bath fit on the
            host) with the FP64 GPU solver and with the oracle: iterations,
            time per solve and per iteration.

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dmft-ed-zt_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def measured_traffic():
    """HBM bytes per H.v launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE profile (calibrated against a device copy), or None."""
    path = os.path.join(ROOT, "profiles", "hv_ns14_fp32_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_solver(edgpu, prec, U, mu):
    return edgpu.Solver(norb=1, nbath=13, U=U, mu=mu, maxnstep=200, maxnstep_green=150,
                        precision=prec, both_spins=True, seed=1)


def solve(s, em, ek, vmk, om):
    gs = s.ground_state(em, ek, vmk, [(7, 7)])
    occ, docc = s.observables()
    G, co = s.cluster_green_ftn(em, ek, vmk, om)
    return gs.val, G, co, occ, docc


def oracle_fixture():
    """The FP64 CPU oracle's solves of the Ns=14 problems (U = 1, 2, 4 and the
    two-orbital Kanamori one), written by tests/golden/make_ns14.py."""
    path = os.path.join(ROOT, "tests", "golden", "ns14_oracle.npz")
    try:
        return dict(np.load(path))
    except OSError:
        return None


def errors_vs_oracle(ref, name, res):
    """|dE0|, max|dG|, max|dn|, |d docc| of a solve (solve()'s tuple) against
    the oracle's solve `name` of the fixture (None when the fixture lacks it)."""
    if ref is None or f"{name}_E0" not in ref:
        return None
    e0, G, _, occ, docc = res
    return dict(dE0=float(abs(e0 - float(ref[f"{name}_E0"]))),
                max_dG=float(np.abs(G[:, :, 0] - ref[f"{name}_G"]).max()),
                max_d_occ=float(np.abs(occ - ref[f"{name}_occ"]).max()),
                max_d_docc=float(np.abs(docc - ref[f"{name}_docc"]).max()))


def cpu_baseline(em, ek, vmk, U, mu, nsteps, threads, repeats=3):
    """Oracle (FP64 CPU restatement) Lanczos steps at the same size, scaled to
    the reference's work per solve: 2*200+1 ground-state H.v (two-pass
    eigenvector, lanczos_solver.f90:20-90) + 4 x 150 chain H.v.  Each figure
    is the median of `repeats` timed runs (nsteps steps on `threads` host
    threads; one step on one core)."""
    from oracle import oracle as O
    O.build()
    used = O.set_threads(threads)
    m = O.Model(1, 13, U=U, mu=mu)
    b = O.Basis(14, 7, 7)
    v0 = O.start_vector(1, b.ntot)
    O.lanczos_iteration(m, em, ek, vmk, b, v0, 2)  # warm caches / page in
    per = []
    for _ in range(repeats):
        t = time.perf_counter()
        O.lanczos_iteration(m, em, ek, vmk, b, v0, nsteps)
        per.append((time.perf_counter() - t) / nsteps)
    dt = float(np.median(per))
    O.set_threads(1)
    per1 = []
    for _ in range(repeats):
        t = time.perf_counter()
        O.lanczos_iteration(m, em, ek, vmk, b, v0, 1)
        per1.append(time.perf_counter() - t)
    dt1 = float(np.median(per1))
    O.set_threads(threads)
    hv_per_solve = 2 * 200 + 1 + 4 * 150
    return dict(value=1.0 / (dt * hv_per_solve), unit="solves/s", cores=used, kind="port",
                sample=f"median of {repeats} runs of {nsteps} FP64 Lanczos steps of the oracle at "
                       f"Ns=14 sector (7,7) (D=11778624) on {used} threads, {dt * 1e3:.1f} ms/step, "
                       f"scaled to the {hv_per_solve} H.v of one solve",
                ms_per_lanczos_step=dt * 1e3, ms_per_step_runs=[x * 1e3 for x in per],
                one_core=dict(value=1.0 / (dt1 * hv_per_solve), unit="solves/s",
                              ms_per_lanczos_step=dt1 * 1e3, ms_per_step_runs=[x * 1e3 for x in per1]),
                host_cpus=os.cpu_count())


def cpu_full_solve(em, ek, vmk, U, mu, threads, repeats=3):
    """The oracle's whole solve at Ns=14, timed end to end (no scaling): the
    reference's ground_state (200 Lanczos steps, tridiagonal eigensolve,
    second pass for the eigenvector, lanczos_solver.f90:20-90) and
    cluster_green_ftn with four chains of 150 steps (both spins), on
    `threads` host threads; median and all runs, seconds per phase."""
    from oracle import oracle as O
    O.build()
    used = O.set_threads(threads)
    m = O.Model(1, 13, U=U, mu=mu)
    om = O.matsubara(2048, 100.0)
    runs = []
    for _ in range(repeats):
        t = time.perf_counter()
        gs = O.ground_state(m, em, ek, vmk, [(7, 7)], 200, seed=1)
        t1 = time.perf_counter()
        O.cluster_green_ftn(m, em, ek, vmk, gs, 150, om, both_spins=True)
        t2 = time.perf_counter()
        runs.append((t1 - t, t2 - t1))
        log(f"  oracle solve: ground state {t1 - t:.1f} s, Green's function {t2 - t1:.1f} s")
    tot = [a + b for a, b in runs]
    k = int(np.argsort(tot)[len(tot) // 2])
    return dict(value=1.0 / tot[k], unit="solves/s", cores=used, kind="port",
                solve_s=tot[k], gs_phase_s=runs[k][0], green_phase_s=runs[k][1],
                runs_s=[list(r) for r in runs],
                sample=f"median of {repeats} whole oracle solves at Ns=14 sector (7,7): 200 GS "
                       f"steps + second eigenvector pass + 4 chains x 150, on {used} threads")


def cpu_kanamori_step_ms(W, nsteps, threads):
    """Oracle FP64 Lanczos steps of the two-orbital Kanamori problem (Ns=14,
    sector (7,7)): ms per step on the host cores."""
    from oracle import oracle as O
    O.build()
    O.set_threads(threads)
    em, ek, vmk, mu, J, Up = W.two_orbital_kanamori(6, 3.0, seed=1)
    m = O.Model(2, 12, U=3.0, J=J, Up=Up, Jp=J, mu=mu)
    b = O.Basis(14, 7, 7)
    v0 = O.start_vector(1, b.ntot)
    O.lanczos_iteration(m, em, ek, vmk, b, v0, 2)
    t = time.perf_counter()
    O.lanczos_iteration(m, em, ek, vmk, b, v0, nsteps)
    return (time.perf_counter() - t) / nsteps * 1e3


def bench_dmft_ns6(edgpu, W, cpu, threads):
    """The Ns=6 single-orbital Bethe impurity (U=2, half filling,
    2048 Matsubara points, beta=100) through the DMFT loop (edgpu.dmft.run_dmft)
    with the FP64 GPU solver; with cpu, the same loop with the oracle solver
    on the host cores.  Times per impurity solve and per DMFT iteration."""
    from edgpu import dmft
    om = W.matsubara(2048, 100.0)
    em, ek, vmk, mu = W.single_orbital(5, 2.0, 1)
    s = edgpu.Solver(norb=1, nbath=5, U=2.0, mu=mu, maxnstep=80, precision="fp64", seed=1)
    solves = []

    def gpu_solve(em, ek, vmk):
        t = time.perf_counter()
        s.ground_state(em, ek, vmk, [(3, 3)])
        G, _ = s.cluster_green_ftn(em, ek, vmk, om)
        solves.append(time.perf_counter() - t)
        return G

    gpu_solve(em, ek, vmk)  # warm-up
    solves.clear()
    t = time.perf_counter()
    r = dmft.run_dmft(gpu_solve, em, ek, vmk, om, nloop=20, scf_tol=1e-6)
    wall = time.perf_counter() - t
    s.close()
    out = dict(workload="single-orbital Bethe impurity Ns=6, sector (3,3), D=400, U=2, "
                        "half filling, 2048 Matsubara points, beta=100; DMFT loop (bath fit on the "
                        "host) to |G_cl - G_loc|^2 < 1e-6",
               iterations=r["iterations"], converged=r["converged"], final_scf=r["history"][-1],
               gpu_solve_ms=float(np.median(solves)) * 1e3, loop_s=wall,
               loop_ms_per_iteration=wall / r["iterations"] * 1e3)
    if cpu:
        from oracle import oracle as O
        O.build()
        O.set_threads(threads)
        m = O.Model(1, 5, U=2.0, mu=mu)
        cs = []

        def cpu_solve(em, ek, vmk):
            t = time.perf_counter()
            gs = O.ground_state(m, em, ek, vmk, [(3, 3)], 80, seed=1)
            G, _ = O.cluster_green_ftn(m, em, ek, vmk, gs, 80, om)
            cs.append(time.perf_counter() - t)
            return G

        cpu_solve(em, ek, vmk)
        cs.clear()
        rc = dmft.run_dmft(cpu_solve, em, ek, vmk, om, nloop=20, scf_tol=1e-6)
        out.update(cpu_solve_ms=float(np.median(cs)) * 1e3, cpu_iterations=rc["iterations"],
                   cpu_cores=threads, cpu_kind="port")
    return out


def profile_hv(edgpu, em, ek, vmk, U, mu):
    """H.v launches of the hot kernels plus one plain device copy of 2*4*D
    bytes (read 4D, write 4D) whose counter readings calibrate FETCH_SIZE /
    WRITE_SIZE: the single-orbital k_hv1 (FP32, Ns=14, sector (7,7)) and the
    streaming reference kernels on vectors of the same size, the
    two-orbital k_hv2 (FP32 vectors, Ns=14 Kanamori, sector (7,7)), and one
    batched run of 64 single-orbital problems (Ns in {8,10,12}; k_bhv and
    k_baxpy)."""
    import ctypes as C
    from edgpu import workloads as W
    s32 = make_solver(edgpu, "fp32", U, mu)
    ms, by = s32.time_hv(em, ek, vmk, 7, 7, warmup=2, reps=10)
    stream = streaming_ceiling(s32, reps=10)
    s32.close()
    emk, ekk, vmkk, muk, J, Up = W.two_orbital_kanamori(6, 3.0, seed=1)
    sk = edgpu.Solver(norb=2, nbath=12, U=3.0, J=J, Up=Up, Jp=J, mu=muk, maxnstep=200,
                      maxnstep_green=150, precision="fp32", both_spins=True, seed=1)
    ms_k, by_k = sk.time_hv(emk, ekk, vmkk, 7, 7, warmup=2, reps=10)
    sk.close()
    om = W.matsubara(2048, 100.0)
    b = edgpu.Batch(om, precision="fp32", graphs=False)
    for p in W.batched_problems(edgpu, n=64):
        b.submit(p)
    b.run()
    st = b.stats()
    b.close()
    nbytes = 4 * 3432 * 3432
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
        raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"hv_us": ms * 1e3, "hv_bytes": by, "streaming_ceiling": stream, "kanamori_hv_us": ms_k * 1e3,
                      "kanamori_hv_bytes": by_k, "batched_chain_steps": st["chain_steps"],
                      "copy_bytes_read": nbytes, "copy_bytes_written": nbytes}))


def bench_batched(edgpu, W, S, graphs=True):
    """512 independent single-orbital problems (Ns in {8,10,12}) through the
    batched solver: whole-batch solves/s, the median of three timed batches
    after one untimed batch of the same problems (which maps the device
    memory the batch holds; a first batch in a process also pays for that)."""
    from edgpu import autotune
    om = W.matsubara(2048, 100.0)
    probs = W.batched_problems(edgpu, n=512)
    sched, timings = autotune.tune_batch(edgpu, probs, om, graphs=graphs)
    S.run_batched(edgpu, probs, om, schedule=sched, graphs=graphs)
    runs = []
    for _ in range(3):
        res, wall, st = S.run_batched(edgpu, probs, om, schedule=sched, graphs=graphs)
        runs.append((wall, res, st))
    runs.sort(key=lambda r: r[0])
    wall, res, st = runs[1]
    ns = [p.nbath + 1 for p in probs]
    return dict(workload="512 single-orbital problems, Ns uniform in {8,10,12}, U~U[0,6], "
                         "n~U[0.8,1.2] (N=round(n Ns), N_up=ceil(N/2)), eps~U[-1.5,1.5], "
                         "V~U[0.1,0.7]; 200 GS steps + c+/c chains of 150 steps, FP32 vectors, "
                         "2048 Matsubara points",
                problems=len(probs), ns_counts={str(k): ns.count(k) for k in (8, 10, 12)},
                schedule=[list(e) for e in sched],
                solves_per_s=len(probs) / wall, wall_s=wall, runs_solves_per_s=[len(probs) / r[0] for r in runs],
                host_s={k[5:-2]: st[k] for k in ("host_admit_s", "host_progress_s", "host_steps_s", "host_wait_s")},
                batched_steps=st["steps"],
                chain_steps_per_s=st["chain_steps"] / wall,
                e0_finite=bool(np.all(np.isfinite([r["e0"] for r in res]))))


def bench_served(edgpu, W, S, seconds, graphs=True):
    """Poisson arrivals at 200 req/s for `seconds` (Ns 10/12/14 with weights
    0.5/0.35/0.15, U~U[0,8], 200 GS steps + 4 chains of 150) replayed
    against the wall clock; latency = result time - arrival time, solves/s
    over the whole run (arrivals + drain), and the queue depth over time:
    whether the GPU keeps up is read from the backlog's growth over the
    second half of the arrival window."""
    from edgpu import autotune
    om = W.matsubara(2048, 100.0)
    req = W.served_requests(edgpu, rate=200.0, duration=seconds)
    sched, _ = autotune.tune_batch(edgpu, [p for _, p in req], om, sample=8, graphs=graphs)
    r = S.serve(edgpu, req, om, drain_limit=600.0, log=log, schedule=sched, graphs=graphs)
    depth = r["depth"]
    win = [d for d in depth if d[0] <= seconds] or depth[:1]
    half = [d for d in win if d[0] >= 0.5 * seconds]
    slope = float(np.polyfit([d[0] for d in half], [d[3] for d in half], 1)[0]) if len(half) >= 3 else None
    offered = r["submitted"] / seconds if seconds else 0.0
    ns = np.array([req[k][1].nbath + 1 for k in r["index"]])
    lat = r["latencies"]
    by_ns = {str(n): dict(count=int((ns == n).sum()), p50_s=float(np.percentile(lat[ns == n], 50)),
                          p99_s=float(np.percentile(lat[ns == n], 99)))
             for n in (10, 12, 14) if (ns == n).any()}
    every = max(1, int(round(5.0 / 1.0)))
    return dict(workload=f"Poisson 200 req/s for {seconds:g} s, Ns in {{10,12,14}} (0.5/0.35/0.15), "
                         "U~U[0,8], half filling, 200 GS steps + 4 chains x 150 steps (both spins), "
                         "FP32 vectors; admission by device bytes (stored Krylov basis + rings)",
                arrival_window_s=seconds, schedule=[list(e) for e in sched],
                submitted=r["submitted"], completed=r["completed"], unfinished=r["unfinished"],
                wall_s=r["wall"], drain_s=max(0.0, r["wall"] - seconds),
                solves_per_s=r["solves_per_s"], p50_s=r["p50"], p99_s=r["p99"], by_ns=by_ns,
                offered_per_s=offered,
                completed_in_window_per_s=win[-1][2] / seconds if seconds else 0.0,
                backlog_at_window_end=win[-1][3], backlog_max=max(d[3] for d in depth),
                backlog_growth_per_s=slope,
                keeps_up=bool(slope is not None and slope < 0.02 * offered),
                depth_every_5s=[[round(d[0], 2), d[1], d[2], d[3]] for d in depth[::every]])


def streaming_ceiling(solver, n=3432 * 3432, reps=100):
    """What plain streaming reaches on this GPU, for the roofline reading:
    the library's hand-written 16-byte-access copy (read 4n bytes, write 4n)
    and add (read 8n, write 4n) of Ns=14-sized FP32 vectors, timed with HIP
    events on the solver's stream (edg_time_stream); TB/s."""
    cm, am = solver.time_stream(n, reps=reps)
    return dict(copy_us=cm * 1e3, copy_tbs=8 * n / (cm * 1e-3) / 1e12,
                add_us=am * 1e3, add_tbs=12 * n / (am * 1e-3) / 1e12, floats_per_vector=n)


def bench_reorth(edgpu, W, ref, U, mu, em, ek, vmk):
    """The Ns=14 FP32 solve with each reorthogonalization mode of the
    Green's-function chains (0: none, the reference's plain recurrence; 1:
    partial, Simon's omega recurrence; 2: every step against every stored
    chain vector, FP64 Gram-Schmidt): Green's-phase time, reorthogonalized
    steps and G's error against the FP64 oracle's plain solve."""
    om = W.matsubara(2048, 100.0)
    out = {}
    for mode in (0, 1, 2):
        s = make_solver(edgpu, "fp32", U, mu)
        s.set_reorthogonalization(mode)
        solve(s, em, ek, vmk, om)
        res = solve(s, em, ek, vmk, om)
        st = s.last_solve_stats()
        rs = s.reorth_stats()
        err = errors_vs_oracle(ref, f"u{U:g}", res)
        out[str(mode)] = dict(green_phase_ms=st["green_ms"], gs_phase_ms=st["gs_ms"],
                              reorth_events=rs["events"], tested_steps=rs["steps"],
                              max_dG_vs_oracle=None if err is None else err["max_dG"])
        s.close()
    out["workload"] = ("Ns=14 single orbital, sector (7,7), FP32 vectors, 200 GS steps + 4 chains x 150 steps; "
                       "modes 0 none, 1 partial (omega recurrence), 2 full")
    return out


def bench_kanamori(edgpu, W, ref):
    """Two-orbital Kanamori impurity, Ns=14 (6 bath sites per orbital),
    sector (7,7), U=3, J=U/4, U'=U-2J; FP32-vector vs FP64 solve with the
    2x2 G per spin from 8 chains, each against the oracle's solve."""
    em, ek, vmk, mu, J, Up = W.two_orbital_kanamori(6, 3.0, seed=1)
    om = W.matsubara(2048, 100.0)
    out = {}
    for prec in ("fp32", "fp64"):
        s = edgpu.Solver(norb=2, nbath=12, U=3.0, J=J, Up=Up, Jp=J, mu=mu, maxnstep=200,
                         maxnstep_green=150, precision=prec, both_spins=True, seed=1)
        s.ground_state(em, ek, vmk, [(7, 7)])
        t = time.perf_counter()
        res = solve(s, em, ek, vmk, om)
        out[prec] = (time.perf_counter() - t, res)
        ms, by = s.time_hv(em, ek, vmk, 7, 7, warmup=3, reps=20)
        out[prec + "_hv"] = (ms, by)
        s.close()
    t32, r32 = out["fp32"]
    t64, r64 = out["fp64"]
    return dict(workload="two-orbital Kanamori, Ns=14 (6 bath sites per orbital), sector (7,7), "
                         "D=11778624, U=3, J=0.75, U'=1.5, Jp=J; 200 GS steps + 8 chains x 150",
                fp32_solve_ms=t32 * 1e3, fp64_solve_ms=t64 * 1e3, fp32_vs_fp64_speedup=t64 / t32,
                e0=r64[0], errors_vs_oracle=dict(fp32=errors_vs_oracle(ref, "kan", r32),
                                                 fp64=errors_vs_oracle(ref, "kan", r64)),
                fp32_vs_fp64_gpu=dict(dE0=abs(r32[0] - r64[0]),
                                      max_dG=float(np.abs(r32[1] - r64[1]).max())),
                hv_fp32_us=out["fp32_hv"][0] * 1e3,
                hv_fp32_gbs=out["fp32_hv"][1] / (out["fp32_hv"][0] * 1e-3) / 1e9)


def bench_batched_kanamori(edgpu, W, ref):
    """Two-orbital Kanamori problems through the batched solver (every
    problem's 8 c+/c chains in the shared launches), FP32 vectors, against
    the same problems solved one by one on the single-problem path:
    4 problems at Ns=14 (U = 3 with the oracle fixture's bath, and U = 2, 4, 5)
    and 48 at Ns=10 (4 bath sites per orbital, U~U[1,5])."""
    om = W.matsubara(2048, 100.0)

    def kan(nb_per_orb, U, seed):
        em, ek, vmk, mu, J, Up = W.two_orbital_kanamori(nb_per_orb, U, seed=seed)
        ns = 2 * nb_per_orb + 2
        return edgpu.Problem(2 * nb_per_orb, U, mu, em, ek, vmk, ns // 2, ns // 2, maxnstep=200,
                             maxnstep_green=150, both_spins=True, seed=seed, norb=2, J=J, Up=Up, Jp=J)

    def single(prs):
        t = time.perf_counter()
        for pr in prs:
            s = edgpu.Solver(norb=2, nbath=pr.nbath, U=pr.U, J=pr.J, Up=pr.Up, Jp=pr.Jp, mu=pr.mu,
                             maxnstep=200, maxnstep_green=150, precision="fp32", both_spins=True, seed=pr.seed)
            em = pr.em.reshape(2, 2, order="F")
            ek = pr.ek.reshape(pr.nbath, 2, order="F")
            vmk = pr.vmk.reshape(2, pr.nbath, 2, order="F")
            s.ground_state(em, ek, vmk, [(pr.nup, pr.ndn)])
            s.cluster_green_ftn(em, ek, vmk, om)
            s.close()
        return time.perf_counter() - t

    def batched(prs):
        b = edgpu.Batch(om, precision="fp32")
        t = time.perf_counter()
        ids = [b.submit(p) for p in prs]
        b.run()
        wall = time.perf_counter() - t
        res = {r["id"]: r for r in b.poll()}
        b.close()
        return wall, [res[i] for i in ids]

    rng = np.random.Generator(np.random.PCG64(5))
    big = [kan(6, 3.0, 1), kan(6, 2.0, 2), kan(6, 4.0, 3), kan(6, 5.0, 4)]
    small = [kan(4, float(rng.uniform(1.0, 5.0)), 10 + k) for k in range(48)]
    batched(small[:4])  # warm-up: tables, LDS attributes, memory chunks
    out = {}
    for name, prs in (("ns14", big), ("ns10", small)):
        wb, res = batched(prs)
        ws = single(prs)
        out[name] = dict(problems=len(prs), batched_wall_s=wb, batched_solves_per_s=len(prs) / wb,
                         single_path_wall_s=ws, single_path_solves_per_s=len(prs) / ws,
                         batched_vs_single=ws / wb)
        if name == "ns14" and ref is not None and "kan_E0" in ref:
            out[name]["errors_vs_oracle_u3"] = dict(dE0=float(abs(res[0]["e0"] - float(ref["kan_E0"]))),
                                                    max_dG=float(np.abs(res[0]["G"][:, :, 0] - ref["kan_G"]).max()))
    out["workload"] = ("two-orbital Kanamori, half filling, 200 GS steps + 8 chains x 150 steps, FP32 vectors, "
                       "2048 Matsubara points; 4 problems at Ns=14 (sector (7,7)) and 48 at Ns=10 (sector (5,5))")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--U", type=float, default=2.0)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="seconds of Poisson arrivals for the served-stream figures (0: skip)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time three whole oracle solves at Ns=14 end to end on the host "
                         "(about 2 minutes each on 16 cores)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-kanamori", action="store_true")
    ap.add_argument("--no-reorth", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="batched and served sections launch each step's kernels instead of "
                         "replaying a captured HIP graph (for kernel-trace profiles)")
    ap.add_argument("--profile-hv", action="store_true",
                    help="only run the hot H.v kernels (FP32 k_hv1 and k_hv2 at Ns=14, one "
                         "batched run of 64 problems) and one device copy, for hardware-counter "
                         "profiles")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (--gpus {args.gpus} "
              "refused). Independent impurity problems shard across GPUs with no exchange.",
              file=sys.stderr)
        sys.exit(2)

    import edgpu
    from edgpu import workloads as W
    if edgpu.device_count() < 1:
        raise SystemExit("no HIP device visible")

    U = args.U
    em, ek, vmk, mu = W.single_orbital(13, U, seed=1)
    if args.profile_hv:
        profile_hv(edgpu, em, ek, vmk, U, mu)
        return
    om = W.matsubara(2048, 100.0)
    D = 3432 * 3432

    from edgpu import autotune
    s32 = make_solver(edgpu, "fp32", U, mu)
    log("H.v schedule autotuning and timing (FP32)")
    best, sched_us = autotune.tune_single(s32, em, ek, vmk, 7, 7)
    hv_ms, hv_bytes = s32.time_hv(em, ek, vmk, 7, 7, warmup=20, reps=200)
    hv_gbs = hv_bytes / (hv_ms * 1e-3) / 1e9

    log("warmup solves")
    for _ in range(args.warmup):
        solve(s32, em, ek, vmk, om)
    log(f"timed solves: {args.steps}")
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        res32 = solve(s32, em, ek, vmk, om)
        lat.append(time.perf_counter() - t)
    wall = time.perf_counter() - t0
    e32, G32 = res32[0], res32[1]
    st = s32.last_solve_stats()
    info32 = s32.gs_info()
    steps_per_solve = st["n_steps"]

    ref = oracle_fixture()
    uname = f"u{U:g}"
    out = dict(errors_vs_oracle={f"{U:g}": dict(fp32=errors_vs_oracle(ref, uname, res32))})
    if not args.no_fp64:
        log("FP64 solves")
        s64 = make_solver(edgpu, "fp64", U, mu)
        autotune.tune_single(s64, em, ek, vmk, 7, 7)
        solve(s64, em, ek, vmk, om)
        lat64 = []
        for _ in range(max(1, min(args.steps, 3))):
            t = time.perf_counter()
            res64 = solve(s64, em, ek, vmk, om)
            lat64.append(time.perf_counter() - t)
        hv64_ms, hv64_bytes = s64.time_hv(em, ek, vmk, 7, 7, warmup=10, reps=100)
        out.update(fp64_solve_ms=float(np.median(lat64)) * 1e3,
                   fp32_vs_fp64_speedup=float(np.median(lat64) / np.median(lat)),
                   fp32_vs_fp64_gpu=dict(dE0=float(abs(e32 - res64[0])),
                                         max_dG=float(np.abs(G32 - res64[1]).max())),
                   hv_fp64_gbs=hv64_bytes / (hv64_ms * 1e-3) / 1e9, hv_fp64_us=hv64_ms * 1e3)
        out["errors_vs_oracle"][f"{U:g}"]["fp64"] = errors_vs_oracle(ref, uname, res64)
        s64.close()
        # the other interaction strengths of the Ns=14 workload (one solve
        # per precision each), against the oracle's solves
        for U2 in (1.0, 2.0, 4.0):
            if U2 == U:
                continue
            log(f"FP32 and FP64 solves at U={U2:g}")
            em2, ek2, vmk2, mu2 = W.single_orbital(13, U2, seed=1)
            e = {}
            for prec in ("fp32", "fp64"):
                s2 = make_solver(edgpu, prec, U2, mu2)
                e[prec] = errors_vs_oracle(ref, f"u{U2:g}", solve(s2, em2, ek2, vmk2, om))
                s2.close()
            out["errors_vs_oracle"][f"{U2:g}"] = e
    head = out["errors_vs_oracle"][f"{U:g}"]["fp32"] or {}
    out.update(dE0=head.get("dE0"), max_dG=head.get("max_dG"), max_d_occ=head.get("max_d_occ"),
               max_d_docc=head.get("max_d_docc"))

    extra = {}
    from edgpu import serve as S
    if not args.no_batched:
        log("batched problems")
        extra["batched"] = bench_batched(edgpu, W, S, graphs=not args.no_graphs)
    if args.serve_seconds > 0:
        log("served stream")
        extra["served"] = bench_served(edgpu, W, S, args.serve_seconds, graphs=not args.no_graphs)
    threads = min(os.cpu_count() or 1, 16)
    log("streaming ceiling")
    extra["streaming_ceiling"] = streaming_ceiling(s32)
    if not args.no_reorth:
        log("reorthogonalization modes")
        extra["reorth"] = bench_reorth(edgpu, W, ref, U, mu, em, ek, vmk)
    if not args.no_kanamori:
        log("two-orbital Kanamori")
        extra["kanamori"] = bench_kanamori(edgpu, W, ref)
        if not args.no_batched:
            log("batched two-orbital Kanamori")
            extra["batched_kanamori"] = bench_batched_kanamori(edgpu, W, ref)
        if not args.no_cpu_baseline:
            ms = cpu_kanamori_step_ms(W, 3, threads)
            extra["kanamori"].update(cpu_ms_per_lanczos_step=ms, cpu_cores=threads,
                                     cpu_solves_per_s=1e3 / (ms * (2 * 200 + 1 + 8 * 150)))
    log("Ns=6 DMFT loop")
    extra["dmft_ns6"] = bench_dmft_ns6(edgpu, W, not args.no_cpu_baseline, threads)

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle)")
        cpu = cpu_baseline(em, ek, vmk, U, mu, args.cpu_steps, threads)
        if args.cpu_full:
            log("CPU baseline: whole oracle solves")
            cpu["full_solve"] = cpu_full_solve(em, ek, vmk, U, mu, threads)

    lat_ms = np.array(lat) * 1e3
    res = {
        "metric": "H·v HBM GB/s + ED solves/s at Ns=14; p99 latency, FP32-vec vs FP64 at ΔE0≤1e-6",
        "value": args.steps / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-vectors/f64-reductions",
        "data": "synthetic seeded bath (PCG64 seed 1), no external data",
        "config": {"workload": "single-orbital Bethe impurity Ns=14 (Nbath=13), sector (7,7), "
                               "D=11778624, U=%g; 200 GS Lanczos steps + 4 chains x 150 steps, "
                               "2048 Matsubara points, beta=100" % U,
                   "parallelism": "single GPU"},
        "hv_gbs": hv_gbs,
        "hv_us": hv_ms * 1e3,
        "hv_schedule": best,
        "hv_schedule_us": {str(k): round(v, 2) for k, v in sched_us.items()},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "gs_phase_ms": st["gs_ms"],
        "green_phase_ms": st["green_ms"],
        "lanczos_steps_per_solve": steps_per_solve,
        "lanczos_steps_per_s": steps_per_solve * args.steps / wall,
        "e0": e32,
        "gs_residual": info32["residual"],
        "roofline": {"bound": "hbm", "achieved": hv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": hv_gbs / HBM_PEAK_GBS, "traffic": measured_traffic()},
        "cpu_baseline": cpu,
    }
    res.update(out)
    res.update(extra)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
