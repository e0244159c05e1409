"""Autotuner of the H.v schedules, by measurement on the device.

Single-problem path: the sector columns per workgroup (1, 2, 4) and the
persistent table-in-LDS forms (16, 32, 64) of edg_set_tuning, timed with
edg_time_hv (HIP events on the solver's stream) for the sector in use.

Batched path: per size class (the number of sites of the class's problems),
the tile shape of edg_batch_set_schedule: short-column tiles of 1024 / 2048
/ 4096 elements or one sector column per 512-thread tile, timed by solving a
sample of the class's problems through edgpu.Batch.

Results can be kept in a JSON file keyed by the problem shape, so a caller
tunes once per device and workload.
"""
import json
import os
import time

SINGLE_CANDIDATES = (1, 2, 4, 16, 32, 64)
BATCH_TILES = (1024, 2048, 4096)


def tune_single(solver, em, ek, vmk, ne_up, ne_dn, candidates=SINGLE_CANDIDATES, warmup=3, reps=20,
                fp64_accumulate=None):
    """Time each schedule of the H.v on sector (ne_up, ne_dn), install the
    fastest on `solver` and return (best, {schedule: microseconds})."""
    acc = fp64_accumulate
    if acc is None:  # the context default: FP64 row sums for FP32 vectors with norb >= 3
        acc = solver.precision == "fp32" and solver.norb >= 3
    times = {}
    for c in candidates:
        solver.set_tuning(c, acc)
        ms, _ = solver.time_hv(em, ek, vmk, ne_up, ne_dn, warmup=warmup, reps=reps)
        times[c] = ms * 1e3
    best = min(times, key=times.get)
    solver.set_tuning(best, acc)
    return best, times


def _classes(problems):
    """Problems grouped by their number of sites."""
    out = {}
    for p in problems:
        out.setdefault(p.nbath + 1, []).append(p)
    return dict(sorted(out.items()))


def tune_batch(edgpu, problems, omega, precision="fp32", sample=8, tiles=BATCH_TILES, steps=(80, 60),
               log=None, graphs=True):
    """Per size class (number of sites) of `problems`, time the candidate
    tile shapes on up to `sample` of its problems (with `steps` =
    (ground-state, chain) Lanczos lengths) and return the schedule
    [(nsite_max, tile_elems, wide), ...] plus the timings
    {nsite: {candidate: seconds}}."""
    classes = _classes(problems)
    timings, chosen = {}, {}
    for nsite, probs in classes.items():
        sub = []
        for p in probs[:sample]:
            q = edgpu.Problem(p.nbath, p.U, p.mu, p.em.reshape(1, 2, order="F"),
                              p.ek.reshape(p.nbath, 2, order="F"), p.vmk.reshape(1, p.nbath, 2, order="F"),
                              p.nup, p.ndn, maxnstep=steps[0], maxnstep_green=steps[1], seed=p.seed)
            sub.append(q)
        cands = [(t, 0) for t in tiles] + ([(2048, 1)] if nsite >= 10 else [])
        res = {}
        for tile, wide in cands:
            sched = [(64, tile, wide)]
            best_t = None
            for _ in range(2):
                t = _timed(edgpu, sub, omega, precision, sched, graphs)
                best_t = t if best_t is None else min(best_t, t)
            res[(tile, wide)] = best_t
            if log:
                log(f"autotune batch nsite={nsite}: tile={tile} wide={wide}: {best_t * 1e3:.1f} ms")
        timings[nsite] = res
        chosen[nsite] = min(res, key=res.get)
    keys = sorted(chosen)
    sched = [(k if i + 1 < len(keys) else 64, chosen[k][0], chosen[k][1]) for i, k in enumerate(keys)]
    return sched, timings


def _timed(edgpu, problems, omega, precision, sched, graphs):
    b = edgpu.Batch(omega, precision=precision, graphs=graphs)
    b.set_schedule(sched)
    t = time.perf_counter()
    for p in problems:
        b.submit(p)
    b.run()
    b.poll()
    dt = time.perf_counter() - t
    b.close()
    return dt


def load(path):
    """The tuning cache at `path` ({} when absent)."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save(path, cache):
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def single_key(solver, ne_up, ne_dn):
    return f"single:norb={solver.norb}:nbath={solver.nbath}:sector={ne_up},{ne_dn}:{solver.precision}"


def cached_single(solver, em, ek, vmk, ne_up, ne_dn, path):
    """tune_single through the cache at `path`: measured once per shape."""
    cache = load(path)
    key = single_key(solver, ne_up, ne_dn)
    if key in cache:
        solver.set_tuning(cache[key]["best"], cache[key]["fp64_accumulate"])
        return cache[key]["best"], cache[key]["times_us"]
    best, times = tune_single(solver, em, ek, vmk, ne_up, ne_dn)
    acc = solver.precision == "fp32" and solver.norb >= 3
    cache[key] = dict(best=best, fp64_accumulate=acc, times_us={str(k): v for k, v in times.items()})
    save(path, cache)
    return best, times
