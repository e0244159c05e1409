"""Request loop over the batched solver (edgpu.Batch): problems arrive on a
clock, are admitted by device bytes as memory frees up, and every problem in
flight advances through shared batched Lanczos launches (continuous
batching).  Latency of a request = time its result is returned - its arrival
time.
"""
import time

import numpy as np


def run_batched(edgpu, problems, omega, precision="fp32", budget_bytes=0, max_steps=32, graphs=True,
                schedule=None):
    """Solve a list of problems submitted at once; returns (results, wall
    seconds, engine stats).  schedule: tile schedule by size class
    (edgpu.autotune.tune_batch), default the engine's."""
    b = edgpu.Batch(omega, precision=precision, budget_bytes=budget_bytes, graphs=graphs)
    if schedule:
        b.set_schedule(schedule)
    t0 = time.perf_counter()
    ids = [b.submit(p) for p in problems]
    pending = len(ids)
    res = {}
    while pending:
        pending = b.step(max_steps)
        for r in b.poll():
            res[r["id"]] = r
    for r in b.poll():
        res[r["id"]] = r
    wall = time.perf_counter() - t0
    st = b.stats()
    b.close()
    return [res[i] for i in ids], wall, st


def serve(edgpu, requests, omega, precision="fp32", budget_bytes=0, max_steps=8, graphs=True,
          drain_limit=120.0, log=None, schedule=None, sample_every=1.0, results=None):
    """Replay [(t_arrival, Problem)] against the wall clock.  Returns a dict
    with per-request latencies (seconds, in completion order; index[i] is the
    position in `requests` of the i-th completed one), solves/s over
    the run, completion counts, engine stats and the queue depth over time:
    depth = [(t, submitted, completed, pending)] every sample_every seconds,
    pending = requests submitted and not yet finished (queued or on the GPU).
    With results a dict, each finished request's result (edgpu.Batch.poll)
    is stored there under its index in `requests`."""
    b = edgpu.Batch(omega, precision=precision, budget_bytes=budget_bytes, graphs=graphs)
    if schedule:
        b.set_schedule(schedule)
    arrival, index = {}, {}
    lat, order = [], []
    depth = []
    nxt = 0
    t0 = time.perf_counter()
    last_log = t0
    next_sample = 0.0
    duration = requests[-1][0] if requests else 0.0
    pending = 0
    while True:
        now = time.perf_counter() - t0
        while nxt < len(requests) and requests[nxt][0] <= now:
            i = b.submit(requests[nxt][1])
            arrival[i], index[i] = requests[nxt][0], nxt
            nxt += 1
        if now >= next_sample:
            depth.append((now, nxt, len(lat), len(arrival)))
            next_sample += sample_every
        if nxt == len(requests) and pending == 0 and not arrival:
            break
        if now > duration + drain_limit:
            break
        pending = b.step(max_steps)
        now = time.perf_counter() - t0
        for r in b.poll():
            lat.append(now - arrival.pop(r["id"]))
            k = index.pop(r["id"])
            order.append(k)
            if results is not None:
                results[k] = r
        if pending == 0 and nxt < len(requests):
            time.sleep(max(0.0, min(requests[nxt][0] - now, 0.002)))
        if log and time.perf_counter() - last_log > 5.0:
            last_log = time.perf_counter()
            log(f"served t={now:.1f}s submitted={nxt} done={len(lat)} pending={pending}")
    wall = time.perf_counter() - t0
    depth.append((wall, nxt, len(lat), len(arrival)))
    st = b.stats()
    b.close()
    lat = np.array(lat)
    return dict(latencies=lat, index=np.array(order, dtype=np.int64), wall=wall, submitted=nxt,
                completed=len(lat),
                unfinished=len(arrival), stats=st, depth=depth,
                solves_per_s=len(lat) / wall if wall > 0 else 0.0,
                p50=float(np.percentile(lat, 50)) if len(lat) else None,
                p99=float(np.percentile(lat, 99)) if len(lat) else None)
