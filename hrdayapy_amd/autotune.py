"""Autotuner: the brick shape and the steps per graph launch of a monodomain
engine, chosen by timing the engine itself on the GPU.

No counterpart in the reference, which has a single sparse-matrix path.  The
tile shape decides the LDS tile, the halo share and the padding of partly
filled bricks; the steps per launch decide how many coupled steps one HIP
graph replay covers (longer graphs amortise the replay's fixed cost, shorter
ones let host work such as ECG recording interleave).
"""
from __future__ import annotations

import time

import torch

from .layout import BRICK_SHAPES


def time_steps(eng, steps: int, chunk: int) -> float:
    """Seconds per step of eng.run with `chunk` steps per graph launch (one
    untimed chunk first: capture and warm-up)."""
    eng.run(chunk, graph_chunk=chunk)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.run(steps, graph_chunk=chunk)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def tune(build, shapes=BRICK_SHAPES, chunks=(50,), steps: int = 200, log=None):
    """build(brick) -> a fresh engine.  Every brick shape is timed at chunks[0]
    steps per launch; the fastest shape is then timed at every chunk in
    `chunks` (each even, dividing `steps`).  Returns (brick, chunk, table) with
    table = {"shape_us": {"4x4x4": us/step, ...}, "chunk_us": {chunk: us/step}}."""
    if any(c <= 0 or c % 2 or steps % c for c in chunks):
        raise ValueError("every chunk must be even (V ping-pong) and divide the timed step count")
    shape_us, best = {}, None
    for b in shapes:
        eng = build(b)
        us = time_steps(eng, steps, chunks[0]) * 1e6
        shape_us["x".join(map(str, b))] = round(us, 1)
        if log:
            log(f"[autotune] brick {b}: {us:.1f} us/step")
        if best is None or us < best[1]:
            best = (b, us, eng)
        else:
            del eng
        torch.cuda.empty_cache()
    brick, _, eng = best
    chunk_us = {}
    for c in chunks:
        chunk_us[c] = round(time_steps(eng, steps, c) * 1e6, 1)
    chunk = min(chunk_us, key=chunk_us.get)
    del eng
    torch.cuda.empty_cache()
    return brick, chunk, {"shape_us": shape_us, "chunk_us": chunk_us}
