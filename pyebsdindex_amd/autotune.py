"""Autotuner for the launch shape of one plan (pattern size × dtype × background mode).

Sweeps the gather's patterns per LDS word, work-group size, row segments (integer sums), LDS
row pitch and line unit, then the peak kernel that runs: the streamed kernel's theta block (columns per
wavefront, hence wavefronts per pattern), or -- where only the tiled kernel applies -- its
theta tile width and work-group size.  Each shape is timed on a sample of the caller's
patterns with HIP events; the fastest is applied to the plan and remembered per (device, H,
W, nTheta, nRho, dtype, background mode).  Every shape gives the same results.
"""
import itertools

_CACHE = {}


def _key(bd, plan, dtype):
    import torch
    name = torch.cuda.get_device_name(plan.device)
    return (name, plan.H, plan.W, plan.nT, plan.nR, str(dtype), bd.background_mode,
            bd.backgroundsub is not None)


def _time(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def autotune(bd, pats, gpu_id=None, sample=2048, reps=3, gather_packs=(1, 2, 4, 8),
             gather_threads=(512, 1024), row_segments=(1, 2, 3, 4), peak_tiles=(24, 30, 36, 45, 60, 90),
             peak_threads=(256, 512), line_units=(1, 8, 16), verbose=False):
    """Tune ``bd``'s device plan on (a sample of) the device tensor ``pats``; returns
    ``{"gather_pack", "gather_threads", "row_segments", "stage_pitch", "line_unit", "peak_block",
    "peak_tile", "peak_threads", "gather_ms", "peak_ms"}`` (``row_segments`` 0: the plan's
    automatic choice; ``peak_block`` 0 where the tiled peak kernel runs, whose
    ``peak_tile``/``peak_threads`` are 0 where the streamed kernel runs)."""
    import torch
    from . import _hip
    plan = bd.device_plan(gpu_id)
    x = pats[:sample]
    key = _key(bd, plan, x.dtype)
    if key in _CACHE:
        best = _CACHE[key]
        plan.set_tuning(best["gather_pack"], best["gather_threads"], best["peak_tile"], best["peak_threads"])
        plan.set_row_segments(best["row_segments"])
        plan.set_stage_pitch(best["stage_pitch"])
        plan.set_line_unit(best["line_unit"])
        if best["peak_block"]:
            plan.set_peak_block(best["peak_block"])
        return dict(best)
    xx, flags = bd._prepare(plan, x)
    rad = torch.empty((xx.shape[0], plan.nR, plan.nT), dtype=torch.float32, device=plan.device)
    out = plan.alloc_out(xx.shape[0])
    best_g = None
    for gp, gt in itertools.product(gather_packs, gather_threads):
        try:
            plan.set_tuning(gp, gt, 0, 0)
            ms = _time(lambda: plan.radon(xx, flags=flags, out=rad), reps)
        except _hip.HipLibraryError:
            continue
        if verbose:
            print(f"  gather pack={gp} threads={gt}: {ms:.4f} ms")
        if best_g is None or ms < best_g[0]:
            best_g = (ms, gp, gt)
    plan.set_tuning(best_g[1], best_g[2], 0, 0)
    # row segments of the integer gather (more patterns per LDS word in fewer rows), per pack
    best_g = best_g + (0,)
    if x.dtype != torch.float32 and not (bd.background_mode == 'exact' and bd.backgroundsub is not None):
        for S, gp in itertools.product(row_segments, gather_packs):
            try:
                plan.set_row_segments(S)
                plan.set_tuning(gp, 0, 0, 0)
                ms = _time(lambda: plan.radon(xx, flags=flags, out=rad), reps)
            except _hip.HipLibraryError:
                continue
            if verbose:
                print(f"  gather segments={S} pack={gp}: {ms:.4f} ms")
            if ms < best_g[0]:
                best_g = (ms, gp, best_g[2], S)
    plan.set_row_segments(best_g[3])
    plan.set_tuning(best_g[1], 0, 0, 0)
    # LDS row pitch of the staged patterns (bank spread of the lines' pixels)
    best_s = None
    for pitch in sorted({plan.W, plan.W + 1, plan.W + 2, plan.W + 3, plan.W + 5}):
        try:
            plan.set_stage_pitch(pitch)
            ms = _time(lambda: plan.radon(xx, flags=flags, out=rad), reps)
        except _hip.HipLibraryError:
            continue
        if verbose:
            print(f"  gather pitch={pitch}: {ms:.4f} ms")
        if best_s is None or ms < best_s[0]:
            best_s = (ms, pitch)
    plan.set_stage_pitch(best_s[1])
    best_g = (min(best_g[0], best_s[0]),) + best_g[1:]
    # line order in units of neighbouring cells (whole output sectors against padding)
    best_u = None
    for unit in line_units:
        try:
            plan.set_line_unit(unit)
            ms = _time(lambda: plan.radon(xx, flags=flags, out=rad), reps)
        except _hip.HipLibraryError:
            continue
        if verbose:
            print(f"  gather line unit={unit}: {ms:.4f} ms")
        if best_u is None or ms < best_u[0]:
            best_u = (ms, unit)
    plan.set_line_unit(best_u[1])
    best_g = (min(best_g[0], best_u[0]),) + best_g[1:]
    plan.radon(xx, flags=flags, out=rad)
    best_p = None
    block = plan.stream_block() if getattr(plan, "peak_kind", 0) != 1 else None
    if block is not None:
        # the streamed kernel: its theta block sets the wavefronts per pattern
        default, widest = block
        n0 = -(-plan.nT // widest)
        cols = sorted({-(-plan.nT // nb) for nb in range(n0, min(16, n0 + 8) + 1)} | {default}, reverse=True)
        for c in cols:
            try:
                plan.set_peak_block(c)
                ms = _time(lambda: plan.peaks(rad, out=out), reps)
            except _hip.HipLibraryError:
                continue
            if verbose:
                print(f"  peaks (streamed) block={c} waves={-(-plan.nT // c)}: {ms:.4f} ms")
            if best_p is None or ms < best_p[0]:
                best_p = (ms, c)
        plan.set_peak_block(best_p[1])
        best = {"gather_pack": best_g[1], "gather_threads": best_g[2], "row_segments": best_g[3],
                "stage_pitch": best_s[1], "line_unit": best_u[1], "peak_block": best_p[1], "peak_tile": 0,
                "peak_threads": 0,
                "gather_ms": best_g[0], "peak_ms": best_p[0]}
        _CACHE[key] = best
        return dict(best)
    for pt, pth in itertools.product(peak_tiles, peak_threads):
        try:
            plan.set_tuning(0, 0, pt, pth)
            ms = _time(lambda: plan.peaks(rad, out=out), reps)
        except _hip.HipLibraryError:
            continue
        if verbose:
            print(f"  peaks tile={pt} threads={pth}: {ms:.4f} ms")
        if best_p is None or ms < best_p[0]:
            best_p = (ms, pt, pth)
    plan.set_tuning(0, 0, best_p[1], best_p[2])
    best = {"gather_pack": best_g[1], "gather_threads": best_g[2], "row_segments": best_g[3],
            "stage_pitch": best_s[1], "line_unit": best_u[1], "peak_block": 0, "peak_tile": best_p[1],
            "peak_threads": best_p[2], "gather_ms": best_g[0], "peak_ms": best_p[0]}
    _CACHE[key] = best
    return dict(best)
