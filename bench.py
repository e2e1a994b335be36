"""Band-detection throughput on one MI355X (patterns/s), with roofline and CPU baseline.

Workload (default): 16,384 synthetic 120x120 uint16 Kikuchi-like FCC patterns, background
subtraction, Radon 180x90, 9x5 Radon-space kernel (tSigma=1, rSigma=2), 9 bands with
sub-pixel fit -- BandDetect.find_bands on patterns already resident in HBM.  One step =
one find_bands pass over the whole batch.

Prints ONE JSON line.  ``--gpus N`` with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "patterns/s (Radon+band detect), batched vs plain, at ≤1e-3-bin peak error"
LDS_NOMINAL_GBS = 256 * 128 * 2.4       # 256 CU x 128 B/clk x 2.4 GHz (the nominal figure once quoted)
HBM_PEAK_GBS = 8000.0
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4   # wave64 VALU instructions/s: 256 CU x 4 SIMD, 4 clocks each
PUBLISHED_C1 = 6020.0                   # find_bands, 1,000 60x60 patterns, OpenCL GPU (2019 Mac Pro),
                                        # doc/tutorials/ebsd_index_demo.ipynb:206-210


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # steady state: the first steps of a fresh process run slower (10 timed steps after 2
    # warm-up steps measured 2.36 ms/step, 100 after 10 measured 2.19 ms/step, same work)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--npats", type=int, default=16384)
    ap.add_argument("--size", type=int, default=120)
    ap.add_argument("--dtype", default="uint16")
    ap.add_argument("--no-background", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline duration")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the mixed-size batched measurement")
    ap.add_argument("--no-streaming", action="store_true", help="skip the pinned-host streaming measurement")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiation measurement")
    ap.add_argument("--no-nlpar", action="store_true", help="skip the NLPAR measurement")
    ap.add_argument("--no-tripvote", action="store_true", help="skip the triplet-vote measurement")
    ap.add_argument("--no-c1", action="store_true", help="skip the 1,000 x 60x60 uint8 measurement")
    ap.add_argument("--stream-patterns", type=int, default=1048576,
                    help="patterns streamed in the scan-scale measurement")
    ap.add_argument("--roofline-reps", type=int, default=20)
    ap.add_argument("--pmc-driver", action="store_true",
                    help="short fixed launch sequence for rocprofv3 --pmc: gather x3, peaks x3 on one "
                         "chunk, then a 256 MiB device copy x3 (the counter calibration)")
    ap.add_argument("--pmc-driver-probe", action="store_true",
                    help="the VALU issue probe alone for rocprofv3 --pmc: each of its four instruction "
                         "mixes once (an untimed and a timed launch of 2,048 work-groups x 4,096 steps)")
    ap.add_argument("--pmc-driver-nlpar", action="store_true",
                    help="short fixed NLPAR sequence for rocprofv3 --pmc: the 256x256 scan of 60x60 uint8 "
                         "patterns of the NLPAR measurement, sigma once, NLPAR twice (uint8 output)")
    ap.add_argument("--tune", default="", help="gather_pack,gather_threads,peak_tile,peak_threads (0 = keep)")
    ap.add_argument("--line-unit", type=int, default=0,
                    help="gather line order in units of this many neighbouring cells (0 = the plan's default)")
    ap.add_argument("--chunk", type=int, default=0,
                    help="patterns per Radon+peak pass (0: the library's automatic size, all "
                         "patterns while their Radons fit 1 GiB)")
    return ap.parse_args()


def lds_ceiling(reps=2):
    """Conflict-free LDS read bandwidth of this GPU (ds_read_b32/b64/b128, bytes/s), from
    the measurement-only probe library, the best of ``reps`` runs per width."""
    import ctypes
    import torch
    from pyebsdindex_amd.build import PROBE_LIB
    L = ctypes.CDLL(PROBE_LIB)
    L.ebsd_probe_lds_read.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
    s = torch.cuda.current_stream().cuda_stream
    out = {}
    for w in (4, 8, 16):
        best = 0.0
        for _ in range(reps):
            v = ctypes.c_double()
            if L.ebsd_probe_lds_read(w, 1024, 2048, ctypes.byref(v), s) != 0:
                raise RuntimeError("LDS probe failed")
            best = max(best, v.value)
        out[f"ds_read_b{8 * w}"] = round(best / 1e9, 1)
    return out


def valu_ceiling(reps=2):
    """Vector-instruction issue rate of this GPU (wave-instructions/s) for independent chains
    of v_fma_f32, of v_xor_b32/v_add_u32, of v_max_i32 with a DPP source and v_sub_u32, and of
    v_fma_f64, from the measurement-only probe library, the best of ``reps`` runs per kind."""
    import ctypes
    import torch
    from pyebsdindex_amd.build import PROBE_LIB
    L = ctypes.CDLL(PROBE_LIB)
    L.ebsd_probe_valu_issue.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
    s = torch.cuda.current_stream().cuda_stream
    out = {}
    for kind, name in ((0, "v_fma_f32"), (1, "v_xor_b32+v_add_u32"), (2, "v_max_i32_dpp+v_sub_u32"),
                       (3, "v_fma_f64")):
        best = 0.0
        for _ in range(reps):
            v = ctypes.c_double()
            # 2,048 work-groups of 256 threads: 8 waves on every SIMD of 256 CUs
            if L.ebsd_probe_valu_issue(kind, 2048, 4096, ctypes.byref(v), s) != 0:
                raise RuntimeError("VALU probe failed")
            best = max(best, v.value)
        out[name] = round(best / 1e9, 1)
    return out


def valu_counts(H, W, dtype, chunk):
    """VALU instructions per launch of the gather and the peak kernel from the committed
    rocprof counter summary of this workload, or None."""
    f = os.path.join(ROOT, "profiles", f"valu_counts_{H}x{W}_{dtype}.json")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        j = json.load(fh)
    return j if j.get("patterns_per_launch") == chunk else None


def c1_measure(dev, steps=50, warmup=10, cpu_seconds=8.0, seed=31):
    """The reference's own CPU-runnable case: 1,000 synthetic 60x60 uint8 FCC patterns, Radon
    180x90, 9 bands, no background -- on the device (resident), through the drop-in API
    (numpy in, structured array out, as EBSDIndexer calls it), and the oracle's
    restatement of the reference CPU path on the host cores."""
    import torch
    from pyebsdindex_amd import BandDetect
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    params = dict(nTheta=180, nRho=90, tSigma=1, rSigma=2, rhoMaskFrac=0.1, nBands=9)
    bd = BandDetect(patDim=(60, 60), **params)
    x = kikuchi_patterns(1000, (60, 60), "uint8", seed=seed, device=dev)
    host = x.cpu().numpy()
    out = bd.device_plan().alloc_out(1000)
    work = bd.device_plan().workspace(1000, 1000)
    for _ in range(warmup):
        bd.find_bands_device(x, out=out, work=work, chunk=1000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        bd.find_bands_device(x, out=out, work=work, chunk=1000)
    torch.cuda.synchronize()
    dev_rate = 1000 * steps / (time.perf_counter() - t0)
    bd.find_bands(host)
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        res = bd.find_bands(host)
    api_rate = 1000 * reps / (time.perf_counter() - t0)
    cpu = None
    if cpu_seconds > 0:
        pps, done, dt = cpu_baseline(host, None, params, cpu_seconds)
        cpu = {"value": round(pps, 1), "unit": "patterns/s", "cores": os.cpu_count(), "kind": "port",
               "sample": f"{done} patterns (the 1,000, cycled) in {dt:.1f} s, chunks of 512"}
    return {"workload": "1000 x 60x60 uint8 FCC patterns, Radon 180x90, kernel 9x5, 9 bands, no background",
            "device_patterns_per_s": round(dev_rate, 1), "api_patterns_per_s": round(api_rate, 1),
            "published_opencl_patterns_per_s": PUBLISHED_C1,
            "published_source": "doc/tutorials/ebsd_index_demo.ipynb:206-210 (OpenCL GPU of a 2019 Mac Pro)",
            "api_vs_published": round(api_rate / PUBLISHED_C1, 1),
            "valid_band_fraction": round(float((res["valid"] == 1).mean()), 4), "cpu_baseline": cpu}


def cpu_baseline(pats_host, bg, params, target_s):
    """Reference-faithful CPU restatement (oracle.ref_cpu) on a bounded sample."""
    from oracle import ref_cpu
    H, W = pats_host.shape[1:]
    ref = ref_cpu.RefBandDetect((H, W), **params)
    ref.backgroundsub = bg
    ref.find_bands(pats_host[:64])  # warm-up (also builds the C library if needed)
    done, off, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 < target_s:
        n = min(512, pats_host.shape[0] - off)
        ref.find_bands(pats_host[off:off + n], chunksize=512)
        done += n
        off = (off + n) % pats_host.shape[0]
    dt = time.perf_counter() - t0
    return done / dt, done, dt


def streaming_measure(dev, total=1048576, batch=8192, ring=4, seed=11):
    """Scan-scale streaming: 240x240 uint16 HCP (c/a 1.587) patterns, Radon 360x180, 12 bands,
    background subtraction; ``total`` patterns in ``batch``-pattern batches copied from pinned
    host memory on a side stream, overlapped with band detection.  The host holds a ring of
    ``ring`` distinct pinned batches that the stream cycles through."""
    import torch
    from pyebsdindex_amd import BandDetect
    from pyebsdindex_amd.streaming import StreamingBandDetect
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    shape = (240, 240)
    bd = BandDetect(patDim=shape, nTheta=360, nRho=180, tSigma=1, rSigma=2, rhoMaskFrac=0.1, nBands=12)
    sb = StreamingBandDetect(bd, batch=batch, dtype="uint16", depth=2)
    host = []
    for k in range(ring):
        x = kikuchi_patterns(batch, shape, "uint16", seed=seed + k, family="HCP", device=dev, chunk=1024)
        if k == 0:
            bd.collect_background(patsIn=x[:1024].cpu().numpy().astype(np.float64), nsample=1024)
        t = sb.pinned_batch()
        t.copy_(x)
        host.append(t)
        del x
    torch.cuda.synchronize()
    nb = max(1, total // batch)
    seq = [host[j % ring] for j in range(nb)]
    blocks = sb.pinned_blocks(nb)
    sb.run(seq[:2], blocks=blocks[:2])  # warm-up
    t0 = time.perf_counter()
    res = sb.run(seq, blocks=blocks)
    dt = time.perf_counter() - t0
    n = nb * batch
    # the same batches resident in HBM (no copy) and a plain pinned host->device copy
    dev_in = torch.empty_like(sb.slots[0])
    dev_in.copy_(host[0])
    out = sb.outs[0]
    bd.find_bands_device(dev_in, out=out, work=sb.work)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for j in range(4):
        bd.find_bands_device(dev_in, out=out, work=sb.work)
    torch.cuda.synchronize()
    resident = 4 * batch / (time.perf_counter() - t1)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for j in range(4):
        dev_in.copy_(host[j % ring], non_blocking=True)
    torch.cuda.synchronize()
    h2d = 4 * host[0].numel() * 2 / (time.perf_counter() - t2) / 1e9
    bytes_per = shape[0] * shape[1] * 2
    achieved_h2d = n * bytes_per / dt / 1e9
    same = all(bool(torch.equal(res[0][k], res[ring][k])) for k in ("maxloc", "valid")) if nb > ring else None
    return {"workload": f"{n} patterns of 240x240 uint16 HCP (c/a 1.587), Radon 360x180, 12 bands, background "
                        f"subtraction; {nb} batches of {batch} from pinned host memory (a ring of {ring} "
                        f"distinct batches), H2D copy on a side stream overlapped with band detection",
            "patterns_per_s": round(n / dt, 1), "seconds": round(dt, 3),
            "resident_patterns_per_s": round(resident, 1),
            "h2d_ceiling_patterns_per_s": round(h2d * 1e9 / bytes_per, 1),
            "overlap_efficiency": round((n / dt) / min(resident, h2d * 1e9 / bytes_per), 4),
            "roofline": {"bound": "h2d", "achieved": round(achieved_h2d, 2),
                         "peak": round(h2d, 2), "unit": "GB/s", "frac": round(achieved_h2d / h2d, 4),
                         "peak_kind": "plain pinned host->device copy measured in the same run"},
            "repeat_batches_identical": same,
            "from_file": file_streaming_measure(bd, host, batch=2048)}


def file_streaming_measure(bd, host, npats=16384, batch=2048, threads=16):
    """The same 240x240 uint16 workload fed from a UP2 file: ``npats`` patterns (the first
    pinned batches of ``host``, repeated) are written to a file in the temporary directory
    (so the page cache holds them), then streamed with ``StreamingBandDetect.run_file`` in
    ``batch``-pattern batches read straight into a pinned ring.  Also times the file reads
    alone (into the ring, no GPU work)."""
    import tempfile
    import torch
    from pyebsdindex_amd.ebsd_pattern import UPFile, get_pattern_file_obj
    from pyebsdindex_amd.streaming import StreamingBandDetect
    H, W = int(bd.patDim[-2]), int(bd.patDim[-1])
    with tempfile.TemporaryDirectory() as td:
        fn = os.path.join(td, "scan.up2")
        w = UPFile(fn)
        w.patternW, w.patternH, w.nCols, w.nRows, w.xStep, w.yStep = W, H, npats, 1, 1.0, 1.0
        w.write_header()
        off, k = 0, 0
        while off < npats:
            n = min(int(host[k % len(host)].shape[0]), npats - off)
            w.pat_writer(host[k % len(host)].numpy(), off, n, np.uint16)
            off, k = off + n, k + 1
        f = get_pattern_file_obj(fn)
        sb = StreamingBandDetect(bd, batch=batch, dtype="uint16", depth=3)
        ring = [sb.pinned_batch() for _ in range(sb.depth + 1)]
        blocks = sb.pinned_blocks(-(-npats // batch))
        sb.run_file(f, 0, 2 * batch, blocks=blocks, ring=ring, threads=threads)      # warm-up
        t0 = time.perf_counter()
        for j in range(npats // batch):
            f.read_into(ring[j % len(ring)], j * batch, batch, threads=threads)
        t_read = time.perf_counter() - t0
        t0 = time.perf_counter()
        res = sb.run_file(f, 0, npats, blocks=blocks, ring=ring, threads=threads)
        dt = time.perf_counter() - t0
        valid = float(torch.cat([r["valid"] for r in res]).float().mean())
    bytes_per = H * W * 2
    return {"workload": f"{npats} patterns of {H}x{W} uint16 in a UP2 file (page-cache resident), read in "
                        f"{batch}-pattern batches into a pinned ring with {threads} reader threads, then the "
                        f"same pipeline",
            "patterns_per_s": round(npats / dt, 1),
            "file_read_GBps": round(npats * bytes_per / t_read / 1e9, 2),
            "file_read_patterns_per_s": round(npats / t_read, 1),
            "valid_band_fraction": round(valid, 4)}


VALU_F32_PEAK_TFLOPS = 157.3            # MI355X published fp32 vector peak


def nlpar_measure(dev, nrows=256, ncols=256, shape=(60, 60), sr=3, reps=5, cpu_seconds=8.0, seed=17):
    """NLPAR (non-local pattern averaging) of a 256x256 scan of 60x60 uint8 patterns (16
    grains, Poisson noise), search radius 3, automask: the sigma and averaging kernels on
    the device-resident scan, timed with HIP events; the oracle's C restatement of the
    reference's Numba kernels timed on a 32x32 sub-scan on the host cores."""
    import torch
    from oracle import nlpar_ref
    from pyebsdindex_amd.nlpar import makeautomask, nlpar_device, pattern_max, sigma_device
    from pyebsdindex_amd.synthetic import grain_scan
    pats = grain_scan(nrows, ncols, shape, "uint8", seed=seed, grains=16, device=dev)
    idx = np.nonzero(makeautomask(*shape).ravel())[0]
    mx = pattern_max(pats)
    sig = sigma_device(pats, nrows, ncols, idx, mxval=mx)
    out = nlpar_device(pats, nrows, ncols, idx, sig, sr=sr, mxval=mx, out_dtype=torch.uint8)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t_sig = t_avg = 0.0
    for _ in range(reps):
        ev[0].record()
        sigma_device(pats, nrows, ncols, idx, mxval=mx)
        ev[1].record()
        nlpar_device(pats, nrows, ncols, idx, sig, sr=sr, mxval=mx, out_dtype=torch.uint8)
        ev[2].record()
        torch.cuda.synchronize()
        t_sig += ev[0].elapsed_time(ev[1]) / reps
        t_avg += ev[1].elapsed_time(ev[2]) / reps
    n = nrows * ncols
    M, P, win = idx.size, shape[0] * shape[1], (2 * sr + 1) ** 2
    flops = (win - 1) * M * 3 + win * P * 2          # distances (sub, mul, add) + weighted sum (mul, add)
    tflops = n * flops / (t_avg * 1e-3) / 1e12
    # CPU: the restated Numba kernels (OpenMP over all host cores) on a 32x32 sub-scan
    sub = pats.reshape(nrows, ncols, -1)[:32, :32].reshape(32 * 32, -1).float().cpu().numpy()
    nlpar_ref.nlpar(sub, 32, 32, idx, nlpar_ref.sigma(sub, 32, 32, idx)[0], sr=sr)   # warm-up
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < cpu_seconds:
        s = nlpar_ref.sigma(sub, 32, 32, idx)[0]
        nlpar_ref.nlpar(sub, 32, 32, idx, s, sr=sr)
        done += 32 * 32
    cpu = None
    if done:
        cpu = {"value": round(done / (time.perf_counter() - t0), 1), "unit": "patterns/s", "cores": os.cpu_count(),
               "kind": "port", "sample": f"{done} patterns: the 32x32 corner of the same scan, sigma + NLPAR, "
                                         f"OpenMP over rows/columns as numba.prange"}
    return {"workload": f"NLPAR of a {nrows}x{ncols} scan of {shape[0]}x{shape[1]} uint8 patterns (16 grains, "
                        f"Poisson noise), search radius {sr}, lambda 0.7, automask ({M} pixels), uint8 output",
            "patterns_per_s": round(n / ((t_sig + t_avg) * 1e-3), 1),
            "kernels_ms": {"k_nlpar_sigma": round(t_sig, 4), "k_nlpar": round(t_avg, 4)},
            "roofline": {"bound": "valu", "kernel": "k_nlpar", "achieved": round(tflops, 2),
                         "peak": VALU_F32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tflops / VALU_F32_PEAK_TFLOPS, 4),
                         "algorithmic_flops_per_pattern": flops},
            "cpu_baseline": cpu,
            "output_checksum": int(out.to(torch.int64).sum().item())}


def tripvote_measure(dev, npats=65536, nbands=9, reps=20, cpu_seconds=6.0, seed=23):
    """Triplet vote of 65,536 patterns x 9 bands against the FCC library (pole families
    {002} {111} {022} {113}, angTol 2): band normals of 4,096 random orientations with
    0.3 degree noise, 3 % of the bands invalid; the device vote timed with HIP events on
    the launching stream; the oracle's C restatement of the reference's Numba kernel timed
    on a sample of the same normals on the host cores."""
    import torch
    from oracle import tripvote_ref
    from pyebsdindex_amd import tripletvote as tv
    from pyebsdindex_amd.synthetic import cubic_band_normals
    ind = tv.addphase("FCC")
    bn, fam_true = cubic_band_normals(ind.families, npats, nbands, seed=seed, noise_deg=0.3)
    rng = np.random.default_rng(seed)
    inten = rng.uniform(0.2, 1.0, (npats, nbands)).astype(np.float32)
    inten[rng.random((npats, nbands)) < 0.03] = 0.0
    bnd = torch.as_tensor(bn, device=dev)
    intd = torch.as_tensor(inten, device=dev)
    out = ind.tripvote(bnd, intd)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        out = ind.tripvote(bnd, intd)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    ntrip = int(ind.angtriplets["angles"].shape[0])
    tests = math.comb(nbands, 3) * ntrip
    agree = float((out[1].cpu().numpy() == fam_true).mean())
    # CPU: the restated Numba kernel (OpenMP over patterns, all host cores) on a sample
    sample = 2048
    tripvote_ref.tripvote(bn[:64], inten[:64], ind.lut, 2.0, ind.angtriplets["angles"],
                          ind.angtriplets["familyid"], ind.npolefamilies)
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < cpu_seconds:
        s0 = done % (npats - sample)
        tripvote_ref.tripvote(bn[s0:s0 + sample], inten[s0:s0 + sample], ind.lut, 2.0, ind.angtriplets["angles"],
                              ind.angtriplets["familyid"], ind.npolefamilies)
        done += sample
    cpu = None
    if done:
        cpu = {"value": round(done / (time.perf_counter() - t0), 1), "unit": "patterns/s", "cores": os.cpu_count(),
               "kind": "port", "sample": f"{done} patterns in blocks of {sample}, OpenMP over patterns as numba.prange"}
    # the whole BandIndexer.bandindex on the device: vote, band assignment, QUEST, cm
    bi = ind.bandindex(bnd, intd)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        bi = ind.bandindex(bnd, intd)
    e1.record(st)
    torch.cuda.synchronize()
    bi_ms = e0.elapsed_time(e1) / reps
    nmatch = bi[4].cpu().numpy()
    cpu_bi = None
    if cpu_seconds > 0:
        sample_bi = 1024
        tripvote_ref.bandindex(ind, bn[:64], inten[:64])
        done_bi, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            s0 = done_bi % (npats - sample_bi)
            tripvote_ref.bandindex(ind, bn[s0:s0 + sample_bi], inten[s0:s0 + sample_bi])
            done_bi += sample_bi
        cpu_bi = {"value": round(done_bi / (time.perf_counter() - t0), 1), "unit": "patterns/s",
                  "cores": os.cpu_count(), "kind": "port",
                  "sample": f"{done_bi} patterns in blocks of {sample_bi}: the oracle's vote, assignment, QUEST "
                            f"weights and fit, cm (C, OpenMP over patterns)"}
    bandindex = {"workload": "BandIndexer.bandindex on the same bands: vote, band assignment, QUEST orientation, "
                             "confidence metric (nband_earlyexit 8)",
                 "patterns_per_s": round(npats / (bi_ms * 1e-3), 1), "ms": round(bi_ms, 4),
                 "indexed_fraction": round(float((nmatch >= 3).mean()), 4),
                 "mean_matched_bands": round(float(nmatch.mean()), 3), "cpu_baseline": cpu_bi}
    return {"workload": f"triplet vote of {npats} patterns x {nbands} bands, FCC library ({ntrip} triplets), "
                        f"angTol 2, 0.3 degree noise, 3 % invalid bands",
            "patterns_per_s": round(npats / (ms * 1e-3), 1), "kernel_ms": round(ms, 4),
            "library_row_tests_per_pattern": tests,
            "row_tests_per_s": round(npats * tests / (ms * 1e-3), 1),
            "band_family_agreement": round(agree, 4),
            "cpu_baseline": cpu, "bandindex": bandindex}


def nlpar_pmc_driver(dev, nrows=256, ncols=256, shape=(60, 60), sr=3, seed=17):
    """The launches of nlpar_measure without timing or CPU work, for counter runs."""
    import torch
    from pyebsdindex_amd.nlpar import makeautomask, nlpar_device, pattern_max, sigma_device
    from pyebsdindex_amd.synthetic import grain_scan
    pats = grain_scan(nrows, ncols, shape, "uint8", seed=seed, grains=16, device=dev)
    idx = np.nonzero(makeautomask(*shape).ravel())[0]
    mx = pattern_max(pats)
    sig = sigma_device(pats, nrows, ncols, idx, mxval=mx)
    sums = []
    for _ in range(2):
        out = nlpar_device(pats, nrows, ncols, idx, sig, sr=sr, mxval=mx, out_dtype=torch.uint8)
        sums.append(int(out.to(torch.int64).sum().item()))
    torch.cuda.synchronize()
    print(json.dumps({"pmc_driver_nlpar": True, "patterns": nrows * ncols, "mask_pixels": int(idx.size),
                      "output_checksums": sums}), flush=True)


def differentiation_measure(dev, n=256, reps=5, seed=13):
    """256 float32 120x120 FCC patterns at a hidden pattern centre: forward (Radon,
    convolution, 9-band sub-pixel peaks) + backward (band-fit Jacobian, transposed
    convolution, backprojection, Radon recomputed) throughput, the PC gradient of the
    band-position loss, and Gauss-Newton refinement from a centre perturbed by 0.01."""
    import torch
    from pyebsdindex_amd import BandDetect
    from pyebsdindex_amd.differentiable import BandPeaks, refine_pc
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    shape = (120, 120)
    bd = BandDetect(patDim=shape, nTheta=180, nRho=90, tSigma=1, rSigma=2, rhoMaskFrac=0.1, nBands=9)
    true = np.array([0.512, 0.487, 0.634])
    pats, meta = kikuchi_patterns(n, shape, "float32", seed=seed, family="FCC", pc=true, device=dev,
                                  return_meta=True)
    normals = torch.as_tensor(meta["g_det"], device=dev)
    w = torch.randn((n, 9), device=dev, generator=torch.Generator(device=dev).manual_seed(seed))

    def step():
        x = pats.detach().requires_grad_(True)
        rho, theta, valid = BandPeaks.apply(x, bd)
        ((w * rho).sum() + (w * theta).sum()).backward()
        return x.grad

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    fb = reps * n / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for _ in range(reps):
        with torch.no_grad():
            BandPeaks.apply(pats, bd)
    torch.cuda.synchronize()
    fw = reps * n / (time.perf_counter() - t0)
    pc0 = true + np.array([0.01, -0.01, 0.01])
    t0 = time.perf_counter()
    pc, hist = refine_pc(bd, pats, normals, pc0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    pc = pc.cpu().numpy()
    return {"workload": f"{n} float32 120x120 FCC patterns at a hidden pattern centre, Radon 180x90, 9 bands",
            "forward_backward_patterns_per_s": round(fb, 1), "forward_patterns_per_s": round(fw, 1),
            "pc_true": true.tolist(), "pc_start": pc0.tolist(), "pc_refined": [round(float(v), 6) for v in pc],
            "pc_error_start": float(np.abs(pc0 - true).max()), "pc_error_refined": float(np.abs(pc - true).max()),
            "refine_iterations": len(hist) - 1, "refine_seconds": round(dt, 4),
            "loss_start": hist[0][1], "loss_refined": hist[-1][1]}


def batched_measure(dev, steps, warmup, npats=32768, seed=7, autotune_sets=True):
    """Mixed sizes {60,80,120,160}^2 x {uint8, uint16} (equal shares), BCC bands, one pattern
    centre per size (x*, y* ~ U(.45,.55), z* ~ U(.55,.7)), a plan per size: per-size plain
    launches vs one batched persistent launch pair replayed as a hipGraph."""
    import torch
    from pyebsdindex_amd import BandDetect
    from pyebsdindex_amd.autotune import autotune
    from pyebsdindex_amd.batch import BatchBandDetect
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    rng = np.random.default_rng(seed)
    sizes = [60, 80, 120, 160]
    dtypes = ["uint8", "uint16"]
    per = npats // (len(sizes) * len(dtypes))
    params = dict(nTheta=180, nRho=90, tSigma=1, rSigma=2, rhoMaskFrac=0.1, nBands=9)
    dets, pats, tuned = [], [], {}
    for si, sz in enumerate(sizes):
        pc = (rng.uniform(0.45, 0.55), rng.uniform(0.45, 0.55), rng.uniform(0.55, 0.7))
        for di, dt in enumerate(dtypes):
            x = kikuchi_patterns(per, (sz, sz), dt, seed=seed + 10 * si + di, family="BCC", pc=pc,
                                 device=dev, chunk=1024)
            bd = BandDetect(patDim=(sz, sz), **params)
            if autotune_sets:
                tuned[f"{sz}x{sz} {dt}"] = autotune(bd, x, sample=1024)
            dets.append(bd)
            pats.append(x)
    torch.cuda.synchronize()
    outs = [bd.device_plan().alloc_out(x.shape[0]) for bd, x in zip(dets, pats)]
    works = [bd.device_plan().workspace(x.shape[0]) for bd, x in zip(dets, pats)]

    def plain(chunk=None):
        for bd, x, o, w in zip(dets, pats, outs, works):
            bd.find_bands_device(x, out=o, work=w, chunk=chunk)

    def time_plain(chunk):
        for _ in range(warmup):
            plain(chunk)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            plain(chunk)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    plain_s = time_plain(None)      # the library's pass size (all of a set's patterns)
    # the same per-size launches at the reference's default chunk (band_detect.py:427-434)
    plain512_s = time_plain(512)
    batch = BatchBandDetect(dets, pats)
    gs = torch.cuda.Stream(dev)
    for _ in range(warmup):
        batch.graph_launch(gs)
    gs.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.graph_launch(gs)
    gs.synchronize()
    graph_s = (time.perf_counter() - t0) / steps
    for _ in range(warmup):
        batch.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run()
    torch.cuda.synchronize()
    run_s = (time.perf_counter() - t0) / steps
    # the batched bins and sub-pixel positions against the per-size launches
    dmax = 0.0
    same = True
    for o, ob in zip(outs, batch.outs):
        same &= bool(torch.equal(o["maxloc"], ob["maxloc"]))
        dmax = max(dmax, float((o["aveloc"] - ob["aveloc"]).abs().max()))
    n = sum(int(x.shape[0]) for x in pats)
    return {"workload": f"{n} patterns, {{60,80,120,160}}^2 x {{uint8,uint16}} in equal shares, BCC, "
                        f"one pattern centre and plan per size, Radon 180x90, 9 bands",
            "plain_patterns_per_s": round(n / plain_s, 1), "batched_patterns_per_s": round(n / run_s, 1),
            "graph_patterns_per_s": round(n / graph_s, 1), "graph_vs_plain": round(plain_s / graph_s, 3),
            "plain_ms": round(plain_s * 1e3, 4), "graph_ms": round(graph_s * 1e3, 4),
            "plain_chunk512_patterns_per_s": round(n / plain512_s, 1),
            "graph_vs_plain_chunk512": round(plain512_s / graph_s, 3),
            "peak_bins_identical": same, "max_subpixel_diff_bins": dmax, "autotuned": tuned}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build targets one GPU; --gpus {args.gpus} is not supported "
              f"(patterns are independent: N GPUs would each run this bench on 1/N of the patterns)", file=sys.stderr)
        sys.exit(2)
    import torch
    from pyebsdindex_amd import BandDetect
    from pyebsdindex_amd.synthetic import kikuchi_patterns

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.pmc_driver_nlpar:
        nlpar_pmc_driver(dev)
        return
    if args.pmc_driver_probe:
        torch.zeros(1, device=dev)
        print(json.dumps({"pmc_driver_probe": True, "valu_ceiling_GIPS": valu_ceiling(reps=1)}), flush=True)
        return
    H = W = args.size
    params = dict(nTheta=180, nRho=90, tSigma=1, rSigma=2, rhoMaskFrac=0.1, nBands=9)
    bd = BandDetect(patDim=(H, W), **params)
    print(f"[bench] generating {args.npats} {H}x{W} {args.dtype} patterns", file=sys.stderr, flush=True)
    pats = kikuchi_patterns(args.npats, (H, W), args.dtype, seed=2024, device=dev, chunk=1024)
    torch.cuda.synchronize()
    sample = pats[:1024].cpu().numpy()
    bg = None
    if not args.no_background:
        bd.collect_background(patsIn=sample.astype(np.float64), nsample=1024)
        bg = bd.backgroundsub
    plan = bd.device_plan()
    if args.tune:
        plan.set_tuning(*[int(v) for v in args.tune.split(",")])
    if args.line_unit:
        plan.set_line_unit(args.line_unit)
    info = plan.info()
    N = pats.shape[0]
    out = plan.alloc_out(N)
    chunk = plan.auto_chunk(N) if args.chunk <= 0 else max(1, min(args.chunk, N))
    work = plan.workspace(N, chunk)
    print(f"[bench] plan {info}", file=sys.stderr, flush=True)

    if args.pmc_driver:
        x, flags = bd._prepare(plan, pats[:chunk])
        rad = torch.empty((chunk, bd.nRho, bd.nTheta), dtype=torch.float32, device=dev)
        o2 = plan.alloc_out(chunk)
        for _ in range(3):
            plan.radon(x, flags=flags, out=rad)
        for _ in range(3):
            plan.peaks(rad, out=o2)
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_driver": True, "chunk": chunk, "radon_bytes_per_gather": rad.numel() * 4,
                          "input_bytes_per_gather": x.numel() * x.element_size(),
                          "copy_bytes": src.numel() * 4}), flush=True)
        return

    for _ in range(args.warmup):
        bd.find_bands_device(pats, out=out, work=work, chunk=chunk)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bd.find_bands_device(pats, out=out, work=work, chunk=chunk)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    value = N * args.steps / (t1 - t0)
    print(f"[bench] find_bands {ms:.3f} ms/step, {value:.1f} patterns/s", file=sys.stderr, flush=True)

    # ---- per-kernel durations with HIP events on the launch stream ----
    # one launch = one chunk, as inside find_bands
    s = torch.cuda.current_stream(dev)
    rad = torch.empty((chunk, bd.nRho, bd.nTheta), dtype=torch.float32, device=dev)
    x, flags = bd._prepare(plan, pats[:chunk])
    out_c = plan.alloc_out(chunk)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    plan.radon(x, flags=flags, out=rad)
    torch.cuda.synchronize()
    ev[0].record(s)
    for _ in range(args.roofline_reps):
        plan.radon(x, flags=flags, out=rad)
    ev[1].record(s)
    for _ in range(args.roofline_reps):
        plan.peaks(rad, out=out_c)
    ev[2].record(s)
    torch.cuda.synchronize()
    gather_ms = ev[0].elapsed_time(ev[1]) / args.roofline_reps
    peaks_ms = ev[1].elapsed_time(ev[2]) / args.roofline_reps
    pairs = int(info["pairs"])
    lds_bytes = 4.0 * pairs * chunk                   # algorithmic LDS bytes per gather launch
    lds_gbs = lds_bytes / (gather_ms * 1e-3) / 1e9
    print("[bench] LDS read ceiling (conflict-free ds_read, measured now)", file=sys.stderr, flush=True)
    ceiling = lds_ceiling()
    lds_peak = ceiling["ds_read_b128"]
    # the bytes the gather's ds_read_b64 actually move: every padded entry slot, one 8-byte
    # word per 4 uint16 patterns (8 uint8 patterns)
    pack = {"uint16": 4, "uint8": 8}.get(args.dtype)
    slots = int(info["slots"])      # the plan's own padded entry slots per pattern
    actual = {"bytes_per_launch": slots * 8.0 * chunk / pack if pack else None}
    if pack:
        actual["achieved"] = round(actual["bytes_per_launch"] / (gather_ms * 1e-3) / 1e9, 1)
        actual["peak"] = ceiling["ds_read_b64"]
        actual["frac"] = round(actual["achieved"] / actual["peak"], 4)
        actual["entry_slots_per_pattern"] = slots
    vc = valu_counts(H, W, args.dtype, chunk)
    print("[bench] VALU issue ceiling (independent chains, measured now)", file=sys.stderr, flush=True)
    valu_ceil = valu_ceiling()
    valu_peak = max(valu_ceil.values()) * 1e9

    # throughput variants of the headline workload: the exact background mode (the
    # reference's per-pixel float32 rounding, float64 sums), and the drop-in API from host
    # memory (numpy in, structured array out)
    exact = None
    if bg is not None:
        bdx = BandDetect(patDim=(H, W), background_mode="exact", **params)
        bdx.backgroundsub = bg
        outx = bdx.device_plan().alloc_out(N)
        workx = bdx.device_plan().workspace(N, chunk)
        for _ in range(3):
            bdx.find_bands_device(pats, out=outx, work=workx, chunk=chunk)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k = max(5, args.steps // 4)
        for _ in range(k):
            bdx.find_bands_device(pats, out=outx, work=workx, chunk=chunk)
        torch.cuda.synchronize()
        exact = {"patterns_per_s": round(N * k / (time.perf_counter() - t0), 1),
                 "note": "background_mode='exact': float32(x - b) per pixel, float64 line sums in plan "
                         "order -- the reference's arithmetic, Radon bit-identical to it"}
        del bdx, outx, workx
    print("[bench] drop-in API from host memory", file=sys.stderr, flush=True)
    host_pats = pats.cpu().numpy()
    bd.find_bands(host_pats)
    t0 = time.perf_counter()
    for _ in range(3):
        api_res = bd.find_bands(host_pats)
    api_rate = 3 * N / (time.perf_counter() - t0)
    api = {"patterns_per_s": round(api_rate, 1),
           "workload": f"BandDetect.find_bands on a numpy ({N}, {H}, {W}) {args.dtype} array -> structured "
                       f"array (the call EBSDIndexer makes): pinned-ring pipeline, host copies on "
                       f"{8} threads overlapped with uploads and band detection, records packed on "
                       f"the device into the returned array's pinned memory",
           "valid_band_fraction": round(float((api_res["valid"] == 1).mean()), 4)}
    del host_pats, api_res
    esz = {"uint8": 1, "uint16": 2, "float32": 4}[args.dtype]
    hbm_bytes = N * (H * W * esz + 9 * 28 + 12)       # input + band output per launch
    hbm_gbs = hbm_bytes / (ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_radon_gather_{H}x{W}_{args.dtype}.json")  # 2,048-pattern launches
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        if tj.get("patterns_per_launch") == chunk:
            traffic = tj.get("hbm_bytes_per_launch")

    # the streamed measurement runs before the batched one, so that it starts from an idle
    # device (an earlier version lost its copy overlap after the batched workload)
    streaming = None
    if not args.no_streaming:
        print("[bench] scan-scale streaming workload", file=sys.stderr, flush=True)
        streaming = streaming_measure(dev, total=args.stream_patterns)
        print(f"[bench] streaming {streaming['patterns_per_s']:.1f} patterns/s "
              f"(resident {streaming['resident_patterns_per_s']:.1f}, H2D "
              f"{streaming['roofline']['achieved']:.1f} of {streaming['roofline']['peak']:.1f} GB/s)",
              file=sys.stderr, flush=True)
        ff = streaming["from_file"]
        print(f"[bench] from a UP2 file {ff['patterns_per_s']:.1f} patterns/s (file reads alone "
              f"{ff['file_read_GBps']:.1f} GB/s)", file=sys.stderr, flush=True)

    batched = None
    if not args.no_batched:
        print("[bench] mixed-size batched workload", file=sys.stderr, flush=True)
        batched = batched_measure(dev, args.steps, args.warmup)
        print(f"[bench] batched {batched['graph_patterns_per_s']:.1f} patterns/s, plain per-size "
              f"{batched['plain_patterns_per_s']:.1f} patterns/s (chunks of 512: "
              f"{batched['plain_chunk512_patterns_per_s']:.1f})", file=sys.stderr, flush=True)
        sweep = []
        for n in (512, 2048, 8192):
            small = batched_measure(dev, args.steps, args.warmup, npats=n, autotune_sets=False)
            sweep.append({k: small[k] for k in ("plain_patterns_per_s", "graph_patterns_per_s", "graph_vs_plain",
                                                "plain_chunk512_patterns_per_s", "graph_vs_plain_chunk512",
                                                "plain_ms", "graph_ms", "peak_bins_identical")})
            sweep[-1]["patterns"] = n
            print(f"[bench] {n} mixed patterns: graph {small['graph_patterns_per_s']:.1f} vs plain "
                  f"{small['plain_patterns_per_s']:.1f} patterns/s", file=sys.stderr, flush=True)
        batched["smaller_batches"] = {"workload": "the same mixture at fewer patterns (launch defaults, no autotune)",
                                      "sweep": sweep}

    diff = None
    if not args.no_diff:
        print("[bench] differentiation workload", file=sys.stderr, flush=True)
        diff = differentiation_measure(dev)
        print(f"[bench] forward+backward {diff['forward_backward_patterns_per_s']:.1f} patterns/s, PC error "
              f"{diff['pc_error_start']:.4f} -> {diff['pc_error_refined']:.5f}", file=sys.stderr, flush=True)

    tripvote = None
    if not args.no_tripvote:
        print("[bench] triplet-vote workload", file=sys.stderr, flush=True)
        tripvote = tripvote_measure(dev, cpu_seconds=0.0 if args.no_cpu else 6.0)
        cpu_tv = tripvote["cpu_baseline"]["value"] if tripvote["cpu_baseline"] else float("nan")
        print(f"[bench] triplet vote {tripvote['patterns_per_s']:.1f} patterns/s ({tripvote['kernel_ms']} ms per "
              f"65,536), CPU {cpu_tv:.1f} patterns/s, family agreement "
              f"{tripvote['band_family_agreement']}", file=sys.stderr, flush=True)
        bi = tripvote["bandindex"]
        print(f"[bench] bandindex {bi['patterns_per_s']:.1f} patterns/s ({bi['ms']} ms per 65,536), CPU "
              f"{(bi['cpu_baseline'] or {}).get('value')} patterns/s, indexed {bi['indexed_fraction']}",
              file=sys.stderr, flush=True)

    nlpar = None
    if not args.no_nlpar:
        print("[bench] NLPAR workload", file=sys.stderr, flush=True)
        nlpar = nlpar_measure(dev, cpu_seconds=0.0 if args.no_cpu else 8.0)
        print(f"[bench] NLPAR {nlpar['patterns_per_s']:.1f} patterns/s ({nlpar['kernels_ms']}), "
              f"{nlpar['roofline']['achieved']} TFLOP/s; CPU {(nlpar['cpu_baseline'] or {}).get('value')} patterns/s",
              file=sys.stderr, flush=True)

    c1 = None
    if not args.no_c1:
        print("[bench] C1: 1,000 60x60 uint8 patterns", file=sys.stderr, flush=True)
        c1 = c1_measure(dev, cpu_seconds=0.0 if args.no_cpu else 6.0)
        print(f"[bench] C1 device {c1['device_patterns_per_s']:.1f}, API {c1['api_patterns_per_s']:.1f} patterns/s "
              f"(published OpenCL {PUBLISHED_C1})", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline (oracle restatement of the reference CPU path)", file=sys.stderr, flush=True)
        pps, done, dt = cpu_baseline(sample, bg, params, args.cpu_seconds)
        cpu = {"value": round(pps, 2), "unit": "patterns/s", "cores": os.cpu_count(),
               "kind": "port",
               "sample": f"{done} patterns ({sample.shape[0]} distinct, cycled) of the same {H}x{W} "
                         f"{args.dtype} workload in {dt:.1f} s "
                         f"(chunks of 512, FFT workers = os.cpu_count(), OMP_NUM_THREADS="
                         f"{os.environ.get('OMP_NUM_THREADS', 'unset')})"}

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "patterns/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u16" if esz == 2 else
        ("u8" if esz == 1 else "f32"), "data": "synthetic",
        "config": {"workload": f"plain (per-size, device-resident) find_bands, {N} x {H}x{W} {args.dtype} FCC "
                               f"patterns, {'background subtraction (fast mode), ' if bg is not None else ''}"
                               f"Radon 180x90, kernel 9x5, 9 bands; the metric's batched-vs-plain comparison "
                               f"is in 'batched'", "global_batch": N, "chunk": chunk, "parallelism": "none"},
        "roofline": {"bound": "lds", "kernel": "k_radon_gather", "achieved": round(lds_gbs, 1),
                     "peak": lds_peak, "unit": "GB/s", "frac": round(lds_gbs / lds_peak, 4),
                     "traffic": traffic, "algorithmic_bytes_per_launch": lds_bytes,
                     "avg_launch_ms": round(gather_ms, 4), "pairs_per_pattern": pairs,
                     "peak_kind": "conflict-free ds_read_b128 measured in this run (the LDS read ceiling)",
                     "lds_ceiling_measured_GBps": ceiling, "peak_nominal": LDS_NOMINAL_GBS,
                     "lds_actual": actual,
                     "valu": None if vc is None else {
                         "instructions_per_launch": vc["k_radon_gather"],
                         "achieved": round(vc["k_radon_gather"] / (gather_ms * 1e-3) / 1e9, 1),
                         "peak": round(valu_peak / 1e9, 1), "unit": "G wave-instructions/s",
                         "frac": round(vc["k_radon_gather"] / (gather_ms * 1e-3) / valu_peak, 4),
                         "peak_kind": "VALU issue ceiling measured in this run (the best of the "
                                      "probe's instruction mixes)",
                         "peak_nominal": VALU_ISSUE_PEAK / 1e9}},
        "roofline_peaks": None if vc is None else {
            "bound": "valu", "kernel": "k_peaks_stream", "unit": "G wave-instructions/s",
            "achieved": round(vc["k_peaks_stream"] / (peaks_ms * 1e-3) / 1e9, 1),
            "peak": round(valu_peak / 1e9, 1),
            "frac": round(vc["k_peaks_stream"] / (peaks_ms * 1e-3) / valu_peak, 4),
            "peak_kind": "VALU issue ceiling measured in this run (the best of the probe's "
                         "instruction mixes)",
            "valu_ceiling_measured": valu_ceil, "peak_nominal": VALU_ISSUE_PEAK / 1e9,
            "instructions_per_launch": vc["k_peaks_stream"], "avg_launch_ms": round(peaks_ms, 4),
            "counts_from": f"profiles/valu_counts_{H}x{W}_{args.dtype}.json (rocprofv3 SQ_INSTS_VALU)"},
        "exact_mode": exact,
        "api": api,
        "c1": c1,
        "roofline_hbm": {"bound": "hbm", "achieved": round(hbm_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(hbm_gbs / HBM_PEAK_GBS, 5), "algorithmic_bytes_per_step": hbm_bytes},
        "kernels_ms_per_chunk": {"k_radon_gather": round(gather_ms, 4), "k_peaks": round(peaks_ms, 4),
                                 "chunk": chunk},
        "cpu_baseline": cpu,
        "batched": batched,
        "streaming": streaming,
        "differentiation": diff,
        "nlpar": nlpar,
        "tripvote": tripvote,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
