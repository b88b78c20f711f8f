#!/usr/bin/env python3
"""Frames/s of the ProRes slice engine on one MI355X.

Workload (the decode case of the metric): 16 frames of 3840x2160 ProRes
4444 XQ, 12-bit 4:4:4 plus 16-bit alpha, 8-macroblock slices, default
matrix, the profile's target scaled per macroblock. Colour planes are
seeded sinusoids plus noise (sigma 24 codes at 12 bits, 256-3760); alpha is
soft-edged ellipses. The bitstreams are encoded on the GPU by this library,
which is byte-identical to the CPU oracle (checked below on one frame).

A step decodes all 16 packets, already resident in device memory, into
device-resident planes. "batched" is one launch sequence for the 16 frames,
"plain" one launch sequence per frame. The timed region is `--steps` steps
after `--warmup` untimed ones, each step ending with the stream synchronized.
The batched steps are timed twice: one call at a time, and (the value) with
two batches in flight, two host threads calling the same synchronous entry
point, each with its own output planes (the library runs them on its two
per-call streams, so one batch's planning and launches overlap the other's
kernels). The plain steps run one call at a time.

Prints one JSON line. `roofline` is for the dominant kernel, from its HIP
event time on the library's stream; `cpu_baseline` is the CPU oracle (a C++
restatement of the reference, with its slice-parallel thread pool) decoding
the same packets on this host's cores.
"""

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rivet-prores_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "frames/s on 1 MI355X, batched vs plain (bit-exact); serving p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(P, synth, n, width, height):
    frames = [synth.frame(width, height, P.ChromaFormat.Yuv444, 12, seed=1000 + i, alpha=True) for i in range(n)]
    cfg = P.Config(P.Profile.P4444Xq)
    bufs = []
    items = []
    for f in frames:
        dd = P.DeviceBuffer(f.data.nbytes + 64)
        dd.upload(f.data)
        da = P.DeviceBuffer(f.alpha.nbytes + 64)
        da.upload(f.alpha)
        bufs.append((dd, da))
        items.append((cfg, f, dd.ptr, da.ptr))
    t0 = time.perf_counter()
    res = P.encode_device_batch(items)
    enc_s = time.perf_counter() - t0
    packets, dev_packets = [], []
    for r in res:
        if isinstance(r, Exception):
            raise r
        ptr, n_bytes = r
        host = P.download_ptr(ptr, n_bytes).tobytes()
        dp = P.DeviceBuffer(n_bytes + 128)
        P.copy_device(dp.ptr, ptr, n_bytes)
        packets.append(host)
        dev_packets.append(dp)
    return frames, cfg, bufs, packets, dev_packets, enc_s


def alpha_stream_len(packet):
    """Bytes of alpha data in a progressive packet's slices (slice header
    sizes, src/decode.rs:301-338)."""
    b = memoryview(packet)
    hsize = int.from_bytes(b[8:10], "big")
    pos = 8 + hsize
    psize = int.from_bytes(b[pos + 1:pos + 5], "big")
    phs = b[pos] >> 3
    # slices: the table is between the picture header and the first slice
    total = 0
    t = pos + phs
    end = pos + psize
    sizes = []
    while True:
        # the table ends where the first slice begins: sum of sizes + table == picture bytes
        sizes.append(int.from_bytes(b[t:t + 2], "big"))
        t += 2
        if t - pos - phs + sum(sizes) >= psize - phs:
            break
    s = t
    for sz in sizes:
        hs = b[s] >> 3
        y, cb, cr = (int.from_bytes(b[s + 2 + 2 * k:s + 4 + 2 * k], "big") for k in range(3))
        total += sz - hs - y - cb - cr
        s += sz
    assert s == end
    return total


def profile_block(name, workload):
    """The machine-readable block of a committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/<name>) when it covers `workload`, or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        text = open(path).read()
        block = text[text.index("```json") + 7:]
        d = json.loads(block[:block.index("```")])
    except (OSError, ValueError):
        return None
    return d if d.get("workload") == workload else None


def profiled_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on the decode `workload`, or None."""
    d = profile_block("decode_c2_hbm_traffic.md", workload)
    return None if d is None else d.get("traffic_bytes", {}).get(kernel)


def host_cpu():
    """(model name, logical CPUs of the machine, CPUs this process may use)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return model, os.cpu_count(), allowed


def best_rate(fn, units, budget_s, max_runs):
    """Best-of-N rate (units per second) of fn(), N runs within budget_s (at
    least one), as the reference's bench keeps its best run
    (examples/prores_bench.rs:153-172)."""
    best, runs, t_start = None, 0, time.perf_counter()
    while runs == 0 or (time.perf_counter() - t_start < budget_s and runs < max_runs):
        t0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
        runs += 1
    return units / best, runs


def cpu_baseline(O, packets, c4, c1, budget_s):
    """The CPU oracle (the reference's algorithm restated in C++, with its
    slice-parallel pool: one task per macroblock row for decode, chunks of
    slices for encode) on this host's cores and on one thread, for the
    headline decode (C2), the 2160p HQ encode (C4) and the 1080p HQ round
    trip (C1). Samples are bounded so the whole baseline takes about
    budget_s seconds."""
    model, nproc, allowed = host_cpu()
    env = os.environ.get("OMP_NUM_THREADS")
    threads = min(allowed, int(env)) if env and env.isdigit() and int(env) > 0 else allowed
    share = budget_s / 6.0

    def dec(ps, t):
        return lambda: [O.decode(p, 0, t) for p in ps]

    def enc(frames, t):
        return lambda: [O.encode(of, O.config(3, luma=l, chroma=c, threads=t)) for of, l, c in frames]

    c2_all, runs = best_rate(dec(packets, threads), len(packets), share, 20)
    one = packets[:2]
    c2_one, runs1 = best_rate(dec(one, 1), len(one), share, 3)
    c4_all, runs4 = best_rate(enc(c4[:4], threads), len(c4[:4]), share, 10)
    c4_one, runs41 = best_rate(enc(c4[:1], 1), 1, share, 2)
    c1_frame, c1_cfg = c1

    def rt(t):
        cfg = O.config(3, threads=t)
        return lambda: O.decode(O.encode(c1_frame, cfg), 0, t)

    c1_all, _ = best_rate(rt(threads), 1, share / 2, 20)
    c1_one, _ = best_rate(rt(1), 1, share / 2, 5)
    return {"value": round(c2_all, 3), "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"decode of the {len(packets)} headline packets (16 x 3840x2160 4444 XQ + 16-bit alpha) "
                      f"by the C++ oracle's slice-parallel pool, {threads} threads, best of {runs} runs",
            "cpu_model": model, "nproc": nproc, "cpus_allowed": allowed,
            "one_thread": {"value": round(c2_one, 3), "unit": "frames/s", "cores": 1,
                           "sample": f"decode of {len(one)} headline packets, best of {runs1} runs"},
            "encode_2160p_hq": {"all_cores_fps": round(c4_all, 3), "one_thread_fps": round(c4_one, 3),
                                "cores": threads,
                                "sample": f"encode of {len(c4[:4])} (all cores, best of {runs4}) and 1 (one "
                                          f"thread, best of {runs41}) of the 32 x 3840x2160 422 HQ frames"},
            "roundtrip_1080p_hq": {"all_cores_fps": round(c1_all, 3), "one_thread_fps": round(c1_one, 3),
                                   "cores": threads,
                                   "sample": "encode then decode of the 1920x1080 422 HQ frame"}}


def fnv_packets(P, packets):
    """FNV-1a 64 over the packets in order (examples/prores_bench.rs:173-176)."""
    h = 0xCBF29CE484222325
    for p in packets:
        h = P.fnv1a64(h, p)
    return f"{h:016x}"


def fnv_pictures(P, pictures):
    """FNV-1a 64 over each picture's samples then its alpha, little-endian
    u16 (examples/prores_bench.rs:177-183)."""
    h = 0xCBF29CE484222325
    for data, alpha in pictures:
        h = P.fnv1a64(h, np.ascontiguousarray(data, dtype="<u2"))
        if alpha is not None:
            h = P.fnv1a64(h, np.ascontiguousarray(alpha, dtype="<u2"))
    return f"{h:016x}"


def calibration_copy(P, mb):
    """One device-to-device copy of `mb` MiB: a dispatch of known traffic,
    for scaling the profiler's byte counters."""
    if mb:
        nbytes = mb << 20
        ca, cb = P.DeviceBuffer(nbytes), P.DeviceBuffer(nbytes)
        P.copy_device(cb.ptr, ca.ptr, nbytes)
        log(f"bench: calibration copy of {nbytes} bytes")
        ca.free()
        cb.free()


def two_in_flight(dec, jobs_a, jobs_b, steps):
    """`steps` batches in all, alternate ones on each of two host threads
    (jobs_a and jobs_b write planes of their own); the library runs them on
    its two per-call streams. Returns the seconds they took together."""
    import threading
    failed = []

    def run(j, k):
        try:
            for _ in range(k):
                errs = dec.decode_device_batch(j)
                if any(errs):
                    raise RuntimeError(f"decode failed: {errs}")
        except Exception as e:  # reported after the join
            failed.append(e)
    ts = [threading.Thread(target=run, args=(jobs_a, (steps + 1) // 2)),
          threading.Thread(target=run, args=(jobs_b, steps // 2))]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    dt = time.perf_counter() - t0
    if failed:
        raise failed[0]
    return dt


def timed(fn, steps):
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    return (time.perf_counter() - t0) / steps


def leg_mixed_decode(P, synth, steps, warmup, check):
    """96 mixed frames in one batch: 32 x 720x486 interlaced 422 HQ (top
    field first, 2-MB slices), 32 x 1280x720 LT (4-MB), 24 x 1920x1080 Proxy
    (1-MB), 8 x 4096x2160 4444 with 8-bit alpha and loaded matrices (entries
    4-32); per-frame noise sigma uniform in 2-40 codes. The batch tuning is
    picked by the autotuner; "batched" keeps two batches in flight (one call
    at a time beside it), "plain" decodes the frames one call each."""
    rng = np.random.default_rng(7)
    kinds = ([(720, 486, 2, 10, P.Profile.Hq, 1, 1, 0, False)] * 32 +
             [(1280, 720, 2, 10, P.Profile.Lt, 2, 0, 0, False)] * 32 +
             [(1920, 1080, 2, 10, P.Profile.Proxy, 0, 0, 0, False)] * 24 +
             [(4096, 2160, 3, 12, P.Profile.P4444, 3, 0, 1, True)] * 8)
    frames, items, bufs = [], [], []
    for i, (w, h, ch, bd, prof, log2, il, atype, mats) in enumerate(kinds):
        f = synth.frame(w, h, ch, bd, seed=5000 + i, sigma=float(rng.uniform(2, 40)), alpha=atype > 0,
                        interlace=P.Interlace(il))
        luma = chroma = None
        if mats:
            luma = [int(v) for v in rng.integers(4, 33, 64)]
            chroma = [int(v) for v in rng.integers(4, 33, 64)]
        cfg = P.Config(prof, log2_slice_mbs=log2, alpha=P.AlphaType(atype) if atype else P.AlphaType.Bits16,
                       luma_matrix=luma, chroma_matrix=chroma)
        dd = P.DeviceBuffer(f.data.nbytes + 64)
        dd.upload(f.data)
        da = None
        if f.alpha is not None:
            da = P.DeviceBuffer(f.alpha.nbytes + 64)
            da.upload(f.alpha)
        frames.append((f, cfg))
        bufs.append((dd, da))
        items.append((cfg, f, dd.ptr, da.ptr if da else 0))
    res = P.encode_device_batch(items)
    packets, dev_packets, outs = [], [], []
    for r, (f, cfg) in zip(res, frames):
        if isinstance(r, Exception):
            raise r
        host = P.download_ptr(*r).tobytes()
        dp = P.DeviceBuffer(len(host) + 128)
        P.copy_device(dp.ptr, r[0], len(host))
        packets.append(host)
        dev_packets.append(dp)
        n = f.data.size
        outs.append((P.DeviceBuffer(n * 2 + 64), P.DeviceBuffer(f.width * f.height * 2 + 64) if f.alpha is not None
                     else None))
    del bufs
    n = len(packets)
    jobs = P.DecodeJobs([(packets[i], dev_packets[i].ptr, outs[i][0].ptr, outs[i][1].ptr if outs[i][1] else 0)
                         for i in range(n)])
    # The second batch in flight writes planes of its own.
    outs2 = [(P.DeviceBuffer(b.nbytes), P.DeviceBuffer(a.nbytes) if a is not None else None) for b, a in outs]
    jobs2 = P.DecodeJobs([(packets[i], dev_packets[i].ptr, outs2[i][0].ptr, outs2[i][1].ptr if outs2[i][1] else 0)
                          for i in range(n)])
    singles = [P.DecodeJobs([(packets[i], dev_packets[i].ptr, outs[i][0].ptr,
                              outs[i][1].ptr if outs[i][1] else 0)]) for i in range(n)]
    dec = P.Decoder()
    default = P.DecodeTuning()
    tuning, tuned_us = dec.autotune(jobs, reps=3)

    def batched():
        if any(dec.decode_device_batch(jobs)):
            raise RuntimeError("mixed decode failed")

    def plain():
        for j in singles:
            if dec.decode_device_batch(j, default)[0]:
                raise RuntimeError("mixed decode failed")

    for _ in range(warmup):
        batched()
        plain()
    two_in_flight(dec, jobs, jobs2, max(2, 2 * warmup))
    one_s = timed(batched, steps)
    # Batched: two batches in flight, as the headline decode is timed.
    b_s = two_in_flight(dec, jobs, jobs2, steps) / steps
    p_s = timed(plain, steps)
    out = {"workload": "decode 96 mixed frames: 32x 720x486i HQ, 32x 1280x720 LT, 24x 1920x1080 Proxy, "
                       "8x 4096x2160 4444 + 8-bit alpha, loaded matrices",
           "frames": n, "slices": None, "batched_fps": round(n / b_s, 2), "batches_in_flight": 2,
           "batched_one_in_flight_fps": round(n / one_s, 2), "plain_fps": round(n / p_s, 2),
           "batched_over_plain": round(p_s / b_s, 3), "ms_per_batch": round(b_s * 1e3, 3),
           "tuning": dataclasses.asdict(tuning), "autotune_median_us": round(tuned_us, 1),
           "bytes_per_batch": sum(len(p) for p in packets)}
    if check:
        import oracle_lib as O
        ok = True
        for i in (0, 32, 64, 88):
            f, cfg = frames[i]
            oc = O.config(int(cfg.profile), log2_slice_mbs=cfg.log2_slice_mbs, alpha=int(cfg.alpha),
                          luma=cfg.luma_matrix, chroma=cfg.chroma_matrix)
            of = O.Frame(f.width, f.height, int(f.chroma), f.bit_depth, f.data, f.alpha, int(f.interlace))
            ok &= O.encode(of, oc) == packets[i]
            od = O.decode(packets[i])
            ok &= bool(np.array_equal(od.data, outs[i][0].download(f.data.size * 2, dtype=np.uint16)))
            if f.alpha is not None:
                ok &= bool(np.array_equal(od.alpha, outs[i][1].download(f.width * f.height * 2, dtype=np.uint16)))
        out["bit_exact_sample"] = {"frames": [0, 32, 64, 88], "encode_and_decode_equal_oracle": bool(ok)}
    return out


def leg_batch_encode(P, synth, steps, warmup, check, keep_for_cpu):
    """32 x 3840x2160 422 HQ from 10-bit input, 8-MB slices; frames 16-31
    use seeded loaded matrices (entries 4-32); every fourth frame adds
    hard-edged rectangles. "batched" is one encode call for the 32 frames,
    "plain" one call per frame. Every packet of the last timed batch is
    compared with the oracle's encode of the same frame and Config.
    `roofline` is the whole launch sequence's: algorithmic bytes (planes read,
    packets written) over the batch's kernel time; `stage_us` breaks that
    time down kernel by kernel (HIP events on the library's stream).
    `traffic` is the batch's HBM bytes from the committed FETCH_SIZE /
    WRITE_SIZE summary (profiles/encode_c4_hbm_traffic.md); `dominant` is the
    longest stage, its kernels, their traffic and the rate they move it at
    (the size passes read the engine's coefficient intermediate: they have no
    algorithmic bytes of their own)."""
    rng = np.random.default_rng(11)
    frames, items, bufs = [], [], []
    for i in range(32):
        f = synth.frame(3840, 2160, P.ChromaFormat.Yuv422, 10, seed=9000 + i, rectangles=(i % 4 == 0))
        luma = chroma = None
        if i >= 16:
            luma = [int(v) for v in rng.integers(4, 33, 64)]
            chroma = [int(v) for v in rng.integers(4, 33, 64)]
        cfg = P.Config(P.Profile.Hq, luma_matrix=luma, chroma_matrix=chroma)
        dd = P.DeviceBuffer(f.data.nbytes + 64)
        dd.upload(f.data)
        frames.append((f, cfg))
        bufs.append(dd)
        items.append((cfg, f, dd.ptr, 0))
    last = {}

    def batched():
        res = P.encode_device_batch(items)
        for r in res:
            if isinstance(r, Exception):
                raise r
        last["res"] = res

    def plain():
        for it in items:
            r = P.encode_device_batch([it])[0]
            if isinstance(r, Exception):
                raise r

    for _ in range(warmup):
        batched()
        plain()
    stages = {k: 0.0 for k in P.ENCODE_STAGES}
    t0 = time.perf_counter()
    for _ in range(steps):
        batched()
        for k, v in P.last_encode_stages().items():
            stages[k] += v / steps
    b_s = (time.perf_counter() - t0) / steps
    sizes = [r[1] for r in last["res"]]
    # The output arena is reused by the next encode call: download now.
    got = [P.download_ptr(*r).tobytes() for r in last["res"]]
    target = P.Profile.Hq.target_frame_bytes(3840, 2160)
    p_s = timed(plain, steps)
    in_bytes = sum(f.data.nbytes for f, _ in frames)
    out_bytes = sum(sizes)
    kernels_us = sum(stages[k] for k in P.ENCODE_STAGES[:-1])
    dom = max(P.ENCODE_STAGES[:-1], key=lambda k: stages[k])
    gbs = (in_bytes + out_bytes) / (kernels_us * 1e-6) / 1e9
    workload = ("encode 32 x 3840x2160 422 HQ 10-bit, 8-MB slices, half loaded matrices, "
                "a quarter with hard-edged rectangles")
    # HBM bytes a batch, whole sequence and dominant stage, from the committed
    # FETCH_SIZE / WRITE_SIZE summary of this workload.
    prof = profile_block("encode_c4_hbm_traffic.md", workload)
    traffic = dom_traffic = None
    if prof:
        tb, per = prof["traffic_bytes"], prof["launches_per_batch"]
        traffic = int(sum(tb[k] * per[k] for k in tb))
        dom_traffic = int(sum(tb[k] * per[k] for k in prof["stages"].get(dom, []) if k in tb))
    dominant = {"stage": dom, "us": round(stages[dom], 1),
                "kernels": prof["stages"].get(dom) if prof else None, "traffic": dom_traffic,
                "traffic_gbs": round(dom_traffic / (stages[dom] * 1e-6) / 1e9, 1) if dom_traffic else None,
                "algorithmic_bytes": {"fdct": in_bytes, "pack": out_bytes}.get(dom, 0)}
    out = {"workload": workload,
           "frames": 32, "batched_fps": round(32 / b_s, 2), "plain_fps": round(32 / p_s, 2),
           "batched_over_plain": round(p_s / b_s, 3), "ms_per_batch": round(b_s * 1e3, 3),
           "stage_us": {k: round(v, 1) for k, v in stages.items()},
           "rate_control_us": round(stages["probes"] + stages["window"] + stages["shareout"] + stages["offsets"], 1),
           "kernels_us": round(kernels_us, 1),
           "roofline": {"bound": "hbm", "kernel": "the whole launch sequence",
                        "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                        "algorithmic_bytes": int(in_bytes + out_bytes), "dominant": dominant},
           "target_frame_bytes": target, "max_frame_bytes": max(sizes), "min_frame_bytes": min(sizes),
           "all_at_or_under_target": all(s <= target for s in sizes),
           "packets_fnv1a": fnv_packets(P, got)}
    if check:
        import oracle_lib as O
        want, oracle_in = [], []
        for f, cfg in frames:
            of = O.Frame(f.width, f.height, int(f.chroma), f.bit_depth, f.data, None, 0)
            oracle_in.append((of, cfg.luma_matrix, cfg.chroma_matrix))
            want.append(O.encode(of, O.config(3, luma=cfg.luma_matrix, chroma=cfg.chroma_matrix)))
        same = [g == w for g, w in zip(got, want)]
        out["bit_exact"] = {"frames": 32, "bytes_equal_oracle": all(same),
                            "frames_differing": [i for i, x in enumerate(same) if not x]}
        out["oracle_packets_fnv1a"] = fnv_packets(P, want)
        keep_for_cpu.extend(oracle_in)
    return out


def leg_roundtrip_1080p(P, synth, steps, warmup, check):
    """The plain round trip of one 1920x1080 progressive 422 HQ frame from
    10-bit input, 8-MB slices, default matrix (target 917,583 bytes): encode
    then decode, device-resident, one call each. Both must reproduce the
    oracle's bytes and samples."""
    f = synth.frame(1920, 1080, P.ChromaFormat.Yuv422, 10, seed=1)
    cfg = P.Config(P.Profile.Hq)
    dd = P.DeviceBuffer(f.data.nbytes + 64)
    dd.upload(f.data)
    dp = P.DeviceBuffer(4 << 20)
    do = P.DeviceBuffer(f.data.nbytes + 64)
    dec = P.Decoder()
    last = {}

    def step():
        r = P.encode_device_batch([(cfg, f, dd.ptr, 0)])[0]
        if isinstance(r, Exception):
            raise r
        P.copy_device(dp.ptr, r[0], r[1])
        last["n"] = r[1]
        jobs = P.DecodeJobs([(r[1], dp.ptr, do.ptr, 0)])
        if dec.decode_device_batch(jobs)[0] is not None:
            raise RuntimeError("1080p round trip decode failed")

    for _ in range(warmup):
        step()
    rt_s = timed(step, steps)
    packet = dp.download(last["n"]).tobytes()
    picture = do.download(f.data.nbytes, dtype=np.uint16)
    out = {"workload": "round trip (encode, then decode) of one 1920x1080 422 HQ 10-bit frame, 8-MB slices",
           "roundtrip_fps": round(1 / rt_s, 2), "ms_per_roundtrip": round(rt_s * 1e3, 3),
           "packet_bytes": len(packet), "target_frame_bytes": P.Profile.Hq.target_frame_bytes(1920, 1080),
           "packets_fnv1a": fnv_packets(P, [packet]), "pictures_fnv1a": fnv_pictures(P, [(picture, None)])}
    c1 = None
    if check:
        import oracle_lib as O
        of = O.Frame(1920, 1080, 2, 10, f.data, None, 0)
        want = O.encode(of, O.config(3))
        od = O.decode(want)
        out["bit_exact"] = {"bytes_equal_oracle": want == packet,
                            "samples_equal_oracle": bool(np.array_equal(od.data, picture))}
        out["oracle_packets_fnv1a"] = fnv_packets(P, [want])
        out["oracle_pictures_fnv1a"] = fnv_pictures(P, [(od.data, None)])
        c1 = (of, None)
    return out, c1


SERVE_SIZES = [(720, 486, 1), (1280, 720, 0), (1920, 1080, 0), (2048, 1080, 0), (3840, 2160, 0), (4096, 2160, 0)]


def leg_serving(P, synth, seconds, rate, check):
    """The serving loop under Poisson arrivals at `rate` requests/s for
    `seconds`: 70% decode, 30% encode; sizes from SERVE_SIZES, profiles from
    all six (alpha on half of the 4444 requests), 1-8 macroblock slices,
    seeded sinusoid+noise content. Requests draw from a pool of 24 decode
    packets and 12 encode frames made up front. A poller thread records each
    result's latency (submit to done, on the server's clock) and releases it;
    a 5% sample is checked against the oracle."""
    import threading
    rng = np.random.default_rng(23)

    def spec():
        w, h, il = SERVE_SIZES[int(rng.integers(0, len(SERVE_SIZES)))]
        prof = P.Profile(int(rng.integers(0, 6)))
        ch = P.ChromaFormat.Yuv444 if prof >= P.Profile.P4444 else P.ChromaFormat.Yuv422
        bd = 12 if ch == P.ChromaFormat.Yuv444 else 10
        alpha = ch == P.ChromaFormat.Yuv444 and rng.random() < 0.5
        atype = P.AlphaType(int(rng.integers(1, 3))) if alpha else P.AlphaType.Bits16
        log2 = int(rng.integers(0, 4))
        f = synth.frame(w, h, ch, bd, seed=int(rng.integers(1 << 30)), sigma=float(rng.uniform(2, 40)),
                        alpha=alpha, interlace=P.Interlace(il))
        return f, P.Config(prof, log2_slice_mbs=log2, alpha=atype)

    log("bench: serving pool")
    dec_pool = [spec() for _ in range(24)]
    enc_pool = [spec() for _ in range(12)]
    packets = [P.Encoder(cfg).encode(f) for f, cfg in dec_pool]
    n_req = max(1, int(rate * seconds))
    kinds = rng.random(n_req) < 0.7
    picks = [int(rng.integers(0, 24 if k else 12)) for k in kinds]
    gaps = rng.exponential(1.0 / rate, n_req)
    sample = set(int(i) for i in np.flatnonzero(rng.random(n_req) < 0.05))

    srv = P.Server()
    results, sampled, lock = {}, {}, threading.Lock()
    id_to_req = {}
    stop = threading.Event()

    def poller():
        while not stop.is_set() or len(results) < n_req:
            for r in srv.poll(256, 50):
                with lock:
                    k = id_to_req.get(r.id)
                while k is None:  # submitted but not yet recorded by the submitter
                    time.sleep(0)
                    with lock:
                        k = id_to_req.get(r.id)
                err = r.error
                if k in sample and err is None:
                    try:
                        sampled[k] = r.samples().copy() if r.kind == "decode" else r.packet()
                        if r.kind == "decode" and r.dev_alpha:
                            sampled[k] = (sampled[k], r.alpha().copy())
                    except P.DeviceError as e:  # counted as a failed request; the leg goes on
                        log(f"bench: serving: reading result {r.id} ({r.kind}): {e}")
                        sampled.pop(k, None)
                        err = e
                results[k] = (r.kind, r.latency_us, r.batch_frames, err)
                srv.release(r.id)
            if stop.is_set() and len(results) >= n_req:
                break

    th = threading.Thread(target=poller, daemon=True)
    th.start()
    log(f"bench: serving {n_req} requests at {rate}/s")
    t0 = time.perf_counter()
    t_next = t0
    for i in range(n_req):
        t_next += gaps[i]
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        if kinds[i]:
            rid = srv.submit_decode(packets[picks[i]])
        else:
            f, cfg = enc_pool[picks[i]]
            rid = srv.submit_encode(cfg, f)
        with lock:
            id_to_req[rid] = i
    stop.set()
    th.join(timeout=120)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array([v[1] for v in results.values()])
    lat_d = np.array([v[1] for v in results.values() if v[0] == "decode"])
    lat_e = np.array([v[1] for v in results.values() if v[0] == "encode"])
    errors = sum(1 for v in results.values() if v[3] is not None)
    pct = lambda a, q: round(float(np.percentile(a, q)) / 1e3, 3) if a.size else None
    out = {"workload": "serving: Poisson arrivals, 70% decode / 30% encode, sizes 720x486i-4096x2160, "
                       "all six profiles (alpha on half of 4444), 1-8 MB slices",
           "duration_s": round(wall, 2), "offered_rps": rate, "requests": n_req, "completed": len(results),
           "errors": errors, "frames_per_s": round(len(results) / wall, 2),
           "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": round(float(lat.max()) / 1e3, 3) if lat.size else None,
                          "decode_p50": pct(lat_d, 50), "decode_p99": pct(lat_d, 99),
                          "encode_p50": pct(lat_e, 50), "encode_p99": pct(lat_e, 99)},
           "mean_batch_frames": round(float(np.mean([v[2] for v in results.values()])), 2) if results else None,
           "decode_batches": st["decode_batches"], "encode_batches": st["encode_batches"],
           "peak_held_bytes": st["peak_held_bytes"], "hbm_budget": st["hbm_budget"]}
    if check:
        import oracle_lib as O
        ok, n, want = True, 0, {}
        for k, got in sampled.items():
            key = (bool(kinds[k]), picks[k])
            if key not in want:  # the oracle runs once per pool item
                if kinds[k]:
                    want[key] = O.decode(packets[picks[k]])
                else:
                    f, cfg = enc_pool[picks[k]]
                    oc = O.config(int(cfg.profile), log2_slice_mbs=cfg.log2_slice_mbs, alpha=int(cfg.alpha))
                    of = O.Frame(f.width, f.height, int(f.chroma), f.bit_depth, f.data, f.alpha, int(f.interlace))
                    want[key] = O.encode(of, oc)
            w = want[key]
            if kinds[k]:
                if isinstance(got, tuple):
                    ok &= bool(np.array_equal(got[0][:w.data.size], w.data) and
                               np.array_equal(got[1][:w.alpha.size], w.alpha))
                else:
                    ok &= bool(np.array_equal(got[:w.data.size], w.data))
            else:
                ok &= w == got
            n += 1
        out["bit_exact_sample"] = {"requests": n, "equal_oracle": bool(ok)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--frames", type=int, default=16)
    ap.add_argument("--width", type=int, default=3840)
    ap.add_argument("--height", type=int, default=2160)
    ap.add_argument("--cpu-seconds", type=float, default=18.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="roundtrip,mixed,encode,serve",
                    help="extra legs after the headline decode: roundtrip (1080p HQ encode + decode), mixed "
                         "(96 mixed frames, batched vs plain decode), encode (32-frame batched encode), serve "
                         "(the serving loop), comma-separated, or none")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=300.0, help="offered requests per second")
    ap.add_argument("--serve-peak-seconds", type=float, default=15.0)
    ap.add_argument("--serve-peak-rate", type=float, default=1000.0,
                    help="a second serving run at this offered load (0: none)")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="after the timed steps, one device-to-device copy of this many MiB (a kernel of "
                         "known traffic, for scaling profiler byte counters)")
    ap.add_argument("--in-flight", type=int, default=2, choices=(1, 2),
                    help="headline batches in flight: 2 (the value; two host threads) or 1 (one call at a time "
                         "only: for profiling, so every launch runs without another batch beside it)")
    ap.add_argument("--legs-only", action="store_true",
                    help="run only the --legs (no headline decode, no CPU baseline): for profiling one leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import prores_amd as P
    from prores_amd import synth

    if args.legs_only:
        legs = [x for x in args.legs.split(",") if x and x != "none"]
        check = not args.no_cpu
        out = {}
        if "roundtrip" in legs:
            out["roundtrip_1080p"] = leg_roundtrip_1080p(P, synth, args.steps, args.warmup, check)[0]
        if "mixed" in legs:
            out["mixed_decode"] = leg_mixed_decode(P, synth, args.steps, args.warmup, check)
        if "encode" in legs:
            out["batch_encode"] = leg_batch_encode(P, synth, args.steps, args.warmup, check, [])
        if "serve" in legs:
            out["serving"] = leg_serving(P, synth, args.serve_seconds, args.serve_rate, check)
        calibration_copy(P, args.calibration_copy_mb)
        print(json.dumps(out), flush=True)
        return

    log(f"bench: generating and GPU-encoding {args.frames} frames of {args.width}x{args.height} 4444 XQ")
    frames, cfg, in_bufs, packets, dev_packets, enc_s = make_inputs(P, synth, args.frames, args.width, args.height)
    enc_stage = P.last_encode_timing()
    w, h = args.width, args.height
    n = len(frames)
    samples = frames[0].data.size
    outs = [(P.DeviceBuffer(samples * 2 + 64), P.DeviceBuffer(w * h * 2 + 64)) for _ in range(n)]
    jobs = P.DecodeJobs([(packets[i], dev_packets[i].ptr, outs[i][0].ptr, outs[i][1].ptr) for i in range(n)])
    # The second batch in flight writes planes of its own.
    outs2 = [(P.DeviceBuffer(samples * 2 + 64), P.DeviceBuffer(w * h * 2 + 64)) for _ in range(n)]
    jobs2 = P.DecodeJobs([(packets[i], dev_packets[i].ptr, outs2[i][0].ptr, outs2[i][1].ptr) for i in range(n)])
    singles = [P.DecodeJobs([(packets[i], dev_packets[i].ptr, outs[i][0].ptr, outs[i][1].ptr)]) for i in range(n)]
    dec = P.Decoder()

    def batched(j=jobs):
        errs = dec.decode_device_batch(j)
        if any(errs):
            raise RuntimeError(f"decode failed: {errs}")

    def plain():
        for j in singles:
            errs = dec.decode_device_batch(j)
            if errs[0]:
                raise RuntimeError(f"decode failed: {errs}")

    log("bench: warmup")
    for _ in range(args.warmup):
        batched()
        plain()
    if args.in_flight == 2:
        two_in_flight(dec, jobs, jobs2, max(2, 2 * args.warmup))  # the second stream's workspace is made here
    # One batch at a time: the stage times (HIP events, no other batch on
    # the GPU) and the one-in-flight rate.
    stage = np.zeros(5)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batched()
        stage += np.array(P.last_batch_timing())
    one_s = (time.perf_counter() - t0) / args.steps
    stage /= args.steps
    log(f"bench: batched, one in flight {one_s * 1e3:.3f} ms/step")
    batched_s = one_s
    if args.in_flight == 2:
        batched_s = two_in_flight(dec, jobs, jobs2, args.steps) / args.steps
        log(f"bench: batched, two in flight {batched_s * 1e3:.3f} ms/step")
    # Every frame of the last timed batches, before the plain loop reuses
    # the output buffers; the two sets of planes must be equal.
    batched_out = [(outs[i][0].download(samples * 2, dtype=np.uint16),
                    outs[i][1].download(w * h * 2, dtype=np.uint16)) for i in range(n)]
    for i in range(n if args.in_flight == 2 else 0):
        if not (np.array_equal(batched_out[i][0], outs2[i][0].download(samples * 2, dtype=np.uint16)) and
                np.array_equal(batched_out[i][1], outs2[i][1].download(w * h * 2, dtype=np.uint16))):
            raise RuntimeError(f"frame {i}: the two batches in flight decoded different planes")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plain()
    plain_s = (time.perf_counter() - t0) / args.steps
    log(f"bench: plain {plain_s * 1e3:.2f} ms/step")

    calibration_copy(P, args.calibration_copy_mb)

    # Algorithmic bytes (the packet read, the planes written).
    packet_bytes = sum(len(p) for p in packets)
    colour_bytes = n * samples * 2
    alpha_bytes = n * w * h * 2
    alpha_stream_bytes = sum(alpha_stream_len(p) for p in packets)
    # Per kernel launch: the slice tables read; the colour streams read; the
    # alpha streams read and the alpha plane written; the colour planes written.
    names = ["pd_slice_table", "pd_entropy", "pd_alpha", "pd_idct"]
    kernel_bytes = [2 * n * 4050, packet_bytes - alpha_stream_bytes, alpha_stream_bytes + alpha_bytes, colour_bytes]
    # The dominant kernel is the longest on the main stream (slice table,
    # colour entropy, IDCT); alpha runs beside them on a second stream and its
    # stage is the span from the end of the slice table to its own end.
    dom = max((0, 1, 3), key=lambda i: stage[i])
    achieved = kernel_bytes[dom] / (stage[dom] * 1e-6) / 1e9
    # The whole job: bytes moved a frame at the measured frame rate.
    pipeline_gbs = (packet_bytes + colour_bytes + alpha_bytes) / batched_s / 1e9

    workload = f"decode {n} x {w}x{h} ProRes 4444 XQ, 12-bit 4:4:4 + 16-bit alpha, 8-MB slices"
    result = {
        "metric": METRIC,
        "value": round(n / batched_s, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batched_s * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload,
                   "frames": n, "profile": "ap4x", "bit_depth": 12, "alpha_bits": 16, "log2_slice_mbs": 3},
        "batches_in_flight": args.in_flight,
        "batched_fps": round(n / batched_s, 3),
        "batched_one_in_flight_fps": round(n / one_s, 3),
        "ms_per_step_one_in_flight": round(one_s * 1e3, 4),
        "plain_fps": round(n / plain_s, 3),
        "batched_over_plain": round(plain_s / batched_s, 3),
        "stage_us": {"slice_table": round(stage[0], 2), "entropy": round(stage[1], 2), "alpha": round(stage[2], 2),
                     "idct": round(stage[3], 2), "kernels_total": round(stage[4], 2)},
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(workload, names[dom]),
                     "algorithmic_bytes": int(kernel_bytes[dom])},
        "pipeline_roofline": {"achieved": round(pipeline_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(pipeline_gbs / HBM_PEAK_GBS, 4),
                              "bytes_per_frame": round((packet_bytes + colour_bytes + alpha_bytes) / n)},
        "bytes_per_packet": round(packet_bytes / n),
        "encode_s_for_inputs": round(enc_s, 3),
        "encode_stage_us": {"fdct_alpha": round(enc_stage[0], 1), "rate_control": round(enc_stage[1], 1),
                            "pack": round(enc_stage[2], 1), "all": round(enc_stage[3], 1)},
    }

    # The host-buffer entry points, PCIe transfers included: prores_decode
    # (packet from host memory, planes back to host memory, one frame a
    # call) and prores_decode_host_batch (all 16 packets in one call, chunks
    # overlapped on two streams). Frames land in page-locked memory the
    # library keeps for reuse; a warm call first.
    frames_out = [dec.decode(packets[0])]
    del frames_out
    t0 = time.perf_counter()
    for i in range(4):
        dec.decode(packets[i])
    result["host_api_fps"] = round(4 / (time.perf_counter() - t0), 2)
    hb = dec.decode_batch(packets)
    if any(isinstance(x, Exception) for x in hb):
        raise RuntimeError(f"host batch decode failed: {hb}")
    del hb
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        hb = dec.decode_batch(packets)
        del hb
    host_batch_s = (time.perf_counter() - t0) / reps
    result["host_batch_fps"] = round(n / host_batch_s, 2)
    result["host_batch_gbs"] = round((colour_bytes + alpha_bytes + packet_bytes) / host_batch_s / 1e9, 2)

    result["hashes"] = {"packets_fnv1a": fnv_packets(P, packets), "pictures_fnv1a": fnv_pictures(P, batched_out)}

    legs = [x for x in args.legs.split(",") if x and x != "none"]
    check = not args.no_cpu
    c1 = None
    c4_frames = []
    if "roundtrip" in legs:
        log("bench: 1080p round trip leg")
        result["roundtrip_1080p"], c1 = leg_roundtrip_1080p(P, synth, args.steps, args.warmup, check)
    if "mixed" in legs:
        log("bench: mixed decode leg")
        result["mixed_decode"] = leg_mixed_decode(P, synth, args.steps, args.warmup, check)
    if "encode" in legs:
        log("bench: batch encode leg")
        result["batch_encode"] = leg_batch_encode(P, synth, max(1, args.steps // 2), 1, check, c4_frames)
    if "serve" in legs:
        log(f"bench: serving leg, {args.serve_seconds} s at {args.serve_rate}/s")
        result["serving"] = leg_serving(P, synth, args.serve_seconds, args.serve_rate, check)
        if args.serve_peak_rate > 0:
            log(f"bench: serving leg, {args.serve_peak_seconds} s at {args.serve_peak_rate}/s")
            result["serving_peak"] = leg_serving(P, synth, args.serve_peak_seconds, args.serve_peak_rate, check)

    if check:
        import oracle_lib as O
        # The checker: every headline frame's GPU packet against the
        # oracle's encode, and every frame of the timed batch against the
        # oracle's decode of its packet.
        log("bench: checking the headline frames against the oracle")
        oc = O.config(5)
        want_packets, want_pictures, same_enc, same_dec = [], [], [], []
        for i in range(n):
            of = O.Frame(w, h, 3, 12, frames[i].data, frames[i].alpha)
            wp = O.encode(of, oc)
            want_packets.append(wp)
            same_enc.append(wp == packets[i])
            od = O.decode(packets[i])
            want_pictures.append((od.data, od.alpha))
            same_dec.append(bool(np.array_equal(od.data, batched_out[i][0]) and
                                 np.array_equal(od.alpha, batched_out[i][1])))
        result["bit_exact_encode"] = all(same_enc)
        result["bit_exact_decode"] = all(same_dec)
        result["bit_exact_frames"] = {"checked": n, "packets_differing": [i for i, x in enumerate(same_enc) if not x],
                                      "pictures_differing": [i for i, x in enumerate(same_dec) if not x]}
        result["hashes"]["oracle_packets_fnv1a"] = fnv_packets(P, want_packets)
        result["hashes"]["oracle_pictures_fnv1a"] = fnv_pictures(P, want_pictures)
        del want_pictures
        log("bench: cpu baseline")
        if not c4_frames:
            f4 = synth.frame(3840, 2160, P.ChromaFormat.Yuv422, 10, seed=9000, rectangles=True)
            c4_frames = [(O.Frame(3840, 2160, 2, 10, f4.data, None, 0), None, None)]
        if c1 is None:
            f1 = synth.frame(1920, 1080, P.ChromaFormat.Yuv422, 10, seed=1)
            c1 = (O.Frame(1920, 1080, 2, 10, f1.data, None, 0), None)
        result["cpu_baseline"] = cpu_baseline(O, packets, c4_frames, c1, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
