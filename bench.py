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

Prints one JSON line. `roofline` is for the dominant kernel, from its HIP
event time on the library's stream; `cpu_baseline` is the CPU oracle (a C++
restatement of the reference, with its slice-parallel thread pool) decoding
the same packets on this host's cores.
"""

import argparse
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


def cpu_baseline(O, packets, budget_s):
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    best = None
    runs = 0
    t_start = time.perf_counter()
    while runs == 0 or (time.perf_counter() - t_start < budget_s and runs < 20):
        t0 = time.perf_counter()
        for p in packets:
            O.decode(p, 0, threads)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
        runs += 1
    return {"value": round(len(packets) / best, 3), "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"the {len(packets)} benchmark packets decoded by the C++ oracle's slice-parallel pool "
                      f"(one task per macroblock row), best of {runs} runs"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=16)
    ap.add_argument("--width", type=int, default=3840)
    ap.add_argument("--height", type=int, default=2160)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="after the timed steps, one device-to-device copy of this many MiB (a kernel of "
                         "known traffic, for scaling profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import prores_amd as P
    from prores_amd import synth

    log(f"bench: generating and GPU-encoding {args.frames} frames of {args.width}x{args.height} 4444 XQ")
    frames, cfg, in_bufs, packets, dev_packets, enc_s = make_inputs(P, synth, args.frames, args.width, args.height)
    enc_stage = P.last_encode_timing()
    w, h = args.width, args.height
    n = len(frames)
    samples = frames[0].data.size
    outs = [(P.DeviceBuffer(samples * 2 + 64), P.DeviceBuffer(w * h * 2 + 64)) for _ in range(n)]
    jobs = P.DecodeJobs([(packets[i], dev_packets[i].ptr, outs[i][0].ptr, outs[i][1].ptr) for i in range(n)])
    singles = [P.DecodeJobs([(packets[i], dev_packets[i].ptr, outs[i][0].ptr, outs[i][1].ptr)]) for i in range(n)]
    dec = P.Decoder()

    def batched():
        errs = dec.decode_device_batch(jobs)
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
    stage = np.zeros(5)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batched()
        stage += np.array(P.last_batch_timing())
    batched_s = (time.perf_counter() - t0) / args.steps
    stage /= args.steps
    log(f"bench: batched {batched_s * 1e3:.2f} ms/step")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plain()
    plain_s = (time.perf_counter() - t0) / args.steps
    log(f"bench: plain {plain_s * 1e3:.2f} ms/step")

    if args.calibration_copy_mb:
        nbytes = args.calibration_copy_mb << 20
        ca, cb = P.DeviceBuffer(nbytes), P.DeviceBuffer(nbytes)
        P.copy_device(cb.ptr, ca.ptr, nbytes)
        log(f"bench: calibration copy of {nbytes} bytes")
        ca.free()
        cb.free()

    # Algorithmic bytes (the packet read, the planes written).
    packet_bytes = sum(len(p) for p in packets)
    colour_bytes = n * samples * 2
    alpha_bytes = n * w * h * 2
    alpha_stream_bytes = sum(alpha_stream_len(p) for p in packets)
    # Per kernel launch: the slice tables read; the colour streams read; the
    # alpha streams read and the alpha plane written; the colour planes written.
    names = ["pd_slice_table", "pd_entropy", "pd_alpha", "pd_idct"]
    kernel_bytes = [2 * n * 4050, packet_bytes - alpha_stream_bytes, alpha_stream_bytes + alpha_bytes, colour_bytes]
    dom = int(np.argmax(stage[:4]))
    achieved = kernel_bytes[dom] / (stage[dom] * 1e-6) / 1e9
    pipeline_gbs = (packet_bytes + colour_bytes + alpha_bytes) / (stage[4] * 1e-6) / 1e9

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
        "config": {"workload": f"decode {n} x {w}x{h} ProRes 4444 XQ, 12-bit 4:4:4 + 16-bit alpha, 8-MB slices",
                   "frames": n, "profile": "ap4x", "bit_depth": 12, "alpha_bits": 16, "log2_slice_mbs": 3},
        "batched_fps": round(n / batched_s, 3),
        "plain_fps": round(n / plain_s, 3),
        "batched_over_plain": round(plain_s / batched_s, 3),
        "stage_us": {"slice_table": round(stage[0], 2), "entropy": round(stage[1], 2), "alpha": round(stage[2], 2),
                     "idct": round(stage[3], 2), "kernels_total": round(stage[4], 2)},
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "pipeline_roofline": {"achieved": round(pipeline_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(pipeline_gbs / HBM_PEAK_GBS, 4),
                              "bytes_per_frame": round((packet_bytes + colour_bytes + alpha_bytes) / n)},
        "bytes_per_packet": round(packet_bytes / n),
        "encode_s_for_inputs": round(enc_s, 3),
        "encode_stage_us": {"fdct_alpha": round(enc_stage[0], 1), "rate_control": round(enc_stage[1], 1),
                            "pack": round(enc_stage[2], 1), "all": round(enc_stage[3], 1)},
    }

    if not args.no_cpu:
        import oracle_lib as O
        # The checker: one frame's GPU packet against the oracle's bytes, and
        # the GPU's decoded planes against the oracle's decode.
        oc = O.config(5)
        of = O.Frame(w, h, 3, 12, frames[0].data, frames[0].alpha)
        result["bit_exact_encode"] = O.encode(of, oc) == packets[0]
        od = O.decode(packets[0])
        gd = outs[0][0].download(samples * 2, dtype=np.uint16)
        ga = outs[0][1].download(w * h * 2, dtype=np.uint16)
        result["bit_exact_decode"] = bool(np.array_equal(od.data, gd) and np.array_equal(od.alpha, ga))
        log("bench: cpu baseline")
        result["cpu_baseline"] = cpu_baseline(O, packets, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
