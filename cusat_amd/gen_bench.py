"""Test-traffic generation rate: the host transmit chain (cusat_gen_frames,
std::thread workers) against the GPU one (cusat_gen_frames_dev), same
MODCOD and Es/N0, frames per second and info Gbit/s.

    python -m cusat_amd.gen_bench --modcod 0,3/4,8PSK --frames 4096
"""
import argparse
import json
import os
import sys
import time


def run(framesize=0, rate="3/4", mod="8PSK", frames=4096, host_frames=512, es=8.0, reps=5, log=print):
    import torch
    import cusat_amd as ca
    g = ca.geometry(framesize, rate, mod)
    kb = g["k_bch"]
    t0 = time.time()
    ca.gen_frames(framesize, rate, mod, 1, host_frames, es)
    t_host = time.time() - t0
    ctx = ca.FecContext(framesize, rate, mod, max_trials=10, max_batch=frames)
    sym = torch.empty((frames, g["xfec_len"], 2), dtype=torch.float32, device="cuda")
    info = torch.empty((frames, kb // 8), dtype=torch.uint8, device="cuda")
    ctx.gen_frames_dev(1, frames, es, sym, info_out=info)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for r in range(reps):
        ctx.gen_frames_dev(2 + r, frames, es, sym, info_out=info)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    res = {"modcod": [framesize, rate, mod], "es_n0_db": es,
           "host_frames": host_frames, "host_threads": min(16, os.cpu_count() or 1),
           "host_frames_per_s": round(host_frames / t_host, 1),
           "gpu_frames": frames, "gpu_ms_per_launch": round(ms, 3),
           "gpu_frames_per_s": round(frames / (ms * 1e-3), 1),
           "gpu_info_gbps": round(frames * kb / (ms * 1e-3) / 1e9, 2)}
    res["speedup"] = round(res["gpu_frames_per_s"] / res["host_frames_per_s"], 1)
    log(json.dumps(res))
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--modcod", default="0,3/4,8PSK", help="framesize,rate,modulation")
    ap.add_argument("--frames", type=int, default=4096)
    ap.add_argument("--host-frames", type=int, default=512)
    ap.add_argument("--es", type=float, default=8.0)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    fs, rate, mod = a.modcod.split(",")
    r = run(int(fs), rate, mod, a.frames, a.host_frames, a.es, log=lambda s: print(s, flush=True))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(r, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
