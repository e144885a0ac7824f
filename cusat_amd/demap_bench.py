"""Per-code demap time of the APSK constellations (cusat_fec_demap, one
launch per code): normal FECFRAMEs generated on the GPU at each MODCOD's
threshold (lib/cusat_dsp/acm_config.h:54-121) + a margin, demapped with the
scaling of that Es/N0, microseconds per launch (HIP events on the launch
stream).  64/128/256APSK go through the sorted path, 16/32APSK through the
per-symbol kernel.

    python -m cusat_amd.demap_bench --frames 293 --reps 20
"""
import argparse
import json
import sys

# (rate, modulation, threshold Es/N0 dB): the APSK MODCODs of the combined
# S2/S2X stream
APSK = [("4/5", "64APSK", 16.10), ("135/180", "128APSK", 18.00), ("135/180", "256APSK", 23.34),
        ("2/3", "16APSK", 9.17), ("3/4", "32APSK", 12.93)]


def run(frames=293, reps=20, margin=0.3, seed=7, log=print):
    import torch
    import cusat_amd as ca
    from cusat_amd.stream_bench import snr_scales
    out = []
    for rate, mod, thr in APSK:
        es = thr + margin
        c = ca.FecContext(0, rate, mod, max_batch=frames)
        inf = c.info
        sym = torch.empty((frames, inf["xfec_len"], 2), dtype=torch.float32, device="cuda")
        c.gen_frames_dev(seed, frames, es, sym)
        llr = torch.empty((frames, inf["n_ldpc"]), dtype=torch.int16, device="cuda")
        nv, gain = snr_scales(es)
        for _ in range(3):
            c.demap(sym, llr, frames, nv, gain)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            c.demap(sym, llr, frames, nv, gain)
        ev[1].record()
        torch.cuda.synchronize()
        us = ev[0].elapsed_time(ev[1]) * 1000.0 / reps
        sym_total = frames * inf["xfec_len"]
        r = {"modcod": [0, rate, mod], "es_n0_db": round(es, 2), "frames": frames, "us_per_launch": round(us, 1),
             "msym_per_s": round(sym_total / us, 1)}
        log("%-8s %-8s %d frames  %.1f us per launch  %.1f Msym/s" % (mod, rate, frames, us, r["msym_per_s"]))
        out.append(r)
        c.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--frames", type=int, default=293)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--margin", type=float, default=0.3, help="dB above the MODCOD threshold")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    res = run(a.frames, a.reps, a.margin, log=lambda s: print(s, flush=True))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
