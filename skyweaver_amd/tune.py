"""Writes the tuning table that BeamformerPipeline reads at creation.

For each workload below, sw_pipeline_autotune times the MFMA coherent
beamformer with 2, 4 and 8 waves (of 32 time rows) per workgroup and, where
no FFT length is prescribed, the coherent dedisperser's FFT lengths with and
without shared forward transforms, on synthetic input on the current GPU.
The table (JSON) is read back by the library's sw_pipeline_create.

usage: python -m skyweaver_amd.tune [--out PATH] [--chans N]
"""
import argparse
import datetime
import json
import os
import sys

CHBW = 208984.375    # Hz, MeerKAT's 856 MHz / 4096 channels
CFREQ = 1284e6

# (label, antennas, channels, samples, beams, tscrunch, Stokes, coherent DMs, FFT lengths or None)
WORKLOADS = [
    ("64 ant x 64 ch x 8192 samples, 64 beams, ts 4, DM 0", 64, 64, 8192, 64, 4, "I", [0.0], None),
    ("64 ant x 64 ch x 65536 samples, 800 beams, ts 16, DM 50 at FFT 16384", 64, 64, 65536, 800, 16, "I",
     [50.0], [16384]),
    ("64 ant x 64 ch x 65536 samples, 800 beams, ts 16, 8 DMs 0-350", 64, 64, 65536, 800, 16, "I",
     [0.0, 5.0, 20.0, 50.0, 100.0, 180.0, 260.0, 350.0], None),
    ("64 ant x 64 ch x 131072 samples, 512 beams, ts 8, 8 DMs 0-350", 64, 64, 131072, 512, 8, "I",
     [0.0, 5.0, 20.0, 50.0, 100.0, 180.0, 260.0, 350.0], None),
    ("80 ant x 128 ch x 65536 samples, 1024 beams, ts 16, IQUV, DM 0", 80, 128, 65536, 1024, 16, "IQUV",
     [0.0], None),
    ("64 ant x 256 ch x 262144 samples, 512 beams, ts 8, DMs 0 and 100", 64, 256, 262144, 512, 8, "I",
     [0.0, 100.0], None),
]


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=os.path.join(here, "tuning", "gfx950.json"))
    ap.add_argument("--chans", type=int, default=64, help="channels the timings run on (scaled to the workload)")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(here))
    import skyweaver_amd as sw
    entries = []
    for label, A, F, T, B, ts, stokes, dms, fft in WORKLOADS:
        print(f"tuning {label}", file=sys.stderr, flush=True)
        e = sw.pipeline_autotune(A, F, B, T, ts, dms, CFREQ, CHBW * F, stokes=stokes, fft_len=fft,
                                 tune_chans=args.chans)
        e["workload"] = label
        entries.append(e)
        print(json.dumps(e), file=sys.stderr, flush=True)
    table = {"device": "gfx950 (MI355X)",
             "generated": datetime.datetime.utcnow().strftime("%Y-%m-%d %H:%M UTC"),
             "tool": "python -m skyweaver_amd.tune --chans %d" % args.chans,
             "note": "beams_ms_per_gulp: coherent + incoherent beam device time per gulp for 2, 4 and 8 waves "
                     "per workgroup (-1: the shape does not fit), scaled from the tuned channels to the workload",
             "entries": entries}
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
        f.write("\n")
    print(args.out)


if __name__ == "__main__":
    main()
