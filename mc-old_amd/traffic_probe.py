"""HBM traffic probe: run under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE).

A 512 MiB device-to-device copy first, the calibration kernel of known size,
then a forward run of the continuous-energy shield with plain launches (no
graph), so every event-loop kernel is a dispatch of its own; the run's tail
goes to the drain kernel, so the iterations counted run on a full bank:

    rocprofv3 --pmc FETCH_SIZE --kernel-trace --output-format csv -d DIR -o run -- \\
        python3 mc-old_amd/traffic_probe.py [histories] [xs_dir]

It prints the copy size and the event counts the per-event figures need.
"""
import importlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

COPY_BYTES = 512 * 1024 * 1024


def main():
    import torch
    n = int(float(sys.argv[1])) if len(sys.argv) > 1 else 10_000_000
    xs = sys.argv[2] if len(sys.argv) > 2 else os.path.join(os.path.dirname(HERE), "_bench_xs")
    mc = importlib.import_module("mc-old_amd")
    decks = importlib.import_module("mc-old_amd.decks")
    if not os.path.exists(os.path.join(xs, "926056.txt")):
        decks.write_shield_nuclides(xs)
    a = torch.empty(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    p = mc.Problem.from_string(decks.shield_deck(n), xs)
    e = mc.Engine(p, use_graph=False, drain_divisor=1, profile=True)
    f = e.forward(n_histories=n)
    ks = e.kernel_stats()
    print(json.dumps({"copy_bytes": COPY_BYTES, "flights": int(f.flights),
                      "items": {k: int(v["items"]) for k, v in ks.items()},
                      "launches": {k: int(v["launches"]) for k, v in ks.items()}}), flush=True)


if __name__ == "__main__":
    main()
