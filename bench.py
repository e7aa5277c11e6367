"""Benchmark of the MI355X density-fitted JK build.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload large|mid|water]

One step is one fused JK build (J^n, K^n, J_MO, K_MO, H_core for one set of
orbitals C) against a device-resident b, with C already on the device.
Default workload: nbf=1000, naux=3600, nbf5=300, b decay 0.02 (the size the
headline builds/s metric is quoted on). Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "JK builds/s at nbf=1000,naux=3600 split vs FP64 plain @1e-9 Eh; p50/p99 latency"
PEAKS = {"fp64_mfma_tflops": 78.6, "bf16_mfma_tflops": 2500.0, "hbm_gbs": 8000.0}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(w, sample_naux: int):
    """The oracle with the reference's call sequence (JKj_RI then JKH_MO_RI,
    half-transform done twice) on an naux-slice of the workload. Every
    contraction of the build is linear in naux, so the full-size time is the
    sample time times naux / sample_naux."""
    from oracle import jk_oracle as O
    from pynof_amd import synth

    b = synth.make_b(w.nbf, sample_naux, w.seed_b, w.lam)
    C = synth.make_C(w.nbf, w.seed_c)
    H = synth.make_H(w.nbf, w.seed_h)
    t0 = time.perf_counter()
    O.jkj_ri(C, b, w.nbf5)
    O.jkh_mo_ri(C, H, b, w.nbf5)
    dt = time.perf_counter() - t0
    scale = w.naux / sample_naux
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / (dt * scale),
        "unit": "builds/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"numpy FP64 oracle (reference call sequence, integrals.py:208-218 + :348-362) at "
                   f"nbf={w.nbf}, nbf5={w.nbf5}, naux={sample_naux} of {w.naux}: {dt:.2f} s measured, "
                   f"scaled x{scale:.0f} (linear in naux); BLAS threads={cores}; CPU: {cpu_model()}"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="large", choices=["water", "mid", "large"])
    ap.add_argument("--cpu-sample-naux", type=int, default=72)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every box here has one MI355X, so --gpus must be 1 "
              "(a multi-GPU build would shard orbitals q across GPUs; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch

    from pynof_amd import JKEngine, _lib, synth, workcount

    k = {"water": 0, "mid": 1, "large": 2}[args.workload]
    w = synth.workload(k)
    dev = torch.device("cuda", 0)
    print(f"[bench] workload {w.name}: generating b on the device", file=sys.stderr, flush=True)
    eng = JKEngine(0)
    if k == 2:
        btri = synth.device_b_packed(w.nbf, w.naux, w.seed_b, w.lam, JKEngine.packed_ld(w.nbf, w.naux))
        eng.upload_b_packed(btri, w.nbf, w.naux)
        del btri
        torch.cuda.empty_cache()
        data = "synthetic: b ~ N(0,1) symmetrised, x 0.3 exp(-0.02|mu-nu|), drawn on the device (torch Philox)"
    else:
        b, _, _, _ = synth.make_inputs(w)
        eng.upload_b(b)
        data = "synthetic: seeded numpy b/C/H/n generators"
    C0 = synth.make_C(w.nbf, w.seed_c)
    Cs = [C0]
    for j in range(2):
        Cs.append(Cs[-1] @ synth.rotation(w.nbf, 6000 + j))
    Cd = [torch.from_numpy(np.ascontiguousarray(c)).to(dev) for c in Cs]
    Hd = torch.from_numpy(synth.make_H(w.nbf, w.seed_h)).to(dev)
    nbf, naux, nbf5 = w.nbf, w.naux, w.nbf5
    out = {
        "J": torch.empty((nbf5, nbf, nbf), dtype=torch.float64, device=dev),
        "K": torch.empty((nbf5, nbf, nbf), dtype=torch.float64, device=dev),
        "J_MO": torch.empty((nbf5, nbf5), dtype=torch.float64, device=dev),
        "K_MO": torch.empty((nbf5, nbf5), dtype=torch.float64, device=dev),
        "H_core": torch.empty((nbf5,), dtype=torch.float64, device=dev),
    }
    stream = torch.cuda.current_stream(dev).cuda_stream
    L = _lib.lib()
    nst = L.pnjk_stage_count()
    names = [L.pnjk_stage_name(i).decode() for i in range(nst)]
    n_out = ("J", "K", "J_MO", "K_MO", "H_core")

    def step(i):
        eng.build(Cd[i % len(Cd)], nbf5, H=Hd, precision="fp64", outputs=n_out, out=out, stream=stream)

    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    _lib.check(L.pnjk_set_profiling(eng._h, 1))
    import ctypes

    st = (ctypes.c_double * nst)()
    stage_ms = np.zeros((args.steps, nst))
    lat = []
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        step(i)
        _lib.check(L.pnjk_stage_times(eng._h, st))  # waits for this build's last event
        lat.append(time.perf_counter() - ts)
        stage_ms[i] = np.frombuffer(st, dtype=np.float64)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.check(L.pnjk_set_profiling(eng._h, 0))
    ms_per_step = 1e3 * elapsed / args.steps

    flops = workcount.stage_flops(nbf, naux, nbf5)
    avg = stage_ms.mean(axis=0)
    stages = {names[i]: round(float(avg[i]), 4) for i in range(nst)}
    dom = max((n for n in names if n in flops), key=lambda n: stages[n])
    dom_ms = stages[dom]
    achieved = flops[dom] / (dom_ms * 1e-3) / 1e12
    peak = PEAKS["fp64_mfma_tflops"]
    roofline = {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": None,
                "per_stage_tflops": {n: round(flops[n] / (stages[n] * 1e-3) / 1e12, 3)
                                     for n in flops if stages.get(n, 0) > 0}}
    total_flops = workcount.build_flops(nbf, naux, nbf5)
    line = {
        "metric": METRIC,
        "value": round(1e3 / ms_per_step, 4),
        "unit": "builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": data,
        "config": {"workload": w.name, "nbf": nbf, "naux": naux, "nbf5": nbf5, "form": "fp64",
                   "outputs": list(n_out), "device_resident": True},
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
        "build_tflops": round(total_flops / (ms_per_step * 1e-3) / 1e12, 3),
        "stage_ms": stages,
        "roofline": roofline,
    }
    eng.close()
    del out, Cd, Hd
    torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(w, min(args.cpu_sample_naux, w.naux))
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
