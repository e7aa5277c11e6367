"""Benchmark of the MI355X density-fitted JK build.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload large|mid|water|batch64|serving|gemmref]

One step is one fused JK build (J^n, K^n, J_MO, K_MO, H_core for one set of
orbitals C) against a device-resident b, with C already on the device.
Default workload: nbf=1000, naux=3600, nbf5=300, b decay 0.02, the size the
headline builds/s metric is quoted on.

The timed region runs K builds as orbital-optimisation requests of 20 builds:
C is rotated by a seeded exp(kappa), ||kappa||_F = 1e-2, between builds; the
first 19 builds of a request use the split form and the last one, whose energy
the request reports, uses split_refine (J_MO, K_MO, H_core through the FP64
path). `value` is builds/s of that stream. The same run also times every form
alone (fp64, split, split_refine) and states each form's errors against the
FP64 form at this size. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "JK builds/s at nbf=1000,naux=3600 split vs FP64 plain @1e-9 Eh; p50/p99 latency"
PEAK_FP64 = 78.6      # TFLOP/s, MI355X dense FP64 MFMA
PEAK_BF16 = 2500.0    # TFLOP/s, MI355X dense BF16 MFMA
PEAK_HBM = 8000.0     # GB/s, MI355X HBM3E
REQUEST = 20          # builds per orbital-optimisation request
OUTS = ("J", "K", "J_MO", "K_MO", "H_core")
# rocprofv3 summary of this workload (kernel times and --pmc HBM traffic per dispatch)
PROFILE = os.path.join(ROOT, "profiles", "jk_build_large_kernels.md")
SPLIT_KERNEL = {"half_transform": "k_gemm_x3<5, 5, 2, 2, ProbHT>", "quarter_transform": "k_gemm_x3<5, 4, 2, 2, ProbQT>",
                "coulomb_jn": "k_gemm_x3<5, 4, 2, 2, ProbJN>", "exchange_kn": "k_gemm_x3<4, 4, 2, 2, ProbSYRK>"}
FP64_KERNEL = {"half_transform": "k_gemm_f64<5, 4, 8, ProbHT>", "quarter_transform": "k_gemm_f64<5, 2, 16, ProbQT>",
               "coulomb_jn": "k_gemm_f64<5, 4, 8, ProbJN>", "exchange_kn": "k_gemm_f64<4, 4, 8, ProbSYRK>"}


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


def cpu_baseline_batch(ws, flops_total: float, every: int = 8, naux_div: int = 8):
    """The oracle (reference call sequence) on every `every`-th instance of the batch by
    size, each on an naux-slice of 1/naux_div: the sample time is scaled by naux (every
    contraction is linear in it) and then by the batch's total over the sample's useful
    flops. Returns builds/s of the whole batch."""
    from oracle import jk_oracle as O
    from pynof_amd import synth, workcount

    sample = sorted(ws, key=lambda w: w.nbf)[every // 2::every]
    t_full, f_sample, t_meas = 0.0, 0.0, 0.0
    for w in sample:
        na = max(8, w.naux // naux_div)
        b = synth.make_b(w.nbf, na, w.seed_b, w.lam)
        C = synth.make_C(w.nbf, w.seed_c)
        H = synth.make_H(w.nbf, w.seed_h)
        t0 = time.perf_counter()
        O.jkj_ri(C, b, w.nbf5)
        O.jkh_mo_ri(C, H, b, w.nbf5)
        dt = time.perf_counter() - t0
        t_meas += dt
        t_full += dt * w.naux / na
        f_sample += workcount.build_flops(w.nbf, w.naux, w.nbf5)
    t_batch = t_full * flops_total / f_sample
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": len(ws) / t_batch,
        "unit": "builds/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"numpy FP64 oracle (reference call sequence) on {len(sample)} of {len(ws)} instances "
                   f"(every {every}th by nbf: {[w.nbf for w in sample]}), each at naux/{naux_div}: "
                   f"{t_meas:.1f} s measured, scaled by naux per instance and by total/sample useful flops "
                   f"({flops_total / f_sample:.2f}); BLAS threads={cores}; CPU: {cpu_model()}"),
    }


def profiled_traffic(kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 --pmc summary:
    FETCH_SIZE scaled by the copy-kernel calibration plus WRITE_SIZE. None when the
    summary or the kernel's row is absent."""
    try:
        with open(PROFILE) as f:
            text = f.read()
    except OSError:
        return None
    for row in text.splitlines():
        cells = [c.strip() for c in row.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[0] == f"`{kernel}`":
            try:
                fetch, write = float(cells[2]), float(cells[3])
            except ValueError:
                continue
            return {"bytes": round((fetch + write) * 2**30), "fetch_GiB": fetch, "write_GiB": write,
                    "source": "profiles/jk_build_large_kernels.md (rocprofv3 --pmc FETCH_SIZE x2 calibration, WRITE_SIZE)"}
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="large", choices=["water", "mid", "large", "batch64", "serving", "gemmref"])
    ap.add_argument("--requests", type=int, default=120, help="serving: requests per run")
    ap.add_argument("--batch-range", default="40,400",
                    help="batch64: nbf range of the 64 instances (the batch workload is 40,400)")
    ap.add_argument("--lanes", default="128,256",
                    help="serving: ascending nbf bounds of the size-class lanes, or 'none' for lockstep steps")
    ap.add_argument("--form-steps", type=int, default=5, help="builds timed per single form")
    ap.add_argument("--cpu-sample-naux", type=int, default=72,
                    help="naux slice of the CPU baseline for workloads with nbf > 400 (smaller ones run in full)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip", nargs="*", default=[], choices=["b_producer"],
                    help="report sections to leave out (the headline stream is always run)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every box here has one MI355X, so --gpus must be 1 "
              "(a multi-GPU build would shard the orbitals q over GPUs and all-gather d)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch64":
        return bench_batch(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "gemmref":
        return bench_gemmref(args)

    import torch

    from pynof_amd import JKEngine, _lib, synth, workcount

    k = {"water": 0, "mid": 1, "large": 2}[args.workload]
    w = synth.workload(k)
    nbf, naux, nbf5 = w.nbf, w.naux, w.nbf5
    dev = torch.device("cuda", 0)
    log(f"workload {w.name}: b on the device")
    eng = JKEngine(0)
    if k == 2:
        btri = synth.device_b_packed(nbf, naux, w.seed_b, w.lam, JKEngine.packed_ld(nbf, naux))
        eng.upload_b_packed(btri, nbf, naux)
        del btri
        torch.cuda.empty_cache()
        data = "synthetic: b ~ N(0,1) symmetrised x 0.3 exp(-0.02|mu-nu|) drawn on the device; seeded C, H, n"
    else:
        b, _, _, _ = synth.make_inputs(w)
        eng.upload_b(b)
        data = "synthetic: seeded numpy b/C/H/n generators"
    n_occ = synth.make_n(nbf5, w.ndoc, w.seed_n)
    # C sequence of one request: C_{i+1} = C_i exp(kappa_i)
    Cs = [synth.make_C(nbf, w.seed_c)]
    for j in range(REQUEST - 1):
        Cs.append(Cs[-1] @ synth.rotation(nbf, 6000 + j))
    Cd = [torch.from_numpy(np.ascontiguousarray(c)).to(dev) for c in Cs]
    Hd = torch.from_numpy(synth.make_H(nbf, w.seed_h)).to(dev)

    def alloc():
        return {"J": torch.empty((nbf5, nbf, nbf), dtype=torch.float64, device=dev),
                "K": torch.empty((nbf5, nbf, nbf), dtype=torch.float64, device=dev),
                "J_MO": torch.empty((nbf5, nbf5), dtype=torch.float64, device=dev),
                "K_MO": torch.empty((nbf5, nbf5), dtype=torch.float64, device=dev),
                "H_core": torch.empty((nbf5,), dtype=torch.float64, device=dev)}

    out = alloc()
    stream = torch.cuda.current_stream(dev).cuda_stream
    L = _lib.lib()
    nst = L.pnjk_stage_count()
    names = [L.pnjk_stage_name(i).decode() for i in range(nst)]
    st = (ctypes.c_double * nst)()

    def build(i, form, o=None):
        eng.build(Cd[i % REQUEST], nbf5, H=Hd, precision=form, outputs=OUTS, out=o or out, stream=stream)

    def timed(forms, steps):
        """Runs `steps` builds, form forms(i); returns (seconds, latencies, stage ms per build)."""
        _lib.check(L.pnjk_set_profiling(eng._h, 1))
        lat, stages = [], []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            ts = time.perf_counter()
            build(i, forms(i))
            _lib.check(L.pnjk_stage_times(eng._h, st))  # waits for this build's last event
            lat.append(time.perf_counter() - ts)
            stages.append((forms(i), np.frombuffer(st, dtype=np.float64).copy()))
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        _lib.check(L.pnjk_set_profiling(eng._h, 0))
        return el, lat, stages

    def energy(o):
        e1, e2 = synth.energy_terms(n_occ, o["J_MO"].cpu().numpy(), o["K_MO"].cpu().numpy(),
                                    o["H_core"].cpu().numpy())
        return e1 + e2

    # ---- every form alone, and its errors against the FP64 form
    flops = workcount.stage_flops(nbf, naux, nbf5)
    total_flops = workcount.build_flops(nbf, naux, nbf5)
    ref = alloc()
    forms = {}
    for form in ("fp64", "split", "split_refine"):
        log(f"form {form}: warmup {args.warmup}, {args.form_steps} builds")
        for i in range(args.warmup):
            build(i, form)
        el, lat, stg = timed(lambda i, f=form: f, args.form_steps)
        avg = np.mean([s for _, s in stg], axis=0)
        res = {"builds_per_s": round(args.form_steps / el, 3), "ms_per_build": round(1e3 * el / args.form_steps, 3),
               "stage_ms": {names[i]: round(float(avg[i]), 4) for i in range(nst)}}
        o = ref if form == "fp64" else out
        build(REQUEST - 1, form, o)
        torch.cuda.synchronize()
        if form == "fp64":
            e_ref = energy(ref)
            res["energy_Eh"] = e_ref
        else:
            e = energy(out)
            res["energy_err_Eh"] = e - e_ref
            res["rel_err_vs_fp64"] = {
                nm: float(torch.max(torch.abs(out[nm] - ref[nm])) / torch.max(torch.abs(ref[nm]))) for nm in OUTS}
        forms[form] = res
    del ref
    torch.cuda.empty_cache()

    # ---- headline: requests of 20 builds, split x19 + split_refine x1
    def seq_form(i):
        return "split_refine" if (i + 1) % REQUEST == 0 else "split"

    log(f"headline stream: warmup {args.warmup}, {args.steps} builds")
    for i in range(args.warmup):
        build(i, "split")
    el, lat, stg = timed(seq_form, args.steps)
    ms_per_step = 1e3 * el / args.steps
    # final energy of the request (the split_refine build) against the FP64 energy of the same C
    build(REQUEST - 1, "split_refine")
    torch.cuda.synchronize()
    e_final = energy(out)

    # roofline of the dominant kernel of the stream's split builds (events over the timed region)
    split_stages = np.mean([s for f, s in stg if f == "split"] or [stg[0][1]], axis=0)
    sms = {names[i]: float(split_stages[i]) for i in range(nst)}
    dom = max((n for n in names if n in flops), key=lambda n: sms[n])
    achieved = flops[dom] / (sms[dom] * 1e-3) / 1e12
    tr = profiled_traffic(SPLIT_KERNEL.get(dom, ""))
    roofline = {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 3), "peak": PEAK_BF16,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_BF16, 4),
                "traffic": tr["bytes"] if tr else None, "traffic_detail": tr,
                "note": "BF16x3: each useful multiply-add issues 3 bf16 MFMA products; "
                        f"issued rate {3 * achieved:.1f} TFLOP/s = {3 * achieved / PEAK_BF16:.3f} of peak",
                "per_stage_tflops": {n: round(flops[n] / (sms[n] * 1e-3) / 1e12, 3) for n in flops if sms[n] > 0}}
    fp64_stages = forms["fp64"]["stage_ms"]
    fd = max((n for n in flops), key=lambda n: fp64_stages[n])
    tr64 = profiled_traffic(FP64_KERNEL.get(fd, ""))
    roofline_fp64 = {"bound": "mfma", "kernel": fd, "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "achieved": round(flops[fd] / (fp64_stages[fd] * 1e-3) / 1e12, 3),
                     "traffic": tr64["bytes"] if tr64 else None}
    roofline_fp64["frac"] = round(roofline_fp64["achieved"] / PEAK_FP64, 4)

    # ---- the drop-in boundary's host path: host C in, host J and K out (4.8 GB over PCIe into
    # the page-locked arrays the engine allocates for outputs this large); not `value`
    log("host outputs (drop-in path)")
    C_host = Cs[0]
    eng.build(C_host, nbf5, H=None, precision="split", outputs=("J", "K"))
    host_ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        hres = eng.build(C_host, nbf5, H=None, precision="split", outputs=("J", "K"))
        host_ms.append(1e3 * (time.perf_counter() - t0))
    d2h_bytes = hres["J"].nbytes + hres["K"].nbytes
    pinned = bool(torch.from_numpy(hres["J"]).is_pinned())
    del hres
    split_ms = forms["split"]["ms_per_build"]
    host_path = {"form": "split", "ms_per_build": round(float(np.median(host_ms)), 2),
                 "ms_per_build_device_outputs": split_ms, "d2h_bytes": d2h_bytes, "page_locked": pinned,
                 "d2h_GBps": round(d2h_bytes / ((float(np.median(host_ms)) - split_ms) * 1e-3) / 1e9, 1)}

    # ---- computeJKH_MO's build (J_MO, K_MO, H_core only; the orbital optimiser's energy call,
    # pnof.py:1107): the FP64 form takes the triangular route there (T = L C, W = C^T T)
    log("MO-only builds (computeJKH_MO)")
    mo_outs = ("J_MO", "K_MO", "H_core")
    mo_only = {}
    for form in ("fp64", "split"):
        eng.build(Cd[0], nbf5, H=Hd, precision=form, outputs=mo_outs, out=out, stream=stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = max(1, args.form_steps)
        e0.record()
        for _ in range(reps):
            eng.build(Cd[0], nbf5, H=Hd, precision=form, outputs=mo_outs, out=out, stream=stream)
        e1.record()
        torch.cuda.synchronize()
        mo_only[form] = {"ms": round(e0.elapsed_time(e1) / reps, 3)}
    for form, arith in (("fp64", "FP64"), ("split", "BF16x3")):
        mo_only[form]["route"] = (f"triangular in {arith} (T = L C, W = C^T T)" if nbf5 < nbf
                                  else "half + quarter transform")
        fst = forms[form]["stage_ms"]
        mo_only[form]["full_build_ht_plus_qt_ms"] = round(fst["half_transform"] + fst["quarter_transform"], 3)
    # ---- computeJKj's build (J and K only; ENERGY1r with no1 > 0, utils.py:421): the quarter
    # transform forms d alone (window row 0) and J_MO, K_MO are skipped
    log("J, K-only builds (computeJKj)")
    jk_only = {}
    for form in ("fp64", "split"):
        eng.build(Cd[0], nbf5, precision=form, outputs=("J", "K"), out=out, stream=stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = max(1, args.form_steps)
        e0.record()
        for _ in range(reps):
            eng.build(Cd[0], nbf5, precision=form, outputs=("J", "K"), out=out, stream=stream)
        e1.record()
        torch.cuda.synchronize()
        jk_only[form] = {"ms": round(e0.elapsed_time(e1) / reps, 3),
                         "full_build_ms": forms[form]["ms_per_build"]}

    # ---- orbital-gradient MO integrals (JKH_MO_tmp): HT + full-row quarter transform + H_mat
    mo = {}
    mo_flops = 2.0 * nbf5 * nbf * nbf * naux + 2.0 * nbf * nbf5 * nbf * naux + 2.0 * nbf * nbf * nbf5 * 2
    mo_out = {"b_MO": torch.empty((nbf, nbf5, naux), dtype=torch.float64, device=dev),
              "H_mat": torch.empty((nbf, nbf5), dtype=torch.float64, device=dev)}
    mo_ref = None
    for form in ("fp64", "split"):
        log(f"MO integrals {form}")
        eng.mo_integrals(Cd[0], nbf5, H=Hd, precision=form, out=mo_out, stream=stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        reps = max(1, args.form_steps)
        for _ in range(reps):
            eng.mo_integrals(Cd[0], nbf5, H=Hd, precision=form, out=mo_out, stream=stream)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        mo[form] = {"ms": round(ms, 3), "useful_tflops": round(mo_flops / (ms * 1e-3) / 1e12, 2)}
        if form == "fp64":
            mo_ref = mo_out["b_MO"].clone()
        else:
            mo[form]["rel_err_b_MO_vs_fp64"] = float(torch.linalg.norm(mo_out["b_MO"] - mo_ref) /
                                                     torch.linalg.norm(mo_ref))
    mo["flops"] = mo_flops
    mo["speedup_split_vs_fp64"] = round(mo["fp64"]["ms"] / mo["split"]["ms"], 3)
    mo["roofline"] = {"bound": "mfma", "unit": "TFLOP/s", "scope": "whole call (HT + b_MO + H_mat)",
                      "fp64": {"achieved": mo["fp64"]["useful_tflops"], "peak": PEAK_FP64,
                               "frac": round(mo["fp64"]["useful_tflops"] / PEAK_FP64, 4)},
                      "split": {"achieved": mo["split"]["useful_tflops"], "peak": PEAK_BF16,
                                "frac": round(mo["split"]["useful_tflops"] / PEAK_BF16, 4),
                                "issued_frac": round(3 * mo["split"]["useful_tflops"] / PEAK_BF16, 4)}}
    del mo_out, mo_ref
    torch.cuda.empty_cache()

    # ---- orbital-gradient Lagrangian (computeLagrange2): MO integrals + contractions on the device
    rng = np.random.default_rng(7000)
    cj = rng.standard_normal((nbf5, nbf5))
    ck = rng.standard_normal((nbf5, nbf5))
    lag_in = [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (n_occ, cj + cj.T, ck + ck.T)]
    nb = w.ndoc
    lag = {}
    lag_ref = None
    for form in ("fp64", "split"):
        log(f"Lagrangian {form}")
        lo = eng.lagrangian(Cd[0], nbf5, Hd, *lag_in, nb, nb, precision=form, stream=stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        reps = max(1, args.form_steps)
        for _ in range(reps):
            lo = eng.lagrangian(Cd[0], nbf5, Hd, *lag_in, nb, nb, precision=form, out=lo, stream=stream)
        e1.record()
        torch.cuda.synchronize()
        lag[form] = {"ms": round(e0.elapsed_time(e1) / reps, 3)}
        if form == "fp64":
            lag_ref = lo["elag"].clone()
        else:
            lag[form]["rel_err_elag_vs_fp64"] = float(torch.linalg.norm(lo["elag"] - lag_ref) /
                                                      torch.linalg.norm(lag_ref))
    lag["exchange_flops"] = 2.0 * nbf * nbf5 * nbf5 * naux

    # ---- generalised Fock matrices from J^n, K^n (computeF_RC) and elag from F (computeLagrange)
    log("Fock matrices and Lagrangian from F")
    build(0, "fp64")
    F = eng.fock_rc(out["J"], out["K"], lag_in[0], Hd, lag_in[1], lag_in[2], 0, nb, nb, stream=stream)
    elag_f = eng.lagrange_fock(F, Cd[0], nbf, stream=stream)
    reps = max(1, args.form_steps)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record()
    for _ in range(reps):
        eng.fock_rc(out["J"], out["K"], lag_in[0], Hd, lag_in[1], lag_in[2], 0, nb, nb, out=F, stream=stream)
    ev[1].record()
    for _ in range(reps):
        eng.lagrange_fock(F, Cd[0], nbf, out=elag_f, stream=stream)
    ev[2].record()
    torch.cuda.synchronize()
    f_ms, l_ms = ev[0].elapsed_time(ev[1]) / reps, ev[1].elapsed_time(ev[2]) / reps
    f_flops = 2.0 * nbf5 * (2 * nbf5) * nbf * nbf
    fock = {"fock_ms": round(f_ms, 3), "fock_flops": f_flops,
            "fock_tflops": round(f_flops / (f_ms * 1e-3) / 1e12, 2),
            "fock_bytes_min": 8.0 * 3 * nbf5 * nbf * nbf,
            "lagrange_from_F_ms": round(l_ms, 3),
            "lagrange_from_F_GBps": round(8.0 * nbf5 * nbf * nbf / (l_ms * 1e-3) / 1e9, 1)}
    fock["roofline"] = {"fock_gemm": {"bound": "mfma", "unit": "TFLOP/s", "achieved": fock["fock_tflops"],
                                      "peak": PEAK_FP64, "frac": round(fock["fock_tflops"] / PEAK_FP64, 4)},
                        "lagrange_from_F": {"bound": "hbm", "unit": "GB/s", "achieved": fock["lagrange_from_F_GBps"],
                                            "peak": PEAK_HBM, "frac": round(fock["lagrange_from_F_GBps"] / PEAK_HBM, 4),
                                            "scope": "whole call; algorithmic bytes = F read once"}}
    del lag_in, lag_ref, lo, F, elag_f
    torch.cuda.empty_cache()

    # ---- occupation step (calcocce + calcoccg, GNOF, softmax) on this build's J_MO, K_MO, H_core
    log("occupation step")
    occ_ndoc, occ_ncwo = 60, (nbf5 - 60) // 60  # nbf5 = ndoc (1 + ncwo) = 300
    occ = {}
    if occ_ndoc * (1 + occ_ncwo) == nbf5:
        pn = JKEngine.pnof(0, occ_ndoc, 0, occ_ncwo, ipnof=8, occ_method="Softmax")
        gam = torch.from_numpy(np.random.default_rng(7100).uniform(-2.0, 0.5, occ_ndoc * occ_ncwo)).to(dev)
        oo = eng.occupation_step(gam, pn, out["J_MO"], out["K_MO"], out["H_core"], stream=stream)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        reps = 20
        ev[0].record()
        for _ in range(reps):
            eng.occupation_step(gam, pn, out["J_MO"], out["K_MO"], out["H_core"], out=oo, stream=stream)
        ev[1].record()
        torch.cuda.synchronize()
        occ = {"functional": "GNOF", "occ_method": "Softmax", "ndoc": occ_ndoc, "ncwo": occ_ncwo,
               "nv": occ_ndoc * occ_ncwo, "energy_and_gradient_ms": round(ev[0].elapsed_time(ev[1]) / reps, 4)}
        if not args.no_cpu_baseline:  # the oracle's numpy restatement of the same step, for scale
            from oracle import pnof_oracle as PO

            po = PO.Pnof(0, occ_ndoc, 0, occ_ncwo, ipnof=8, occ_method="Softmax")
            Jh, Kh, hh = (out[k].cpu().numpy() for k in ("J_MO", "K_MO", "H_core"))
            gh = gam.cpu().numpy()
            t0 = time.perf_counter()
            PO.calcocce(gh, Jh, Kh, hh, po)
            PO.calcoccg(gh, Jh, Kh, hh, po)
            occ["oracle_numpy_ms"] = round(1e3 * (time.perf_counter() - t0), 1)

    # ---- b from 3-centre integrals (integrals.py:35-49): pack + metric eigenpairs + FP64 GEMM
    bprod = {}
    if "b_producer" not in args.skip:
        bprod = bench_b_producer(JKEngine, nbf, naux, dev)

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
        "dtype": "bf16x3 (fp32 accumulate) + f64 refinement",
        "data": data,
        "config": {"workload": w.name, "nbf": nbf, "naux": naux, "nbf5": nbf5, "ndoc": w.ndoc,
                   "form": f"orbital-optimisation requests of {REQUEST} builds: split x{REQUEST - 1} + "
                           "split_refine x1 (final energy)",
                   "outputs": list(OUTS), "device_resident": True},
        "speedup_vs_fp64": round((1e3 / ms_per_step) / forms["fp64"]["builds_per_s"], 3),
        "final_energy_err_Eh": e_final - forms["fp64"]["energy_Eh"],
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
        "forms": forms,
        "fp64_build_tflops": round(total_flops * forms["fp64"]["builds_per_s"] / 1e12, 3),
        "roofline": roofline,
        "roofline_fp64": roofline_fp64,
        "host_outputs": host_path,
        "mo_only_builds": mo_only,
        "jk_only_builds": jk_only,
        "mo_integrals": mo,
        "lagrangian": lag,
        "fock": fock,
        "occupation_step": occ,
        "b_producer": bprod,
    }
    eng.close()
    del out, Cd, Hd
    torch.cuda.empty_cache()
    # A plain 1 GiB device copy: the reference kernel for scaling FETCH_SIZE/WRITE_SIZE
    # counters of profiled runs (known bytes moved).
    src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    if not args.no_cpu_baseline:
        log("cpu baseline")
        # The water and mid-size workloads are timed in full; the large one on an naux slice.
        line["cpu_baseline"] = cpu_baseline(w, w.naux if w.nbf <= 400 else min(args.cpu_sample_naux, w.naux))
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def bench_b_producer(JKEngine, nbf, naux, dev):
    """b from 3-centre integrals (integrals.py:35-49) at full size: pack + metric eigenpairs +
    FP64 GEMM, with (mu nu|P) and the metric drawn on the device."""
    import torch

    log("b from 3-centre integrals")
    gen = torch.Generator(device=dev)
    gen.manual_seed(7200)
    g3 = torch.randn((nbf, nbf, naux), dtype=torch.float64, device=dev, generator=gen)
    mnk = g3 + g3.transpose(0, 1)
    del g3
    A = torch.randn((naux, naux), dtype=torch.float64, device=dev, generator=gen) / np.sqrt(naux)
    metric = A @ A.T + torch.eye(naux, dtype=torch.float64, device=dev)
    del A
    torch.cuda.synchronize()
    eng3 = JKEngine(0)
    eng3.upload_b_3c(mnk, metric)  # first call: rocBLAS/rocSOLVER set-up
    t0 = time.perf_counter()
    eng3.upload_b_3c(mnk, metric)
    bp_s = time.perf_counter() - t0
    ld = JKEngine.packed_ld(nbf, naux)
    bp_flops = 2.0 * (nbf * (nbf + 1) // 2) * ld * ld
    bprod = {"s": round(bp_s, 3), "gemm_flops": bp_flops, "gemm_tflops_if_all_time": round(bp_flops / bp_s / 1e12, 2),
             "input": "device (mu nu|P) 28.8 GB dense + metric", "eigensolver": "rocsolver_dsyevd"}
    eng3.close()
    del mnk, metric
    torch.cuda.empty_cache()
    return bprod


def bench_batch(args):
    """64 independent builds, nbf ~ U[40, 400] (seeded), naux = ceil8(3.6 nbf), nbf5 = nbf:
    sequential single-instance builds against one grouped pass (eager, and replayed
    from a captured hipGraph), FP64 and split forms."""
    import torch

    from pynof_amd import JKEngine, synth, workcount
    from pynof_amd.batch import JKGroup

    lo, hi = (int(x) for x in args.batch_range.split(","))
    ws = synth.batch_workloads(64, lo=lo, hi=hi)
    dev = torch.device("cuda", 0)
    log(f"batch64: {len(ws)} instances, nbf {min(w.nbf for w in ws)}..{max(w.nbf for w in ws)}")
    engs, Cs, Hs = [], [], []
    for w in ws:
        e = JKEngine(0)
        bt = synth.device_b_packed(w.nbf, w.naux, w.seed_b, w.lam, JKEngine.packed_ld(w.nbf, w.naux))
        e.upload_b_packed(bt, w.nbf, w.naux)
        del bt
        engs.append(e)
        Cs.append(torch.from_numpy(synth.make_C(w.nbf, w.seed_c)).to(dev))
        Hs.append(torch.from_numpy(synth.make_H(w.nbf, w.seed_h)).to(dev))
    torch.cuda.empty_cache()
    n5s = [w.nbf5 for w in ws]
    flops = sum(workcount.build_flops(w.nbf, w.naux, w.nbf5) for w in ws)
    grp = JKGroup(engs, nbf5s=n5s)  # the tuned tiles, for the sequential passes too
    outs = grp.alloc_outputs(n5s)
    stream = torch.cuda.current_stream(dev).cuda_stream
    res = {}
    for form in ("fp64", "split"):
        def seq():
            for e, c, h, o, n5 in zip(engs, Cs, Hs, outs, n5s):
                e.build(c, n5, H=h, precision=form, out=o, stream=stream)

        def grouped(graph):
            grp.build(Cs, n5s, Hs, precision=form, outs=outs, use_graph=graph)

        for name, fn in (("sequential", seq), ("grouped", lambda: grouped(False)),
                         ("grouped_graph", lambda: grouped(True))):
            for _ in range(max(1, args.warmup)):
                fn()
            torch.cuda.synchronize()
            reps = 5
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / reps
            res[f"{form}_{name}"] = {"ms_per_pass": round(1e3 * dt, 3), "builds_per_s": round(len(ws) / dt, 2),
                                     "useful_tflops": round(flops / dt / 1e12, 2)}
            log(f"{form} {name}: {1e3 * dt:.2f} ms per 64 builds")
    line = {
        "metric": "batched JK builds/s, 64 uneven instances in one grouped persistent pass",
        "value": res["split_grouped_graph"]["builds_per_s"],
        "unit": "builds/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "bf16x3 (fp32 accumulate)",
        "data": f"synthetic: device-drawn b, seeded C/H; nbf ~ U[{lo},{hi}] (seed 4000)",
        "config": {"workload": f"batch64 nbf~U[{lo},{hi}] naux=ceil8(3.6 nbf) nbf5=nbf",
                   "sizes": [w.nbf for w in ws], "useful_flops_per_pass": flops},
        "results": res,
        "grouped_graph_speedup_vs_sequential": {
            f: round(res[f"{f}_sequential"]["ms_per_pass"] / res[f"{f}_grouped_graph"]["ms_per_pass"], 3)
            for f in ("fp64", "split")},
    }
    if not getattr(args, "no_cpu_baseline", False):
        log("batch64: CPU baseline (oracle on a sample of instances)")
        line["cpu_baseline"] = cpu_baseline_batch(ws, flops)
    print(json.dumps(line), flush=True)


def bench_gemmref(args):
    """Vendor-library GEMM rates on this GPU (hipBLASLt/rocBLAS through torch.matmul),
    as a practical reference for the hand-written cores: large square GEMMs and the
    half-transform's own shape (M = 320 rows of C^T, N = nbf * naux_p, K = 1024),
    bf16 and fp64. Prints one JSON line."""
    import torch

    dev = torch.device("cuda", 0)

    def rate(M, N, K, dtype, reps):
        a = torch.randn(M, K, device=dev, dtype=dtype)
        b = torch.randn(K, N, device=dev, dtype=dtype)
        for _ in range(2):
            a @ b
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            a @ b
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        del a, b
        torch.cuda.empty_cache()
        res = {"M": M, "N": N, "K": K, "dtype": str(dtype).replace("torch.", ""), "ms": round(ms, 3),
               "tflops": round(2.0 * M * N * K / (ms * 1e-3) / 1e12, 1)}
        log(json.dumps(res))
        return res

    rows = [rate(8192, 8192, 8192, torch.bfloat16, 10), rate(16384, 16384, 16384, torch.bfloat16, 5),
            rate(320, 3648000, 1024, torch.bfloat16, 5), rate(8192, 8192, 8192, torch.float64, 5),
            rate(16384, 16384, 16384, torch.float64, 2), rate(320, 3648000, 1024, torch.float64, 2)]
    print(json.dumps({"metric": "vendor GEMM reference rates (torch.matmul: hipBLASLt / rocBLAS)", "unit": "TFLOP/s",
                      "n_gpus": 1, "results": rows}), flush=True)


def bench_serving(args):
    """Continuous-batching serving loop: saturation throughput, then p50/p99
    latency under Poisson arrivals at 50%, 80% and 100% of it."""
    import torch

    from pynof_amd.serving import DeviceExecutor, Server, poisson_requests

    n = args.requests
    lanes = None if args.lanes in ("none", "") else tuple(int(x) for x in args.lanes.split(","))
    alone = {}  # rid -> service time of the request run by itself (upload + 20 builds)

    def run(rate, max_active=64, use_lanes=True):
        rs = poisson_requests(n, rate=rate, seed=5000)
        srv = Server(DeviceExecutor(0, with_jk=True), hbm_budget=200e9, max_active=max_active,
                     lanes=lanes if use_lanes else None)
        for r in rs:
            srv.submit(r)
        t0 = time.perf_counter()
        srv.run(timeout=600)
        wall = time.perf_counter() - t0
        srv.ex.close()
        st = srv.stats()
        st["wall_s"] = round(wall, 3)
        st["requests_per_s"] = round(st["finished"] / wall, 3)
        st["builds_per_s"] = round(20 * st["finished"] / wall, 2)
        if "p50_s" in st:
            st["p99_over_p50"] = round(st["p99_s"] / st["p50_s"], 3)
        if alone:
            # slowdown = latency / the request's own service time alone: what the
            # scheduler adds, with the spread of request sizes divided out
            sd = np.array([r.latency / alone[r.rid] for r in srv.finished])
            st["slowdown_p50"] = round(float(np.percentile(sd, 50)), 3)
            st["slowdown_p99"] = round(float(np.percentile(sd, 99)), 3)
        torch.cuda.empty_cache()
        return st, srv.finished

    log("serving: warmup (8 requests)")
    small = poisson_requests(8, rate=0.0, seed=77)
    srv = Server(DeviceExecutor(0), hbm_budget=200e9)
    for r in small:
        srv.submit(r)
    srv.run(timeout=300)
    log(f"serving: each request alone ({n} requests, one at a time)")
    _, fin = run(0.0, max_active=1, use_lanes=False)
    for r in fin:
        alone[r.rid] = r.done - r.started
    svc = np.array(list(alone.values()))
    isolated = {"p50_s": round(float(np.percentile(svc, 50)), 4), "p99_s": round(float(np.percentile(svc, 99)), 4),
                "p99_over_p50": round(float(np.percentile(svc, 99) / np.percentile(svc, 50)), 3),
                "total_s": round(float(svc.sum()), 3),
                # no request finishes faster than alone, so p99 latency >= p99 of these times; p99 < 2 p50
                # then needs a median latency above half of it: this many times the median service time
                "median_slowdown_needed_for_p99_under_2x_p50": round(
                    float(np.percentile(svc, 99) / 2.0 / np.percentile(svc, 50)), 3)}
    log(f"serving: saturation ({n} requests at t=0)")
    sat, _ = run(0.0)
    sat_rate = sat["requests_per_s"]
    loads = {}
    for frac in (0.5, 0.8, 1.0):
        log(f"serving: Poisson arrivals at {frac:.0%} of saturation ({frac * sat_rate:.2f} req/s)")
        loads[f"{int(frac * 100)}%"], _ = run(frac * sat_rate)
    line = {
        "metric": "serving: orbital-optimisation requests (20 JK builds each), throughput and p50/p99 latency",
        "value": sat_rate,
        "unit": "requests/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "bf16x3 (fp32 accumulate) + f64 refinement of each request's final build",
        "data": "synthetic: Poisson arrivals (seed 5000), nbf lognormal median 150 in [24,800], naux=ceil8(3.6 nbf)",
        "config": {"workload": "serving", "requests": n, "builds_per_request": 20, "hbm_budget_gb": 200,
                   "max_active": 64, "lanes_nbf_bounds": list(lanes) if lanes else None},
        "isolated_service": isolated,
        "saturation": sat,
        "loads": loads,
        "p99_under_2x_p50_at_80pct": bool(loads["80%"].get("p99_over_p50", 99) < 2.0),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
