"""Benchmark: grid-point-steps/s of the tethered flapping-insect workload
(512^3, RK4 with explicit diffusion, insect mask and us regenerated in every
stage) on one MI355X, FP32-mixed, with the FP64 path and the oracle's FP64
C++ OpenMP restatement on the host cores beside it, plus brief runs of the
ensemble and adjoint workloads.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 512] [--no-cpu] [--no-fp64]
    python bench.py --workload ensemble [--steps K] [--warmup W] [--no-cpu]
    python bench.py --workload adjoint [--adj-steps 400] [--fd 2]

The ensemble workload is 48 independent penalized simulations (16 each at
64^3, 96^3 and 128^3, half spheres, half flapping plates, RK2, FP32-mixed)
advanced as batched launches replayed from one captured hipGraph.  The
adjoint workload is d(mean lift over 400 RK4 steps)/d(12 wing kinematics
coefficients) of the flapping insect at 128^3 (checkpoint every 20 steps), in
FP64 and FP32-mixed, with central differences of the FP64 mean lift for the
largest gradient components beside it.

Prints one JSON line (rank 0).  Data are synthetic: the flow starts from rest
and is driven by the flapping wings (quiescent initial condition), with the
fixed time step dt = 4e-5 so both precisions take identical steps.
"""
import argparse
import json
import math
import os

import numpy as np
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-pt·steps/s at 512^3, FP32-mixed vs FP64 (force rel err≤1e-4); % HBM roofline"
HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
B_PT_RK4 = {4: 895.2, 8: 1790.5}   # algorithmic bytes per grid-point-step, RK4 (penalised)
B_PT_RK2 = {4: 423.0, 8: 846.0}    # the same for RK2 (ensemble members)
ENS_METRIC = "ensemble grid-pt·steps/s (48 sims: 16x64^3+16x96^3+16x128^3, RK2, FP32-mixed, one hipGraph); % HBM roofline"


def make_ini(tmpdir, n, nt):
    src = os.path.join(ROOT, "flusi_amd", "configs", "cfg3_insect_512.ini")
    txt = open(src).read()
    for k in ("nx", "ny", "nz"):
        txt = txt.replace(f"{k}=512;", f"{k}={n};")
    txt = txt.replace("nt=999;", f"nt={nt};").replace("itdrag=10;", "itdrag=99999;")
    path = os.path.join(tmpdir, "bench_insect.ini")
    with open(path, "w") as fh:
        fh.write(txt)
    import shutil
    shutil.copy(os.path.join(ROOT, "flusi_amd", "configs", "kin_cfg3.ini"), tmpdir)
    return path


def pass_bytes(n, s, stepper="RK4"):
    """Algorithmic HBM bytes per launch of each pass at an n^3 grid (complex
    field C = 2*(n/2+1)*n*n*s bytes, real field R = n^3*s)."""
    C = 2 * (n // 2 + 1) * n * n * s
    R = n ** 3 * s
    return {
        "ZINV": 8 * C,            # 3 fields in, 5 out
        "YINV": 11 * C,           # 5 in, 6 out
        "XPASS": 9 * C + R,       # 6 in, 3 out, mask (us only inside the mask)
        "YFWD": 6 * C,            # 3 in, 3 out
        # RK4: average of 12/15/18/12 C (stage 2 rebuilds u_n from its two
        # stage-1 registers); RK2: 12 C in both stages
        "ZSTEP": (14.25 if stepper == "RK4" else 12) * C,
    }


def measured_traffic(n, kernel):
    """HBM bytes per launch of a pass from the committed rocprofv3 PMC summary
    (reads from the TCC request-size counters, writes from WRITE_SIZE, both
    checked against the probe kernels of known size); None when not measured
    for this grid.  Lines of the file: `traffic <grid> <PASS> <bytes>`."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_insect_fp32.md")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        f = ln.split()
        if len(f) == 4 and f[0] == "traffic" and f[1] == str(n) and f[2] == kernel:
            return float(f[3])
    return None


def run_gpu(ini, prec, warmup, steps, timers):
    """Warm up, then time `steps` steps; flusi_timers synchronises the solver's
    stream on both sides of the timed region and collects the per-pass HIP
    event times recorded on that stream."""
    import flusi_amd
    s = flusi_amd.Flusi(ini, prec)
    s.init_fields()
    s.advance(warmup)
    s.timers(1 if timers else 0)
    t0 = time.perf_counter()
    s.advance(steps)
    ms = s.timers(-1)
    t1 = time.perf_counter()
    integ = s.write_integrals(s.get_time()[0])
    s.close()
    return t1 - t0, ms, integ


def accuracy_series(ini, steps, every):
    """Forces and kinetic energy every `every` steps over `steps` RK4 steps
    from rest in FP32-mixed and in FP64 (identical fixed steps, the mask
    redrawn at each sample time as write_integrals does); per component the
    largest error over the series, against a scale (`max_err_over_scale`) and
    pointwise where the FP64 value is at least 1 % of that scale
    (`max_rel_err`, null where no sample is).  The force components share one
    scale, the largest force component over the series: the side force of
    this symmetric insect is zero up to rounding, and an error relative to
    its own size would measure nothing."""
    import flusi_amd
    ser = {}
    for prec in (flusi_amd.FP32_MIXED, flusi_amd.FP64):
        s = flusi_amd.Flusi(ini, prec)
        s.init_fields()
        rows = []
        for _ in range(steps // every):
            s.advance(every)
            t = s.get_time()[0]
            o = s.write_integrals(t)
            rows.append([t, o[0], o[1], o[2], o[18]])  # time, Fx, Fy, Fz, E_kin
        s.close()
        ser[prec] = np.array(rows)
        print(f"accuracy series precision {prec}: {len(rows)} samples to t={rows[-1][0]:.4g}", file=sys.stderr,
              flush=True)
    a, b = ser[flusi_amd.FP32_MIXED], ser[flusi_amd.FP64]

    def comp(i, scale):
        d = np.abs(a[:, i] - b[:, i])
        ref = np.abs(b[:, i])
        big = ref >= 1e-2 * scale
        return {"max_abs_err": float(d.max()), "max_err_over_scale": float(d.max() / scale) if scale > 0 else 0.0,
                "max_rel_err": float((d[big] / ref[big]).max()) if big.any() else None,
                "component_max": float(ref.max()), "scale": scale}

    fscale = float(np.abs(b[:, 1:4]).max())
    return {"steps": steps, "every": every, "samples": int(len(b)), "t_end": float(b[-1, 0]),
            "drag_Fx": comp(1, fscale), "side_Fy": comp(2, fscale), "lift_Fz": comp(3, fscale),
            "energy": comp(4, float(np.abs(b[:, 4]).max()))}


def hbm_probe(gib=4, reps=5):
    """Copy (bytes read + written), read-only and write-only bandwidth of
    16-byte non-temporal streams over two `gib` GiB buffers on this device,
    and of a copy that reads 128-byte segments from a strided side (the tile
    passes' pattern), from flusi_amd/libflusi_probe.so, in bytes/s; None when
    the probe library is missing."""
    import ctypes
    path = os.path.join(ROOT, "flusi_amd", "libflusi_probe.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.flusi_probe_hbm.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.flusi_probe_hbm.restype = ctypes.c_int
    out = (ctypes.c_double * 4)()
    rc = lib.flusi_probe_hbm(int(gib) << 30, int(reps), out)
    if rc != 0:
        return {"error": rc}
    return {"copy": out[0], "read": out[1], "write": out[2], "seg_copy": out[3], "buffer_gib": gib}


def host_info():
    """CPU model, logical CPUs usable by this process and RAM of the host."""
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    ram = None
    try:
        for ln in open("/proc/meminfo"):
            if ln.startswith("MemTotal"):
                ram = round(int(ln.split()[1]) / 2 ** 20, 1)
                break
    except OSError:
        pass
    cpus = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu_model": model, "logical_cpus": cpus, "ram_gib": ram}


def cpu_baseline(n, threads):
    """RK4 steps of the same workload by the oracle's C++ OpenMP restatement
    (oracle/cpu/flusi_cpu: own FFT, FP64, the reference's operation order) on
    `threads` host threads, from the same quiescent start with the mask drawn
    at every stage; returns (seconds per step, set-up seconds, steps) or
    (None, error)."""
    exe = os.path.join(ROOT, "oracle", "cpu", "flusi_cpu")
    if not os.path.exists(exe):
        return None, "oracle/cpu/flusi_cpu is not built"
    d = tempfile.mkdtemp()
    ini = make_ini(d, n, 0)
    steps = 1
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([exe, ini, str(steps), str(threads)], capture_output=True, text=True, env=env, timeout=900)
    if r.returncode != 0:
        return None, r.stderr[-400:]
    info = json.loads(r.stdout.strip().splitlines()[-1])
    return (info["s_per_step"], info["setup_s"], steps), None


def ensemble_cpu_baseline(sizes, per_size, workers):
    """One right-hand side of one plate member per grid size with the FP64
    oracle in a child process; the ensemble's CPU rate is its points over
    sum_size(per_size * 2 * t_RHS(size)) (RK2: two right-hand sides a step)."""
    code = f"""
import sys, time, json, tempfile
sys.path.insert(0, {ROOT!r})
from oracle import flusi_oracle as fo
from flusi_amd.ensemble_configs import write_ensemble
out = {{}}
for n in {list(sizes)!r}:
    paths = write_ensemble(tempfile.mkdtemp(), sizes=(n,), per_size=2)
    sim = fo.Simulation(paths[1], workers={workers})
    sim.init_fields()
    t0 = time.perf_counter()
    sim.cal_nlk(0.0, sim.uk)
    out[n] = time.perf_counter() - t0
print(json.dumps(out))
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        return None, r.stderr[-400:]
    t = {int(k): v for k, v in json.loads(r.stdout.strip().splitlines()[-1]).items()}
    pts = sum(per_size * n ** 3 for n in sizes)
    return pts / sum(per_size * 2.0 * t[n] for n in sizes), t


def run_ensemble(args):
    import flusi_amd
    from flusi_amd.ensemble_configs import write_ensemble
    sizes, per_size = (64, 96, 128), 16
    paths = write_ensemble(tempfile.mkdtemp(), sizes=sizes, per_size=per_size,
                           nt=args.warmup + 2 * args.steps + 20)
    pts = sum(per_size * n ** 3 for n in sizes)
    e = flusi_amd.Ensemble(paths, flusi_amd.FP32_MIXED)
    e.init_fields()
    tune = e.autotune(reps=2)
    print(f"ensemble: {len(paths)} members, batches {e.batches()}, autotune {tune}", file=sys.stderr, flush=True)
    e.advance(args.warmup, flusi_amd.ENSEMBLE_GRAPH)  # first step eager, then the capture
    t0 = time.perf_counter()
    e.advance(args.steps, flusi_amd.ENSEMBLE_GRAPH)     # synchronises before returning
    dt = time.perf_counter() - t0
    value = pts * args.steps / dt
    # per-pass device times of the batched launches (eager, event-timed on
    # each batch's stream) for the roofline of the dominant pass
    leads = e.batches(with_lead=True)
    for b in leads:
        e.member(b[4]).timers(1)
    k_eager = max(2, args.steps // 4)
    t1 = time.perf_counter()
    e.advance(k_eager, flusi_amd.ENSEMBLE_BATCHED)
    dt_eager = time.perf_counter() - t1
    per_pass, per_bytes, per_batch = {}, {}, []
    for (nm, nx, ny, nz, lead) in leads:
        ms = e.member(lead).timers(0)
        per_batch.append({"members": nm, "grid": nx, "lead": lead,
                          "ms_per_step": {k: round(v / k_eager, 4) for k, v in ms.items() if v > 0}})
        pb = pass_bytes(nx, 4, "RK2")
        for k, v in ms.items():
            if k in pb:
                per_pass[k] = per_pass.get(k, 0.0) + v
                per_bytes[k] = per_bytes.get(k, 0.0) + pb[k] * nm * 2 * k_eager  # 2 launches a step
    t2 = time.perf_counter()
    e.advance(2, flusi_amd.ENSEMBLE_MEMBERS)
    dt_members = (time.perf_counter() - t2) / 2
    e.close()
    best = max(per_pass, key=lambda k: per_pass[k])
    nlaunch = 2 * k_eager * len(leads)
    ach = per_bytes[best] / (per_pass[best] / 1e3)
    out = {
        "metric": ENS_METRIC, "value": value, "unit": "grid-pt·steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "48-member ensemble: 16x64^3, 16x96^3, 16x128^3, box 2pi, half spheres "
                               "(r~U(0.3,0.8), default_rng(5)), half flapping rectangular plates, eta=1e-3, "
                               "nu=1e-2, RK2, FP32-mixed, batched launches in one captured hipGraph",
                   "members": len(paths), "points": pts, "batches": len(leads)},
        "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": ach / HBM_PEAK, "traffic": None, "kernel": best,
                     "algorithmic_bytes_per_launch": per_bytes[best] / nlaunch,
                     "avg_launch_ms": per_pass[best] / nlaunch},
        "step_roofline_frac": B_PT_RK2[4] * value / HBM_PEAK,
        "batched_eager_gps": pts * k_eager / dt_eager,
        "members_alone_gps": pts / dt_members,
        "pass_ms_per_step": {k: v / k_eager for k, v in per_pass.items()},
        "batch_ms_per_step": per_batch,
        "autotune": tune,
    }
    if not args.no_cpu:
        workers = min(16, os.cpu_count() or 1)
        v, t = ensemble_cpu_baseline(sizes, per_size, workers)
        if v:
            out["cpu_baseline"] = {"value": v, "unit": "grid-pt·steps/s", "cores": workers, "kind": "port",
                                   "sample": "one FP64 right-hand side of one plate member per grid size by the "
                                             f"numpy/scipy oracle (t_RHS {t}); rate = points / "
                                             "sum(16 * 2 * t_RHS)"}
        else:
            out["cpu_baseline"] = {"value": None, "error": t}
    return out


def run_adjoint(args):
    import shutil
    import flusi_amd
    d = tempfile.mkdtemp()
    for f in ("cfg4_insect_adjoint_128.ini", "kin_cfg4.ini"):
        shutil.copy(os.path.join(ROOT, "flusi_amd", "configs", f), d)
    ini = os.path.join(d, "cfg4_insect_adjoint_128.ini")
    K, every = args.adj_steps, 20
    out = {}
    for name, prec in (("fp64", flusi_amd.FP64), ("fp32", flusi_amd.FP32_MIXED)):
        s = flusi_amd.Flusi(ini, prec)
        th = s.adjoint_controls()
        t0 = time.perf_counter()
        J = s.adjoint_forward(th, K, every)
        t1 = time.perf_counter()
        g = s.adjoint_backward()
        t2 = time.perf_counter()
        out[name] = {"J": J, "grad": list(g), "forward_s": t1 - t0, "backward_s": t2 - t1}
        print(f"adjoint {name}: J={J:.10g} forward {t1 - t0:.2f}s backward {t2 - t1:.2f}s", file=sys.stderr, flush=True)
        if name == "fp64" and args.fd > 0:
            h = 1.0e-4
            fd = {}
            for i in np.argsort(-np.abs(g))[:args.fd]:
                tp, tm = th.copy(), th.copy()
                tp[i] += h
                tm[i] -= h
                fd[int(i)] = (s.adjoint_forward(tp, K, every) - s.adjoint_forward(tm, K, every)) / (2 * h)
            out["fd"] = {"h": h, "components": fd,
                         "max_rel_err": max(abs(out["fp64"]["grad"][i] - v) for i, v in fd.items())
                         / max(abs(v) for v in fd.values())}
            print(f"adjoint fd check {out['fd']}", file=sys.stderr, flush=True)
        s.close()
    n = 128
    g64, g32 = np.array(out["fp64"]["grad"]), np.array(out["fp32"]["grad"])
    t64 = out["fp64"]["forward_s"] + out["fp64"]["backward_s"]
    res = {
        "metric": "adjoint gradient: grid-pt·steps/s of forward + backward sweep, 128^3 insect, 12 controls",
        "value": n ** 3 * K / t64, "unit": "grid-pt·steps/s", "n_gpus": 1, "steps": K, "warmup": 0,
        "ms_per_step": 1e3 * t64 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "d(mean lift)/d(12 Fourier coefficients of the wing kinematics), flapping insect "
                               "128^3, RK4 dt=9e-5, eta=1e-4, checkpoint every 20 steps", "grid": [n, n, n],
                   "controls": len(g64), "checkpoint_every": every},
        "roofline": None,
        "J_fp64": out["fp64"]["J"], "J_fp32": out["fp32"]["J"],
        "grad_fp64": out["fp64"]["grad"], "grad_fp32_rel_l2_vs_fp64": float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)),
        "fd_check": out.get("fd"),
        "fp32": {"forward_s": out["fp32"]["forward_s"], "backward_s": out["fp32"]["backward_s"]},
        "fp64": {"forward_s": out["fp64"]["forward_s"], "backward_s": out["fp64"]["backward_s"]},
    }
    return res


def other_forms():
    """The ensemble and adjoint workloads in brief, so that the default run
    records them too: the 48-member ensemble (5 graph-replayed RK2 steps after
    2 warm-up steps) and the 128^3 adjoint gradient (400 RK4 steps forward and
    backward in both precisions, central differences of its two largest
    components).  Their own modes (--workload ensemble|adjoint) report more."""
    sub = argparse.Namespace(steps=5, warmup=2, no_cpu=True, adj_steps=400, fd=2)
    e = run_ensemble(sub)
    a = run_adjoint(sub)
    return {
        "ensemble_48": {
            "metric": e["metric"], "value": e["value"], "unit": e["unit"], "ms_per_step": e["ms_per_step"],
            "steps": e["steps"], "workload": e["config"]["workload"],
            "roofline": {k: e["roofline"][k] for k in ("kernel", "achieved", "frac", "avg_launch_ms")},
            "step_roofline_frac": e["step_roofline_frac"], "batched_eager_gps": e["batched_eager_gps"],
            "members_alone_gps": e["members_alone_gps"]},
        "adjoint_128": {
            "metric": a["metric"], "value": a["value"], "unit": a["unit"], "ms_per_step": a["ms_per_step"],
            "steps": a["steps"], "workload": a["config"]["workload"], "J_fp64": a["J_fp64"], "J_fp32": a["J_fp32"],
            "grad_fp32_rel_l2_vs_fp64": a["grad_fp32_rel_l2_vs_fp64"],
            "fd_max_rel_err": (a["fd_check"] or {}).get("max_rel_err"), "fp64": a["fp64"], "fp32": a["fp32"]},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--workload", choices=["insect", "ensemble", "adjoint"], default="insect")
    ap.add_argument("--adj-steps", type=int, default=400)
    ap.add_argument("--fd", type=int, default=2, help="central-difference checks of the largest gradient components")
    ap.add_argument("--series-steps", type=int, default=1000,
                    help="RK4 steps of the FP32-mixed vs FP64 force/energy series (0: none)")
    ap.add_argument("--series-every", type=int, default=10, help="steps between samples of that series")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the brief ensemble and adjoint runs of the default workload")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run the HBM probe on 1 GiB buffers (counter calibration for HBM bytes)")
    args = ap.parse_args()
    if args.copy_calib:
        # counter calibration: every launch of probe_read_k and probe_write_k
        # streams exactly 1 GiB, of probe_copy_k 1 GiB each way
        print(json.dumps({"probe": hbm_probe(gib=1, reps=1), "bytes_per_launch": 1 << 30}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the solver keeps the whole domain on one GPU and has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        print(json.dumps(run_ensemble(args)), flush=True)
        return
    if args.workload == "adjoint":
        print(json.dumps(run_adjoint(args)), flush=True)
        return
    import flusi_amd
    n = args.grid
    N = n ** 3
    tmp = tempfile.mkdtemp()
    ini = make_ini(tmp, n, args.warmup + args.steps + 10)
    probe = hbm_probe()

    dt32, ms32, integ32 = run_gpu(ini, flusi_amd.FP32_MIXED, args.warmup, args.steps, True)
    gps32 = N * args.steps / dt32
    out = {
        "metric": METRIC, "value": gps32, "unit": "grid-pt·steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt32 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"tethered flapping insect {n}^3, RK4, mask+us per stage, FP32-mixed",
                   "grid": [n, n, n], "stepper": "RK4", "dt": 4e-5, "eta": 1e-4, "nu": 1.4e-3},
    }
    # roofline of the dominant pass (per launch: 4 launches of each pass per RK4 step)
    pb = pass_bytes(n, 4)
    launches = 4 * args.steps
    best = max((k for k in pb), key=lambda k: ms32.get(k, 0.0))
    avg_s = ms32[best] / launches / 1e3
    ach = pb[best] / avg_s
    out["roofline"] = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                       "frac": ach / HBM_PEAK, "traffic": measured_traffic(n, best), "kernel": best,
                       "algorithmic_bytes_per_launch": pb[best], "avg_launch_ms": avg_s * 1e3}
    if probe and "copy" in probe:
        # the practical ceiling beside the nominal peak: a plain device copy
        out["roofline"]["copy_bw"] = probe["copy"] / 1e9
        out["roofline"]["frac_of_copy"] = ach / probe["copy"]
        out["hbm_probe_GBps"] = {k: round(probe[k] / 1e9, 1) for k in ("copy", "read", "write", "seg_copy")}
    out["pass_ms_per_step"] = {k: v / args.steps for k, v in ms32.items() if v > 0}
    out["pass_achieved_GBps"] = {k: round(pb[k] / (ms32[k] / launches / 1e3) / 1e9, 1)
                                 for k in pb if ms32.get(k, 0.0) > 0}
    out["step_roofline_frac"] = B_PT_RK4[4] * gps32 / HBM_PEAK

    if not args.no_fp64:
        dt64, ms64, integ64 = run_gpu(ini, flusi_amd.FP64, args.warmup, args.steps, True)
        out["fp64_value"] = N * args.steps / dt64
        out["fp32_over_fp64"] = gps32 / out["fp64_value"]
        out["fp64_pass_ms_per_step"] = {k: v / args.steps for k, v in ms64.items() if v > 0} if ms64 else None
        out["forces_fp64"] = list(integ64[0:3])
        out["forces_fp32"] = list(integ32[0:3])
        if args.series_steps > 0:
            sini = make_ini(tempfile.mkdtemp(), n, args.series_steps + 10)
            acc = accuracy_series(sini, args.series_steps, args.series_every)
            out["accuracy_vs_fp64"] = acc
            # the headline figure: the worst of drag and lift over the series
            rel = [acc[c]["max_rel_err"] for c in ("drag_Fx", "lift_Fz") if acc[c]["max_rel_err"] is not None]
            out["force_rel_err_vs_fp64"] = max(rel) if rel else None

    if not args.no_forms:
        out["forms"] = other_forms()

    if not args.no_cpu:
        host = host_info()
        threads = max(1, host["logical_cpus"] or 1)
        res, err = cpu_baseline(n, threads)
        if res:
            sps, setup, k = res
            out["cpu_baseline"] = {"value": N / sps, "unit": "grid-pt·steps/s", "cores": threads,
                                   "kind": "port", "host": host,
                                   "sample": f"{k} FP64 RK4 step of the {n}^3 workload (four right-hand sides, the "
                                             f"insect mask drawn at each stage) by the oracle's C++ OpenMP "
                                             f"restatement (oracle/cpu/flusi_cpu.cpp, own FFT) on {threads} threads: "
                                             f"{sps:.2f} s per step after {setup:.1f} s of set-up"}
        else:
            out["cpu_baseline"] = {"value": None, "error": err}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
