"""Benchmark: 3LPT + PLT initial conditions, bcc lattice, 512^3 grid (768^3
padded), box 1000 Mpc/h, z_start = 24, FP32 bulk with FP64 q + Psi, on one
MI355X.  One "step" is the whole hot path for one box: NGenIC white noise,
phi1, the 2LPT/3LPT sources (Orszag 3/2 products) and inverse Laplacians,
the six displacement / velocity components with PLT corrections, their
c2r transforms and the bcc particle readout (2 x 512^3 particles).  Plan-time
work (seed table, PLT tables, cosmology scalars) runs once before timing.

Prints one JSON line.  Extra objects: "roofline" (dominant kernel, HIP-event
timing on the library stream during the timed steps), "cpu_baseline" (the
CPU FP64 oracle on a bounded 256^3 sample of the same configuration),
"fp64" (same workload with FP64 bulk), "accuracy" (FP32 vs FP64 device
runs of the benchmarked particle path: rel-L2 / max error of Psi = pos - q and
v per component over all 2 x 512^3 particles), "accuracy_by_order" (the 256^3
3LPT workload: FP32 vs FP64 error of Psi1, Psi2, Psi3 and the velocity
separately), "batch" (48 independent 2LPT
boxes of 128^3-256^3 replayed as one HIP graph with autotuned tiles) and
"adjoint" (the hand-written adjoint of the 2LPT map under a CIC density loss
at 256^3: timing, FP64 and finite-difference gradient checks).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512]
       python bench.py --calibrate-copy          (copy kernel for counter scaling)
       python bench.py --summarize-pmc DIR       (traffic table from rocprofv3 runs)
       python bench.py --summarize-counters MATCH DIR...  (counter means of one kernel)
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
METRIC = "3LPT IC-gen cells/s at 512^3, FP32 vs FP64 on 1 GPU; FFT HBM GB/s % of peak"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def device_sync():
    from monofonic_amd.lpt import _hip
    hip = _hip()
    hip.hipDeviceSynchronize.restype = C.c_int
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")


def prof_report(ctx):
    from monofonic_amd import _lib as L
    buf = C.create_string_buffer(1 << 20)
    L.call("mfic_prof_report", ctx.h, buf, len(buf))
    return json.loads(buf.value.decode())


def time_steps(ctx, run, steps, warmup, profile):
    from monofonic_amd import _lib as L
    for _ in range(warmup):
        run.execute(blocking=True)
    device_sync()
    L.call("mfic_prof_enable", ctx.h, 1 if profile else 0)
    t0 = time.perf_counter()
    for _ in range(steps):
        run.execute(blocking=False)
    ctx.sync()
    device_sync()
    t1 = time.perf_counter()
    L.call("mfic_prof_enable", ctx.h, 0)
    rep = prof_report(ctx) if profile else {}
    return (t1 - t0) / steps * 1e3, rep


def roofline(rep, steps):
    """dominant kernel by total time: algorithmic bytes per launch / mean
    launch duration, against the MI355X HBM peak (every kernel on this path is
    a transform or pointwise pass; none is a dense contraction)"""
    fft = {k: v for k, v in rep.items() if k.startswith("fft_") or k.startswith("fused_")}
    name = max(rep, key=lambda k: rep[k]["ms_total"])
    v = rep[name]
    avg_ms = v["ms_total"] / v["count"]
    bytes_per = v["bytes_total"] / v["count"]
    gbs = bytes_per / (avg_ms * 1e-3) / 1e9
    tot = sum(x["ms_total"] for x in rep.values())
    fft_tot = sum(x["ms_total"] for x in fft.values())
    fft_bytes = sum(x["bytes_total"] for x in fft.values())
    return {
        "kernel": name, "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(name),
        "avg_launch_ms": round(avg_ms, 4), "bytes_per_launch": bytes_per, "launches_per_step": v["count"] / steps,
        "fft_passes_GBs": round(fft_bytes / (fft_tot * 1e-3) / 1e9, 1),
        "fft_passes_frac": round(fft_bytes / (fft_tot * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "fft_time_fraction": round(fft_tot / tot, 3),
    }


# HIP kernel symbol (prefix) behind each profiled pass, for the rocprofv3
# counter summary lookup
SYMBOLS = [
    ("fused_src3_", "mfic::ProdSrc3>"),
    ("fused_src2q_", "mfic::ProdSrc2Q>"),
    ("fused_src2_", "mfic::ProdSrc2>"),
    ("fft_c2rC_n", "k_rows_c2r_pruned<"),
    ("fft_r2cC_n", "k_rows_r2c_prod<"),
    ("fft_c2rA_", "LdOperand<"),
    ("fft_c2rB_", "false, mfic::LdAx1<"),
    ("fft_r2cB_", "true, mfic::LdAx1<"),
    ("fft_r2cA_", "LdPadCol<"),
    ("fft_rows_c2r_", "k_rows_c2r<"),
    ("fft_rows_r2c_", "k_rows_r2c<"),
    ("readout3", "k_readout3<"),
    ("assemble_all", "k_assemble_all<"),
]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_3lpt_plt_bcc_512_f32.txt")


def measured_traffic(name):
    """HBM bytes per dispatch of the kernel behind pass `name` from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or
    None when the summary does not cover it"""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    prec = "double" if name.endswith("_f64") else "float"
    sym = next((s for p, s in SYMBOLS if name.startswith(p)), None)
    if sym is None or name.startswith("fft_r2cC_n"):
        return None  # product-row launches of every width share one kernel symbol
    hits = []
    for line in open(TRAFFIC_FILE):
        if line.startswith("#") or not line.strip():
            continue
        kname, _, fetch, write = line.rstrip("\n").split("\t")[:4]
        if sym in kname and f"<{prec}" in kname:
            hits.append(float(fetch) + float(write))
    return round(sum(hits) / len(hits)) if len(hits) == 1 else None


def kernel_table(rep, steps):
    out = {}
    for k, v in sorted(rep.items(), key=lambda kv: -kv[1]["ms_total"]):
        ms = v["ms_total"] / steps
        out[k] = {"ms_per_step": round(ms, 3), "launches_per_step": v["count"] / steps,
                  "GBs": round(v["bytes_total"] / (v["ms_total"] * 1e-3) / 1e9, 1) if v["ms_total"] > 0 else None}
    return out


def cpu_baseline(n_sample=256):
    """CPU FP64 oracle (C + OpenMP, own FFT) on the same configuration at
    n_sample^3 with the same cell size; PLT table build excluded as on the GPU."""
    from oracle import oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    box = 1000.0 * n_sample / 512
    cfg = O.ic_cfg(n_sample, box, 3, 12345, lattice=1, plt=1)
    oc = O.baseline_cosmology(24.0)
    _, _, _, st = O.ic_run(cfg, oc, particles=True)
    t = st[4] - st[2]
    return {"value": round(n_sample ** 3 / t, 1), "unit": "cells/s", "cores": cores, "kind": "port",
            "cpu_model": model, "nproc": os.cpu_count(), "cpus_available": avail,
            "sample": f"oracle (C/OpenMP FP64, own FFT, reference schedule of 33 Orszag products) 3LPT+PLT bcc "
                      f"{n_sample}^3 box {box:g} Mpc/h z=24 (same cell size as 512^3/1000); PLT table build "
                      f"excluded; {t:.2f} s",
            "seconds": round(t, 3)}


def batch_bench(reps=3):
    """48 independent 2LPT FP32 boxes (16 x 128^3, 16 x 192^3, 16 x 256^3,
    seeds 12345+i, box sizes from mt19937_64(20261202)), tiles autotuned per
    size, replayed as one captured HIP graph; compared with the same boxes
    enqueued one after another and with one host-synchronised run per box"""
    from monofonic_amd import F32, Context
    from monofonic_amd.batch import Batch, batch_boxes
    ctx = Context(0)
    t0 = time.perf_counter()
    b = Batch(ctx, batch_boxes(), lpt_order=2, prec=F32, zstart=24.0, tune=True)
    b.execute()
    plan_s = time.perf_counter() - t0

    def timed(fn):
        device_sync()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.sync()
        device_sync()
        return (time.perf_counter() - t) / reps

    seq = timed(lambda: b.execute(blocking=False))
    per_box = timed(lambda: [r.execute(blocking=True) for r in b.runs])
    b.capture()
    b.launch()
    graph = timed(lambda: b.launch(blocking=False))
    cells = b.cells
    out = {"workload": "48 boxes 2LPT FP32 fields (16x128^3, 16x192^3, 16x256^3), seeds 12345+i, "
                       "L_i = 100 + 700 U_i (mt19937_64(20261202)), z_start 24",
           "cells": cells, "value": round(cells / graph, 1), "unit": "cells/s",
           "graph_ms": round(graph * 1e3, 3), "sequential_ms": round(seq * 1e3, 3),
           "sequential_value": round(cells / seq, 1), "synchronised_per_box_ms": round(per_box * 1e3, 3),
           "synchronised_per_box_value": round(cells / per_box, 1),
           "graph_speedup_vs_synchronised": round(per_box / graph, 3),
           "autotuned_ms_per_box": {str(k): round(v, 3) for k, v in b.tuned.items()},
           "plan_and_tune_s": round(plan_s, 2)}
    b.close()
    ctx.close()
    return out


def adjoint_bench(n=256, box=500.0, reps=3, fd_cells=64):
    """Differentiable 2LPT at 256^3, box 500 Mpc/h, z_start 24: loss =
    sum_cells (CIC density of the displaced particles - target)^2 with the
    target from seed 54321 through the same map, noise from seed 12345.
    Times one FP32 loss forward + backward (hand-written adjoint of the map,
    torch autograd for the CIC), and checks the gradient against an FP64
    reference gradient and against FP64 central differences (eps 1e-3) on 64
    cells drawn with mt19937_64(7)."""
    import numpy as np
    import torch

    from monofonic_amd import F32, F64, Context, Grid, lpt
    from monofonic_amd.adjoint import LPT2, cic_density, lpt2_map
    from monofonic_amd.batch import MT19937_64
    ctx = Context(0)

    def real_noise(seed):
        g = Grid(ctx, (n,) * 3, (box,) * 3, prec=F64)
        lpt.ngenic_fill(g, seed)
        g.FourierTransformBackward()
        x = g.download_r().copy()
        g.close()
        return torch.tensor(x, device="cuda")

    d64 = real_noise(12345)
    plans = {F32: LPT2(ctx, n, box, prec=F32), F64: LPT2(ctx, n, box, prec=F64)}
    with torch.no_grad():
        tgt64 = cic_density(plans[F64].forward(real_noise(54321)), box)
    tgt32 = tgt64.float()

    def loss_of(d, prec):
        tgt = tgt64 if prec == F64 else tgt32
        return ((cic_density(lpt2_map(d, plans[prec]), box) - tgt) ** 2).sum()

    def grad_of(prec):
        d = (d64 if prec == F64 else d64.float()).clone().requires_grad_(True)
        loss = loss_of(d, prec)
        loss.backward()
        return loss.item(), d.grad.double()

    grad_of(F32)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        loss32, g32 = grad_of(F32)
    torch.cuda.synchronize()
    ms_loss = (time.perf_counter() - t) / reps * 1e3
    x32 = d64.float()
    cot = torch.ones((3, n, n, n), dtype=torch.float32, device="cuda")
    plans[F32].forward(x32)
    plans[F32].adjoint(x32, cot)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        plans[F32].forward(x32)
    torch.cuda.synchronize()
    ms_fwd = (time.perf_counter() - t) / reps * 1e3
    t = time.perf_counter()
    for _ in range(reps):
        plans[F32].adjoint(x32, cot)
    torch.cuda.synchronize()
    ms_adj = (time.perf_counter() - t) / reps * 1e3
    loss64, g64 = grad_of(F64)
    rel32 = (torch.linalg.norm(g32 - g64) / torch.linalg.norm(g64)).item()
    g = MT19937_64(7)
    idx = [g() % n ** 3 for _ in range(fd_cells)]
    rms = g64.square().mean().sqrt().item()
    eps = 1e-3
    errs = []
    with torch.no_grad():
        for c in idx:
            e = torch.zeros(n ** 3, dtype=torch.float64, device="cuda")
            e[c] = eps
            e = e.reshape(n, n, n)
            fd = (loss_of(d64 + e, F64) - loss_of(d64 - e, F64)).item() / (2 * eps)
            errs.append(abs(fd - g64.reshape(-1)[c].item()) / rms)
    for p in plans.values():
        p.close()
    ctx.close()
    return {"workload": f"2LPT {n}^3 box {box:g} Mpc/h z_start 24, CIC loss vs target (seed 54321), noise seed 12345",
            "fp32_loss_forward_backward_ms": round(ms_loss, 3), "map_forward_ms": round(ms_fwd, 3),
            "map_adjoint_ms": round(ms_adj, 3), "map_cells_per_s": round(n ** 3 / (ms_fwd * 1e-3), 1),
            "loss_fp32": loss32, "loss_fp64": loss64, "grad_fp32_vs_fp64_rel_l2": rel32,
            "fd_cells": fd_cells, "fd_eps": eps, "fd_max_err_over_rms_grad": max(errs),
            "fd_median_err_over_rms_grad": float(np.median(errs))}


class _DeviceArray:
    """__cuda_array_interface__ view of a library-owned device buffer, so
    torch can read it in place"""

    def __init__(self, ptr, count, typestr):
        self.__cuda_array_interface__ = {"shape": (count,), "typestr": typestr, "data": (ptr, False), "version": 3}


def accuracy(ctx, n, box):
    """The benchmarked path itself (bcc particles, FP64 q+Psi) in FP32 and in
    FP64 on the device, same seed: rel-L2 and max error of each displacement
    component Psi = pos - q and each velocity component of all 2 n^3
    particles, computed on the device against the FP64 run"""
    import torch
    from monofonic_amd import F32, F64, _lib as L
    from monofonic_amd.lpt import ICConfig, ICRun
    runs, arr = {}, {}
    for prec in (F64, F32):
        cfg = ICConfig(n=n, box=box, zstart=24.0, lpt_order=3, lattice=L.BCC, plt=True, prec=prec, seed=12345,
                       particles=True)
        r = ICRun(ctx, cfg)
        r.execute()
        pos, vel, npart = r.particles_device()
        arr[prec] = [torch.as_tensor(_DeviceArray(p, 3 * npart, "<f8"), device="cuda").view(npart, 3)
                     for p in (pos, vel)]
        runs[prec] = r
    torch.cuda.synchronize()
    res = {}
    n3 = n ** 3
    ar = torch.arange(n, device="cuda", dtype=torch.float64)
    for c, name in enumerate(("dx", "dy", "dz", "vx", "vy", "vz")):
        d = c % 3
        num = den = 0.0
        dmax = bmax = 0.0
        for s in range(2):
            sl = slice(s * n3, (s + 1) * n3)
            a = arr[F32][c // 3][sl, d]
            b = arr[F64][c // 3][sl, d]
            if c < 3:
                idx = ar.view(-1, 1, 1) if d == 0 else (ar.view(1, -1, 1) if d == 1 else ar.view(1, 1, -1))
                q = ((idx + 0.5 * s) / n).expand(n, n, n).reshape(-1)
                a, b = a - q, b - q
            diff = a - b
            num += float(torch.sum(diff * diff))
            den += float(torch.sum(b * b))
            dmax = max(dmax, float(diff.abs().max()))
            bmax = max(bmax, float(b.abs().max()))
        res[name] = {"rel_l2": (num / den) ** 0.5, "max_abs_rel": dmax / bmax}
    for r in runs.values():
        r.close()
    return res


def accuracy_by_order(ctx, n=256, box=500.0, z=24.0):
    """The 256^3 3LPT workload (Orszag padding to 384^3, sc, no PLT) in FP32
    and FP64 on the device, same seed: rel-L2 and max error of each order's
    displacement (Psi1 = g1 grad phi1, Psi2 = g2 grad phi2, Psi3 = g3 grad
    phi3 + g3c curl A3) and of the velocity sum (ic_generator.cc:939-1047),
    against the FP64 run; the fields are formed on the host from the
    downloaded potentials"""
    import numpy as np
    from monofonic_amd import F32, F64, lpt
    from monofonic_amd.lpt import ICConfig, ICRun
    coef = lpt.lpt_coefficients(lpt.baseline_cosmology(z), n, box, 3)
    kf = 2 * np.pi / box
    ki = np.fft.fftfreq(n, 1.0 / n)
    ki[n // 2] = 0.0                       # gradient is 0 at the Nyquist index (grid_fft.hh:376-385)
    kz = np.arange(n // 2 + 1, dtype=np.float64)
    kz[n // 2] = 0.0
    kvec = (ki[:, None, None] * kf, ki[None, :, None] * kf, kz[None, None, :] * kf)

    def grad(f, d):
        return np.fft.irfftn(1j * kvec[d] * f, s=(n, n, n), axes=(0, 1, 2), norm="ortho")

    psi = {}
    for prec in (F64, F32):
        r = ICRun(ctx, ICConfig(n=n, box=box, zstart=z, lpt_order=3, prec=prec, seed=12345))
        r.execute()
        pot = [r.potential(w).download_k() for w in range(6)]
        r.close()
        out = {}
        for d in range(3):
            dp, dpp = (d + 1) % 3, (d + 2) % 3
            out.setdefault("psi1", []).append(coef.g1 * grad(pot[0], d))
            out.setdefault("psi2", []).append(coef.g2 * grad(pot[1], d))
            out.setdefault("psi3", []).append(coef.g3 * grad(pot[2], d) +
                                              coef.g3c * (grad(pot[3 + dpp], dp) - grad(pot[3 + dp], dpp)))
        out["v"] = [coef.vfac1 * a + coef.vfac2 * b + coef.vfac3 * c
                    for a, b, c in zip(out["psi1"], out["psi2"], out["psi3"])]
        psi[prec] = {k: np.stack(v) for k, v in out.items()}
    res = {}
    for name in ("psi1", "psi2", "psi3", "v"):
        a, b = psi[F32][name], psi[F64][name]
        res[name] = {"rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                     "max_abs_rel": float(np.abs(a - b).max() / np.abs(b).max())}
    return res


def calibrate_copy(n=512, reps=10):
    """plain device-to-device copies of a known size (n^3 FP32 grid): under
    rocprofv3 --pmc they give the FETCH_SIZE / WRITE_SIZE scale factors"""
    from monofonic_amd import F32, Context, Grid
    ctx = Context(0)
    a = Grid(ctx, (n,) * 3, (1.0,) * 3, prec=F32)
    b = Grid(ctx, (n,) * 3, (1.0,) * 3, prec=F32)
    a.zero()
    for _ in range(reps):
        b.copy_from(a)
    ctx.sync()
    nbytes = n * n * (n + 2) * 4
    print(json.dumps({"copy_bytes": nbytes, "copies": reps}), flush=True)
    a.close()
    b.close()
    ctx.close()


def _pmc_rows(d):
    import csv
    import glob
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    return list(csv.DictReader(open(files[0])))


def summarize_pmc(d):
    """d holds rocprofv3 --pmc runs in fetch/, write/ (the workload) and
    copy_fetch/, copy_write/ (calibrate_copy); prints the per-kernel table
    (KiB counters scaled to bytes by the copy kernel's known size)"""
    import collections
    nbytes = 512 * 512 * 514 * 4

    def per_kernel(rows, counter):
        acc, cnt = collections.defaultdict(float), collections.Counter()
        for r in rows:
            if r["Counter_Name"] == counter:
                acc[r["Kernel_Name"]] += float(r["Counter_Value"])
                cnt[r["Kernel_Name"]] += 1
        return {k: (acc[k] / cnt[k], cnt[k]) for k in acc}

    def copy_factor(sub, counter):
        pk = per_kernel(_pmc_rows(os.path.join(d, sub)), counter)
        big = [v for k, (v, c) in pk.items() if "copyBuffer" in k and v * 1024 > 0.25 * nbytes]
        return nbytes / (sum(big) / len(big) * 1024)

    ff, fw = copy_factor("copy_fetch", "FETCH_SIZE"), copy_factor("copy_write", "WRITE_SIZE")
    fetch = per_kernel(_pmc_rows(os.path.join(d, "fetch")), "FETCH_SIZE")
    write = per_kernel(_pmc_rows(os.path.join(d, "write")), "WRITE_SIZE")
    print(f"# HBM traffic per dispatch from rocprofv3 FETCH_SIZE / WRITE_SIZE (KiB), scaled by a plain copy of "
          f"{nbytes} bytes: FETCH x {ff:.4f}, WRITE x {fw:.4f}")
    print("# kernel\tdispatches\tfetch_bytes\twrite_bytes")
    for k in sorted(fetch, key=lambda k: -fetch[k][0] * fetch[k][1]):
        f, c = fetch[k]
        w = write.get(k, (0.0, 0))[0]
        print(f"{k}\t{c}\t{f * 1024 * ff:.0f}\t{w * 1024 * fw:.0f}")


def summarize_counters(dirs, match):
    """mean per dispatch of every counter of rocprofv3 --pmc runs (one
    directory per pass) for the kernels whose name contains `match`"""
    import collections
    acc, cnt = collections.defaultdict(float), collections.Counter()
    names = set()
    for d in dirs:
        for r in _pmc_rows(d):
            if match in r["Kernel_Name"]:
                names.add(r["Kernel_Name"])
                acc[r["Counter_Name"]] += float(r["Counter_Value"])
                cnt[r["Counter_Name"]] += 1
    for k in sorted(names):
        print(f"# kernel: {k}")
    print("# counter\tmean per dispatch\tdispatches")
    for c in sorted(acc):
        print(f"{c}\t{acc[c] / cnt[c]:.6g}\t{cnt[c]}")


def main():
    if "--calibrate-copy" in sys.argv:
        return calibrate_copy()
    if "--summarize-pmc" in sys.argv:
        return summarize_pmc(sys.argv[sys.argv.index("--summarize-pmc") + 1])
    if "--summarize-counters" in sys.argv:
        i = sys.argv.index("--summarize-counters")
        return summarize_counters(sys.argv[i + 2:], sys.argv[i + 1])
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--plans", default="",
                    help="fixed radix plans instead of autotuning: LENGTH:PASS:VARIANT[,...] (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); every box of this workload fits one "
              "MI355X and multi-GPU runs are out of scope for this build.", file=sys.stderr)
        sys.exit(2)

    from monofonic_amd import BCC, F32, F64, Context
    from monofonic_amd.lpt import ICConfig, ICRun

    n, box = args.n, 1000.0 * args.n / 512
    ctx = Context(0)
    cfg = ICConfig(n=n, box=box, zstart=24.0, lpt_order=3, lattice=BCC, plt=True, prec=F32, seed=12345,
                   particles=True)
    if args.plans:
        from monofonic_amd.batch import set_plan
        for item in args.plans.split(","):
            length, pass_kind, variant = (int(x) for x in item.split(":"))
            set_plan(ctx, length, pass_kind, F32, variant)
        args.no_tune = True
    if not args.no_tune:
        from monofonic_amd.batch import autotune_config, chosen_plans, chosen_tiles
        from monofonic_amd.lpt import baseline_cosmology
        log(f"autotune radix plans and tiles for the {n}^3 3LPT+PLT bcc FP32 run (plan time)")
        tuned_ms = autotune_config(ctx, cfg, baseline_cosmology(24.0))
        plans = {f"L{k[0]}_pass{k[1]}": v for k, v in chosen_plans(ctx, n, F32).items()}
        tiles = {f"L{k[0]}_kind{k[1]}": v for k, v in chosen_tiles(ctx, n, F32).items()}
        log(f"tuned: {tuned_ms:.2f} ms per execution, non-default plans {plans}, tiles {tiles}")
    log(f"plan: 3LPT+PLT bcc n={n} box={box} FP32")
    run = ICRun(ctx, cfg)
    log(f"timing {args.steps} steps after {args.warmup} warmup")
    ms, rep = time_steps(ctx, run, args.steps, args.warmup, profile=True)
    run.close()
    log(f"FP32: {ms:.2f} ms/step")
    line = {
        "metric": METRIC, "value": round(n ** 3 / (ms * 1e-3), 1), "unit": "cells/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (NGenIC white noise, seed 12345; Eisenstein-Hu P(k), sigma_8=0.81)",
        "config": {"workload": f"3LPT+PLT bcc {n}^3 ({3 * n // 2}^3 padded), box {box:g} Mpc/h, z_start 24, "
                               "FP32 bulk, FP64 q+Psi", "n": n, "padded": 3 * n // 2, "box": box, "zstart": 24.0,
                   "lpt_order": 3, "lattice": "bcc", "plt": True, "particles": 2 * n ** 3},
        "autotune": None if args.no_tune else {"tuned_ms": round(tuned_ms, 3), "plans": plans, "tiles": tiles},
        "roofline": roofline(rep, args.steps),
        "kernels": kernel_table(rep, args.steps),
    }
    if not args.no_fp64:
        log("FP64 form of the same workload")
        cfg64 = ICConfig(n=n, box=box, zstart=24.0, lpt_order=3, lattice=BCC, plt=True, prec=F64, seed=12345,
                         particles=True)
        tune64 = None
        if not args.no_tune:  # the FP64 plans and tiles are keys of their own
            from monofonic_amd.batch import autotune_config, chosen_plans, chosen_tiles
            from monofonic_amd.lpt import baseline_cosmology
            log("autotune radix plans and tiles for the FP64 run (plan time)")
            t64 = autotune_config(ctx, cfg64, baseline_cosmology(24.0))
            tune64 = {"tuned_ms": round(t64, 3),
                      "plans": {f"L{k[0]}_pass{k[1]}": v for k, v in chosen_plans(ctx, n, F64).items()},
                      "tiles": {f"L{k[0]}_kind{k[1]}": v for k, v in chosen_tiles(ctx, n, F64).items()}}
        run64 = ICRun(ctx, cfg64)
        ms64, rep64 = time_steps(ctx, run64, max(2, args.steps // 2), 1, profile=True)
        run64.close()
        line["fp64"] = {"ms_per_step": round(ms64, 3), "value": round(n ** 3 / (ms64 * 1e-3), 1),
                        "fp32_speedup": round(ms64 / ms, 3), "autotune": tune64,
                        "roofline": roofline(rep64, max(2, args.steps // 2))}
        log(f"FP64: {ms64:.2f} ms/step")
    if not args.no_accuracy:
        log("accuracy: FP32 vs FP64 device particles (Psi = pos - q, v)")
        line["accuracy"] = {"reference": "FP64 device run (same seed, same schedule)", **accuracy(ctx, n, box)}
        log("accuracy by LPT order: 256^3 3LPT FP32 vs FP64")
        line["accuracy_by_order"] = {"workload": "3LPT sc 256^3 (384^3 padded), box 500 Mpc/h, z_start 24, seed 12345",
                                     "reference": "FP64 device run", **accuracy_by_order(ctx)}
    ctx.close()
    if not args.no_batch:
        log("batched boxes: 48 x 2LPT FP32 as one HIP graph")
        line["batch"] = batch_bench()
    if not args.no_adjoint:
        log("differentiable 2LPT: FP32 loss gradient at 256^3, FP64 and finite-difference checks")
        line["adjoint"] = adjoint_bench()
    if not args.no_cpu:
        log("CPU baseline (oracle) on a 256^3 sample")
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
