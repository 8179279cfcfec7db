"""Benchmark of the fused ensemble propagate on one MI355X.

Workload (default `c2`): the standard map at K in {0.97, 2.0, 6.0}, a 1024 x 1024 grid of cell centres
(2^20 orbits) per K, 10^4 steps, k = 2, tau = 1, with MEGNO + FLI + OFLI + SALI on one tangent frame and
QRSpectrum(HouseholderQR()) on a second frame, all FP64, fused in one kernel pass. One timed step is one
full propagate of one K's grid (frames formed on the device, states initialised, 10^4 steps); the steps
cycle through the three K values. Inputs are resident in HBM before the timed region; results stay there.

`--workload c1`: the reference's CPU-runnable case (K = 6, 4096 starts, 10^5 steps, QRSpectrum(MGS())).

`--workload c3`: the coupled standard-map chain, N in {8, 16, 32, 64, 128} sites x 2048 orbits each, ring,
K = 1.5, eps = 0.1, per-orbit step counts uniform on [2000, 20000] (multiples of tau = 10), k = 16,
QRSpectrum(HouseholderQR()); one timed step propagates all five ensembles (longest orbits launched first).

`--workload c5`: surrogate triage on the standard map with per-orbit K ~ U[0.5, 6]. The network is trained on
1M orbits x 10^5 steps (prefix features + full-method labels), checked on 262144 held-out orbits, and then
4M orbits are labelled twice: by the plain method (QR spectrum to N/10 and N for every orbit) and by the
triage (256-step prefix, network, full method only where |logit| is below a margin chosen on the held-out
set). `value` is the triaged rate in orbit-steps/s (4M x 10^5 over the triage wall time); the line also
carries the speedup and the misclassified fraction against the plain labels.

`--workload rec`: the recorder's device path (grid records) on the reference's cost ensemble: 2048 standard-map
orbits x 2e5 steps per method (L1, L3, L4, L4b, L5, L5b), lifted, tau = 1, the study grid.

Prints one JSON line with the metric, the roofline of the orbit kernel (HIP events on the launch stream)
and the CPU baseline (the oracle, glibc math, on the host's cores, on a bounded sample of the workload).
"""
import argparse
import math
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "orbit-steps/s at equal λ error vs Double64 oracle; surrogate-triage speedup"
FP64_PEAK_TFLOPS = 78.6
# Algorithmic FP64 operations per orbit-step (fma = 2, add/mul/div/sqrt = 1, each sin/cos/log = 20):
# c2 = map step and two 2x2 frame products (~70) + column norms and logs (~55) + MEGNO (~55) + FLI (~10)
#      + OFLI on the map displacement (~145) + SALI (~30) + Householder QR spectrum (~150), about 500;
# c1 = map step and one frame product (~57) + MGS, G_w, orthogonality loss and two logs (~117), about 174.
FLOPS_PER_ORBIT_STEP = {"c2": 500.0, "c1": 174.0}
# HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch), scaled by what the
# counters read for a 1 GiB device copy (FETCH x 2.0, WRITE x 1.0): for c2 the two passes of the split chain
# (indicator pass 24792.9 / 590963.3 KiB, QR pass 24693.1 / 164881.8 KiB); summaries under profiles/.
TRAFFIC_BYTES_PER_LAUNCH = {"c2": (2 * (24792.88 + 24693.06) + 590963.25 + 164881.81) * 1024,
                            # the ind workload's bk launch at 2^20 orbits (58419373.6 / 226179433.0 KiB)
                            "ind_bk": (2 * 58419373.56 + 226179433.0) * 1024,
                            # the c4 launch (5.2041e8 / 1.7745e9 KiB): the Gauss kernel's register spills reach HBM
                            "c4": (2 * 5.2041e8 + 1.7745e9) * 1024}
# Mean Gauss fixed-point iterations per step on the c4 workload, measured with the oracle on its sample
# (used when the CPU leg is skipped).
C4_ITERATIONS_DEFAULT = 6.3


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_cores():
    """The CPU the baseline may use: os.cpu_count(), the affinity set, and the cgroup CPU quota (cpu.max) when one is
    set. The oracle runs one thread per usable core: the affinity set, capped by the quota (a GPU box gives one GPU's
    job a share of its cores; threads beyond the quota would only time-share it)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(math.ceil(int(q) / int(per))))
    except (OSError, ValueError):
        pass
    threads = min(aff, quota) if quota else aff
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(cpu_count=os.cpu_count(), affinity=aff, cgroup_quota=quota, threads=threads, model=model)


def nthreads():
    return host_cores()["threads"]


nthreads_ = nthreads


def workload(name):
    from chaos_amd import MEGNO, FLI, OFLI, SALI, GALI, QRSpectrum, HouseholderQR, MGS, GivensQR
    from chaos_amd import StandardMap, HenonHeilesGauss
    from chaos_amd.inputs import grid_centres, torus_uniform, hh_section_starts
    if name == "c2":
        return dict(name="c2", systems=[StandardMap(K) for K in (0.97, 2.0, 6.0)], x0=grid_centres(1024), N=10_000,
                    tau=1, seed=1, d=2, k=2, red=(MEGNO(replace=False), FLI(replace=False), OFLI(replace=False), SALI()),
                    red2=(QRSpectrum(HouseholderQR()),),
                    desc="standard map K in {0.97,2,6}, 1024^2 grid centres per K, 1e4 steps, k=2, tau=1, "
                         "MEGNO+FLI+OFLI+SALI on frame 1 + QR(Householder) on frame 2, FP64")
    if name == "c1":
        return dict(name="c1", systems=[StandardMap(6.0)], x0=torus_uniform(1, 4096), N=100_000, tau=1, seed=1,
                    d=2, k=2, red=(QRSpectrum(MGS()),), red2=(),
                    desc="standard map K=6, 4096 starts uniform on the torus (seed 1), 1e5 steps, k=2, tau=1, "
                         "QRSpectrum(MGS()), FP64")
    if name == "c4":
        m = 65536
        x0 = np.concatenate([hh_section_starts(1, m, E, orbit0=1 + j * m) for j, E in enumerate((1 / 12, 1 / 8, 1 / 6))])
        return dict(name="c4", systems=[HenonHeilesGauss(h=0.01, tol=1e-15)], x0=x0, N=100_000, tau=10, seed=1,
                    d=4, k=4, red=(MEGNO(replace=False), GALI()), red2=(QRSpectrum(GivensQR()),), N_warm=100,
                    desc="Henon-Heiles, Gauss s=3, h=0.01, tol 1e-15; 65536 starts on the q2=0 section (p2>0) at each "
                         "E in {1/12,1/8,1/6}, one launch of 196608 orbits; 1e5 steps, k=4, tau=10, "
                         "MEGNO(columns 1-2)+GALI_2..4 on frame 1 + QR(Givens) on frame 2, FP64")
    raise SystemExit(f"unknown workload {name}")


def gpu_run(wl, steps, warmup, elt=1):
    import torch
    from chaos_amd import QRSpectrum
    from chaos_amd.tangent import make_chain, _context
    lib, ctx = _context()
    x0 = wl["x0"]
    m, s = x0.shape
    d, k = wl["d"], wl["k"]
    chain, allr = make_chain(wl["red"], wl["red2"])
    nf = 2 if wl["red2"] else 1
    sd0 = wl["systems"][0].descriptor()
    words = lib.cl_state_bytes(ctypes.byref(sd0), ctypes.byref(chain), elt, d, k) // 8
    dev = torch.device("cuda")
    x_master = torch.from_numpy(np.ascontiguousarray(x0.T)).to(dev)
    dx = torch.empty_like(x_master)
    dW = torch.empty((d * k, m), dtype=torch.float64, device=dev)
    dW2 = torch.empty_like(dW) if nf > 1 else None
    dst = torch.empty((words, m), dtype=torch.int64, device=dev)
    dcode = torch.zeros(m, dtype=torch.int8, device=dev)
    dh = torch.zeros(m, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    N, tau = wl["N"], wl["tau"]
    kernel_ms = []

    def one_step(i, timed):
        sd = wl["systems"][i % len(wl["systems"])].descriptor()
        n = N if timed else wl.get("N_warm", N)
        dx.copy_(x_master)
        rc = lib.cl_initial_deviations_device(ctx, wl["seed"], 1, m, d, k, dW.data_ptr(), sh)
        assert rc == 0
        if dW2 is not None:
            dW2.copy_(dW)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        rc = lib.cl_propagate_device(ctx, ctypes.byref(sd), ctypes.byref(chain), elt, m, s, d, k, 0, n, n, tau,
                                     dx.data_ptr(), dW.data_ptr(), dW2.data_ptr() if dW2 is not None else None,
                                     dst.data_ptr(), dcode.data_ptr(), dh.data_ptr(), sh)
        e1.record(stream)
        assert rc == 0, lib.cl_last_error(ctx)
        if timed:
            kernel_ms.append((e0, e1))

    for i in range(warmup):
        one_step(i, False)
        torch.cuda.synchronize()
        log(f"warmup step {i + 1}/{warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        one_step(warmup + i, True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"timed {steps} steps in {wall:.3f} s")
    kms = [a.elapsed_time(b) for a, b in kernel_ms]
    bad = int((dcode != 0).sum().item())
    # final QR spectra of a fixed sample of orbits of the last timed launch (for the lambda check)
    lam = None
    last = (warmup + steps - 1) % len(wl["systems"])
    if any(isinstance(r, QRSpectrum) for r in list(wl["red"]) + list(wl["red2"])) and wl["d"] == 2:
        from chaos_amd.tangent import decode_states, finals_of
        idx = np.arange(0, m, max(1, m // 64))[:64]
        raw = dst[:, torch.from_numpy(idx).to(dev)].cpu().numpy().T.copy().view(np.uint8).reshape(len(idx), -1)
        fin = finals_of(decode_states(raw, allr, d, k), allr, d, k)
        qi = [j for j, (r, _) in enumerate(allr) if isinstance(r, QRSpectrum)][0]
        lam = dict(idx=idx, lam=fin[qi]["lambda"], system=wl["systems"][last])
    return dict(wall=wall, kernel_ms=kms, orbit_steps=float(m) * N, m=m, bad=bad, lam=lam,
                x_final=dx.cpu().numpy().T.copy() if wl["name"] == "c4" else None)


def cpu_baseline(wl, budget_orbits):
    """The oracle (glibc sin/cos/log) on the host's cores over a bounded sample of the workload."""
    from chaos_amd.tangent import make_chain
    from tests import oracle_lib as O
    x0 = wl["x0"]
    stride = max(1, len(x0) // budget_orbits)
    xs = x0[::stride]
    m = len(xs)
    W0 = O.frames(wl["seed"], 1, len(x0), wl["d"], wl["k"])[::stride]
    chain, _ = make_chain(wl["red"], wl["red2"])
    nt = nthreads()
    total, t = 0.0, 0.0
    O.gauss_iterations()
    for sysm in wl["systems"]:
        t0 = time.perf_counter()
        O.run(sysm.descriptor(), chain, xs, W0, wl["N"], wl["tau"], wl["k"],
              W2=W0 if wl["red2"] else None, math=O.MATH_LIBM, nthreads=nt)
        t += time.perf_counter() - t0
        total += float(m) * wl["N"]
        log(f"cpu baseline {sysm!r}: {m} orbits done")
    it, st = O.gauss_iterations()
    ns = len(wl["systems"])
    out = dict(value=total / t, unit="orbit-steps/s", cores=nt, kind="port",
               sample=(f"every {stride}th orbit" if stride > 1 else "every orbit") +
                      f" ({m} orbits x {wl['N']} steps x {ns} system(s)), "
                      f"oracle/oracle.cpp with glibc sin/cos/log, {nt} threads, {t:.1f} s",
               host=host_cores())
    if st:
        out["gauss_iterations_per_step"] = it / st
    return out


def lambda_check(wl, g):
    """Same-orbit check of the GPU's FP64 Lyapunov spectra on a sample of the last launch's orbits: the oracle
    recomputes the Jacobians, frame products and QR spectrum in __float128 along the same FP64 orbit (the AC4
    method, test/tangent-core/tc_ac01_04.jl:66-88) and the largest |lambda_gpu - lambda_quad| is reported."""
    from chaos_amd import QRSpectrum
    from chaos_amd.tangent import make_chain
    from tests import oracle_lib as O
    lam = g.get("lam")
    if lam is None:
        return None
    qr = [r for r in list(wl["red"]) + list(wl["red2"]) if isinstance(r, QRSpectrum)][0]
    chain, _ = make_chain((qr,))
    idx = lam["idx"]
    x0 = wl["x0"][idx]
    W0 = O.frames(wl["seed"], 1, len(wl["x0"]), 2, 2)[idx]
    t0 = time.perf_counter()
    lq = O.same_orbit_quad(lam["system"].K, chain, x0, W0, wl["N"], wl["tau"], 2)
    d = np.abs(lam["lam"] - lq)
    return dict(mode="same orbit: FP64 orbit, __float128 Jacobians, frame products and QR (oracle)",
                orbits=int(len(idx)), system=repr(lam["system"]), max_abs=float(d.max()),
                median_abs=float(np.median(d)), seconds=round(time.perf_counter() - t0, 1))


def flops_per_orbit_step(wl, cpu):
    """Algorithmic FP64 operations per orbit-step (fma = 2, add/mul/div/sqrt = 1, sin/cos/log = 20)."""
    if wl["name"] != "c4":
        return FLOPS_PER_ORBIT_STEP[wl["name"]]
    # Gauss s = 3 on Henon-Heiles: per fixed-point iteration 3 vector fields (~10 each) and 12 compensated
    # 3-term dot products (~62 each), ~780; the stage solve of the tangent ~2300 (12 x 12 LU); the handoff
    # (Givens 4 x 4, MEGNO, GALI ~750) over tau = 10. Iterations per step from the oracle sample when it ran.
    it = cpu.get("gauss_iterations_per_step") if cpu else None
    it = it if it else C4_ITERATIONS_DEFAULT
    return 780.0 * it + 2300.0 + 75.0


def c4_block(a):
    """C4 in the default line: one timed launch of the Henon-Heiles workload (196608 orbits x 1e5 steps) with the FP64
    orbit and one with the compensated (double-double) orbit, their energy drift per energy and the census roofline
    (Gauss iterations per step at the oracle's measured mean)."""
    wl = workload("c4")
    log("c4 block:", wl["desc"])
    g = gpu_run(wl, 1, 1)
    ms = float(np.mean(g["kernel_ms"]))
    wdd = dict(wl)
    wdd["x0"] = np.concatenate([wl["x0"], np.zeros_like(wl["x0"])], axis=1)
    gdd = gpu_run(wdd, 1, 1, elt=3)
    ms_dd = float(np.mean(gdd["kernel_ms"]))
    fpos = flops_per_orbit_step(wl, None)
    achieved = fpos * g["orbit_steps"] / (ms * 1e-3) / 1e12
    return {"value": g["orbit_steps"] / g["wall"], "unit": "orbit-steps/s", "workload": wl["desc"], "kernel_ms": ms,
            "nonfinite_orbits": g["bad"],
            "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP64_PEAK_TFLOPS, "traffic": TRAFFIC_BYTES_PER_LAUNCH["c4"],
                         "traffic_note": "register spills of the Gauss kernel (about 1.2e8 B algorithmic)",
                         "flops_per_orbit_step": fpos,
                         "gauss_iterations_per_step": C4_ITERATIONS_DEFAULT},
            "dd": {"value": gdd["orbit_steps"] / gdd["wall"], "kernel_ms": ms_dd, "over_f64": ms / ms_dd,
                   "nonfinite_orbits": gdd["bad"]},
            "energy_drift": {"f64": c4_energy_drift(wl, g["x_final"]), "dd": c4_energy_drift(wl, gdd["x_final"][:, :4])}}


def c4_energy_drift(wl, xf):
    """|H(z_N) - H(z_0)| of the Henon-Heiles launch per energy: max and rms over its orbits (the last timed launch ran
    all N steps from the workload's starts; non-finite orbits excluded and counted)."""
    from chaos_amd import hh_energy
    x0 = wl["x0"][:, :4]
    m = len(x0) // 3
    out = {}
    for j, E in enumerate(("1/12", "1/8", "1/6")):
        sl = slice(j * m, (j + 1) * m)
        H0 = hh_energy(x0[sl])
        H1 = hh_energy(xf[sl])
        d = np.abs(H1 - H0)
        ok = np.isfinite(d)
        out[f"E={E}"] = {"max": float(d[ok].max()) if ok.any() else None,
                         "rms": float(np.sqrt(np.mean(d[ok] ** 2))) if ok.any() else None,
                         "nonfinite": int((~ok).sum())}
    return out


def ks_two_sample(a, b):
    """Two-sample Kolmogorov-Smirnov statistic D and its 1% critical value 1.63 sqrt((n + m) / (n m))."""
    a, b = np.sort(np.asarray(a)), np.sort(np.asarray(b))
    z = np.concatenate([a, b])
    D = float(np.max(np.abs(np.searchsorted(a, z, side="right") / len(a) - np.searchsorted(b, z, side="right") / len(b))))
    return D, 1.63 * math.sqrt((len(a) + len(b)) / (len(a) * len(b)))


def accuracy_protocol(wl, same_stride=1024, ind_orbits=96):
    """The metric's "at equal lambda error vs Double64" on C2, for the Float64 kernels and the Float32-tangent ones.
    (a) same orbit (the AC4 method, test/tangent-core/tc_ac01_04.jl:66-88): every `same_stride`-th grid orbit of each K,
        the GPU's QR exponents against the oracle's __float128 Jacobians, frame products and QR along the same FP64
        orbit; max, median, and the counts above 1e-6 and 1e-10.
    (b) independent orbit (the reference's AC12 protocol, verify/gpu.jl:414-465): `ind_orbits` grid orbits per K run in
        full by the __float128 oracle (113-bit significand, finer than Double64's 106 bits), the whole chain. Orbits
        whose __float128 MEGNO Y (column 1) is below 4 form the regular class (MEGNO tends to 2 on quasi-periodic
        orbits): per-indicator maximum errors and the counts above 1e-10 and 1e-6 of lambda_1; the rest is the chaotic
        class: two-sample KS of lambda_1 (D against 1.63 sqrt((n+m)/(nm))) and the 4-SE test of the means."""
    from chaos_amd import propagate, QRSpectrum
    from chaos_amd.tangent import make_chain, decode_states, finals_of
    from tests import oracle_lib as O
    nt = nthreads()
    x0_all, N, tau = wl["x0"], wl["N"], wl["tau"]
    W_all = O.frames(wl["seed"], 1, len(x0_all), 2, 2)
    red, red2 = wl["red"], wl["red2"]
    chain, allr = make_chain(red, red2)
    qr_chain, _ = make_chain(tuple(r for r in red2 if isinstance(r, QRSpectrum)))
    qi = len(red)  # the QR spectrum's position in the chain (first reducer of the second frame)
    out = {"same_orbit": {}, "independent": {}, "orbits_same": 0, "orbits_independent": 0}
    t0 = time.perf_counter()
    for sysm in wl["systems"]:
        K = sysm.K
        idx = np.arange(0, len(x0_all), same_stride)
        xs, Ws = x0_all[idx], W_all[idx]
        g = {e: propagate(sysm, xs, N, k=2, tau=tau, seed=wl["seed"], reducers=red, reducers2=red2, W0=Ws,
                          check=False, elt=e) for e in ("f64", "f32tan")}
        lq = O.same_orbit_quad(K, qr_chain, xs, Ws, N, tau, 2, nthreads=nt)
        row = {}
        for e, res in g.items():
            dl = np.abs(res.finals[qi]["lambda"] - lq)
            row[e] = {"max": float(dl.max()), "median": float(np.median(dl)), "above_1e-6": int(np.sum(dl.max(1) > 1e-6)),
                      "above_1e-10": int(np.sum(dl.max(1) > 1e-10))}
        out["same_orbit"][f"K={K}"] = row
        out["orbits_same"] += len(idx)
        # independent orbits: the full chain in __float128
        jdx = np.linspace(0, len(x0_all) - 1, ind_orbits).astype(np.int64)
        xi, Wi = x0_all[jdx], W_all[jdx]
        q = O.run(sysm.descriptor(), chain, xi, Wi, N, tau, 2, W2=Wi, quad=True, nthreads=nt)
        fq = finals_of(decode_states(q["states"], allr, 2, 2), allr, 2, 2)
        reg = fq[0]["Y"][:, 0] < 4.0
        rowi = {"regular": int(reg.sum()), "chaotic": int((~reg).sum())}
        for e in ("f64", "f32tan"):
            r = propagate(sysm, xi, N, k=2, tau=tau, seed=wl["seed"], reducers=red, reducers2=red2, W0=Wi,
                          check=False, elt=e).finals
            d = {}
            if reg.any():
                dl = np.abs(r[qi]["lambda"][reg] - fq[qi]["lambda"][reg])
                d["lambda_max"] = float(dl.max())
                d["lambda1_above_1e-10"] = int(np.sum(dl[:, 0] > 1e-10))
                d["lambda1_above_1e-6"] = int(np.sum(dl[:, 0] > 1e-6))
                d["megno_Y_max"] = float(np.max(np.abs(r[0]["Y"][reg] - fq[0]["Y"][reg])))
                d["megno_Ybar_max"] = float(np.max(np.abs(r[0]["Ybar"][reg] - fq[0]["Ybar"][reg])))
                d["fli_max"] = float(np.max(np.abs(r[1]["fli"][reg] - fq[1]["fli"][reg])))
                d["ofli_max"] = float(np.max(np.abs(r[2]["ofli"][reg] - fq[2]["ofli"][reg])))
                d["sali_max"] = float(np.max(np.abs(r[3]["sali"][reg] - fq[3]["sali"][reg])))
            ch = ~reg
            if ch.sum() >= 10:
                a_, b_ = r[qi]["lambda"][ch, 0], fq[qi]["lambda"][ch, 0]
                D, Dc = ks_two_sample(a_, b_)
                se = math.sqrt((np.var(a_) + np.var(b_)) / len(a_))
                d["chaotic_ks"] = {"D": D, "D_crit_1pct": Dc, "pass": D <= Dc,
                                   "mean_gpu": float(np.mean(a_)), "mean_quad": float(np.mean(b_)),
                                   "means_in_SE": float(abs(np.mean(a_) - np.mean(b_)) / se) if se > 0 else 0.0}
            rowi[e] = d
        out["independent"][f"K={K}"] = rowi
        out["orbits_independent"] += len(jdx)
        log(f"accuracy K={K}: same-orbit {len(idx)} orbits, independent {len(jdx)} ({int(reg.sum())} regular)")
    out["seconds"] = round(time.perf_counter() - t0, 1)
    out["threads"] = nt
    out["reference"] = ("__float128 oracle (113-bit significand) standing in for Double64 (106-bit); the reference's "
                        "own AC12(a): 1-20 of 1024 regular orbits above 1e-6 between its FP64 backends and Double64")
    return out


def tune_block(wl, tau_allowed=None):
    """The project's autotuner (chaos_amd/autotune.py) on the workload's own ensemble and chain: it leaves the fastest
    launch setting on the context for the timed run; the record is its choice, the built-in rule's time and the gain."""
    from chaos_amd import autotune
    out = autotune.tune(wl["systems"][0], wl["x0"], k=wl["k"], tau=wl["tau"], reducers=wl["red"],
                        reducers2=wl["red2"], tau_allowed=tau_allowed, seed=wl["seed"], log=log)
    return {"chosen": out["best"], "builtin_rule": out["default"], "gain": out["gain"], "candidates": len(out["table"]),
            "trial_steps": out["trial_steps"], "seconds": out["seconds"]}


def c1_block(a):
    """C1, the reference's CPU-runnable case, measured beside the headline: the GPU rate, its roofline, and the
    oracle on the host's cores over every orbit of the workload at its full step count; the target is >= 50x."""
    from chaos_amd import autotune
    wl = workload("c1")
    tuned = tune_block(wl)
    g = gpu_run(wl, 10, 2)
    value = g["orbit_steps"] * 10 / g["wall"]
    avg_ms = float(np.mean(g["kernel_ms"]))
    achieved = FLOPS_PER_ORBIT_STEP["c1"] * g["orbit_steps"] / (avg_ms * 1e-3) / 1e12
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(wl, 4096)
    # what a caller who allows tau in {1, 2, 4} would get (the QR spectrum alone is tau-invariant up to rounding)
    tau_opt = tune_block(wl, tau_allowed=(2, 4))
    autotune.set_tuning()
    out = {"workload": wl["desc"], "value": value, "unit": "orbit-steps/s", "kernel_ms": avg_ms, "autotune": tuned,
           "autotune_if_tau_allowed_1_2_4": tau_opt,
           "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS}, "cpu_baseline": cpu, "target_gpu_over_cpu": 50}
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    return out


def lognorm_one_thread():
    """The reference's CPU sanity figure: the column-normalising log-sum chain (LogNorm, test/tangent-core/
    tc_helpers.jl:56-70), k = 2, K = 2 pi, tau = 1, on one thread of the oracle with glibc math, 1e3 steps from starts
    uniform on the torus as the reference's record; its own 1-thread figure is 33.09 ns per orbit-step on a 5.4 GHz
    desktop core (gpu-rtx5060-verify-log.txt:11)."""
    from chaos_amd import StandardMap, _lib as L
    from chaos_amd.inputs import torus_uniform
    from tests import oracle_lib as O
    ch = L.cl_chain()
    ch.n = 1
    ch.r[0].kind = 8
    m, N = 16384, 1000  # the reference's record: 1e5 orbits x 1e3 steps (the orbits stay near their starts)
    x0 = torus_uniform(7, m) * 2 * np.pi
    W0 = O.frames(7, 1, m, 2, 2)
    t0 = time.perf_counter()
    O.run(StandardMap(2 * np.pi).descriptor(), ch, x0, W0, N, 1, 2, math=O.MATH_LIBM, nthreads=1)
    dt = time.perf_counter() - t0
    return {"ns_per_orbit_step": dt / (m * N) * 1e9, "reference_ns_1thread": 33.09,
            "sample": f"{m} orbits x {N} steps, LogNorm k=2, K=2pi, tau=1, one thread, glibc sin/cos/log"}


C3_NS = (8, 16, 32, 64, 128)


def c3_flops(N, k=16, tau=10):
    """Algorithmic FP64 operations per orbit-step of the chain: 6N sin/cos (force and tangent, 20 each), the banded
    product ~8Nk, and per handoff (/tau) the Householder sweep, the Q formation and orth_loss, 2(2N)k^2 each, and
    the k x k window inverse ~k^3/3."""
    return 6 * N * 20 + 8 * N * k + (3 * 2 * (2 * N) * k * k + k ** 3 / 3) / tau


def qr_lambdas(dst, k):
    """lambda_i = total(sums_i) / n of QRState{Float64, k} rows (device int64 words [m][2k + 9])."""
    w = dst.cpu().numpy().view(np.float64)
    n = dst[:, 2 * k].cpu().numpy().astype(np.float64)
    s = w[:, 0:2 * k:2] + w[:, 1:2 * k:2]
    return s / np.where(n > 0, n, 1.0)[:, None]


def c3_mixed(a, lib, ctx, chain, sets, m, k, tau, desc, E=1):
    """C3 as one persistent launch over the five lattices per chunk (cl_propagate_mixed), chunks of --c3-chunk
    steps captured in one graph."""
    import torch
    from chaos_amd import _lib as L
    st = torch.cuda.current_stream()
    lats = (L.cl_lattice * len(sets))()
    for j, S in enumerate(sets):
        lats[j].N, lats[j].m, lats[j].Ns = S["N"], m, S["Ns"].ctypes.data
        lats[j].dx, lats[j].dW, lats[j].dstates = S["dx"].data_ptr(), S["dW"].data_ptr(), S["dst"].data_ptr()
        lats[j].dcode, lats[j].dhstep = S["dc"].data_ptr(), S["dh"].data_ptr()
    warm = [np.full(m, 20, dtype=np.int64) for _ in sets]

    def run(Ns_list, chunk):
        for j, S in enumerate(sets):
            S["dx"].copy_(S["xm"])
            assert lib.cl_initial_deviations_device(ctx, 3, 1, m, 2 * S["N"], k, S["dW"].data_ptr(), st.cuda_stream) == 0
            lats[j].Ns = Ns_list[j].ctypes.data
        rc = lib.cl_propagate_mixed(ctx, ctypes.byref(sets[0]["sd"]), ctypes.byref(chain), E, len(sets), lats, k, tau,
                                    chunk, st.cuda_stream)
        assert rc == 0, lib.cl_last_error(ctx)

    run(warm, 0)
    torch.cuda.synchronize()
    log("warmup done")
    Nsl = [S["Ns"] for S in sets]
    t0 = time.perf_counter()
    evs = []
    for _ in range(a.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        run(Nsl, a.c3_chunk)
        e1.record(st)
        evs.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    span = float(np.mean([x.elapsed_time(y) for x, y in evs]))
    steps_n = {S["N"]: float(S["Ns"].sum()) for S in sets}
    total = sum(steps_n.values())
    flops = sum(c3_flops(N) * steps_n[N] for N in steps_n)
    achieved = flops / (span * 1e-3) / 1e12
    bad = int(sum(int((S["dc"] != 0).sum().item()) for S in sets))
    lam8 = qr_lambdas(sets[0]["dst"], k)  # N = 8: k = 16 = 2N, the full spectrum
    pair = np.max(np.abs(lam8 + lam8[:, ::-1]), axis=1)
    nsteps = sets[0]["Ns"].astype(np.float64)
    out = {
        "metric": METRIC, "value": total * a.steps / wall, "unit": "orbit-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": 1, "ms_per_step": wall * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": desc + f"; one persistent launch per chunk of {a.c3_chunk or 'all'} steps",
                   "orbit_steps_per_step": total, "nonfinite_orbits": bad,
                   "pairing_n8": {"max": float(pair.max()), "max_times_n": float(np.max(pair * nsteps)),
                                  "sum_max": float(np.max(np.abs(lam8.sum(axis=1))))}},
        "roofline": {"bound": "fp64-valu" if a.c3_qr == "sites" else "fp64-valu+mfma", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel_ms": span},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return out


def record_main(a):
    """--workload record: the per-machine throughput record a Julia ChaoslabHIPExt reads (as ext/ChaoslabCUDAExt.jl:42-62
    reads data/gpu-benchmark.toml), measured on the reference's record workload: the standard map at K = 2 pi, k = 2,
    tau = 1, QRSpectrum(MGS()) (the log-sum chain's closest device chain), 1e5 orbits x 1e3 steps for gpu_ns_per_step,
    one wave (64 orbits) x 1e4 steps for gpu_ns_per_step_wave, the oracle on the host's cores for cpu_ns_per_step.
    Writes [<device>."Float64"] to --record-path (TOML) and prints the JSON line."""
    import torch
    from chaos_amd import StandardMap, QRSpectrum, MGS, propagate
    from chaos_amd.tangent import device_name
    from chaos_amd.inputs import torus_uniform
    sysm = StandardMap(2 * math.pi)
    red = (QRSpectrum(MGS()),)
    res = {}
    for key, m, N in (("gpu_ns_per_step", 100_000, 1000), ("gpu_ns_per_step_wave", 64, 10_000)):
        x0 = torus_uniform(5, m) * 2 * math.pi
        propagate(sysm, x0, 100, k=2, tau=1, seed=5, reducers=red, check=False)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        propagate(sysm, x0, N, k=2, tau=1, seed=5, reducers=red, check=False)
        torch.cuda.synchronize()
        res[key] = (time.perf_counter() - t0) / (m * N) * 1e9
    cpu = None
    if not a.no_cpu:
        from tests import oracle_lib as O
        from chaos_amd.tangent import make_chain
        m, N = 16384, 1000
        x0 = torus_uniform(5, m) * 2 * math.pi
        nt = nthreads()
        t0 = time.perf_counter()
        O.run(sysm.descriptor(), make_chain(red)[0], x0, O.frames(5, 1, m, 2, 2), N, 1, 2, math=O.MATH_LIBM, nthreads=nt)
        cpu = (time.perf_counter() - t0) / (m * N) * 1e9
        res["cpu_ns_per_step"] = cpu
        res["cpu_threads"] = nt
    res["gpu_faster"] = bool(cpu is None or res["gpu_ns_per_step"] < cpu)
    dev = device_name()
    if a.record_path:
        lines = [f'["{dev}"."Float64"]'] + [f"{k} = {str(v).lower() if isinstance(v, bool) else v}" for k, v in res.items()]
        with open(a.record_path, "a") as fh:
            fh.write("\n".join(lines) + "\n")
    out = {"metric": "gpu-benchmark record", "value": res["gpu_ns_per_step"], "unit": "ns/orbit-step", "n_gpus": 1,
           "steps": 1, "warmup": 1, "higher_is_better": False, "device": dev, "record": res}
    print(json.dumps(out), flush=True)
    return out


def c3_main(a, emit=True):
    import torch
    from chaos_amd import StandardMapChain, QRSpectrum, HouseholderQR, _lib as L
    from chaos_amd.inputs import chain_starts, chain_step_counts
    from chaos_amd.tangent import make_chain, _context
    from chaos_amd import autotune
    lib, ctx = _context()
    autotune.set_tuning(chain_qr={"auto": 0, "blocked": 1, "sites": 2}[a.c3_qr])
    m, k, tau = 2048, 16, 10
    order = ("blocked HouseholderQR on the FP64 matrix cores (two panels of 8, compact WY trailing update)"
             if a.c3_qr != "sites" else "unblocked HouseholderQR, sums over sites as lane butterflies")
    E = {"f64": 1, "f32tan": 2}[a.c3_elt]
    prec = "FP64" if a.c3_elt == "f64" else "FP64 orbit, Float32 tangents, FP64 factorisation and log sums"
    desc = (f"coupled standard-map chain, ring, K=1.5, eps=0.1, N in {list(C3_NS)} x {m} orbits, steps U[2000,20000] "
            f"per orbit (multiples of tau={tau}), k={k}, QRSpectrum(HouseholderQR()), {prec}; {order}")
    log("workload", desc)
    chain, _ = make_chain((QRSpectrum(HouseholderQR()),))
    sets = []
    for j, N in enumerate(C3_NS):
        sysm = StandardMapChain(N, 1.5, 0.1, "ring")
        x0 = chain_starts(3 + j, m, N)
        Ns = chain_step_counts(3 + j, m, 2000, 20000, tau)
        order = np.argsort(-Ns, kind="stable")
        dx = torch.from_numpy(x0).cuda()
        dW = torch.empty((m, k, 2 * N), dtype=torch.float64, device="cuda")
        dst = torch.zeros((m, 2 * k + 9), dtype=torch.int64, device="cuda")
        dc = torch.zeros(m, dtype=torch.int8, device="cuda")
        dh = torch.zeros(m, dtype=torch.int64, device="cuda")
        sets.append(dict(N=N, sys=sysm, sd=sysm.descriptor(), x0=x0, Ns=Ns, order=order, dx=dx, dW=dW, dst=dst,
                         dc=dc, dh=dh, xm=dx.clone()))
    shared = torch.cuda.Stream() if a.c3_mode == "serial" else None
    for S in sets:
        # the five ensembles run concurrently, one stream each (default priority); "serial": one after another
        S["stream"] = shared if shared is not None else torch.cuda.Stream()

    def setup(S):  # the starts and the seeded frames (part of the timed work)
        sh = S["stream"].cuda_stream
        with torch.cuda.stream(S["stream"]):
            S["dx"].copy_(S["xm"])
        assert lib.cl_initial_deviations_device(ctx, 3, 1, m, 2 * S["N"], k, S["dW"].data_ptr(), sh) == 0

    def run(S, Ns, with_setup=True):
        sh = S["stream"].cuda_stream
        if with_setup:
            setup(S)
        rc = lib.cl_propagate_ragged(ctx, ctypes.byref(S["sd"]), ctypes.byref(chain), E, m, 2 * S["N"], 2 * S["N"], k,
                                     Ns.ctypes.data, S["order"].ctypes.data, tau, S["dx"].data_ptr(),
                                     S["dW"].data_ptr(), S["dst"].data_ptr(), S["dc"].data_ptr(), S["dh"].data_ptr(),
                                     sh)
        assert rc == 0, lib.cl_last_error(ctx)

    if a.c3_mode == "mixed":
        return c3_mixed(a, lib, ctx, chain, sets, m, k, tau, desc, E)
    for S in sets:  # warmup: 20 steps per orbit
        run(S, np.full(m, 20, dtype=np.int64))
    torch.cuda.synchronize()
    log("warmup done")
    per_n = {}
    t0 = time.perf_counter()
    for _ in range(a.steps):
        # the previous step's kernels on the five (non-blocking) streams are done before this step's start event
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        for S in sets:
            S["stream"].wait_event(ev0)
        # every lattice's starts and frames first, so that no frame kernel waits behind another lattice's chain
        # kernel for its compute units
        for S in sets:
            setup(S)
        for S in (sets[::-1] if a.c3_order == "descending" else sets):
            es, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            es.record(S["stream"])  # the lattice's chain kernel alone: from its own start on its stream
            run(S, S["Ns"], with_setup=False)
            e1.record(S["stream"])
            per_n.setdefault(S["N"], []).append((ev0, es, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = {N: float(np.mean([x.elapsed_time(z) for _, x, z in v])) for N, v in per_n.items()}
    # the five launches overlap: the span from the common start event to the last one's end
    span = float(np.mean([max(v[i][0].elapsed_time(v[i][2]) for v in per_n.values()) for i in range(a.steps)]))
    steps_n = {S["N"]: float(S["Ns"].sum()) for S in sets}
    total = sum(steps_n.values())
    flops = sum(c3_flops(N) * steps_n[N] for N in steps_n)
    kern_s = span * 1e-3
    achieved = flops / kern_s / 1e12
    bad = int(sum(int((S["dc"] != 0).sum().item()) for S in sets))
    log("per-N ms", ms)
    cpu = None
    if not a.no_cpu:
        from tests import oracle_lib as O
        nthreads = nthreads_()
        tc, done = 0.0, 0.0
        for j, S in enumerate(sets):
            sel = np.arange(0, m, 32)
            W0 = O.frames(3, 1, m, 2 * S["N"], k)[sel]
            t1 = time.perf_counter()
            O.run_chain(S["sd"], chain, S["x0"][sel], W0, 0, tau, k, Ns=S["Ns"][sel], math=O.MATH_LIBM,
                        mode=O.RED_ROWS, nthreads=nthreads)
            tc += time.perf_counter() - t1
            done += float(S["Ns"][sel].sum())
            log(f"cpu baseline N={S['N']}: {len(sel)} orbits done")
        cpu = dict(value=done / tc, unit="orbit-steps/s", cores=nthreads, kind="port",
                   sample=f"every 32nd orbit of each N's set at its own step count ({int(done)} orbit-steps), "
                          f"oracle/oracle_chain.hpp in the reference's row order with glibc sin/cos/log, "
                          f"{nthreads} threads, {tc:.1f} s")
    out = {
        "metric": METRIC, "value": total * a.steps / wall, "unit": "orbit-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": 1, "ms_per_step": wall * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": desc + ("; the five launches in turn on one stream" if a.c3_mode == "serial" else
                                       "; the five launches on five streams"),
                   "orbit_steps_per_step": total, "kernel_ms_per_N": ms, "nonfinite_orbits": bad},
        "roofline": {"bound": "fp64-valu" if a.c3_qr == "sites" else "fp64-valu+mfma", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel_ms": span,
                     "flops_per_orbit_step": {N: c3_flops(N) for N in C3_NS},
                     "frac_per_N": {N: c3_flops(N) * steps_n[N] / (ms[N] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS
                                    for N in steps_n}},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = out["value"] / cpu["value"]
    if emit:
        print(json.dumps(out), flush=True)
    out["_lambdas"] = {S["N"]: qr_lambdas(S["dst"], k) for S in sets}  # for the caller's same-orbit comparison
    return out


def c5_main(a):
    out = c5_run(a)
    print(json.dumps(out), flush=True)
    return out


def c5_run(a, cpu_leg=True):
    import torch
    from chaos_amd import surrogate as S
    n_train, n_ho, n_tri, N = (int(v) for v in a.c5_sizes.split(","))
    prefix = 256
    desc = (f"standard map, per-orbit K ~ U[0.5,6], starts uniform on the torus; train {n_train} x {N} steps, "
            f"held out {n_ho}, triage {n_tri} orbits; prefix {prefix} steps (MEGNO Y, FLI_B, log alpha_1), "
            f"MLP 768-256-256-256-2 BF16/FP32-acc, margin |logit| calibrated on the held-out set for <= 0.4% "
            f"misclassified; full method QRSpectrum(MGS()) k=2 FP64")
    log("workload", desc)
    t0 = time.perf_counter()
    train = S.run_full(S.make_starts(11, n_train), 11, N, prefix)
    t_data = time.perf_counter() - t0
    log(f"training data: {n_train} orbits x {N} steps in {t_data:.1f} s")
    t0 = time.perf_counter()
    net, info = S.fit(train, epochs=a.c5_epochs, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del train
    ho = S.run_full(S.make_starts(12, n_ho, orbit0=1 + n_train), 12, N, prefix)
    margin = S.calibrate_margin(net, ho, target=0.004)
    ev = S.evaluate(net, ho, margin)
    ev["margin"] = margin
    log("held-out", ev)
    del ho
    torch.cuda.empty_cache()
    x0 = S.make_starts(13, n_tri, orbit0=1 + n_train + n_ho)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain = S.run_plain(x0, 13, N)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    lab_plain = S.p1_labels(plain["lam_N"].cpu().numpy(), plain["lam_N10"].cpu().numpy())
    del plain
    torch.cuda.empty_cache()
    log(f"plain: {n_tri} orbits in {t_plain:.2f} s")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tri = S.triage_run(net, x0, 13, N, prefix, margin, log=log)
    torch.cuda.synchronize()
    t_tri = time.perf_counter() - t0
    wrong = float(np.mean(tri["label"] != (lab_plain == 1)))
    # the same triage with the prefix pass on Float32 tangents (the continuation stays Float64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tri32 = S.triage_run(net, x0, 13, N, prefix, margin, log=log, elt=2)
    torch.cuda.synchronize()
    t_tri32 = time.perf_counter() - t0
    wrong32 = float(np.mean(tri32["label"] != (lab_plain == 1)))
    cpu = None
    if cpu_leg and not a.no_cpu:
        from chaos_amd import QRSpectrum, MGS, StandardMapK
        from chaos_amd.tangent import make_chain
        from tests import oracle_lib as O
        ms = 16384
        nthreads = nthreads_()
        chain, _ = make_chain((QRSpectrum(MGS()),))
        xs = x0[:: max(1, n_tri // ms)][:ms]
        W0 = O.frames(13, 1, n_tri, 2, 2)[:: max(1, n_tri // ms)][:ms]
        tc = time.perf_counter()
        O.run(StandardMapK().descriptor(), chain, xs, W0, N, 1, 2, math=O.MATH_LIBM, nthreads=nthreads)
        tc = time.perf_counter() - tc
        cpu = dict(value=float(len(xs)) * N / tc, unit="orbit-steps/s", cores=nthreads, kind="port",
                   sample=f"plain method on {len(xs)} of the triaged orbits x {N} steps, oracle/oracle.cpp with "
                          f"glibc sin/cos/log, {nthreads} threads, {tc:.1f} s")
    value = float(n_tri) * N / t_tri
    # the dominant part of the triage is the continuation of the ambiguous orbits: QRSpectrum(MGS()) k = 2 at the C1
    # census per orbit-step, over the wall time of that pass (host work included, so a lower bound); the network's
    # training throughput beside it, forward + backward at 3 x 2 (768*256 + 2*256*256 + 256*2) flops per sample-epoch
    cont_steps = float((1.0 - tri["confident"].mean()) * n_tri * (N - prefix))
    a_rest = FLOPS_PER_ORBIT_STEP["c1"] * cont_steps / max(tri["t_rest"], 1e-9) / 1e12
    mlp_flops = 3.0 * 2.0 * (768 * 256 + 2 * 256 * 256 + 256 * 2) * n_train * a.c5_epochs
    a_mlp = mlp_flops / max(t_train, 1e-9) / 1e12
    roof = {"bound": "fp64-valu", "achieved": a_rest, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": a_rest / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "continuation of the ambiguous orbits",
            "mlp_training": {"bound": "mfma", "achieved": a_mlp, "peak": 2500.0, "unit": "TFLOP/s (bf16 dense)",
                             "frac": a_mlp / 2500.0, "seconds": t_train}}
    out = {
        "metric": METRIC, "value": value, "unit": "orbit-steps/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": t_tri * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+bf16", "data": "synthetic",
        "config": {"workload": desc, "orbits": n_tri, "steps_per_orbit": N},
        "surrogate": {"speedup": t_plain / t_tri, "plain_s": t_plain, "triage_s": t_tri,
                      "triage_prefix_s": tri["t_prefix"], "triage_rest_s": tri["t_rest"],
                      "continued_fraction": float(1.0 - tri["confident"].mean()),
                      "misclassified_vs_plain": wrong, "heldout": ev, "train_s": t_train, "data_s": t_data,
                      "chaotic_fraction_train": info["chaotic_fraction"],
                      "final_loss": [float(v) for v in info["history"][-1]],
                      "f32tan_prefix": {"speedup": t_plain / t_tri32, "triage_s": t_tri32,
                                        "prefix_s": tri32["t_prefix"], "misclassified_vs_plain": wrong32,
                                        "continued_fraction": float(1.0 - tri32["confident"].mean())}},
        "roofline": roof, "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    return out


# The recorder's cost as the reference measures it (verify/experiments-device.jl:37-58): 2048 orbits x 2e5 steps per
# method, lifted, tau 1, the study grid; ns per orbit-step on its CPU threads and its GPU
# (workspaces/module-build/stages/04-verify/output/gpu-rtx5060-d53-verify-log.txt:168-175).
REC_REFERENCE_NS = {"L1": (4.173, 12.454), "L2": (2.937, 11.187), "L3": (7.906, 25.252), "L4": (8.594, 25.120),
                    "L4b": (8.732, 35.920), "L5": (9.171, 27.082), "L5b": (9.804, 27.927)}
REC_L2_EPS = 1e-8       # the study's primary two-trajectory distance; L2 records the wrapped map (its main_l2 arm)
REC_MONITOR_FLOPS = 47  # RunMonitor per handoff: det (5), M' Omega M - Omega and its maximum (35), max |M| (7)


def rec_main(a):
    """--workload rec: the recorder's device path (src/Experiments/run_device.jl) on the reference's cost ensemble:
    each method's grid records of 2048 standard-map orbits (K = 6, starts uniform on the torus) over 2e5 steps, lifted
    (L2: wrapped, eps = 1e-8, as the reference's cost run takes it), tau = 1, the study grid (20 points per decade,
    every step to 100, fractions 0.1 and 0.5), frame direction pi/4. One step = one method's full recording; value =
    orbit-steps over all methods / wall time (the rows copied back to the host included)."""
    import torch
    from chaos_amd import StandardMap, WrappedStandardMap, RunMonitor, torus_state
    from chaos_amd.recorder import record, record_grid, deviation_frame, estimator, vectors, METHODS
    from chaos_amd.inputs import torus_uniform
    m, N = [int(v) for v in a.rec_sizes.split(",")]
    sysm, sysw = StandardMap(6.0), WrappedStandardMap(6.0)
    x0 = torus_uniform(1, m)
    xw = np.array([torus_state(z) for z in x0])
    grid = record_grid(N, 1, points_per_decade=20, first_steps=100, extra_fractions=(0.1, 0.5))
    log(f"recorder: {m} orbits x {N} steps, {len(grid)} grid rows, methods {', '.join(METHODS)}")
    record(sysm, x0[:64], 200, method="L4", tau=1)  # warm-up: loads the kernels
    record(sysw, xw[:64], 200, method="L2", tau=1, eps=REC_L2_EPS)
    per, wall, dev_ms, bad = {}, 0.0, {}, 0
    for meth in METHODS:
        tm = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if meth == "L2":
            rec = record(sysw, xw, N, method=meth, tau=1, grid=grid, eps=REC_L2_EPS, timing=tm)
        else:
            rec = record(sysm, x0, N, method=meth, tau=1, grid=grid, timing=tm)
        dt = time.perf_counter() - t0
        wall += dt
        per[meth] = dt / (m * N) * 1e9
        dev_ms[meth] = tm["device_ms"]
        bad += int(np.sum(rec.fail > 0))
        log(f"recorder {meth}: {dt:.3f} s, {per[meth]:.3f} ns per orbit-step, {int(np.sum(rec.fail > 0))} failed")
    total = len(METHODS) * m * N
    value = total / wall
    cpu = None
    if not a.no_cpu:
        from chaos_amd.tangent import make_chain
        from tests import oracle_lib as O
        ms_ = max(1, m // 8)
        nt = nthreads()
        tc, steps_c = 0.0, 0
        for meth in METHODS:
            k = vectors(meth)
            W0 = np.broadcast_to(deviation_frame("pi/4", k), (ms_, 2, k))
            t0 = time.perf_counter()
            if meth == "L2":
                O.shadow_record(sysw.descriptor(), xw[:ms_], np.ascontiguousarray(W0[:, :, 0]), N, 1, REC_L2_EPS,
                                grid, math=O.MATH_LIBM, nthreads=nt)
            else:
                chain, _ = make_chain((estimator(meth), RunMonitor(sysm)))
                O.record(sysm.descriptor(), chain, x0[:ms_], W0, N, 1, k, grid, math=O.MATH_LIBM, nthreads=nt)
            tc += time.perf_counter() - t0
            steps_c += ms_ * N
        cpu = dict(value=steps_c / tc, unit="orbit-steps/s", cores=nt, kind="port",
                   sample=f"the first {ms_} orbits of each method's ensemble x {N} steps, the oracle's recording "
                          f"(oracle/oracle.cpp, or_record and or_shadow_record) with glibc sin/cos/log, {nt} threads, "
                          f"{tc:.1f} s")
    fpos = 174 + REC_MONITOR_FLOPS
    achieved = fpos * m * N / (dev_ms["L4"] * 1e-3) / 1e12
    out = {
        "metric": METRIC, "value": value, "unit": "orbit-steps/s", "n_gpus": 1, "steps": len(METHODS), "warmup": 1,
        "ms_per_step": wall * 1e3 / len(METHODS), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"recorder (grid records), standard map K=6, {m} orbits x {N} steps per method, "
                               f"lifted (L2 wrapped, eps={REC_L2_EPS}), tau=1, study grid ({len(grid)} rows), "
                               f"methods {'/'.join(METHODS)}",
                   "ns_per_orbit_step": per, "device_ms": dev_ms, "failed_orbits": bad,
                   "reference_ns_per_orbit_step_cpu23_gpu_rtx5060": REC_REFERENCE_NS},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel_ms": dev_ms["L4"],
                     "flops_per_orbit_step": fpos, "kernel": "method L4 (MGS + RunMonitor)"},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)


# The reference's per-spec cost records of its indicator recorder (verify/experiments-indicators-device.jl:151-176, specs
# st and bk of verify/experiments-indicators-block.toml:123-184; workspaces/module-build/stages/04-verify/output/
# gpu-rtx5060-d86-verify-log.txt:24-29): 1024 wrapped orbits, ns per orbit-step on 23 CPU threads and its GPU (st is
# CPU-only there).
IND_REFERENCE_NS = {"st": (15.34, None), "bk": (29.29, 63.1), "tg": (45.14, 64.88), "pr": (925.4, 267.6),
                    "pa": (29.34, 32.79)}


def ind_pa_case(m, K=2 * math.pi, seed=5):
    """The phase-advance window kind as the reference times it (gpu-rtx5060-d86-verify-log.txt:28, spec pa of
    experiments-indicators-block.toml:139-149): RunMonitor then WeightedPhaseAdvance (windows 250 and 1000, order 2,
    planar (q, p)) with the constant deviation column, 2001 steps, wrapped map. Returns the recording's device time."""
    from chaos_amd import WrappedStandardMap, WeightedPhaseAdvance, PlanarObservable, torus_state
    from chaos_amd.inputs import torus_uniform
    from chaos_amd.recorder import record_indicators
    sysm = WrappedStandardMap(K)
    x0 = np.array([torus_state(v) for v in torus_uniform(seed, m)])
    pa = WeightedPhaseAdvance(window=(250, 1000), observable=PlanarObservable(1, 2), birkhoff_order=2)
    t = {}
    t0 = time.perf_counter()
    rec = record_indicators(sysm, x0, 2001, reader=pa, tau=1, timing=t)
    return dict(m=m, N=2001, wall=time.perf_counter() - t0, kernel_ms=t["device_ms"], rows=len(rec.grid),
                bad=int((rec.code != 0).sum()))


def ind_pr_case(m, N=300, K=2 * math.pi, seed=5):
    """The indicator recorder's pair kind as the reference times it (gpu-rtx5060-d86-verify-log.txt:25, spec pr of
    experiments-indicators-block.toml:109-115): PairMap of the wrapped map, members RLI, MEGNO then RunMonitor,
    separation 1e-12 along the start direction, the study grid of 300 steps. Returns the recording's device time."""
    from chaos_amd import WrappedStandardMap, PairMap, pair_start
    from chaos_amd.inputs import torus_uniform
    from chaos_amd.recorder import record_indicators
    base = WrappedStandardMap(K)
    u = np.array([math.cos(math.pi / 4), math.sin(math.pi / 4)])
    z = np.array([pair_start(base, v, u, 1e-12)[0] for v in torus_uniform(seed, m)])
    t = {}
    t0 = time.perf_counter()
    rec = record_indicators(PairMap(base), z, N, members=("RLI", "MEGNO"), tau=1, timing=t)
    return dict(m=m, N=N, wall=time.perf_counter() - t0, kernel_ms=t["device_ms"], rows=len(rec.grid),
                bad=int((rec.code != 0).sum()))


def ind_tg_case(m, N=4000, K=2 * math.pi, seed=5):
    """The indicator recorder's tangent kind as the reference times it (gpu-rtx5060-d86-verify-log.txt:24, spec tg of
    experiments-indicators-block.toml:97-103): members FLI, MEGNO, SALI then RunMonitor, k = 2, the study grid of
    4000 steps, rows written at every grid step; wrapped map, tau = 1. Returns the device time of the recording."""
    from chaos_amd import WrappedStandardMap, torus_state
    from chaos_amd.inputs import torus_uniform
    from chaos_amd.recorder import record_indicators
    sysm = WrappedStandardMap(K)
    x0 = np.array([torus_state(v) for v in torus_uniform(seed, m)])
    t = {}
    t0 = time.perf_counter()
    rec = record_indicators(sysm, x0, N, members=("FLI", "MEGNO", "SALI"), tau=1, timing=t)
    return dict(m=m, N=N, wall=time.perf_counter() - t0, kernel_ms=t["device_ms"], rows=len(rec.grid),
                bad=int((rec.code != 0).sum()))


def ind_case(lib, ctx, name, m, K=2 * math.pi, seed=5):
    """One indicator workload on device buffers: st = (StretchingNumbers(bins 100, interval 4), RunMonitor), k = 2,
    4000 steps; bk = (RunMonitor, WeightedBirkhoff(windows 250 and 1000, order 2, cos q, sin q, cos 2p, lifted p)),
    k = 1 with the window kinds' constant deviation column, 2000 steps. Wrapped map, tau = 1, starts uniform on the
    torus. Returns the launch's device time (HIP events on the launch stream) and the device buffers."""
    import torch
    from chaos_amd import (WrappedStandardMap, StretchingNumbers, RunMonitor, WeightedBirkhoff, CosObservable,
                           SinObservable, LiftedObservable, stretching_range, torus_state)
    from chaos_amd.tangent import make_chain
    from chaos_amd.inputs import torus_uniform
    sysm = WrappedStandardMap(K)
    z = torus_uniform(seed, m)
    x0 = np.array([torus_state(v) for v in z])
    if name == "st":
        reds = (StretchingNumbers(range=stretching_range(sysm), bins=100, interval=4, min_per_cell=2), RunMonitor(sysm))
        k, N, reset = 2, 4000, False
    else:
        reds = (RunMonitor(sysm), WeightedBirkhoff(window=(250, 1000), birkhoff_order=2, observables=(
            CosObservable(1), SinObservable(1), CosObservable(2, 2), LiftedObservable(2, 4))))
        k, N, reset = 1, 2000, True
    chain, allr = make_chain(reds, reset_frame=reset)
    sd = sysm.descriptor()
    words = lib.cl_state_bytes(ctypes.byref(sd), ctypes.byref(chain), 1, 2, k) // 8
    nws = lib.cl_workspace_entries(ctypes.byref(chain), 2, k)
    dev = torch.device("cuda")
    dx = torch.from_numpy(np.ascontiguousarray(x0.T)).to(dev)
    W = np.eye(2)[:, :k]
    dW = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(W.T.reshape(2 * k, 1), (2 * k, m)))).to(dev)
    dst = torch.empty((words, m), dtype=torch.int64, device=dev)
    dcode = torch.zeros(m, dtype=torch.int8, device=dev)
    dh = torch.zeros(m, dtype=torch.int64, device=dev)
    dws = torch.empty((max(nws, 1), m), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    if nws:
        rc = lib.cl_propagate_device_ws(ctx, ctypes.byref(sd), ctypes.byref(chain), 1, m, 4, 2, k, 0, N, N, 1,
                                        dx.data_ptr(), dW.data_ptr(), None, dst.data_ptr(), dcode.data_ptr(),
                                        dh.data_ptr(), dws.data_ptr(), stream.cuda_stream)
    else:
        rc = lib.cl_propagate_device(ctx, ctypes.byref(sd), ctypes.byref(chain), 1, m, 4, 2, k, 0, N, N, 1,
                                     dx.data_ptr(), dW.data_ptr(), None, dst.data_ptr(), dcode.data_ptr(),
                                     dh.data_ptr(), stream.cuda_stream)
    e1.record(stream)
    assert rc == 0, lib.cl_last_error(ctx)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return dict(name=name, m=m, N=N, wall=wall, kernel_ms=e0.elapsed_time(e1), bad=int((dcode != 0).sum().item()),
                x0=x0, chain=chain, allr=allr, k=k, nws=nws, sysm=sysm)


# algorithmic HBM bytes per orbit-step of the bk launch: the horizon sums a handoff reads and writes (window 1: B and,
# per observable, S and A1, 2 + 3K words; window 2: S and A2, 3K words), summed over the two horizons' 2N handoffs
# (250: 250 x 14 + 250 x 12 words; 1000: 1000 x 14 + 1000 x 12), 8 bytes each way, over 2000 steps
IND_BK_BYTES = (250 * 14 + 250 * 12 + 1000 * 14 + 1000 * 12) * 16 / 2000
HBM_PEAK_GBS = 8000.0


def ind_main(a):
    """--workload ind: the two indicator reducers of the path with no reference device path or a slow one, at the
    reference's own ensemble (1024 orbits) and at 2^20 orbits. value = orbit-steps of the 2^20 runs / their wall time
    (inputs resident, one launch each)."""
    from chaos_amd.tangent import _context
    lib, ctx = _context()
    big = int(a.ind_orbits)
    for name in ("st", "bk"):
        ind_case(lib, ctx, name, 256)   # warm-up: loads the kernels
    res = {}
    for name in ("st", "bk"):
        for m in (1024, big):
            r = ind_case(lib, ctx, name, m)
            res[(name, m)] = r
            log(f"ind {name}: {m} orbits x {r['N']} steps: {r['kernel_ms']:.2f} ms device, "
                f"{r['kernel_ms'] * 1e6 / (m * r['N']):.4f} ns per orbit-step, {r['bad']} failed")
    ind_tg_case(256, N=400)   # warm-up
    for m in (1024, 65536):
        r = ind_tg_case(m)
        res[("tg", m)] = r
        log(f"ind tg (recorded, {r['rows']} rows): {m} orbits x {r['N']} steps: {r['kernel_ms']:.2f} ms device, "
            f"{r['kernel_ms'] * 1e6 / (m * r['N']):.4f} ns per orbit-step, {r['bad']} failed")
    ind_pr_case(256, N=60)   # warm-up
    for m in (1024, 65536):
        r = ind_pr_case(m)
        res[("pr", m)] = r
        log(f"ind pr (recorded, {r['rows']} rows): {m} orbits x {r['N']} steps: {r['kernel_ms']:.2f} ms device, "
            f"{r['kernel_ms'] * 1e6 / (m * r['N']):.4f} ns per orbit-step, {r['bad']} failed")
    ind_pa_case(256)   # warm-up
    for m in (1024, 65536):
        r = ind_pa_case(m)
        res[("pa", m)] = r
        log(f"ind pa (recorded, {r['rows']} rows): {m} orbits x {r['N']} steps: {r['kernel_ms']:.2f} ms device, "
            f"{r['kernel_ms'] * 1e6 / (m * r['N']):.4f} ns per orbit-step, {r['bad']} failed")
    total = sum(res[(n, big)]["m"] * res[(n, big)]["N"] for n in ("st", "bk"))
    wall = sum(res[(n, big)]["wall"] for n in ("st", "bk"))
    value = total / wall
    per = {f"{n}@{m}": res[(n, m)]["kernel_ms"] * 1e6 / (m * res[(n, m)]["N"]) for n in ("st", "bk") for m in (1024, big)}
    per.update({f"{n}@{m}": res[(n, m)]["kernel_ms"] * 1e6 / (m * res[(n, m)]["N"]) for n in ("tg", "pr", "pa")
                for m in (1024, 65536)})
    cpu = None
    if not a.no_cpu:
        from tests import oracle_lib as O
        nt = nthreads()
        tc, steps_c = 0.0, 0
        for n in ("st", "bk"):
            r = res[(n, 1024)]
            ms_ = 256
            W0 = np.broadcast_to(np.eye(2)[:, :r["k"]], (ms_, 2, r["k"]))
            t0 = time.perf_counter()
            if r["nws"]:
                O.run_ws(r["sysm"].descriptor(), r["chain"], r["nws"], r["x0"][:ms_], W0, r["N"], 1, r["k"],
                         math=O.MATH_LIBM, nthreads=nt)
            else:
                O.run(r["sysm"].descriptor(), r["chain"], r["x0"][:ms_], W0, r["N"], 1, r["k"], math=O.MATH_LIBM,
                      nthreads=nt)
            tc += time.perf_counter() - t0
            steps_c += ms_ * r["N"]
        cpu = dict(value=steps_c / tc, unit="orbit-steps/s", cores=nt, kind="port",
                   sample=f"the first 256 orbits of each 1024-orbit case, the oracle (or_propagate_ws, or_propagate) "
                          f"with glibc sin/cos/log/exp, {nt} threads, {tc:.2f} s")
    bk = res[("bk", big)]
    achieved = IND_BK_BYTES * bk["m"] * bk["N"] / (bk["kernel_ms"] * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": value, "unit": "orbit-steps/s", "n_gpus": 1, "steps": 2, "warmup": 1,
        "ms_per_step": wall * 1e3 / 2, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"indicator reducers on the wrapped standard map K=2pi, tau=1: st = StretchingNumbers "
                               f"(100 cells, interval 4) + RunMonitor, k=2, 4000 steps; bk = RunMonitor + "
                               f"WeightedBirkhoff (windows 250/1000, order 2, 4 observables), k=1 constant column, "
                               f"2000 steps; {big} orbits each (and the reference's 1024); tg = the indicator "
                               f"recorder's tangent kind (FLI, MEGNO, SALI + RunMonitor, study grid rows), 4000 steps, "
                               f"1024 and 65536 orbits; pr = its pair kind (PairMap, RLI, MEGNO + RunMonitor), 300 "
                               f"steps, 1024 and 65536 pairs; pa = its phase-advance window kind (RunMonitor + "
                               f"WeightedPhaseAdvance, windows 250/1000, order 2, planar), 2001 steps, 1024 and 65536 "
                               f"orbits (tg, pr and pa outside value)",
                   "ns_per_orbit_step": per, "reference_ns_per_orbit_step_cpu23_gpu_rtx5060": IND_REFERENCE_NS,
                   "failed_orbits": {f"{n}@{m}": res[(n, m)]["bad"] for n in ("st", "bk") for m in (1024, big)}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": TRAFFIC_BYTES_PER_LAUNCH["ind_bk"] if bk["m"] == 1 << 20 else None,
                     "kernel_ms": bk["kernel_ms"],
                     "bytes_per_orbit_step": IND_BK_BYTES, "kernel": "bk (WeightedBirkhoff horizon sums in HBM)"},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c2", "c1", "c3", "c4", "c5", "rec", "ind", "record"])
    ap.add_argument("--record-path", default=None, help="record: the TOML file the entry is appended to")
    ap.add_argument("--ind-orbits", default=1 << 20, type=int, help="ind: orbits of the large cases")
    ap.add_argument("--rec-sizes", default="2048,200000", help="rec: orbits, steps")
    ap.add_argument("--c5-sizes", default="1048576,262144,4194304,100000",
                    help="c5: training orbits, held-out orbits, triaged orbits, steps N")
    ap.add_argument("--c5-epochs", type=int, default=8)
    ap.add_argument("--cpu-orbits", type=int, default=32768, help="orbits per K in the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="c2: skip the Float32-tangent, accuracy, C1, C5, C3 and C4 blocks of the default run")
    ap.add_argument("--c3-qr", default="auto", choices=["auto", "blocked", "sites"],
                    help="c3: the HouseholderQR order (auto: the library's rule, the blocked one on the matrix cores)")
    ap.add_argument("--c3-elt", default="f64", choices=["f64", "f32tan"],
                    help="c3: FP64 frames or Float32 tangents (FP64 orbit, factorisation and log sums)")
    ap.add_argument("--c3-mode", default="streams", choices=["streams", "mixed", "serial"],
                    help="c3: one launch per lattice on five streams, one persistent launch over all five, or the five "
                         "launches one after another on one stream (each lattice's kernel alone on the device)")
    ap.add_argument("--c3-chunk", type=int, default=0, help="c3 mixed: steps per graph-captured chunk (0: one chunk)")
    ap.add_argument("--c3-order", default="ascending", choices=["ascending", "descending"],
                    help="c3 streams: launch the lattices' chain kernels in ascending or descending N")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first copy 1 GiB device-to-device (the known-size kernel that calibrates FETCH_SIZE/WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; multi-GPU runs are not supported (asked for --gpus {a.gpus}). The ensemble "
              f"would shard by orbit ranges with no exchange during propagation, so throughput is expected to scale "
              f"with the GPU count; that is not measured.", file=sys.stderr)
        sys.exit(2)
    if a.workload == "c5":
        return c5_main(a)
    if a.workload == "c3":
        return c3_main(a)
    if a.workload == "rec":
        return rec_main(a)
    if a.workload == "record":
        return record_main(a)
    if a.workload == "ind":
        return ind_main(a)
    wl = workload(a.workload)
    log("workload", wl["desc"])
    if a.copy_probe:
        import torch
        src = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 1 GiB read + 1 GiB written by one device copy kernel")
        del src, dst
    extra = {}
    if wl["name"] in ("c2", "c1") and not a.no_extras:
        extra["autotune"] = tune_block(wl)
    g = gpu_run(wl, a.steps, a.warmup)
    value = g["orbit_steps"] * a.steps / g["wall"]
    avg_ms = float(np.mean(g["kernel_ms"]))
    if wl["name"] == "c4":
        extra["energy_drift"] = {"f64": c4_energy_drift(wl, g["x_final"])}
        if not a.no_extras:  # the compensated (double-double) orbit on the same workload
            wdd = dict(wl)
            wdd["x0"] = np.concatenate([wl["x0"], np.zeros_like(wl["x0"])], axis=1)
            gdd = gpu_run(wdd, max(1, min(a.steps, 2)), 1, elt=3)
            ms_dd = float(np.mean(gdd["kernel_ms"]))
            extra["dd"] = {"value": gdd["orbit_steps"] * len(gdd["kernel_ms"]) / gdd["wall"], "kernel_ms": ms_dd,
                           "over_f64": avg_ms / ms_dd, "nonfinite_orbits": gdd["bad"]}
            extra["energy_drift"]["dd"] = c4_energy_drift(wl, gdd["x_final"][:, :4])
        log("energy drift", extra["energy_drift"])
    if wl["name"] == "c2" and not a.no_extras:
        # the Float32-tangent form of the same workload (Float64 orbit and reducers)
        g32 = gpu_run(wl, max(1, min(a.steps, 6)), 1, elt=2)
        ms32 = float(np.mean(g32["kernel_ms"]))
        a32 = FLOPS_PER_ORBIT_STEP["c2"] * g32["orbit_steps"] / (ms32 * 1e-3) / 1e12
        extra["c2_f32tan"] = {"value": g32["orbit_steps"] * len(g32["kernel_ms"]) / g32["wall"], "kernel_ms": ms32,
                              "over_f64": avg_ms / ms32, "nonfinite_orbits": g32["bad"],
                              "roofline": {"bound": "fp64-valu", "achieved": a32, "peak": FP64_PEAK_TFLOPS,
                                           "unit": "TFLOP/s", "frac": a32 / FP64_PEAK_TFLOPS}}
        log("c2 f32tan", extra["c2_f32tan"])
        if not a.no_cpu:
            extra["accuracy"] = accuracy_protocol(wl)
            extra["cpu_sanity"] = lognorm_one_thread()
            log("cpu sanity", extra["cpu_sanity"])
        extra["c1"] = c1_block(a)
        log("c1", {k: v for k, v in extra["c1"].items() if k != "cpu_baseline"})
        c5 = c5_run(a, cpu_leg=False)
        extra["c5"] = {"value": c5["value"], "unit": c5["unit"], "workload": c5["config"]["workload"],
                       **c5["surrogate"], "roofline": c5["roofline"]}
        log("c5", {k: extra["c5"][k] for k in ("speedup", "misclassified_vs_plain", "continued_fraction")})
        # the coupled chains (C3) on five streams in the library's QR order (the blocked one on the matrix cores)
        a3 = argparse.Namespace(**vars(a))
        a3.steps, a3.c3_mode, a3.c3_qr, a3.no_cpu, a3.c3_order, a3.c3_elt = 2, "streams", "auto", True, "ascending", "f64"
        c3 = c3_main(a3, emit=False)
        lam8 = c3["_lambdas"][8]  # N = 8: k = 16 = 2N, the full spectrum
        pair = np.max(np.abs(lam8 + lam8[:, ::-1]), axis=1)
        extra["c3"] = {"value": c3["value"], "unit": c3["unit"], "workload": c3["config"]["workload"],
                       "kernel_ms_per_N": c3["config"]["kernel_ms_per_N"], "roofline": c3["roofline"],
                       "nonfinite_orbits": c3["config"]["nonfinite_orbits"],
                       "pairing_n8": {"max": float(pair.max()), "median": float(np.median(pair)),
                                      "sum_max": float(np.max(np.abs(lam8.sum(axis=1))))}}
        log("c3", {"value": c3["value"], "span_ms": c3["roofline"]["kernel_ms"], "frac": c3["roofline"]["frac"]})
        # the Float32-tangent variant of the same ensembles: rate and the same-orbit error of the spectrum
        a3.c3_elt = "f32tan"
        c3f = c3_main(a3, emit=False)
        err = {}
        for N, l64 in c3["_lambdas"].items():
            d = np.abs(c3f["_lambdas"][N] - l64)
            err[N] = {"lambda_1_max": float(d[:, 0].max()), "lambda_1_median": float(np.median(d[:, 0])),
                      "all_max": float(d.max()), "all_median": float(np.median(d))}
        extra["c3"]["f32tan"] = {"value": c3f["value"], "over_f64": c3f["value"] / c3["value"],
                                 "span_ms": c3f["roofline"]["kernel_ms"], "frac": c3f["roofline"]["frac"],
                                 "kernel_ms_per_N": c3f["config"]["kernel_ms_per_N"],
                                 "nonfinite_orbits": c3f["config"]["nonfinite_orbits"],
                                 "lambda_error_vs_f64_same_orbit": err}
        log("c3 f32tan", {k: v for k, v in extra["c3"]["f32tan"].items() if k != "kernel_ms_per_N"})
        extra["c4"] = c4_block(a)
        log("c4", {k: extra["c4"][k] for k in ("value", "kernel_ms", "energy_drift")})
    budget = {"c2": a.cpu_orbits, "c1": 4096, "c4": 2048}[wl["name"]]
    cpu = None if a.no_cpu else cpu_baseline(wl, budget)
    if cpu is not None:
        cpu["lambda_check"] = lambda_check(wl, g)
    fpos = flops_per_orbit_step(wl, cpu)
    achieved = fpos * g["orbit_steps"] / (avg_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC, "value": value, "unit": "orbit-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": g["wall"] * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl["desc"], "orbits_per_step": g["m"], "steps_per_orbit": wl["N"],
                   "tau": wl["tau"], "k": wl["k"], "nonfinite_orbits": g["bad"]},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": TRAFFIC_BYTES_PER_LAUNCH.get(wl["name"]),
                     "kernel_ms": avg_ms, "flops_per_orbit_step": fpos},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
