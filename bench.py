"""Benchmark: direct-SCF Fock builds/s and SCF wall time on one MI355X, FP64 vs mixed precision.

Workload: linear all-trans C20H42 in cc-pVDZ with its general contractions expanded (510
Cartesian functions, 246 shells), seeded +-0.02 Bohr jitter.  One step = one full Fock
build (one hfx_jk call, dD = D_conv, D_prev = 0) from device-resident inputs.  The run:
  1. FP64 SCF (Schwarz 1e-10, density 1e-10, DIIS 6, energy tolerance 1e-8) -> D_conv, E64
  2. autotuned FP32 cut at D_conv, then the mixed-precision SCF: mixed builds until |dE| < 1e-4, then FP64
     builds from a full FP64 build (DIIS history kept) to convergence
  3. W warm-up + K timed builds each of: FP64 eager, FP64 graph-replayed, mixed eager,
     mixed graph-replayed (the headline value); then K more builds of each precision with
     HIP events around every class kernel group (per-class times, roofline of the
     dominant group)
  4. batched mode: 64 seeded water clusters (n ~ U{1..8}, 6-31G) through the batched SCF
     driver, FP64 and mixed + graph, per-molecule |dE| of mixed against FP64
  5. gradient: RHF nuclear gradient of (H2O)8 / 6-31G**, FP64 and mixed
  6. CPU baseline: the oracle's OpenMP port (reference release flags) on a bounded slice
     of the same build, scaled by the model FLOPs of the slice
  7. conventional path: the Schwarz-screened integral blocks stored in HBM once, then K
     graph-replayed FP64 builds digesting them (HBM-bound), and the conventional SCF
Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]

Profiling aids: --profile-gradient runs only the (H2O)8 gradient (W + K times, FP64 and mixed);
--save-density PATH stores the converged FP64 density (and stops); a later
  python bench.py --profile-builds --density PATH --cut C --steps K --warmup W
runs only the W + K mixed builds with per-class HIP-event timing (no SCF), so a profiler's
per-kernel averages cover exactly the launches the roofline is computed from; with --stored
instead of --cut, the same for W + K builds of the conventional path (store, then stored builds).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hartree-fock_amd"))

import numpy as np  # noqa: E402

from hfx import Basis, autotune, device, geometry, lib  # noqa: E402

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}  # MI355X listed vector peaks (FP32 counts packed FP32)
PEAK_HBM_GBPS = 8000.0  # MI355X HBM3E
CLASS_NAMES = ["ssss", "psss", "psps", "ppss", "ppps", "pppp", "dsss", "dsps", "dspp", "dsds", "dpss", "dpps",
               "dppp", "dpds", "dpdp", "ddss", "ddps", "ddpp", "ddds", "dddp", "dddd"]
_L = {"s": 0, "p": 1, "d": 2}
CLASS_L = [tuple(_L[c] for c in name) for name in CLASS_NAMES]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def timed_builds(L, eng, bufs, n, steps, warmup, precision, cut, Q, flags=lib.HFX_JK_CLASS_TIMING):
    dDt, dDa, J, Ka = bufs
    kw = dict(Q=Q, precision=precision, fp32_cut=cut, flags=flags)
    for _ in range(warmup):
        lib.jk_device(eng, dDt, dDa, J, Ka, **kw)
    device.synchronize()
    ms64 = np.zeros(21)
    ms32 = np.zeros(21)
    dev_ms = []
    t0 = time.perf_counter()
    st = None
    for _ in range(steps):
        st = lib.jk_device(eng, dDt, dDa, J, Ka, **kw)
        ms64 += st["class_ms64"]
        ms32 += st["class_ms32"]
        dev_ms.append(st["device_ms"])
    device.synchronize()
    dt = time.perf_counter() - t0
    st = dict(st, device_ms_median=float(np.median(dev_ms)))
    return dt, st, ms64 / steps, ms32 / steps


def cpu_baseline(basis, Da, Q, total_gflop, budget_s):
    """Oracle port (OpenMP, reference release flags) on bra shells [a0, a1) of the same build."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as C

    import pyoracle
    O = pyoracle.load(fast=True)
    e = lib.Engine(O, basis)
    n = basis.n_aos
    nS = basis.n_shells
    Dt = np.ascontiguousarray((2 * Da).reshape(-1, order="F"))
    Dal = np.ascontiguousarray(Da.reshape(-1, order="F"))
    Qf = np.ascontiguousarray(Q.reshape(-1, order="F"))
    f64p = C.POINTER(C.c_double)
    a = lib.HfxJkArgs()
    a.dD_tot, a.dD_alpha, a.dD_beta = Dt.ctypes.data_as(f64p), Dal.ctypes.data_as(f64p), None
    a.Q = Qf.ctypes.data_as(f64p)
    a.schwarz_thresh, a.density_thresh, a.precision = 1e-10, 1e-10, lib.HFX_FP64
    J = np.zeros(n * n)
    K = np.zeros(n * n)

    def run(a0, a1):
        st = lib.HfxJkStats()
        t = time.perf_counter()
        rc = O.dll.oracle_jk_range(e.h, C.byref(a), a0, a1, J.ctypes.data_as(f64p), K.ctypes.data_as(f64p), None,
                                   C.byref(st))
        O.check(rc, "oracle_jk_range")
        return time.perf_counter() - t, st.model_gflop

    mid = nS // 2
    t, g = run(mid, mid + 1)  # calibrate on one bra shell
    width = 1
    while t < budget_s / 3 and mid + width < nS:
        width *= 2
        t, g = run(mid, min(nS, mid + width))
    rate = g / t if t > 0 else 0.0  # model GFLOP/s on this slice
    builds = rate / total_gflop if total_gflop > 0 else 0.0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return dict(value=builds, unit="builds/s", cores=threads, kind="port", cpu_model=model,
                gflops_model=rate,
                sample="bra shells %d..%d of %d (%.1f of %.1f model GFLOP) of the same FP64 build, %.1f s" %
                       (mid, min(nS, mid + width) - 1, nS, g, total_gflop, t))


def batched_run(L, cut, n_mol=64):
    mols = geometry.water_batch(n_mol=n_mol, seed=1004)
    bases = [Basis("6-31g", Z, x) for Z, x in mols]
    kw = dict(energy_tol=1e-8, schwarz_thresh=1e-10, density_thresh=1e-10, diis_size=6, max_iter=60)
    t = time.perf_counter()
    r64 = lib.scf_run_batched(L, bases, options=L.default_options(**kw))
    t64 = time.perf_counter() - t
    t = time.perf_counter()
    rmx = lib.scf_run_batched(L, bases, options=L.default_options(precision=lib.HFX_MIXED, fp32_cut=cut,
                                                                  final_fp64=1, jk_flags=lib.HFX_JK_USE_GRAPH, **kw))
    tmx = time.perf_counter() - t
    dE = [abs(a["e_total"] - b["e_total"]) for a, b in zip(rmx, r64)]
    dEm = [abs(a["e_total_mixed"] - b["e_total"]) for a, b in zip(rmx, r64)]
    return dict(molecules=n_mol, functions_min_max=[min(b.n_aos for b in bases), max(b.n_aos for b in bases)],
                total_functions=sum(b.n_aos for b in bases), scf_wall_s_fp64=t64, scf_wall_s_mixed_graph=tmx,
                builds_fp64=r64[0]["n_jk_builds"], builds_mixed=rmx[0]["n_jk_builds"],
                converged=sum(a["converged"] for a in rmx), max_dE_final=max(dE), max_dE_at_switch=max(dEm))


def eri_model_flop(la, lb, lc, ld):
    """Model FLOPs of one primitive quartet of the J/K build (the engine's count: Boys values and
    R^0_tuv to level L, the bra-side Hermite sums for every ket Hermite index, the ket-side sums
    for every bra component, and the final products), with the cheaper of the two orientations."""
    def nherm(l):
        return (l + 1) * (l + 2) * (l + 3) // 6

    def ncart(l):
        return (l + 1) * (l + 2) // 2

    def hsum(l1, l2):
        s = 0
        for i1 in range(l1, -1, -1):
            for j1 in range(l1 - i1, -1, -1):
                for i2 in range(l2, -1, -1):
                    for j2 in range(l2 - i2, -1, -1):
                        s += (i1 + i2 + 1) * (j1 + j2 + 1) * ((l1 - i1 - j1) + (l2 - i2 - j2) + 1)
        return s

    def one(a, b, c, d):
        L = a + b + c + d
        c4 = (L + 4) * (L + 3) * (L + 2) * (L + 1) / 24.0
        return (20 + 3 * (L + 1)) + 3 * c4 + 2.0 * nherm(a + b) * hsum(c, d) + \
            2.0 * ncart(c) * ncart(d) * hsum(a, b) + 2.0 * ncart(a) * ncart(b) * ncart(c) * ncart(d)
    return min(one(la, lb, lc, ld), one(lc, ld, la, lb))


def grad_model_flop(la, lb, lc, ld):
    """Model FLOPs of one primitive quartet of the two-electron gradient: the derivative of (ab|cd)
    on centre A is 2a (a+1 b|cd) - a_x (a-1 b|cd), and likewise on B and C (D follows from
    translational invariance), so the gradient needs the three classes with one shell raised, each
    counted with the J/K build's model (eri_model_flop; the lowered classes are their subsets), plus
    the contraction of the nine derivative blocks with Gamma (2 FLOPs per element).  A count of the
    algorithm, not of the kernel's instructions (the kernels re-form the ket sums per component pair)."""
    ncomp = ((la + 1) * (la + 2) // 2) * ((lb + 1) * (lb + 2) // 2) * ((lc + 1) * (lc + 2) // 2) * \
        ((ld + 1) * (ld + 2) // 2)
    return (eri_model_flop(la + 1, lb, lc, ld) + eri_model_flop(la, lb + 1, lc, ld) +
            eri_model_flop(la, lb, lc + 1, ld) + 18.0 * ncomp)


def gradient_roofline(st, device_ms):
    """Roofline of the gradient's dominant class (its kernels, one class, timed alone): model FLOPs
    of its primitive quartets (grad_model_flop) over its device time, against the FP64 vector peak;
    the whole two-electron part over device_ms (the classes overlapping on four streams)."""
    cms, prims = st["class_ms"], st["class_prims"]
    k = int(np.argmax(cms))
    gfl = float(prims[k]) * grad_model_flop(*CLASS_L[k]) * 1e-9
    achieved = gfl / cms[k] if cms[k] > 0 else 0.0
    total = sum(float(prims[i]) * grad_model_flop(*CLASS_L[i]) for i in range(21)) * 1e-9
    kern = "k_grad_eri<double> (generic)" if CLASS_NAMES[k] in ("dddp", "dddd") else "k_grad_eri_t<double, c0>"
    per = {CLASS_NAMES[i]: round(float(prims[i]) * grad_model_flop(*CLASS_L[i]) * 1e-9 / cms[i], 3)
           for i in range(21) if cms[i] > 0 and prims[i] > 0}
    return dict(bound="valu", pipe="VALU f64", kernel="%s %s" % (kern, CLASS_NAMES[k]), achieved=achieved,
                peak=PEAK_TFLOPS["f64"], unit="TFLOP/s", frac=achieved / PEAK_TFLOPS["f64"], traffic=None,
                launch_ms=float(cms[k]), launch_model_gflop=gfl, prim_quartets=int(prims[k]),
                two_electron_model_gflop=total, two_electron_tflops=total / device_ms,
                class_model_tflops=per)


def gradient_run(L):
    Z, xyz = geometry.water_cluster_grid(shape=(2, 2, 2), spacing=5.6, jitter=0.4, seed=1002)
    b = Basis("6-31g**", Z, xyz)
    r = lib.scf_run(L, b, options=L.default_options(energy_tol=1e-9, density_tol=1e-7))
    P = 2 * r["D_alpha"]
    W = lib.energy_weighted_density(r["C_alpha"], r["eps_alpha"], sum(b.Z) // 2)
    e = lib.Engine(L, b)
    lib.rhf_gradient(e, b, P, W)  # warm-up
    t = time.perf_counter()
    g64, _, s64 = lib.rhf_gradient(e, b, P, W)
    t64 = time.perf_counter() - t
    t = time.perf_counter()
    gmx, _, smx = lib.rhf_gradient(e, b, P, W, precision=lib.HFX_MIXED, fp32_cut=1e-5)
    tmx = time.perf_counter() - t
    _, _, stc = lib.rhf_gradient(e, b, P, W, flags=lib.HFX_JK_CLASS_TIMING)  # per-class device times
    return dict(system="(H2O)8/6-31G** (%d functions)" % b.n_aos, seconds_fp64=t64, seconds_mixed=tmx,
                two_electron_ms_fp64=s64["device_ms"], two_electron_ms_mixed=smx["device_ms"],
                quartets=s64["n_quartets"], fp32_quartets=smx["n_fp32"],
                max_dgrad_mixed=float(np.abs(gmx - g64).max()), max_abs_grad=float(np.abs(g64).max()),
                class_ms_fp64={CLASS_NAMES[i]: round(float(stc["class_ms"][i]), 3) for i in range(21)},
                roofline=gradient_roofline(stc, s64["device_ms"]))


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_eri_c20h42_mixed_traffic.json")
CAL_COPY_DOUBLES = 1 << 26  # 512 MiB calibration copy of --profile-builds


def measured_traffic(group):
    """HBM bytes per launch of a class kernel group from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of the C20H42 mixed build (scaled by the copy-kernel calibration recorded there)."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)["per_class_launch_group"].get(group)
    except (OSError, ValueError, KeyError):
        return None
    return None if t is None else t["read_bytes"] + t["write_bytes"]


def stored_run(L, basis, eng, bufs, Q, steps, warmup, J64, K64, Da, scf_kw):
    """The conventional path at the same density: store, K graph-replayed stored builds, per-class
    times, the HBM roofline of the dominant stored-digest group, and the conventional SCF."""
    eng.set_launch(2, 8192, 4)  # the stored builds' defaults (longer claimed runs: fewer row flushes)
    t0 = time.perf_counter()
    nq, nbytes = eng.store_eris(1e-10, Q=Q)
    t_store = time.perf_counter() - t0
    flags = lib.HFX_JK_STORED | lib.HFX_JK_NO_STATS | lib.HFX_JK_USE_GRAPH
    dt, _, _, _ = timed_builds(L, eng, bufs, None, steps, warmup, lib.HFX_FP64, 0.0, Q, flags)
    Js, Ks = bufs[2].download(), bufs[3].download()
    _, st, cms, _ = timed_builds(L, eng, bufs, None, steps, 1, lib.HFX_FP64, 0.0, Q,
                                 lib.HFX_JK_STORED | lib.HFX_JK_CLASS_TIMING)
    eng.release_eris()
    # bytes a stored build reads per class: the padded blocks and the list entries
    groups = []
    for k in range(21):
        n_k = st["class_count"][k]
        if n_k == 0 or cms[k] <= 0:
            continue
        la, lb, lc, ld = CLASS_L[k]
        blk = ((la + 1) * (la + 2) // 2) * ((lb + 1) * (lb + 2) // 2) * ((lc + 1) * (lc + 2) // 2) * \
            ((ld + 1) * (ld + 2) // 2)
        # two passes read the blocks: the stored digest (bra-major) and the Coulomb pass (ket-major copy)
        groups.append((cms[k], k, 2 * 8.0 * ((n_k + 63) // 64 * 64 * blk + n_k)))
    ms, k, byt = max(groups)
    achieved = byt / ms * 1e-6  # GB/s
    roof = dict(bound="hbm", kernel="k_eri %s f64 stored digest + k_jket" % CLASS_NAMES[k], achieved=achieved,
                peak=PEAK_HBM_GBPS, unit="GB/s", frac=achieved / PEAK_HBM_GBPS, launch_ms=ms, launch_bytes=byt)
    rconv = lib.scf_run(L, basis, options=L.default_options(direct=0, **scf_kw))
    dE = abs(float(np.sum(Da.reshape(-1, order="F") * ((Js - Ks) - (J64 - K64)))))
    return dict(quartets=nq, stored_bytes=nbytes, store_s=t_store, builds_per_s=steps / dt, ms_per_build=1e3 * dt / steps,
                read_GBps_build=nbytes / (dt / steps) * 1e-9, build_dE_vs_direct=dE,
                build_max_dJK_vs_direct=float(max(np.abs(Js - J64).max(), np.abs(Ks - K64).max())),
                class_ms={CLASS_NAMES[i]: round(cms[i], 3) for i in range(21) if cms[i] > 0}, roofline=roof,
                scf_wall_s=rconv["time_total_s"], scf_iterations=rconv["iterations"], scf_e=rconv["e_total"])


def roofline_of(stmx, cm64, cm32):
    """Roofline of the dominant class kernel group (one class, one precision) of a build.

    `achieved` counts the model FLOPs of the primitive quartets the kernel executes (the engine
    prunes primitive pairs below 1e-20 / 1e-11 in FP64 / FP32, which an unpruned model count
    would still credit); the kernels are VALU-bound (no MFMA), so the bound is the
    vector pipe, priced at MI355X's listed vector peak (FP32 157.3 TFLOP/s counts packed FP32,
    which these kernels do not use: their non-packed ceiling is 78.6, given as peak_nonpacked)."""
    groups = [(cm64[k], "f64", k) for k in range(21)] + [(cm32[k], "f32", k) for k in range(21)]
    ms, prec, k = max(groups)
    gfl = (stmx["class_exec_gflop64"] if prec == "f64" else stmx["class_exec_gflop32"])[k]
    gfl_model = (stmx["class_gflop64"] if prec == "f64" else stmx["class_gflop32"])[k]
    achieved = gfl / ms if ms > 0 else 0.0  # GFLOP/ms = TFLOP/s
    group = "%s %s" % (CLASS_NAMES[k], prec)
    return dict(bound="valu", pipe="VALU %s" % prec, kernel="k_eri " + group,
                achieved=achieved, peak=PEAK_TFLOPS[prec], unit="TFLOP/s", frac=achieved / PEAK_TFLOPS[prec],
                peak_nonpacked=78.6, frac_nonpacked=achieved / 78.6,
                traffic=measured_traffic(group), traffic_source="rocprofv3 FETCH_SIZE + WRITE_SIZE bytes per launch, "
                "profiles/k_eri_c20h42_mixed_traffic.md", launch_ms=ms, launch_exec_gflop=gfl,
                launch_model_gflop_unpruned=gfl_model,
                build_exec_tflops=stmx["exec_gflop"] / stmx["device_ms"] if stmx["device_ms"] > 0 else 0.0)


def profile_builds(args):
    L = lib.load()
    Z, xyz = geometry.alkane_chain(n_carbon=args.carbons, seed=1003)
    basis = Basis("cc-pvdz-expanded", Z, xyz)
    n = basis.n_aos
    Da = np.load(args.density)
    eng = lib.Engine(L, basis)
    Q = eng.schwarz()
    bufs = (device.DeviceArray.from_host(2 * Da), device.DeviceArray.from_host(Da), device.DeviceArray(n * n),
            device.DeviceArray(n * n))
    # A plain device-to-device copy of known size, so that a counter run can calibrate its
    # FETCH_SIZE / WRITE_SIZE readings (the runtime's copy kernel, spread over the whole chip).
    src, dst = device.DeviceArray(CAL_COPY_DOUBLES), device.DeviceArray(CAL_COPY_DOUBLES)
    dst.copy_from(src)
    device.synchronize()
    src.free()
    dst.free()
    if args.stored:
        nq, nbytes = eng.store_eris(1e-10, Q=Q)
        _, st, cms, _ = timed_builds(L, eng, bufs, n, args.steps, args.warmup, lib.HFX_FP64, 0.0, Q,
                                     lib.HFX_JK_STORED | lib.HFX_JK_CLASS_TIMING)
        print(json.dumps({"profile_stored": args.steps + args.warmup, "quartets": nq, "stored_bytes": nbytes,
                          "class_ms": {CLASS_NAMES[i]: round(cms[i], 4) for i in range(21) if cms[i] > 0},
                          "device_ms": st["device_ms"]}), flush=True)
        return
    _, stmx, cm64, cm32 = timed_builds(L, eng, bufs, n, args.steps, args.warmup, lib.HFX_MIXED, args.cut, Q)
    roof = roofline_of(stmx, cm64, cm32)
    per_class = {CLASS_NAMES[i]: [round(cm64[i], 4), round(cm32[i], 4)] for i in range(21) if cm64[i] + cm32[i] > 0}
    print(json.dumps({"profile_builds": args.steps + args.warmup, "fp32_cut": args.cut, "roofline": roof,
                      "class_ms_f64_f32": per_class, "device_ms": stmx["device_ms"]}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--carbons", type=int, default=20)
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--save-density", default=None)
    ap.add_argument("--density", default=None)
    ap.add_argument("--cut", type=float, default=None)
    ap.add_argument("--profile-builds", action="store_true")
    ap.add_argument("--stored", action="store_true")
    ap.add_argument("--profile-gradient", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this engine has no multi-GPU path (the shell-quartet work "
              "list would be sharded with one all-reduce of J/K per build); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.profile_gradient:
        # only the (H2O)8 / 6-31G** gradient, FP64 then mixed, W + K times each with per-class
        # device times (for a profiler's per-kernel averages)
        L = lib.load()
        Z, xyz = geometry.water_cluster_grid(shape=(2, 2, 2), spacing=5.6, jitter=0.4, seed=1002)
        b = Basis("6-31g**", Z, xyz)
        r = lib.scf_run(L, b, options=L.default_options(energy_tol=1e-9, density_tol=1e-7))
        P = 2 * r["D_alpha"]
        W = lib.energy_weighted_density(r["C_alpha"], r["eps_alpha"], sum(b.Z) // 2)
        e = lib.Engine(L, b)
        out = {}
        for name, prec, cut in (("fp64", lib.HFX_FP64, 0.0), ("mixed", lib.HFX_MIXED, 1e-5)):
            ms = []
            for it in range(args.warmup + args.steps):
                _, _, st = lib.rhf_gradient(e, b, P, W, precision=prec, fp32_cut=cut, flags=lib.HFX_JK_CLASS_TIMING)
                if it >= args.warmup:
                    ms.append(st["class_ms"])
            cm = np.mean(ms, axis=0)
            out[name] = {CLASS_NAMES[i]: round(float(cm[i]), 3) for i in range(21) if cm[i] > 0}
            out[name + "_two_electron_ms"] = round(float(np.sum(cm)), 3)
        print(json.dumps({"profile_gradient": args.steps + args.warmup, "quartets": st["n_quartets"], **out}),
              flush=True)
        return
    if args.profile_builds:
        if not args.density or (args.cut is None and not args.stored):
            print("--profile-builds needs --density and --cut (or --stored)", file=sys.stderr)
            sys.exit(2)
        profile_builds(args)
        return

    L = lib.load()
    Z, xyz = geometry.alkane_chain(n_carbon=args.carbons, seed=1003)
    basis = Basis("cc-pvdz-expanded", Z, xyz)
    n = basis.n_aos
    log("C%dH%d/cc-pVDZ: %d shells, %d functions" % (args.carbons, 2 * args.carbons + 2, basis.n_shells, n))
    scf_kw = dict(energy_tol=1e-8, schwarz_thresh=1e-10, density_thresh=1e-10, diis_size=6, max_iter=60)

    r64 = lib.scf_run(L, basis, options=L.default_options(**scf_kw))
    log("FP64 SCF: E=%.10f it=%d conv=%d wall=%.2fs (J/K %.2fs, diag %.2fs)" % (
        r64["e_total"], r64["iterations"], r64["converged"], r64["time_total_s"], r64["time_jk_s"], r64["time_diag_s"]))
    Da = r64["D_alpha"]
    if args.save_density:
        np.save(args.save_density, Da)
        log("saved the converged FP64 density to %s" % args.save_density)
        return

    eng = lib.Engine(L, basis)
    Q = eng.schwarz()
    # the cut: the fastest whose build-level |dE2| against FP64 stays within half the 1e-6 Eh budget
    cut, table = autotune.tune_cut(eng, Da, target=5e-7, Q=Q)
    for row in table:
        log("cut %.0e: |dE2| %.2e, %.3fs, fp32 %d / %d" % row)
    if cut == 0.0:
        cut = 1e-6
    # launch parameters (blocks per CU, claimed run length, streams), tuned for each precision
    launch64, lt64 = autotune.tune_launch(eng, Da, Q=Q, precision=lib.HFX_FP64)
    launchmx, ltmx = autotune.tune_launch(eng, Da, Q=Q, precision=lib.HFX_MIXED, cut=cut)
    log("launch tuning FP64 %s, mixed %s" % (launch64, launchmx))
    for (cfg, ms64_), (_, msmx_) in zip(lt64, ltmx):
        log("  blocks/CU %d, chunk %d, streams %d: FP64 %.2f ms, mixed %.2f ms" % (cfg + (ms64_, msmx_)))
    rmx = lib.scf_run(L, basis, options=L.default_options(precision=lib.HFX_MIXED, fp32_cut=cut, final_fp64=1,
                                                          **scf_kw))
    dE_mixed = abs(rmx["e_total"] - r64["e_total"])
    dE_mixed_stop = abs(rmx["e_total_mixed"] - r64["e_total"])
    dF = float(np.abs(rmx["F_alpha"] - r64["F_alpha"]).max())
    log("mixed SCF (cut %.0e): E=%.10f |dE|=%.2e (at the switch to FP64 %.2e) max|dF|=%.2e it=%d+%d wall=%.2fs" % (
        cut, rmx["e_total"], dE_mixed, dE_mixed_stop, dF, rmx["iterations_mixed"],
        rmx["iterations"] - rmx["iterations_mixed"], rmx["time_total_s"]))

    # the same two SCFs with the iteration's linear algebra on the device (rocSOLVER eigensolver, DIIS
    # and densities in HBM); the first such run in the process loads rocBLAS/rocSOLVER (reported apart)
    r64d = lib.scf_run(L, basis, options=L.default_options(device_linalg=1, **scf_kw), outputs=False)
    rmxd = lib.scf_run(L, basis, options=L.default_options(device_linalg=1, precision=lib.HFX_MIXED, fp32_cut=cut,
                                                           final_fp64=1, jk_flags=lib.HFX_JK_USE_GRAPH, **scf_kw),
                       outputs=False)
    scf_dev = dict(wall_s_fp64=r64d["time_total_s"] - r64d["time_linalg_load_s"],
                   wall_s_mixed=rmxd["time_total_s"] - rmxd["time_linalg_load_s"],
                   linalg_load_s=r64d["time_linalg_load_s"] + rmxd["time_linalg_load_s"],
                   jk_s_mixed=rmxd["time_jk_s"], rest_s_mixed=rmxd["time_diag_s"], init_s_mixed=rmxd["time_init_s"],
                   iterations_fp64=r64d["iterations"],
                   iterations_mixed=[rmxd["iterations_mixed"], rmxd["iterations"] - rmxd["iterations_mixed"]],
                   dE_fp64_vs_host=abs(r64d["e_total"] - r64["e_total"]),
                   dE_mixed_vs_fp64=abs(rmxd["e_total"] - r64["e_total"]),
                   converged=[r64d["converged"], rmxd["converged"]])
    log("device-linalg SCF: FP64 %.2fs (%d it), mixed %.2fs (%d+%d it; J/K %.2fs, rest %.2fs), library load %.2fs, "
        "|dE| mixed vs FP64 %.2e" % (scf_dev["wall_s_fp64"], r64d["iterations"], scf_dev["wall_s_mixed"],
                                      *scf_dev["iterations_mixed"], rmxd["time_jk_s"], rmxd["time_diag_s"],
                                      scf_dev["linalg_load_s"], scf_dev["dE_mixed_vs_fp64"]))

    bufs = (device.DeviceArray.from_host(2 * Da), device.DeviceArray.from_host(Da), device.DeviceArray(n * n),
            device.DeviceArray(n * n))
    plain = lib.HFX_JK_NO_STATS
    graph = lib.HFX_JK_NO_STATS | lib.HFX_JK_USE_GRAPH
    eng.set_launch(*launch64)
    t64e, _, _, _ = timed_builds(L, eng, bufs, n, args.steps, args.warmup, lib.HFX_FP64, 0.0, Q, plain)
    t64, sg64, _, _ = timed_builds(L, eng, bufs, n, args.steps, args.warmup, lib.HFX_FP64, 0.0, Q, graph)
    J64, K64 = bufs[2].download(), bufs[3].download()
    eng.set_launch(*launchmx)
    tmxe, _, _, _ = timed_builds(L, eng, bufs, n, args.steps, args.warmup, lib.HFX_MIXED, cut, Q, plain)
    tmx, sgmx, _, _ = timed_builds(L, eng, bufs, n, args.steps, args.warmup, lib.HFX_MIXED, cut, Q, graph)
    Jmx, Kmx = bufs[2].download(), bufs[3].download()
    log("builds/s: FP64 eager %.3f graph %.3f | mixed eager %.3f graph %.3f" % (
        args.steps / t64e, args.steps / t64, args.steps / tmxe, args.steps / tmx))
    # Per-class times (and the roofline's kernel duration) one class at a time, at the engine's
    # default launch (2 blocks per CU, runs of 2048), the configuration of --profile-builds and of
    # the rocprofv3 summaries: a lone kernel at the tuner's 1 block per CU would leave three of the
    # four blocks per CU that the overlapped build gives it empty.
    eng.set_launch(2, 2048, 4)
    _, st64, c64, _ = timed_builds(L, eng, bufs, n, args.steps, 1, lib.HFX_FP64, 0.0, Q)
    _, stmx, cm64, cm32 = timed_builds(L, eng, bufs, n, args.steps, 1, lib.HFX_MIXED, cut, Q)
    eng.set_launch(*launchmx)
    dF_build = float(max(np.abs(Jmx - J64).max(), np.abs(Kmx - K64).max()))
    dE_build = abs(float(np.sum(Da.reshape(-1, order="F") * ((Jmx - Kmx) - (J64 - K64)))))
    b64 = args.steps / t64
    bmx = args.steps / tmx
    log("FP64 builds/s %.3f (%.2f ms, device %.2f ms, %.1f model GFLOP, %d quartets)" % (
        b64, 1e3 * t64 / args.steps, st64["device_ms"], st64["model_gflop"], st64["n_quartets"]))
    for i in range(21):
        if st64["class_count"][i]:
            log("  %s: %10d quartets %8.1f GFLOP  fp64 %7.3f ms (%.2f TFLOP/s)" % (
                CLASS_NAMES[i], st64["class_count"][i], st64["class_gflop64"][i], c64[i],
                st64["class_gflop64"][i] / c64[i] if c64[i] > 0 else 0))
    log("mixed builds/s %.3f (%.2f ms, device %.2f ms, fp32 quartets %d)" % (
        bmx, 1e3 * tmx / args.steps, stmx["device_ms"], stmx["n_fp32"]))
    for i in range(21):
        if stmx["class_count"][i]:
            log("  %s: fp64 %7.3f ms %8.1f GFLOP | fp32 %7.3f ms %8.1f GFLOP" % (
                CLASS_NAMES[i], cm64[i], stmx["class_gflop64"][i], cm32[i], stmx["class_gflop32"][i]))

    # dominant kernel group of the headline (mixed) build
    roof = roofline_of(stmx, cm64, cm32)
    roof["timing_launch"] = "one class at a time, 2 blocks/CU, runs of 2048 (the --profile-builds configuration)"
    per_class = {CLASS_NAMES[i]: [round(cm64[i], 3), round(cm32[i], 3)] for i in range(21) if cm64[i] + cm32[i] > 0}

    stored = stored_run(L, basis, eng, bufs, Q, args.steps, args.warmup, J64, K64, Da, scf_kw)
    log("stored: %s" % stored)

    batched = batched_run(L, cut)
    log("batched: %s" % batched)
    grad = gradient_run(L)
    log("gradient: %s" % grad)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(basis, Da, Q, st64["model_gflop"], args.cpu_budget)
        log("cpu baseline: %s" % cpu)

    out = {
        "metric": "Fock builds/s (C20H42/cc-pVDZ, mixed FP32/FP64, |dE| <= 1e-6 Eh after the FP64 restart)",
        "value": bmx,
        "unit": "builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tmx / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": "C%dH%d/cc-pVDZ expanded (%d Cartesian functions), full direct Fock build at the "
                               "converged FP64 density, Schwarz 1e-10, density 1e-10" % (
                                   args.carbons, 2 * args.carbons + 2, n),
                   "fp32_cut": cut, "quartets": st64["n_quartets"], "model_gflop": st64["model_gflop"],
                   "exec_gflop_fp64": st64["exec_gflop"], "exec_gflop_mixed": stmx["exec_gflop"],
                   "launch_fp64": dict(zip(("blocks_per_cu", "chunk_len", "streams"), launch64)),
                   "launch_mixed": dict(zip(("blocks_per_cu", "chunk_len", "streams"), launchmx))},
        "ms_per_build_event_median_mixed": sgmx["device_ms_median"],
        "ms_per_build_event_median_fp64": sg64["device_ms_median"],
        "fp64_builds_per_s": b64,
        "fp64_eager_builds_per_s": args.steps / t64e,
        "mixed_eager_builds_per_s": args.steps / tmxe,
        "speedup_mixed_vs_fp64": bmx / b64,
        "scf_wall_s_fp64": r64["time_total_s"],
        "scf_wall_s_mixed": rmx["time_total_s"],
        "scf_device_linalg": scf_dev,
        "scf_iterations_fp64": r64["iterations"],
        "scf_iterations_mixed": [rmx["iterations_mixed"], rmx["iterations"] - rmx["iterations_mixed"]],
        "e_fp64": r64["e_total"],
        "dE_mixed_scf": dE_mixed,
        "dE_mixed_at_switch": dE_mixed_stop,
        "max_dF_mixed_scf": dF,
        "build_dE_mixed": dE_build,
        "build_max_dJK_mixed": dF_build,
        "class_ms_f64_f32": per_class,
        "class_ms_fp64_build": {CLASS_NAMES[i]: round(c64[i], 3) for i in range(21) if c64[i] > 0},
        "stored": stored,
        "batched": batched,
        "gradient": grad,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
