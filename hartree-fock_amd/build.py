"""Build the HIP engine library ``lib/libhfx_mi355x.so`` for gfx950.

Translation units:
  csrc/engine.hip             host engine, screening / one-electron / symmetrization kernels
  csrc/k_grad.hip             RHF nuclear-gradient kernels (generic two-electron, one-electron, reduction)
  csrc/k_grad_eri.hip (many)  the unrolled two-electron gradient: one object per light quartet class,
                              several (chunks of bra components) per heavier one
  csrc/k_jket.hip (many)      one object per quartet class: the ket-major Coulomb pass of stored builds
  csrc/hfx_scf_gpu.hip        the SCF iteration on the device (rocBLAS/rocSOLVER loaded on first use)
  csrc/k_eri.hip (many)       one object per (quartet class, chunk range): the class kernels
                              are fully unrolled, so they are compiled in parallel slices
  csrc/hfx_scf.cpp            host SCF driver (g++ with OpenMP)
  tests/c_caller/hfx_c_caller a C++ test program calling the library through include/hfx.h (g++)

Usage: python build.py [-j N] [--only engine hfx_scf ...]
  --only recompiles just the named objects (a name prefixes the objects it selects; with a
  trailing '=', as in k_grad=, it selects that object alone) and links them with the existing others.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libhfx_mi355x.so")
INC = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HFX_ARCH", "gfx950")
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++20", "-fPIC", "-I" + INC,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXXFLAGS = ["-O3", "-std=c++20", "-fPIC", "-fopenmp", "-I" + INC, "-Wall"]

ACC_MAX = 36          # must match k_eri.hip
CHUNKS_PER_TU = 4     # chunks per translation unit (one launch each; FP64, FP32 and stored-digest kernels,
                      # each with FP64-atomic and fixed-point accumulation)
GRAD_LMAX_UNROLLED = 3  # gradient classes compiled unrolled (k_grad_eri.hip) with one thread per quartet
# heavier gradient classes compiled unrolled in chunks of bra components, FP64 only (combo:
# components per thread, the most that keeps the kernel free of scratch spills; dpdp and ddpp
# spill a little even with one); dddp and dddd use the generic kernel of k_grad.hip (unrolled,
# dddp took 2.5x less time on the device but 90 minutes of compiler time)
GRAD_CHUNKED = {5: 1, 8: 1, 9: 1, 11: 3, 12: 1, 13: 1, 14: 1, 15: 6, 16: 2, 17: 1, 18: 1}
GRAD_CHUNKS_PER_TU = 3
CHUNKS_PER_TU_HEAVY = 2  # the same for the classes with la+lb+lc+ld >= 6, whose units compile longest


def ncart(l):
    return (l + 1) * (l + 2) // 2


def combo_classes():
    """(combo index, la, lb, lc, ld, number of chunk kernels) for the 21 classes."""
    out = []
    for k in range(21):
        p1 = 0
        while (p1 + 1) * (p1 + 2) // 2 <= k:
            p1 += 1
        p2 = k - p1 * (p1 + 1) // 2

        def pl(pc):
            a = 0
            while (a + 1) * (a + 2) // 2 <= pc:
                a += 1
            return a, pc - a * (a + 1) // 2

        la, lb = pl(p1)
        lc, ld = pl(p2)
        nab, ncd = ncart(la) * ncart(lb), ncart(lc) * ncart(ld)
        cbmax = max(1, ACC_MAX // ncd)
        nch = (nab + cbmax - 1) // cbmax
        out.append((k, la, lb, lc, ld, nch))
    return out


def jobs():
    srcs = []
    srcs.append(("engine", [HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, "engine.hip")]))
    srcs.append(("k_grad", [HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, "k_grad.hip")]))
    srcs.append(("hfx_scf_gpu", [HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, "hfx_scf_gpu.hip")]))
    for (k, la, lb, lc, ld, _) in combo_classes():
        srcs.append(("k_jket_%02d" % k, [HIPCC] + HIPFLAGS + ["-DHFX_COMBO=%d" % k, "-c",
                                                               os.path.join(CSRC, "k_jket.hip")]))
        if la + lb + lc + ld <= GRAD_LMAX_UNROLLED:
            srcs.append(("k_grad_eri_%02d" % k, [HIPCC] + HIPFLAGS + ["-DHFX_GRAD_COMBO=%d" % k, "-c",
                                                                      os.path.join(CSRC, "k_grad_eri.hip")]))
        elif k in GRAD_CHUNKED:
            cb = GRAD_CHUNKED[k]
            nch = (ncart(la) * ncart(lb) + cb - 1) // cb
            for b in range(0, nch, GRAD_CHUNKS_PER_TU):
                srcs.append(("k_grad_eri_%02d_%02d" % (k, b),
                             [HIPCC] + HIPFLAGS + ["-DHFX_GRAD_COMBO=%d" % k, "-DHFX_GRAD_FP64_ONLY",
                                                   "-DHFX_GRAD_CB=%d" % cb,
                                                   "-DHFX_GRAD_CHUNK_BEGIN=%d" % b,
                                                   "-DHFX_GRAD_CHUNKS=%d" % min(GRAD_CHUNKS_PER_TU, nch - b),
                                                   "-c", os.path.join(CSRC, "k_grad_eri.hip")]))
    for (k, la, lb, lc, ld, nch) in combo_classes():
        step = CHUNKS_PER_TU_HEAVY if la + lb + lc + ld >= 6 else CHUNKS_PER_TU
        for b in range(0, nch, step):
            e = min(nch, b + step)
            name = "k_eri_%02d_%02d" % (k, b)
            srcs.append((name, [HIPCC] + HIPFLAGS + ["-DHFX_COMBO=%d" % k, "-DHFX_CHUNK_BEGIN=%d" % b,
                                                     "-DHFX_CHUNK_END=%d" % e, "-c",
                                                     os.path.join(CSRC, "k_eri.hip")]))
    srcs.append(("hfx_scf", ["g++"] + CXXFLAGS + ["-c", os.path.join(CSRC, "hfx_scf.cpp")]))
    return srcs


def _run(name, cmd):
    obj = os.path.join(BUILD, name + ".o")
    t = time.time()
    r = subprocess.run(cmd + ["-o", obj], capture_output=True, text=True)
    msg = "\n".join(x for x in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in x)
    return name, obj, r.returncode, msg, time.time() - t


def build(jobs_n=None, verbose=False, only=None):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    todo = jobs()
    keep = []
    if only:
        # a name matches its object and, unless it ends with '=', every object it prefixes
        def sel(name):
            return any(name == o[:-1] if o.endswith("=") else name.startswith(o) for o in only)
        keep = [os.path.join(BUILD, name + ".o") for name, _ in todo if not sel(name)]
        missing = [o for o in keep if not os.path.exists(o)]
        if missing:
            raise RuntimeError("--only needs a full build first; missing " + ", ".join(missing[:4]))
        todo = [j for j in todo if sel(j[0])]
    n = jobs_n or min(16, os.cpu_count() or 8)
    objs, failed = [], []
    t0 = time.time()
    # the biggest classes first, so the long compiles start early
    first = ("k_eri_2", "k_grad_eri_1", "k_eri_1")
    order = sorted(todo, key=lambda j: (next((i for i, f in enumerate(first) if j[0].startswith(f)), len(first)), j[0]))
    with cf.ThreadPoolExecutor(max_workers=n) as ex:
        futs = [ex.submit(_run, name, cmd) for name, cmd in order]
        for f in cf.as_completed(futs):
            name, obj, rc, msg, dt = f.result()
            if rc != 0:
                failed.append((name, msg))
            else:
                objs.append(obj)
                if verbose:
                    print("  %-16s %6.1fs" % (name, dt), flush=True)
    if failed:
        for name, msg in failed:
            print("---- %s ----\n%s" % (name, msg[-4000:]), file=sys.stderr)
        raise RuntimeError("HIP build failed: " + ", ".join(n for n, _ in failed))
    objs = sorted(objs + keep)
    link = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs + \
        ["-lgomp", "-ldl", "-Wl,-Bsymbolic", "-Wl,--no-undefined"]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(LIB + ".tmp", LIB)
    build_c_caller()
    if verbose:
        print("built %s in %.1fs (%d objects)" % (LIB, time.time() - t0, len(objs)), flush=True)
    return LIB


C_CALLER_SRC = os.path.join(os.path.dirname(HERE), "tests", "c_caller", "hfx_c_caller.cpp")
C_CALLER_BIN = os.path.join(os.path.dirname(HERE), "tests", "c_caller", "hfx_c_caller")


def build_c_caller():
    """The test program that calls the library through include/hfx.h from C++ (g++, no HIP headers)."""
    if not os.path.exists(C_CALLER_SRC):
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + INC, C_CALLER_SRC, "-o", C_CALLER_BIN + ".tmp",
           "-L" + LIB_DIR, "-lhfx_mi355x", "-Wl,-rpath,$ORIGIN/../../hartree-fock_amd/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("C caller build failed:\n" + r.stderr[-4000:])
    os.replace(C_CALLER_BIN + ".tmp", C_CALLER_BIN)
    return C_CALLER_BIN


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--only", nargs="*", default=None)
    a = ap.parse_args()
    build(a.j, verbose=True, only=a.only)
