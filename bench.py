"""Benchmark: AMOEBA water force evaluations per second on one MI355X.

Workload: 4,096 rigid AMOEBA waters (water03 parameters, 12,288 atoms) on a
16^3 lattice with random orientations (seed 4096) in a cubic box of
49.68 A; Ewald multipoles with PME (48^3 grid, order 5, alpha 0.5446) and a
7 A real-space cutoff; mutual induced dipoles by preconditioned CG to
1e-6 D rms.  One step = one full energy + gradient evaluation through the C
ABI (neighbor sort and tiles, frames, real-space multipoles, PME, direct
field, the PCG solve, polarization energy and gradient, torques), with the
coordinates already resident in HBM.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--precision mixed|fp64] [--no-cpu]

Both forms are timed on the same system: the plain FP64 form first (the
reference point), then the mixed form (FP32 pair and tensor arithmetic,
64-bit fixed-point accumulation, FP64 reductions), whose rate is `value`;
the line also carries the FP64 rate, the speedup and the RMS force error of
the mixed result against the FP64 one.  Prints one JSON line.  `roofline` is the dominant kernel's achieved rate:
algorithmic FLOPs per launch (per-pair operation count x pairs inside the
cutoff) over its average launch time, from HIP events on the library's own
stream around every launch in the timed region.  `cpu_baseline` is the
FP64 CPU restatement (the oracle) on the host cores: one single-thread
evaluation of the same system per core, all at once (each core one frame),
run before the GPU is touched.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

WORKLOAD = {"workload": "amoeba-water-4096 (12288 atoms)", "box_A": 49.68, "pme_grid": 48, "pme_order": 5,
            "aewald": 0.5446, "cutoff_A": 7.0, "polarization": "mutual, PCG 1e-6 D", "seed": 4096}

# Algorithmic operation counts per in-cutoff pair (each + - * / sqrt exp erfc
# counts 1; an FMA counts 2), from an operation census of the FP64 kernel
# source; the derivation is kept with the design notes.
PAIR_FLOPS = {
    "mpole_tiles": 552,       # image 6, r2 5, Ewald series 37, 5 terms + energy 70, gradient 150, torques 284
    "pol_field_tiles": 206,   # image 6, r2 5, Thole 28, Ewald+Thole series 68, two fields 84, accumulate 15
    "pol_grad_tiles": 1480,   # image 6, r2 5, Thole 28, series 68, four pair gradients ~330 each after zero folding
}
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def hbm_bytes_per_launch(family: str, n: int, npad: int, nnz: int, grid: tuple, w: int, gw: int = 8):
    """Algorithmic HBM bytes of one launch of the memory-bound kernel families:
    every array the launch needs read once and every output written once.
    pol_mutual_spmv: column index (4 B) and the two tensor scalars (w B each)
    per stored entry, the row pointers, the sorted positions (3 x 8 B per
    atom, from which each entry's pair vector is rebuilt), both dipole sets
    read and both field sets written (6 x 8 B per atom each).  pme_spread
    (induced-dipole form): coordinates and both dipole sets read (9 x 8 B per
    atom), two grids written (gw B per point: 4 for FP32 grids, else 8).
    pme_fft (one two-set convolution of K0 x K1 x K2 grids, H0 = K0/2 + 1
    half-complex columns): per set the real grid read and the half-complex
    grid written (forward), the half-complex grid read and written (z
    transforms with the kernel, which is read once), the half-complex grid
    read and the real grid written (inverse).  pme_gather_field: both grids
    read, positions read (3 x 8 B per atom), both dipole sets read for the
    self term and both field sets read and written (3 x 8 B per atom each)."""
    K0, K1, K2 = grid
    nreal, ncplx = K0 * K1 * K2, (K0 // 2 + 1) * K1 * K2
    if family == "pol_mutual_spmv":
        return nnz * (4 + 2 * w) + (npad + 1) * 4 + 3 * 8 * npad + 2 * 6 * 8 * npad
    if family == "pme_spread":
        return n * 9 * 8 + 2 * nreal * gw
    if family == "pme_fft":
        return 2 * (2 * nreal * gw + 4 * 2 * ncplx * gw) + ncplx * gw
    if family == "pme_gather_field":
        return 2 * nreal * gw + n * 3 * 8 + 2 * n * 3 * 8 + 2 * 2 * n * 3 * 8
    if family == "pol_spmv_spread":  # the two in one launch (both read the dipoles)
        return (hbm_bytes_per_launch("pol_mutual_spmv", n, npad, nnz, grid, w, gw)
                + hbm_bytes_per_launch("pme_spread", n, npad, nnz, grid, w, gw))
    return None


HBM_FAMILIES = ("pol_mutual_spmv", "pol_spmv_spread", "pme_spread", "pme_fft", "pme_gather_field")


def copy_calibration(mb: int) -> int:
    """One device-to-device copy of `mb` MiB through the HIP runtime (a copy
    kernel that reads and writes each byte once): the reference point that
    scales the HBM counters of a profiling run.  Returns bytes moved."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    nbytes = mb << 20
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)):
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    return 2 * nbytes


def count_pairs(xyz: np.ndarray, box: float, cutoff: float) -> int:
    from scipy.spatial import cKDTree

    tree = cKDTree(np.mod(xyz, box), boxsize=box)
    return int(len(tree.query_pairs(cutoff, output_type="ndarray")))


def build_system():
    from polmdqc_amd import water

    xyz = water.lattice_box(16, 49.68, 4096)
    s = water.water_system(xyz, "water03", box=49.68,
                           ewald={"aewald": 0.5446, "nfft": 48, "order": 5, "cutoff": 7.0})
    return xyz, s


def _oracle_eval(s):
    import oracle

    t = time.perf_counter()
    o = oracle.elec(s, want_grad=True)
    return time.perf_counter() - t, o["iters"]


def cpu_baseline(s) -> dict:
    """The FP64 CPU restatement on the host cores: W single-thread
    evaluations of the same system at once, one per core (W = the cores this
    process may run on, at most 16, the host share of one GPU), each worker a
    forked process; evals/s = W / wall time.  Must run before the GPU is
    initialized in this process (the workers are forked)."""
    import multiprocessing as mp

    cores = min(len(os.sched_getaffinity(0)), 16)
    t = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_oracle_eval, [s] * cores)
    wall = time.perf_counter() - t
    each = [r[0] for r in res]
    return {"value": cores / wall, "unit": "evals/s", "cores": cores, "kind": "port",
            "sample": f"{cores} concurrent single-thread energy+gradient evaluations of the same 12,288-atom box "
                      f"(FP64 oracle, {res[0][1]} PCG iterations; {min(each):.1f}-{max(each):.1f} s each, "
                      f"{wall:.1f} s wall)",
            "per_core_evals_per_s": round(1.0 / (sum(each) / len(each)), 4)}


PIN_FFT = None  # (FP64-form engine, mixed-form engine) from --pin-fft, or None (timed choice)


def pin_fft_values(spec: str):
    """`--pin-fft FP64/GRIDS-ENGINE`, e.g. `library/fp32-direct`: the PME
    transform engine of the FP64 form, then the mixed form's grid precision
    and engine (pmq_set_fft values).  For profiled runs, whose per-launch
    profiler cost would otherwise steer the timed choice away from the one an
    unprofiled run makes."""
    from polmdqc_amd.engine import FFT_AUTO, FFT_DIRECT, FFT_GRIDS_FP32, FFT_GRIDS_FP64, FFT_LIBRARY

    eng = {"auto": FFT_AUTO, "library": FFT_LIBRARY, "direct": FFT_DIRECT}
    f64, mixed = spec.split("/")
    grids, engm = mixed.split("-")
    return eng[f64], eng[engm] | {"fp32": FFT_GRIDS_FP32, "fp64": FFT_GRIDS_FP64}[grids]


def timed(eng, L, terms, prec, steps, warmup):
    """ms per evaluation over `steps` evaluations (no instrumentation), then
    the per-family kernel times from HIP events over a second pass of the
    same evaluations."""
    from polmdqc_amd.engine import profile_enable, profile_read, set_fft

    if PIN_FFT is not None:
        set_fft(eng, PIN_FFT[0] if prec == L.PLAIN_FP64 else PIN_FFT[1])

    for _ in range(warmup):
        eng.compute(L.GRADIENT, terms, prec)
        eng.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        # each evaluation is completed and checked (as a caller that needs
        # the forces every step would)
        eng.compute(L.GRADIENT, terms, prec)
        eng.sync()
    t1 = time.perf_counter()
    res = eng.fetch()
    profile_enable(eng, True)
    eng.sync()
    for _ in range(steps):
        eng.compute(L.GRADIENT, terms, prec)
        eng.sync()
    prof = profile_read(eng)
    profile_enable(eng, False)
    return (t1 - t0) * 1e3 / steps, prof, res


def bench_clusters(eng, L, steps, warmup) -> dict:
    """The batched workload: 512 non-periodic clusters of U{2..64} waters
    (seed 512), no cutoff, mutual PCG to 1e-6 D per cluster with the
    polarizability preconditioner (pcgprec 0: on these clusters the sparse
    one, pairs within usolvcut, takes 5.9 instead of 9.7 iterations but its
    extra near-field pass per iteration costs more, 672 against 830 evals/s
    mixed); energy, forces and parameter gradients of every cluster per
    evaluation (one graph launch), both forms, with the autotuner's shape per
    size bucket; the same without parameter gradients beside it."""
    from polmdqc_amd import water

    xyz, off = water.cluster_batch(512, 512, 2, 64)
    s = water.water_system(xyz, "water03")
    s["pcgprec"] = 0  # the faster preconditioner on these clusters
    eng.set_system(s)
    eng.set_batch(off)
    terms = L.TERM_MPOLE | L.TERM_POLAR
    out = {"workload": "512 AMOEBA water clusters, 2-64 molecules, seed 512", "atoms": int(s["n"])}
    res, pgs = {}, {}
    for pgrad in (True, False):
        eng.set_batch_param_grad(pgrad)
        tag = "" if pgrad else "_no_param_grad"
        for name, prec in (("fp64", L.PLAIN_FP64), ("mixed", L.MIXED)):
            for _ in range(warmup):
                eng.compute(L.GRADIENT, terms, prec)
                eng.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                eng.compute(L.GRADIENT, terms, prec)
                eng.sync()
            ms = (time.perf_counter() - t0) * 1e3 / steps
            if pgrad:
                res[name] = eng.fetch()
                pgs[name] = eng.param_grad()
                out[f"tuned_shapes_{name}"] = eng.batch_tuning(prec)
            if not np.all(eng.fetch_batch()["pcg_iters"] > 0):
                raise RuntimeError("bench: a cluster's induced dipoles did not converge")
            out[f"{name}_evals_per_s{tag}"] = round(1e3 / ms, 3)
            out[f"{name}_ms_per_eval{tag}"] = round(ms, 4)
    d = res["mixed"]["desum"] - res["fp64"]["desum"]
    out["rms_force_err_mixed_vs_fp64"] = float(np.sqrt(np.mean(np.sum(d * d, axis=1))))
    dp = pgs["mixed"]["dpole"] - pgs["fp64"]["dpole"]
    out["max_dpole_err_mixed_vs_fp64_rel"] = float(np.abs(dp).max() / np.abs(pgs["fp64"]["dpole"]).max())
    out["speedup_mixed_vs_fp64"] = round(out["fp64_ms_per_eval"] / out["mixed_ms_per_eval"], 3)
    eng.set_batch_param_grad(False)
    eng.set_batch(None)
    return out


def bench_hippo(eng, L, steps, warmup) -> dict:
    """The HIPPO workload: 32,000 waters (hippo19, 96,000 atoms) on a 32^3
    lattice with 768 sites left empty (seed 32000), L = 98.58 A; multipole and
    dispersion PME on one 100^3 grid of order 5, Ewald coefficient 0.5446 for
    both (the erfc criterion of kewald.cpp:286-318 at 7 A), 7 A real-space
    cutoff, repulsion to 6 A (switched from 5.4 A), charge penetration and
    HIPPO polarization to 1e-6 D.  Energy + forces of all four terms per
    evaluation, both forms, with per-term energy errors of the mixed form."""
    from polmdqc_amd import water

    xyz = water.lattice_box(32, 98.58, 32000, 32000)
    s = water.water_system(xyz, "hippo19", box=98.58,
                           ewald={"aewald": 0.5446, "nfft": 100, "order": 5, "cutoff": 7.0, "dewald": 0.5446})
    eng.set_system(s)
    terms = L.TERM_MPOLE | L.TERM_POLAR | L.TERM_REPEL | L.TERM_DISP
    out = {"workload": "hippo-water-32000 (96000 atoms)", "box_A": 98.58, "pme_grid": 100, "pme_order": 5,
           "aewald": 0.5446, "dewald": 0.5446, "cutoff_A": 7.0, "repcut_A": 6.0, "seed": 32000}
    ms64, prof64, r64 = timed(eng, L, terms, L.PLAIN_FP64, steps, warmup)
    ms32, prof32, r32 = timed(eng, L, terms, L.MIXED, steps, warmup)
    out["fp64_evals_per_s"] = round(1e3 / ms64, 3)
    out["mixed_evals_per_s"] = round(1e3 / ms32, 3)
    out["speedup_mixed_vs_fp64"] = round(ms64 / ms32, 3)
    d = r32["desum"] - r64["desum"]
    out["rms_force_err_mixed_vs_fp64"] = float(np.sqrt(np.mean(np.sum(d * d, axis=1))))
    out["energy_kcal_fp64"] = {k: r64[k] for k in ("em", "ep", "er", "edsp")}
    out["energy_abs_err_mixed"] = {k: abs(r32[k] - r64[k]) for k in ("em", "ep", "er", "edsp")}
    out["energy_rel_err_mixed"] = {k: abs(r32[k] - r64[k]) / abs(r64[k]) for k in ("em", "ep", "er", "edsp")}
    out["pcg_iters"] = {"fp64": r64["pcg_iters"], "mixed": r32["pcg_iters"]}
    out["kernel_ms_per_step_fp64"] = {k: round(v[0] / steps, 4) for k, v in prof64.items() if v[1]}
    out["kernel_ms_per_step_mixed"] = {k: round(v[0] / steps, 4) for k, v in prof32.items() if v[1]}
    from polmdqc_amd.engine import fft_choice
    out["pme_transforms"] = fft_choice(eng)
    return out


# Kernels of each timed family (demangled-name prefixes, after "void pmq::"),
# for the HBM counter summary; every one runs once per launch of its family.
FAMILY_KERNELS = {
    "pol_spmv_spread": ("k_spread_spmv<",),
    "pme_fft": ("k_fft_planes_fwd<", "k_fft_columns<", "k_fft_planes_inv<"),
    "pme_gather_field": ("k_pme_gather_field<",),
    "pme_spread": ("k_pme_spread_brick<5, 1,",),
    "mpole_tiles": ("k_mpole_tiles<",),
    "pol_field_tiles": ("k_pol_field_tiles<",),
    "pol_grad_tiles": ("k_pol_grad_tiles<",),
    "pol_pair_fill": ("k_pair_fill<",),
}

PROFILE_COMMANDS = {
    "trace": "rocprofv3 --kernel-trace --stats --output-format csv -d gpurun_out/prof -o bench -- python3 bench.py "
             "--steps 5 --warmup 2 --no-cpu --parts amoeba --pin-fft library/fp32-direct > gpurun_out/prof_amoeba.log 2>&1",
    "fetch": "rocprofv3 --pmc FETCH_SIZE --output-format csv -d gpurun_out/pmcF -o p -- python3 bench.py --steps 2 "
             "--warmup 1 --no-cpu --parts amoeba --calibrate-copy 256 --pin-fft library/fp32-direct > gpurun_out/pmcF.log 2>&1",
    "write": "rocprofv3 --pmc WRITE_SIZE --output-format csv -d gpurun_out/pmcW -o p -- python3 bench.py --steps 2 "
             "--warmup 1 --no-cpu --parts amoeba --calibrate-copy 256 --pin-fft library/fp32-direct > gpurun_out/pmcW.log 2>&1",
    "hippo": "rocprofv3 --kernel-trace --stats --output-format csv -d gpurun_out/profh -o hippo -- python3 bench.py "
             "--steps 3 --warmup 1 --no-cpu --parts hippo --pin-fft library/fp32-library > gpurun_out/prof_hippo.log 2>&1",
}


def _short(name: str) -> str:
    return name.split("(")[0].replace("void ", "").replace("pmq::", "")


def _last_json(path: str):
    with open(path) as fh:
        lines = [ln for ln in fh if ln.startswith("{")]
    return json.loads(lines[-1]) if lines else None


def summarize_profiles(src: str, dst: str) -> None:
    """Turn one profiling pass (the four commands of PROFILE_COMMANDS, run
    from the repository root with their output under `src`) into the
    summaries under `dst`: kernel time of the AMOEBA and HIPPO workloads,
    HBM bytes per launch of every kernel, and traffic.json (bytes per launch
    of each timed family, which the bench line reports as roofline.traffic)."""
    import csv
    from collections import defaultdict

    cmds = "\n".join(["export TMPDIR=/tmp"] + [PROFILE_COMMANDS[k] for k in ("trace", "fetch", "write", "hippo")])
    cmds += f"\npython3 bench.py --summarize-profiles {src}"

    def stats_table(path, top):
        rows = list(csv.DictReader(open(path)))
        rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
        out = ["| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
        for r in rows[:top]:
            out.append(f"| `{_short(r['Name'])[:90]}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                       f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
        return out, {_short(r["Name"]): (int(r["Calls"]), float(r["AverageNs"]) / 1e3) for r in rows}

    # kernel time, AMOEBA 12,288 atoms
    amo = _last_json(os.path.join(src, "prof_amoeba.log"))
    tab, avg = stats_table(os.path.join(src, "prof", "bench_kernel_stats.csv"), 40)
    dom = amo["roofline"]["kernel"] if amo else None
    fam_us = ""
    if dom in FAMILY_KERNELS:
        parts = [(k, v[1]) for k, v in avg.items() if k.startswith(FAMILY_KERNELS[dom]) and "float" in k]
        fam_us = (f"\n\nThe roofline family `{dom}` (mixed form), average kernel time per launch from this trace: "
                  + ", ".join(f"`{k}` {us:.2f} us" for k, us in parts)
                  + f"; sum {sum(us for _, us in parts):.2f} us per family launch (the library's HIP events, which "
                  f"also count each launch's dispatch: {amo['roofline']['avg_launch_us']} us).")
    md = ["# Kernel time, AMOEBA water 12,288 atoms (bench workload), FP64 then mixed", "",
          "Commands (one MI355X, from the repository root; this file is the last command's output):", "",
          "```", cmds, "```", "",
          "The trace covers 2+5 FP64 evaluations, then 2+5 mixed evaluations, each set twice (timed, then with the "
          "library's own HIP-event timing on).  PME transforms pinned (`--pin-fft`) to an unprofiled run's choice: "
          "FP64 form on rocFFT, mixed form on FP32 grids with the direct LDS transforms."]
    if amo:
        md += ["", f"Profiled run's bench line: {amo['value']} mixed evals/s, {amo['fp64_evals_per_s']} FP64 evals/s "
               f"(kernel tracing slows the launch-bound parts); PCG iterations {amo['pcg_iters']}; roofline family "
               f"`{dom}`."]
    md += ["", "Top kernels by total time (`bench_kernel_stats.csv`):", ""] + tab
    with open(os.path.join(dst, "amoeba_water_12288_kernel_stats.md"), "w") as fh:
        fh.write("\n".join(md) + fam_us + "\n")

    # kernel time, HIPPO 96,000 atoms
    hp = _last_json(os.path.join(src, "prof_hippo.log"))
    tab, _ = stats_table(os.path.join(src, "profh", "hippo_kernel_stats.csv"), 30)
    md = ["# Kernel time, HIPPO water 96,000 atoms (the HIPPO bench workload), FP64 then mixed", "",
          "Commands: as for `amoeba_water_12288_kernel_stats.md`; this file comes from", "", "```",
          "export TMPDIR=/tmp", PROFILE_COMMANDS["hippo"], f"python3 bench.py --summarize-profiles {src}", "```", "",
          "Multipoles with charge penetration, HIPPO polarization, repulsion and dispersion (PME 100^3, order 5); "
          "transforms pinned to rocFFT for both forms, FP32 grids in the mixed form."]
    if hp and "hippo_96k" in hp:
        h = hp["hippo_96k"]
        md += ["", f"Profiled run: {h['fp64_evals_per_s']} FP64 / {h['mixed_evals_per_s']} mixed evals/s; "
               f"PCG iterations {h['pcg_iters']}."]
    md += ["", "Top kernels by total time (`hippo_kernel_stats.csv`):", ""] + tab
    with open(os.path.join(dst, "hippo_water_96000_kernel_stats.md"), "w") as fh:
        fh.write("\n".join(md) + "\n")

    # HBM bytes per launch from the two counter runs
    def counters(path, counter):
        per = defaultdict(list)  # kernel -> [(grid, KiB)]
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                per[_short(r["Kernel_Name"])].append((int(r["Grid_Size"]), float(r["Counter_Value"])))
        return per

    fetch = counters(os.path.join(src, "pmcF", "p_counter_collection.csv"), "FETCH_SIZE")
    write = counters(os.path.join(src, "pmcW", "p_counter_collection.csv"), "WRITE_SIZE")
    copy_bytes = 256 << 20
    cf = max(fetch["__amd_rocclr_copyBuffer"])  # the 256 MiB calibration copy has the largest grid
    cw = max(write["__amd_rocclr_copyBuffer"])
    f_fac = copy_bytes / (cf[1] * 1024.0)
    w_fac = copy_bytes / (cw[1] * 1024.0)

    def modal(vals):  # dispatches of the most common grid size (one workload shape)
        grids = defaultdict(list)
        for g, v in vals:
            grids[g].append(v)
        g = max(grids, key=lambda k: len(grids[k]))
        return sum(grids[g]) / len(grids[g]), len(grids[g])

    per_kernel = {}
    for k in fetch:
        if k in write and k != "__amd_rocclr_copyBuffer":
            fk, nf = modal(fetch[k])
            wk, _ = modal(write[k])
            per_kernel[k] = (nf, fk, wk, (fk * f_fac + wk * w_fac) * 1024.0)
    traffic = {"fp64": {}, "mixed": {}, "kernels": {}}
    for fam, pref in FAMILY_KERNELS.items():
        for form in ("fp64", "mixed"):
            # the mixed form's kernels are the ones instantiated on float
            ks = sorted(k for k in per_kernel if k.startswith(pref) and ("float" in k) == (form == "mixed"))
            if not ks:
                continue
            # variants of one kernel (e.g. with and without fused vector steps): their mean per dispatch
            by_base = defaultdict(list)
            for k in ks:
                by_base[pref[[k.startswith(p) for p in pref].index(True)]].append(per_kernel[k][3])
            traffic[form][fam] = int(round(sum(sum(v) / len(v) for v in by_base.values())))
            traffic["kernels"][f"{fam}/{form}"] = ks
    traffic["note"] = (f"HBM bytes per family launch: FETCH_SIZE x {f_fac:.3f} + WRITE_SIZE x {w_fac:.3f} "
                       f"(KiB x 1024), the factors that make a 256 MiB device-to-device copy read and write 256 MiB; "
                       f"dispatches of each kernel's most common grid size")
    with open(os.path.join(dst, "traffic.json"), "w") as fh:
        json.dump(traffic, fh, indent=1)
    md = ["# HBM traffic per launch, AMOEBA water 12,288 atoms (bench workload)", "",
          "Commands: the `fetch` and `write` counter runs and the summary step of", "", "```", "export TMPDIR=/tmp",
          PROFILE_COMMANDS["fetch"], PROFILE_COMMANDS["write"], f"python3 bench.py --summarize-profiles {src}", "```",
          "", f"Calibration: the 256 MiB device-to-device copy (`__amd_rocclr_copyBuffer`) reads FETCH_SIZE = "
          f"{cf[1]:.0f} KiB and writes WRITE_SIZE = {cw[1]:.0f} KiB, so FETCH_SIZE is scaled by {f_fac:.3f} and "
          f"WRITE_SIZE by {w_fac:.3f}.  Each kernel's figures are the mean over its dispatches of the most common "
          f"grid size.", "", "Per family launch (`traffic.json`, the bench line's `roofline.traffic`): "
          + "; ".join(f"{fam} {form} {v / 1e6:.2f} MB" for form in ("fp64", "mixed")
                      for fam, v in traffic[form].items()) + ".", "",
          "| kernel | dispatches | FETCH KiB (raw) | WRITE KiB (raw) | HBM MB/launch (scaled) |",
          "|---|---:|---:|---:|---:|"]
    for k, (nd, fk, wk, b) in sorted(per_kernel.items(), key=lambda kv: -kv[1][3] * kv[1][0])[:40]:
        md.append(f"| `{k[:90]}` | {nd} | {fk:.1f} | {wk:.1f} | {b / 1e6:.2f} |")
    with open(os.path.join(dst, "amoeba_water_12288_hbm_traffic.md"), "w") as fh:
        fh.write("\n".join(md) + "\n")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pin-fft", default=None, metavar="FP64/GRIDS-ENGINE",
                    help="fix the PME transform choice (e.g. library/fp32-direct) instead of timing it")
    ap.add_argument("--parts", default="amoeba,clusters,hippo",
                    help="workloads to run (profiling runs take one at a time); the JSON line needs amoeba")
    ap.add_argument("--pair-form", choices=["cartesian", "qi"], default="cartesian",
                    help="multipole pair algebra: Cartesian or in each pair's quasi-internal frame")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="also run one device-to-device copy kernel of MB MiB (HBM counter calibration)")
    ap.add_argument("--summarize-profiles", default=None, metavar="DIR",
                    help="write profiles/*.md and profiles/traffic.json from the rocprofv3 output under DIR and exit")
    args = ap.parse_args()
    if args.summarize_profiles:
        summarize_profiles(args.summarize_profiles, os.path.join(ROOT, "profiles"))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        return 2

    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import PAIR_CARTESIAN, PAIR_QI, fft_choice, set_pair_form
    from polmdqc_amd.engine import Engine as _Engine

    def Engine(dev):
        e = _Engine(dev)
        set_pair_form(e, PAIR_QI if args.pair_form == "qi" else PAIR_CARTESIAN)
        return e

    global PIN_FFT
    if args.pin_fft:
        PIN_FFT = pin_fft_values(args.pin_fft)

    parts = set(args.parts.split(","))
    if "amoeba" not in parts:
        # a profiling run of the other workloads alone
        eng = Engine(0)
        out = {}
        if "clusters" in parts:
            out["batched_clusters"] = bench_clusters(eng, L, args.steps, args.warmup)
        if "hippo" in parts:
            out["hippo_96k"] = bench_hippo(eng, L, args.steps, args.warmup)
        eng.close()
        if args.calibrate_copy:
            out["calibration_copy_bytes"] = copy_calibration(args.calibrate_copy)
        print(json.dumps(out))
        return 0

    xyz, s = build_system()
    # the CPU baseline first: its workers are forked before this process
    # touches the GPU
    cpu = cpu_baseline(s) if not args.no_cpu else None
    terms = L.TERM_MPOLE | L.TERM_POLAR
    eng = Engine(0)
    eng.set_system(s)
    # the FP64 form is always run: it is the reference point of the mixed one
    ms64, prof64, res64 = timed(eng, L, terms, L.PLAIN_FP64, args.steps, args.warmup)
    if args.precision == "mixed":
        ms, prof, res = timed(eng, L, terms, L.MIXED, args.steps, args.warmup)
    else:
        ms, prof, res = ms64, prof64, res64
    value = 1e3 / ms

    pairs = count_pairs(xyz, 49.68, 7.0)
    mixed = args.precision == "mixed"
    n = len(xyz)
    npad = (n + 63) // 64 * 64
    w = 4 if mixed else 8
    # the dominant kernel family: largest total time among those with an
    # algorithmic model (pair flops, or HBM bytes)
    cand = {k: v for k, v in prof.items() if v[1] > 0 and (k in PAIR_FLOPS or k in HBM_FAMILIES)}
    dom = max(cand, key=lambda k: cand[k][0])
    tot_ms, launches = cand[dom]
    avg_s = tot_ms / launches * 1e-3
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            tr = json.load(fh)
        traffic = tr.get(args.precision, {}).get(dom)
    if dom in PAIR_FLOPS:
        peak = PEAK_FP32_TFLOPS if mixed else PEAK_FP64_TFLOPS
        achieved = PAIR_FLOPS[dom] * pairs / avg_s / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                    "frac": round(achieved / peak, 4), "traffic": traffic, "kernel": dom,
                    "avg_launch_us": round(avg_s * 1e6, 2), "flops_per_launch": PAIR_FLOPS[dom] * pairs,
                    "note": "VALU pair kernel (no MFMA: pair terms are not a dense contraction); vector peak"}
    else:
        gw = 4 if mixed and fft_choice(eng)["mixed_grids"] == "fp32" else 8
        nbytes = hbm_bytes_per_launch(dom, n, npad, 2 * pairs, (48, 48, 48), w, gw)
        achieved = nbytes / avg_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic, "kernel": dom,
                    "avg_launch_us": round(avg_s * 1e6, 2), "bytes_per_launch": nbytes}
    frms = float(np.sqrt(np.mean(np.sum((res["desum"] - res64["desum"]) ** 2, axis=1))))
    line = {
        "metric": "AMOEBA 12k-atom water: force evals/s mixed vs FP64 GPU @ rms force err<1e-3",
        "value": round(value, 3), "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+i64" if mixed else "f64", "data": "synthetic (seeded water lattice)",
        "config": dict(WORKLOAD, precision=args.precision, pair_form=args.pair_form),
        "roofline": roofline,
        "fp64_evals_per_s": round(1e3 / ms64, 3),
        "speedup_vs_fp64": round(ms64 / ms, 3),
        "rms_force_err_vs_fp64": frms,
        "energy_rel_err_vs_fp64": {k: abs(res[k] - res64[k]) / abs(res64[k]) for k in ("em", "ep")},
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in prof.items() if v[1]},
        "kernel_launches_per_step": {k: round(v[1] / args.steps, 2) for k, v in prof.items() if v[1]},
        "kernel_us_per_launch": {k: round(v[0] * 1e3 / v[1], 2) for k, v in prof.items() if v[1]},
        "kernel_ms_per_step_fp64": {k: round(v[0] / args.steps, 4) for k, v in prof64.items() if v[1]},
        "pairs_in_cutoff": pairs, "pcg_iters": res["pcg_iters"],
        "pme_transforms": fft_choice(eng),
        "energy_kcal": {"em": res["em"], "ep": res["ep"]},
    }
    if "clusters" in parts:
        line["batched_clusters"] = bench_clusters(eng, L, args.steps, args.warmup)
    if "hippo" in parts:
        line["hippo_96k"] = bench_hippo(eng, L, args.steps, args.warmup)
    line["cpu_baseline"] = cpu
    eng.close()
    if args.calibrate_copy:
        line["calibration_copy_bytes"] = copy_calibration(args.calibrate_copy)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
