#!/usr/bin/env python3
"""Benchmark of the Chebyshev-expansion path on one MI355X.

Default workload (one "step" = one LDOS evaluation at the lattice centre,
summed over spin, i.e. two diagonal Green's functions G_ii):
  2D square lattice 1024 x 1024 with spin (2,097,152 states, 10,477,568
  non-zeros, complex128 CSR), t = 1, Zeeman J = 0.5, on-site disorder
  U[-1, 1] (seed 2), scale factor 12, 4000 Chebyshev coefficients,
  Lorentz damping 1e-6 (the reference's default), 2000 energies on
  [-11.88, 11.88], plain FP64 path (one start vector per recurrence),
  Hamiltonian resident in HBM.
`value` = LDOS sites/s, one site = one diagonal element G_ii to 4000 moments
contracted to 2000 energies.

The same line carries the batched comparison ("batched FP32 vs plain FP64
at fixed error"): 512 x 512 Nambu s-wave superconductor (1,048,576 states),
256 LDOS sites drawn by seed as one block of unit start vectors, 2000
coefficients, FP32 storage with FP64 accumulation and the doubling identities,
against the plain FP64 path on the same model; error = max over compared sites
of max_E |LDOS - LDOS_FP64| / max_E |LDOS_FP64|.

Every configuration carries a "cpu_baseline": the reference's own CPU path
(compiled from its sources, oracle/_ref/ref_driver) timed on a bounded sample
of the same workload on this host, on one core and in one process per core
(the throughput mode), with the GPU/CPU ratios beside it.

Further objects on the same line: "uneven_batch" (64 models of uneven size
in one launch per step), "stochastic_dos" (128^3 cubic, 512
random-phase vectors, M=4096) and "contraction_mfma" (1e5 x 4096 x 8192
split-BF16 Green's-function GEMM), each with its own roofline fraction and
error against the FP64 path.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
                       [--parts plain,batched,uneven,dos,mfma,cpu]
                       [--skip-cpu-baseline] [--calibrate-copy] [--complex] [--no-palette]
                       [--no-patterns | --row-pattern-bytes] [--slot-rows R]
                       [--no-graph] [--uneven-moments-div D] [--uneven-tile-rows G]
                       [--uneven-grid-share S] [--dos-moments-div D]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tbtk_amd import engine as X  # noqa: E402
from tbtk_amd import kpm  # noqa: E402
from tbtk_amd import lattice as L  # noqa: E402

METRIC = "LDOS sites/s + SpMV GB/s, batched FP32 vs plain FP64 at fixed error"
HBM_PEAK_GBPS = 8000.0
BF16_PEAK_TFLOPS = 2500.0

# Default workload
SCALE = 12.0
M_PLAIN = 4000
E_PTS = 2000
WINDOW = (-0.99 * SCALE, 0.99 * SCALE)
BROADENING = 1e-6
CENTRE = (512, 512)

# Batched comparison
M_BATCH = 2000
B_SITES = 256
PLAIN_CMP_SITES = 16

PARTS = ("plain", "batched", "uneven", "dos", "mfma", "cpu")


def rel_errors(got, ref):
    """Error of got against ref, both measures the reports carry: max|got - ref|
    over max|ref|, and the largest pointwise |got - ref| / |ref| over the points
    where |ref| > 1e-3 max|ref|."""
    got, ref = np.asarray(got), np.asarray(ref)
    d, a = np.abs(got - ref), np.abs(ref)
    big = a > 1e-3 * a.max()
    return float(d.max() / a.max()), float((d[big] / a[big]).max())


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def plain_ldos(eng, model, site_xy, M):
    """LDOS at a site summed over spin: per spin one plain FP64 recurrence,
    Lorentz damping on the host, FP64 NonPrincipal contraction on the GPU."""
    ldos = np.zeros(E_PTS)
    kernel_ms = 0.0
    launches = 0
    for s in (0, 1):
        b = model.basis_index((*site_xy, s))
        c = eng.moments(b, [b], M)
        ms, n = eng.last_timing()
        kernel_ms += ms
        launches += n
        c = kpm.lorentz(c, BROADENING)
        ldos += kpm.ldos_from_gf(eng.contract(c, X.NONPRINCIPAL))
    return ldos, kernel_ms, launches


def cpu_info():
    """CPU model and logical CPU count of this host, and the worker count the
    throughput mode uses (the OMP_NUM_THREADS share, at most 16)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", nproc) or nproc)
    return model, nproc, max(1, min(16, share, nproc))


def ref_runs(procs, **kw):
    """Start `procs` copies of the compiled reference driver (oracle/_ref/
    ref_driver, one OpenMP thread each) with the same arguments at once and
    return each one's JSON timing lines."""
    from oracle import oracle as O
    args = [O.REF_DRIVER] + [f"{k}={v}" for k, v in kw.items()]
    env = dict(os.environ, OMP_NUM_THREADS="1")
    ps = [subprocess.Popen(args, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True,
                           env=env) for _ in range(procs)]
    rows = []
    for p in ps:
        _, err = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"reference driver failed ({p.returncode}): {err[-500:]}")
        rows.append([json.loads(x) for x in err.splitlines() if x.startswith("{")])
    return rows


def ref_ldos_fit(path, scale, E, patterns, Ms, procs):
    """calculateLDOS(patterns) of the reference, timed at the moment counts
    Ms in each of `procs` concurrent processes: per process (fixed, per
    moment) seconds of a linear fit."""
    fits = []
    for rows in ref_runs(procs, mode="timeldos", model=path, scale=scale, M=Ms[0],
                         broadening=BROADENING, window=f"{-0.99 * scale},{0.99 * scale},{E}",
                         lookup=1, patterns=patterns, Ms=",".join(map(str, Ms))):
        t = {int(x["M"]): x["seconds"] for x in rows}
        slope = (t[Ms[1]] - t[Ms[0]]) / (Ms[1] - Ms[0])
        fits.append((t[Ms[0]] - Ms[0] * slope, slope, t))
    return fits


def baseline_ldos(model, scale, E, patterns, sites_per_call, M_target, label, procs,
                  workdir, Ms=(20, 60), factor=None):
    """Reference CPU baseline of an LDOS workload: one call of
    PropertyExtractor::ChebyshevExpander::calculateLDOS(patterns)
    (Lib/src/PropertyExtractor/ChebyshevExpander.cpp:455-505; one sequential
    recurrence per site, Lib/src/Solver/ChebyshevExpander.cpp:103-199) at two
    small moment counts, extrapolated linearly to M_target, on one core. The
    throughput mode (one process per core) is measured by running the same
    sample in `procs` processes at once, unless `factor` (aggregate over
    single-core rate, measured on another configuration) is given."""
    path = os.path.join(workdir, label + ".bin")
    L.write_amplitudes(model, path)
    cpu, nproc, _ = cpu_info()
    (f1, s1, t1), = ref_ldos_fit(path, scale, E, patterns, Ms, 1)
    t_call = f1 + M_target * s1
    single = sites_per_call / t_call
    if factor is None:
        # the parallel samples at the larger moment count, against the
        # single-process one: each process's slowdown under full load
        par = ref_ldos_fit(path, scale, E, patterns, Ms, procs)
        factor = sum(t1[Ms[1]] / t[Ms[1]] for _, _, t in par)
        mode = f"{procs} processes at once running the same sample"
    else:
        mode = (f"{procs} processes at once: the aggregate/single-core factor {factor:.2f} "
                "measured on the 1024x1024 spin lattice sample")
    return {"value": single, "unit": "LDOS sites/s", "cores": 1,
            "kind": "reference",
            "sample": (f"reference calculateLDOS({patterns}) on the same model, "
                       f"M={Ms[0]} ({t1[Ms[0]]:.2f} s) and M={Ms[1]} ({t1[Ms[1]]:.2f} s), "
                       f"1 thread, extrapolated linearly to M={M_target}: {t_call:.1f} s per "
                       f"call of {sites_per_call} site(s)"),
            "seconds_per_call": t_call,
            "throughput": {"value": single * factor, "unit": "LDOS sites/s", "cores": procs,
                           "aggregate_over_single": factor, "mode": mode},
            "cpu_model": cpu, "nproc": nproc}


def baseline_uneven(batch, procs, workdir, factor):
    """Reference CPU time of the uneven batch: every model's sites run as
    the reference runs them, one calculateLDOS recurrence per site. The
    per-call fixed time and per-moment time are measured (as in
    baseline_ldos) on the models with the fewest and the most non-zeros and
    interpolated linearly in the non-zero count for the others."""
    cpu, nproc, _ = cpu_info()
    order = sorted(range(len(batch)), key=lambda i: batch[i][0].nnz)
    pick = [order[0], order[-1]]
    fit = {}
    for i in pick:
        model, sites, _ = batch[i]
        path = os.path.join(workdir, f"uneven{i}.bin")
        L.write_amplitudes(model, path)
        s = model.gershgorin_bound() * 1.05
        idx = ",".join(map(str, model.physical_index(int(sites[0]))))
        (f1, s1, _), = ref_ldos_fit(path, s, E_PTS, idx, (20, 60), 1)
        fit[i] = (f1, s1)
    (za, (fa, sa)), (zb, (fb, sb)) = [(batch[i][0].nnz, fit[i]) for i in pick]

    def lin(a, b, z):
        return a + (b - a) * (z - za) / (zb - za) if zb != za else a
    total = 0.0
    sites = 0
    for model, st, M in batch:
        z = model.nnz
        total += len(st) * (lin(fa, fb, z) + M * lin(sa, sb, z))
        sites += len(st)
    return {"value": sites / total, "unit": "LDOS sites/s", "cores": 1, "kind": "reference",
            "seconds": total,
            "sample": (f"reference calculateLDOS of one site on the models with the fewest "
                       f"({za}) and most ({zb}) non-zeros at M=20 and 60, 1 thread; per-call "
                       f"and per-moment times interpolated in the non-zero count over the 64 "
                       f"models and summed over their {sites} sites: {total:.0f} s"),
            "throughput": {"value": sites / total * factor, "unit": "LDOS sites/s",
                           "cores": procs, "aggregate_over_single": factor,
                           "mode": (f"{procs} processes at once: the aggregate/single-core "
                                    f"factor measured on the 1024x1024 spin lattice sample")},
            "cpu_model": cpu, "nproc": nproc}


def baseline_contraction(K, M, E, scale, procs, rows=2):
    """Reference CPU time of the Green's-function contraction: the FP64
    lookup-table loop of generateGreensFunctionCPU (Lib/src/Solver/
    ChebyshevExpander.cpp:324-370; one vector per call, the table built
    once) timed for `rows` vectors and extrapolated to K."""
    cpu, nproc, _ = cpu_info()
    kw = dict(mode="timegf", M=M, window=f"{-0.99 * scale},{0.99 * scale},{E}", scale=scale,
              type="Retarded", rows=rows)
    (r1,), = ref_runs(1, **kw)
    per = r1["seconds"] / rows
    par = ref_runs(procs, **kw)
    agg = sum(rows / p[0]["seconds"] for p in par)   # vectors per second, all processes
    return {"value": K * per, "unit": "s", "cores": 1, "kind": "reference",
            "sample": (f"reference lookup-table contraction, {rows} vectors of M={M} on E={E} "
                       f"({r1['seconds']:.2f} s, table built apart in "
                       f"{r1['table_and_first_seconds']:.1f} s), 1 thread, x {K} vectors"),
            "throughput": {"value": K / agg, "unit": "s", "cores": procs,
                           "aggregate_over_single": agg * per,
                           "mode": f"{procs} processes at once, one vector sequence each"},
            "cpu_model": cpu, "nproc": nproc}


def load_traffic(kernel):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    entry = data.get(kernel)
    return None if entry is None else entry.get("hbm_bytes_per_launch")


def plain_step_bytes(eng, N, Z, csr=None, row_pattern_bytes=False):
    """Algorithmic HBM bytes of one launch of the plain FP64 step in the
    form the engine's last plain call ran (options read back from the
    engine), and the flags that name the form."""
    real = eng.get_option(X.OPT_REAL) == 1
    col16 = eng.get_option(X.OPT_COL16) == 1
    palette = eng.get_option(X.OPT_PALETTE) == 1
    slot = eng.get_option(X.OPT_PLAIN_KERNEL) == 2
    patterns = eng.get_option(X.OPT_ROW_PATTERNS) == 1
    # Wave-uniform patterns (the default): only the rows outside 64-row
    # chunks of one shared pattern read their pattern byte.
    uniform = patterns and not row_pattern_bytes
    byte_rows = eng.get_option(X.OPT_PATTERN_BYTE_ROWS) if uniform else N
    cb = 2 if col16 else 4
    e = 8 if real else 16
    if patterns:
        step_bytes = 4 * e * N + byte_rows + (N // 64 if uniform else 0)
    elif slot:
        rowptr = (csr if csr is not None else eng.csr_copy())[0]
        W = max(4, int(np.diff(rowptr).max()))
        step_bytes = W * N * (2 + (1 if palette else e)) + N * (3 * e + (e if palette else 0))
    elif palette:
        step_bytes = 3 * Z + 36 * N if real else 3 * Z + 68 * N
    else:
        step_bytes = (8 + cb) * Z + 28 * N if real else (16 + cb) * Z + 52 * N
    form = {"real": real, "col16": col16, "palette": palette, "slot": slot,
            "patterns": patterns, "uniform": uniform, "byte_rows": byte_rows}
    return step_bytes, form


def batched_comparison(eng):
    """Config: 512x512 Nambu, 256 sites, M=2000; batched FP32 vs plain FP64."""
    # H and its Gershgorin bound (FP64) built on the device from the lattice
    spec = L.config_nambu_spec(512)
    s = eng.lattice_gershgorin(spec) * 1.05
    eng.set_lattice(spec, s)
    N, Z = eng.csr_size()
    # Launch shape (block width, L2 strip, grid share, rows in flight) from
    # the engine's autotuner, which times the candidates on this model once.
    eng.set_option(X.OPT_BLOCK_WIDTH, 0)
    rng = np.random.default_rng(3)
    sites = np.sort(rng.choice(N, B_SITES, replace=False))
    lo, hi = -0.99 * s, 0.99 * s
    eng.lut_generate(M_BATCH, E_PTS, lo, hi, s)
    # Jackson damping in FP64 on the device, applied by the contraction call
    eng.set_damping(X.DAMP_JACKSON)

    def contract(mu):
        return eng.contract(mu, X.NONPRINCIPAL).imag / np.pi

    # warm-up (allocations and the shape autotuner for 256 sites), then timed
    eng.moments_block(sites, 16, X.DOUBLING)
    t0 = time.perf_counter()
    mu = eng.moments_block(sites, M_BATCH, X.DOUBLING)
    ms_rec, launches = eng.last_timing()
    ld_batched = contract(mu)
    t_batched = time.perf_counter() - t0

    packed = eng.get_option(X.OPT_REAL) == 1
    eng.moments(int(sites[0]), [int(sites[0])], 16)
    plain_real = eng.get_option(X.OPT_REAL) == 1
    t0 = time.perf_counter()
    ref = []
    plain_ms = 0.0
    for b in sites[:PLAIN_CMP_SITES]:
        c = eng.moments(int(b), [int(b)], M_BATCH)
        plain_ms += eng.last_timing()[0]
        ref.append(c[0])
    plain_bytes, plain_form = plain_step_bytes(eng, N, Z)
    ld_plain = contract(np.array(ref))
    t_plain = (time.perf_counter() - t0) / PLAIN_CMP_SITES

    errs = [rel_errors(ld_batched[i], ld_plain[i]) for i in range(PLAIN_CMP_SITES)]
    err = max(e[0] for e in errs)
    err_pt = max(e[1] for e in errs)
    eng.set_damping(X.DAMP_NONE)
    B = eng.get_option(X.OPT_BLOCK_WIDTH)
    shape = {"block_width": B, "strip_rows": eng.get_option(X.OPT_STRIP_ROWS),
             "grid_share": eng.get_option(X.OPT_GRID_SHARE),
             "rows_in_flight_x": eng.get_option(X.OPT_ROWS_IN_FLIGHT)}
    step_bytes = 12 * Z + 4 * N + 24 * B * N
    per_step = ms_rec * 1e-3 / launches
    sps_b = B_SITES / t_batched
    sps_p = 1.0 / t_plain
    return {
        "workload": ("nambu s-wave 512x512x4 (1,048,576 states, t=1, mu=-1, Delta=0.3, "
                     "J=0.5, disorder U[-0.5,0.5] seed 3), 256 seeded LDOS sites, "
                     f"M={M_BATCH}, E={E_PTS}, Jackson damping (FP64, on the device), "
                     f"s=Gershgorin*1.05={s:.6f} (H and bound built on the device)"),
        "nnz": int(Z),
        "sites_per_s": sps_b,
        "plain_fp64_sites_per_s": sps_p,
        "speedup_vs_plain_fp64": sps_b / sps_p,
        "max_ldos_rel_err_vs_fp64": float(err),
        "max_ldos_pointwise_rel_err_vs_fp64": err_pt,
        "error_sites_compared": PLAIN_CMP_SITES,
        "batched_dtype": ("fp32 vectors; real H as one fp64 value per 8-byte entry, fp64 row "
                          "sums; fp64 moments, doubling"
                          + (", real H: two real vectors per complex column" if packed else "")),
        "spmm_block_width": B,
        "vectors_per_block": 2 * B if packed else B,
        "spmm_shape": shape,
        "spmm_step_ms": per_step * 1e3,
        "spmm_GBps": step_bytes / per_step / 1e9,
        "spmm_hbm_frac": step_bytes / per_step / 1e9 / HBM_PEAK_GBPS,
        "plain_arithmetic": "real fp64" if plain_real else "complex fp64",
        "plain_form": ("slot-major, wave-uniform row patterns" if plain_form["uniform"] else
                       "slot-major, row patterns" if plain_form["patterns"] else
                       "slot-major" if plain_form["slot"] else "row blocks"),
        "plain_step_us": plain_ms * 1e3 / (PLAIN_CMP_SITES * (M_BATCH - 1)),
        "plain_spmv_GBps": plain_bytes
        / (plain_ms * 1e-3 / (PLAIN_CMP_SITES * (M_BATCH - 1))) / 1e9,
    }


def uneven_batch(eng, graph=True, moments_div=1, tile_rows=0, grid_share=0):
    """64 independent models of uneven size (honeycomb ribbons and square
    lattices, 20k-1M states, 1-64 sites and 500-8000 moments each, seed 4) in
    one launch per recurrence step: FP32 storage, FP64 moments,
    doubling, autotuned launch shape, replayed as hipGraphs. The models
    are built on the device from their lattice descriptions
    (tbtkx_multi_set_lattices). For counter runs, moments_div > 1 divides
    every model's moment count (at least 50: the same phases, a tenth of the
    steps at 10) and tile_rows fixes the tile (row groups per wavefront
    tile) instead of the autotuned shape, on the grid grid_share names (0:
    one tile per wavefront, s > 0: persistent grid of occupancy x CUs / s
    workgroups)."""
    t0 = time.perf_counter()
    batch = L.config_uneven_specs(seed=4)
    if moments_div > 1:
        batch = [(sp, st, max(50, M // moments_div)) for sp, st, M in batch]
    eng.set_option(X.OPT_TILE_ROWS, tile_rows)
    eng.set_option(X.OPT_GRID_SHARE, grid_share if tile_rows > 0 else 0)
    specs = [b[0] for b in batch]
    scales = [eng.lattice_gershgorin(sp) * 1.05 for sp in specs]
    eng.multi_set_lattices(specs, scales)
    t_gen = time.perf_counter() - t0
    sites = [b[1] for b in batch]
    Ms = [b[2] for b in batch]
    N = np.array([sp.basis_size for sp in specs], dtype=np.int64)
    B = np.array([len(x) for x in sites], dtype=np.int64)
    M = np.array(Ms, dtype=np.int64)
    S = (M + 1) // 2
    # ELL slots per row of each model (its longest row, diagonal included,
    # rounded up to a kernel width: 4, 5, 6 or 8)
    W = np.array([min(w for w in (4, 5, 6, 8) if w >= sp.max_row_entries()) for sp in specs],
                 dtype=np.int64)
    # 1: autotune + direct launches; 2: hipGraph capture; 3: graph replay (timed)
    t0 = time.perf_counter()
    mu = eng.multi_moments(sites, Ms, X.DOUBLING)
    t_first = time.perf_counter() - t0
    ms_direct, launches = eng.last_timing()
    shape = {"tile_row_groups": eng.get_option(X.OPT_TILE_ROWS),
             "grid_share": eng.get_option(X.OPT_GRID_SHARE)}
    # complex columns per model: two real vectors per column for a real batch
    packed = eng.get_option(X.OPT_REAL) == 1
    cols = (B + 1) // 2 if packed else B
    step_bytes_total = float(((12 * W * N + 24 * cols * N) * S).sum())
    # graph=False (--no-graph): a second direct run is timed instead of the
    # graph replay (for profilers that do not follow graph launches)
    eng.set_option(X.OPT_GRAPH, 1 if graph else 0)
    try:
        if graph:
            eng.multi_moments(sites, Ms, X.DOUBLING)
        t0 = time.perf_counter()
        mu_g = eng.multi_moments(sites, Ms, X.DOUBLING)
        t_graph = time.perf_counter() - t0
        ms_graph, _ = eng.last_timing()
    finally:
        eng.set_option(X.OPT_GRAPH, 0)
    graph_diff = max(float(np.abs(a - b).max()) for a, b in zip(mu, mu_g))
    # error against the plain FP64 path: the two smallest and the two
    # largest models, their first and last sites
    errs = []
    order = np.argsort(N)
    for i in list(order[:2]) + list(order[-2:]):
        spec, st, Mi = batch[i]
        eng.set_lattice(spec, scales[i])
        for k in sorted({0, len(st) - 1}):
            ref = eng.moments(int(st[k]), [int(st[k])], Mi)[0]
            lo, hi = -0.99 * scales[i], 0.99 * scales[i]
            eng.lut_generate(Mi, E_PTS, lo, hi, scales[i])
            eng.set_damping(X.DAMP_JACKSON)
            l_ref = kpm.ldos_from_gf(eng.contract(ref, X.NONPRINCIPAL))
            l_got = kpm.ldos_from_gf(eng.contract(mu[i][k], X.NONPRINCIPAL))
            eng.set_damping(X.DAMP_NONE)
            errs.append(rel_errors(l_got, l_ref))
    eng.lut_destroy()
    eng.set_option(X.OPT_TILE_ROWS, 0)
    eng.set_option(X.OPT_GRID_SHARE, 0)
    dev_s = ms_graph * 1e-3
    extra = {"launch_shape": shape}
    if moments_div > 1:
        extra["moments_divided_by"] = moments_div
    return {**extra,
        "workload": ("64 models, seed 4: honeycomb ribbons (t=2.7) and square lattices (t=1), "
                     "width 20-400, length 50-2000 (<= 1e6 states), disorder U[-0.5,0.5], "
                     "1-64 LDOS sites and 500-8000 moments per model, Gershgorin*1.05 scales"),
        "models": len(batch), "states": int(N.sum()), "sites": int(B.sum()),

        "ell_slots": {str(w): int((W == w).sum()) for w in (4, 5, 6, 8) if (W == w).any()},
        "recurrence_steps": int(S.max()),
        "dtype": ("fp32 vectors; real H as one fp64 value per 8-byte entry, fp64 row sums; "
                  "fp64 moments, doubling"
                  + (", real H: two real vectors per complex column" if packed else "")),
        "timed_run": "hipGraph replay" if graph else "direct launches (--no-graph)",
        "sites_per_s": float(B.sum()) / (t_graph),
        "device_s": dev_s,
        "direct_device_s": ms_direct * 1e-3,
        "first_call_s_incl_autotune": t_first,
        "graph_replay_wall_s": t_graph,
        "step_launches": int(launches),
        "algorithmic_bytes": step_bytes_total,
        "GBps": step_bytes_total / dev_s / 1e9,
        "hbm_frac": step_bytes_total / dev_s / 1e9 / HBM_PEAK_GBPS,
        "graph_vs_direct_max_abs_diff": graph_diff,
        "max_ldos_rel_err_vs_fp64": max(e[0] for e in errs),
        "max_ldos_pointwise_rel_err_vs_fp64": max(e[1] for e in errs),
        "error_sites_compared": len(errs),
        "model_build_s": t_gen,
        "model_build": "lattice descriptions on the host, CSR/ELL generated on the device",
    }


def stochastic_dos(eng, moments_div=1):
    """128^3 cubic lattice (t=1, disorder U[-1,1], seed 5), R=512 random-phase
    start vectors, M=4096, FP32 storage, FP64 moment sums, doubling. For
    counter runs, moments_div > 1 divides M (per-launch figures unchanged,
    fewer launches)."""
    spec = L.config_cubic_spec(128)
    s = eng.lattice_gershgorin(spec) * 1.05
    eng.set_lattice(spec, s)
    N, Z = eng.csr_size()
    R, M, seed = 512, max(64, 4096 // moments_div), 17
    eng.set_option(X.OPT_BLOCK_WIDTH, 0)
    eng.moments_trace(R, seed, 8)  # allocations and the shape autotuner
    log("dos: shape chosen, timed run")
    t0 = time.perf_counter()
    mu = eng.moments_trace(R, seed, M)
    wall = time.perf_counter() - t0
    log(f"dos: timed run done ({wall:.2f} s), FP32/FP64 error reference")
    ms, launches = eng.last_timing()
    Bw = eng.get_option(X.OPT_BLOCK_WIDTH)
    shape = {"block_width": Bw, "strip_rows": eng.get_option(X.OPT_STRIP_ROWS),
             "grid_share": eng.get_option(X.OPT_GRID_SHARE),
             "rows_in_flight_x": eng.get_option(X.OPT_ROWS_IN_FLIGHT)}
    # error: 8 vectors in FP32 storage vs the same 8 in FP64 (H and vectors)
    Rc = 8
    a = eng.moments_trace(Rc, seed, M)
    b = eng.moments_trace(Rc, seed, M, X.VEC_FP64 | X.H_FP64)
    eng.lut_generate(M, E_PTS, -0.99 * s, 0.99 * s, s)
    eng.set_damping(X.DAMP_JACKSON)
    dos_a = kpm.ldos_from_gf(eng.contract(a, X.NONPRINCIPAL))
    dos_b = kpm.ldos_from_gf(eng.contract(b, X.NONPRINCIPAL))
    eng.set_damping(X.DAMP_NONE)
    eng.lut_destroy()
    blocks = (R + Bw - 1) // Bw
    step_bytes = 12 * Z + 4 * N + 24 * Bw * N
    per_step = ms * 1e-3 / launches
    extra = {"moments_divided_by": moments_div} if moments_div > 1 else {}
    return {
        **extra,
        "workload": (f"cubic 128^3 (2,097,152 states, t=1, disorder U[-1,1] seed 5), R={R} "
                     f"random-phase vectors (seed {seed}), M={M}, s=Gershgorin*1.05={s:.6f} "
                     "(H and bound built on the device)"),
        "nnz": int(Z),
        "dtype": ("fp32 vectors; real H as one fp64 value per 8-byte entry, fp64 row sums; "
                  "fp64 moment sums, doubling"),
        "seconds": ms * 1e-3,
        "wall_s": wall,
        "block_width": Bw, "spmm_shape": shape, "blocks": blocks,
        "step_launches": int(launches),
        "spmm_step_ms": per_step * 1e3,
        "spmm_GBps": step_bytes / per_step / 1e9,
        "spmm_hbm_frac": step_bytes / per_step / 1e9 / HBM_PEAK_GBPS,
        "mu0_over_RN": float(mu[0].real / (R * N)),
        "max_dos_rel_err_fp32_vs_fp64_same_8_vectors": rel_errors(dos_a, dos_b)[0],
        "max_dos_pointwise_rel_err_fp32_vs_fp64_same_8_vectors": rel_errors(dos_a, dos_b)[1],
    }


def contraction_mfma(eng):
    """G(E) for K=1e5 coefficient vectors x M=4096 moments x E=8192 energies
    (Retarded) as a split-BF16 complex GEMM on MFMA, FP32 accumulate."""
    K, M, E = 100_000, 4096, 8192
    s = 10.0
    rng = np.random.default_rng(21)
    # Moments of single levels at random energies (Jackson-damped), 2048
    # distinct rows repeated to K.
    x = rng.uniform(-0.9, 0.9, 2048)
    n = np.arange(M)
    base = np.cos(np.outer(np.arccos(x), n)) * kpm.jackson(M)[None, :]
    base = base * np.exp(1j * rng.uniform(0, 2 * np.pi, (2048, 1)))
    coeffs = np.tile(base, (K // 2048 + 1, 1))[:K]
    eng.lut_generate(M, E, -0.99 * s, 0.99 * s, s)
    eng.coeffs_upload(coeffs)
    eng.contract_mfma(X.RETARDED, keep_on_device=True)  # table planes, warm-up
    times = []
    for _ in range(3):
        eng.contract_mfma(X.RETARDED, keep_on_device=True)
        times.append(eng.last_timing()[0])
    ms = min(times)
    # NonPrincipal (the type every LDOS/density/magnetization extraction
    # uses): the table operand is Im L only, 6 BF16 products per MAC
    times_np = []
    for _ in range(3):
        eng.contract_mfma(X.NONPRINCIPAL, keep_on_device=True)
        times_np.append(eng.last_timing()[0])
    ms_np = min(times_np)
    # error on 64 rows against the FP64 contraction
    rows = coeffs[:64]
    eng.coeffs_upload(rows)
    err = {}
    for t in (X.RETARDED, X.NONPRINCIPAL):
        got = eng.contract_mfma(t)
        ref = eng.contract(rows, t)
        err[t] = rel_errors(got, ref)
    eng.lut_destroy()
    executed = 2.0 * K * E * M * 12
    algorithmic = 8.0 * K * E * M
    executed_np = 2.0 * K * E * M * 6
    lib_tflops = library_bf16_gemm_tflops(K, M, E)
    return {
        "workload": f"K={K} coefficient vectors x M={M} x E={E}, Retarded, s={s}",
        "dtype": "split-bf16 (3 terms) x complex, fp32 accumulate",
        "ms": ms,
        "executed_TFLOPs": executed / (ms * 1e-3) / 1e12,
        "algorithmic_TFLOPs": algorithmic / (ms * 1e-3) / 1e12,
        "mfma_frac_of_bf16_dense_peak": executed / (ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS,
        "max_rel_err_vs_fp64": err[X.RETARDED][0],
        "max_pointwise_rel_err_vs_fp64": err[X.RETARDED][1],
        "nonprincipal": {
            "ms": ms_np,
            "executed_TFLOPs": executed_np / (ms_np * 1e-3) / 1e12,
            "mfma_frac_of_bf16_dense_peak":
                executed_np / (ms_np * 1e-3) / 1e12 / BF16_PEAK_TFLOPS,
            "max_rel_err_vs_fp64": err[X.NONPRINCIPAL][0],
            "max_pointwise_rel_err_vs_fp64": err[X.NONPRINCIPAL][1],
        },
        "library_bf16_gemm_same_shape_TFLOPs": lib_tflops,
        "library_bf16_gemm_frac_of_peak": lib_tflops / BF16_PEAK_TFLOPS,
    }


def library_bf16_gemm_tflops(K, M, E):
    """Best of three torch.matmul (hipBLASLt) launches of one real BF16 GEMM
    (K x M) x (M x E), timed in the same run: the dense-BF16 rate the vendor
    library reaches on this GPU at this shape, for comparison with the
    contraction's executed rate."""
    import torch
    a = torch.randn(K, M, device="cuda", dtype=torch.bfloat16)
    b = torch.randn(M, E, device="cuda", dtype=torch.bfloat16)
    c = a @ b
    best = float("inf")
    for _ in range(3):
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        torch.matmul(a, b, out=c)
        t1.record()
        torch.cuda.synchronize()
        best = min(best, t0.elapsed_time(t1))
    del a, b, c
    torch.cuda.empty_cache()
    return 2.0 * K * M * E / (best * 1e-3) / 1e12


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--parts", default=",".join(PARTS),
                    help="comma list of " + ",".join(PARTS) + " (default: all); without "
                         "'plain' the line carries no headline value (profiling runs)")
    ap.add_argument("--skip-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a 1 GiB read + 1 GiB write elementwise kernel (torch) "
                         "whose known bytes scale the rocprofv3 memory counters")
    ap.add_argument("--no-graph", action="store_true",
                    help="uneven part: time a direct run instead of the hipGraph replay")
    ap.add_argument("--uneven-moments-div", type=int, default=1,
                    help="uneven part, for counter runs: divide every model's moment count "
                         "by this (at least 50 moments)")
    ap.add_argument("--dos-moments-div", type=int, default=1,
                    help="stochastic DOS part, for counter runs: divide M by this "
                         "(at least 64 moments)")
    ap.add_argument("--uneven-tile-rows", type=int, default=0,
                    help="uneven part: fixed row groups per tile instead of the autotuned "
                         "launch shape")
    ap.add_argument("--uneven-grid-share", type=int, default=0,
                    help="uneven part with --uneven-tile-rows: 0 = one tile per wavefront, "
                         "s > 0 = persistent grid of occupancy x CUs / s workgroups")
    ap.add_argument("--no-patterns", action="store_true",
                    help="slot-major plain step with per-slot offsets and codes instead of "
                         "row patterns")
    ap.add_argument("--row-pattern-bytes", action="store_true",
                    help="row patterns read as one byte per row everywhere, without the "
                         "wave-uniform scalar path (TBTKX_OPT_ROW_PATTERNS = 2)")
    ap.add_argument("--slot-rows", type=int, default=0,
                    help="rows per thread of the slot-major plain step (1, 2, 4; 0 = default)")
    ap.add_argument("--no-palette", action="store_true",
                    help="plain step with full values instead of the value palette")
    ap.add_argument("--complex", action="store_true",
                    help="run every recurrence in complex arithmetic, also for a real H "
                         "(TBTKX_OPT_REAL = 0), to measure the complex kernels")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs the Chebyshev path on a "
              "single MI355X (independent sites/vectors shard across GPUs with no "
              "communication; see DESIGN.md). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    parts = [x for x in args.parts.split(",") if x]
    for x in parts:
        if x not in PARTS:
            ap.error(f"unknown part {x!r}")
    if args.skip_cpu_baseline and "cpu" in parts:
        parts.remove("cpu")

    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.add(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b

    out = {"metric": METRIC}
    if "mfma" in parts and not args.calibrate_copy:
        # torch (the library GEMM of the MFMA part) brings its own HIP
        # runtime, which initializes only when it comes before the engine's.
        import torch
        torch.zeros(1, device="cuda")
        torch.cuda.synchronize()
    eng = None
    if any(p != "cpu" for p in parts):
        eng = X.Engine(0)
        if args.complex:
            eng.set_option(X.OPT_REAL, 0)
        if args.no_palette:
            eng.set_option(X.OPT_PALETTE, 0)
        if args.no_patterns:
            eng.set_option(X.OPT_ROW_PATTERNS, 0)
        elif args.row_pattern_bytes:
            eng.set_option(X.OPT_ROW_PATTERNS, 2)
        if args.slot_rows:
            eng.set_option(X.OPT_SLOT_ROWS, args.slot_rows)
    model = None
    if "plain" in parts:
        log("building 1024x1024 spin lattice")
        model = L.config_plain_large(seed=2)
    if "plain" in parts:
        csr = model.csr()
        N, Z = model.basis_size, csr[1].size
        eng.set_csr(csr, SCALE)
        kpm.check_window(*WINDOW, SCALE)
        eng.lut_generate(M_PLAIN, E_PTS, WINDOW[0], WINDOW[1], SCALE)
        for _ in range(args.warmup):
            plain_ldos(eng, model, CENTRE, M_PLAIN)
        log(f"timing {args.steps} steps")
        kernel_ms = 0.0
        launches = 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ldos, kms, n = plain_ldos(eng, model, CENTRE, M_PLAIN)
            kernel_ms += kms
            launches += n
        elapsed = time.perf_counter() - t0
        eng.lut_destroy()
        sites_per_step = 2
        value = args.steps * sites_per_step / elapsed
        # Bytes of the form the calls ran: real or complex FP64, 16- or
        # 32-bit columns, palette codes or values, slot-major or row blocks,
        # row patterns (wave-uniform or a byte for every row).
        step_bytes, form = plain_step_bytes(eng, N, Z, csr, args.row_pattern_bytes)
        real, col16, palette = form["real"], form["col16"], form["palette"]
        slot, patterns, uniform = form["slot"], form["patterns"], form["uniform"]
        byte_rows = form["byte_rows"]
        avg_kernel_s = kernel_ms * 1e-3 / launches
        achieved = step_bytes / avg_kernel_s / 1e9
        dE = (WINDOW[1] - WINDOW[0]) / (E_PTS - 1)
        out.update({
            "value": value,
            "unit": "LDOS sites/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": elapsed / args.steps * 1e3,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f64",
            "data": "synthetic",
            "config": {
                "workload": ("square lattice 1024x1024 with spin, t=1, J=0.5, disorder "
                             "U[-1,1] seed 2; plain FP64 single start vector; LDOS at "
                             "(512,512,sum over spin) = 2 sites per step; s=12, M=4000, "
                             "E=2000 on [-11.88,11.88], Lorentz 1e-6"),
                "states": int(N), "nnz": int(Z), "moments": M_PLAIN, "energies": E_PTS,
                "sites_per_step": sites_per_step,
            },
            "physical_sites_per_s": value / 2,
            "spmv_GBps": achieved,
            "roofline": {
                "kernel": ((("plain_step_slot<double> (real FP64 slot-major ELL SpMV "
                             "recurrence step, one thread per row; H is real)" if real else
                             "plain_step_slot<double2> (complex FP64 slot-major ELL SpMV "
                             "recurrence step)") if slot else
                            ("plain_step_lds<double> (real FP64 CSR SpMV recurrence step; "
                             "H is real)" if real else
                             "plain_step_lds<double2> (complex FP64 CSR SpMV recurrence step)"))
                           + (("; row patterns (the model's distinct rows: 16-bit offsets, "
                               "palette codes): wavefronts whose 64-row chunks share one read "
                               "it through scalar loads, the other rows one byte per row; "
                               "diagonal array" if uniform else
                               "; one byte per row into the model's distinct rows (16-bit "
                               "offsets, palette codes) + diagonal array") if patterns else
                              (("; 16-bit column offsets" if col16 else "; 32-bit columns")
                               + ("; 1-byte value codes (palette) + diagonal array"
                                  if palette else "")))),
                "bound": "hbm",
                "achieved": achieved,
                "peak": HBM_PEAK_GBPS,
                "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBPS,
                "traffic": load_traffic(("plain_step_slot" if slot else "plain_step_lds")
                                        + ("_real" if real else "")
                                        + ("_col16" if col16 and not slot else "")
                                        + ("_palette" if palette else "")
                                        + ("_patterns" if patterns else "")
                                        + ("_uniform" if uniform else "")),
                "algorithmic_bytes_per_launch": step_bytes,
                **({"pattern_byte_rows": byte_rows} if patterns else {}),
                "avg_launch_us": avg_kernel_s * 1e6,
            },
            "checks": {"ldos_integral": float(ldos.sum() * dE),
                       "ldos_integral_expected": 2.0,
                       "ldos_min": float(ldos.min())},
        })
    if "batched" in parts:
        log("batched comparison (512x512 Nambu, 256 sites)")
        out["batched"] = batched_comparison(eng)
    if "uneven" in parts:
        log("uneven batch (64 models, one launch per step)")
        out["uneven_batch"] = uneven_batch(eng, graph=not args.no_graph,
                                           moments_div=args.uneven_moments_div,
                                           tile_rows=args.uneven_tile_rows,
                                           grid_share=args.uneven_grid_share)
    if "dos" in parts:
        log("stochastic DOS (128^3 cubic, R=512, M=4096)")
        out["stochastic_dos"] = stochastic_dos(eng, moments_div=args.dos_moments_div)
    if "mfma" in parts:
        log("MFMA contraction (1e5 x 4096 x 8192)")
        out["contraction_mfma"] = contraction_mfma(eng)
    if eng is not None:
        eng.close()
    if "cpu" in parts:
        cpu_baselines(out, parts, model)
    elif "plain" in parts:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def cpu_baselines(out, parts, model):
    """Reference CPU baselines (oracle/_ref/ref_driver, the reference's own
    CPU path compiled from its sources) for every configuration whose GPU
    part ran (all of them with --parts cpu alone), each on a bounded sample,
    one core and one process per core; the GPU/CPU ratios beside them."""
    from oracle import oracle as O
    if not O.ref_available():
        log("no reference build (oracle/_ref/ref_driver): CPU baselines skipped")
        out["cpu_baseline"] = None
        return
    only_cpu = parts == ["cpu"]
    procs = cpu_info()[2]
    factor = None   # throughput mode: aggregate over single-core rate
    with tempfile.TemporaryDirectory() as td:
        if "plain" in parts or only_cpu:
            log("CPU baseline: plain 1024x1024 spin lattice")
            if model is None:
                model = L.config_plain_large(seed=2)
            out["cpu_baseline"] = baseline_ldos(model, SCALE, E_PTS, "512,512,+", 2, M_PLAIN,
                                                "plain", procs, td)
            factor = out["cpu_baseline"]["throughput"]["aggregate_over_single"]
        if "batched" in out or only_cpu:
            log("CPU baseline: 512x512 Nambu, one site")
            m = L.config_nambu(512)
            site = np.sort(np.random.default_rng(3).choice(m.basis_size, B_SITES,
                                                           replace=False))[0]
            idx = ",".join(map(str, m.physical_index(int(site))))
            cb = baseline_ldos(m, m.gershgorin_bound() * 1.05, E_PTS, idx, 1, M_BATCH,
                               "nambu", procs, td, factor=factor)
            factor = cb["throughput"]["aggregate_over_single"]
            del m
            sub = out.setdefault("batched", {})
            sub["cpu_baseline"] = cb
            if "sites_per_s" in sub:
                sub["vs_cpu_1core"] = sub["sites_per_s"] / cb["value"]
                sub["vs_cpu_throughput"] = sub["sites_per_s"] / cb["throughput"]["value"]
        if "uneven_batch" in out or only_cpu:
            log("CPU baseline: uneven batch (two sample models)")
            if factor is None:
                factor = baseline_ldos(L.config_nambu(64), 6.4, E_PTS, "0,0,0", 1, 500,
                                       "factor", procs, td)["throughput"]["aggregate_over_single"]
            cb = baseline_uneven(L.config_uneven(seed=4), procs, td, factor)
            sub = out.setdefault("uneven_batch", {})
            sub["cpu_baseline"] = cb
            if "sites_per_s" in sub:
                sub["vs_cpu_1core"] = sub["sites_per_s"] / cb["value"]
                sub["vs_cpu_throughput"] = sub["sites_per_s"] / cb["throughput"]["value"]
        if "stochastic_dos" in out or only_cpu:
            # The reference's DOS is sampleDOS (Lib/src/PropertyExtractor/
            # PropertyExtractor.cpp:234-290): one LDOS recurrence per sampled
            # site, so 512 vectors cost 512 single-site LDOS calls.
            log("CPU baseline: 128^3 cubic, one site")
            m = L.config_cubic(128)
            cl = baseline_ldos(m, m.gershgorin_bound() * 1.05, E_PTS, "64,64,64", 1, 4096,
                               "cubic", procs, td, factor=factor)
            factor = cl["throughput"]["aggregate_over_single"]
            del m
            R = 512
            cb = {"value": R / cl["value"], "unit": "s", "cores": 1, "kind": "reference",
                  "sample": (cl["sample"] + f"; x {R} samples (sampleDOS sums one "
                             "single-site LDOS per sample)"),
                  "throughput": {"value": R / cl["throughput"]["value"], "unit": "s",
                                 "cores": cl["throughput"]["cores"],
                                 "aggregate_over_single": factor,
                                 "mode": cl["throughput"]["mode"]},
                  "cpu_model": cl["cpu_model"], "nproc": cl["nproc"]}
            sub = out.setdefault("stochastic_dos", {})
            sub["cpu_baseline"] = cb
            if "seconds" in sub:
                sub["vs_cpu_1core"] = cb["value"] / sub["seconds"]
                sub["vs_cpu_throughput"] = cb["throughput"]["value"] / sub["seconds"]
        if "contraction_mfma" in out or only_cpu:
            log("CPU baseline: FP64 lookup-table contraction")
            cb = baseline_contraction(100_000, 4096, 8192, 10.0, procs)
            sub = out.setdefault("contraction_mfma", {})
            sub["cpu_baseline"] = cb
            if "ms" in sub:
                sub["vs_cpu_1core"] = cb["value"] / (sub["ms"] * 1e-3)
                sub["vs_cpu_throughput"] = cb["throughput"]["value"] / (sub["ms"] * 1e-3)


if __name__ == "__main__":
    main()
