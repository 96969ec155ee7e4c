"""Benchmark: density-matrix (P) evolution steps/s on one MI355X.

Workload (N = 10^4): SquareLattice(100, 100) with periodic boundaries along
both axes, t1 = 1, uniform Landau-gauge field ramped linearly to 0.01 flux
quanta per plaquette over tau = 10, on-site Anderson disorder uniform in
[-0.5, 0.5] from a fixed seed, dt = 0.05 (500 steps: t in 0:0.05:25), complex
FP64 dense P.  One "step" is one pass of the iterator: H(t) rebuilt on the
device when the field changed, the Taylor order selected by the reference
rule, and P <- U P U' -- exactly what `Evolution(t -> H(t), P0)` does per
time point.  P0 is a rank-64 orthogonal projector (synthetic; the step's cost
does not depend on the values of P, and K depends only on H).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Prints one JSON line.  `roofline` refers to the dominant kernel: its
algorithmic bytes per launch divided by its average duration from HIP events
recorded on the library's stream.  A per-term launch carries one Horner term
(each of the K terms of a side reads Y and X and writes Y at 16 B per
complex, the first reads X once: (48 K - 16) N^2 per side over K launches);
a fused launch carries a whole side and is measured against what it must
move at least (read P, write the result: 32 N^2), with the per-term
accounting reported beside it as `per_term_equivalent_GBs`.
`cpu_baseline` times the oracle (the reference algorithm restated in C:
sparse U built by myexp! with droptol, then U*P and (U*P)*U') on the host
cores for one full step of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "latticemodels.jl_amd"))

import numpy as np  # noqa: E402

L = 100
TAU = 10.0
BMAX = 0.01
DT = 0.05
NSTEPS = 500
SEED = 20240601
RANK = 64
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6                  # MI355X FP64, vector and matrix
LDS_PEAK_TBS = 128 * 256 * 2.4e9 / 1e12  # 128 B/clk per CU, 256 CUs, 2.4 GHz
# HBM bytes per launch of a fused kernel at N = 10^4 come from two rocprofv3
# --pmc runs (FETCH_SIZE, then WRITE_SIZE, copy-calibrated), recorded in the
# kernel's summary under profiles/ as a line
#     traffic_bytes_per_launch: <variant> <bytes>
# `roofline.traffic` is that figure for the variant that ran, with the file
# it came from; null when no summary covers the variant.
TRAFFIC_FILES = ("strip_fused_steps_n10000.md", "strip_kernel_n10000.md", "fused_value_classes_n10000.md",
                 "fused_column_pass_n10000.md")
# H bytes each fused variant with one column per workgroup reads from L2 per
# site and Horner term (zmax <= 4 layouts): the packed record (diagonal +
# four 16-bit columns) and four complex values, or, with value classes, the
# record alone (the class values are read from LDS); the strip kernel (32)
# holds its bond values in registers and streams the diagonal only.
H_STREAM_BYTES = {20: 80.0, 21: 80.0, 27: 80.0, 28: 80.0, 29: 16.0, 30: 16.0, 32: 8.0}


def profiled_traffic(kernel, variant):
    if kernel != "fused_left":
        return None, None
    for name in TRAFFIC_FILES:
        path = os.path.join(ROOT, "profiles", name)
        try:
            with open(path) as f:
                for line in f:
                    parts = line.split()
                    if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch:" and parts[1] == str(variant):
                        return float(parts[2]), "profiles/" + name
        except OSError:
            continue
    return None, None

# batched workload: 64 disorder realisations (W=2) on open L x L squares with
# L uniform in 24..96, constant Landau field 0.005, dt=0.05, 400 steps
BATCH_MEMBERS = 64
BATCH_LMIN, BATCH_LMAX = 24, 96
BATCH_W = 2.0
BATCH_B = 0.005
BATCH_STEPS = 400
BATCH_SEED = 20240602


def build_inputs():
    rng = np.random.default_rng(SEED)
    V = rng.uniform(-0.5, 0.5, L * L)
    Q = rng.normal(size=(L * L, RANK)) + 1j * rng.normal(size=(L * L, RANK))
    Q, _ = np.linalg.qr(Q)
    P0 = np.asfortranarray(Q @ Q.conj().T)
    return V, P0


def timeline():
    # Julia's 0:0.05:25 (correctly rounded k*step)
    from fractions import Fraction
    return [float(k * Fraction("0.05")) for k in range(NSTEPS + 1)]


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(V, P0, threads):
    """The reference algorithm restated in C (the oracle) on the host cores:
    one full step at t = 5 (host H assembly, myexp! sparse U with droptol,
    U*P and (U*P)*U') with `threads` OpenMP threads -- the reported value --
    plus a 1-thread estimate of the same step (a bounded column sample) and
    the step as Horner passes (the device's algorithm on the CPU)."""
    sys.path.insert(0, ROOT)
    from oracle import fields, geometry, hamiltonian, numerics
    lat = geometry.SquareLattice(L, L, periodic=(True, True))
    t = 5.0
    t0 = time.perf_counter()
    H = hamiltonian.tightbinding_hamiltonian(lat, onsite=V, field=fields.LandauGauge(BMAX * min(t, TAU) / TAU))
    t_h = time.perf_counter() - t0
    t0 = time.perf_counter()
    U, K, s, _ = numerics.myexp(H, DT)
    t_u = time.perf_counter() - t0
    P = np.array(P0, order="F", copy=True)
    t0 = time.perf_counter()
    numerics.step(U, P, nthreads=threads)
    t_step = time.perf_counter() - t0
    el = t_h + t_u + t_step
    # 1 thread: U * X for 2000 of the N columns, scaled to the step's two
    # sparse products over all N columns
    m1 = 2000
    X1 = np.asfortranarray(P0[:, :m1])
    t0 = time.perf_counter()
    numerics.apply(U, X1, nthreads=1)
    t_step1 = (time.perf_counter() - t0) * 2.0 * (L * L) / m1
    P2 = np.array(P0, order="F", copy=True)
    t0 = time.perf_counter()
    numerics.horner_step(H, DT, K, s, P2, nthreads=threads)
    t_horner = time.perf_counter() - t0
    return {"value": 1.0 / el, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": "1 full step at t=5 of the N=10^4 workload (host H assembly %.2f s + myexp! sparse U, K=%d, "
                      "%.2f s + U*P*U' %.2f s), %d OpenMP threads on %s (nproc %d)"
                      % (t_h, K, t_u, t_step, threads, _cpu_model(), os.cpu_count() or 0),
            "single_thread_step_s": t_h + t_u + t_step1,
            "single_thread_note": "U*X on 2000 columns with 1 thread, scaled by 2 N / 2000 (both products)",
            "horner_cpu_step_s": t_h + t_horner,
            "horner_cpu_note": "same order-K polynomial as K Horner passes per side on the host (%d threads)" % threads}


# differentiation workload: SquareLattice 32x32 open, 100-step piecewise-constant
# PointFlux schedule B_k ~ U[0, 0.2] through the lattice centre, V ~ N(0, 0.1^2),
# dt = 0.1; loss = sum of localdensity over the 5x5 corner + (sum of
# DensityCurrents across the column cut x = 16 | 17)^2; gradient w.r.t. the 100
# B_k and the 1024 V_i, checkpoint every 10 steps
ADJ_L, ADJ_STEPS, ADJ_DT, ADJ_SEED, ADJ_EVERY = 32, 100, 0.1, 20240603, 10


def adjoint_workload(lmhip, _lib, reps=3):
    """Time one plain forward evolution and one forward (recorded) + loss +
    adjoint gradient of the differentiation workload; returns the ratio of
    host wall times (min over reps) and of summed device kernel times."""
    import ctypes
    rng = np.random.default_rng(ADJ_SEED)
    L_ = ADJ_L
    B = rng.uniform(0.0, 0.2, ADJ_STEPS)
    V = rng.normal(0.0, 0.1, L_ * L_)
    lat = lmhip.SquareLattice(L_, L_)
    pt = (L_ / 2 + 0.5, L_ / 2 + 0.5)
    hams = [lmhip.tightbinding_hamiltonian(lat, V, field=lmhip.PointFlux(b, pt)) for b in B]
    h = lambda t: hams[min(int(round(t / ADJ_DT)), ADJ_STEPS - 1)]
    xy = lat.coords
    corner = np.nonzero((xy[:, 0] <= 5) & (xy[:, 1] <= 5))[0]
    ctx = lmhip.Context(0)
    P0 = lmhip.densitymatrix(lmhip.tightbinding_hamiltonian(lat, V, field=lmhip.PointFlux(0.0, pt)), mu=0.0,
                             ctx=ctx)
    times = [ADJ_DT * k for k in range(ADJ_STEPS + 1)]
    bnd = hams[0].bonds
    xi, xj = xy[bnd.i, 0], xy[bnd.j, 0]
    cut = (((xi == L_ // 2) & (xj == L_ // 2 + 1)) | ((xj == L_ // 2) & (xi == L_ // 2 + 1))).astype(float)
    lam_rho = lmhip.localdensity_cotangent(L_ * L_, 1.0, sites=corner)
    lib = _lib.lib()

    def device_ms():
        tot = 0.0
        for kind in range(10):
            n_ = ctypes.c_int64()
            ms_ = ctypes.c_double()
            _lib.check(lib.lmh_get_profile(ctx.h, kind, ctypes.byref(n_), ctypes.byref(ms_)))
            tot += ms_.value
        return tot

    def run(grad, profile):
        ev = lmhip.Evolution(lmhip.HIPExp(ctx=ctx), h, P0)
        _lib.check(lib.lmh_set_profiling(ctx.h, 1 if profile else 0))
        t0 = time.perf_counter()
        tape = lmhip.EvolutionTape(ev, ADJ_EVERY) if grad else None
        H_last = None
        for ts in ev(times):
            H_last = ts.H
        rho, J = lmhip.observables(ev.states[0], H_last)
        S = float(np.sum(J * cut))
        loss = float(np.sum(rho[corner])) + S * S
        g = None
        if grad:
            lamJ, gdir = lmhip.currents_cotangent(H_last, ev.states[0], 2.0 * S * cut)
            g = tape.backward(lam_rho + lamJ)
            tape.close()
        _lib.check(lib.lmh_sync(ctx.h))
        el = time.perf_counter() - t0
        dev = device_ms() if profile else None
        _lib.check(lib.lmh_set_profiling(ctx.h, 0))
        ev.states[0].close()
        return el, dev, loss, g

    run(False, False)
    run(True, False)
    t_f = min(run(False, False)[0] for _ in range(reps))
    t_g = min(run(True, False)[0] for _ in range(reps))
    _, d_f, loss, _ = run(False, True)
    _, d_g, loss_g, g = run(True, True)
    return {"workload": "SquareLattice 32x32 open, 100-step PointFlux schedule + 1024 on-site potentials, "
                        "corner density + squared cut current, checkpoint every 10",
            "ratio_wall": t_g / t_f, "ratio_device": d_g / d_f,
            "forward_ms": t_f * 1e3, "forward_plus_adjoint_ms": t_g * 1e3,
            "forward_device_ms": d_f, "forward_plus_adjoint_device_ms": d_g,
            "loss": loss, "grad_norm_B": float(np.linalg.norm(g.field[:, 0])),
            "grad_norm_V": float(np.linalg.norm(g.onsite)), "target_ratio": 3.0}


# large workload: HoneycombLattice 120x120 unit cells (N = 28800), Haldane model
# t1 = 1, t2 = 0.1 (i t2 on the next-nearest neighbours), sublattice mass 0.2,
# PointFlux through the hexagon centre nearest the lattice centre ramped to
# 0.5 flux quanta over tau = 10, dt = 0.02, 1000 steps
LARGE_L, LARGE_DT, LARGE_STEPS, LARGE_TAU, LARGE_FLUX, LARGE_SEED = 120, 0.02, 1000, 10.0, 0.5, 20240604


def large_point():
    """Hexagon centre nearest the lattice centre (no site or bond on it)."""
    c = LARGE_L // 2
    rx, ry = c * 1.0 + c * 0.5, c * np.sqrt(3.0) / 2.0
    return (rx - 0.5, ry - np.sqrt(3.0) / 6.0)


def pin_schedule(lib, _lib, ctx, spec):
    """MODE:LEFT_COLS:RIGHT_PASS:RIGHT_COLS[:FUSED_VARIANT[:LEFT_PASS]] -> library options."""
    parts = spec.split(":")
    m_, lc_, rp_, rc_ = parts[:4]
    fv_ = int(parts[4]) if len(parts) > 4 else -1
    lv_ = int(parts[5]) if len(parts) > 5 else 0
    pinned = {"perpass": _lib.STEP_PER_PASS, "fusedT": _lib.STEP_FUSED_T, "fusedN": _lib.STEP_FUSED_N}[m_]
    for key, v in ((_lib.OPT_STEP_MODE, pinned), (_lib.OPT_LEFT_COLS, int(lc_)), (_lib.OPT_RIGHT_PASS, int(rp_)),
                   (_lib.OPT_RIGHT_COLS, int(rc_)), (_lib.OPT_FUSED_VARIANT, fv_), (_lib.OPT_LEFT_PASS, lv_)):
        _lib.check(lib.lmh_set_option(ctx.h, key, v))
    return {"pinned": spec}


def bench_large(args, lmhip, _lib):
    import ctypes
    lat = lmhip.HoneycombLattice(LARGE_L, LARGE_L)
    n = len(lat)
    pt = large_point()
    h = lambda t: lmhip.haldane(lat, 1.0, 0.1, 0.2, field=lmhip.PointFlux(LARGE_FLUX * min(t, LARGE_TAU) / LARGE_TAU, pt))
    ctx = lmhip.Context(0)
    H0 = h(0.0)
    ctx.set_hamiltonian(H0)
    t0 = time.perf_counter()
    if args.p0 == "fermi":
        P0 = lmhip.densitymatrix(H0, mu=0.0, ctx=ctx)
    else:
        occ = np.random.default_rng(LARGE_SEED).permutation(n)[: n // 2]
        P0 = lmhip.DensityMatrix(ctx, n).set_entries(occ, occ, np.ones(len(occ)))
    t_p0 = time.perf_counter() - t0
    ev = lmhip.Evolution(lmhip.HIPExp(ctx=ctx), h, P0)
    if args.adjoint_steps <= 0:
        P0.close()                        # else kept: the adjoint leg starts from it too
    from fractions import Fraction
    ts = [float(k * Fraction("0.02")) for k in range(args.warmup + args.steps + 1)]
    it = ev(ts)
    next(it)
    for _ in range(args.warmup):
        next(it)
    lib = _lib.lib()
    tuned = None
    if args.schedule:
        tuned = pin_schedule(lib, _lib, ctx, args.schedule)
        ev.solver._tuned = True           # keep the pinned options (no per-step mode reset)
    elif not args.no_autotune:
        tuned = ev.solver.autotune(ev.states[0], reps=1)
    _lib.check(lib.lmh_sync(ctx.h))
    _lib.check(lib.lmh_set_profiling(ctx.h, 1))
    t0 = time.perf_counter()
    Ks = []
    for _ in range(args.steps):
        next(it)
        Ks.append(ev.solver.K)
    _lib.check(lib.lmh_sync(ctx.h))
    el = time.perf_counter() - t0
    prof = {}
    for kind, name in ((_lib.KERNEL_HORNER_LEFT, "horner_left"), (_lib.KERNEL_HORNER_RIGHT, "horner_right"),
                       (_lib.KERNEL_BUILD_H, "build_h"), (_lib.KERNEL_TAYLOR_ORDER, "taylor_order")):
        nl_ = ctypes.c_int64()
        ms_ = ctypes.c_double()
        _lib.check(lib.lmh_get_profile(ctx.h, kind, ctypes.byref(nl_), ctypes.byref(ms_)))
        prof[name] = (nl_.value, ms_.value)
    k_sum = float(sum(Ks))
    bytes_side = (48.0 * k_sum - 16.0 * len(Ks)) * n * n
    # every per-term launch (left or right, either side) carries one term
    dom = "horner (all per-term launches)"
    nl = prof["horner_left"][0] + prof["horner_right"][0]
    ms = prof["horner_left"][1] + prof["horner_right"][1]
    avg_ms = ms / max(nl, 1)
    achieved = 2.0 * bytes_side / max(nl, 1) / (avg_ms * 1e-3) / 1e9
    if args.copy_calibration:
        ev.states[0].copy().close()
        _lib.check(lib.lmh_sync(ctx.h))
    rho, _J = lmhip.observables(ev.states[0])
    adj = None
    if args.adjoint_steps > 0:
        # edge density gradient w.r.t. all on-site potentials and the flux of
        # every step: a fresh evolution of adjoint_steps steps with a tape
        ev.states[0].close()
        _lib.check(lib.lmh_set_profiling(ctx.h, 0))
        H0 = h(0.0)
        deg = np.zeros(n, dtype=int)
        np.add.at(deg, H0.bonds.i, 1)
        np.add.at(deg, H0.bonds.j, 1)
        edge = np.nonzero(deg < np.max(deg))[0]
        ev2 = lmhip.Evolution(lmhip.HIPExp(ctx=ctx), h, P0)
        P0.close()
        ts2 = [float(k * Fraction("0.02")) for k in range(args.adjoint_steps + 1)]
        t0 = time.perf_counter()
        tape = lmhip.EvolutionTape(ev2, args.checkpoint_every)
        for k, _ in enumerate(ev2(ts2)):
            if k and k % args.checkpoint_every == 0:
                _lib.check(lib.lmh_sync(ctx.h))
                print("forward with tape: %d of %d steps, %.1f s" % (k, args.adjoint_steps, time.perf_counter() - t0),
                      file=sys.stderr, flush=True)
        _lib.check(lib.lmh_sync(ctx.h))
        t_fwd = time.perf_counter() - t0
        loss = float(np.sum(lmhip.localdensity(ev2.states[0])[edge]))
        t0 = time.perf_counter()
        report = lambda done, total: print("backward: %d of %d reverse steps, %.1f s" % (done, total, time.perf_counter() - t0),
                                           file=sys.stderr, flush=True)
        g = tape.backward(lmhip.localdensity_cotangent(n, 1.0, sites=edge), progress=report)
        t_bwd = time.perf_counter() - t0
        cp_dev, cp_host = tape.checkpoints
        adj = {"steps": args.adjoint_steps, "checkpoint_every": args.checkpoint_every, "p0": args.p0,
               "checkpoints": -(-args.adjoint_steps // args.checkpoint_every),
               "checkpoints_in_device_memory": cp_dev, "checkpoints_in_host_memory": cp_host,
               "loss": "edge density (%d sites)" % len(edge), "loss_value": loss,
               "forward_with_tape_s": t_fwd, "backward_s": t_bwd, "ratio": (t_fwd + t_bwd) / t_fwd,
               "grad_norm_V": float(np.linalg.norm(g.onsite)), "grad_norm_flux": float(np.linalg.norm(g.field))}
        tape.close()
    out = {
        "metric": "P-evolution steps/s at N=28800 (honeycomb Haldane) FP64",
        "value": args.steps / el, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 (complex FP64)", "data": "synthetic (P0 %s)" % args.p0,
        "config": {"workload": "HoneycombLattice 120x120 open, Haldane t1=1 t2=0.1 M=0.2, PointFlux ramp to 0.5 "
                               "over tau=10 at the central hexagon, dt=0.02", "N": n, "K": sorted(set(Ks)),
                   "p0_seconds": t_p0, "autotuned": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom, "avg_launch_ms": avg_ms,
                     "launches": nl, "accounting": "per-term Horner bytes: (48 K - 16) N^2 per side over K launches",
                     "step_per_term_frac": 2.0 * bytes_side / len(Ks) / (el / args.steps) / (HBM_PEAK_GBS * 1e9),
                     "step_per_term_note": "the per-term bytes of a whole step (both sides) over the whole step "
                                           "time, transposes and order search included"},
        "kernels_ms_total": {k: v[1] for k, v in prof.items()},
        "trace_P": float(np.sum(rho)),
        "adjoint": adj,
    }
    print(json.dumps(out), flush=True)


# ket workload (the reference's time-dependent evolution benchmark,
# benchmarks/benchmark_evolution_dynamic.jl with benchmarks/models.jl): a
# honeycomb lattice of ~10^5 sites (224 x 224 cells; the benchmark's largest
# size, on a rectangle instead of a disc), t1 = -2.8, a 0.2 potential inside
# the central disc of a quarter of the radius squared, SymmetricGauge ramped
# to 3 * 1.519e-3 over t = 5, one ket, 100 steps of 0.1 with localdensity
# every step, Krylov propagation (KrylovKitExp's defaults).  The ket is a
# seeded random normalised vector (the benchmark's ground state needs a sparse
# eigensolver; a random ket spans the whole spectrum, the harder case).
KETS_L, KETS_T1, KETS_B, KETS_RAMP, KETS_DT, KETS_STEPS, KETS_SEED = 224, -2.8, 3.0 * 1.519e-3, 5.0, 0.1, 100, 20240605


def kets_model(lmhip):
    lat = lmhip.HoneycombLattice(KETS_L, KETS_L)
    xy = lat.coords
    c = 0.5 * (xy.min(axis=0) + xy.max(axis=0))
    rad = 0.5 * float(np.min(xy.max(axis=0) - xy.min(axis=0)))
    V = np.where(np.sum((xy - c) ** 2, axis=1) < rad * rad / 4.0, 0.2, 0.0)
    h = lambda t: lmhip.tightbinding_hamiltonian(lat, V, t1=KETS_T1,
                                                 field=lmhip.SymmetricGauge(KETS_B * min(t / KETS_RAMP, 1.0)))
    return lat, V, h


def kets_cpu_baseline(lat, V, psi, nsteps):
    """The oracle's Krylov restatement (NumPy, scipy.sparse H from the
    oracle's Hamiltonian builder; one thread for the sparse products) for
    `nsteps` steps at the start of the same timeline."""
    sys.path.insert(0, ROOT)
    import scipy.sparse as sp
    from oracle import fields, geometry, hamiltonian, krylov
    olat = geometry.HoneycombLattice(KETS_L, KETS_L)
    mv = 0
    t_h = t_k = 0.0
    x = psi.copy()
    for k in range(nsteps):
        t = k * KETS_DT
        t0 = time.perf_counter()
        Ho = hamiltonian.tightbinding_hamiltonian(olat, onsite=V, t1=KETS_T1,
                                                  field=fields.SymmetricGauge(KETS_B * min(t / KETS_RAMP, 1.0)))
        c = Ho.csc
        Hs = sp.csc_matrix((c.nzval, c.rowval, c.colptr), shape=(c.n, c.n)).tocsr()
        t_h += time.perf_counter() - t0
        t0 = time.perf_counter()
        x, m_ = krylov.lanczos_expv(Hs, KETS_DT, x)
        _ = np.abs(x) ** 2
        t_k += time.perf_counter() - t0
        mv += m_
    return {"value": nsteps / (t_h + t_k), "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": "%d steps of the same timeline: host H assembly by the oracle's builder (%.2f s per step) "
                      "and the Lanczos exponential in NumPy with scipy.sparse products (%.3f s per step, %d H "
                      "applications), %s" % (nsteps, t_h / nsteps, t_k / nsteps, mv, _cpu_model()),
            "lanczos_only_steps_per_s": nsteps / t_k}


def bench_kets(args, lmhip, _lib):
    import ctypes
    from fractions import Fraction
    lat, V, h = kets_model(lmhip)
    n = len(lat)
    rng = np.random.default_rng(KETS_SEED)
    psi = rng.normal(size=n) + 1j * rng.normal(size=n)
    psi /= np.linalg.norm(psi)
    solver = lmhip.HIPKrylovExp(ctx=lmhip.Context(0))
    ev = lmhip.Evolution(solver, h, psi)
    ctx = solver.ctx
    lib = _lib.lib()
    ts = [float(k * Fraction(str(KETS_DT))) for k in range(args.warmup + args.steps + 1)]
    it = ev(ts)
    next(it)
    for _ in range(args.warmup):
        st = next(it)
        st.states[0].observables(J=False)
    _lib.check(lib.lmh_sync(ctx.h))
    _lib.check(lib.lmh_set_profiling(ctx.h, 1))
    mvs = []
    dens = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = next(it)
        mvs.append(solver.matvecs)
        rho, _ = st.states[0].observables(J=False)       # localdensity, as the benchmark's TimeSequence
        dens.append(float(rho[n // 2]))
    _lib.check(lib.lmh_sync(ctx.h))
    el = time.perf_counter() - t0
    prof = {}
    for kind, name in ((_lib.KERNEL_KRYLOV, "krylov"), (_lib.KERNEL_BUILD_H, "build_h"),
                       (_lib.KERNEL_OBSERVABLES, "observables")):
        nl_ = ctypes.c_int64()
        ms_ = ctypes.c_double()
        _lib.check(lib.lmh_get_profile(ctx.h, kind, ctypes.byref(nl_), ctypes.byref(ms_)))
        prof[name] = (nl_.value, ms_.value)
    # algorithmic bytes of the Krylov launches: per Lanczos iteration and
    # site, H's row (diagonal 8 B + zmax x (4 B column + 16 B value)), v_j
    # and v_{j-1} read and w written (matvec), v_j and w read and w written
    # (orthogonalisation), w read and v_{j+1} written (scaling); per step the
    # combination reads the basis and writes the ket
    zmax = 3
    per_iter = 8 + zmax * 20 + 3 * 16 + 3 * 16 + 2 * 16
    mv_total = float(sum(mvs))
    bytes_total = n * (per_iter * mv_total + 16.0 * (mv_total + len(mvs)) + 16.0 * len(mvs))
    nl, ms = prof["krylov"]
    achieved = bytes_total / (ms * 1e-3) / 1e9 if ms > 0 else None
    out = {
        "metric": "ket evolution steps/s (Krylov, honeycomb N=%d, time-dependent H, localdensity per step)" % n,
        "value": args.steps / el, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 (complex FP64)", "data": "synthetic (seeded random normalised ket)",
        "config": {"workload": "HoneycombLattice %dx%d open, t1=-2.8, 0.2 potential in the central disc, "
                               "SymmetricGauge ramp to 3*1.519e-3 over t=5, dt=0.1, one ket, HIPKrylovExp "
                               "(krylovdim 30, tol 1e-12)" % (KETS_L, KETS_L), "N": n,
                   "matvecs_per_step": [min(mvs), max(mvs)], "density_center_last": dens[-1]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None,
                     "kernel": "krylov (Lanczos iterations + combination)", "launches": nl,
                     "kernel_ms_per_step": ms / args.steps,
                     "accounting": "per iteration and site: H row (8 + 20 zmax B) + 8 x 16 B of vectors; "
                                   "combination: the basis read, the ket written"},
        "kernels_ms_total": {k: v[1] for k, v in prof.items()},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = kets_cpu_baseline(lat, V, psi, 5)
    print(json.dumps(out), flush=True)


def bench_batched(args, lmhip):
    """All realisations advanced together (lmh_batch_step: one persistent strip
    launch per side over all members -- or, for the record, one fused launch
    per side and size class -- three-step HIP graphs replayed).  P0 per
    realisation: the Fermi projector at mu=0 (--p0 fermi, rocSOLVER on the
    device) or, by default, a half-filled random product state (same cost per
    step; K depends only on H)."""
    rng = np.random.default_rng(BATCH_SEED)
    sizes = rng.integers(BATCH_LMIN, BATCH_LMAX + 1, BATCH_MEMBERS)
    members = []
    for r, Ls in enumerate(sizes):
        Ls = int(Ls)
        rr = np.random.default_rng(BATCH_SEED + 1 + r)
        V = rr.uniform(-BATCH_W / 2, BATCH_W / 2, Ls * Ls)
        H = lmhip.tightbinding_hamiltonian(lmhip.SquareLattice(Ls, Ls), V, field=lmhip.LandauGauge(BATCH_B))
        if args.p0 == "fermi":
            P0 = lmhip.densitymatrix(H, mu=0.0, ctx=lmhip.Context(0))
        else:
            P0 = np.zeros((Ls * Ls, Ls * Ls), dtype=np.complex128, order="F")
            occ = rr.permutation(Ls * Ls)[: Ls * Ls // 2]
            P0[occ, occ] = 1.0
        members.append((H, P0))
    def timed(strip):
        t_setup = time.perf_counter()
        batch = lmhip.BatchedEvolution(members, DT, strip=strip)
        t_setup = time.perf_counter() - t_setup
        kernels = batch.member_kernels()
        batch.step(args.warmup, graph=True)
        t0 = time.perf_counter()
        batch.step(args.steps, graph=True)
        el = time.perf_counter() - t0
        return batch, el, t_setup, kernels

    # the per-size-class launches first (then freed), for the record
    classes = None
    if args.batch_kernels == "both":
        b2, el2, _s2, k2 = timed(False)
        classes = {"steps_per_s": args.steps / el2, "ms_per_step": el2 * 1e3 / args.steps,
                   "launches_per_step": 2 * len(set(k2)), "member_kernels": k2}
        b2.close()
        del b2
    batch, el, t_setup, kernels = timed(args.batch_kernels != "classes")
    del members
    md = batch.mean_density()
    Ns = np.array([int(L) ** 2 for L in sizes], dtype=np.float64)
    bytes_step = float(np.sum(2 * 32.0 * Ns * Ns))       # both sides: read P, write result
    achieved = bytes_step * args.steps / el / 1e9
    out = {
        "metric": "batched P-evolution steps/s (64 realisations, N 576..9216) FP64",
        "value": args.steps / el,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (complex FP64)",
        "data": "synthetic (seeded disorder per realisation, P0 %s)" % args.p0,
        "config": {"workload": "64 realisations, open LxL squares, L uniform 24..96, W=2, Landau 0.005, "
                               "dt=0.05, graph-replayed batch steps",
                   "sizes_L": [int(x) for x in sizes], "sum_N": int(Ns.sum()),
                   "K": [int(s.K) for s in batch.solvers], "setup_s": t_setup,
                   "member_kernels": kernels,
                   "launches_per_step": 2 * len(set(kernels)),
                   "batch_kernels": "persistent strip launch (one per side over all members)"
                   if set(kernels) == {32} else "per-size-class fused launches"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "batched_strip_kernel (both sides)" if set(kernels) == {32} else "batched_fused (all launches)",
                     "bytes_per_step": bytes_step},
        "per_class_launches": classes,
        "mean_density": [float(x) for x in md],
    }
    print(json.dumps(out), flush=True)
    batch.close()


# small workload: SquareLattice 10x10 open, nearest-neighbour t = 1, a
# PointFlux (axial gauge) through (5.5, 5.5) ramped 0.2 min(t, 10) / 10,
# dt = 0.1, t in 0:0.1:20 (200 steps), localdensity and DensityCurrents on
# every bond at every step; P0 is the committed fixture (the B = 0 spectrum
# has ten exactly degenerate zero modes, so the mu = 0 projector is fixed by
# the fixture, not by an eigensolver)
SMALL_L, SMALL_DT, SMALL_STEPS, SMALL_TAU, SMALL_BMAX = 10, 0.1, 200, 10.0, 0.2


def bench_small(args, lmhip, _lib):
    from fractions import Fraction
    g = np.load(os.path.join(ROOT, "tests", "golden", "cfg1.npz"))
    lat = lmhip.SquareLattice(SMALL_L, SMALL_L)
    field = lambda t: lmhip.PointFlux(SMALL_BMAX * min(t, SMALL_TAU) / SMALL_TAU, (5.5, 5.5))
    h = lambda t: lmhip.tightbinding_hamiltonian(lat, field=field(t))
    ts = [float(k * Fraction("0.1")) for k in range(SMALL_STEPS + 1)]
    reps = max(1, args.steps)
    rho_last = J_last = None
    t_best = None
    for _ in range(reps + 1):                     # the first run warms up
        ev = lmhip.Evolution(lmhip.HIPExp(), h, g["P0"])
        t0 = time.perf_counter()
        for (P, H, t) in ev(ts):
            rho_last, J_last = lmhip.observables(P, H)
        el = time.perf_counter() - t0
        t_best = el if t_best is None else min(t_best, el)
        ev.states[0].close()
    n_steps = SMALL_STEPS
    out = {
        "metric": "P-evolution steps/s at N=100 FP64 (with localdensity and DensityCurrents every step)",
        "value": n_steps / t_best, "unit": "steps/s", "n_gpus": 1, "steps": n_steps, "warmup": 1,
        "ms_per_step": t_best * 1e3 / n_steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 (complex FP64)", "data": "P0 from the committed fixture tests/golden/cfg1.npz",
        "config": {"workload": "SquareLattice 10x10 open, t1=1, PointFlux axial through (5.5,5.5) ramped to 0.2 "
                               "over tau=10, dt=0.1, t in 0:0.1:20, observables every step",
                   "N": SMALL_L * SMALL_L, "runs": reps, "timing": "best of the runs after one warm-up run"},
        "roofline": {"bound": "launch latency", "note": "N = 100: every step is a handful of small launches and "
                     "one read-back of the observables; no HBM or FLOP roofline applies"},
        "max_abs_rho_vs_fixture": float(np.max(np.abs(rho_last - g["rho"][-1]))),
        "max_abs_J_vs_fixture": float(np.max(np.abs(J_last - g["J"][-1]))),
    }
    if not args.no_cpu_baseline:
        # the oracle (C sparse U by myexp! and U*P*U') over the same 200 steps,
        # with the observables, on one host thread
        sys.path.insert(0, ROOT)
        from oracle import fields as OF, geometry as OG, hamiltonian as OH, numerics as ON
        olat = OG.SquareLattice(SMALL_L, SMALL_L)
        oh = lambda t: OH.tightbinding_hamiltonian(
            olat, field=OF.PointFlux(SMALL_BMAX * min(t, SMALL_TAU) / SMALL_TAU, (5.5, 5.5), gauge="axial"))
        t0 = time.perf_counter()
        for (Ho, Po, to) in ON.Evolution(oh, g["P0"])(ts):
            ON.localdensity(Po)
            ON.density_currents(Ho, Po)
        el_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n_steps / el_cpu, "unit": "steps/s", "cores": 1, "kind": "port",
                               "sample": "the whole 200-step run with the observables (host H assembly, myexp! "
                                         "sparse U, U*P*U'), one thread on %s" % _cpu_model()}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (480 single, 30 batched)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (20 single, 3 batched)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--mode", choices=("auto", "perpass", "fusedT", "fusedN"), default="auto",
                    help="how the step applies U P U' (see lmh_set_step_mode)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default schedule")
    ap.add_argument("--no-defer", action="store_true",
                    help="single workload: both passes of a strip step at every lmh_step (LMH_OPT_DEFER 0)")
    ap.add_argument("--blocking-order", action="store_true",
                    help="single workload: every H update waits for its Taylor order (LMH_OPT_ASYNC_ORDER 0)")
    ap.add_argument("--schedule", default=None,
                    metavar="MODE:LEFT_COLS:RIGHT_PASS:RIGHT_COLS[:FUSED_VARIANT[:LEFT_PASS]]",
                    help="pin a schedule instead of autotuning, e.g. perpass:16:4:4 or fusedT:32:3:16:12 "
                         "(as the autotuner reports it)")
    ap.add_argument("--workload", choices=("single", "batched", "adjoint", "large", "kets", "small"), default="single",
                    help="single: the N=10^4 ramp (headline, also reports the adjoint cost ratio); batched: 64 "
                         "disorder realisations of uneven size; adjoint: the differentiation workload alone; "
                         "large: the N=28800 honeycomb; kets: Krylov ket evolution at N~10^5; small: the "
                         "N=100 point-flux ramp with observables every step")
    ap.add_argument("--no-adjoint", action="store_true", help="single workload: skip the adjoint cost ratio")
    ap.add_argument("--quench", action="store_true",
                    help="single workload: the quench variant (constant Landau field 0.01 from t=0) instead of the ramp")
    ap.add_argument("--adjoint-steps", type=int, default=0,
                    help="large workload: also record this many steps and run the edge-density adjoint")
    ap.add_argument("--checkpoint-every", type=int, default=100, help="large workload adjoint: checkpoint interval")
    ap.add_argument("--p0", choices=("product", "fermi"), default=None,
                    help="initial states: single workload: the Fermi projector at B=0, mu=0 (default) or a "
                         "seeded rank-64 projector; batched / large: random half-filled product states (default) "
                         "or Fermi projectors (the step cost does not depend on P0)")
    ap.add_argument("--batch-kernels", choices=("strip", "classes", "both"), default="both",
                    help="batched workload: time the persistent strip launch, the per-size-class launches, "
                         "or both (the line's value is the strip launch's)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also copy P once with the library's streaming copy kernel (16 N^2 B read + write), "
                         "the known-size kernel for scaling HBM counters")
    args = ap.parse_args()
    if args.p0 is None:
        args.p0 = "fermi" if args.workload == "single" else "product"
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    batched = args.workload == "batched"
    large = args.workload == "large"
    kets = args.workload == "kets"
    small = args.workload == "small"
    if args.steps is None:       # small: timed runs of its 200-step evolution
        args.steps = 30 if batched else 10 if large else 90 if kets else 5 if small else 480
    if args.warmup is None:
        args.warmup = 3 if batched else 2 if large else 10 if kets else 20
    if kets and args.warmup + args.steps > KETS_STEPS:
        print("warmup + steps must not exceed the %d-step ket workload" % KETS_STEPS, file=sys.stderr)
        sys.exit(2)
    if args.workload == "single" and args.warmup + args.steps > NSTEPS or batched and args.warmup + args.steps > BATCH_STEPS:
        print("warmup + steps must not exceed the %d-step workload" % NSTEPS, file=sys.stderr)
        sys.exit(2)

    import lmhip
    from lmhip import _lib

    if batched:
        return bench_batched(args, lmhip)
    if large:
        return bench_large(args, lmhip, _lib)
    if kets:
        return bench_kets(args, lmhip, _lib)
    if small:
        return bench_small(args, lmhip, _lib)
    if args.workload == "adjoint":
        a = adjoint_workload(lmhip, _lib)
        print(json.dumps({"metric": "fwd+adjoint cost / plain fwd (differentiation workload)", "value": a["ratio_wall"],
                          "unit": "x", "n_gpus": 1, "higher_is_better": False, "vs_baseline": None,
                          "dtype": "c128 (complex FP64)", "data": "synthetic (seeded schedule and potentials)",
                          "config": a}), flush=True)
        return
    V, P0 = build_inputs()
    lat = lmhip.SquareLattice(L, L, boundaries=(("axis1", True), ("axis2", True)))
    # the ramp, or the quench variant: the final field from t = 0
    bfield = (lambda t: BMAX) if args.quench else (lambda t: BMAX * min(t, TAU) / TAU)
    h = lambda t: lmhip.tightbinding_hamiltonian(lat, V, field=lmhip.LandauGauge(bfield(t)))
    mode = {"auto": _lib.STEP_AUTO, "perpass": _lib.STEP_PER_PASS, "fusedT": _lib.STEP_FUSED_T,
            "fusedN": _lib.STEP_FUSED_N}[args.mode]
    solver = lmhip.HIPExp(step_mode=mode)
    t_p0 = None
    if args.p0 == "fermi":
        # P0 = densitymatrix(h at B = 0, mu = 0): the Fermi projector of the
        # disordered lattice without field (rocSOLVER on the device)
        t0 = time.perf_counter()
        ctx0 = lmhip.default_context(0, L * L)
        Pd = lmhip.densitymatrix(lmhip.tightbinding_hamiltonian(lat, V), mu=0.0, ctx=ctx0)
        t_p0 = time.perf_counter() - t0
        solver.ctx = ctx0
        P0 = Pd.to_numpy() if not args.no_cpu_baseline else Pd
    ev = lmhip.Evolution(solver, h, P0)
    ctx = ev.solver.ctx
    if args.no_defer:
        _lib.check(_lib.lib().lmh_set_option(ctx.h, _lib.OPT_DEFER, 0))
    if args.blocking_order:
        _lib.check(_lib.lib().lmh_set_option(ctx.h, _lib.OPT_ASYNC_ORDER, 0))
    ts = timeline()
    it = ev(ts[: args.warmup + args.steps + 1])
    next(it)                     # t = 0: no step (dt = 0)
    for _ in range(args.warmup):
        next(it)
    lib = _lib.lib()
    import ctypes
    tuned = None
    if args.schedule:
        tuned = pin_schedule(lib, _lib, ctx, args.schedule)
        ev.solver._tuned = True           # keep the pinned options (no per-step mode reset)
    elif args.mode == "auto" and not args.no_autotune:
        # the library's autotuner picks step mode and tile shapes on this
        # problem (scratch copy of the state; outside the timed region)
        best = ctypes.c_double()
        _lib.check(lib.lmh_autotune(ctx.h, ev.states[0].h, 3, ctypes.byref(best)))
        tuned = {}
        for key, name in ((_lib.OPT_STEP_MODE, "step_mode"), (_lib.OPT_LEFT_COLS, "left_cols"),
                          (_lib.OPT_RIGHT_COLS, "right_cols"), (_lib.OPT_RIGHT_PASS, "right_pass"),
                          (_lib.OPT_FUSED_VARIANT, "fused_variant"), (_lib.OPT_LEFT_PASS, "left_pass")):
            v = ctypes.c_int64()
            _lib.check(lib.lmh_get_option(ctx.h, key, ctypes.byref(v)))
            tuned[name] = v.value
        tuned["step_mode"] = {1: "perpass", 2: "fusedT", 3: "fusedN"}.get(tuned["step_mode"], "auto")
        tuned["best_ms"] = best.value
    _lib.check(lib.lmh_sync(ctx.h))
    _lib.check(lib.lmh_set_profiling(ctx.h, 1))
    t0 = time.perf_counter()
    # (K is not read inside the loop: the updates leave it on the device,
    # and the library counts the steps' terms)
    for _ in range(args.steps):
        next(it)
    _lib.check(lib.lmh_sync(ctx.h))
    el = time.perf_counter() - t0
    n_steps = ctypes.c_int64()
    n_terms = ctypes.c_int64()
    _lib.check(lib.lmh_get_option(ctx.h, _lib.OPT_STEP_COUNT, ctypes.byref(n_steps)))
    _lib.check(lib.lmh_get_option(ctx.h, _lib.OPT_STEP_TERMS, ctypes.byref(n_terms)))
    assert n_steps.value == args.steps, (n_steps.value, args.steps)
    Ks = [n_terms.value / n_steps.value] * n_steps.value     # the mean K per step

    prof = {}
    kinds = ((_lib.KERNEL_HORNER_LEFT, "horner_left"), (_lib.KERNEL_HORNER_RIGHT, "horner_right"),
             (_lib.KERNEL_BUILD_H, "build_h"), (_lib.KERNEL_TAYLOR_ORDER, "taylor_order"),
             (_lib.KERNEL_FUSED, "fused_left"), (_lib.KERNEL_TRANSPOSE, "conj_transpose"))
    for kind, name in kinds:
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        _lib.check(lib.lmh_get_profile(ctx.h, kind, ctypes.byref(n), ctypes.byref(ms)))
        prof[name] = (n.value, ms.value)
    N = L * L
    k_sum = float(sum(Ks))
    # Per-term Horner accounting (fixed by the algorithm): per step and side,
    # K terms move (48 K - 16) N^2 bytes (each term reads Y and X and writes
    # Y; the first reads X once).  A per-term launch carries one term.  A
    # fused launch carries a whole side but must move only 32 N^2 bytes (read
    # P, write the result): its roofline is that floor, and the per-term
    # figure is reported beside it as an equivalent rate (it can exceed the
    # HBM peak, since the fused kernel does not move those bytes).
    bytes_side = (48.0 * k_sum - 16.0 * len(Ks)) * N * N
    per_term_equiv = None
    dv = ctypes.c_int64()
    _lib.check(lib.lmh_get_option(ctx.h, _lib.OPT_DEFER, ctypes.byref(dv)))
    defer = bool(dv.value)
    _lib.check(lib.lmh_get_option(ctx.h, _lib.OPT_ASYNC_ORDER, ctypes.byref(dv)))
    async_order = bool(dv.value)
    terms_per_launch = None
    if prof["fused_left"][0] > 0:
        dom = "fused_left"
        nl, ms = prof[dom]
        bytes_per_launch = 32.0 * N * N
        terms_per_launch = 2.0 * k_sum / max(nl, 1)
        accounting = ("fused floor: read P + write result = 32 N^2 per launch (%.1f Horner terms per launch: "
                      "%s)" % (terms_per_launch,
                               "each step's first side fused with the previous step's second side, one launch "
                               "per step" if terms_per_launch > 1.5 * k_sum / len(Ks)
                               else "one side of a step per launch"))
        per_term_equiv = 2.0 * bytes_side / max(nl, 1) / (ms / max(nl, 1) * 1e-3) / 1e9
    else:
        # every per-term launch (left or right, either side) carries one term
        dom = "horner (all per-term launches)"
        nl = prof["horner_left"][0] + prof["horner_right"][0]
        ms = prof["horner_left"][1] + prof["horner_right"][1]
        bytes_per_launch = 2.0 * bytes_side / max(nl, 1)
        accounting = "per-term Horner bytes: (48 K - 16) N^2 per side over K launches"
    avg_ms = ms / max(nl, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    # the fused variant the library actually launched for the last step
    fv = ctypes.c_int64()
    _lib.check(lib.lmh_get_option(ctx.h, _lib.OPT_FUSED_LAUNCHED, ctypes.byref(fv)))
    variant = fv.value
    fused_steps = dom == "fused_left" and terms_per_launch > 1.5 * k_sum / len(Ks)
    traffic, traffic_src = profiled_traffic(dom, "%d+defer" % variant if fused_steps else variant)
    h_stream = None
    if dom == "fused_left" and variant in H_STREAM_BYTES:
        hb = terms_per_launch * N * N * H_STREAM_BYTES[variant]
        h_stream = {"bytes_per_launch": hb, "GBs": hb / (avg_ms * 1e-3) / 1e9,
                    "note": "H read from L2 per site and term (%d B)" % H_STREAM_BYTES[variant]}
    # the strip kernel is not HBM-bound (its 16 Horner terms per launch stay
    # in registers and LDS): its FP64 and LDS rates beside the HBM roofline.
    # Per site and term: the diagonal product (2 multiplies), the two y bonds
    # (4 complex FMAs each), the two x bonds (real when the Landau gauge
    # keeps them real: 2 FMAs each) and the Horner update (2 FMAs) = 30
    # FLOP; LDS: one 16-byte write and two 16-byte reads
    on_chip = None
    if dom == "fused_left" and variant == _lib.FUSED_STRIP:
        site_terms = terms_per_launch * N * N
        flops = 30.0 * site_terms
        lds_b = 48.0 * site_terms
        t = avg_ms * 1e-3
        on_chip = {"fp64_TFLOPs": flops / t / 1e12, "fp64_peak_TFLOPs": FP64_PEAK_TFLOPS,
                   "fp64_frac": flops / t / 1e12 / FP64_PEAK_TFLOPS,
                   "lds_TBs": lds_b / t / 1e12, "lds_peak_TBs": LDS_PEAK_TBS, "lds_frac": lds_b / t / 1e12 / LDS_PEAK_TBS,
                   "accounting": "30 FLOP and 48 B of LDS traffic per site and Horner term (real x bonds); "
                                 "LDS peak 128 B/clk/CU x 256 CUs x 2.4 GHz"}
    # the whole step against the lower of the two floors a fused step has:
    # moving P through HBM twice (read P, write W, read W, write P': 64 N^2 B)
    # or its FP64 work, 2 K N^2 (8 z + 6) FLOP (8 per complex MAC over z = 4
    # off-diagonal neighbours, 4 for the real diagonal, 2 for the Horner add)
    k_mean = k_sum / len(Ks)
    t_hbm = 64.0 * N * N / (HBM_PEAK_GBS * 1e9)
    t_fp64 = 2.0 * k_mean * N * N * (8 * 4 + 6) / (FP64_PEAK_TFLOPS * 1e12)
    t_roof = max(t_hbm, t_fp64)
    step_roofline = {"t_roof_ms": t_roof * 1e3, "hbm_ms": t_hbm * 1e3, "fp64_ms": t_fp64 * 1e3,
                     "frac": t_roof * 1e3 / (el * 1e3 / args.steps),
                     "definition": "max(64 N^2 B / 8 TB/s, 2 K N^2 (8 z + 6) FLOP / 78.6 TFLOP/s) over ms_per_step"}
    if args.copy_calibration:
        ev.states[0].copy().close()
        _lib.check(lib.lmh_sync(ctx.h))
    # accuracy report: Tr P (conserved) and the idempotency error of the
    # projector after the timed steps
    rho, _J = lmhip.observables(ev.states[0])
    trace = float(np.sum(rho))
    idem = ev.states[0].idempotency_error()

    ms_per_step = el * 1e3 / args.steps
    out = {
        "metric": "P-evolution steps/s at N=10^4 FP64",
        "value": args.steps / el,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (complex FP64)",
        "data": "synthetic (seeded disorder, %s)" % ("P0 = Fermi projector at B=0, mu=0 (%.1f s)" % t_p0
                                                      if t_p0 is not None else "rank-64 projector P0"),
        "config": {"workload": "SquareLattice 100x100 PBC, %s, W=1 disorder, dt=0.05, steps %d..%d of 500"
                               % ("quench to Landau 0.01 at t=0" if args.quench else "Landau ramp to 0.01 over tau=10",
                                  args.warmup + 1, args.warmup + args.steps),
                   "N": N, "K_mean": k_sum / len(Ks), "step_mode": args.mode, "autotuned": tuned,
                   "deferred_second_pass": defer, "order_on_device": async_order},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": dom,
                     "avg_launch_ms": avg_ms, "launches": nl, "accounting": accounting,
                     "per_term_equivalent_GBs": per_term_equiv, "fused_variant": variant if dom == "fused_left"
                     else None, "l2_h_stream": h_stream, "on_chip": on_chip, "step": step_roofline},
        "kernels_ms_total": {k: v[1] for k, v in prof.items()},
        "trace_P": trace,
        "idempotency_error": idem,
    }
    if not args.no_adjoint:
        out["adjoint"] = adjoint_workload(lmhip, _lib)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(V, P0, min(args.cpu_threads, os.cpu_count() or 1))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
