"""H.psi band-application throughput and all-band solve time on one MI355X.

Workload (default): the 216-atom Si cell (3x3x3 supercell of the 8-atom
diamond cell with seeded +-0.03 A positional noise), Gamma, 245.5 eV, 37,705
plane waves, 96^3 box, 1,728 projectors, PBE, 512 bands; seeded synthetic
density and PAW datasets (no real inputs exist offline).  `--workload hfo2`
selects the 12-atom HfO2-like cell (3,407 plane waves, 48^3, 64 bands).

One step = one application of H and S (FP64, complex128) to the resident
512-band block: scatter, backward FFT, Veff multiply, forward FFT, gather +
kinetic, projector contraction B^H psi on the matrix cores, D/Q mixing,
expansion B(Dc), B(Qc).  value = band applications per second over the timed
steps (inputs resident in HBM, no host traffic in the timed region).

Also reported: the FP32 operator's rate; the all-band solve times of the FP64
and the mixed form under the same solver settings (locking, preconditioner)
and of the plain FP64 form (the reference's iteration: full 2m subspace,
reference preconditioner, Cholesky-reduced Rayleigh-Ritz steps), each with
its eigenvalue error against a tightly
converged FP64 solve (or the reference's golden eigenvalues where they cover
the bands); the dominant kernel's roofline and every operator kernel's; and
the oracle's CPU rate on a bounded sample of the same workload with all host
threads.
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "deepaw-half_amd"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
MFMA_FP64_TFLOPS = 78.6        # dense FP64 matrix
MFMA_FP32_TFLOPS = 157.3


def algorithmic_bytes(nr, n, fp64=True):
    """Per band application: each FFT pass pair reads and writes the box
    (2 transforms x (read + write)), Veff read once, and the plane-wave side
    psi, H psi, S psi, kinetic and the int32 index map."""
    sc, sr = (16, 8) if fp64 else (8, 4)
    return 4 * nr * sc + nr * sr + n * (3 * sc + sr + 4)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="si216")
    ap.add_argument("--bands", type=int, default=0)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 solve (the reference's iteration) beside the chosen form")
    ap.add_argument("--extra", type=int, default=-1, help="buffer bands of the all-band solves (-1: bands / 4, "
                                                          "at least 8)")
    ap.add_argument("--tol", type=float, default=1e-5, help="residual tolerance (eV) of the all-band solves")
    ap.add_argument("--switch-tol", type=float, default=0.0,
                    help="residual (eV) at which the mixed solve's FP32 phase hands over to FP64 sweeps "
                         "(0: the library's default)")
    ap.add_argument("--ref-tol", type=float, default=1e-7,
                    help="residual tolerance (eV) of the untimed FP64 solve the eigenvalue errors are measured "
                         "against, where no golden eigenvalues cover the bands")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="length of the CPU baseline sample")
    ap.add_argument("--precond", type=int, default=1,
                    help="preconditioner of the all-band solves (solver.precond: 0 the reference's, 1 "
                         "Teter-Payne-Allan)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="set a kernel tunable (half_amd_set_tuning) before measuring; repeatable")
    ap.add_argument("--kmesh", type=int, default=0, metavar="NK",
                    help="only time the batched all-band solve of the workload's first NK mesh k points "
                         "(-1: all) against per-k solves, and print that line")
    ap.add_argument("--serial", type=int, default=-1, metavar="NK",
                    help="with --kmesh: per-k comparison solves on the first NK points (-1: all)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the byte-counter calibration "
                         "for profiles) and print their rate")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: one GPU only; --gpus {args.gpus} is not supported (the library runs on one "
              "GPU; sharding k points and bands across GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import half_amd as H
    from half_amd import hip, synth

    H.set_tuning("solver.precond", args.precond)
    for kv in args.tune:
        key, value = kv.split("=")
        H.set_tuning(key, int(value))

    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        src, dst = hip.Buffer((nbytes,), np.uint8), hip.Buffer((nbytes,), np.uint8)
        stream = hip.Stream()
        hip.memset(src, 1)
        t0, t1 = hip.Event(), hip.Event()
        t0.record(stream)
        for _ in range(5):
            hip.copy_async(dst, src, nbytes, stream)
        t1.record(stream)
        ms = t0.elapsed_ms(t1) / 5
        print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": ms, "GBps": 2 * nbytes / ms / 1e6}))
        return

    if args.kmesh:
        kmesh_bench(args)
        return

    w = synth.workload(args.workload)
    nb = args.bands or w.nbands
    # buffer bands: 216-atom cell, 512 bands, locking mode 2, TPA, no buffer
    # correction columns (solver.bufcols 0): 96 of them took 50 iterations
    # (FP64 1.23 s, mixed 1.12 s), 128 took 39 (1.19 s, 1.09 s), 160 36
    # (1.19, 1.09 s), 192 28 (1.15, 1.09 s), 256 27 (1.16, 1.13 s); with
    # every buffer band's column 128 took 25 (1.39, 1.27 s), 64 took 63 and
    # 48 left the FP64 solve stagnating above tol for 600 iterations
    extra = args.extra if args.extra >= 0 else max(8, nb // 4)
    tmp = tempfile.mkdtemp(prefix="halfbench")
    chg, pot, _ = synth.make_inputs(args.workload, tmp)
    dev_name, cus = H.device_info()
    print(f"bench: {dev_name} ({cus} CUs), workload {args.workload}", file=sys.stderr, flush=True)
    ctx = H.Context(chg, pot, w.encut, w.xc, H.BACKEND_CUDA, H.SOLVER_EVD)
    k = tuple(w.kpoints[0])
    info = ctx.system_info()
    nr = int(np.prod(info["grid"]))
    stream = hip.Stream()
    rng = np.random.default_rng(11)

    rates = {}
    per_step = {}
    kernels = {}
    for prec, dt in ((64, np.complex128), (32, np.complex64)):
        op = ctx.operator(k, prec)
        n = op.npw
        x = hip.Buffer.from_numpy((rng.standard_normal((nb, n)) + 1j * rng.standard_normal((nb, n))).astype(dt))
        hx, sx = hip.Buffer(x.shape, dt), hip.Buffer(x.shape, dt)

        def step():
            op.apply_device(nb, x.ptr, hx.ptr, sx.ptr, ld=n, stream=stream.ptr)

        for _ in range(args.warmup):
            step()
        stream.synchronize()
        t0, t1 = hip.Event(), hip.Event()
        t0.record(stream)
        for _ in range(args.steps):
            step()
        t1.record(stream)
        ms = t0.elapsed_ms(t1) / args.steps
        per_step[prec] = ms
        rates[prec] = nb / (ms * 1e-3)
        print(f"bench: fp{prec} H.psi {ms:.3f} ms/step, {rates[prec]:.1f} band-applications/s",
              file=sys.stderr, flush=True)
        # same steps again with the launch-boundary timeline on: per-kernel durations
        H.trace(True)
        for _ in range(args.steps):
            step()
        stream.synchronize()
        kernels[prec] = H.trace_read()
        H.trace(False)
        op.close()
        x.free(), hx.free(), sx.free()
    n = ctx.basis_size(k)
    nproj = ctx.operator(k, 64).nproj

    solve = {}
    evals = {}
    if not args.no_solve:
        ref, ref_kind = None, None
        gpath = os.path.join(ROOT, "tests", "golden", f"{w.extra.get('inputs', w.name)}_pbe.json")
        if os.path.exists(gpath):
            with open(gpath) as fh:
                g = json.load(fh)
            if len(g["eigenvalues"]) >= nb:
                ref, ref_kind = np.array(g["eigenvalues"][:nb]), "reference CPU dense solve (golden)"
        # warm-up: first-call library initialisation (rocBLAS/rocSOLVER kernels,
        # operator caches) stays out of the timed solves
        for mixed in (False, True):
            ctx.solve_acc(nb, k, mixed=mixed, tol=args.tol, maxiter=2, extra=extra)
        if ref is None:
            t = time.perf_counter()
            ref, _, st = ctx.solve_acc(nb, k, tol=args.ref_tol, maxiter=600, extra=extra)
            ref_kind = f"FP64 all-band solve to residual {args.ref_tol:g} eV"
            solve["reference"] = {"kind": ref_kind, "seconds": time.perf_counter() - t,
                                  "iterations": st["iterations"], "max_residual_eV": st["max_residual_eV"]}
            print(f"bench: reference solve {solve['reference']}", file=sys.stderr, flush=True)
        for mixed in (False, True):
            t = time.perf_counter()
            wv, _, st = ctx.solve_acc(nb, k, mixed=mixed, tol=args.tol, switch_tol=args.switch_tol, maxiter=400,
                                      extra=extra)
            el = time.perf_counter() - t
            key = "mixed" if mixed else "fp64"
            # the same solve again on the launch-boundary timeline: where the time goes
            H.trace(True)
            ctx.solve_acc(nb, k, mixed=mixed, tol=args.tol, switch_tol=args.switch_tol, maxiter=400, extra=extra)
            phases = H.trace_read()
            H.trace(False)
            err = np.abs(wv - ref)
            solve[key] = {"seconds": el, "iterations": st["iterations"], "fp64_iterations": st["fp64_iterations"],
                          "max_residual_eV": st["max_residual_eV"], "band_applications": st["band_applications"],
                          "rr_order_mean": st["rr_order_mean"],
                          "max_abs_err_eV": float(err.max()), "mean_abs_err_eV": float(err.mean()),
                          "bands_above_1e-8_eV": int((err > 1e-8).sum()),
                          "phase_ms": {p: round(v[0], 3) for p, v in sorted(phases.items(), key=lambda x: -x[1][0])}}
            print(f"bench: {key} all-band solve {el:.3f} s {solve[key]}", file=sys.stderr, flush=True)
            evals[key] = wv
        solve["errors_against"] = ref_kind
        if "fp64" in solve and "mixed" in solve:
            solve["speedup_mixed_vs_fp64"] = solve["fp64"]["seconds"] / solve["mixed"]["seconds"]
            solve["max_abs_diff_mixed_vs_fp64_eV"] = float(np.abs(evals["mixed"] - evals["fp64"]).max())
        solve["controls"] = {"tol_eV": args.tol, "extra_bands": extra, "maxiter": 400,
                             "switch_tol_eV": args.switch_tol or "library default",
                             "lock": H.get_tuning("solver.lock"), "precond": H.get_tuning("solver.precond")}
        if not args.no_plain and "mixed" in solve:
            # the plain FP64 form: the reference's iteration (every band's
            # correction column, the full 2m-wide Rayleigh-Ritz subspace, the
            # reference preconditioner, every Rayleigh-Ritz step through the
            # Cholesky reduction as the reference's zhegv) with the FP64
            # operator, same tolerance
            saved = {key: H.get_tuning(key) for key in ("solver.lock", "solver.precond", "solver.identity")}
            H.set_tuning("solver.lock", 0)
            H.set_tuning("solver.precond", 0)
            H.set_tuning("solver.identity", 0)
            try:
                # warm-up: the reduction's library kernels load on first use
                ctx.solve_acc(nb, k, tol=args.tol, maxiter=2, extra=extra)
                t = time.perf_counter()
                wp, _, st = ctx.solve_acc(nb, k, tol=args.tol, maxiter=800, extra=extra)
                el = time.perf_counter() - t
            finally:
                for key, v in saved.items():
                    H.set_tuning(key, v)
            err = np.abs(wp - ref)
            solve["plain_fp64"] = {"kind": "FP64 operator, the reference's iteration: full 2m subspace (solver.lock "
                                           "0), reference preconditioner (solver.precond 0), Cholesky-reduced "
                                           "Rayleigh-Ritz steps (solver.identity 0)",
                                   "seconds": el, "iterations": st["iterations"],
                                   "max_residual_eV": st["max_residual_eV"], "rr_order_mean": st["rr_order_mean"],
                                   "max_abs_err_eV": float(err.max())}
            # the chosen form (mixed precision with the locking and
            # preconditioner of `controls`) against it; the like-for-like
            # ratio is speedup_mixed_vs_fp64
            solve["speedup_chosen_vs_plain_fp64"] = el / solve["mixed"]["seconds"]
            print(f"bench: plain FP64 all-band solve {solve['plain_fp64']}", file=sys.stderr, flush=True)

    if not args.no_solve and n <= 6000:
        # dense FP64 solve (assembly, overlap spectrum, generalized eigensolve),
        # the reference CPU path's algorithm, for comparison with its timing
        ctx.solve(nb, k, vectors=False)
        t = time.perf_counter()
        wd, _, _, _ = ctx.solve(nb, k, vectors=False)
        solve["dense_evd"] = {"seconds": time.perf_counter() - t, "bands": nb}
        print(f"bench: dense solve {solve['dense_evd']}", file=sys.stderr, flush=True)
    gv, _ = ctx.basis(k)
    nlines, nplanes = sphere_counts(gv, info["grid"])
    roofline = kernel_roofline(kernels[64], nr, info["grid"], n, nproj, nb, args.steps, nlines, nplanes, True,
                               args.workload)
    roofline32 = kernel_roofline(kernels[32], nr, info["grid"], n, nproj, nb, args.steps, nlines, nplanes, False,
                                 args.workload)
    # every operator kernel against its own bound (MFMA for the projector
    # contractions, HBM for the local stage)
    per_kernel = {}
    for prec in kernels:
        per_kernel[str(prec)] = {}
        for kname in kernels[prec]:
            r = kernel_roofline(kernels[prec], nr, info["grid"], n, nproj, nb, args.steps, nlines, nplanes,
                                prec == 64, args.workload, name=kname)
            per_kernel[str(prec)][kname] = {q: r.get(q) for q in ("bound", "avg_launch_ms", "achieved", "unit", "frac",
                                                              "traffic")}
    b_alg = algorithmic_bytes(nr, n, True)
    f_proj = 24.0 * n * nproj
    t_band = per_step[64] * 1e-3 / nb
    roofline["operator_frac"] = (b_alg / (HBM_PEAK_GBS * 1e9) + f_proj / (MFMA_FP64_TFLOPS * 1e12)) / t_band
    # the same convention per stage and precision, from the traced kernel
    # times: the local stage against HBM, the projector stage against MFMA
    stage_fracs = {}
    for prec, fp64 in ((64, True), (32, False)):
        kt = kernels[prec]
        t_loc = sum(v[0] for q, v in kt.items() if q.startswith("local_")) * 1e-3 / args.steps
        t_prj = sum(v[0] for q, v in kt.items() if q.startswith("proj_")) * 1e-3 / args.steps
        peak = MFMA_FP64_TFLOPS if fp64 else MFMA_FP32_TFLOPS
        stage_fracs[f"fp{prec}"] = {
            "local_vs_hbm": (nb * algorithmic_bytes(nr, n, fp64) / (HBM_PEAK_GBS * 1e9)) / t_loc if t_loc else None,
            "projector_vs_mfma": (nb * f_proj / (peak * 1e12)) / t_prj if t_prj else None}
    roofline["stage_fracs"] = stage_fracs

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(chg, pot, w, k, args.cpu_seconds)
        if not args.no_solve:
            cpu["acc_iteration"] = cpu_acc_iteration(chg, pot, w, k, nb + extra)
            plain = solve.get("plain_fp64")
            if plain:
                cpu["acc_iteration"]["gpu_plain_fp64_s"] = plain["seconds"] / max(1, plain["iterations"])

    out = {
        "metric": "H.psi band-applications/s + all-band solve time at 1e-8 eV, mixed vs FP64, 1 GPU",
        "value": rates[64], "unit": "band-applications/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": per_step[64], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (seeded Gaussian density, synthetic PAW datasets)",
        "config": {"workload": f"{args.workload}: {sum(w.counts)} atoms, {n} plane waves, box {info['grid']}, "
                               f"{nproj} projectors, {nb} bands (+{extra} buffer bands in the solves), Gamma, "
                               f"{w.xc.upper()}",
                   "npw": n, "box": list(info["grid"]), "nproj": nproj, "bands": nb, "extra_bands": extra,
                   "active_lines": nlines, "active_planes": nplanes},
        "fp32_value": rates[32], "fp32_ms_per_step": per_step[32],
        "all_band_solve": solve,
        "roofline": roofline,
        "roofline_fp32": roofline32,
        "kernels_ms_per_step": {str(p): {k: v[0] / args.steps for k, v in kernels[p].items()} for p in kernels},
        "kernel_rooflines": per_kernel,
        "cpu_baseline": cpu,
    }
    if args.tune:
        out["tuning"] = dict(kv.split("=") for kv in args.tune)
    print(json.dumps(out))


def sphere_counts(gvec, grid):
    """Active (i1, i2) lines and i1 planes of the plane-wave sphere."""
    import numpy as np
    i1 = np.mod(gvec[:, 0], grid[0])
    i2 = np.mod(gvec[:, 1], grid[1])
    return len(set(zip(i1.tolist(), i2.tolist()))), len(set(i1.tolist()))


def kernel_bytes(name, nr, grid, n, nlines, nplanes, nb, sc, sr):
    """Bytes a kernel must move per step (nb bands) in the pruned scheme:
    each active line/plane read and written once, Veff read once per band
    (transform form of the x pass) or its i1-axis DFT once per launch
    (convolution forms), plane-wave vectors read/written once.  The fused
    y-x-y kernel keeps the planes in LDS: its HBM bytes are the active
    lines in and out."""
    n1, n2, n3 = grid
    line = nlines * n3 * sc
    plane = nplanes * n2 * n3 * sc
    table = {
        "local_z_bwd": n * (sc + 8) + line,
        "local_y_bwd": line + plane,
        "local_x_veff": 2 * plane + nr * sr,
        "local_x_conv": 2 * plane,
        "local_x_mfma": 2 * plane,
        "local_y_fwd": plane + line,
        "local_z_fwd": line + n * (3 * sc + sr + 8),
        # fused y-x-y pass: each active line's values read and written once
        "local_yxy": 2 * line,
    }
    # the convolution forms read the i1-axis DFT of Veff (nr complex) once per launch
    once = {"local_x_conv": nr * sc, "local_x_mfma": nr * sc, "local_yxy": nr * sc}
    return None if name not in table else nb * table[name] + once.get(name, 0)


def measured_traffic(workload, nb, fp64, name):
    """HBM bytes per launch of a kernel from the committed counter summary
    (profiles/hbm_traffic_<workload>.json: FETCH_SIZE / WRITE_SIZE scaled by
    a copy-kernel calibration) when it covers this workload and band count."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    if t.get("bands") != nb:
        return None
    return t["bytes_per_launch"].get("fp64" if fp64 else "fp32", {}).get(name)


def kernel_roofline(kern, nr, grid, n, nproj, nb, steps, nlines, nplanes, fp64=True, workload=None, name=None):
    """Roofline of one traced kernel (default: the one with the largest share
    of the traced time): its algorithmic bytes (or flops) over its summed
    launch durations."""
    sc, sr = (16, 8) if fp64 else (8, 4)
    per_step_flops = {"proj_coeff_mfma": 8.0 * n * nproj * nb, "proj_expand_mfma": 16.0 * n * nproj * nb}
    if name is None:
        name = max(kern, key=lambda k: kern[k][0])
    total_ms, launches = kern[name]
    t = total_ms * 1e-3
    out = {"kernel": name, "avg_launch_ms": total_ms / launches, "launches_per_step": launches / steps}
    if name in per_step_flops:
        peak = MFMA_FP64_TFLOPS if fp64 else MFMA_FP32_TFLOPS
        ach = per_step_flops[name] * steps / t / 1e12
        out.update(bound="mfma", achieved=ach, peak=peak, unit="TFLOP/s", frac=ach / peak,
                   traffic=measured_traffic(workload, nb, fp64, name) if workload else None)
        return out
    b = kernel_bytes(name, nr, grid, n, nlines, nplanes, nb, sc, sr)
    ach = b * steps / t / 1e9 if b else None
    out.update(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
               frac=(ach / HBM_PEAK_GBS) if ach else None,
               traffic=measured_traffic(workload, nb, fp64, name) if workload else None,
               bytes_per_launch=(b * steps / launches) if b else None)
    return out


def kmesh_bench(args):
    """All-band solves of every k point of the workload's mesh in one batched
    call (operator stages one launch for all k, strided-batched
    Rayleigh-Ritz, graph-replayed correction block), FP64 and mixed, against
    the same solves done k point by k point.  seconds_solver excludes the
    operator construction; seconds_wall includes it."""
    import numpy as np
    import half_amd as H
    from half_amd import synth

    w = synth.workload(args.workload)
    nb = args.bands or w.nbands
    extra = args.extra if args.extra >= 0 else 8
    ks = np.asarray(w.kpoints, float)
    if args.kmesh > 0:
        ks = ks[: args.kmesh]
    nk = len(ks)
    tmp = tempfile.mkdtemp(prefix="halfbench")
    chg, pot, _ = synth.make_inputs(args.workload, tmp)
    dev_name, cus = H.device_info()
    print(f"bench: {dev_name} ({cus} CUs), workload {args.workload}, {nk} k points", file=sys.stderr, flush=True)
    ctx = H.Context(chg, pot, w.encut, w.xc, H.BACKEND_CUDA, H.SOLVER_EVD)
    sizes = [ctx.basis_size(tuple(k)) for k in ks]
    for mixed in (False, True):  # library initialisation outside the timed solves
        ctx.solve_kmesh(nb, ks[:2], mixed=mixed, tol=args.tol, maxiter=2, extra=extra)
    out = {"metric": "k-mesh all-band solve seconds, batched vs per-k, 1 GPU", "unit": "s",
           "higher_is_better": False, "n_gpus": 1, "data": "synthetic (seeded Gaussian density, synthetic PAW "
           "datasets)", "config": {"workload": f"{args.workload}: {sum(w.counts)} atoms, {nk} k points, "
                                               f"{min(sizes)}-{max(sizes)} plane waves, box {list(w.grid)}, "
                                               f"{nb} bands per k",
                                   "nk": nk, "bands": nb, "npw_min": min(sizes), "npw_max": max(sizes)},
           "controls": {"tol_eV": args.tol, "extra_bands": extra, "maxiter": 400}}
    mesh = {}
    for mixed in (False, True):
        key = "mixed" if mixed else "fp64"
        t = time.perf_counter()
        wv, st = ctx.solve_kmesh(nb, ks, mixed=mixed, tol=args.tol, maxiter=400, extra=extra)
        el = time.perf_counter() - t
        mesh[key] = (wv, st)
        out[f"batched_{key}"] = {"seconds_wall": el, "seconds_solver": st["seconds"],
                                 "iterations": st["iterations"], "fp64_iterations": st["fp64_iterations"],
                                 "max_residual_eV": st["max_residual_eV"],
                                 "band_applications": st["band_applications"],
                                 "rr_order_mean": st["rr_order_mean"],
                                 "band_applications_per_s": st["band_applications"] / st["seconds"]}
        print(f"bench: batched {key} {out[f'batched_{key}']}", file=sys.stderr, flush=True)
    out["max_abs_diff_mixed_vs_fp64_eV"] = float(np.abs(mesh["mixed"][0] - mesh["fp64"][0]).max())
    # one mixed solve on the launch-boundary timeline (eager launches): phases
    H.trace(True)
    ctx.solve_kmesh(nb, ks, mixed=True, tol=args.tol, maxiter=400, extra=extra, graph=False)
    phases = H.trace_read()
    H.trace(False)
    out["batched_mixed_phase_ms"] = {p: round(v[0], 2) for p, v in sorted(phases.items(), key=lambda x: -x[1][0])}
    ns = nk if args.serial < 0 else min(args.serial, nk)
    if ns > 0:
        for mixed in (False, True):
            key = "mixed" if mixed else "fp64"
            solver_s = 0.0
            diff = 0.0
            t = time.perf_counter()
            for j in range(ns):
                wj, _, st = ctx.solve_acc(nb, tuple(ks[j]), mixed=mixed, tol=args.tol, maxiter=400,
                                          extra=extra)
                solver_s += st["seconds"]
                diff = max(diff, float(np.abs(wj - mesh[key][0][j]).max()))
            el = time.perf_counter() - t
            out[f"per_k_{key}"] = {"k_points": ns, "seconds_wall": el, "seconds_solver": solver_s,
                                   "max_abs_diff_vs_batched_eV": diff}
            out[f"speedup_batched_vs_per_k_{key}"] = (solver_s / ns) / (out[f"batched_{key}"]["seconds_solver"] / nk)
            print(f"bench: per-k {key} {out[f'per_k_{key}']}", file=sys.stderr, flush=True)
    out["value"] = out["batched_mixed"]["seconds_solver"]
    ctx.close()
    print(json.dumps(out))


def host_threads():
    """Host threads of this run: the box's CPU share (OMP_NUM_THREADS where
    set), at most the visible cores."""
    n = os.cpu_count() or 1
    try:
        n = min(n, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        pass
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() else n


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:  # pragma: no cover
        pass
    return "unknown"


def cpu_baseline(chg, pot, w, k, seconds):
    """The oracle's matrix-free H.psi (numpy/scipy, FP64) on the host cores
    for a bounded sample: repeated applications to a block of up to 64 bands
    until `seconds` have elapsed, FFTs (scipy.fft) and BLAS on all host
    threads of the run.  The block is as wide as memory allows comfortably:
    the projector products are matrix-matrix there (on the 216-atom cell,
    8 cores: 7.8 band-applications/s with 8-band blocks, 28 with 64)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import half_oracle as O
    from threadpoolctl import threadpool_info, threadpool_limits
    threads = host_threads()
    ctx = O.Context(chg, pot, w.encut, w.xc)
    basis, paw = ctx.prepare(k)
    rng = np.random.default_rng(5)
    nb = 64
    psi = rng.standard_normal((basis.kinetic.size, nb)) + 1j * rng.standard_normal((basis.kinetic.size, nb))
    with threadpool_limits(threads):
        blas = ", ".join(sorted({f"{i.get('internal_api')} {i.get('version')}" for i in threadpool_info()
                                 if i.get("user_api") == "blas"})) or "unknown"
        O.apply_hs(ctx.veff, basis, paw, psi, workers=threads)  # warm-up
        t0 = time.perf_counter()
        done = 0
        while time.perf_counter() - t0 < seconds:
            O.apply_hs(ctx.veff, basis, paw, psi, workers=threads)
            done += nb
        el = time.perf_counter() - t0
    return {"value": done / el, "unit": "band-applications/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "blas": blas, "fft": f"scipy.fft, {threads} workers",
            "sample": f"{done} band applications of the oracle's matrix-free H.psi (FP64, numpy/scipy) in "
                      f"{el:.1f} s, same workload, {nb}-band blocks, {threads} host threads"}


def cpu_acc_iteration(chg, pot, w, k, m):
    """One iteration of the reference's all-band iteration (the oracle's
    restatement: full 2m subspace, reference preconditioner, FP64) on the
    host cores, at the solve's band count m: the time from the start of the
    first iteration's H.psi to that of the second (corrections' H.psi in
    64-band blocks, orthogonalisation, metric eigen-decomposition, the 2m x 2m
    generalized eigensolve, rotations), all host threads for FFTs and BLAS.
    The GPU figure beside it is the plain FP64 form's time per iteration."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import half_oracle as O
    from threadpoolctl import threadpool_limits
    threads = host_threads()
    ctx = O.Context(chg, pot, w.encut, w.xc)
    basis, paw = ctx.prepare(k)
    starts = []

    def apply(x):
        starts.append(time.perf_counter())
        hx, sx = np.empty_like(x), np.empty_like(x)
        for j in range(0, x.shape[1], 64):
            hx[:, j:j + 64], sx[:, j:j + 64] = O.apply_hs(ctx.veff, basis, paw, x[:, j:j + 64], workers=threads)
        return hx, sx

    with threadpool_limits(threads):
        t0 = time.perf_counter()
        O.acc_solve(apply, basis.kinetic, m, tol=1e-12, maxiter=2)
        total = time.perf_counter() - t0
    it = starts[2] - starts[1] if len(starts) >= 3 else None
    return {"seconds": it, "cores": threads, "kind": "port",
            "sample": f"one iteration of the oracle's restatement of the reference's all-band iteration, {m} bands "
                      f"(FP64, numpy/scipy, {threads} host threads; the bounded run took {total:.1f} s)"}


if __name__ == "__main__":
    main()
