"""Autotuner for the tile shapes and forms of the H.psi operator.

For one workload it times the FP64 and FP32 operators on a resident band
block while sweeping, per precision:
  local.<p>.tile.{x,y,z}   FFT points per workgroup tile of each pass
  local.<p>.xgroup         bands sharing one Veff tile in the x pass
  local.<p>.xconv          x pass as transforms (0), a per-column convolution on
                           the vector units (1) or on the matrix cores (2: four
                           waves per workgroup, 3: eight, three-MFMA products)
  local.<p>.fused          y, x, y passes fused per (band, i3) slice or not
  local.<p>.ygroup         bands per workgroup of the fused kernel
  local.<p>.vxlds          fused kernel with the Vx rows staged in LDS or read from L2
  operator.overlap         projector stage on a second stream or in order
  operator.projector       skinny projector kernels (B D, B Q formed at build)
                           or the LDS-tiled contractions
  contract.<p>.gram.wt     tile width of the coefficient contraction B^H psi
  contract.<p>.gram.fill   round fill its row-chunk count must reach
  contract.<p>.tall.wt     tile width of the expansion B (D c), B (Q c)
  contract.<p>.tall.order  its workgroup order: row tiles first (0), column
                           tiles first (1), 1 with an XCD-contiguous renumbering (2)
  contract.fp64.*.gauss    FP64 complex products with three MFMAs or four
The tiles are swept with the separate passes (fused off, so every pass runs):
each candidate is applied to all three passes at once and every pass keeps
the value that minimised its own traced time (the passes are independent
kernels); xgroup and xconv are swept afterwards with the chosen tiles, the
contraction keys by the traced time of their own kernel (generic projector
path).  The forms are then chosen by the whole operator's event-timed step
(no launch timeline, so the overlap is in effect).

    python -m half_amd.autotune --workload hfo2 --bands 64 --out tuning.json

The result maps each key to its best value and records all timings; the
library's defaults are set from it (csrc/local.hip, struct Tuning).
"""
import argparse
import json
import os
import sys
import tempfile

import numpy as np

from . import BACKEND_CUDA, SOLVER_EVD, Context, get_tuning, hip, set_tuning, synth, trace, trace_read

PASSES = {"x": ("local_x_veff", "local_x_conv", "local_x_mfma"), "y": ("local_y_bwd", "local_y_fwd"), "z": ("local_z_bwd", "local_z_fwd")}


def _pass_ms(kern, axis, steps):
    names = PASSES[axis]
    names = (names,) if isinstance(names, str) else names
    return sum(kern[n][0] for n in names if n in kern) / steps


def tune(workload="hfo2", bands=0, steps=10, verbose=True):
    w = synth.workload(workload)
    nb = bands or w.nbands
    tmp = tempfile.mkdtemp(prefix="halftune")
    chg, pot, _ = synth.make_inputs(workload, tmp)
    ctx = Context(chg, pot, w.encut, w.xc, BACKEND_CUDA, SOLVER_EVD)
    k = tuple(w.kpoints[0])
    stream = hip.Stream()
    rng = np.random.default_rng(3)
    result = {"workload": workload, "bands": nb, "best": {}, "timings": {}}
    for prec, dt, cands in ((64, np.complex128, (256, 384, 512, 768, 1024, 1536)),
                            (32, np.complex64, (512, 768, 1024, 1536, 2048, 3072))):
        p = f"fp{prec}"
        op = ctx.operator(k, prec)
        n = op.npw
        x = hip.Buffer.from_numpy((rng.standard_normal((nb, n)) + 1j * rng.standard_normal((nb, n))).astype(dt))
        hx, sx = hip.Buffer(x.shape, dt), hip.Buffer(x.shape, dt)

        def step_ms():
            for _ in range(2):
                op.apply_device(nb, x.ptr, hx.ptr, sx.ptr, stream=stream.ptr)
            stream.synchronize()
            t0, t1 = hip.Event(), hip.Event()
            t0.record(stream)
            for _ in range(steps):
                op.apply_device(nb, x.ptr, hx.ptr, sx.ptr, stream=stream.ptr)
            t1.record(stream)
            return t0.elapsed_ms(t1) / steps

        def run():
            for _ in range(2):
                op.apply_device(nb, x.ptr, hx.ptr, sx.ptr, stream=stream.ptr)
            stream.synchronize()
            trace(True)
            for _ in range(steps):
                op.apply_device(nb, x.ptr, hx.ptr, sx.ptr, stream=stream.ptr)
            stream.synchronize()
            kern = trace_read()
            trace(False)
            return kern

        set_tuning(f"local.{p}.fused", 0)
        best = {a: (None, float("inf")) for a in "xyz"}
        for c in cands:
            for a in "xyz":
                set_tuning(f"local.{p}.tile.{a}", c)
            try:
                kern = run()
            except Exception as exc:  # a tile too large for this box is refused, not run
                result["timings"][f"{p}.tile={c}"] = str(exc)
                continue
            for a in "xyz":
                ms = _pass_ms(kern, a, steps)
                result["timings"][f"{p}.tile.{a}={c}"] = ms
                if ms < best[a][1]:
                    best[a] = (c, ms)
        for a in "xyz":
            set_tuning(f"local.{p}.tile.{a}", best[a][0])
            result["best"][f"local.{p}.tile.{a}"] = best[a][0]
        gbest = (1, float("inf"))
        for g in (1, 2, 4, 8):
            set_tuning(f"local.{p}.xgroup", g)
            kern = run()
            ms = _pass_ms(kern, "x", steps)
            result["timings"][f"{p}.xgroup={g}"] = ms
            if ms < gbest[1]:
                gbest = (g, ms)
        set_tuning(f"local.{p}.xgroup", gbest[0])
        result["best"][f"local.{p}.xgroup"] = gbest[0]
        xbest = (1, float("inf"))
        for xc in (0, 1, 2, 3):
            set_tuning(f"local.{p}.xconv", xc)
            ms = _pass_ms(run(), "x", steps)
            result["timings"][f"{p}.xconv={xc}"] = ms
            if ms < xbest[1]:
                xbest = (xc, ms)
        set_tuning(f"local.{p}.xconv", xbest[0])
        result["best"][f"local.{p}.xconv"] = xbest[0]
        # contraction tiles on the generic projector path
        set_tuning("operator.projector", 0)
        # FP64 also chooses between four- and three-MFMA complex products
        gauss = (0, 1) if prec == 64 else (None,)
        cbest = ((None, 0, 90), float("inf"))
        for ga in gauss:
            for wt in (0, 1, 2):
                for fill in (70, 90, 100):
                    if ga is not None:
                        set_tuning("contract.fp64.gram.gauss", ga)
                    set_tuning(f"contract.{p}.gram.wt", wt)
                    set_tuning(f"contract.{p}.gram.fill", fill)
                    kern = run()
                    ms = kern["proj_coeff_mfma"][0] / steps if "proj_coeff_mfma" in kern else float("inf")
                    result["timings"][f"{p}.gram.gauss={ga},wt={wt},fill={fill}"] = ms
                    if ms < cbest[1]:
                        cbest = ((ga, wt, fill), ms)
        ga, wt, fill = cbest[0]
        if ga is not None:
            set_tuning("contract.fp64.gram.gauss", ga)
            result["best"]["contract.fp64.gram.gauss"] = ga
        set_tuning(f"contract.{p}.gram.wt", wt)
        set_tuning(f"contract.{p}.gram.fill", fill)
        result["best"][f"contract.{p}.gram.wt"], result["best"][f"contract.{p}.gram.fill"] = wt, fill
        tbest = ((None, 0), float("inf"))
        for ga in gauss:
            for wt in (0, 1, 2):
                if ga is not None:
                    set_tuning("contract.fp64.tall.gauss", ga)
                set_tuning(f"contract.{p}.tall.wt", wt)
                kern = run()
                ms = kern["proj_expand_mfma"][0] / steps if "proj_expand_mfma" in kern else float("inf")
                result["timings"][f"{p}.tall.gauss={ga},wt={wt}"] = ms
                if ms < tbest[1]:
                    tbest = ((ga, wt), ms)
        ga, wt = tbest[0]
        if ga is not None:
            set_tuning("contract.fp64.tall.gauss", ga)
            result["best"]["contract.fp64.tall.gauss"] = ga
        set_tuning(f"contract.{p}.tall.wt", wt)
        result["best"][f"contract.{p}.tall.wt"] = wt
        obest = (1, float("inf"))
        for order in (0, 1, 2):
            set_tuning(f"contract.{p}.tall.order", order)
            kern = run()
            ms = kern["proj_expand_mfma"][0] / steps if "proj_expand_mfma" in kern else float("inf")
            result["timings"][f"{p}.tall.order={order}"] = ms
            if ms < obest[1]:
                obest = (order, ms)
        set_tuning(f"contract.{p}.tall.order", obest[0])
        result["best"][f"contract.{p}.tall.order"] = obest[0]
        # the Vx rows go to LDS only with band groups of two or more
        forms = ([(0, 1, 0, ov, pj) for ov in (0, 1) for pj in (0, 1)]
                 + [(1, 1, 0, ov, pj) for ov in (0, 1) for pj in (0, 1)]
                 + [(1, g, v, ov, pj) for g in (2, 4) for v in (0, 1) for ov in (0, 1) for pj in (0, 1)])
        fbest = (None, float("inf"))
        for fused, g, v, ov, pj in forms:
            set_tuning(f"local.{p}.fused", fused)
            set_tuning(f"local.{p}.ygroup", g)
            set_tuning(f"local.{p}.vxlds", v)
            set_tuning("operator.overlap", ov)
            set_tuning("operator.projector", pj)
            ms = step_ms()
            result["timings"][f"{p}.fused={fused},ygroup={g},vxlds={v},overlap={ov},projector={pj}"] = ms
            if ms < fbest[1]:
                fbest = ((fused, g, v, ov, pj), ms)
        (fused, g, v, ov, pj) = fbest[0]
        set_tuning(f"local.{p}.fused", fused)
        set_tuning(f"local.{p}.ygroup", g)
        set_tuning(f"local.{p}.vxlds", v)
        set_tuning("operator.overlap", ov)
        set_tuning("operator.projector", pj)
        result["best"][f"local.{p}.fused"] = fused
        result["best"][f"local.{p}.ygroup"] = g
        result["best"][f"local.{p}.vxlds"] = v
        result["best"][f"operator.overlap.{p}"] = ov
        result["best"][f"operator.projector.{p}"] = pj
        result["timings"][f"{p}.tuned_step_ms"] = fbest[1]
        if verbose:
            print(f"autotune {p}: {result['best']} step {result['timings'][f'{p}.tuned_step_ms']:.4f} ms",
                  file=sys.stderr, flush=True)
        op.close()
        x.free(), hx.free(), sx.free()
    ctx.close()
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="hfo2")
    ap.add_argument("--bands", type=int, default=0)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    res = tune(args.workload, args.bands, args.steps)
    text = json.dumps(res, indent=1)
    if args.out:
        with open(args.out, "w") as fh:
            fh.write(text)
    print(text)


if __name__ == "__main__":
    main()
