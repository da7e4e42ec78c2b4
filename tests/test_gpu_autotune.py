"""The autotuner (half_amd.autotune) on the 2-atom Si cell: every key it
chooses is a valid tuning key holding the chosen value, every sweep was
timed, and the operator applied with the chosen settings still matches the
golden H psi, S psi images (reference Fortran).  All tuning keys are restored afterwards."""
import numpy as np
import pytest
from conftest import golden, psi_block

import half_amd as H
from half_amd import autotune

pytestmark = pytest.mark.gpu

SI2 = ("tests/golden/inputs/si2.CHGCAR", "tests/golden/inputs/si2.POTCAR")
PER_PREC = ["local.{p}.tile.x", "local.{p}.tile.y", "local.{p}.tile.z", "local.{p}.xgroup", "local.{p}.xconv",
            "local.{p}.fused", "local.{p}.ygroup", "local.{p}.vxlds", "contract.{p}.gram.wt",
            "contract.{p}.gram.fill", "contract.{p}.tall.wt", "contract.{p}.tall.order"]


def test_autotune_si2_keys_and_result():
    keys = [k.format(p=p) for p in ("fp64", "fp32") for k in PER_PREC] + ["operator.overlap", "operator.projector",
                                                                         "contract.fp64.gram.gauss",
                                                                         "contract.fp64.tall.gauss"]
    saved = {k: H.get_tuning(k) for k in keys}
    try:
        print("[gpu] autotune si2", flush=True)
        res = autotune.tune("si2", steps=2, verbose=False)
        best = res["best"]
        for p in ("fp64", "fp32"):
            for k in PER_PREC:
                assert k.format(p=p) in best
            assert f"{p}.tuned_step_ms" in res["timings"] and res["timings"][f"{p}.tuned_step_ms"] > 0
            assert f"{p}.xconv=2" in res["timings"]
            assert all(f"{p}.tall.order={o}" in res["timings"] for o in (0, 1, 2))
        assert "fp64.tall.gauss=1,wt=2" in res["timings"] and "fp32.tall.gauss=None,wt=2" in res["timings"]
        assert best["contract.fp64.gram.gauss"] in (0, 1) and best["contract.fp64.tall.gauss"] in (0, 1)
        # the FP32 choices are in effect now; the shared operator keys hold the FP32 sweep's
        for k in PER_PREC:
            assert H.get_tuning(k.format(p="fp32")) == best[k.format(p="fp32")]
        assert H.get_tuning("operator.projector") == best["operator.projector.fp32"]
        # with every choice applied, the FP64 operator still gives the golden images
        for k in PER_PREC:
            H.set_tuning(k.format(p="fp64"), best[k.format(p="fp64")])
        g = golden("si2_pbe")
        ctx = H.Context(*SI2, g["encut"], g["xc"], H.BACKEND_CUDA, H.SOLVER_EVD)
        try:
            hp, sp = ctx.apply_hs(psi_block(g["npw"]), g["kpoint"])
        finally:
            ctx.close()
        ref = np.load("tests/golden/si2_pbe.apply.npy")
        assert np.abs(hp - ref[0]).max() / np.abs(ref[0]).max() < 1e-11
        assert np.abs(sp - ref[1]).max() < 1e-12
    finally:
        for k, v in saved.items():
            H.set_tuning(k, v)
