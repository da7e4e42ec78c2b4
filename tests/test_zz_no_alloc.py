"""half_amd_operator_apply_noalloc: H.psi / S.psi with every temporary in the
library's per-device workspace array (no device allocation), through the C
ABI, against half_amd_operator_apply on the same device states -- two calls
in a row, the second with fewer bands, FP64 and FP32, on the 2-atom Si cell,
the HfO2-like cell (bands in two batches of the 128 MiB array) and the
216-atom cell (96^3 boxes, 1,728 projectors: nine FP64 bands per batch).
Tolerances: FP64 1e-13 and FP32 2e-6 relative to the largest |H psi| (the
allocation-free form always takes the LDS-tiled projector contractions, the
other entry point the skinny kernels where they apply: a different summation
order)."""
import ctypes as C

import numpy as np
import pytest
from conftest import workload_files

import half_amd as H
from half_amd import hip, synth

pytestmark = pytest.mark.gpu

SI2 = ("tests/golden/inputs/si2.CHGCAR", "tests/golden/inputs/si2.POTCAR")


def _call(op, nb, x, hp, sp, ld, stream):
    e = C.create_string_buffer(256)
    st = op.lib.half_amd_operator_apply_noalloc(op.handle, nb, C.c_void_p(x), ld, C.c_void_p(hp), C.c_void_p(sp),
                                                ld, C.c_void_p(stream), e, 256)
    return st, e.value.decode()


@pytest.mark.parametrize("case,counts", [("si2", (24, 9)), ("hfo2", (100, 37)), ("si216", (20, 7))])
def test_noalloc_apply_matches_apply(case, counts, tmp_path_factory):
    if case == "si2":
        chg, pot, encut, xc = *SI2, 400.0, "pbe"
    else:
        w = synth.workload(case)
        chg, pot = workload_files(case, str(tmp_path_factory.mktemp(case)))
        encut, xc = w.encut, w.xc
    print(f"[gpu] create {case}", flush=True)
    ctx = H.Context(chg, pot, encut, xc, H.BACKEND_CUDA, H.SOLVER_EVD)
    stream = hip.Stream()
    rng = np.random.default_rng(17)
    try:
        for prec, dt, tol in ((64, np.complex128, 1e-13), (32, np.complex64, 2e-6)):
            op = ctx.operator((0, 0, 0), prec)
            n = op.npw
            nmax = counts[0]
            psi = (rng.standard_normal((nmax, n)) + 1j * rng.standard_normal((nmax, n))).astype(dt)
            x = hip.Buffer.from_numpy(psi)
            bufs = [hip.Buffer(x.shape, dt) for _ in range(4)]
            for call, nb in enumerate(counts):
                # outputs start from garbage: the entry point overwrites them
                for b in bufs[:2]:
                    hip.memset(b, 0x7f)
                st, msg = _call(op, nb, x.ptr, bufs[0].ptr, bufs[1].ptr, n, stream.ptr)
                assert st == H.SUCCESS, msg
                stream.synchronize()
                op.apply_device(nb, x.ptr, bufs[2].ptr, bufs[3].ptr, stream=stream.ptr)
                stream.synchronize()
                h0, s0 = bufs[2].numpy()[:nb], bufs[3].numpy()[:nb]
                h1, s1 = bufs[0].numpy()[:nb], bufs[1].numpy()[:nb]
                scale = np.abs(h0).max()
                eh = np.abs(h1 - h0).max() / scale
                es = np.abs(s1 - s0).max() / np.abs(s0).max()
                print(f"[gpu] {case} fp{prec} call {call} ({nb} bands): rel |dH psi| {eh:.2e} |dS psi| {es:.2e}",
                      flush=True)
                assert eh < tol and es < tol
            # the calls take no device memory (the other entry point's lazily
            # grown workspace exists by now and does not change either)
            free0 = hip.mem_info()[0]
            for _ in range(3):
                st, msg = _call(op, counts[0], x.ptr, bufs[0].ptr, bufs[1].ptr, n, stream.ptr)
                assert st == H.SUCCESS, msg
            stream.synchronize()
            assert hip.mem_info()[0] == free0
            # argument checks
            st, msg = _call(op, 0, x.ptr, bufs[0].ptr, bufs[1].ptr, n, stream.ptr)
            assert st == H.INVALID_ARGUMENT
            e = C.create_string_buffer(256)
            st = op.lib.half_amd_operator_apply_noalloc(op.handle, 2, C.c_void_p(x.ptr), n - 1,
                                                        C.c_void_p(bufs[0].ptr), C.c_void_p(bufs[1].ptr), n, None,
                                                        e, 256)
            assert st == H.CAPACITY and b"leading dimension" in e.value
            st = op.lib.half_amd_operator_apply_noalloc(987654321, 2, C.c_void_p(x.ptr), n, C.c_void_p(bufs[0].ptr),
                                                        C.c_void_p(bufs[1].ptr), n, None, e, 256)
            assert st == H.INVALID_HANDLE
            op.close()
            for b in [x] + bufs:
                b.free()
    finally:
        ctx.close()
