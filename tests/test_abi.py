"""Drop-in boundary checks that need no GPU: the library loads, exports every
symbol the public headers declare, and honours the reference ABI's argument
validation, status codes and error-buffer contract (reference:
tests/test_c_api.c, src/half_c_api.F90:213-288, 627-664)."""
import ctypes as C
import os
import re
import subprocess

import pytest
from conftest import ROOT

import half_amd as H

SI2 = (os.path.join(ROOT, "tests/golden/inputs/si2.CHGCAR"), os.path.join(ROOT, "tests/golden/inputs/si2.POTCAR"))


def declared_symbols():
    names = set()
    for hdr in ("half.h", "half_amd.h"):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        names |= set(re.findall(r"^\s*(?:int|const char \*)\s*(half_\w+)\s*\(", text, re.M))
    return names


def test_headers_and_binding_agree():
    assert declared_symbols() == set(H.ABI_SYMBOLS) | set(H.EXT_SYMBOLS)
    assert len(H.ABI_SYMBOLS) == 18


def test_library_exports_every_declared_symbol():
    lib = H.load()
    out = subprocess.run(["nm", "-D", "--defined-only", H.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (half_\w+)", out))
    missing = declared_symbols() - exported
    assert not missing, missing
    for name in declared_symbols():
        assert getattr(lib, name)


def test_c_boundary_program(tmp_path):
    """tests/abi_check.c: version words, a stale handle on every entry point
    that takes one, error-buffer rules, creation argument checks."""
    exe = tmp_path / "abi_check"
    subprocess.run(["gcc", "-std=c11", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "abi_check.c"), "-L", os.path.dirname(H.LIB_PATH), "-lhalf",
                    "-Wl,-rpath," + os.path.dirname(H.LIB_PATH), "-o", str(exe)], check=True)
    inputs = os.path.join(ROOT, "tests", "golden", "inputs")
    r = subprocess.run([str(exe), os.path.join(inputs, "si2.CHGCAR"), os.path.join(inputs, "si2.POTCAR")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "C ABI boundary: ok" in r.stdout


def test_reference_c_api_test_against_this_library():
    """The reference's own tests/test_c_api.c, compiled unchanged against
    include/half.h and libhalf.so by oracle/Makefile where the reference tree
    is present (build()), passes: the drop-in boundary as the reference
    tests it."""
    exe = os.path.join(ROOT, "oracle", "_ref", "test_c_api")
    if not os.path.exists(exe):
        pytest.skip("built by oracle/Makefile only where the reference sources are present")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "HALF C API 0.6.0: ok" in r.stdout


def _create(chg, pot, backend=H.BACKEND_AUTO, solver=H.SOLVER_EVD, xc=H.XC_PBE, uspp=0, encut=400.0, cap=256):
    lib = H.load()
    h = C.c_int64(7)
    e = C.create_string_buffer(cap)
    st = lib.half_create_from_files_backend(chg.encode(), pot.encode(), encut, xc, uspp, backend, solver,
                                            C.byref(h), e, cap)
    return st, h.value, e.value.decode()


def test_create_argument_validation():
    assert _create("", "")[0] == H.INVALID_ARGUMENT
    assert _create(*SI2, backend=99)[0] == H.INVALID_ARGUMENT
    assert _create(*SI2, solver=0)[0] == H.INVALID_ARGUMENT
    assert _create(*SI2, solver=5)[0] == H.INVALID_ARGUMENT
    assert _create(*SI2, xc=3)[0] == H.INVALID_ARGUMENT
    st, h, msg = _create(*SI2, encut=-1.0)
    assert st == H.INVALID_ARGUMENT and h == 0 and "ENCUT" in msg
    st, h, msg = _create(*SI2, encut=float("nan"))
    assert st == H.INVALID_ARGUMENT
    st, h, msg = _create("/nonexistent", SI2[1])
    assert st == H.IO and "charge" in msg
    st, h, msg = _create(SI2[0], "/nonexistent")
    assert st == H.IO and "potential" in msg
    # this library carries no CPU numerics
    assert _create(*SI2, backend=H.BACKEND_CPU)[0] == H.UNAVAILABLE


def test_qdep_without_partial_waves():
    """use_uspp_dij = 1 needs the datasets' PAW radial grid and partial
    waves (src/half_uspp.F90:142-170 integrates over them); si2.POTCAR has
    none.  With a device the create fails with INVALID_ARGUMENT naming what
    is missing; without one the missing device is reported first."""
    from half_amd import hip
    st, h, msg = _create(*SI2, uspp=1)
    assert h == 0
    if hip.device_count() == 0:
        assert st == H.UNAVAILABLE and "device" in msg
    else:
        assert st == H.INVALID_ARGUMENT and "partial waves" in msg


def test_escn_entry_points():
    lib = H.load()
    e = C.create_string_buffer(128)
    h = C.c_int64()
    assert lib.half_escn_health(b"http://127.0.0.1:9", 1, e, 128) == H.UNAVAILABLE
    assert lib.half_create_from_escn(b"", b"", None, 400.0, 2, 0, 0, 1, 0, 1, C.byref(h), e, 128) == H.INVALID_ARGUMENT
    g = H.RequestGeometry()
    assert lib.half_create_from_escn(b"http://x", SI2[1].encode(), C.byref(g), 400.0, 2, 0, 0, 1, 0, 1, C.byref(h), e,
                                     128) == H.UNAVAILABLE


def test_invalid_handles_everywhere():
    lib = H.load()
    e = C.create_string_buffer(64)
    k = (C.c_double * 3)(0, 0, 0)
    n = C.c_int64()
    for bad in (0, -1, 12345, 32, 33):
        assert lib.half_destroy(bad, e, 64) == H.INVALID_HANDLE
        assert lib.half_get_basis_size(bad, k, C.byref(n), e, 64) == H.INVALID_HANDLE
        assert e.value == b"invalid HALF context handle"
    # capacity 0 leaves the buffer untouched
    buf = C.create_string_buffer(b"keep", 16)
    assert lib.half_destroy(5, buf, 0) == H.INVALID_HANDLE
    assert buf.value == b"keep"


def test_no_device_is_reported_not_aborted():
    """Without a HIP device a valid create returns UNAVAILABLE (CPU-only
    containers); with a device it succeeds."""
    from half_amd import hip
    st, h, msg = _create(*SI2, solver=H.SOLVER_ACC)
    if hip.device_count() == 0:
        assert st == H.UNAVAILABLE and h == 0 and "device" in msg
    else:
        assert st == H.SUCCESS and h > 0
        e = C.create_string_buffer(64)
        assert H.load().half_destroy(h, e, 64) == H.SUCCESS


def test_solver_tuning_keys_host_only():
    """solver.lock (0 off, 1 soft locking, 2 soft locking plus a frozen
    prefix of converged bands, the default), solver.stall (FP32-phase
    stagnation window, 0..1000), solver.bufcols (percent of the buffer
    bands given correction columns, default 0), solver.identity (0, or
    6..15: the exponent of the unit-metric threshold, default 11) and
    solver.identity32 (0, or 3..15, the same for FP32-product iterations,
    default 5) are host settings: set, read back, restored; out-of-range
    values are rejected with INVALID_ARGUMENT."""
    saved = {k: H.get_tuning(k) for k in ("solver.lock", "solver.stall", "solver.hetrd")}
    assert saved["solver.lock"] == 2 and saved["solver.stall"] == 4
    try:
        H.set_tuning("solver.lock", 0)
        assert H.get_tuning("solver.lock") == 0
        H.set_tuning("solver.stall", 7)
        assert H.get_tuning("solver.stall") == 7
        H.set_tuning("solver.lock", 1)
        assert H.get_tuning("solver.lock") == 1
        assert H.get_tuning("solver.smalleig") == 1 and H.get_tuning("solver.bufcols") == 0
        assert H.get_tuning("solver.identity") == 11 and H.get_tuning("solver.identity32") == 5
        for key, good in (("solver.precond", 1), ("solver.refresh", 8), ("solver.fp32algebra", 0),
                          ("solver.smalleig", 2), ("solver.smalleig", 0), ("solver.smalleig", 3),
                          ("solver.bufcols", 100), ("solver.bufcols", 25), ("solver.identity", 0),
                          ("solver.identity", 6), ("solver.identity", 15), ("solver.identity32", 0),
                          ("solver.identity32", 3)):
            before = H.get_tuning(key)
            H.set_tuning(key, good)
            assert H.get_tuning(key) == good
            H.set_tuning(key, before)
        for key, bad in (("solver.lock", 3), ("solver.lock", -1), ("solver.stall", -1), ("solver.stall", 1001),
                         ("solver.precond", 2), ("solver.refresh", -1), ("solver.fp32algebra", 2),
                         ("solver.smalleig", 4), ("solver.smalleig", -1), ("solver.bufcols", 101),
                         ("solver.bufcols", -1), ("solver.identity", 5), ("solver.identity", 16),
                         ("solver.identity32", 2), ("solver.identity32", 16),
                         ("solver.nope", 0)):
            with pytest.raises(Exception, match="unknown tuning key or value out of range"):
                H.set_tuning(key, bad)
        assert H.get_tuning("solver.stall") == 7
    finally:
        for k, v in saved.items():
            H.set_tuning(k, v)


def test_contract_tuning_keys_host_only():
    """contract.<fp64|fp32>.gram.wt / tall.wt (0 = shape rule, 1, 2),
    contract.<fp64|fp32>.gram.fill (10..100 percent),
    contract.<fp64|fp32>.tall.order (0, 1, 2; default 1) and
    contract.fp64.gram.gauss / tall.gauss (0, 1) are host settings: set, read
    back, restored; out-of-range values and unknown keys are rejected."""
    keys = [f"contract.{p}.{k}" for p in ("fp64", "fp32") for k in ("gram.wt", "tall.wt", "gram.fill", "tall.order")]
    keys += ["contract.fp64.gram.gauss", "contract.fp64.tall.gauss"]
    saved = {k: H.get_tuning(k) for k in keys}
    assert all(saved[k] == 0 for k in keys if k.endswith(".wt"))
    assert all(saved[k] == 1 for k in keys if k.endswith(".gauss") or k.endswith(".order"))
    assert all(saved[k] == 90 for k in keys if k.endswith(".fill"))
    try:
        for k in keys:
            good = 75 if k.endswith(".fill") else 0 if k.endswith(".gauss") else 2 if k.endswith(".order") else 1
            H.set_tuning(k, good)
            assert H.get_tuning(k) == good
        for key, bad in (("contract.fp64.gram.wt", 3), ("contract.fp32.tall.wt", -1),
                         ("contract.fp64.gram.fill", 9), ("contract.fp32.gram.fill", 101), ("contract.fp16.gram.wt", 1),
                         ("contract.fp64.tall.gauss", 2), ("contract.fp32.tall.gauss", 1),
                         ("contract.fp64.tall.order", 3), ("contract.fp32.tall.order", -1)):
            with pytest.raises(Exception, match="unknown tuning key or value out of range"):
                H.set_tuning(key, bad)
    finally:
        for k, v in saved.items():
            H.set_tuning(k, v)
