"""Python host binding of libhalf.so (the MI355X PAW-operator backend).

Mirrors the reference's C ABI (include/half.h) and Fortran binding
(reference: include/half_api.f90:32-162) through ctypes: the same entry
points, argument meaning and status codes, raised as HalfError.  The
device-level extension (include/half_amd.h) is exposed as Operator and
Context.solve_acc / nonlocal_forces.

The library is loaded from this package's lib/ directory; there is no
fallback: importing the binding without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libhalf.so")

SUCCESS, INVALID_ARGUMENT, IO, INVALID_HANDLE, CAPACITY, UNAVAILABLE, INTERNAL = range(7)
XC_LDA, XC_PBE = 1, 2
BACKEND_AUTO, BACKEND_CPU, BACKEND_CUDA = 0, 1, 2
SOLVER_EVD, SOLVER_EVJ, SOLVER_VASP, SOLVER_ACC = 1, 2, 3, 4
CAP_CPU, CAP_CUDA, CAP_MPI, CAP_HDF5, CAP_SPGLIB, CAP_ESCN_API = 1, 2, 4, 8, 16, 32

# every symbol the two public headers declare
ABI_SYMBOLS = (
    "half_get_abi_version", "half_get_version_string", "half_get_capabilities",
    "half_escn_health", "half_escn_predict", "half_create_from_files",
    "half_create_from_files_backend", "half_create_from_escn", "half_destroy",
    "half_set_request_geometry", "half_get_request_geometry", "half_get_system_info",
    "half_get_basis_size", "half_get_basis", "half_assemble_hs", "half_apply_hs",
    "half_solve_kpoint", "half_solve_kpoint_mapped",
)
EXT_SYMBOLS = (
    "half_amd_operator_create", "half_amd_operator_apply", "half_amd_operator_destroy",
    "half_amd_solve_acc", "half_amd_set_solver", "half_amd_energy_terms",
    "half_amd_get_veff", "half_amd_nonlocal_forces", "half_amd_device_info",
    "half_amd_trace", "half_amd_trace_read", "half_amd_set_tuning", "half_amd_get_tuning",
    "half_amd_set_projector_positions", "half_amd_solve_history", "half_amd_solve_kmesh",
    "half_amd_force_set", "half_amd_write_vaspwave", "half_amd_harris_energy",
    "half_amd_operator_apply_noalloc",
)


class HalfError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"HALF status {status}: {message}")
        self.status = status
        self.message = message


class RequestGeometry(C.Structure):
    """half_request_geometry_v1 (reference: include/half.h:64-75)."""
    _fields_ = [("struct_size", C.c_uint32), ("flags", C.c_uint32), ("nions", C.c_int32),
                ("ntypes", C.c_int32), ("grid", C.c_int32 * 3), ("reserved_i32", C.c_int32),
                ("lattice", C.c_double * 9), ("species", C.POINTER(C.c_int32)),
                ("positions_fractional", C.POINTER(C.c_double)), ("reserved", C.c_uint64 * 8)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("fp64_iterations", C.c_int32),
                ("max_residual_eV", C.c_double), ("seconds", C.c_double),
                ("band_applications", C.c_int64), ("metric_fallbacks", C.c_int32),
                ("rr_order_mean", C.c_int32)]


class HarrisTerms(C.Structure):
    _fields_ = [(name, C.c_double) for name in (
        "band_energy", "fermi", "entropy", "e_hartree", "e_xc", "e_xc_potential", "ewald", "local_g0",
        "atomic_reference", "harris_energy", "free_energy", "nelect")]


_lib = None


def load(path: str | None = None):
    """Load libhalf.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"libhalf.so is not built ({p}); run __graft_entry__.build()")
    lib = C.CDLL(p)
    d, i, i64, vp, cp = C.c_double, C.c_int, C.c_int64, C.c_void_p, C.c_char_p
    pd, pi64, pi32 = C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int32)
    lib.half_get_version_string.restype = C.c_char_p
    lib.half_get_abi_version.restype = i
    lib.half_get_capabilities.restype = i
    sig = {
        "half_escn_health": [cp, i, cp, i],
        "half_escn_predict": [cp, vp, vp, i, cp, i],
        "half_create_from_files": [cp, cp, d, i, i, pi64, cp, i],
        "half_create_from_files_backend": [cp, cp, d, i, i, i, i, pi64, cp, i],
        "half_create_from_escn": [cp, cp, vp, d, i, i, i, i, i, i, pi64, cp, i],
        "half_destroy": [i64, cp, i],
        "half_set_request_geometry": [i64, vp, cp, i],
        "half_get_request_geometry": [i64, pi32, pi32, pi32, pd, i64, pi32, i64, pd, cp, i],
        "half_get_system_info": [i64, C.POINTER(i), C.POINTER(i), C.POINTER(i), pd, pd, cp, i],
        "half_get_basis_size": [i64, pd, pi64, cp, i],
        "half_get_basis": [i64, pd, i64, pi32, pd, cp, i],
        "half_assemble_hs": [i64, pd, i64, vp, vp, cp, i],
        "half_apply_hs": [i64, pd, i, i64, vp, vp, vp, cp, i],
        "half_solve_kpoint": [i64, pd, i, pd, vp, i64, pd, pd, cp, i],
        "half_solve_kpoint_mapped": [i64, pd, i, i64, pi32, pd, vp, i64, pd, pd, cp, i],
        "half_amd_operator_create": [i64, pd, i, pi64, pi64, pi32, cp, i],
        "half_amd_operator_apply": [i64, i, i, vp, i64, vp, vp, i64, i, vp, cp, i],
        "half_amd_operator_destroy": [i64, cp, i],
        "half_amd_operator_apply_noalloc": [i64, i, vp, i64, vp, vp, i64, vp, cp, i],
        "half_amd_solve_acc": [i64, pd, i, i, d, d, i, i, i, pd, vp, i64, C.POINTER(SolveStats), cp, i],
        "half_amd_set_solver": [i64, i, cp, i],
        "half_amd_energy_terms": [i64, pd, pd, pd, cp, i],
        "half_amd_get_veff": [i64, i64, pd, cp, i],
        "half_amd_nonlocal_forces": [i64, pd, i, pd, pd, d, vp, i64, pd, cp, i],
        "half_amd_device_info": [cp, i, pi32, cp, i],
        "half_amd_trace": [i, cp, i],
        "half_amd_trace_read": [i, cp, pd, pi64, pi32, cp, i],
        "half_amd_set_tuning": [cp, i, cp, i],
        "half_amd_get_tuning": [cp, C.POINTER(i), cp, i],
        "half_amd_set_projector_positions": [i64, i, pd, cp, i],
        "half_amd_solve_history": [i64, i, pd, pi32, cp, i],
        "half_amd_solve_kmesh": [i64, i, pd, i, i, d, d, i, i, i, pd, C.POINTER(SolveStats), cp, i],
        "half_amd_force_set": [i64, i, pd, pd, i, pd, pd, C.POINTER(C.c_void_p), C.POINTER(C.c_int64), pd, pd,
                               pd, pd, cp, i],
        "half_amd_write_vaspwave": [i64, cp, i, pd, i, pd, pd, d, C.POINTER(C.c_void_p), C.POINTER(C.c_int64),
                                    cp, i],
        "half_amd_harris_energy": [i64, i, pd, i, pd, d, i, pd, C.POINTER(HarrisTerms), cp, i],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = i
    _lib = lib
    return lib


def _err():
    return C.create_string_buffer(512)


def _check(status, buf):
    if status != SUCCESS:
        raise HalfError(status, buf.value.decode(errors="replace"))


def _kp(k):
    return (C.c_double * 3)(*[float(x) for x in k])


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def device_info():
    lib = load()
    name = C.create_string_buffer(256)
    cu = C.c_int32()
    e = _err()
    _check(lib.half_amd_device_info(name, 256, C.byref(cu), e, 512), e)
    return name.value.decode(), cu.value


def set_tuning(key: str, value: int):
    lib = load()
    e = _err()
    _check(lib.half_amd_set_tuning(key.encode(), int(value), e, 512), e)


def get_tuning(key: str) -> int:
    lib = load()
    v = C.c_int()
    e = _err()
    _check(lib.half_amd_get_tuning(key.encode(), C.byref(v), e, 512), e)
    return v.value


def trace(enable: bool):
    """Start (clearing earlier marks) or stop the launch-boundary timeline."""
    lib = load()
    e = _err()
    _check(lib.half_amd_trace(int(bool(enable)), e, 512), e)


def trace_read():
    """{kernel name: (total ms, launches)} of the recorded timeline."""
    lib = load()
    cap = 64
    names = C.create_string_buffer(64 * cap)
    ms = np.zeros(cap)
    cnt = np.zeros(cap, np.int64)
    nk = C.c_int32()
    e = _err()
    _check(lib.half_amd_trace_read(cap, names, _pd(ms), cnt.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(nk),
                                   e, 512), e)
    out = {}
    for i in range(min(nk.value, cap)):
        nm = names.raw[64 * i:64 * (i + 1)].split(b"\0", 1)[0].decode()
        out[nm] = (float(ms[i]), int(cnt[i]))
    return out


class Context:
    """A HALF context (half_create_from_files_backend ... half_destroy)."""

    def __init__(self, chgcar, potcar, encut, xc="pbe", backend=BACKEND_AUTO, solver=SOLVER_EVD,
                 use_uspp=False):
        self.lib = load()
        h = C.c_int64()
        e = _err()
        st = self.lib.half_create_from_files_backend(
            chgcar.encode(), potcar.encode(), float(encut), XC_PBE if xc == "pbe" else XC_LDA,
            int(bool(use_uspp)), int(backend), int(solver), C.byref(h), e, 512)
        _check(st, e)
        self.handle = h.value
        self.chgcar, self.potcar = chgcar, potcar

    def close(self):
        if self.handle:
            e = _err()
            _check(self.lib.half_destroy(self.handle, e, 512), e)
            self.handle = 0

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- queries -----------------------------------------------------------
    def system_info(self):
        n, t = C.c_int(), C.c_int()
        grid = (C.c_int * 3)()
        lat = np.zeros(9)
        vol = C.c_double()
        e = _err()
        _check(self.lib.half_get_system_info(self.handle, C.byref(n), C.byref(t), grid, _pd(lat),
                                             C.byref(vol), e, 512), e)
        return {"nions": n.value, "ntypes": t.value, "grid": tuple(grid), "lattice": lat.reshape(3, 3),
                "volume": vol.value}

    def basis_size(self, k=(0, 0, 0)):
        n = C.c_int64()
        e = _err()
        _check(self.lib.half_get_basis_size(self.handle, _kp(k), C.byref(n), e, 512), e)
        return n.value

    def basis(self, k=(0, 0, 0)):
        n = self.basis_size(k)
        g = np.zeros((n, 3), np.int32)
        kin = np.zeros(n)
        e = _err()
        _check(self.lib.half_get_basis(self.handle, _kp(k), n, g.ctypes.data_as(C.POINTER(C.c_int32)),
                                       _pd(kin), e, 512), e)
        return g, kin

    def energy_terms(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        e = _err()
        _check(self.lib.half_amd_energy_terms(self.handle, C.byref(a), C.byref(b), C.byref(c), e, 512), e)
        return a.value, b.value, c.value

    def harris_energy(self, eigenvalues, kweights=None, sigma=0.0, ismear=0):
        """Harris energy terms of solved bands (eigenvalues nk x nbands, or one
        k point's nbands) and their occupations (reference:
        app/half_cli.F90:896-920).  Returns (terms dict, occupations)."""
        eig = np.ascontiguousarray(np.atleast_2d(np.asarray(eigenvalues, float)))
        nk, nb = eig.shape
        w = np.ascontiguousarray(np.full(nk, 1.0 / nk) if kweights is None else np.asarray(kweights, float))
        occ = np.zeros((nk, nb))
        t = HarrisTerms()
        e = _err()
        _check(self.lib.half_amd_harris_energy(self.handle, nk, _pd(w), nb, _pd(eig), float(sigma), int(ismear),
                                               _pd(occ), C.byref(t), e, 512), e)
        return {name: getattr(t, name) for name, _ in HarrisTerms._fields_}, occ

    def veff(self):
        info = self.system_info()
        n = int(np.prod(info["grid"]))
        out = np.zeros(n)
        e = _err()
        _check(self.lib.half_amd_get_veff(self.handle, n, _pd(out), e, 512), e)
        return out

    # -- matrices and states ----------------------------------------------
    def assemble_hs(self, k=(0, 0, 0), ld=None):
        n = self.basis_size(k)
        ld = ld or n
        h = np.zeros((n, ld), np.complex128)  # column-major: column j at h[j]
        s = np.zeros((n, ld), np.complex128)
        e = _err()
        _check(self.lib.half_assemble_hs(self.handle, _kp(k), ld, h.ctypes.data, s.ctypes.data, e, 512), e)
        return h[:, :n].T.copy(), s[:, :n].T.copy()

    def apply_hs(self, psi, k=(0, 0, 0)):
        psi = np.asarray(psi, np.complex128)
        if psi.ndim == 1:
            psi = psi[:, None]
        n, nb = psi.shape
        colmajor = np.ascontiguousarray(psi.T)
        hp = np.zeros_like(colmajor)
        sp = np.zeros_like(colmajor)
        e = _err()
        _check(self.lib.half_apply_hs(self.handle, _kp(k), nb, n, colmajor.ctypes.data, hp.ctypes.data,
                                      sp.ctypes.data, e, 512), e)
        return hp.T.copy(), sp.T.copy()

    def set_solver(self, solver):
        e = _err()
        _check(self.lib.half_amd_set_solver(self.handle, int(solver), e, 512), e)

    def solve(self, nbands, k=(0, 0, 0), vectors=True, ld=None):
        n = self.basis_size(k)
        ld = ld or n
        w = np.zeros(nbands)
        smin, smax = C.c_double(), C.c_double()
        v = np.zeros((nbands, ld), np.complex128) if vectors else None
        e = _err()
        _check(self.lib.half_solve_kpoint(self.handle, _kp(k), nbands, _pd(w),
                                          v.ctypes.data if vectors else None, ld, C.byref(smin),
                                          C.byref(smax), e, 512), e)
        return w, (v[:, :n].T.copy() if vectors else None), smin.value, smax.value

    def solve_mapped(self, nbands, host_gvec, k=(0, 0, 0), ld=None):
        g = np.ascontiguousarray(host_gvec, np.int32)
        npw = g.shape[0]
        ld = ld or npw
        w = np.zeros(nbands)
        smin, smax = C.c_double(), C.c_double()
        v = np.zeros((nbands, ld), np.complex128)
        e = _err()
        _check(self.lib.half_solve_kpoint_mapped(self.handle, _kp(k), nbands, npw,
                                                 g.ctypes.data_as(C.POINTER(C.c_int32)), _pd(w),
                                                 v.ctypes.data, ld, C.byref(smin), C.byref(smax), e, 512), e)
        return w, v[:, :npw].T.copy(), smin.value, smax.value

    def solve_acc(self, nbands, k=(0, 0, 0), mixed=False, tol=1e-4, switch_tol=0.0, maxiter=40,
                  block=0, extra=0, vectors=False):
        n = self.basis_size(k)
        w = np.zeros(nbands)
        v = np.zeros((nbands, n), np.complex128) if vectors else None
        st = SolveStats()
        e = _err()
        _check(self.lib.half_amd_solve_acc(self.handle, _kp(k), nbands, int(mixed), float(tol),
                                           float(switch_tol), int(maxiter), int(block), int(extra), _pd(w),
                                           v.ctypes.data if vectors else None, n, C.byref(st), e, 512), e)
        stats = {f: getattr(st, f) for f, _ in SolveStats._fields_}
        return w, (v.T.copy() if vectors else None), stats

    def solve_kmesh(self, nbands, kpoints, mixed=False, tol=1e-4, switch_tol=0.0, maxiter=40, extra=0,
                    graph=True):
        """All-band solves at every k point of `kpoints` (nk x 3, fractional)
        in lockstep; returns (eigenvalues nk x nbands, stats)."""
        ks = np.ascontiguousarray(np.asarray(kpoints, np.float64).reshape(-1, 3))
        w = np.zeros((len(ks), nbands))
        st = SolveStats()
        e = _err()
        _check(self.lib.half_amd_solve_kmesh(self.handle, len(ks), _pd(ks), nbands, int(mixed), float(tol),
                                             float(switch_tol), int(maxiter), int(extra), int(graph), _pd(w),
                                             C.byref(st), e, 512), e)
        return w, {f: getattr(st, f) for f, _ in SolveStats._fields_}

    def nonlocal_forces(self, eigenvalues, occupations, vectors, k=(0, 0, 0), kweight=1.0):
        vec = np.ascontiguousarray(np.asarray(vectors, np.complex128).T)
        nb, n = vec.shape
        info = self.system_info()
        f = np.zeros((info["nions"], 3))
        eps = np.ascontiguousarray(eigenvalues, float)
        occ = np.ascontiguousarray(occupations, float)
        e = _err()
        _check(self.lib.half_amd_nonlocal_forces(self.handle, _kp(k), nb, _pd(eps), _pd(occ), float(kweight),
                                                 vec.ctypes.data, n, _pd(f), e, 512), e)
        return f

    FORCE_PARTS = ("ewald", "local", "nonlocal", "projector", "augmentation", "nlcc", "harris")

    def force_set(self, kpoints, kweights, eigenvalues, occupations, vectors, density=False):
        """Analytic force set (half_amd_force_set) for the bands of each k
        point: vectors[j] is npw_j x nbands.  Returns a dict with "total"
        and the FORCE_PARTS (natoms x 3 each), "smooth_charge",
        "aug_charge" and, with density, "density" (CHGCAR order)."""
        ks = np.ascontiguousarray(np.asarray(kpoints, float).reshape(-1, 3))
        nk = len(ks)
        kw = np.ascontiguousarray(np.asarray(kweights, float).reshape(nk))
        eps = np.ascontiguousarray(np.asarray(eigenvalues, float).reshape(nk, -1))
        nb = eps.shape[1]
        occ = np.ascontiguousarray(np.asarray(occupations, float).reshape(nk, nb))
        cols = [np.asfortranarray(np.asarray(v, np.complex128)[:, :nb]) for v in vectors]
        ptrs = (C.c_void_p * nk)(*[c.ctypes.data for c in cols])
        lds = (C.c_int64 * nk)(*[c.shape[0] for c in cols])
        nat = self.system_info()["nions"]
        total = np.zeros((nat, 3))
        parts = np.zeros((7, nat, 3))
        charges = np.zeros(2)
        grid = self.system_info()["grid"]
        dens = np.zeros(int(np.prod(grid))) if density else None
        e = _err()
        _check(self.lib.half_amd_force_set(self.handle, nk, _pd(ks), _pd(kw), nb, _pd(eps), _pd(occ), ptrs, lds,
                                           _pd(total), _pd(parts), _pd(dens) if density else None, _pd(charges),
                                           e, 512), e)
        out = {"total": total, "smooth_charge": charges[0], "aug_charge": charges[1]}
        out.update({k: parts[i] for i, k in enumerate(self.FORCE_PARTS)})
        if density:
            out["density"] = dens
        return out

    def write_vaspwave(self, path, kpoints, eigenvalues, occupations, vectors, fermi):
        """vaspwave.h5 export (half_amd_write_vaspwave): vectors[j] is
        npw_j x nbands for k point j."""
        ks = np.ascontiguousarray(np.asarray(kpoints, float).reshape(-1, 3))
        nk = len(ks)
        eps = np.ascontiguousarray(np.asarray(eigenvalues, float).reshape(nk, -1))
        nb = eps.shape[1]
        occ = np.ascontiguousarray(np.asarray(occupations, float).reshape(nk, nb))
        cols = [np.asfortranarray(np.asarray(v, np.complex128)[:, :nb]) for v in vectors]
        ptrs = (C.c_void_p * nk)(*[c.ctypes.data for c in cols])
        lds = (C.c_int64 * nk)(*[c.shape[0] for c in cols])
        e = _err()
        _check(self.lib.half_amd_write_vaspwave(self.handle, str(path).encode(), nk, _pd(ks), nb, _pd(eps),
                                                _pd(occ), float(fermi), ptrs, lds, e, 512), e)

    def solve_history(self):
        """Max residual (eV) per iteration of the last solve_acc."""
        cnt = C.c_int32()
        e = _err()
        self.lib.half_amd_solve_history(self.handle, 0, None, C.byref(cnt), e, 512)  # sizes only
        out = np.zeros(max(cnt.value, 1))
        _check(self.lib.half_amd_solve_history(self.handle, out.size, _pd(out), C.byref(cnt), e, 512), e)
        return out[:cnt.value]

    def set_projector_positions(self, positions_fractional):
        """Frozen-potential displacement of the atoms seen by the projectors."""
        pos = np.ascontiguousarray(positions_fractional, float)
        e = _err()
        _check(self.lib.half_amd_set_projector_positions(self.handle, pos.shape[0], _pd(pos.ravel()), e, 512), e)

    def operator(self, k=(0, 0, 0), precision=64):
        return Operator(self, k, precision)

    # -- request geometry --------------------------------------------------
    def set_request_geometry(self, lattice, species, positions, grid, ntypes=None):
        """The host's cell (rows = lattice vectors, A), one-based type of
        every ion, fractional positions and FFT grid; the context then builds
        everything on it (half_set_request_geometry)."""
        sp = np.ascontiguousarray(species, np.int32)
        pos = np.ascontiguousarray(positions, float).ravel()
        g = RequestGeometry()
        g.struct_size = C.sizeof(RequestGeometry)
        g.flags = 0
        g.nions = sp.size
        g.ntypes = int(ntypes) if ntypes is not None else (int(sp.max()) if sp.size else 0)
        g.grid[:] = list(grid)
        g.lattice[:] = list(np.asarray(lattice, float).ravel())
        g.species = sp.ctypes.data_as(C.POINTER(C.c_int32))
        g.positions_fractional = pos.ctypes.data_as(C.POINTER(C.c_double))
        e = _err()
        _check(self.lib.half_set_request_geometry(self.handle, C.byref(g), e, 512), e)

    def get_request_geometry(self):
        n, t = C.c_int32(), C.c_int32()
        grid = (C.c_int32 * 3)()
        lat = np.zeros(9)
        e = _err()
        _check(self.lib.half_get_request_geometry(self.handle, C.byref(n), C.byref(t), grid, _pd(lat), 0, None,
                                                  0, None, e, 512), e)
        sp = np.zeros(n.value, np.int32)
        pos = np.zeros(3 * n.value)
        _check(self.lib.half_get_request_geometry(self.handle, C.byref(n), C.byref(t), grid, _pd(lat), n.value,
                                                  sp.ctypes.data_as(C.POINTER(C.c_int32)), 3 * n.value, _pd(pos),
                                                  e, 512), e)
        return {"nions": n.value, "ntypes": t.value, "grid": tuple(grid), "lattice": lat.reshape(3, 3),
                "species": sp, "positions": pos.reshape(-1, 3)}


class Operator:
    """Device-resident H.psi / S.psi at one k point (half_amd_operator_*)."""

    def __init__(self, ctx: Context, k, precision):
        self.ctx, self.lib = ctx, ctx.lib
        op, n, np_ = C.c_int64(), C.c_int64(), C.c_int32()
        e = _err()
        _check(self.lib.half_amd_operator_create(ctx.handle, _kp(k), int(precision), C.byref(op), C.byref(n),
                                                 C.byref(np_), e, 512), e)
        self.handle, self.npw, self.nproj, self.precision = op.value, n.value, np_.value, precision

    def apply_device(self, nstates, d_psi, d_hpsi, d_spsi, ld=None, stage=0, box_batch=0, stream=None):
        """Apply on device pointers (integers), e.g. torch tensors' data_ptr()."""
        ld = ld or self.npw
        e = _err()
        _check(self.lib.half_amd_operator_apply(self.handle, int(stage), int(nstates), C.c_void_p(d_psi), ld,
                                                C.c_void_p(d_hpsi), C.c_void_p(d_spsi), ld, int(box_batch),
                                                C.c_void_p(stream) if stream else None, e, 512), e)

    def apply_device_noalloc(self, nstates, d_psi, d_hpsi, d_spsi, ld=None, stream=None):
        """The full operator on device pointers through the entry point that
        allocates no device memory (half_amd_operator_apply_noalloc)."""
        ld = ld or self.npw
        e = _err()
        _check(self.lib.half_amd_operator_apply_noalloc(self.handle, int(nstates), C.c_void_p(d_psi), ld,
                                                        C.c_void_p(d_hpsi), C.c_void_p(d_spsi), ld,
                                                        C.c_void_p(stream) if stream else None, e, 512), e)

    def close(self):
        if self.handle:
            e = _err()
            self.lib.half_amd_operator_destroy(self.handle, e, 512)
            self.handle = 0
