"""ctypes binding of ``libbse_mi355x.so`` (C ABI in include/bse_mi355x.h).

There is no CPU fallback: if the library is missing or no GPU is visible,
construction raises.  Arrays cross the boundary as host numpy buffers (or raw
device pointers where an entry point takes ``*_on_device``).
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(os.environ.get(
    "BSE_MI355X_LIB", pathlib.Path(__file__).resolve().parent / "libbse_mi355x.so"))

PREC_FP64, PREC_FP32, PREC_BF16X2 = 0, 1, 2
ORDER_REFERENCE, ORDER_KMAJOR = 0, 1
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x2": PREC_BF16X2}

# Symbols every build must export (mirrors include/bse_mi355x.h).
ABI_SYMBOLS = (
    "bse_create", "bse_destroy", "bse_last_error", "bse_abi_version", "bse_set_problem",
    "bse_problem_sizes", "bse_tb_states", "bse_build", "bse_get_matrix", "bse_get_dipoles",
    "bse_eigh", "bse_get_eigenvalues", "bse_get_eigenvectors", "bse_oscillator",
    "bse_spectrum", "bse_eigh_dense", "bse_stage_times", "bse_profile", "bse_apply", "bse_grad_energies",
    "bse_batch_prepare", "bse_batch_run", "bse_batch_results", "bse_batch_info",
    "bse_grad_spectrum", "bse_refine_info", "bse_set_problem_klist", "bse_tb_states_at",
    "bse_eigh_dense_noalloc", "bse_noalloc_info", "bse_build_full", "bse_eigh_full", "bse_get_coupling",
    "bse_wannier_integrals", "bse_probe_rate",
)


class BSEError(RuntimeError):
    pass


class TBModelC(C.Structure):
    _fields_ = [("norb", C.c_int32), ("nr", C.c_int32), ("r_cart", C.POINTER(C.c_double)),
                ("h_r", C.POINTER(C.c_double)), ("tau", C.POINTER(C.c_double))]


class SystemC(C.Structure):
    _fields_ = [("bravais", C.c_double * 9), ("n1", C.c_int32), ("n2", C.c_int32),
                ("shift", C.c_double * 2), ("n_occ", C.c_int32), ("nv", C.c_int32),
                ("nc", C.c_int32), ("scissor", C.c_double), ("q", C.c_double * 3),
                ("r0", C.c_double), ("eps_env", C.c_double), ("g_cut", C.c_double),
                ("f_s", C.c_double), ("exchange_scale", C.c_double),
                ("direct_scale", C.c_double), ("dipole_dir", C.c_double * 3),
                ("dimension", C.c_int32), ("n3", C.c_int32), ("shift3", C.c_double),
                ("radius", C.c_double), ("screening", C.c_int32), ("rpa_nv", C.c_int32),
                ("rpa_nc", C.c_int32), ("dipole_model", C.c_int32), ("dipole_nr", C.c_int32),
                ("dipole_r", C.POINTER(C.c_double)), ("dipole_d", C.POINTER(C.c_double))]


class WannierGridC(C.Structure):
    _fields_ = [("nw", C.c_int32), ("nspin", C.c_int32), ("n_sc", C.c_int32 * 3), ("p", C.c_int32 * 3),
                ("origin", C.c_double * 3), ("axis", C.c_double * 9), ("w", C.POINTER(C.c_double))]


class BatchOptionsC(C.Structure):
    _fields_ = [("precision", C.c_int32), ("n_lowest", C.c_int32), ("refine", C.c_int32),
                ("n_omega", C.c_int32), ("omega", C.POINTER(C.c_double)), ("eta", C.c_double)]


def _system_to_c(s):
    """(TBModelC, SystemC, arrays to keep alive) for a host BSESystem."""
    m = s.model
    r_cart = np.ascontiguousarray(m.r_cart, dtype=np.float64)
    h_r = np.ascontiguousarray(m.h_r, dtype=np.complex128)
    tau = np.ascontiguousarray(m.tau, dtype=np.float64)
    mc = TBModelC(m.norb, m.nr, _dptr(r_cart), h_r.ctypes.data_as(C.POINTER(C.c_double)),
                  _dptr(tau))
    sc = SystemC()
    sc.bravais[:] = list(np.asarray(m.lattice.bravais, dtype=np.float64).T.ravel())
    sc.n1, sc.n2 = s.grid.n1, s.grid.n2
    sc.shift[:] = [float(x) for x in list(s.grid.shift)[:2]]
    sc.n_occ, sc.nv, sc.nc = s.n_occ, s.nv, s.nc
    sc.scissor = s.scissor
    sc.q[:] = list(s.q_opt)
    sc.r0, sc.eps_env, sc.g_cut = s.r0, s.eps_env, s.g_cut
    sc.f_s = s.spin_factor
    sc.exchange_scale, sc.direct_scale = s.exchange_scale, s.direct_scale
    sc.dipole_dir[:] = [1.0, 0.0, 0.0]
    sc.dimension = int(getattr(s, "dimension", 2))
    sc.n3 = int(getattr(s.grid, "n3", 1))
    sc.shift3 = float(getattr(s.grid, "shift3", 0.0))
    sc.radius = float(getattr(s, "radius", 0.0))
    scr = getattr(s, "screening", "model")
    if scr not in ("model", "rpa"):
        raise ValueError(f"unknown screening {scr!r}")
    sc.screening = 1 if scr == "rpa" else 0
    sc.rpa_nv = int(getattr(s, "rpa_nv", 0))
    sc.rpa_nc = int(getattr(s, "rpa_nc", 0))
    keep = [r_cart, h_r, tau]
    dr = getattr(s, "dipole_r", None)
    if dr is not None:
        dr = np.ascontiguousarray(dr, dtype=np.float64).reshape(-1, 3)
        dd = np.ascontiguousarray(s.dipole_d, dtype=np.float64)
        if dd.shape != (dr.shape[0], m.norb, m.norb, 3):
            raise ValueError(f"dipole_d must be (nR, norb, norb, 3), got {dd.shape}")
        sc.dipole_model, sc.dipole_nr = 1, dr.shape[0]
        sc.dipole_r, sc.dipole_d = _dptr(dr), _dptr(dd)
        keep += [dr, dd]
    return mc, sc, keep


def _model_to_c(m):
    r_cart = np.ascontiguousarray(m.r_cart, dtype=np.float64)
    h_r = np.ascontiguousarray(m.h_r, dtype=np.complex128)
    tau = np.ascontiguousarray(m.tau, dtype=np.float64)
    mc = TBModelC(m.norb, m.nr, _dptr(r_cart), h_r.ctypes.data_as(C.POINTER(C.c_double)),
                  _dptr(tau))
    return mc, [r_cart, h_r, tau]


_lib = None


def load_library(path: pathlib.Path | str | None = None):
    """Load the shared library once; raises if it is absent (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise BSEError(f"{p} is missing: build it with build.py (hipcc, gfx950)")
    lib = C.CDLL(str(p))
    vp, i32, i64, dp = C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_double)
    sig = {
        "bse_create": (i32, [C.POINTER(vp), i32]),
        "bse_destroy": (i32, [vp]),
        "bse_last_error": (C.c_char_p, [vp]),
        "bse_abi_version": (i32, []),
        "bse_set_problem": (i32, [vp, C.POINTER(TBModelC), C.POINTER(SystemC)]),
        "bse_problem_sizes": (i32, [vp, C.POINTER(C.c_int64)]),
        "bse_tb_states": (i32, [vp, dp, dp]),
        "bse_build": (i32, [vp, i32]),
        "bse_get_matrix": (i32, [vp, vp, i64, i32, i32]),
        "bse_get_dipoles": (i32, [vp, dp]),
        "bse_eigh": (i32, [vp, i32, i32]),
        "bse_get_eigenvalues": (i32, [vp, dp]),
        "bse_get_eigenvectors": (i32, [vp, dp, i64, i32, i32]),
        "bse_oscillator": (i32, [vp, dp]),
        "bse_spectrum": (i32, [vp, dp, i32, C.c_double, dp, dp]),
        "bse_eigh_dense": (i32, [vp, i32, dp, i64, i32, i32, i32, dp, dp, i64]),
        "bse_stage_times": (i32, [vp, dp, i32]),
        "bse_profile": (i32, [vp, i32, dp]),
        "bse_apply": (i32, [vp, i32, dp, i64, dp, i64]),
        "bse_grad_energies": (i32, [vp, i32, C.POINTER(C.c_int), dp]),
        "bse_batch_prepare": (i32, [vp, i32, C.POINTER(TBModelC), C.POINTER(SystemC),
                                    C.POINTER(BatchOptionsC)]),
        "bse_batch_run": (i32, [vp, dp, dp]),
        "bse_batch_results": (i32, [vp, dp, dp, dp]),
        "bse_batch_info": (i32, [vp, C.POINTER(C.c_int64)]),
        "bse_grad_spectrum": (i32, [vp, dp, i32, C.c_double, dp, C.POINTER(C.c_int), dp, dp]),
        "bse_refine_info": (i32, [vp, dp, i32, C.POINTER(C.c_int)]),
        "bse_set_problem_klist": (i32, [vp, C.POINTER(TBModelC), C.POINTER(SystemC), dp, i32]),
        "bse_tb_states_at": (i32, [vp, C.POINTER(TBModelC), dp, i32, i32, i32, i32, C.c_double,
                                   dp, dp]),
        "bse_build_full": (i32, [vp]),
        "bse_eigh_full": (i32, [vp, i32]),
        "bse_get_coupling": (i32, [vp, dp, i64]),
        "bse_wannier_integrals": (i32, [vp, C.POINTER(WannierGridC), i32, C.POINTER(C.c_int32), i32, dp,
                                        dp, dp, dp]),
        "bse_probe_rate": (i32, [i32, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


PROBE_FORMS = {"mfma_f32_16x16x4": 0, "mfma_f64_16x16x4": 1, "fma_f64": 2}


def probe_rate(form: str) -> float:
    """TFLOP/s one instruction form reaches on the current GPU with no memory
    traffic (bse_probe_rate): the measured ceiling beside the published peak."""
    lib = load_library()
    out = C.c_double(0.0)
    rc = lib.bse_probe_rate(PROBE_FORMS[form], C.byref(out))
    if rc != 0:
        raise BSEError(f"bse_probe_rate failed ({rc})")
    return float(out.value)


STAGES = ("tb", "tables", "assembly", "tridiagonalisation", "tridiagonal_values",
          "tridiagonal_vectors", "back_transform", "refinement", "spectrum",
          "band_reduction", "bulge_chase")


class Context:
    """One device context (stream + device buffers) for one host thread."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.bse_create(C.byref(h), int(device))
        if rc != 0:
            raise BSEError(f"bse_create failed ({rc}): no usable GPU")
        self.h = h
        self.system = None
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            self.lib.bse_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.bse_last_error(self.h)
            raise BSEError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # ------------------------------------------------------------------
    def set_system(self, s):
        """Upload a host-side system description (BSESystem)."""
        mc, sc, keep = _system_to_c(s)
        self._keep = keep
        self._check(self.lib.bse_set_problem(self.h, C.byref(mc), C.byref(sc)), "bse_set_problem")
        self.system = s
        return self

    def set_system_klist(self, s, k_cryst):
        """Upload a system whose k-points are the caller's list (crystal
        coordinates, 3 x nk, any order of a complete regular grid); the BSE
        index k of every reference-order output is the list position."""
        mc, sc, keep = _system_to_c(s)
        kc = np.asfortranarray(np.asarray(k_cryst, dtype=np.float64).reshape(3, -1))
        self._keep = keep + [kc]
        self._check(self.lib.bse_set_problem_klist(self.h, C.byref(mc), C.byref(sc), _dptr(kc),
                                                   kc.shape[1]), "bse_set_problem_klist")
        self.system = s
        return self

    def tb_states_at(self, kpts, n_occ=0, nv=0, nc=0, scissor=0.0, model=None):
        """States at Cartesian k (3 x nk): every band ascending when nv = nc = 0
        (pull_ks_states), else the subset window (pull_ks_states_subset).
        Returns (eps [nk, ns], coef [nk, ns, norb])."""
        kp = np.asfortranarray(np.asarray(kpts, dtype=np.float64).reshape(3, -1))
        nk = kp.shape[1]
        mc_p, keep = None, []
        if model is not None:
            mc, keep = _model_to_c(model)
            mc_p = C.byref(mc)
            norb = model.norb
        else:
            norb = self.system.norb
        ns = norb if (nv == 0 and nc == 0) else nv + nc
        eps = np.zeros((nk, ns))
        coef = np.zeros((nk, ns, norb), dtype=np.complex128)
        self._check(self.lib.bse_tb_states_at(self.h, mc_p, _dptr(kp), nk, int(n_occ), int(nv), int(nc),
                                              float(scissor), _dptr(eps),
                                              coef.ctypes.data_as(C.POINTER(C.c_double))),
                    "bse_tb_states_at")
        del keep
        return eps, coef

    def bands(self, segments, model=None, lattice=None, crystal=True, n_occ=0, nv=0, nc=0, scissor=0.0):
        """Band energies along a path (Hamiltonian_TB::pull_bands, 823-875):
        segments from lattice.read_k_path, points by band_path_points, states
        by bse_tb_states_at (every band ascending when nv = nc = 0, else the
        window: valence top-down, conduction bottom-up + scissor).  Returns
        (k 3 x nk Cartesian, energies nk x bands); lattice.write_bands writes
        the reference's bands file."""
        from .lattice import band_path_points
        m = model if model is not None else self.system.model
        lat = lattice if lattice is not None else m.lattice
        k = band_path_points(segments, lat, crystal)
        eps, _ = self.tb_states_at(k, n_occ=n_occ, nv=nv, nc=nc, scissor=scissor, model=model)
        return k, eps

    def sizes(self):
        out = (C.c_int64 * 4)()
        self._check(self.lib.bse_problem_sizes(self.h, out), "bse_problem_sizes")
        return dict(nk=out[0], n=out[1], norb=out[2], ng=out[3])

    def tb_states(self):
        s = self.system
        ns = s.nv + s.nc
        eps = np.zeros((s.nk, ns))
        coef = np.zeros((s.nk, ns, s.norb), dtype=np.complex128)
        self._check(self.lib.bse_tb_states(self.h, _dptr(eps), coef.ctypes.data_as(C.POINTER(C.c_double))),
                    "bse_tb_states")
        return eps, coef

    def build(self, precision="fp64"):
        p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self._check(self.lib.bse_build(self.h, p), "bse_build")
        self.precision = p
        return self

    def matrix(self, order=ORDER_REFERENCE):
        n = self.sizes()["n"]
        dt = np.complex128 if self.precision == PREC_FP64 else np.complex64
        out = np.zeros((n, n), dtype=dt, order="F")
        self._check(self.lib.bse_get_matrix(self.h, out.ctypes.data, n, 0, order), "bse_get_matrix")
        return out

    def dipoles(self):
        n = self.sizes()["n"]
        out = np.zeros(n, dtype=np.complex128)
        self._check(self.lib.bse_get_dipoles(self.h, out.ctypes.data_as(C.POINTER(C.c_double))),
                    "bse_get_dipoles")
        return out

    def apply(self, x):
        """Matrix-free FP64 product H x (reference order); x: (N,) or (N, k)."""
        x = np.asarray(x, dtype=np.complex128)
        vec = x.ndim == 1
        xf = np.asfortranarray(x.reshape(x.shape[0], -1))
        n, k = xf.shape
        y = np.zeros((n, k), dtype=np.complex128, order="F")
        self._check(self.lib.bse_apply(self.h, k, xf.ctypes.data_as(C.POINTER(C.c_double)), n,
                                       y.ctypes.data_as(C.POINTER(C.c_double)), n), "bse_apply")
        return y[:, 0] if vec else y

    def grad_energies(self, n_exc, which=None):
        """(n_exc, 5) dE_m/dtheta for theta = (hopping scale, scissor, r0, eps_env,
        exchange field); requires eigh(n_vectors >= n_exc)."""
        out = np.zeros((5, n_exc))
        w = None
        if which is not None:
            w = (C.c_int * 5)(*[int(bool(x)) for x in which])
        self._check(self.lib.bse_grad_energies(self.h, int(n_exc), w, _dptr(out)),
                    "bse_grad_energies")
        return out.T.copy()

    def grad_spectrum(self, omegas, eta, target, which=None):
        """(L, dL/dtheta[5]) of the spectrum misfit (needs eigh(-1))."""
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        tg = np.ascontiguousarray(target, dtype=np.float64)
        if tg.shape != om.shape:
            raise ValueError("target must have one value per frequency")
        w = None
        if which is not None:
            w = (C.c_int * 5)(*[1 if x else 0 for x in which])
        L = C.c_double(0.0)
        dL = np.zeros(5)
        self._check(self.lib.bse_grad_spectrum(self.h, _dptr(om), om.size, float(eta), _dptr(tg), w,
                                               C.byref(L), _dptr(dL)), "bse_grad_spectrum")
        return L.value, dL

    def wannier_integrals(self, grid, r_cells, kvec=None):
        """Real-space Wannier integrals on the GPU (bse_wannier_integrals):
        A (nspin, nw, nw, nR, nK) complex, D (nspin, nw, nw, nR, 3), S
        (nspin, nw, nw, nR) for the lattice translations r_cells (nR x 3,
        unit cells) and Cartesian vectors kvec (nK x 3, 1/A)."""
        w = np.ascontiguousarray(grid.values, dtype=np.float64)
        gc = WannierGridC()
        gc.nw, gc.nspin = grid.nw, grid.nspin
        gc.n_sc[:] = [int(x) for x in grid.n_sc]
        gc.p[:] = [int(x) for x in grid.p]
        gc.origin[:] = [float(x) for x in grid.origin]
        gc.axis[:] = list(np.asarray(grid.axes, dtype=np.float64).T.ravel())
        gc.w = _dptr(w)
        rc = np.ascontiguousarray(np.asarray(r_cells, dtype=np.int32).reshape(-1, 3))
        kv = np.ascontiguousarray(np.zeros((0, 3)) if kvec is None else np.asarray(kvec, dtype=np.float64)
                                  ).reshape(-1, 3)
        nr, nk, ns, nw = rc.shape[0], kv.shape[0], grid.nspin, grid.nw
        A = np.zeros(max(ns * nw * nw * nr * nk, 1), dtype=np.complex128)
        D = np.zeros(ns * nw * nw * nr * 3)
        S = np.zeros(ns * nw * nw * nr)
        self._check(self.lib.bse_wannier_integrals(
            self.h, C.byref(gc), nr, rc.ctypes.data_as(C.POINTER(C.c_int32)), nk,
            _dptr(kv) if nk else None, A.ctypes.data_as(C.POINTER(C.c_double)) if nk else None,
            _dptr(D), _dptr(S)), "bse_wannier_integrals")
        A = A[:ns * nw * nw * nr * nk].reshape(ns, nw, nw, nr, nk)
        return A, D.reshape(ns, nw, nw, nr, 3), S.reshape(ns, nw, nw, nr)

    def eigh(self, n_vectors=-1, refine=False):
        self._check(self.lib.bse_eigh(self.h, int(n_vectors), int(bool(refine))), "bse_eigh")
        self.n_vectors = n_vectors
        self._full = False
        return self

    def build_full(self):
        """Assemble A and the coupling block B of the full BSE (FP64)."""
        self._check(self.lib.bse_build_full(self.h), "bse_build_full")
        self._full = False
        return self

    def coupling(self):
        """The coupling block B (N x N complex symmetric), reference order."""
        n = self.sizes()["n"]
        out = np.zeros((n, n), dtype=np.complex128, order="F")
        self._check(self.lib.bse_get_coupling(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), n),
                    "bse_get_coupling")
        return out

    def eigh_full(self, n_vectors=-1):
        """Structure-preserving eig of [[A, B], [-conj B, -conj A]]: afterwards
        eigenvalues() are the N positive energies, oscillator() f = rho~^H x,
        eigenvectors(k) the 2N-row (x; y), spectrum() the full sum."""
        self._check(self.lib.bse_eigh_full(self.h, int(n_vectors)), "bse_eigh_full")
        self.n_vectors = n_vectors
        self._full = True
        return self

    def refine_info(self, n):
        """(residual norms in eV of the first n refined vectors, rounds, status)."""
        out = np.zeros(max(int(n), 1))
        rounds = C.c_int(0)
        rc = self.lib.bse_refine_info(self.h, _dptr(out), int(n), C.byref(rounds))
        if rc == -1:
            self._check(rc, "bse_refine_info")
        return out[:n], rounds.value, rc

    def eigenvalues(self):
        n = self.sizes()["n"] if self.system is not None else self._n_dense
        out = np.zeros(n)
        self._check(self.lib.bse_get_eigenvalues(self.h, _dptr(out)), "bse_get_eigenvalues")
        return out

    def eigenvectors(self, nvec):
        n = self.sizes()["n"] * (2 if getattr(self, "_full", False) else 1)
        out = np.zeros((n, nvec), dtype=np.complex128, order="F")
        self._check(self.lib.bse_get_eigenvectors(self.h, out.ctypes.data_as(C.POINTER(C.c_double)),
                                                  n, int(nvec), 0), "bse_get_eigenvectors")
        return out

    def oscillator(self):
        n = self.sizes()["n"]
        out = np.zeros(n, dtype=np.complex128)
        self._check(self.lib.bse_oscillator(self.h, out.ctypes.data_as(C.POINTER(C.c_double))),
                    "bse_oscillator")
        return out

    def spectrum(self, omegas, eta):
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        re = np.zeros_like(om)
        im = np.zeros_like(om)
        self._check(self.lib.bse_spectrum(self.h, _dptr(om), om.size, float(eta), _dptr(re), _dptr(im)),
                    "bse_spectrum")
        return re + 1j * im

    def eigh_dense(self, a, precision="fp64", n_vectors=-1):
        a = np.asfortranarray(a, dtype=np.complex128)
        n = a.shape[0]
        p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        ev = np.zeros(n)
        nv = n if n_vectors < 0 else int(n_vectors)
        z = np.zeros((n, max(nv, 1)), dtype=np.complex128, order="F")
        self._check(self.lib.bse_eigh_dense(self.h, n, a.ctypes.data_as(C.POINTER(C.c_double)), n, 0, p,
                                            int(n_vectors), _dptr(ev),
                                            z.ctypes.data_as(C.POINTER(C.c_double)), n),
                    "bse_eigh_dense")
        self.system = None
        self._n_dense = n
        return ev, z[:, :nv]

    def profile(self, enable=None):
        """Profile of the last eig's dominant kernel: the one-stage reduction's
        Hermitian matrix-vector product (algorithmic bytes), or the two-stage
        band reduction's rank-2k update and block product (algorithmic flops;
        FP32 matrix cores for FP32 storage, FP64 arithmetic otherwise)."""
        out = np.zeros(9)
        en = 1 if enable else 0
        if enable is None:
            en = 1 if getattr(self, "_prof", False) else 0
        self._prof = bool(en)
        self._check(self.lib.bse_profile(self.h, en, _dptr(out)), "bse_profile")
        if int(out[4]) == 1:
            f32 = int(out[1]) == 1
            tag = "FP32 MFMA" if f32 else "FP64 products"
            look = int(out[8]) >= 3
            packed = int(out[8]) == 4
            kname = ((f"k_her2k_part<float2, {4 if packed else 3}, 2>" if look else "k_her2k_mfma32")
                     if f32 else "her2k")
            part = ", tiles beyond the next panel's tile column" if look else ""
            if packed:
                part += ", operands packed in FP32 planes"
            return dict(kind="flops", kernel=f"band rank-2k update ({kname}, {tag}{part})",
                        ms=out[0], flops=out[3], launches=int(out[2]), peak=157.3 if f32 else 78.6,
                        second=dict(kernel=f"{'k_band_hemm32' if f32 else 'k_band_hemm'} (X = A22 V, {tag})",
                                    ms=out[5], flops=out[6], launches=int(out[7])))
        return dict(kind="bytes", kernel="kb_hemv (every variant the tuner picks)", ms=out[0],
                    bytes=out[1], launches=int(out[2]), hemv_ms=out[0], hemv_bytes=out[1],
                    hemv_launches=int(out[2]))

    def stage_times(self):
        out = np.zeros(len(STAGES))
        self._check(self.lib.bse_stage_times(self.h, _dptr(out), len(STAGES)), "bse_stage_times")
        return dict(zip(STAGES, out.tolist()))

    # ---- batched sweep ---------------------------------------------------
    def batch_prepare(self, systems, precision="fp32", n_lowest=16, refine=True, omegas=None,
                      eta=0.05):
        """Upload many systems and record the sweep graph (bse_batch_prepare)."""
        n = len(systems)
        models = (TBModelC * n)()
        syss = (SystemC * n)()
        keep = []
        for i, s in enumerate(systems):
            mc, sc, k = _system_to_c(s)
            models[i] = mc
            syss[i] = sc
            keep.extend(k)
        om = np.ascontiguousarray(omegas if omegas is not None else np.zeros(0), dtype=np.float64)
        opt = BatchOptionsC(PRECISIONS[precision] if isinstance(precision, str) else int(precision),
                            int(n_lowest), 1 if refine else 0, int(om.size),
                            _dptr(om) if om.size else None, float(eta))
        self._check(self.lib.bse_batch_prepare(self.h, n, models, syss, C.byref(opt)),
                    "bse_batch_prepare")
        self._batch = dict(n=n, n_lowest=int(n_lowest), n_omega=int(om.size))
        return self

    def batch_run(self):
        """One graph launch: (total ms, per-system completion ms in input order)."""
        b = self._batch
        tot = C.c_double(0.0)
        per = np.zeros(b["n"])
        self._check(self.lib.bse_batch_run(self.h, C.byref(tot), _dptr(per)), "bse_batch_run")
        return tot.value, per

    def batch_results(self):
        """(evals [n, n_lowest], eps [n, n_omega] complex or None)."""
        b = self._batch
        ev = np.zeros((b["n"], b["n_lowest"]))
        if b["n_omega"] > 0:
            re = np.zeros((b["n"], b["n_omega"]))
            im = np.zeros((b["n"], b["n_omega"]))
            self._check(self.lib.bse_batch_results(self.h, _dptr(ev), _dptr(re), _dptr(im)),
                        "bse_batch_results")
            return ev, re + 1j * im
        self._check(self.lib.bse_batch_results(self.h, _dptr(ev), None, None), "bse_batch_results")
        return ev, None

    def batch_info(self):
        out = (C.c_int64 * 6)()
        self._check(self.lib.bse_batch_info(self.h, out), "bse_batch_info")
        return dict(graph_nodes=out[0], arena_bytes=out[1], max_n=out[2], systems=out[3],
                    hemv_ch_top=out[4], her2k_valu=out[5])
