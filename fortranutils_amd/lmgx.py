"""Python binding of liblmgx (ctypes over the C ABI in ``include/lmgx.h``).

Mirrors the reference's pressure-solve interface, ``lmg_solve(ap,aw,ae,as,an,al,ar,f,b,residu,
ia,ra)`` (reference ``code/lmg_solve.F:5``): coefficient arrays, solution ``f``, source ``b`` and
residual ``residu`` in the reference's multi-level layout.  The reference takes its topology and
multigrid knobs from common blocks (``/cblgrid1_1/``, ``/cbmgdata1/``, ``/cbmgdata2/``); here they
are the :class:`Topology` given to :class:`Solver` and the keyword arguments of
:meth:`Solver.lmg_solve`.  The reference stops the program on a bad configuration
(``code/solvers.F:35-36``, ``code/smooth.F:77-79``); this binding raises :class:`LmgxError`.

The HIP library is required: nothing here falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblmgx.so")

LMGX_SMOOTH_EXACT, LMGX_SMOOTH_LINE, LMGX_SMOOTH_SPORLX2 = 0, 1, 2
LMGX_PREC_FP64, LMGX_PREC_MIXED = 0, 1
MAXL = 8


class LmgxError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [("cycle", C.c_int), ("numpre", C.c_int), ("numpost", C.c_int),
                ("numcits", C.c_int), ("max_cycles", C.c_int), ("omega", C.c_double),
                ("omega_c", C.c_double), ("tolf", C.c_double), ("rtol", C.c_double),
                ("smoother", C.c_int), ("precision", C.c_int), ("check_every", C.c_int),
                ("sweeps_per_iter", C.c_int), ("coarse_op", C.c_int), ("coarse_solve", C.c_int),
                ("fine_post", C.c_int)]


class _Topo(C.Structure):
    _fields_ = [("nblk", C.c_int), ("nlev", C.c_int), ("nijkp1", C.POINTER(C.c_int)),
                ("bc", C.POINTER(C.c_int)), ("bl2bl", C.POINTER(C.c_int)),
                ("bl2side", C.POINTER(C.c_int))]


_lib = None


def lib():
    """Load liblmgx.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LmgxError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        # PyTorch ships its own HIP runtime with the same SONAME (libamdhip64.so.7).  Loading it
        # first makes liblmgx bind to that one runtime, so torch tensors and liblmgx share a
        # single HIP context; without torch, liblmgx uses /opt/rocm's runtime.
        try:
            import torch  # noqa: F401
            import torch.cuda  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.lmgx_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise LmgxError(f"lmgx error {rc}: {lib().lmgx_last_error().decode()}")


def _hp(a, n: int, what: str = "array") -> C.POINTER(C.c_double):
    """Pointer to a host float64 numpy array of at least ``n`` elements (C order); anything else
    raises :class:`LmgxError` before the library is called."""
    if not isinstance(a, np.ndarray):
        raise LmgxError(f"{what}: a numpy float64 array is required, got {type(a).__name__}")
    if a.dtype != np.float64:
        raise LmgxError(f"{what}: dtype {a.dtype}, float64 is required")
    if not a.flags.c_contiguous:
        raise LmgxError(f"{what}: the array must be C-contiguous")
    if a.size < n:
        raise LmgxError(f"{what}: {a.size} elements, at least {n} are required")
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _dp(a, n: int, what: str = "tensor") -> C.POINTER(C.c_double):
    """Pointer to a contiguous float64 CUDA tensor of at least ``n`` elements; a CPU tensor, a
    numpy array, another dtype or a short tensor raises :class:`LmgxError` before the library is
    called (the library itself cannot check device pointers).

    The library works on its own HIP stream, so a device tensor is handed over only after the
    work torch has queued on its current stream (which may still be writing the tensor) is done.
    The library synchronises its own stream before it returns, so its results are complete for
    torch."""
    import torch
    if not isinstance(a, torch.Tensor):
        raise LmgxError(f"{what}: a float64 CUDA tensor is required, got {type(a).__name__}")
    if not a.is_cuda:
        raise LmgxError(f"{what}: the tensor is on {a.device}; device entry points need a CUDA "
                        "tensor (host arrays go as numpy)")
    if a.dtype != torch.float64:
        raise LmgxError(f"{what}: dtype {a.dtype}, torch.float64 is required")
    if not a.is_contiguous():
        raise LmgxError(f"{what}: the tensor must be contiguous")
    if a.numel() < n:
        raise LmgxError(f"{what}: {a.numel()} elements, at least {n} are required")
    torch.cuda.current_stream(a.device).synchronize()
    return C.cast(C.c_void_p(a.data_ptr()), C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


@dataclass
class Topology:
    """Block dims (halos included), side codes and links, reference ``code/cb.lgrid.inc``."""
    nijkp1: np.ndarray     # [nblk, 3]
    bc: np.ndarray         # [nblk, 6]
    bl2bl: np.ndarray      # [nblk, 6]
    bl2side: np.ndarray    # [nblk, 6]
    nlev: int = 0

    @classmethod
    def from_problem(cls, p) -> "Topology":
        return cls(np.array([b.nijkp1 for b in p.blocks], np.int32),
                   np.array([b.bc for b in p.blocks], np.int32),
                   np.array([b.bl2bl for b in p.blocks], np.int32),
                   np.array([b.bl2side for b in p.blocks], np.int32), p.nummgl)


def default_params(**kw) -> Params:
    p = Params()
    lib().lmgx_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class Solver:
    """One LESter block topology resident on the GPU (the reference's commons + arena)."""

    def __init__(self, topo: Topology):
        L = lib()
        self._keep = [np.ascontiguousarray(a, np.int32).ravel()
                      for a in (topo.nijkp1, topo.bc, topo.bl2bl, topo.bl2side)]
        t = _Topo(len(topo.nijkp1), topo.nlev, *[_ip(a) for a in self._keep])
        h = C.c_void_p()
        _check(L.lmgx_init(C.byref(t), C.byref(h)))
        self.h = h
        nlev, ntot, n1 = C.c_int(), C.c_long(), C.c_long()
        nb = len(topo.nijkp1)
        dims = np.zeros((MAXL, nb, 3), np.int32)
        off = np.zeros((MAXL, nb), np.int64)
        _check(L.lmgx_levels(h, C.byref(nlev), C.byref(ntot), C.byref(n1), _ip(dims),
                             off.ctypes.data_as(C.POINTER(C.c_long))))
        self.nlev, self.ntot, self.n1, self.nblk = nlev.value, ntot.value, n1.value, nb
        self.nijkp1 = dims[: self.nlev].copy()
        self.off = off[: self.nlev].copy()
        # multi-level 2-D (ij) and 1-D (k) metric arrays, reference code/set_offset.F:166-234
        self.n2d = int(np.sum(self.nijkp1[..., 0].astype(np.int64) * self.nijkp1[..., 1]))
        self.n1d = int(np.sum(self.nijkp1[..., 2]))

    def _end(self, level: int) -> int:
        """Length of a multi-level array up to the end of ``level`` (the size a single pass on
        that level needs)."""
        if not 0 <= int(level) < self.nlev:
            raise LmgxError(f"level {level} out of range (0..{self.nlev - 1})")
        return int(self.off[level, -1] + np.prod(self.nijkp1[level, -1]))

    def close(self):
        if getattr(self, "h", None):
            lib().lmgx_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- operator -----------------------------------------------------------------------
    def set_operator(self, coeffs: Sequence):
        """ap, aw, ae, as, an, al, ar (ntot each): numpy (host) or torch CUDA tensors."""
        if len(coeffs) != 7:
            raise LmgxError(f"set_operator: 7 coefficient arrays are required, got {len(coeffs)}")
        names = ("ap", "aw", "ae", "as", "an", "al", "ar")
        if isinstance(coeffs[0], np.ndarray):
            host = [np.ascontiguousarray(a, np.float64) for a in coeffs]
            _check(lib().lmgx_set_operator(self.h, *[_hp(a, self.ntot, q)
                                                     for a, q in zip(host, names)]))
        else:
            arr = (C.POINTER(C.c_double) * 7)(*[_dp(a, self.ntot, q)
                                                for a, q in zip(coeffs, names)])
            _check(lib().lmgx_set_operator_device(self.h, arr))

    def set_operator_cached(self, coeffs: Sequence[np.ndarray], force: bool = False) -> bool:
        """Host coefficients as :meth:`set_operator`, uploaded only when ``force``, when the
        arrays are not those of the last upload through this method, or when a sampled
        fingerprint of them changed; returns whether they were uploaded."""
        if len(coeffs) != 7:
            raise LmgxError(f"set_operator_cached: 7 coefficient arrays are required")
        up = C.c_int()
        _check(lib().lmgx_set_operator_cached(self.h, *[_hp(a, self.ntot) for a in coeffs],
                                              int(bool(force)), C.byref(up)))
        return bool(up.value)

    def counters(self):
        """(operator uploads, fast-path operator preparations, graph captures, solves)."""
        out = (C.c_long * 4)()
        _check(lib().lmgx_counters(self.h, out))
        return tuple(out)

    def build_operator(self, problem, rho1: Optional[np.ndarray] = None, dirichlet=None):
        """Metrics + coefficients of every level on the device from the level-1 vertices."""
        xv = np.concatenate([np.broadcast_to(b.xv[None, :], (b.nijkp1[1] - 1, b.nijkp1[0] - 1)).ravel()
                             for b in problem.blocks]).astype(np.float64)
        yv = np.concatenate([np.broadcast_to(b.yv[:, None], (b.nijkp1[1] - 1, b.nijkp1[0] - 1)).ravel()
                             for b in problem.blocks]).astype(np.float64)
        zv = np.concatenate([b.zv for b in problem.blocks]).astype(np.float64)
        r = None if rho1 is None else np.ascontiguousarray(rho1, np.float64)
        d = None if dirichlet is None else np.ascontiguousarray(dirichlet, np.int32).ravel()
        _check(lib().lmgx_build_operator(self.h, _hp(xv, xv.size), _hp(yv, yv.size),
                                         _hp(zv, zv.size),
                                         None if r is None else _hp(r, self.n1, "rho1"),
                                         None if d is None else _ip(d)))

    def update_density(self, rho1=None):
        """Rebuild the coefficients of every level on the device for a new level-1 density (a
        device tensor of n1 with halos; None = 1), keeping the metrics and Dirichlet sides."""
        _check(lib().lmgx_update_density_device(
            self.h, None if rho1 is None else _dp(rho1, self.n1, "rho1")))

    def set_dirichlet(self, flags):
        f = None if flags is None else np.ascontiguousarray(flags, np.int32).ravel()
        _check(lib().lmgx_set_dirichlet(self.h, None if f is None else _ip(f)))

    def get_operator(self):
        a = [np.zeros(self.ntot) for _ in range(7)]
        _check(lib().lmgx_get_operator(self.h, *[_hp(x, self.ntot) for x in a]))
        return a

    def source(self, u, v, w, rho, rhonp1, dtime, b):
        n = self.n1
        _check(lib().lmgx_source_device(self.h, _dp(u, n, "u"), _dp(v, n, "v"), _dp(w, n, "w"),
                                        _dp(rho, n, "rho"), _dp(rhonp1, n, "rhonp1"),
                                        C.c_double(dtime), _dp(b, n, "b")))

    def uvw_adv(self, ivar, u, u_int, rho, p, dtime, diu=None, rp=None, dkw=None):
        """Velocity correction after the solve (reference ``code/uvw_adv.F:5``): component ``ivar``
        (1 U, 2 V, 3 W) of ``u`` from ``u_int``, ``rho`` (rho^{n+1}) and the pressure ``p``, then
        the interface/periodic swap of ``u``.  numpy arrays -> host entry point with the
        caller's metrics ``diu`` (djv for V), ``rp``, ``dkw``; torch CUDA tensors -> device entry
        point, metrics optional (default: those of ``build_operator``)."""
        if isinstance(u, np.ndarray):
            if diu is None or rp is None or dkw is None:
                raise LmgxError("host uvw_adv needs the metrics diu (djv for V), rp and dkw")
            n = self.n1
            _check(lib().lmgx_uvw_adv(self.h, int(ivar), _hp(diu, self.n2d, "diu"),
                                      _hp(rp, self.n2d, "rp"), _hp(dkw, self.n1d, "dkw"),
                                      _hp(u, n, "u"), _hp(u_int, n, "u_int"), _hp(rho, n, "rho"),
                                      _hp(p, n, "p"), C.c_double(dtime)))
        else:
            m = [None if a is None else _dp(a, sz, q)
                 for a, sz, q in ((diu, self.n2d, "diu"), (rp, self.n2d, "rp"),
                                  (dkw, self.n1d, "dkw"))]
            n = self.n1
            _check(lib().lmgx_uvw_adv_device(self.h, int(ivar), *m, _dp(u, n, "u"),
                                             _dp(u_int, n, "u_int"), _dp(rho, n, "rho"),
                                             _dp(p, n, "p"), C.c_double(dtime)))
        return u

    # ---- field dumps (reference code/file_3.F, code/read_3.F) ------------------------------
    def file_3(self, filename, f, filefrmt=0, level=0, mg_or_sg=1, real_kind=4):
        """Write level ``level`` of the device field ``f`` (multi-level natural layout) in the
        reference's 3-D file format: ``filefrmt`` 0 unformatted (REAL*4 values, or REAL*8 with
        ``real_kind=8``), 1 text."""
        need = self.n1 if int(mg_or_sg) == 2 else self._end(level)
        _check(lib().lmgx_file_3_device(self.h, os.fsencode(filename), int(filefrmt),
                                        _dp(f, need, "f"),
                                        int(level), int(mg_or_sg), int(real_kind)))

    def read_3(self, filename, f, filefrmt=0, real_kind=4) -> bool:
        """Read such a file into level 1 of the device field ``f``; a missing file gives zeros
        and returns False, as the reference does."""
        found = C.c_int()
        _check(lib().lmgx_read_3_device(self.h, os.fsencode(filename), int(filefrmt),
                                        _dp(f, self.n1, "f"),
                                        int(real_kind), C.byref(found)))
        return bool(found.value)

    # ---- solve --------------------------------------------------------------------------
    def lmg_solve(self, f, b, residu, *, max_cycles=5, history=True, **kw):
        """Run up to ``max_cycles`` V-cycles.  numpy arrays -> host entry point, torch CUDA
        tensors -> device entry point.  Returns the history [ncycles, nlev, 4] of
        (rnorm, fnorm, enorm, vnorm)."""
        p = default_params(max_cycles=max_cycles, **kw)
        hist = np.zeros((max(1, max_cycles), self.nlev, 4))
        nc = C.c_int()
        n = self.ntot
        hp = _hp(hist, hist.size)
        if isinstance(f, np.ndarray):
            _check(lib().lmgx_solve(self.h, _hp(f, n, "f"), _hp(b, n, "b"),
                                    _hp(residu, n, "residu"), C.byref(p), hp, C.byref(nc)))
        else:
            _check(lib().lmgx_solve_device(self.h, _dp(f, n, "f"), _dp(b, n, "b"),
                                           _dp(residu, n, "residu"), C.byref(p), hp,
                                           C.byref(nc)))
        return hist[: nc.value]

    def solve_adjoint(self, lam, g, res, *, max_cycles=50, history=True, **kw):
        """A^T lam = g with the transposed-operator multigrid (device tensors of ntot); same
        keywords and history as :meth:`lmg_solve`."""
        p = default_params(max_cycles=max_cycles, **kw)
        hist = np.zeros((max(1, max_cycles), self.nlev, 4))
        nc = C.c_int()
        n = self.ntot
        _check(lib().lmgx_solve_adjoint_device(self.h, _dp(lam, n, "lam"), _dp(g, n, "g"),
                                               _dp(res, n, "res"), C.byref(p),
                                               _hp(hist, hist.size), C.byref(nc)))
        return hist[: nc.value]

    def transpose_operator(self):
        """The transposed operator, 7 device tensors of ntot (ap aw ae as an al ar)."""
        import torch
        out = [torch.zeros(self.ntot, dtype=torch.float64, device="cuda") for _ in range(7)]
        arr = (C.POINTER(C.c_double) * 7)(*[_dp(a, self.ntot) for a in out])
        _check(lib().lmgx_transpose_operator_device(self.h, arr))
        return out

    def rho_adjoint(self, lam, p):
        """dJ/dlog(rho) (device tensor of n1) from level-1 p (A p = b) and lam (A^T lam = dJ/dp)."""
        import torch
        g = torch.zeros(self.n1, dtype=torch.float64, device="cuda")
        _check(lib().lmgx_rho_adjoint_device(self.h, _dp(lam, self.n1, "lam"),
                                             _dp(p, self.n1, "p"), _dp(g, self.n1)))
        return g

    @property
    def jumble(self) -> int:
        return lib().lmgx_get_jumble(self.h)

    @jumble.setter
    def jumble(self, v: int):
        _check(lib().lmgx_set_jumble(self.h, int(v)))

    # ---- single passes (device tensors) -------------------------------------------------
    def norms(self, level, x, b, res):
        out = np.zeros(4)
        n = self._end(level)
        _check(lib().lmgx_norms_device(self.h, level, _dp(x, n, "x"), _dp(b, n, "b"),
                                       _dp(res, n, "res"), _hp(out, 4)))
        return out

    def sporlx(self, level, f, b, omega):
        n = self._end(level)
        _check(lib().lmgx_sporlx_device(self.h, level, _dp(f, n, "f"), _dp(b, n, "b"),
                                        C.c_double(omega)))

    def sporlx2(self, level, f, b, omega):
        n = self._end(level)
        _check(lib().lmgx_sporlx2_device(self.h, level, _dp(f, n, "f"), _dp(b, n, "b"),
                                         C.c_double(omega)))

    def swap(self, level, var):
        _check(lib().lmgx_swap_device(self.h, level, _dp(var, self._end(level), "var")))

    def restrict(self, coarse_level, res, b, f):
        n = self._end(coarse_level)
        _check(lib().lmgx_restrict_device(self.h, coarse_level, _dp(res, n, "res"),
                                          _dp(b, n, "b"), _dp(f, n, "f")))

    def prolong(self, fine_level, f):
        _check(lib().lmgx_prolong_device(self.h, fine_level,
                                         _dp(f, self._end(int(fine_level) + 1), "f")))

    def extrap_p(self, p):
        _check(lib().lmgx_extrap_p_device(self.h, _dp(p, self.n1, "p")))

    def p_ref_subtract(self, p):
        _check(lib().lmgx_p_ref_subtract_device(self.h, _dp(p, self.n1, "p")))

    def p_mean_subtract(self, p) -> float:
        """Mean-pressure gauge: subtract the level-1 interior mean from every level-1 point of
        the device tensor ``p`` (halos included); returns the mean."""
        m = C.c_double()
        _check(lib().lmgx_p_mean_subtract_device(self.h, _dp(p, self.n1, "p"), C.byref(m)))
        return m.value

    def p_mean_subtract_noalloc(self, p, mean=None):
        """The mean-pressure gauge through the entry point that allocates nothing (static device
        scratch); ``mean``: optional float64 CUDA tensor of >= 1 element receiving the mean."""
        m = None if mean is None else _dp(mean, 1, "mean")
        _check(lib().lmgx_p_mean_subtract_device_noalloc(self.h, _dp(p, self.n1, "p"), m))

    def global_sum(self, level, f):
        out = np.zeros(3)
        _check(lib().lmgx_global_sum_device(self.h, level, _dp(f, self._end(level), "f"),
                                            _hp(out, 3)))
        return out

    def tile_widths(self):
        out = np.zeros(2 * MAXL, np.int32)
        _check(lib().lmgx_tile_widths(self.h, _ip(out)))
        return out.reshape(2, MAXL)[:, : self.nlev]

    def get_tuning(self):
        """The autotuner's choices [precision][kernel, halo stores][level] (zeros: not tuned)."""
        out = np.zeros(4 * MAXL, np.int32)
        _check(lib().lmgx_get_tuning(self.h, _ip(out)))
        return out.reshape(2, 2, MAXL)[:, :, : self.nlev].copy()

    def set_tuning(self, table):
        """Install a table from :meth:`get_tuning` (same shape) for the next fast solves."""
        t = np.zeros((2, 2, MAXL), np.int32)
        a = np.asarray(table, np.int32)
        t[:, :, : a.shape[2]] = a
        t = np.ascontiguousarray(t).ravel()
        _check(lib().lmgx_set_tuning(self.h, _ip(t)))

    def fused_halos(self):
        """[precision][level]: 1 where the line kernel stores the face-interior halos itself."""
        out = np.zeros(2 * MAXL, np.int32)
        _check(lib().lmgx_fused_halos(self.h, _ip(out)))
        return out.reshape(2, MAXL)[:, : self.nlev]

    def coarsest_info(self):
        """(exact solve, banded factor, half-bandwidth, components) of the coarsest level."""
        out = np.zeros(4, np.int32)
        _check(lib().lmgx_coarsest_info(self.h, _ip(out)))
        return tuple(int(v) for v in out)

    def coef16_levels(self):
        """Per level: 1 where the FP32 line relaxation reads binary16 coefficients."""
        out = np.zeros(MAXL, np.int32)
        _check(lib().lmgx_coef16_levels(self.h, _ip(out)))
        return out[: self.nlev]

    def finished_in_cycle(self) -> bool:
        """True when the last fast solve ended on a cycle that wrote the returned solution and
        residual itself (no final conversion and residual pass)."""
        out = C.c_int(0)
        _check(lib().lmgx_finished_in_cycle(self.h, C.byref(out)))
        return bool(out.value)

    def metric_form(self):
        """(FP32 relaxation, FP64 residual): whether the last solve took level 1's builder
        operator in its metric form (couplings recomputed from the densities and face factors)."""
        out = np.zeros(2, np.int32)
        _check(lib().lmgx_metric_form(self.h, _ip(out)))
        return bool(out[0]), bool(out[1])

    # ---- profiling ----------------------------------------------------------------------
    def profile(self, enable=True):
        _check(lib().lmgx_profile(self.h, int(enable)))

    def profile_read(self):
        ms, n, cells = C.c_double(), C.c_long(), C.c_double()
        _check(lib().lmgx_profile_read(self.h, C.byref(ms), C.byref(n), C.byref(cells)))
        return ms.value, n.value, cells.value
