"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of ``oracle/liblmg_oracle.so``, the plain-C
restatement of the reference pressure path.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg import it, as the oracle the library is checked against.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblmg_oracle.so")
MAXB, MAXL = 64, 8

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.lmgo_sizeof.restype = C.c_ulong
    return _lib


class Params(C.Structure):
    _fields_ = [("cycle", C.c_int), ("numpre", C.c_int), ("numpost", C.c_int),
                ("numcits", C.c_int), ("ncycles", C.c_int), ("omega", C.c_double),
                ("omega_c", C.c_double), ("tolf", C.c_double), ("meth_smooth", C.c_int)]


class _G(C.Structure):
    _fields_ = [("nblk", C.c_int), ("nlev", C.c_int),
                ("nijkp1", C.c_int * 3 * MAXB * MAXL),
                ("off", C.c_long * MAXB * MAXL), ("off2d", C.c_long * MAXB * MAXL),
                ("off1d", C.c_long * MAXB * MAXL), ("offxy", C.c_long * MAXB * MAXL),
                ("offz", C.c_long * MAXB * MAXL),
                ("nijkse", C.c_int * 6 * MAXB * MAXL), ("mult", C.c_int * 3 * MAXB * MAXL),
                ("npts", C.c_double * MAXL),
                ("bc", C.c_int * 6 * MAXB), ("bl2bl", C.c_int * 6 * MAXB),
                ("bl2side", C.c_int * 6 * MAXB),
                ("ntot", C.c_long), ("n1", C.c_long), ("n2d", C.c_long), ("n1d", C.c_long),
                ("nxy", C.c_long), ("nz", C.c_long), ("jumble", C.c_int)]


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Oracle:
    """One problem topology (the reference's commons) plus its SAVEd sweep counter."""

    def __init__(self, problem):
        L = lib()
        assert L.lmgo_sizeof() == C.sizeof(_G), (L.lmgo_sizeof(), C.sizeof(_G))
        self.g = _G()
        self.problem = problem
        p = problem
        nijk = np.array([b.nijkp1 for b in p.blocks], np.int32).ravel()
        bc = np.array([b.bc for b in p.blocks], np.int32).ravel()
        l2b = np.array([b.bl2bl for b in p.blocks], np.int32).ravel()
        l2s = np.array([b.bl2side for b in p.blocks], np.int32).ravel()
        rc = L.lmgo_init(C.byref(self.g), p.nblk, p.nummgl, _pi(nijk), _pi(bc), _pi(l2b), _pi(l2s))
        if rc != 0:
            raise ValueError(f"lmgo_init failed: {rc}")
        g = self.g
        self.nlev, self.nblk = g.nlev, g.nblk
        self.ntot, self.n1, self.n2d, self.n1d = g.ntot, g.n1, g.n2d, g.n1d
        self.nxy, self.nz = g.nxy, g.nz
        self.nijkp1 = np.array(g.nijkp1)[: self.nlev, : self.nblk]
        self.off = np.array(g.off)[: self.nlev, : self.nblk]
        self.npts = np.array(g.npts)[: self.nlev]

    # ---- fields ---------------------------------------------------------------------------
    def level1(self, per_block):
        return np.concatenate([np.ascontiguousarray(a, np.float64).ravel() for a in per_block])

    def geometry(self):
        p = self.problem
        xv = np.concatenate([np.broadcast_to(b.xv[None, :], (b.nijkp1[1] - 1, b.nijkp1[0] - 1)).ravel()
                             for b in p.blocks]).astype(np.float64)
        yv = np.concatenate([np.broadcast_to(b.yv[:, None], (b.nijkp1[1] - 1, b.nijkp1[0] - 1)).ravel()
                             for b in p.blocks]).astype(np.float64)
        zv = np.concatenate([b.zv for b in p.blocks]).astype(np.float64)
        x = np.zeros(self.nxy); y = np.zeros(self.nxy); z = np.zeros(self.nz)
        met = np.zeros(7 * self.n2d); dkp = np.zeros(self.n1d); dkw = np.zeros(self.n1d)
        rc = lib().lmgo_geometry(C.byref(self.g), _p(xv), _p(yv), _p(zv), _p(x), _p(y), _p(z),
                                 _p(met), _p(dkp), _p(dkw))
        if rc != 0:
            raise ValueError(f"lmgo_geometry failed: {rc}")
        return dict(x=x, y=y, z=z, met=met.reshape(7, self.n2d), dkp=dkp, dkw=dkw)

    def rho_levels(self, rho1):
        rho = np.ones(self.ntot)
        rho[: self.n1] = rho1
        lib().lmgo_rho_levels(C.byref(self.g), _p(rho))
        return rho

    def coeff(self, geo, rho=None, dirichlet=None):
        a = [np.zeros(self.ntot) for _ in range(7)]
        met = np.ascontiguousarray(geo["met"]).ravel()
        d = None if dirichlet is None else np.ascontiguousarray(dirichlet, np.int32).ravel()
        lib().lmgo_coeff(C.byref(self.g), _p(met), _p(geo["dkp"]), _p(geo["dkw"]),
                         None if rho is None else _p(rho), None if d is None else _pi(d),
                         *[_p(x) for x in a])
        return a

    def sour(self, geo, u, v, w, rho, rhonp1, dtime):
        b = np.zeros(self.ntot)
        met = np.ascontiguousarray(geo["met"]).ravel()
        lib().lmgo_sour(C.byref(self.g), _p(met), _p(geo["dkp"]), _p(geo["dkw"]), _p(u), _p(v),
                        _p(w), _p(rho), _p(rhonp1), C.c_double(dtime), _p(b))
        return b

    def uvw_adv(self, ivar, geo, u, u_int, rho, p, dtime):
        """Velocity correction of component ivar (1 U, 2 V, 3 W) into u (level-1 layout), then the
        level-1 swap of u.  p may be the multi-level array (its level-1 part is read)."""
        met = np.ascontiguousarray(geo["met"]).ravel()
        lib().lmgo_uvw_adv(C.byref(self.g), int(ivar), _p(met), _p(geo["dkw"]), _p(u),
                           _p(u_int), _p(rho), _p(p), C.c_double(dtime))
        return u

    def solve(self, coeffs, f, b, ncycles, cycle=1, tolf=0.0, res=None, meth_smooth=1):
        p = self.problem
        prm = Params(cycle, p.numpre, p.numpost, p.numcits, ncycles, p.omega, p.omega_c, tolf,
                     meth_smooth)
        if res is None:
            res = np.zeros(self.ntot)
        hist = np.zeros((ncycles, self.nlev, 4))
        lib().lmgo_solve(C.byref(self.g), C.byref(prm), *[_p(a) for a in coeffs], _p(f), _p(b),
                         _p(res), _p(hist))
        return hist, res

    def norms(self, l, coeffs, x, b, res):
        out = np.zeros(4)
        arr = (C.POINTER(C.c_double) * 7)(*[_p(a) for a in coeffs])
        lib().lmgo_norms(C.byref(self.g), l, arr, _p(x), _p(b), _p(res), _p(out))
        return out

    def sporlx(self, l, coeffs, f, b, omega, inout=None, soltol=0.0):
        io = np.array([0.0, 0.0, 1.0, 1.0]) if inout is None else inout
        arr = (C.POINTER(C.c_double) * 7)(*[_p(a) for a in coeffs])
        lib().lmgo_sporlx(C.byref(self.g), l, arr, _p(f), _p(b), C.c_double(omega), _p(io),
                          C.c_double(soltol))
        return io

    def sporlx2(self, l, coeffs, f, b, omega, inout=None, soltol=0.0):
        io = np.array([0.0, 0.0, 1.0, 1.0]) if inout is None else inout
        arr = (C.POINTER(C.c_double) * 7)(*[_p(a) for a in coeffs])
        lib().lmgo_sporlx2(C.byref(self.g), l, arr, _p(f), _p(b), C.c_double(omega), _p(io),
                           C.c_double(soltol))
        return io

    def swap(self, l, var):
        lib().lmgo_swap(C.byref(self.g), l, _p(var))

    def restrict(self, lc, fine, coarse):
        lib().lmgo_restrict(C.byref(self.g), lc, _p(fine), _p(coarse))

    def prolong(self, lf, f):
        lib().lmgo_prolong(C.byref(self.g), lf, _p(f))

    def extrap_p(self, p):
        lib().lmgo_extrap_p(C.byref(self.g), _p(p))

    def p_ref_subtract(self, p):
        lib().lmgo_p_ref_subtract(C.byref(self.g), _p(p))

    def global_sum(self, l, f):
        out = np.zeros(3)
        lib().lmgo_global_sum(C.byref(self.g), l, _p(f), _p(out))
        return out

    def p_mean_subtract(self, p):
        """Subtract the level-1 interior mean from every level-1 point; returns the mean."""
        m = C.c_double()
        lib().lmgo_p_mean_subtract(C.byref(self.g), _p(p), C.byref(m))
        return m.value

    @property
    def jumble(self):
        return self.g.jumble

    @jumble.setter
    def jumble(self, v):
        self.g.jumble = v
