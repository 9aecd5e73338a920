"""TEST INFRASTRUCTURE ONLY: runs the reference Fortran pressure path (``oracle/_ref/ref_driver``,
built by ``make -C oracle ref`` from the reference sources) on a synthetic problem and parses the
dump.  Only ``tests/``, ``tests/golden/`` and ``bench.py``'s CPU-baseline leg use this module.

Input/output formats are documented at the top of ``oracle/ref_driver.F``.
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile
from typing import Dict, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "_ref", "ref_driver")
# the same driver with the reference's lmg_solve replaced by the product's Fortran drop-in
DRIVER_GPU = os.path.join(HERE, "_ref", "ref_driver_gpu")


def available(gpu: bool = False) -> bool:
    return os.path.exists(DRIVER_GPU if gpu else DRIVER)


def write_input(path: str, problem, *, ncycles: int, coeffs: Optional[Sequence[np.ndarray]] = None,
                b1: Optional[np.ndarray] = None, fields: Optional[dict] = None,
                f0: Optional[np.ndarray] = None, n1: Optional[int] = None, do_post: int = 1,
                ntime: int = 1, tolf: float = 0.0, cycle: int = 1, meth_smooth: int = 1) -> None:
    p = problem
    with open(path, "wb") as fh:
        w = fh.write
        w(struct.pack("<13i", 4711, p.nblk, p.nummgl, cycle, p.numpre, p.numpost, p.numcits, ncycles,
                      0 if coeffs is None else 1, 0 if b1 is None else 1, do_post, ntime,
                      meth_smooth))
        w(struct.pack("<4d", p.omega, p.omega_c, tolf, p.dtime))
        for b in p.blocks:
            w(struct.pack("<21i", *b.nijkp1, *b.bc, *b.bl2bl, *b.bl2side))
        for b in p.blocks:
            nip1, njp1, nkp1 = b.nijkp1
            xx = np.broadcast_to(b.xv[None, :], (njp1 - 1, nip1 - 1))
            yy = np.broadcast_to(b.yv[:, None], (njp1 - 1, nip1 - 1))
            w(np.ascontiguousarray(xx, "<f8").tobytes())
            w(np.ascontiguousarray(yy, "<f8").tobytes())
            w(np.ascontiguousarray(b.zv, "<f8").tobytes())
        if coeffs is not None:
            for a in coeffs:
                w(np.ascontiguousarray(a, "<f8").tobytes())
        if b1 is None:
            for key in ("u", "v", "w", "rho", "rhonp1"):
                w(np.ascontiguousarray(fields[key], "<f8").tobytes())
        else:
            w(np.ascontiguousarray(b1, "<f8").tobytes())
        if f0 is None:
            f0 = np.zeros(n1)
        w(np.ascontiguousarray(f0, "<f8").tobytes())


def read_output(path: str) -> Dict[str, np.ndarray]:
    raw = open(path, "rb").read()
    pos = 0

    def ints(n):
        nonlocal pos
        v = np.frombuffer(raw, "<i4", n, pos).copy()
        pos += 4 * n
        return v

    def reals(n):
        nonlocal pos
        v = np.frombuffer(raw, "<f8", n, pos).copy()
        pos += 8 * n
        return v

    hdr = ints(10)
    assert hdr[0] == 4712, "bad output magic"
    nbl, L, ntot, n1, n2d, n1d, nxy, nz, ncyc = (int(x) for x in hdr[1:])
    out = dict(numbl=nbl, nummgl=L, ntot=ntot, n1=n1, ncycles=ncyc)
    out["nijkp1"] = ints(3 * nbl * L).reshape(L, nbl, 3)
    out["mg_off"] = ints(nbl * L).reshape(L, nbl)
    out["mg_off_2d_ij"] = ints(nbl * L).reshape(L, nbl)
    out["mg_off_1d_k"] = ints(nbl * L).reshape(L, nbl)
    out["mg_off_2d_xy"] = ints(nbl * L).reshape(L, nbl)
    out["mg_off_1d_z"] = ints(nbl * L).reshape(L, nbl)
    out["nijkse"] = ints(6 * nbl * L).reshape(L, nbl, 6)
    out["npts_all"] = reals(L)
    out["tsolve"] = float(reals(1)[0])
    out["x"], out["y"], out["z"] = reals(nxy), reals(nxy), reals(nz)
    out["metrics"] = reals(n2d * 35).reshape(35, n2d)
    out["dkp"], out["dkw"] = reals(n1d), reals(n1d)
    for k in ("ap", "aw", "ae", "as", "an", "al", "ar"):
        out[k] = reals(ntot)
    out["b"], out["f"], out["res"] = reals(ntot), reals(ntot), reals(ntot)
    out["hist"] = reals(4 * L * ncyc).reshape(ncyc, L, 4)
    # global_sum of (f, b, residu) per level: [level][field][sum, max, min]; global_sum_block of
    # each block: [level][block][field][sum, max, min]
    out["gsum"] = reals(9 * L).reshape(L, 3, 3)
    out["gsum_blk"] = reals(9 * nbl * L).reshape(L, nbl, 3, 3)
    out["p_post"] = reals(n1)
    nvel = int(ints(1)[0])
    if nvel:
        out["u_post"], out["v_post"], out["w_post"] = reals(nvel), reals(nvel), reals(nvel)
    ndump, out["dump_ierr"] = (int(x) for x in ints(2))
    if ndump:
        out["p_read_bin"], out["p_read_txt"], out["p_read_none"] = (reals(ndump), reals(ndump),
                                                                     reals(ndump))
    assert pos == len(raw), (pos, len(raw))
    return out


def run(problem, *, ncycles: int, coeffs=None, b1=None, fields=None, f0=None, n1=None,
        do_post: int = 1, ntime: int = 1, tolf: float = 0.0, timeout: float = 3600.0,
        cycle: int = 1, gpu: bool = False, env_extra: Optional[dict] = None,
        dump: bool = False, meth_smooth: int = 1):
    """dump: also write the final p with the reference's file_3 (binary and text, in the
    driver's working directory) and read both back with read_3; the files' bytes are returned
    as out["file_p_bin"] / out["file_p_txt"].  meth_smooth: the reference's smoother switch
    (1 = sporlx, 2 = sporlx2).  out["stdout"] is what the driver printed."""
    if not available(gpu):
        raise RuntimeError("reference driver not built: make -C oracle ref ref-gpu")
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_input(fin, problem, ncycles=ncycles, coeffs=coeffs, b1=b1, fields=fields, f0=f0,
                    n1=n1, do_post=do_post, ntime=ntime, tolf=tolf, cycle=cycle,
                    meth_smooth=meth_smooth)
        env = dict(os.environ)
        env.update(env_extra or {})
        args = [DRIVER_GPU if gpu else DRIVER, fin, fout] + (["dump"] if dump else [])
        r = subprocess.run(args, capture_output=True, text=True, errors="replace", env=env,
                           timeout=timeout, cwd=td)
        if r.returncode != 0 or not os.path.exists(fout):
            raise RuntimeError(f"ref_driver failed ({r.returncode}): {r.stdout}\n{r.stderr}")
        out = read_output(fout)
        out["stdout"] = r.stdout
        if dump:
            for name in ("p_bin", "p_txt"):
                with open(os.path.join(td, name), "rb") as fh:
                    out["file_" + name] = fh.read()
        return out


def run_timing_concurrent(problems, *, ncycles: int, b1s, timeout: float = 900.0):
    """Time the reference's lmg_solve on independent problems at once, one single-rank driver
    process per problem (the reference's own coeff_p on each problem's grid, the given level-1
    sources, zero initial guess).  Returns [(tsolve seconds, hist[ncycles, nummgl, 4])] in
    order.  Independent problems in concurrent processes do what an MPI run with one rank per
    problem does, minus the global norm reductions."""
    if not available():
        raise RuntimeError("reference driver not built: make -C oracle ref")
    with tempfile.TemporaryDirectory() as td:
        procs = []
        for q, (prob, b1) in enumerate(zip(problems, b1s)):
            fin, fout = os.path.join(td, f"in{q}.bin"), os.path.join(td, f"out{q}.bin")
            write_input(fin, prob, ncycles=ncycles, b1=b1, n1=b1.size, do_post=-1)
            log = open(os.path.join(td, f"log{q}.txt"), "wb")     # no pipe can stall a run
            procs.append((subprocess.Popen([DRIVER, fin, fout], stdout=log,
                                           stderr=subprocess.STDOUT, cwd=td), fout, log))
        out = []
        for pr, fout, log in procs:
            pr.wait(timeout=timeout)
            log.close()
            if pr.returncode != 0 or not os.path.exists(fout):
                msg = open(log.name, "rb").read()[-400:]
                raise RuntimeError(f"ref_driver failed ({pr.returncode}): {msg!r}")
            raw = open(fout, "rb").read()
            magic, ncyc, L = np.frombuffer(raw, "<i4", 3, 0)
            assert magic == 4713, "bad timing output magic"
            t = float(np.frombuffer(raw, "<f8", 1, 12)[0])
            hist = np.frombuffer(raw, "<f8", 4 * L * ncyc, 20).reshape(ncyc, L, 4).copy()
            out.append((t, hist))
        return out


def level1_fields(problem, fields_per_block: dict) -> dict:
    """Concatenate per-block [k,j,i] arrays into the level-1 mg layout (blocks in order)."""
    return {k: np.concatenate([a.ravel() for a in v]) for k, v in fields_per_block.items()}
