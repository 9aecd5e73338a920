"""Benchmark: LESter pressure-Poisson multigrid on one MI355X.

Workload (default): problem "c3" - 16 uneven blocks (seeded nx 128..640, ny 64..256), 119.5M cells, nz = 128,
tanh wall-stretched y, wall-layer density with 5% lognormal noise, 5 Gaussian source blobs on top
of the divergence of a seeded N(0,1) velocity field, face-Dirichlet outflow, 6 levels.  Synthetic
data, generated on the GPU before the timed region.

One "step" = one complete solve from p = 0 to a 1e-10 relative residual reduction (rnorm_k /
rnorm_0, the reference's FP64 RMS residual, code/norms.F:118).  metric value = Mcell x V-cycles
per second over the timed steps; also reported: time to 1e-10, the same for the plain FP64 GPU
path, the speed-up, and the pressure difference between the two.

Prints one JSON line: metric, value (Mcell x V-cycles/s of the mixed-precision solve), the
time to 1e-10, the plain FP64 comparison, the roofline of the dominant kernel (the finest-level
FP32 line-relaxation half-sweep, HIP events on the solver's stream) and the CPU baseline.

Profiling helpers (run on rocprofv3 output):
  --pmc-summary CSV [CSV ...]   per-kernel dispatch count and average counter value
  --traffic DIR                 HBM bytes per launch of the finest line kernel from FETCH_SIZE /
                                WRITE_SIZE runs of this benchmark and of --copy-calib
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--scale", type=int, default=1)
    ap.add_argument("--mode", default="mixed", choices=["mixed", "fp64", "exact"])
    ap.add_argument("--rtol", type=float, default=1e-10)
    # V(1,1) with 3 over-relaxed line sweeps per smoothing iteration (the reference's smoothing
    # iteration is 3 SOR sweeps, its default V(1,1)); the fastest schedule to 1e-10 measured on c3
    ap.add_argument("--numpre", type=int, default=1)
    ap.add_argument("--numpost", type=int, default=1)
    ap.add_argument("--spi", type=int, default=3, help="line-relaxation sweeps per smoothing iteration")
    ap.add_argument("--fine-post", type=int, default=2,
                    help="post-smoothing sweeps on level 1 in a lean cycle (0 = numpost * spi)")
    ap.add_argument("--numcits", type=int, default=20)
    ap.add_argument("--omega", type=float, default=1.35)
    ap.add_argument("--omega-c", type=float, default=1.0)
    ap.add_argument("--max-cycles", type=int, default=100)
    ap.add_argument("--no-compare", action="store_true", help="skip the plain FP64 GPU path")
    ap.add_argument("--compare-per-block", action="store_true",
                    help="also time the same solve with the level work launched block by block "
                         "(a per-block launch loop) instead of once per level for all blocks")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cycles", type=int, default=3,
                    help="cycles of the single-chain CPU baseline (the fallback)")
    ap.add_argument("--cpu-cycles-all", type=int, default=1,
                    help="cycles of the whole-workload CPU baseline (all x-chains at once)")
    ap.add_argument("--cpu-one-chain", action="store_true",
                    help="CPU baseline on the first x-chain only, one core")
    ap.add_argument("--no-stream-ceiling", action="store_true",
                    help="skip the measured streaming rate of this GPU (torch kernels)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates rocprof FETCH/WRITE_SIZE)")
    ap.add_argument("--pmc-summary", nargs="+", metavar="CSV")
    ap.add_argument("--stats-summary", metavar="CSV", help="markdown table of a rocprofv3 kernel_stats.csv")
    ap.add_argument("--stats-diff", nargs=3, metavar=("CSV_A", "CSV_B", "N"),
                    help="per-solve kernel time: (B - A) / N for two kernel_stats.csv of runs that "
                         "differ by N timed solves")
    ap.add_argument("--traffic", metavar="DIR")
    ap.add_argument("--step-parts", action="store_true",
                    help="also time the parts of a projection step around the solve (density "
                         "update, source, velocity correction, field dump)")
    return ap.parse_args()


def copy_calib():
    import torch
    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calib_bytes": 4 * n, "dispatches": 3}))


# read/write mixes of the streaming probe: pure reads, the level-1 FP64 residual's (40 B read,
# 4 B written per cell), 9:1, the finest line relaxation's (24 B read, 4 B written) and a copy;
# the last three again with every load and store carrying the non-temporal hint, as the finest
# line kernels' once-read streams do on a level larger than the Infinity Cache
STREAM_MIXES = ((10, 0, 0), (10, 1, 0), (9, 1, 0), (6, 1, 0), (1, 1, 0),
                (9, 1, 1), (6, 1, 1), (1, 1, 1))


def stream_ceiling(reps=5):
    """What plain streaming kernels reach on this GPU, as the practical ceilings next to the
    published 8 TB/s: fortranutils_amd/libstreamprobe.so reads R and writes W arrays of 16-byte
    elements (512 MiB each) in one grid-stride pass, best of ``reps`` timed launches, for each
    mix of STREAM_MIXES, with ordinary or non-temporal accesses.  HBM3E loses bandwidth when a
    read stream carries a small share of writes, so a kernel is compared with the ceiling of its
    own read/write mix and access kind."""
    import ctypes as C
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortranutils_amd",
                        "libstreamprobe.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: python -c 'import __graft_entry__ as g; g.build()'")
    lib = C.CDLL(path)
    lib.lmgx_stream_probe_hint.argtypes = [C.c_int, C.c_int, C.c_int, C.c_long, C.c_int,
                                           C.POINTER(C.c_double)]
    out = {}
    for r, w, nt in STREAM_MIXES:
        v = C.c_double()
        rc = lib.lmgx_stream_probe_hint(r, w, nt, 32 << 20, reps, C.byref(v))
        if rc:
            raise RuntimeError(f"stream probe R{r} W{w} NT{nt} failed ({rc})")
        out[f"r{r}w{w}{'nt' if nt else ''}_GBps"] = v.value
    return out


def solve_bytes(s, args, mode, cycles, post0=None):
    """Algorithmic HBM bytes of one solve of ``cycles`` cycles with this schedule, per kernel and
    level (halo swaps, O(surface), count 0; symmetric operator, as the solver checks):
      half-sweep: 10 words per updated cell (9 with omega = 1); on a level whose FP32 segmented
        solve reads binary16 coefficients the 6 operator words are 2 bytes each (28 / 24 bytes);
        the two zero-start half-sweeps that open a level's correction on the segmented kernel
        read 2 and 1 fewer words of x;
      zeroing a level's correction where no zero start applies: 1 word per cell;
      restriction: residual from b, x, ap, ae, an, ar (6 words per fine cell; binary16 d, ae,
        an, ar where the level's line solve reads those) + the coarse b;
      prolongation + correction: fine x read and written + coarse x;
      coarsest direct solve: the dense inverse (n^2 FP64) and two vectors;
      level 1, mixed: FP64 residual + norms (6 FP64 words) writing the FP32 right-hand side, and
        x64 += e32 (FP64 x read and written, e read); FP64: the residual without the rhs;
      per solve: f and b to the split layout, the first residual, the solution back and its
        natural-layout residual (ap, ae, an, ar, b, x read, r written).
    post0: level 1's post-smoothing sweeps per cycle, on average over the solve's lean and full
    cycles (default numpost * spi)."""
    L = s.nlev
    N = [float(np.prod(np.asarray(s.nijkp1[l]) - 2, axis=1).sum()) for l in range(L)]
    mixed = mode == "mixed"
    w = [4 if mixed and l < L - 1 else 8 for l in range(L)]
    hw = 10 if args.omega != 1.0 else 9
    tw = s.tile_widths()[1 if mixed else 0]
    c16 = s.coef16_levels() if mixed else np.zeros(L, np.int32)
    npre = args.numpre * args.spi
    cyc = 0.0
    for l in range(L - 1):
        hb = hw * w[l] - (6 * w[l] - 12 if c16[l] and int(tw[l]) in (8, 16) else 0)
        npost = post0 if (l == 0 and post0 is not None) else args.numpost * args.spi
        cyc += (npre + npost) * N[l] * hb
        if mixed and npre >= 1 and int(tw[l]) in (8, 16):
            cyc -= 0.5 * N[l] * ((hw - 8) + (hw - 9)) * w[l]     # zero-start half-sweeps
        elif mixed or l > 0:
            cyc += N[l] * w[l]                                     # zeroed before use
        h16 = bool(c16[l]) and int(tw[l]) in (8, 16)
        cyc += (2 * w[l] + 8 if h16 else 6 * w[l]) * N[l] + N[l + 1] * w[l + 1]   # restriction
        cyc += 2 * N[l] * w[l] + N[l + 1] * w[l + 1]               # prolongation + correction
    nc = N[L - 1]
    cyc += 8 * (nc * nc + 2 * nc)                                  # coarsest direct solve
    n1 = N[0]
    cyc += n1 * (6 * 8 + (4 if mixed else 0))                      # FP64 residual + norms
    if mixed:
        cyc += n1 * 12                                             # x + e32 read, x + e written
    setup = n1 * (32 + 6 * 8 + (4 if mixed else 0) + 16 + 56)
    return cycles * cyc + setup


def make_problem(args):
    import torch
    from fortranutils_amd import lmgx, problems as P
    p = P.config(args.config, args.scale)
    s = lmgx.Solver(lmgx.Topology.from_problem(p))
    fl = P.device_fields(p)
    dirich = np.array([b.dirichlet for b in p.blocks], np.int32)
    rho1 = fl["rho"].cpu().numpy() if p.rho_kind != "uniform" else None
    s.build_operator(p, rho1=rho1, dirichlet=dirich)
    b = torch.zeros(s.ntot, dtype=torch.float64, device="cuda")
    s.source(fl["u"], fl["v"], fl["w"], fl["rho"], fl["rhonp1"], p.dtime, b)
    P.device_blobs(p, b[: s.n1])
    m = P.interior_mask_torch(p)
    b1 = b[: s.n1]
    b1[m] -= b1[m].mean()
    del fl
    torch.cuda.synchronize()
    return p, s, b


def solve(s, b, args, mode, rtol, profile=False):
    import torch
    f = torch.zeros_like(b)
    res = torch.zeros_like(b)
    kw = dict(tolf=0.0, rtol=rtol, max_cycles=args.max_cycles, numcits=args.numcits)
    if mode == "exact":
        kw.update(smoother=0, precision=0, numpre=2, numpost=2, omega=1.6, omega_c=1.6)
    else:
        kw.update(smoother=1, precision=1 if mode == "mixed" else 0, numpre=args.numpre,
                  numpost=args.numpost, sweeps_per_iter=args.spi, fine_post=args.fine_post,
                  omega=args.omega,
                  omega_c=args.omega_c)
    if profile:
        s.profile(True)
    hist = s.lmg_solve(f, b, res, **kw)
    if profile:
        ms, n, cells = s.profile_read()
        s.profile(False)
        return f, hist, (ms, n, cells)
    return f, hist, None


def step_parts(p, s, args, reps=5):
    """Wall time (ms, mean of reps; every library call synchronises its stream) of the parts of
    one projection step of the LES time step (code/step.F:353-425) on the resident problem:
    coefficients for a new density, source, the solve, extrap_p + p_ref_subtract, the three
    velocity corrections; and a REAL*4 binary dump of p (code/file_3.F) to local disk."""
    import tempfile
    import torch
    from fortranutils_amd import problems as P
    fl = P.device_fields(p)
    n1 = s.n1

    def clock(fn, n=reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t0) / n

    b = torch.zeros(s.ntot, dtype=torch.float64, device="cuda")
    f = torch.zeros_like(b)
    uo = fl["u"].clone()
    out = {}
    out["update_density"] = clock(lambda: s.update_density(fl["rhonp1"]))
    out["source"] = clock(lambda: s.source(fl["u"], fl["v"], fl["w"], fl["rho"], fl["rhonp1"],
                                           p.dtime, b))
    out["solve_same_operator"] = clock(lambda: solve(s, b, args, "mixed", args.rtol))
    f = solve(s, b, args, "mixed", args.rtol)[0]

    def new_op_solve():
        s.update_density(fl["rhonp1"])
        solve(s, b, args, "mixed", args.rtol)
    out["update_density+solve"] = clock(new_op_solve)
    out["extrap_p+p_ref_subtract"] = clock(lambda: (s.extrap_p(f), s.p_ref_subtract(f)))
    out["uvw_adv (one component)"] = clock(lambda: s.uvw_adv(1, uo, fl["u"], fl["rhonp1"], f,
                                                             p.dtime))
    with tempfile.TemporaryDirectory() as td:
        out["file_3 binary REAL*4"] = clock(lambda: s.file_3(os.path.join(td, "p"), f), n=2)
    # the differentiable solve's backward (fortranutils_amd/adjoint.py): the transposed-operator
    # FP64 multigrid to a 1e-13 reduction and the density gradient, on the same operator
    from fortranutils_amd.adjoint import ADJ_SOLVE
    g = torch.zeros_like(b)
    g[:n1] = b[:n1]
    lam = torch.zeros_like(b)
    res = torch.zeros_like(b)
    cyc = []

    def adjoint():
        lam.zero_()
        cyc.append(len(s.solve_adjoint(lam, g, res, **ADJ_SOLVE)))
        s.rho_adjoint(lam[:n1].contiguous(), f[:n1].contiguous())
    out["adjoint solve + density gradient"] = clock(adjoint, n=2)
    out["adjoint cycles"] = float(cyc[-1])
    for k, v in out.items():
        print(f"# step part {k}: {v:.2f}{'' if 'cycles' in k else ' ms'}", file=sys.stderr,
              flush=True)
    del fl
    return out


def timed(s, b, args, mode, steps, warmup):
    import torch
    for _ in range(warmup):
        solve(s, b, args, mode, args.rtol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cycles = 0
    f = hist = None
    for _ in range(steps):
        f, hist, _ = solve(s, b, args, mode, args.rtol)
        cycles += len(hist)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return dt, cycles, f, hist


def x_chains(p):
    """The problem's x-chains (blocks linked east-west), each independent of the others."""
    chains = []
    for i, blk in enumerate(p.blocks):
        if blk.bl2bl[0] == 0:
            c = [i]
            while p.blocks[c[-1]].bl2bl[1]:
                c.append(p.blocks[c[-1]].bl2bl[1] - 1)
            chains.append(c)
    return chains


def host_cores():
    """Cores this process may run on (the host's share given to it)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count()


def exact_path_rate(s, b, prob):
    """Seconds per V-cycle of the reference's own schedule (FP64, ``prob``'s V(numpre, numpost),
    lexicographic point SOR, ``numcits`` iterations on the coarsest level) on the GPU through
    the bitwise exact path: the cycle the CPU baseline times, on the whole workload."""
    import torch
    f = torch.zeros_like(b)
    res = torch.zeros_like(b)
    b = b.clone()                               # the exact path writes coarse-level sources
    kw = dict(tolf=0.0, rtol=0.0, max_cycles=1, smoother=0, precision=0, numpre=prob.numpre,
              numpost=prob.numpost, numcits=prob.numcits, omega=prob.omega,
              omega_c=prob.omega_c)
    s.lmg_solve(f, b, res, **kw)                # first call builds the natural-layout operator
    f.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.lmg_solve(f, b, res, **kw)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def cpu_baseline_all_chains(args):
    """The reference Fortran (oracle/_ref/ref_driver) on the whole workload: every x-chain in its
    own single-rank process, all at once, one core each.  The chains share no halo, so this is
    the reference's MPI decomposition by chain, without the global norm reductions.  Each
    process runs the reference's own coeff_p on the chain's grid (density 1, no face-Dirichlet
    outflow: the same sweeps per cycle as the product's operator) and a seeded source, FP64
    V(2,2) point SOR, ``--cpu-cycles`` cycles (1 by default here).  Throughput = all cells x
    cycles / the slowest process's solve time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from fortranutils_amd import problems as P
    import ref_io
    if not ref_io.available():
        raise RuntimeError("reference driver not built")
    p = P.config(args.config, args.scale)
    chains = x_chains(p)
    subs = [P.subproblem(p, c) for c in chains]
    rng = np.random.default_rng(11)
    b1s = []
    for q in subs:
        n1 = sum(int(np.prod(b.nijkp1)) for b in q.blocks)
        b1s.append(rng.standard_normal(n1))
    ncyc = args.cpu_cycles_all
    out = ref_io.run_timing_concurrent(subs, ncycles=ncyc, b1s=b1s)
    del b1s
    cells = [q.interior_cells() for q in subs]
    tmax = max(t for t, _ in out)
    big = max(cells) / sum(cells)
    sample = (f"{args.config} whole: {p.nblk} blocks, {sum(cells) / 1e6:.1f}M cells, as its "
              f"{len(chains)} independent x-chains in {len(chains)} concurrent single-rank "
              f"reference processes (largest chain {100 * big:.0f}% of the cells); the "
              f"reference's own coeff_p, FP64 V(2,2) point SOR, {ncyc} cycle(s)")
    return dict(value=sum(cells) * ncyc / tmax / 1e6, unit="Mcell*V-cycles/s", cores=len(chains),
                ranks=len(chains), host_cores=host_cores(), kind="reference", sample=sample,
                seconds=tmax,
                per_process=[dict(cells=c, seconds=t) for c, (t, _) in zip(cells, out)])


def cpu_baseline(args):
    """The reference Fortran (oracle/_ref/ref_driver, built from the reference sources) when it
    runs here, else the C restatement: FP64 V(2,2) with the reference's point-SOR schedule on the
    first x-chain of the same workload (2 blocks, full size), ``--cpu-cycles`` cycles, 1 core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from fortranutils_amd import problems as P
    p = P.config(args.config, args.scale)
    first = [0]
    while p.blocks[first[-1]].bl2bl[1]:
        first.append(p.blocks[first[-1]].bl2bl[1] - 1)
    q = P.subproblem(p, first)
    cells = q.interior_cells()
    import ref_io
    import oracle as O
    o = O.Oracle(q)
    geo = o.geometry()
    us, vs, ws = P.velocities(q)
    rho = P.density(q)
    fl = dict(u=o.level1(us), v=o.level1(vs), w=o.level1(ws), rho=o.level1(rho), rhonp1=o.level1(rho))
    dirich = np.array([b.dirichlet for b in q.blocks], np.int32)
    a = o.coeff(geo, rho=o.rho_levels(fl["rho"]), dirichlet=dirich)
    bb = o.sour(geo, fl["u"], fl["v"], fl["w"], fl["rho"], fl["rhonp1"], q.dtime)
    sample = (f"{args.config} first x-chain, blocks {[i + 1 for i in first]}, {cells / 1e6:.2f}M "
              f"cells, {o.nlev} levels, FP64 V(2,2) point-SOR (reference schedule), "
              f"{args.cpu_cycles} cycles, 1 core")
    if ref_io.available():
        try:
            out = ref_io.run(q, ncycles=args.cpu_cycles, coeffs=a, b1=bb[: o.n1], n1=o.n1,
                             do_post=0, timeout=600)
            return dict(value=cells * args.cpu_cycles / out["tsolve"] / 1e6,
                        unit="Mcell*V-cycles/s", cores=1, ranks=1, host_cores=host_cores(),
                        kind="reference", sample=sample, seconds=out["tsolve"])
        except Exception as e:  # the reference build may not run on this host
            print("reference CPU baseline unavailable:", str(e)[:200], file=sys.stderr)
    f = np.zeros(o.ntot)
    t0 = time.perf_counter()
    o.solve(a, f, bb, args.cpu_cycles)
    dt = time.perf_counter() - t0
    return dict(value=cells * args.cpu_cycles / dt / 1e6, unit="Mcell*V-cycles/s", cores=1,
                ranks=1, host_cores=host_cores(), kind="port", sample=sample, seconds=dt)


def seg_dispatch(p):
    """Whether the library runs level 1's segmented line solve with its short-segment bodies:
    some block's lines cut into segments of at most 14 rows (the fewest power-of-two segments
    of at most 16 rows), unless LMGX_SEG_FULL is set."""
    full = os.environ.get("LMGX_SEG_FULL", "")
    if full and int(full) != 0:
        return False
    for blk in p.blocks:
        nj = blk.nijkp1[1] - 2
        ns = 1
        while ns < 16 and (nj + ns - 1) // ns > 16:
            ns *= 2
        if (nj + ns - 1) // ns <= 14:
            return True
    return False


def traffic_from_profiles(workload, kernel):
    """HBM bytes per launch of ``kernel`` on ``workload`` from profiles/traffic.json (written by
    --traffic from rocprofv3 FETCH_SIZE / WRITE_SIZE runs); None when not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path)).get(workload)
    except (OSError, ValueError):
        return None
    if not d or kernel not in d:
        return None
    return d[kernel].get("bytes_per_launch")


def _counter_rows(path):
    import csv
    for r in csv.DictReader(open(path)):
        name = r.get("Kernel_Name") or r.get("Kernel-Name") or "?"
        yield name, float(r.get("Counter_Value") or r.get("Counter-Value") or 0.0)


def pmc_summary(paths):
    """Per-kernel dispatch count / total / average of a rocprofv3 counter_collection.csv."""
    import collections
    for path in paths:
        agg = collections.defaultdict(lambda: [0, 0.0])
        for name, v in _counter_rows(path):
            agg[name][0] += 1
            agg[name][1] += v
        print(f"# {path}")
        print("kernel,dispatches,total,avg_per_dispatch")
        for k, (n, t) in sorted(agg.items(), key=lambda x: -x[1][1]):
            print(f'"{k[:160]}",{n},{t:.1f},{t / n:.1f}')


def stats_summary(path, top=25):
    """The top kernels of a rocprofv3 --stats kernel_stats.csv as a markdown table."""
    import csv
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    print("| kernel | calls | total ms | avg us | min us | max us | % |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for r in rows[:top]:
        name = r["Name"].replace("void ", "").split("(")[0].replace("|", "/")
        print(f"| `{name[:90]}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
              f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
              f"{float(r['MaxNs']) / 1e3:.1f} | {100 * float(r['TotalDurationNs']) / tot:.1f} |")
    print(f"\n{len(rows)} kernels, {tot / 1e6:.1f} ms in total")


def stats_diff(pa, pb, n, top=30):
    """Kernel time per solve from two rocprofv3 --stats runs of the same workload that differ only
    in the number of timed solves (by n): setup, tuning and profiling cancel."""
    import csv

    def load(path):
        out = {}
        for r in csv.DictReader(open(path)):
            name = r["Name"].replace("void ", "").split("(")[0].replace("|", "/")
            c, t = out.get(name, (0, 0.0))
            out[name] = (c + int(r["Calls"]), t + float(r["TotalDurationNs"]))
        return out
    a, b = load(pa), load(pb)
    rows = []
    for k, (cb, tb) in b.items():
        ca, ta = a.get(k, (0, 0.0))
        dc, dt = (cb - ca) / n, (tb - ta) / n / 1e6
        if dc > 0 or dt > 1e-3:
            rows.append((dt, dc, k))
    rows.sort(reverse=True)
    tot = sum(r[0] for r in rows)
    print("| kernel | calls per solve | ms per solve | avg us | % |")
    print("|---|---:|---:|---:|---:|")
    for dt, dc, k in rows[:top]:
        print(f"| `{k[:90]}` | {dc:.1f} | {dt:.3f} | {1e3 * dt / max(dc, 1):.1f} | "
              f"{100 * dt / tot:.1f} |")
    print(f"\n{tot:.2f} ms of kernel time per solve")


def traffic(d):
    """profiles/traffic.json entry from four rocprofv3 --pmc runs under directory ``d``:
    copy_FETCH_SIZE/, copy_WRITE_SIZE/ (``--copy-calib``: 3 copies of 1 GiB) and
    <config>_<mode>_FETCH_SIZE/, _WRITE_SIZE/ (one solve).  Counters are in KiB per dispatch;
    each is scaled so the copy kernel reads / writes exactly 1 GiB."""
    import glob
    def vals(sub, pred):
        f = glob.glob(os.path.join(d, sub, "**", "*counter_collection.csv"), recursive=True)
        return [v for name, v in _counter_rows(f[0]) if pred(name)]

    def avg(sub, pred):
        v = vals(sub, pred)
        return (sum(v) / len(v) if v else float("nan")), len(v)
    # the three 1 GiB copies, however the runtime splits them into dispatches
    big = lambda name: "copy" in name.lower()
    cf = sum(vals("copy_FETCH_SIZE", big)) / 3
    cw = sum(vals("copy_WRITE_SIZE", big)) / 3
    kf_scale, kw_scale = (1 << 30) / (cf * 1024), (1 << 30) / (cw * 1024)
    out = {}
    for sub in sorted(glob.glob(os.path.join(d, "*_FETCH_SIZE"))):
        wl = os.path.basename(sub)[: -len("_FETCH_SIZE")]
        if wl == "copy":
            continue
        cfg, mode = wl.split("_", 1)
        word = "float" if mode == "mixed" else "double"
        fetch_rows = list(_counter_rows(glob.glob(os.path.join(sub, "**", "*counter_collection.csv"),
                                                  recursive=True)[0]))
        # every finest-level line-kernel instance (tile width, relaxation) separately, the
        # level-1 FP64 residual passes and the mixed-precision correction
        names = sorted({n.replace("void ", "").split("(")[0] for n, _ in fetch_rows
                        if n.replace("void ", "").startswith((f"k_zebra_line<{word}, 1,",
                                                              f"k_zebra_seg<{word}, 1,",
                                                              f"k_zebra_seg8<{word}, 1,",
                                                              "k_resid_cols<double, float, true",
                                                              "k_correct"))})
        ent = {}
        for kname in names:
            pick = lambda name, k=kname: name.replace("void ", "").split("(")[0] == k
            fetch, nf = avg(os.path.basename(sub), pick)
            write, nw = avg(wl + "_WRITE_SIZE", pick)
            ent[kname] = dict(dispatches=nf, fetch_kib_raw=fetch, write_kib_raw=write,
                              fetch_scale=kf_scale, write_scale=kw_scale,
                              bytes_per_launch=fetch * 1024 * kf_scale + write * 1024 * kw_scale)
        out[f"{cfg}:{mode}"] = ent
    print(json.dumps(out, indent=1))


def main():
    args = parse()
    if args.copy_calib:
        copy_calib()
        return
    if args.pmc_summary:
        pmc_summary(args.pmc_summary)
        return
    if args.stats_summary:
        stats_summary(args.stats_summary)
        return
    if args.stats_diff:
        stats_diff(args.stats_diff[0], args.stats_diff[1], int(args.stats_diff[2]))
        return
    if args.traffic:
        traffic(args.traffic)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is "
              "not built: every block of the problem is resident on one GPU", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "needs a GPU"
    t_setup = time.perf_counter()
    p, s, b = make_problem(args)
    n1cells = p.interior_cells()
    print(f"# setup {time.perf_counter() - t_setup:.1f}s: {args.config} {n1cells / 1e6:.1f}M cells, "
          f"{s.nlev} levels, {p.nblk} blocks", file=sys.stderr, flush=True)

    dt, cycles, f, hist = timed(s, b, args, args.mode, args.steps, args.warmup)
    ms_step = 1e3 * dt / args.steps
    value = n1cells * cycles / dt / 1e6
    red = hist[-1, 0, 0] / hist[0, 0, 0] if len(hist) else float("nan")
    print(f"# {args.mode}: {cycles / args.steps:.1f} cycles/solve, {ms_step:.1f} ms/solve, "
          f"rnorm last / first cycle {red:.2e}", file=sys.stderr, flush=True)

    extra = {}
    if not args.no_compare and args.mode != "fp64":
        dt64, cyc64, f64, h64 = timed(s, b, args, "fp64", 1, 0)
        err = (torch.linalg.norm(f[: s.n1] - f64[: s.n1]) / torch.linalg.norm(f64[: s.n1])).item()
        sb64 = solve_bytes(s, args, "fp64", cyc64)
        extra = dict(fp64_ms_per_solve=1e3 * dt64, fp64_cycles=cyc64,
                     fp64_solve_roofline_frac=sb64 / dt64 / PEAK_HBM,
                     speedup_vs_fp64=dt64 / (dt / args.steps),
                     p_rel_l2_vs_fp64=err)
        print(f"# fp64: {cyc64} cycles, {1e3 * dt64:.1f} ms; speed-up {extra['speedup_vs_fp64']:.2f}x,"
              f" |p_mixed - p_fp64|/|p_fp64| = {err:.2e}", file=sys.stderr, flush=True)

    if args.compare_per_block and args.mode != "exact":
        os.environ["LMGX_PER_BLOCK"] = "1"
        dtpb, cycpb, fpb, _ = timed(s, b, args, args.mode, args.steps, 1)
        os.environ["LMGX_PER_BLOCK"] = "0"
        same = bool(torch.equal(fpb, f))
        extra.update(per_block_ms_per_solve=1e3 * dtpb / args.steps,
                     batched_speedup_vs_per_block=dtpb / dt, per_block_same_pressure=same)
        print(f"# per-block launch loop: {1e3 * dtpb / args.steps:.1f} ms/solve, batched form "
              f"{dtpb / dt:.2f}x faster, same pressure bitwise: {same}", file=sys.stderr,
              flush=True)

    # dominant kernel: the finest-level line-relaxation half-sweep; HIP events on the solver's
    # stream.  Algorithmic bytes per updated cell, whatever the kernel: the right-hand side b,
    # 6 operator words (ap, aw, ae, al, ar and the line coupling as; an is the next row's as),
    # 1 word of the other colour's x, 1 word written (x) - 9 words; 10 when over-relaxed (the
    # old x is read).  With binary16 coefficients (mixed precision) the 6 operator words are 2
    # bytes each (ap replaced by the diagonal excess d): 28 bytes over-relaxed, 24 without.
    roof = None
    post0 = None
    if args.mode != "exact":
        _, hist_p, (ms, nl, cells) = solve(s, b, args, args.mode, args.rtol, profile=True)
        # level 1's plain half-sweeps: per cycle 2 numpre spi (less the 2 zero-start ones of the
        # segmented kernel) and 2 per post-smoothing sweep, lean or full
        nzs = 2 if args.mode == "mixed" and int(s.tile_widths()[1][0]) in (8, 16) else 0
        ncp = max(len(hist_p), 1)
        post0 = max(0.0, (nl - ncp * (2 * args.numpre * args.spi - nzs)) / (2.0 * ncp))
        word = 4 if args.mode == "mixed" else 8
        words = 9 if args.omega == 1.0 else 10
        tm = int(s.tile_widths()[1 if word == 4 else 0][0])
        h16 = word == 4 and tm in (8, 16) and bool(s.coef16_levels()[0])
        # algorithmic bytes of an SOR sweep, 10 words per updated cell (7 coefficients, b, x read
        # and written); the kernel moves fewer: binary16 couplings, an from the next row's as
        alg_per_cell = 10 * word
        moved_per_cell = words * word - (6 * word - 12 if h16 else 0)
        avg_s = ms / max(nl, 1) / 1e3
        cells_launch = cells / max(nl, 1)
        per_launch_bytes = cells_launch * alg_per_cell
        achieved = per_launch_bytes / avg_s
        moved_bytes = cells_launch * moved_per_cell
        rl = "true" if args.omega != 1.0 else "false"
        fp = "float" if word == 4 else "double"
        fz = "true" if s.fused_halos()[1 if word == 4 else 0][0] else "false"
        segk = {16: "k_zebra_seg", 8: "k_zebra_seg8"}.get(tm)
        # the finest level's instances: 1 where a half-sweep moves more than the 256 MiB
        # Infinity Cache (streamed past the caches), else 2 (the library's rule, LMGX_STREAM_MIB)
        lim = float(os.environ.get("LMGX_STREAM_MIB", "256")) * 1048576.0
        fine = 1 if 0.5 * cells_launch * 2 * (28 if word == 4 else 104) > lim else 2
        # (the segmented kernels' last arguments: level 1 in metric form, off by default; the
        # short-segment bodies, chosen per block, where some block's segments hold <= 14 rows)
        dsp = "true" if (h16 or word == 8) and rl == "true" and seg_dispatch(p) else "false"
        key = (f"{segk}<{fp}, {fine}, {rl}, 0, {'true' if h16 else 'false'}, false, {dsp}>"
               if segk else f"k_zebra_line<{fp}, {fine}, {tm}, {rl}, {fz}>")
        tr = traffic_from_profiles(f"{args.config}:{args.mode}", key)
        roof = dict(bound="hbm", achieved=achieved / 1e9, peak=PEAK_HBM / 1e9, unit="GB/s",
                    frac=achieved / PEAK_HBM, traffic=tr,
                    kernel=key, avg_launch_us=1e6 * avg_s, launches=nl,
                    bytes_per_launch=per_launch_bytes, bytes_per_cell=alg_per_cell,
                    moved_bytes_per_cell=moved_per_cell, moved_bytes_per_launch=moved_bytes,
                    achieved_moved=moved_bytes / avg_s / 1e9,
                    frac_moved=moved_bytes / avg_s / PEAK_HBM,
                    coef16=h16, gcell_updates_per_s=cells_launch / avg_s / 1e9,
                    cells_per_launch=cells_launch)
        print(f"# smoother: {nl} launches, avg {1e6 * avg_s:.1f} us, {achieved / 1e9:.0f} GB/s "
              f"at {alg_per_cell} algorithmic B per cell ({roof['frac']:.3f} of 8 TB/s), "
              f"{roof['achieved_moved']:.0f} GB/s at the {moved_per_cell} B it moves "
              f"({roof['frac_moved']:.3f}); tile widths fp64/fp32 {s.tile_widths().tolist()}",
              file=sys.stderr, flush=True)
        if not args.no_stream_ceiling:
            sc = stream_ceiling()
            roof["stream_ceiling"] = sc
            # the kernel's own mix: b, the other colour's x, the old x and the couplings read,
            # x written (6:1 with binary16 couplings, 9:1 without), with the non-temporal hint
            # where the instance streams past the caches (all but the other colour's x)
            mix = "r6w1" if h16 else "r9w1"
            roof["frac_of_stream"] = roof["achieved_moved"] / sc[mix + ("nt" if fine == 1 else "")
                                                                + "_GBps"]
            roof["frac_of_stream_plain"] = roof["achieved_moved"] / sc[mix + "_GBps"]
            roof["streamed_nontemporal"] = fine == 1
            print("# streaming on this GPU (reads:writes): " +
                  ", ".join(f"{k[1:-5].replace('w', ':').replace('nt', ' nt')} {v:.0f} GB/s"
                            for k, v in sc.items()) +
                  f"; smoother's moved bytes at {roof['frac_of_stream']:.2f} of its own mix's"
                  f" ({roof['frac_of_stream_plain']:.2f} of ordinary accesses')",
                  file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        if not args.cpu_one_chain:
            try:
                cpu = cpu_baseline_all_chains(args)
            except Exception as e:
                print("whole-workload CPU baseline unavailable:", str(e)[:300], file=sys.stderr)
        if cpu is None:
            try:
                cpu = cpu_baseline(args)
            except Exception as e:
                print("cpu baseline failed:", e, file=sys.stderr)
        if cpu:
            print(f"# CPU baseline ({cpu['kind']}, {cpu['cores']} of {cpu['host_cores']} host "
                  f"cores): {cpu['value']:.2f} {cpu['unit']}", file=sys.stderr, flush=True)
            # the same cycle (the reference's V(2,2) point-SOR schedule) on the GPU, bitwise the
            # reference's arithmetic: the like-for-like rate next to the CPU's
            dt_ex = exact_path_rate(s, b, p)
            cpu["gpu_same_cycle"] = dict(value=n1cells / dt_ex / 1e6, unit="Mcell*V-cycles/s",
                                         seconds_per_cycle=dt_ex,
                                         ratio_to_cpu=n1cells / dt_ex / 1e6 / cpu["value"])
            print(f"# the reference's own cycle on the GPU (exact path): {n1cells / dt_ex / 1e6:.1f}"
                  f" Mcell*V-cycles/s, {cpu['gpu_same_cycle']['ratio_to_cpu']:.0f}x the CPU's",
                  file=sys.stderr, flush=True)

    # the whole solve against the HBM roofline: algorithmic bytes of the schedule / time
    sb = (solve_bytes(s, args, args.mode, cycles / args.steps, post0) if args.mode != "exact"
          else None)
    solve_roof = None
    if sb:
        ach = sb / (dt / args.steps)
        solve_roof = dict(bytes_per_solve=sb, bytes_per_cell=sb / n1cells, achieved=ach / 1e9,
                          unit="GB/s", frac=ach / PEAK_HBM)
        print(f"# whole solve: {sb / n1cells:.0f} algorithmic B per fine cell, {ach / 1e9:.0f} GB/s "
              f"= {ach / PEAK_HBM:.2f} of 8 TB/s", file=sys.stderr, flush=True)
    out = {
        "metric": "Mcell\u00b7V-cycles/s + time to 1e-10 residual, mixed-prec vs FP64, 1 GPU",
        "value": value,
        "unit": "Mcell*V-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if args.mode == "mixed" else "fp64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {p.nblk} blocks, {n1cells / 1e6:.1f}M cells, "
                               f"{s.nlev} levels, solve to {args.rtol:g} residual reduction",
                   "mode": args.mode, "smoother": "zebra y-line relaxation" if args.mode != "exact" else
                   "reference lexicographic SOR", "numpre": args.numpre, "numpost": args.numpost,
                   "sweeps_per_iter": args.spi, "fine_post": args.fine_post, "omega": args.omega,
                   "numcits": args.numcits},
        "time_to_rtol_ms": ms_step,
        "cycles_per_solve": cycles / args.steps,
        "roofline": roof,
        "solve_roofline": solve_roof,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    if args.step_parts:
        out["step_parts_ms"] = step_parts(p, s, args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
