"""Synthetic pressure-Poisson workloads in LESter's multiblock layout.

LESter reads its grid from ``x.bin``/``y.bin``/``z.bin`` and its block connectivity from
``block.dat`` (reference ``code/read_xyz.F:67-141``, ``code/read_bl_info.F:79-130``); it has no
grid generator and ships no input decks.  This module builds the same information in memory:

* per block ``nijkp1`` (interior + one halo layer per side), six boundary codes, and the
  ``bl2bl``/``bl2side`` links (``code/cb.lgrid.inc:13-24``);
* level-1 vertex coordinates of every block (interior vertices ``i=2..nip1``, ``j=2..njp1``,
  ``k=2..nkp1`` as in ``x.bin``);
* cell densities (halos included) and face velocities for the source ``sour_p_block``;
* which physical sides carry a face-Dirichlet ``p=0`` (an extension: the reference treats every
  physical side as homogeneous Neumann, ``code/coeff_p_block.F:160-228``).

Block and side numbering is the reference's: blocks are 1-based, sides 1..6 are
west, east, south, north, left, right (``code/coeff_p_block.F:143-226``).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

# Boundary-condition codes, reference code/bc_type.inc:17-27.
BCT_WALL = 1
BCT_ZERO_GRAD = 2
BCT_CONVECTIVE = 3
BCT_FIXED = 4
BCT_PERIODIC = 12
BCT_INTERFACE = 13


@dataclass
class Block:
    nijkp1: tuple                   # (nip1, njp1, nkp1) incl. halos
    bc: list                        # 6 codes
    bl2bl: list                     # 6 partner blocks (1-based, 0 = none)
    bl2side: list                   # 6 partner sides
    xv: np.ndarray                  # vertex x, i = 2..nip1  (len nip1-1)
    yv: np.ndarray                  # vertex y, j = 2..njp1  (len njp1-1)
    zv: np.ndarray                  # vertex z, k = 2..nkp1  (len nkp1-1)
    dirichlet: list = field(default_factory=lambda: [0] * 6)


@dataclass
class Problem:
    name: str
    blocks: List[Block]
    nummgl: int = 0                 # 0 = as many levels as the dims allow
    rho_kind: str = "uniform"       # uniform | mixing | wall
    rho_seed: int = 0
    src_seed: int = 1234
    blobs: int = 0                  # Gaussian blobs added to the source
    dtime: float = 2.0 / 3.0
    numpre: int = 2
    numpost: int = 2
    numcits: int = 20
    omega: float = 1.6
    omega_c: float = 1.6

    @property
    def nblk(self) -> int:
        return len(self.blocks)

    def interior_cells(self) -> int:
        return int(sum((b.nijkp1[0] - 2) * (b.nijkp1[1] - 2) * (b.nijkp1[2] - 2)
                       for b in self.blocks))


def tanh_vertices(n_cells: int, lo: float, hi: float, beta: float = 2.0) -> np.ndarray:
    """``n_cells+1`` vertices on [lo, hi] clustered at both ends: y(eta)=tanh(b(2eta-1))/tanh(b)."""
    eta = np.linspace(0.0, 1.0, n_cells + 1)
    s = np.tanh(beta * (2.0 * eta - 1.0)) / np.tanh(beta)
    return lo + (hi - lo) * 0.5 * (s + 1.0)


def wall_vertices(n_cells: int, lo: float, hi: float, beta: float = 2.0) -> np.ndarray:
    """One-sided tanh clustering towards ``lo`` (a wall at y=lo)."""
    eta = np.linspace(0.0, 1.0, n_cells + 1)
    s = 1.0 + np.tanh(beta * (eta - 1.0)) / np.tanh(beta)
    return lo + (hi - lo) * s


def _chain_x(dims: Sequence[tuple], x0: float, x1: float, yv: np.ndarray, zv: np.ndarray,
             first_blk: int, outflow_dirichlet: bool = True) -> List[Block]:
    """Blocks joined along x by face-matched interfaces (east of b -> west of b+1)."""
    ntot = sum(d[0] for d in dims)
    xs = np.linspace(x0, x1, ntot + 1)
    blocks = []
    i0 = 0
    nb = len(dims)
    for ib, (nx, ny, nz) in enumerate(dims):
        gb = first_blk + ib
        bc = [BCT_FIXED, BCT_CONVECTIVE, BCT_WALL, BCT_WALL, BCT_PERIODIC, BCT_PERIODIC]
        l2b = [0, 0, 0, 0, gb, gb]
        l2s = [0, 0, 0, 0, 6, 5]
        dirich = [0] * 6
        if ib > 0:
            bc[0] = BCT_INTERFACE
            l2b[0], l2s[0] = gb - 1, 2
        if ib < nb - 1:
            bc[1] = BCT_INTERFACE
            l2b[1], l2s[1] = gb + 1, 1
        elif outflow_dirichlet:
            dirich[1] = 1
        blocks.append(Block((nx + 2, ny + 2, nz + 2), bc, l2b, l2s,
                            xs[i0:i0 + nx + 1].copy(), yv.copy(), zv.copy(), dirich))
        i0 += nx
    return blocks


def config(name: str, scale: int = 1) -> Problem:
    """Named workloads.  ``scale`` divides every cell count (parity tests at reduced size).

    c0  1 block 64x64x32, tanh-stretched y, rho = 1, 4 levels, V(2,2)
    c1  4 uneven blocks (192/128/96/64 x 96/96/48/48 x 64) in two x-chains, mixing layer rho
    c2  1 block 512x256x256 on the 8x2x4 box, mixing layer, 5 source blobs, 6 levels
    c3  16 uneven blocks (~120M cells), nz = 128, wall-layer rho with lognormal noise, 6 levels
    c4  2 blocks 128x96x64 + 96x96x64 (one x-interface), mixing layer, 4 levels
    c0 keeps its 8x2x4 box; c1, c3, c4 use dx = dz (z in [0, 4]) with x extents from the cell
    counts, so cells are isotropic in x-z and the tanh-stretched y is the stiff direction.
    """
    s = scale
    if name == "c0":
        nx, ny, nz = 64 // s, 64 // s, 32 // s
        yv = tanh_vertices(ny, -1.0, 1.0)
        zv = np.linspace(0.0, 4.0, nz + 1)
        blocks = _chain_x([(nx, ny, nz)], 0.0, 8.0, yv, zv, 1)
        return Problem("c0", blocks, nummgl=4 if s == 1 else 0, rho_kind="uniform")
    if name == "c1":
        zv = np.linspace(0.0, 4.0, 64 // s + 1)
        ya = tanh_vertices(96 // s, -1.0, 1.0)
        yb = tanh_vertices(48 // s, -1.0, 1.0)
        h = 4.0 / 64  # dx = dz: x-z isotropic cells, the wall-normal y is the stiff direction
        blocks = _chain_x([(192 // s, 96 // s, 64 // s), (128 // s, 96 // s, 64 // s)],
                          0.0, 320 * h, ya, zv, 1)
        blocks += _chain_x([(96 // s, 48 // s, 64 // s), (64 // s, 48 // s, 64 // s)],
                           0.0, 160 * h, yb, zv, 3)
        return Problem("c1", blocks, nummgl=4 if s == 1 else 0, rho_kind="mixing")
    if name == "c2":
        nx, ny, nz = 512 // s, 256 // s, 256 // s
        yv = tanh_vertices(ny, -1.0, 1.0)
        zv = np.linspace(0.0, 4.0, nz + 1)
        blocks = _chain_x([(nx, ny, nz)], 0.0, 8.0, yv, zv, 1)
        return Problem("c2", blocks, nummgl=6 if s == 1 else 0, rho_kind="mixing", blobs=5)
    if name == "c3":
        return _config3(s)
    if name == "c4":
        zv = np.linspace(0.0, 4.0, 64 // s + 1)
        yv = tanh_vertices(96 // s, -1.0, 1.0)
        h = 4.0 / 64
        blocks = _chain_x([(128 // s, 96 // s, 64 // s), (96 // s, 96 // s, 64 // s)],
                          0.0, 224 * h, yv, zv, 1)
        return Problem("c4", blocks, nummgl=4 if s == 1 else 0, rho_kind="mixing")
    raise ValueError(f"unknown config {name!r}")


def c3_block_dims(seed: int = 2024, target: float = 120e6, tol: float = 0.05):
    """16 block sizes, nx in 128..640 and ny in 64..256 (steps of 32), nz = 128, redrawn until the
    total is within ``tol`` of ``target``.  Blocks sharing (ny, nz) are later chained along x."""
    rng = np.random.default_rng(seed)
    nxs = np.arange(128, 641, 32)
    nys = np.arange(64, 257, 32)
    while True:
        nx = rng.choice(nxs, 16)
        ny = rng.choice(nys, 16)
        tot = float(np.sum(nx * ny * 128))
        if abs(tot - target) <= tol * target:
            return [(int(a), int(b), 128) for a, b in zip(nx, ny)]


def _config3(s: int) -> Problem:
    dims = c3_block_dims()
    # group equal-ny blocks into x-chains (the only face-matched neighbours, nz is common)
    groups = {}
    for d in dims:
        groups.setdefault(d[1], []).append(d)
    blocks: List[Block] = []
    zv = np.linspace(0.0, 4.0, 128 // s + 1)
    for ny in sorted(groups):
        chain = [(nx // s, ny // s, nz // s) for (nx, ny_, nz) in groups[ny]]
        yv = wall_vertices(ny // s, 0.0, 2.0)
        h = 4.0 / 128  # dx = dz
        blocks += _chain_x(chain, 0.0, h * sum(c[0] for c in groups[ny]), yv, zv, len(blocks) + 1)
    return Problem("c3", blocks, nummgl=6 if s == 1 else 0, rho_kind="wall", blobs=5)


# --------------------------------------------------------------------------------------------
# Fields on level 1 (numpy, per block arrays shaped [nkp1, njp1, nip1], i fastest)
# --------------------------------------------------------------------------------------------

def _ghosted(v: np.ndarray) -> np.ndarray:
    """Vertex line with one linearly extrapolated ghost at each end (code/xyz_extrap0.F:30-56)."""
    return np.concatenate([[2 * v[0] - v[1]], v, [2 * v[-1] - v[-2]]])


def cell_centres(b: Block):
    """Cell-centre coordinates of every cell incl. halos (halo centres from ghost vertices)."""
    xg, yg, zg = _ghosted(b.xv), _ghosted(b.yv), _ghosted(b.zv)
    return 0.5 * (xg[1:] + xg[:-1]), 0.5 * (yg[1:] + yg[:-1]), 0.5 * (zg[1:] + zg[:-1])


def density(p: Problem) -> List[np.ndarray]:
    rng = np.random.default_rng(p.rho_seed + 77)
    out = []
    for b in p.blocks:
        nip1, njp1, nkp1 = b.nijkp1
        _, yc, _ = cell_centres(b)
        if p.rho_kind == "uniform":
            r = np.ones((nkp1, njp1, nip1))
        elif p.rho_kind == "mixing":
            ry = 1.0 + 3.0 * (1.0 + np.tanh(yc / 0.1))
            r = np.broadcast_to(ry[None, :, None], (nkp1, njp1, nip1)).copy()
        elif p.rho_kind == "wall":
            ry = 1.0 + 2.0 * np.exp(-np.maximum(yc, 0.0) / 0.05)
            r = ry[None, :, None] * rng.lognormal(0.0, 0.05, (nkp1, njp1, nip1))
        else:
            raise ValueError(p.rho_kind)
        out.append(np.ascontiguousarray(r))
    return out


def velocities(p: Problem):
    """Seeded i.i.d. N(0,1) face velocities u, v, w per block (halo faces included)."""
    rng = np.random.default_rng(p.src_seed)
    us, vs, ws = [], [], []
    for b in p.blocks:
        nip1, njp1, nkp1 = b.nijkp1
        us.append(rng.standard_normal((nkp1, njp1, nip1)))
        vs.append(rng.standard_normal((nkp1, njp1, nip1)))
        ws.append(rng.standard_normal((nkp1, njp1, nip1)))
    return us, vs, ws


def blob_source(p: Problem, amp: float = 10.0, width: float = 0.05) -> List[np.ndarray]:
    """Sum of ``p.blobs`` Gaussian blobs (seeded centres) times the cell volume, level-1 layout."""
    out = []
    if p.blobs == 0:
        return [np.zeros(tuple(reversed(b.nijkp1))) for b in p.blocks]
    rng = np.random.default_rng(p.src_seed + 5)
    xs = np.concatenate([b.xv for b in p.blocks])
    ys = np.concatenate([b.yv for b in p.blocks])
    zs = np.concatenate([b.zv for b in p.blocks])
    cen = np.stack([rng.uniform(xs.min(), xs.max(), p.blobs),
                    rng.uniform(ys.min(), ys.max(), p.blobs),
                    rng.uniform(zs.min(), zs.max(), p.blobs)], 1)
    for b in p.blocks:
        xc, yc, zc = cell_centres(b)
        dx = np.diff(_ghosted(b.xv))
        dy = np.diff(_ghosted(b.yv))
        dz = np.diff(_ghosted(b.zv))
        vol = dz[:, None, None] * dy[None, :, None] * dx[None, None, :]
        g = np.zeros(vol.shape)
        for c in cen:
            r2 = ((xc[None, None, :] - c[0]) ** 2 + (yc[None, :, None] - c[1]) ** 2
                  + (zc[:, None, None] - c[2]) ** 2)
            g += amp * np.exp(-r2 / (2.0 * width * width))
        out.append(g * vol)
    return out


def subproblem(p: Problem, blocks: Sequence[int], name: Optional[str] = None) -> Problem:
    """The given 0-based blocks (closed under their links) as a problem of their own."""
    import copy
    keep = list(blocks)
    remap = {b + 1: i + 1 for i, b in enumerate(keep)}
    out = []
    for b in keep:
        blk = copy.deepcopy(p.blocks[b])
        for s in range(6):
            if blk.bl2bl[s]:
                if blk.bl2bl[s] not in remap:
                    raise ValueError("sub-problem is not closed under its links")
                blk.bl2bl[s] = remap[blk.bl2bl[s]]
        out.append(blk)
    q = copy.copy(p)
    q.blocks = out
    q.name = name or (p.name + "_sub")
    return q


def device_fields(p: Problem, device="cuda", seed_offset: int = 0):
    """Level-1 fields (concatenated blocks, [k,j,i] per block) generated on the GPU with torch:
    u, v, w ~ N(0,1) (seeded), rho per ``p.rho_kind`` (lognormal noise seeded), rhonp1 = rho."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(p.src_seed + seed_offset)
    n1 = sum(int(np.prod(b.nijkp1)) for b in p.blocks)
    u = torch.randn(n1, generator=g, device=device, dtype=torch.float64)
    v = torch.randn(n1, generator=g, device=device, dtype=torch.float64)
    w = torch.randn(n1, generator=g, device=device, dtype=torch.float64)
    rho = torch.empty(n1, device=device, dtype=torch.float64)
    o = 0
    for b in p.blocks:
        nip1, njp1, nkp1 = b.nijkp1
        n = nip1 * njp1 * nkp1
        _, yc, _ = cell_centres(b)
        if p.rho_kind == "uniform":
            ry = np.ones_like(yc)
        elif p.rho_kind == "mixing":
            ry = 1.0 + 3.0 * (1.0 + np.tanh(yc / 0.1))
        else:
            ry = 1.0 + 2.0 * np.exp(-np.maximum(yc, 0.0) / 0.05)
        r = torch.from_numpy(ry).to(device)[None, :, None].expand(nkp1, njp1, nip1).reshape(-1)
        if p.rho_kind == "wall":
            noise = torch.exp(0.05 * torch.randn(n, generator=g, device=device, dtype=torch.float64))
            r = r * noise
        rho[o:o + n] = r
        o += n
    return dict(u=u, v=v, w=w, rho=rho, rhonp1=rho)


def device_blobs(p: Problem, b1, device="cuda", amp: float = 10.0, width: float = 0.05):
    """Add ``p.blobs`` Gaussian blobs (seeded centres) x cell volume to the level-1 source b1."""
    import torch
    if p.blobs == 0:
        return
    rng = np.random.default_rng(p.src_seed + 5)
    xs = np.concatenate([b.xv for b in p.blocks])
    ys = np.concatenate([b.yv for b in p.blocks])
    zs = np.concatenate([b.zv for b in p.blocks])
    cen = np.stack([rng.uniform(xs.min(), xs.max(), p.blobs),
                    rng.uniform(ys.min(), ys.max(), p.blobs),
                    rng.uniform(zs.min(), zs.max(), p.blobs)], 1)
    o = 0
    for b in p.blocks:
        nip1, njp1, nkp1 = b.nijkp1
        n = nip1 * njp1 * nkp1
        xc, yc, zc = [torch.from_numpy(a).to(device) for a in cell_centres(b)]
        dx = torch.from_numpy(np.diff(_ghosted(b.xv))).to(device)
        dy = torch.from_numpy(np.diff(_ghosted(b.yv))).to(device)
        dz = torch.from_numpy(np.diff(_ghosted(b.zv))).to(device)
        vol = dz[:, None, None] * dy[None, :, None] * dx[None, None, :]
        gsum = torch.zeros((nkp1, njp1, nip1), device=device, dtype=torch.float64)
        for c in cen:
            r2 = ((xc[None, None, :] - c[0]) ** 2 + (yc[None, :, None] - c[1]) ** 2
                  + (zc[:, None, None] - c[2]) ** 2)
            gsum += amp * torch.exp(-r2 / (2.0 * width * width))
        blk = b1[o:o + n].view(nkp1, njp1, nip1)
        blk[1:-1, 1:-1, 1:-1] += (gsum * vol)[1:-1, 1:-1, 1:-1]
        o += n


def interior_mask_torch(p: Problem, device="cuda"):
    import torch
    ms = []
    for b in p.blocks:
        nip1, njp1, nkp1 = b.nijkp1
        m = torch.zeros((nkp1, njp1, nip1), dtype=torch.bool, device=device)
        m[1:-1, 1:-1, 1:-1] = True
        ms.append(m.reshape(-1))
    return torch.cat(ms)
