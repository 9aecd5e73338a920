// Shared device/host structures of liblmgx (MI355X / gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LMGX_MAXB 64
#define LMGX_MAXLEV 8
#define LMGX_WG 256
#define LMGX_RMID 256   // middle-stage workgroups of a long norm reduction
#ifndef LMGX_RCH
#define LMGX_RCH 1      // coarse rows per thread of the residual column kernel
#endif
#define LMGX_RCHP 8     // coarse rows per thread of the prolongation column kernel
#ifndef LMGX_CONV_NT
#define LMGX_CONV_NT 1
#endif
#define LMGX_SMAX 16    // segmented line solve: rows per segment held in registers, 16 segments
                        // at most, so lines up to 256 rows

#ifndef LMGX_SEGLA
#define LMGX_SEGLA 6    // segmented line relaxation on binary16 couplings: rows of loads issued ahead
#endif
#ifndef LMGX_MFLA
#define LMGX_MFLA 3     // metric-form line relaxation: rows of loads issued ahead
#endif

#define BCT_PERIODIC 12
#define BCT_INTERFACE 13

// One block on one level, natural (reference) layout.
struct LevBlk {
    long off;        // mg_off: start of the block in the multi-level array
    long off_s;      // start of the block in each colour array of the split layout
    int n[3];        // nip1, njp1, nkp1 (halos included)
    int hp;          // row pitch of the split layout: ceil(nip1/2) slots rounded up to 32
    int ifc[6];      // side is an INTERFACE (correction range covers the halo)
    int swp[6];      // side is INTERFACE or PERIODIC (halo swapped)
    int dir[6];      // side carries a face-Dirichlet p = 0 (fast-path prolongation: odd extension)
    long t_int;      // first tile of this block in the interior (i,j) tiling
    long t_all;      // first tile of this block in the full (i,j) tiling (halos included)
    long t_rb;       // first tile of this block in the split-layout (slot,j) tiling
    long t_ln;       // first tile of this block in the split-layout (slot,k) line tiling 64x4
    long t_z64;      // line-solve tiling: 64 slots x 2 k-planes, two wavefronts per line
    long t_z128;     // line-solve tiling: 128 slots x 1 k-plane, two wavefronts per line
    long t_rr;       // residual column tiling: 64 coarse i x 4 coarse k x LMGX_RCH coarse j
    long t_rp;       // prolongation column tiling: 64 coarse i x 4 coarse k x LMGX_RCHP coarse j
    long t_sg;       // segmented line-solve tiling: 64 slots x sg_s k-planes per 16 wavefronts
    int sg_n;        // segmented line solve: segments per line (power of two <= 16)
    int sg_r;        // rows per segment (<= LMGX_SMAX)
    int sg_s;        // line sets (k-planes) per workgroup = 16 / sg_n
    long t_sg8;      // the same tiling for 8-wavefront workgroups (valid where sg_n <= 8)
    int sg_s8;       // line sets per 8-wavefront workgroup = 8 / sg_n (0 when sg_n = 16)
    int sg_v;        // rows of the segmented body instance that runs the block: sg_r rounded up
                     // to 10, 12, 14 or 16 (16 with LMGX_SEG_FULL=1)
    long s_jk;       // first (j,k) interior line of this block (hyperplane sweep)
    long o2;         // offset of the block's 2-D (i,j) metric arrays
    long o1;         // offset of the block's 1-D k metric arrays
    long oxy;        // offset of the block's vertex array (nip1+1)*(njp1+1)
    long oz;         // offset of the block's k-vertex array (nkp1+1)
    long o2s;        // offset of the block's parity-split 2-D tables (njp1 rows of hp, per parity)
    // Halo stores made by the segmented line solve itself.  Side q = 0..3 is the interior plane
    // i = 1, i = nip1-2, k = 1, k = nkp1-2; when it is the source of a partner's face-interior
    // halo, fz_off[q] >= 0 is that block's off_s and the halo point of source point (I,J,K) is
    // (I,J,K) + fz_d[q] in a block of row pitch fz_hp[q] and fz_n1[q] rows; -1 = no store.
    long fz_off[4];
    int fz_hp[4], fz_n1[4];
    int fz_d[4][3];
};

// One swapped face of one block on one level (reference swap_sc0 tables).
struct Face {
    long pack_base;    // first element of the packed plane (plane 2 or n-1), natural layout
    long unpack_base;  // first element of the halo plane (plane 1 or n)
    long jo, ko;       // strides along the two in-plane axes (cyclic order)
    int nj, nk;        // face extents (halos included)
    long buf;          // start of this face's packed data in the exchange buffer
    long src;          // start of the partner face's packed data (what lands in our halo)
    long e0;           // first element index of this face in the flattened element list
    int skip;          // bit0: j==1 overwritten later, bit1: j==nj, bit2: k==1, bit3: k==nk
    int blk;           // block (for the split layout)
    int pack_i[3];     // natural (i,j,k)-0 based of the pack plane origin
    int unpack_i[3];   // natural origin of the halo plane
    int ax_j, ax_k;    // 0-based axes of the in-plane directions
};

__host__ __device__ inline int xcd_remap(int b, int G)
{
    // Workgroups are dealt round-robin over the 8 XCDs; give each XCD a contiguous tile range so
    // that neighbouring tiles (which share halo lines) meet in the same L2.
    const int nx = 8;
    int q = G / nx, r = G % nx;
    int x = b % nx, i = b / nx;
    return x * q + (x < r ? x : r) + i;
}

// Workgroup i runs on XCD i % 8.  Give each XCD runs of G consecutive tiles, dealt round-robin:
// neighbouring k-planes of one block (which share the other colour's x) meet in one L2, and
// every XCD still gets an even share of every block.  The tail that does not fill 8 runs keeps
// the identity mapping.
__device__ inline long xcd_chunk(long i, long n, int G)
{
    const long span = 8L * G;
    if (i >= n / span * span) return i;
    const long s = i >> 3;
    return (s / G) * span + (i & 7) * G + (s % G);
}

// Find the block owning flattened tile t (start[] ascending, nblk <= 64).
__device__ inline int find_block(const LevBlk *lb, int nblk, long t, int which)
{
    int lo = 0, hi = nblk - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        long s = which == 0 ? lb[mid].t_int : which == 1 ? lb[mid].t_all : which == 2 ? lb[mid].t_rb
                 : which == 3 ? lb[mid].s_jk : which == 4 ? lb[mid].t_ln
                 : which == 5 ? lb[mid].t_z64 : which == 6 ? lb[mid].t_z128
                 : which == 7 ? lb[mid].t_rr : which == 8 ? lb[mid].t_rp
                 : which == 9 ? lb[mid].t_sg : lb[mid].t_sg8;
        if (s <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}
