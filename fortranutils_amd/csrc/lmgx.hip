// liblmgx: MI355X-native pressure-Poisson multigrid for LESter's multiblock layout.
// One HIP stream per handle; every block of a level is processed by one launch driven by the
// device block tables (LevBlk); no host round trip inside a V-cycle except the per-cycle norm
// read of the reference-order path.
#include "lmgx_common.h"
#include "../../include/lmgx.h"

#include <algorithm>
#include <charconv>
#include <mutex>
#include <thread>
#include <type_traits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kernels_exact.inc"
#include "kernels_fast.inc"

// ---------------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------------
static thread_local std::string g_err;

extern "C" const char *lmgx_last_error(void) { return g_err.c_str(); }

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(LMGX_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

#define CHK(expr)                                                                                 \
    do {                                                                                          \
        int rc_ = (expr);                                                                         \
        if (rc_ != 0) return rc_;                                                                 \
    } while (0)

static inline unsigned nblocks(long n) { return (unsigned)((n + LMGX_WG - 1) / LMGX_WG); }

// ---------------------------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------------------------
struct LevelDev {
    std::vector<LevBlk> h_lb;
    LevBlk *d_lb = nullptr;
    // per-block launches (the fast path's block-loop mode): d_lb1[b] is block b alone, its tile
    // offsets rebased to 0; bt_*[b] its tile counts in each tiling
    LevBlk *d_lb1 = nullptr;
    std::vector<long> bt_sg, bt_sg8, bt_z64, bt_z128, bt_rr, bt_rp;
    long tiles_int = 0, tiles_all = 0, tiles_rb = 0, tiles_ln = 0, nlines = 0;
    long tiles_z64 = 0, tiles_z128 = 0, tiles_rr = 0, tiles_rp = 0, tiles_sg = 0, tiles_sg8 = 0;
    long n2s = 0;                       // slots per parity of the parity-split 2-D tables
    int njmax = 0;                      // longest j-line (interior rows) of the level
    int smax = 0;                       // largest hyperplane index i+j+k (1-based)
    std::vector<Face> h_fc;
    Face *d_fc = nullptr;
    int nfaces = 0;
    long nel = 0;                       // swapped elements (all faces)
    long npts_int = 0;                  // interior cells
    long nsplit = 0;                    // slots per colour array in the split layout
};

struct FastBufs;  // split-layout storage of the fast path (kernels_fast.inc)

struct lmgx_handle {
    int nblk = 0, nlev = 0;
    // the operator came from the builder (lmgx_build_operator / lmgx_update_density_device), and
    // whether the adjoint has swapped in its transpose for the moment
    bool op_builder = false, op_flipped = false;
    int nijkp1[LMGX_MAXLEV][LMGX_MAXB][3];
    long off[LMGX_MAXLEV][LMGX_MAXB];
    double npts[LMGX_MAXLEV];
    int bc[LMGX_MAXB][6], bl2bl[LMGX_MAXB][6], bl2side[LMGX_MAXB][6];
    int dirich[LMGX_MAXB][6];          // face-Dirichlet sides (operator extension)
    long ntot = 0, n1 = 0, n2d = 0, n1d = 0, nxy = 0, nz = 0;
    long off2d[LMGX_MAXLEV][LMGX_MAXB], off1d[LMGX_MAXLEV][LMGX_MAXB];
    long offxy[LMGX_MAXLEV][LMGX_MAXB], offz[LMGX_MAXLEV][LMGX_MAXB];
    LevelDev lev[LMGX_MAXLEV];
    hipStream_t stream = nullptr;
    double *d_a[7] = {nullptr};         // operator, natural layout, FP64
    double *d_aT[7] = {nullptr};        // transposed operator (adjoint solves), built on demand
    bool aT_valid = false;
    double *d_rho_op = nullptr;         // densities the operator was built from (all levels)
    double *d_f = nullptr, *d_b = nullptr, *d_res = nullptr;  // host-API staging
    double *d_swap = nullptr;           // exchange buffer (FP64 elements)
    long swap_cap = 0;
    double *d_partial = nullptr;
    long partial_cap = 0;
    double *d_partial2 = nullptr;       // middle-stage partials of long reductions
    double *d_scalar = nullptr;         // 64 doubles of small device scratch
    double *d_hist = nullptr;           // [cap][nlev][4]
    long hist_cap = 0;
    double *h_pinned = nullptr;         // pinned mirror of the history
    long pinned_cap = 0;
    int jumble = 0;
    bool have_op = false;
    // d_a (natural layout) holds the current operator; false after a density update that wrote
    // only the fast path's split copy (ensure_natural rebuilds it on demand)
    bool nat_valid = true;
    // geometry (lmgx_build_operator)
    double *d_met = nullptr;            // [7][n2d]: diu djv rdiv rdju didjp rdidjp rp
    double *d_dkp = nullptr, *d_dkw = nullptr;
    // level 1's metric form (see MetF): packed parity-split face factors [2 n2s] times mf_gsc
    // (0: not available), dkp(k) and 1 / dkw(k) [2][n1d]; written with the metrics
    uint2 *d_mft = nullptr;
    float *d_mfk = nullptr;
    double mf_gsc = 0.0;
    // the same factors in FP64 (see MetD), exact quotients as the builder forms them [2 n2s]
    double4 *d_mfd = nullptr;
    FastBufs *fast = nullptr;
    // profiling of the smoother
    bool prof = false;
    double prof_ms = 0.0;
    long prof_launch = 0;
    double prof_cells = 0.0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev;
    std::vector<double> prof_ev_cells;
    size_t prof_used = 0;
    // work counters (lmgx_counters): operator uploads, fast-path operator preparations, cycle
    // graph captures, solves
    long n_uploads = 0, n_prepares = 0, n_captures = 0, n_solves = 0;
    // line-kernel choices given by lmgx_set_tuning: [precision][kernel, halo stores][level]
    int tune_req[2][2][LMGX_MAXLEV] = {};
    bool tune_have[2] = {false, false};
    // device address of the static scratch of the allocation-free entry point on this handle's
    // GPU (looked up once per GPU by lmgx_init)
    double *scratch = nullptr;
    // lmgx_set_operator_cached: the host arrays of its last upload and their sampled fingerprint
    const double *cached_src[7] = {nullptr};
    unsigned long long cached_fp = 0;
    bool cached_valid = false;
};

static const int CYC[2][3] = {{2, 3, 1}, {3, 1, 2}};

static bool is_swap(int bc) { return bc == BCT_INTERFACE || bc == BCT_PERIODIC; }

// ---------------------------------------------------------------------------------------------
// tables: level count and dims (reference code/set_offset.F:94-140), offsets (:166-234),
// interior counts (code/set_loop_bound.F:87-104), face tables (code/set_interfaces.F:127-188)
// ---------------------------------------------------------------------------------------------
static int build_tables(lmgx_handle *h, const lmgx_topology *t)
{
    if (!t || t->nblk < 1 || t->nblk > LMGX_MAXB || !t->nijkp1 || !t->bc || !t->bl2bl ||
        !t->bl2side)
        return fail(LMGX_E_ARG, "bad topology");
    h->nblk = t->nblk;
    for (int b = 0; b < h->nblk; ++b) {
        for (int d = 0; d < 3; ++d) {
            h->nijkp1[0][b][d] = t->nijkp1[3 * b + d];
            if (h->nijkp1[0][b][d] < 3) return fail(LMGX_E_ARG, "block dims must be >= 3");
        }
        for (int s = 0; s < 6; ++s) {
            h->bc[b][s] = t->bc[6 * b + s];
            h->bl2bl[b][s] = t->bl2bl[6 * b + s];
            h->bl2side[b][s] = t->bl2side[6 * b + s];
        }
    }
    int found = 1;
    for (int l = 0; l < LMGX_MAXLEV - 1; ++l) {
        bool ok = true;
        for (int b = 0; b < h->nblk && ok; ++b) {
            const int *n = h->nijkp1[l][b];
            if (n[0] % 2 || n[1] % 2 || n[2] % 2) { ok = false; break; }
            int c[3] = {n[0] / 2 + 1, n[1] / 2 + 1, n[2] / 2 + 1};
            if (c[0] < 4 || c[1] < 4 || c[2] < 4) { ok = false; break; }
            for (int d = 0; d < 3; ++d) h->nijkp1[l + 1][b][d] = c[d];
        }
        if (!ok) break;
        found = l + 2;
    }
    int L = t->nlev == 0 ? found : t->nlev;
    if (L < 1 || L > found)
        return fail(LMGX_E_LEVELS, "requested " + std::to_string(t->nlev) + " levels, dims allow " +
                                       std::to_string(found));
    h->nlev = L;
    long o = 0, o2 = 0, o1 = 0, oxy = 0, oz = 0;
    for (int l = 0; l < L; ++l)
        for (int b = 0; b < h->nblk; ++b) {
            const int *n = h->nijkp1[l][b];
            h->off[l][b] = o;
            h->off2d[l][b] = o2;
            h->off1d[l][b] = o1;
            h->offxy[l][b] = oxy;
            h->offz[l][b] = oz;
            o += (long)n[0] * n[1] * n[2];
            o2 += (long)n[0] * n[1];
            o1 += n[2];
            oxy += (long)(n[0] + 1) * (n[1] + 1);
            oz += n[2] + 1;
            if (l == 0) h->n1 = o;
        }
    h->ntot = o;
    h->n2d = o2;
    h->n1d = o1;
    h->nxy = oxy;
    h->nz = oz;
    // face consistency (the reference prints a mismatch and carries on; here it is an error)
    for (int b = 0; b < h->nblk; ++b)
        for (int s = 1; s <= 6; ++s) {
            int nb = h->bl2bl[b][s - 1], ns = h->bl2side[b][s - 1];
            bool sw = is_swap(h->bc[b][s - 1]);
            if (!sw) continue;
            if (nb < 1 || nb > h->nblk || ns < 1 || ns > 6)
                return fail(LMGX_E_FACE, "swapped side without a valid partner");
            if (s % 2 == ns % 2) return fail(LMGX_E_FACE, "interface coupling invalid");
            if (!is_swap(h->bc[nb - 1][ns - 1]))
                return fail(LMGX_E_FACE, "partner side is not INTERFACE/PERIODIC");
            for (int l = 0; l < L; ++l) {
                int d1 = (s + 1) / 2, e1 = (ns + 1) / 2;
                if (h->nijkp1[l][b][CYC[0][d1 - 1] - 1] != h->nijkp1[l][nb - 1][CYC[0][e1 - 1] - 1] ||
                    h->nijkp1[l][b][CYC[1][d1 - 1] - 1] != h->nijkp1[l][nb - 1][CYC[1][e1 - 1] - 1])
                    return fail(LMGX_E_FACE, "block " + std::to_string(b + 1) + " face " +
                                                 std::to_string(s) + " mismatch");
            }
        }
    for (int l = 0; l < L; ++l) {
        LevelDev &D = h->lev[l];
        D.h_lb.resize(h->nblk);
        long ti = 0, ta = 0, tr = 0, tl = 0, tl2 = 0, tl4 = 0, trr = 0, trp = 0, tsg = 0, tsg8 = 0;
        long t2s = 0;
        long sl = 0, ns = 0;
        double np = 0.0;
        long npi = 0;
        int smax = 0;
        for (int b = 0; b < h->nblk; ++b) {
            const int *n = h->nijkp1[l][b];
            LevBlk &B = D.h_lb[b];
            memset(&B, 0, sizeof(B));
            for (int q = 0; q < 4; ++q) B.fz_off[q] = -1;
            B.off = h->off[l][b];
            B.n[0] = n[0]; B.n[1] = n[1]; B.n[2] = n[2];
            // row pitch of the split layout: ceil(nip1/2) slots rounded up to 32, so every row
            // of every array starts on a 128-byte line (FP32) and a wavefront's 64 slots cover
            // exactly two lines
            B.hp = ((n[0] + 1) / 2 + 31) & ~31;
            B.off_s = ns;
            ns += (long)B.hp * n[1] * n[2];
            for (int s = 0; s < 6; ++s) {
                B.ifc[s] = h->bc[b][s] == BCT_INTERFACE;
                B.swp[s] = is_swap(h->bc[b][s]);
            }
            B.t_int = ti;
            ti += (long)((n[0] - 2 + 63) / 64) * ((n[1] - 2 + 3) / 4);
            B.t_all = ta;
            ta += (long)((n[0] + 63) / 64) * ((n[1] + 3) / 4);
            B.t_rb = tr;
            tr += (long)((B.hp + 63) / 64) * ((n[1] - 2 + 3) / 4);
            B.t_ln = tl;
            tl += (long)((B.hp + 63) / 64) * ((n[2] - 2 + 3) / 4);
            B.t_z64 = tl2;
            tl2 += (long)((B.hp + 63) / 64) * ((n[2] - 2 + 1) / 2);
            B.t_z128 = tl4;
            tl4 += (long)((B.hp + 127) / 128) * (n[2] - 2);
            B.t_rr = trr;
            trr += (long)((((n[0] - 2) >> 1) + 63) / 64) * ((((n[2] - 2) >> 1) + 3) / 4) *
                   ((((n[1] - 2) >> 1) + LMGX_RCH - 1) / LMGX_RCH);
            {
                // segmented line solve: fewest power-of-two segments of <= LMGX_SMAX rows
                const int nj = n[1] - 2;
                int ns = 1;
                while (ns < 16 && (nj + ns - 1) / ns > LMGX_SMAX) ns *= 2;
                B.sg_n = ns;
                B.sg_r = (nj + ns - 1) / ns;
                B.sg_s = 16 / ns;
                // the over-relaxed solves on binary16 couplings and in FP64 run the body
                // instance with the fewest rows that holds a segment (10, 12, 14 or 16: fewer
                // repeated rows); the others and LMGX_SEG_FULL=1 take the 16-row one, which
                // gives the same values
                const char *full = getenv("LMGX_SEG_FULL");
                B.sg_v = (full && atoi(full)) ? LMGX_SMAX
                         : B.sg_r <= 10 ? 10 : B.sg_r <= 12 ? 12 : B.sg_r <= 14 ? 14 : LMGX_SMAX;
            }
            B.t_sg = tsg;
            tsg += (long)((B.hp + 63) / 64) * ((n[2] - 2 + B.sg_s - 1) / B.sg_s);
            B.sg_s8 = B.sg_n <= 8 ? 8 / B.sg_n : 0;
            B.t_sg8 = tsg8;
            if (B.sg_s8) tsg8 += (long)((B.hp + 63) / 64) * ((n[2] - 2 + B.sg_s8 - 1) / B.sg_s8);
            B.t_rp = trp;
            trp += (long)((((n[0] - 2) >> 1) + 63) / 64) * ((((n[2] - 2) >> 1) + 3) / 4) *
                   ((((n[1] - 2) >> 1) + LMGX_RCHP - 1) / LMGX_RCHP);
            B.s_jk = sl;
            sl += (long)(n[1] - 2) * (n[2] - 2);
            B.o2 = h->off2d[l][b];
            B.o1 = h->off1d[l][b];
            B.oxy = h->offxy[l][b];
            B.oz = h->offz[l][b];
            B.o2s = t2s;
            t2s += (long)n[1] * B.hp;
            np = np + (double)((long)(n[0] - 2) * (n[1] - 2) * (n[2] - 2));
            npi += (long)(n[0] - 2) * (n[1] - 2) * (n[2] - 2);
            smax = std::max(smax, n[0] - 1 + n[1] - 1 + n[2] - 1);
            D.njmax = std::max(D.njmax, n[1] - 2);
        }
        D.tiles_int = ti;
        D.tiles_all = ta;
        D.tiles_rb = tr;
        D.tiles_ln = tl;
        D.tiles_z64 = tl2;
        D.tiles_z128 = tl4;
        D.tiles_rr = trr;
        D.tiles_rp = trp;
        D.tiles_sg = tsg;
        D.tiles_sg8 = tsg8;
        D.n2s = t2s;
        D.nlines = sl;
        D.smax = smax;
        D.npts_int = npi;
        D.nsplit = ns;
        h->npts[l] = np;
        // faces, in block order and side order (the pack order of swap_sc0.F:64-151)
        D.h_fc.clear();
        long e = 0, buf = 0;
        std::vector<long> slot(h->nblk * 6, -1);
        for (int b = 0; b < h->nblk; ++b)
            for (int s = 1; s <= 6; ++s) {
                if (!is_swap(h->bc[b][s - 1])) continue;
                const int *n = h->nijkp1[l][b];
                int d1 = (s + 1) / 2, d2 = CYC[0][d1 - 1], d3 = CYC[1][d1 - 1];
                long ijko[3] = {1, n[0], (long)n[0] * n[1]};
                int ni = n[d1 - 1];
                int ipk = (s % 2 == 0) ? ni - 1 : 2, iun = (s % 2 == 0) ? ni : 1;
                Face F;
                memset(&F, 0, sizeof(F));
                F.pack_base = h->off[l][b] + ijko[d1 - 1] * (ipk - 1);
                F.unpack_base = h->off[l][b] + ijko[d1 - 1] * (iun - 1);
                F.jo = ijko[d2 - 1];
                F.ko = ijko[d3 - 1];
                F.nj = n[d2 - 1];
                F.nk = n[d3 - 1];
                F.buf = buf;
                F.e0 = e;
                F.blk = b;
                F.ax_j = d2 - 1;
                F.ax_k = d3 - 1;
                for (int a = 0; a < 3; ++a) { F.pack_i[a] = 0; F.unpack_i[a] = 0; }
                F.pack_i[d1 - 1] = ipk - 1;
                F.unpack_i[d1 - 1] = iun - 1;
                // points of this halo plane that a later swapped side of the block overwrites
                int sk = 0;
                if (2 * d2 - 1 > s && is_swap(h->bc[b][2 * d2 - 2])) sk |= 1;
                if (2 * d2 > s && is_swap(h->bc[b][2 * d2 - 1])) sk |= 2;
                if (2 * d3 - 1 > s && is_swap(h->bc[b][2 * d3 - 2])) sk |= 4;
                if (2 * d3 > s && is_swap(h->bc[b][2 * d3 - 1])) sk |= 8;
                F.skip = sk;
                slot[b * 6 + s - 1] = buf;
                buf += (long)F.nj * F.nk;
                e += (long)F.nj * F.nk;
                D.h_fc.push_back(F);
            }
        for (auto &F : D.h_fc) {
            // which (block, side) is this?
            int b = F.blk, s = 0;
            for (int q = 1; q <= 6; ++q)
                if (is_swap(h->bc[b][q - 1]) && slot[b * 6 + q - 1] == F.buf) s = q;
            int nb = h->bl2bl[b][s - 1], nsd = h->bl2side[b][s - 1];
            F.src = slot[(nb - 1) * 6 + nsd - 1];
        }
        D.nfaces = (int)D.h_fc.size();
        D.nel = e;
    }
    return 0;
}

// the level's block table, and every block alone with its tile offsets rebased to 0
static int copy_tables(lmgx_handle *h, int l)
{
    LevelDev &D = h->lev[l];
    HIPCHK(hipMemcpy(D.d_lb, D.h_lb.data(), sizeof(LevBlk) * h->nblk, hipMemcpyHostToDevice));
    std::vector<LevBlk> one(D.h_lb);
    for (LevBlk &B : one) {
        B.t_int = B.t_all = B.t_rb = B.t_ln = B.t_z64 = B.t_z128 = B.t_rr = B.t_rp = B.t_sg = 0;
        B.t_sg8 = 0;
    }
    HIPCHK(hipMemcpy(D.d_lb1, one.data(), sizeof(LevBlk) * h->nblk, hipMemcpyHostToDevice));
    return 0;
}

static int upload_tables(lmgx_handle *h)
{
    long swap_need = 1, part_need = 64;
    for (int l = 0; l < h->nlev; ++l) {
        LevelDev &D = h->lev[l];
        HIPCHK(hipMalloc(&D.d_lb, sizeof(LevBlk) * h->nblk));
        HIPCHK(hipMalloc(&D.d_lb1, sizeof(LevBlk) * h->nblk));
        CHK(copy_tables(h, l));
        auto span = [&](long LevBlk::*f, long total, std::vector<long> &out) {
            out.resize(h->nblk);
            for (int b = 0; b < h->nblk; ++b)
                out[b] = (b + 1 < h->nblk ? D.h_lb[b + 1].*f : total) - D.h_lb[b].*f;
        };
        span(&LevBlk::t_sg, D.tiles_sg, D.bt_sg);
        span(&LevBlk::t_sg8, D.tiles_sg8, D.bt_sg8);
        span(&LevBlk::t_z64, D.tiles_z64, D.bt_z64);
        span(&LevBlk::t_z128, D.tiles_z128, D.bt_z128);
        span(&LevBlk::t_rr, D.tiles_rr, D.bt_rr);
        span(&LevBlk::t_rp, D.tiles_rp, D.bt_rp);
        if (D.nfaces) {
            HIPCHK(hipMalloc(&D.d_fc, sizeof(Face) * D.nfaces));
            HIPCHK(hipMemcpy(D.d_fc, D.h_fc.data(), sizeof(Face) * D.nfaces, hipMemcpyHostToDevice));
        }
        swap_need = std::max(swap_need, D.nel);
        part_need = std::max(part_need, 4 * std::max(std::max(D.tiles_int, D.tiles_ln),
                                                     std::max(D.tiles_all, D.tiles_rb)));
        part_need = std::max(part_need, 4 * D.tiles_rr);
    }
    HIPCHK(hipMalloc(&h->d_swap, sizeof(double) * swap_need));
    h->swap_cap = swap_need;
    HIPCHK(hipMalloc(&h->d_partial, sizeof(double) * part_need));
    h->partial_cap = part_need;
    HIPCHK(hipMalloc(&h->d_partial2, sizeof(double) * 4 * LMGX_RMID));
    HIPCHK(hipMalloc(&h->d_scalar, sizeof(double) * 64));
    return 0;
}

extern "C" void lmgx_default_params(lmgx_params *p)
{
    // code/defaults.F:59-71
    p->cycle = 1;
    p->numpre = 1;
    p->numpost = 1;
    p->numcits = 20;
    p->max_cycles = 5;
    p->omega = 1.6;
    p->omega_c = 1.6;
    p->tolf = 5.0e-13;
    p->rtol = 0.0;
    p->smoother = LMGX_SMOOTH_EXACT;
    p->precision = LMGX_PREC_FP64;
    p->check_every = 1;
    p->sweeps_per_iter = 3;
    p->coarse_op = 1;
    p->coarse_solve = 1;
    p->fine_post = 0;
}

static void free_fast(lmgx_handle *h);

extern "C" int lmgx_free(lmgx_handle *h)
{
    if (!h) return 0;
    for (int l = 0; l < h->nlev; ++l) {
        hipFree(h->lev[l].d_lb);
        hipFree(h->lev[l].d_lb1);
        hipFree(h->lev[l].d_fc);
    }
    for (int q = 0; q < 7; ++q) { hipFree(h->d_a[q]); hipFree(h->d_aT[q]); }
    hipFree(h->d_rho_op);
    hipFree(h->d_f); hipFree(h->d_b); hipFree(h->d_res);
    hipFree(h->d_swap); hipFree(h->d_partial); hipFree(h->d_partial2); hipFree(h->d_scalar);
    hipFree(h->d_hist);
    hipFree(h->d_met); hipFree(h->d_dkp); hipFree(h->d_dkw);
    hipFree(h->d_mft); hipFree(h->d_mfk); hipFree(h->d_mfd);
    if (h->h_pinned) hipHostFree(h->h_pinned);
    free_fast(h);
    for (auto &e : h->prof_ev) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

// Scratch of lmgx_p_mean_subtract_device_noalloc: the workgroup partials (sum, max, min) of the
// two-stage level-1 interior sum, then the three final sums and the mean.  A file-scope device
// array exists on every GPU the library runs on from its load to the end of the process, so
// that entry point allocates nothing.  In host code `scratch` is only a stand-in for the symbol:
// kernels get the device address that hipGetSymbolAddress returns for the current GPU.
#define LMGX_NOALLOC_PARTS 8192
#define LMGX_MAXDEV 64
static __device__ __align__(256) double scratch[3 * LMGX_NOALLOC_PARTS + 8];
static double *g_scratch[LMGX_MAXDEV];
static std::mutex g_scratch_mu;

// the device address of `scratch` on the current GPU, looked up once per GPU
static int scratch_address(double **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= LMGX_MAXDEV) return fail(LMGX_E_UNSUPPORTED, "device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(LMGX_E_HIP, "hipGetSymbolAddress returned NULL");
        g_scratch[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch[dev];
    return 0;
}

extern "C" int lmgx_init(const lmgx_topology *topo, lmgx_handle **out)
{
    if (!out) return fail(LMGX_E_ARG, "out is NULL");
    *out = nullptr;
    lmgx_handle *h = new lmgx_handle();
    int rc = build_tables(h, topo);
    if (rc == 0) rc = scratch_address(&h->scratch);
    if (rc == 0) {
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
            rc = fail(LMGX_E_HIP, "hipStreamCreate failed");
        } else {
            rc = upload_tables(h);
        }
    }
    if (rc != 0) {
        std::string msg = g_err;
        lmgx_free(h);
        g_err = msg;
        return rc;
    }
    *out = h;
    return 0;
}

extern "C" int lmgx_levels(const lmgx_handle *h, int *nlev, long *ntot, long *n1,
                           int *nijkp1_all, long *off_all)
{
    if (!h) return fail(LMGX_E_ARG, "NULL handle");
    if (nlev) *nlev = h->nlev;
    if (ntot) *ntot = h->ntot;
    if (n1) *n1 = h->n1;
    for (int l = 0; l < h->nlev; ++l)
        for (int b = 0; b < h->nblk; ++b) {
            if (nijkp1_all)
                for (int d = 0; d < 3; ++d)
                    nijkp1_all[(l * h->nblk + b) * 3 + d] = h->nijkp1[l][b][d];
            if (off_all) off_all[l * h->nblk + b] = h->off[l][b];
        }
    return 0;
}

extern "C" int lmgx_set_dirichlet(lmgx_handle *h, const int *flags)
{
    if (!h) return fail(LMGX_E_ARG, "NULL handle");
    for (int b = 0; b < h->nblk; ++b)
        for (int s = 0; s < 6; ++s) {
            h->dirich[b][s] = flags ? (flags[6 * b + s] != 0 && !is_swap(h->bc[b][s])) : 0;
            for (int l = 0; l < h->nlev; ++l) h->lev[l].h_lb[b].dir[s] = h->dirich[b][s];
        }
    for (int l = 0; l < h->nlev; ++l) CHK(copy_tables(h, l));
    return 0;
}

extern "C" int lmgx_get_jumble(const lmgx_handle *h) { return h ? h->jumble : 0; }
extern "C" int lmgx_set_jumble(lmgx_handle *h, int j)
{
    if (!h) return fail(LMGX_E_ARG, "NULL handle");
    h->jumble = j;
    return 0;
}

static int ensure_operator_storage(lmgx_handle *h)
{
    for (int q = 0; q < 7; ++q)
        if (!h->d_a[q]) HIPCHK(hipMalloc(&h->d_a[q], sizeof(double) * h->ntot));
    return 0;
}

static int fast_operator_changed(lmgx_handle *h);
static int ensure_natural(lmgx_handle *h);

extern "C" int lmgx_set_operator(lmgx_handle *h, const double *ap, const double *aw,
                                 const double *ae, const double *as, const double *an,
                                 const double *al, const double *ar)
{
    if (!h) return fail(LMGX_E_ARG, "NULL handle");
    const double *src[7] = {ap, aw, ae, as, an, al, ar};
    for (int q = 0; q < 7; ++q)
        if (!src[q]) return fail(LMGX_E_ARG, "NULL coefficient array");
    CHK(ensure_operator_storage(h));
    for (int q = 0; q < 7; ++q)
        HIPCHK(hipMemcpyAsync(h->d_a[q], src[q], sizeof(double) * h->ntot, hipMemcpyHostToDevice,
                              h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->have_op = true;
    h->nat_valid = true;
    h->aT_valid = false;
    h->op_builder = false;
    h->cached_valid = false;
    ++h->n_uploads;
    hipFree(h->d_rho_op);
    h->d_rho_op = nullptr;
    return fast_operator_changed(h);
}

// 64-bit FNV-1a over the bits of every 4093rd element (and the last) of the seven arrays
static unsigned long long operator_fingerprint(const double *const src[7], long n)
{
    unsigned long long fp = 1469598103934665603ull;
    for (int q = 0; q < 7; ++q)
        for (long i = 0; i < n; i = (i + 4093 < n || i == n - 1) ? i + 4093 : n - 1) {
            unsigned long long v;
            memcpy(&v, src[q] + i, sizeof(v));
            fp = (fp ^ v) * 1099511628211ull;
        }
    return fp;
}

extern "C" int lmgx_set_operator_cached(lmgx_handle *h, const double *ap, const double *aw,
                                        const double *ae, const double *as, const double *an,
                                        const double *al, const double *ar, int force,
                                        int *uploaded)
{
    if (!h) return fail(LMGX_E_ARG, "NULL handle");
    const double *src[7] = {ap, aw, ae, as, an, al, ar};
    for (int q = 0; q < 7; ++q)
        if (!src[q]) return fail(LMGX_E_ARG, "NULL coefficient array");
    if (uploaded) *uploaded = 0;
    const unsigned long long fp = operator_fingerprint(src, h->ntot);
    bool same = h->cached_valid && h->have_op && !h->op_flipped && fp == h->cached_fp && !force;
    for (int q = 0; q < 7; ++q) same = same && src[q] == h->cached_src[q];
    if (same) return 0;
    CHK(lmgx_set_operator(h, ap, aw, ae, as, an, al, ar));
    for (int q = 0; q < 7; ++q) h->cached_src[q] = src[q];
    h->cached_fp = fp;
    h->cached_valid = true;
    if (uploaded) *uploaded = 1;
    return 0;
}

extern "C" int lmgx_counters(const lmgx_handle *h, long out[4])
{
    if (!h || !out) return fail(LMGX_E_ARG, "NULL argument");
    out[0] = h->n_uploads;
    out[1] = h->n_prepares;
    out[2] = h->n_captures;
    out[3] = h->n_solves;
    return 0;
}

extern "C" int lmgx_set_operator_device(lmgx_handle *h, const double *const d_coeff[7])
{
    if (!h || !d_coeff) return fail(LMGX_E_ARG, "NULL argument");
    CHK(ensure_operator_storage(h));
    for (int q = 0; q < 7; ++q)
        HIPCHK(hipMemcpyAsync(h->d_a[q], d_coeff[q], sizeof(double) * h->ntot,
                              hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->have_op = true;
    h->nat_valid = true;
    h->aT_valid = false;
    h->op_builder = false;
    h->cached_valid = false;
    ++h->n_uploads;
    hipFree(h->d_rho_op);
    h->d_rho_op = nullptr;
    return fast_operator_changed(h);
}

extern "C" int lmgx_get_operator(const lmgx_handle *h, double *ap, double *aw, double *ae,
                                 double *as, double *an, double *al, double *ar)
{
    if (!h || !h->have_op) return fail(LMGX_E_STATE, "no operator");
    CHK(ensure_natural(const_cast<lmgx_handle *>(h)));
    double *dst[7] = {ap, aw, ae, as, an, al, ar};
    for (int q = 0; q < 7; ++q)
        if (dst[q])
            HIPCHK(hipMemcpyAsync(dst[q], h->d_a[q], sizeof(double) * h->ntot,
                                  hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// ---------------------------------------------------------------------------------------------
// reference-order passes
// ---------------------------------------------------------------------------------------------
static int ensure_hist(lmgx_handle *h, long n)
{
    if (n > h->hist_cap) {
        hipFree(h->d_hist);
        HIPCHK(hipMalloc(&h->d_hist, sizeof(double) * n));
        h->hist_cap = n;
    }
    if (n > h->pinned_cap) {
        if (h->h_pinned) hipHostFree(h->h_pinned);
        HIPCHK(hipHostMalloc(&h->h_pinned, sizeof(double) * n, hipHostMallocDefault));
        h->pinned_cap = n;
    }
    return 0;
}

// The 4 norms from nparts workgroup partials in h->d_partial (RMS with the reference's var = 0 -> 1
// rule) into d_out: long lists go through LMGX_RMID middle-stage workgroups first, so no single
// workgroup walks tens of thousands of partials.
static void reduce_norms(lmgx_handle *h, long nparts, double scale, double *d_out)
{
    if (nparts <= 4 * LMGX_WG) {
        hipLaunchKernelGGL(k_reduce_final<4>, dim3(1), dim3(LMGX_WG), 0, h->stream,
                           (const double *)h->d_partial, nparts, scale, d_out, 0);
        return;
    }
    const long chunk = (nparts + LMGX_RMID - 1) / LMGX_RMID;
    hipLaunchKernelGGL(k_reduce_mid<4>, dim3(LMGX_RMID), dim3(LMGX_WG), 0, h->stream,
                       (const double *)h->d_partial, nparts, chunk, h->d_partial2);
    hipLaunchKernelGGL(k_reduce_final<4>, dim3(1), dim3(LMGX_WG), 0, h->stream,
                       (const double *)h->d_partial2, (long)LMGX_RMID, scale, d_out, 0);
}

// norms of level l into d_out[4] (device), residual into res
static int norms_exact(lmgx_handle *h, int l, const double *x, const double *b, double *res,
                       double *d_out)
{
    LevelDev &D = h->lev[l];
    double *const *a = h->d_a;
    hipLaunchKernelGGL(k_norms_exact, dim3(D.tiles_int), dim3(LMGX_WG), 0, h->stream, D.d_lb,
                       h->nblk, a[0], a[1], a[2], a[3], a[4], a[5], a[6], x, b, res, h->d_partial);
    reduce_norms(h, D.tiles_int, 1.0 / h->npts[l], d_out);
    HIPCHK(hipGetLastError());
    return 0;
}

template <typename T>
static int swap_level(lmgx_handle *h, int l, T *x)
{
    LevelDev &D = h->lev[l];
    if (D.nel == 0) return 0;
    T *buf = reinterpret_cast<T *>(h->d_swap);
    hipLaunchKernelGGL(k_swap_pack<T>, dim3(nblocks(D.nel)), dim3(LMGX_WG), 0, h->stream, D.d_fc,
                       D.nfaces, D.nel, x, buf);
    hipLaunchKernelGGL(k_swap_unpack<T>, dim3(nblocks(D.nel)), dim3(LMGX_WG), 0, h->stream, D.d_fc,
                       D.nfaces, D.nel, buf, x);
    HIPCHK(hipGetLastError());
    return 0;
}

static void prof_begin(lmgx_handle *h);
static void prof_end(lmgx_handle *h, double cells);

static void sor_sweep(lmgx_handle *h, int l, double *f, const double *b, double omega, int d1)
{
    LevelDev &D = h->lev[l];
    double *const *a = h->d_a;
    prof_begin(h);
    for (int s = 6; s <= D.smax; ++s)
        hipLaunchKernelGGL(k_sor_hyperplane, dim3(nblocks(D.nlines)), dim3(LMGX_WG), 0, h->stream,
                           D.d_lb, h->nblk, D.nlines, s, d1, omega, a[0], a[1], a[2], a[3], a[4],
                           a[5], a[6], f, b);
    prof_end(h, (double)D.npts_int);
}

// one sporlx2 call (meth_smooth = 2): 7 sweeps in natural order, i.e. spor with
// (io, jo, ko) = (1, nip1, nip1*njp1) and the d1 = 1 summation order, a halo swap after each;
// the SAVEd jumble is not used (code/sporlx2.F:64-146)
static int sporlx2_exact(lmgx_handle *h, int l, double *f, const double *b, double omega)
{
    for (int il = 0; il < 7; ++il) {
        sor_sweep(h, l, f, b, omega, 1);
        CHK(swap_level<double>(h, l, f));
    }
    HIPCHK(hipGetLastError());
    return 0;
}

// one sporlx call: 3 sweeps, halo swap after each (code/sporlx.F:74-159)
static int sporlx_exact(lmgx_handle *h, int l, double *f, const double *b, double omega)
{
    int jumfor = (h->jumble % 16 >= 8) ? 1 : 2;
    int jumst = (h->jumble % 24) / 8 + 1;
    int d[3] = {jumst, CYC[(jumfor + 1) % 2][jumst - 1], CYC[jumfor % 2][jumst - 1]};
    for (int il = 0; il < 3; ++il) {
        sor_sweep(h, l, f, b, omega, d[il]);
        CHK(swap_level<double>(h, l, f));
    }
    HIPCHK(hipGetLastError());
    h->jumble = h->jumble + 13;
    if (h->jumble > 99999) h->jumble = h->jumble % 99999;
    return 0;
}

static int restrict_exact(lmgx_handle *h, int lc, const double *res, double *b, double *f)
{
    LevelDev &C = h->lev[lc];
    hipLaunchKernelGGL(k_restrict_zero<double>, dim3(C.tiles_all), dim3(LMGX_WG), 0, h->stream,
                       C.d_lb, h->lev[lc - 1].d_lb, h->nblk, res, b, f);
    HIPCHK(hipGetLastError());
    return 0;
}

static int prolong_exact(lmgx_handle *h, int lf, double *f)
{
    LevelDev &F = h->lev[lf], &C = h->lev[lf + 1];
    hipLaunchKernelGGL(k_zero_grad<double>, dim3(C.tiles_all), dim3(LMGX_WG), 0, h->stream, C.d_lb,
                       h->nblk, f, 0);
    hipLaunchKernelGGL(k_prolong_exact<double>, dim3(F.tiles_all), dim3(LMGX_WG), 0, h->stream,
                       F.d_lb, C.d_lb, h->nblk, (const double *)f, f);
    HIPCHK(hipGetLastError());
    return swap_level<double>(h, lf, f);
}

// code/smooth.F:34-100 (the smoothing stop test never fires with soltol = 0 unless a sweep
// leaves f exactly unchanged; this path always runs the full count).  smoother selects
// meth_smooth: LMGX_SMOOTH_EXACT = sporlx (1), LMGX_SMOOTH_SPORLX2 = sporlx2 (2)
static int smooth_exact(lmgx_handle *h, int l, double *f, double *b, double *res, int nsmooth,
                        double omega, double *d_nrm, int smoother)
{
    for (int it = 0; it < nsmooth; ++it) {
        if (smoother == LMGX_SMOOTH_SPORLX2) CHK(sporlx2_exact(h, l, f, b, omega));
        else CHK(sporlx_exact(h, l, f, b, omega));
    }
    return norms_exact(h, l, f, b, res, d_nrm);
}

static int read_norms(lmgx_handle *h, const double *d_src, double *host4)
{
    HIPCHK(hipMemcpyAsync(h->h_pinned, d_src, 4 * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    memcpy(host4, h->h_pinned, 4 * sizeof(double));
    return 0;
}

// code/lmg_solve.F:44-160 with nummgits = 1 per call, repeated up to max_cycles times
static int solve_exact(lmgx_handle *h, double *f, double *b, double *res, const lmgx_params *p,
                       double *history, int *ncyc)
{
    const int L = h->nlev;
    // device norm slots: current[l][4] in d_scalar (l < 8 -> 32 doubles)
    double *d_nrm = h->d_scalar;
    double n0[4];
    double rnorm0 = -1.0;
    int done = 0;
    HIPCHK(hipMemsetAsync(d_nrm, 0, sizeof(double) * 4 * LMGX_MAXLEV, h->stream));
    CHK(norms_exact(h, 0, f, b, res, d_nrm));
    CHK(read_norms(h, d_nrm, n0));
    rnorm0 = n0[0];
    double cur[4] = {n0[0], n0[1], n0[2], n0[3]};
    for (int call = 0; call < p->max_cycles; ++call) {
        if (!(cur[2] / cur[3] > p->tolf)) break;
        if (p->rtol > 0.0 && rnorm0 > 0.0 && cur[0] / rnorm0 <= p->rtol) break;
        const bool fcycle = p->cycle == 2;
        int gamma = fcycle ? 2 : 1;
        int nummg[LMGX_MAXLEV];
        nummg[0] = 1;
        int il = 0;
        // the per-call entry norms (lmg_solve.F:60) equal the previous exit norms bit for bit
        while (nummg[0] > 0) {
            bool goto_fine = (il == L - 1) || (nummg[il] == 0);
            if (goto_fine) {
                if (il == L - 1) {
                    CHK(norms_exact(h, il, f, b, res, d_nrm + 4 * il));
                    CHK(smooth_exact(h, il, f, b, res, p->numcits, p->omega_c, d_nrm + 4 * il,
                                     p->smoother));
                    if (fcycle) gamma = 1;
                }
                if (il > 0 && L != 1) {
                    il = il - 1;
                    CHK(prolong_exact(h, il, f));
                    if (il == 0) CHK(norms_exact(h, il, f, b, res, d_nrm));
                    CHK(smooth_exact(h, il, f, b, res, p->numpost, p->omega, d_nrm + 4 * il,
                                     p->smoother));
                    if (il == 0 && fcycle) gamma = 2;
                }
                nummg[il] = nummg[il] - 1;
            } else {
                if (il == 0) CHK(norms_exact(h, il, f, b, res, d_nrm));
                CHK(smooth_exact(h, il, f, b, res, p->numpre, p->omega, d_nrm + 4 * il,
                                 p->smoother));
                il = il + 1;
                nummg[il] = gamma;
                CHK(restrict_exact(h, il, res, b, f));
            }
        }
        HIPCHK(hipMemcpyAsync(h->d_hist + (long)call * L * 4, d_nrm, sizeof(double) * 4 * L,
                              hipMemcpyDeviceToDevice, h->stream));
        CHK(read_norms(h, d_nrm, cur));
        ++done;
        if (!std::isfinite(cur[0])) {
            *ncyc = done;
            return fail(LMGX_E_NUMERIC, "non-finite residual norm");
        }
    }
    if (history && done) {
        HIPCHK(hipMemcpyAsync(h->h_pinned, h->d_hist, sizeof(double) * 4 * L * done,
                              hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        memcpy(history, h->h_pinned, sizeof(double) * 4 * L * done);
    }
    *ncyc = done;
    return 0;
}

static int solve_fast(lmgx_handle *h, double *f, double *b, double *res, const lmgx_params *p,
                      double *history, int *ncyc);

static int check_params(const lmgx_handle *h, const lmgx_params *p)
{
    if (!p) return fail(LMGX_E_ARG, "NULL params");
    if (p->cycle != 1 && p->cycle != 2) return fail(LMGX_E_ARG, "cycle must be 1 (V) or 2 (F)");
    if (p->numpre < 0 || p->numpost < 0 || p->numcits < 0 || p->max_cycles < 0 ||
        p->fine_post < 0)
        return fail(LMGX_E_ARG, "negative iteration count");
    if (p->smoother != LMGX_SMOOTH_EXACT && p->smoother != LMGX_SMOOTH_LINE &&
        p->smoother != LMGX_SMOOTH_SPORLX2)
        return fail(LMGX_E_UNSUPPORTED, "smoothing method " + std::to_string(p->smoother) +
                                            " not implemented (code/smooth.F:77-79)");
    if (p->precision != LMGX_PREC_FP64 && p->precision != LMGX_PREC_MIXED)
        return fail(LMGX_E_ARG, "unknown precision");
    if (p->precision == LMGX_PREC_MIXED && p->smoother != LMGX_SMOOTH_LINE)
        return fail(LMGX_E_ARG, "mixed precision needs the line smoother");
    if (p->smoother == LMGX_SMOOTH_LINE && p->cycle != 1)
        return fail(LMGX_E_UNSUPPORTED, "the line-smoother path runs V-cycles only");
    if (!h->have_op) return fail(LMGX_E_STATE, "lmgx_set_operator has not been called");
    return 0;
}

extern "C" int lmgx_solve_device(lmgx_handle *h, double *d_f, double *d_b, double *d_res,
                                 const lmgx_params *prm, double *history, int *ncycles_out)
{
    if (!h || !d_f || !d_b || !d_res) return fail(LMGX_E_ARG, "NULL argument");
    // the solve reads b and writes f and residu throughout (the fast path reads b where the caller
    // holds it and may write f and residu in a cycle that turns out not to be the last)
    auto overlap = [&](const double *x, const double *y) {
        return x < y + h->ntot && y < x + h->ntot;
    };
    if (overlap(d_f, d_b) || overlap(d_f, d_res) || overlap(d_b, d_res))
        return fail(LMGX_E_ARG, "f, b and residu must not overlap");
    CHK(check_params(h, prm));
    int nc = 0;
    CHK(ensure_hist(h, (long)std::max(1, prm->max_cycles) * h->nlev * 4 + 64));
    ++h->n_solves;
    if (prm->smoother != LMGX_SMOOTH_LINE) CHK(ensure_natural(h));
    int rc = prm->smoother != LMGX_SMOOTH_LINE ? solve_exact(h, d_f, d_b, d_res, prm, history, &nc)
                                               : solve_fast(h, d_f, d_b, d_res, prm, history, &nc);
    if (ncycles_out) *ncycles_out = nc;
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int lmgx_solve(lmgx_handle *h, double *f, double *b, double *residu,
                          const lmgx_params *prm, double *history, int *ncycles_out)
{
    if (!h || !f || !b || !residu) return fail(LMGX_E_ARG, "NULL argument");
    CHK(check_params(h, prm));
    const size_t bytes = sizeof(double) * h->ntot;
    if (!h->d_f) {
        HIPCHK(hipMalloc(&h->d_f, bytes));
        HIPCHK(hipMalloc(&h->d_b, bytes));
        HIPCHK(hipMalloc(&h->d_res, bytes));
    }
    // The reference-order path writes f, b (levels >= 2) and residu on every level.  The line
    // path reads f and b and writes f and residu on level 1 only (its coarse levels live in the
    // split layout), so only level 1 travels, and b does not come back.
    const bool line = prm->smoother == LMGX_SMOOTH_LINE;
    const size_t part = line ? sizeof(double) * h->n1 : bytes;
    HIPCHK(hipMemcpyAsync(h->d_f, f, part, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_b, b, part, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_res, residu, part, hipMemcpyHostToDevice, h->stream));
    int rc = lmgx_solve_device(h, h->d_f, h->d_b, h->d_res, prm, history, ncycles_out);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(f, h->d_f, part, hipMemcpyDeviceToHost, h->stream));
    if (!line) HIPCHK(hipMemcpyAsync(b, h->d_b, part, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(residu, h->d_res, part, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// ---------------------------------------------------------------------------------------------
// single passes
// ---------------------------------------------------------------------------------------------
static int check_level(const lmgx_handle *h, int l)
{
    if (!h) return fail(LMGX_E_ARG, "NULL handle");
    if (l < 0 || l >= h->nlev) return fail(LMGX_E_ARG, "level out of range");
    return 0;
}

extern "C" int lmgx_norms_device(lmgx_handle *h, int level, const double *d_x, const double *d_b,
                                 double *d_res, double norms_out[4])
{
    CHK(check_level(h, level));
    if (!h->have_op) return fail(LMGX_E_STATE, "no operator");
    CHK(ensure_natural(h));
    CHK(ensure_hist(h, 64));
    CHK(norms_exact(h, level, d_x, d_b, d_res, h->d_scalar));
    double tmp[4];
    CHK(read_norms(h, h->d_scalar, tmp));
    if (norms_out) memcpy(norms_out, tmp, sizeof(tmp));
    return 0;
}

extern "C" int lmgx_sporlx_device(lmgx_handle *h, int level, double *d_f, const double *d_b,
                                  double omega)
{
    CHK(check_level(h, level));
    if (!h->have_op) return fail(LMGX_E_STATE, "no operator");
    CHK(ensure_natural(h));
    CHK(sporlx_exact(h, level, d_f, d_b, omega));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int lmgx_sporlx2_device(lmgx_handle *h, int level, double *d_f, const double *d_b,
                                   double omega)
{
    CHK(check_level(h, level));
    if (!h->have_op) return fail(LMGX_E_STATE, "no operator");
    CHK(ensure_natural(h));
    CHK(sporlx2_exact(h, level, d_f, d_b, omega));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int lmgx_swap_device(lmgx_handle *h, int level, double *d_var)
{
    CHK(check_level(h, level));
    CHK(swap_level<double>(h, level, d_var));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int lmgx_restrict_device(lmgx_handle *h, int lc, const double *d_res, double *d_b,
                                    double *d_f)
{
    CHK(check_level(h, lc));
    if (lc == 0) return fail(LMGX_E_ARG, "coarse level must be >= 1");
    CHK(restrict_exact(h, lc, d_res, d_b, d_f));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int lmgx_prolong_device(lmgx_handle *h, int lf, double *d_f)
{
    CHK(check_level(h, lf));
    if (lf >= h->nlev - 1) return fail(LMGX_E_ARG, "fine level must be < nlev-1");
    CHK(prolong_exact(h, lf, d_f));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int lmgx_extrap_p_device(lmgx_handle *h, double *d_p)
{
    if (!h || !d_p) return fail(LMGX_E_ARG, "NULL argument");
    LevelDev &D = h->lev[0];
    hipLaunchKernelGGL(k_zero_grad<double>, dim3(D.tiles_all), dim3(LMGX_WG), 0, h->stream, D.d_lb,
                       h->nblk, d_p, 1);
    HIPCHK(hipGetLastError());
    CHK(swap_level<double>(h, 0, d_p));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int lmgx_p_ref_subtract_device(lmgx_handle *h, double *d_p)
{
    if (!h || !d_p) return fail(LMGX_E_ARG, "NULL argument");
    const int *n = h->nijkp1[0][0];
    long idx = h->off[0][0] + 1 + (long)n[0] + (long)n[0] * n[1];
    hipLaunchKernelGGL(k_pref_read, dim3(1), dim3(LMGX_WG), 0, h->stream, d_p, idx,
                       h->d_scalar + 40);
    hipLaunchKernelGGL(k_sub_scalar, dim3(nblocks(h->n1)), dim3(LMGX_WG), 0, h->stream, d_p, h->n1,
                       (const double *)(h->d_scalar + 40));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int lmgx_global_sum_device(lmgx_handle *h, int level, const double *d_f,
                                      double out[3])
{
    CHK(check_level(h, level));
    LevelDev &D = h->lev[level];
    hipLaunchKernelGGL(k_gsum, dim3(D.tiles_int), dim3(LMGX_WG), 0, h->stream, D.d_lb, h->nblk, d_f,
                       h->d_partial);
    hipLaunchKernelGGL(k_gsum_final, dim3(1), dim3(LMGX_WG), 0, h->stream,
                       (const double *)h->d_partial, D.tiles_int, h->d_scalar + 44);
    HIPCHK(hipGetLastError());
    CHK(ensure_hist(h, 64));
    HIPCHK(hipMemcpyAsync(h->h_pinned, h->d_scalar + 44, 3 * sizeof(double),
                          hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    memcpy(out, h->h_pinned, 3 * sizeof(double));
    return 0;
}

extern "C" int lmgx_p_mean_subtract_device(lmgx_handle *h, double *d_p, double *mean_out)
{
    if (!h || !d_p) return fail(LMGX_E_ARG, "NULL argument");
    LevelDev &D = h->lev[0];
    // the level-1 interior sum in the fixed two-stage order of k_gsum, then mean = sum / npts
    // and p -= mean over the whole level, halos included (as p_ref_subtract shifts them)
    hipLaunchKernelGGL(k_gsum, dim3(D.tiles_int), dim3(LMGX_WG), 0, h->stream, D.d_lb, h->nblk,
                       (const double *)d_p, h->d_partial);
    hipLaunchKernelGGL(k_gsum_final, dim3(1), dim3(LMGX_WG), 0, h->stream,
                       (const double *)h->d_partial, D.tiles_int, h->d_scalar + 44);
    hipLaunchKernelGGL(k_mean_of, dim3(1), dim3(64), 0, h->stream,
                       (const double *)(h->d_scalar + 44), h->npts[0], h->d_scalar + 47);
    hipLaunchKernelGGL(k_sub_scalar, dim3(nblocks(h->n1)), dim3(LMGX_WG), 0, h->stream, d_p, h->n1,
                       (const double *)(h->d_scalar + 47));
    HIPCHK(hipGetLastError());
    CHK(ensure_hist(h, 64));
    HIPCHK(hipMemcpyAsync(h->h_pinned, h->d_scalar + 47, sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (mean_out) memcpy(mean_out, h->h_pinned, sizeof(double));
    return 0;
}

extern "C" int lmgx_p_mean_subtract_device_noalloc(lmgx_handle *h, double *d_p, double *d_mean)
{
    if (!h || !d_p) return fail(LMGX_E_ARG, "NULL argument");
    if (!h->scratch) return fail(LMGX_E_STATE, "no scratch address for this handle's GPU");
    LevelDev &D = h->lev[0];
    if (D.tiles_int > LMGX_NOALLOC_PARTS)
        return fail(LMGX_E_UNSUPPORTED,
                    "level 1 needs " + std::to_string(D.tiles_int) +
                        " workgroup partials; the static scratch holds " +
                        std::to_string(LMGX_NOALLOC_PARTS) +
                        " (use lmgx_p_mean_subtract_device)");
    // the same kernels as lmgx_p_mean_subtract_device, on the static scratch: every partial
    // the final sum reads is written first by k_gsum, so nothing needs clearing
    double *part = h->scratch, *sums = h->scratch + 3 * LMGX_NOALLOC_PARTS;
    double *mean = d_mean ? d_mean : sums + 3;
    hipLaunchKernelGGL(k_gsum, dim3(D.tiles_int), dim3(LMGX_WG), 0, h->stream, D.d_lb, h->nblk,
                       (const double *)d_p, part);
    hipLaunchKernelGGL(k_gsum_final, dim3(1), dim3(LMGX_WG), 0, h->stream, (const double *)part,
                       D.tiles_int, sums);
    hipLaunchKernelGGL(k_mean_of, dim3(1), dim3(64), 0, h->stream, (const double *)sums,
                       h->npts[0], mean);
    hipLaunchKernelGGL(k_sub_scalar, dim3(nblocks(h->n1)), dim3(LMGX_WG), 0, h->stream, d_p, h->n1,
                       (const double *)mean);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// ---------------------------------------------------------------------------------------------
// smoother profiling (HIP events on the handle's stream around every smoother launch group)
// ---------------------------------------------------------------------------------------------
extern "C" int lmgx_profile(lmgx_handle *h, int enable)
{
    if (!h) return fail(LMGX_E_ARG, "NULL handle");
    h->prof = enable != 0;
    h->prof_ms = 0.0;
    h->prof_launch = 0;
    h->prof_cells = 0.0;
    h->prof_used = 0;
    h->prof_ev_cells.clear();
    return 0;
}

static void prof_begin(lmgx_handle *h)
{
    if (!h->prof) return;
    if (h->prof_used == h->prof_ev.size()) {
        hipEvent_t a, b;
        hipEventCreate(&a);
        hipEventCreate(&b);
        h->prof_ev.push_back({a, b});
    }
    hipEventRecord(h->prof_ev[h->prof_used].first, h->stream);
}

static void prof_end(lmgx_handle *h, double cells)
{
    if (!h->prof) return;
    hipEventRecord(h->prof_ev[h->prof_used].second, h->stream);
    h->prof_ev_cells.push_back(cells);
    ++h->prof_used;
}

static void prof_collect(lmgx_handle *h)
{
    if (!h->prof || h->prof_used == 0) return;
    hipStreamSynchronize(h->stream);
    for (size_t q = 0; q < h->prof_used; ++q) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, h->prof_ev[q].first, h->prof_ev[q].second);
        h->prof_ms += ms;
        h->prof_cells += h->prof_ev_cells[q];
        ++h->prof_launch;
    }
    h->prof_used = 0;
    h->prof_ev_cells.clear();
}

extern "C" int lmgx_profile_read(const lmgx_handle *hc, double *ms, long *launches, double *cells)
{
    lmgx_handle *h = const_cast<lmgx_handle *>(hc);
    if (!h) return fail(LMGX_E_ARG, "NULL handle");
    prof_collect(h);
    if (ms) *ms = h->prof_ms;
    if (launches) *launches = h->prof_launch;
    if (cells) *cells = h->prof_cells;
    return 0;
}

#include "fast_path.inc"
#include "operator_build.inc"
#include "adjoint.inc"
#include "field_io.inc"
