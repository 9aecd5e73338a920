/* TEST INFRASTRUCTURE ONLY - plain-C restatement of the reference LESter pressure-Poisson path
 * (code/lmg_solve.F and everything it calls, plus coeff_p/sour_p/extrap_p/p_ref_subtract and the
 * geometry that feeds coeff_p).  It is the oracle for the HIP library: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load it, and never as the measured or
 * shipped path.
 *
 * It follows the reference's loop order and expression order point for point, and is built with
 * -ffp-contract=off, so that on x86-64 (no FMA in the reference build either) every field it
 * produces is bit-identical to the reference Fortran; the tests check exactly that against the
 * golden dumps made by oracle/_ref/ref_driver.
 */
#include "lmg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BCT_PERIODIC 12
#define BCT_INTERFACE 13

/* cyclic / cyclic1 axis tables, code/set_const.F:16-28 (1-based values, index [k-1][d-1]) */
static const int CYC[2][3] = {{2, 3, 1}, {3, 1, 2}};
static const int CYC1[2][3] = {{2, 1, 1}, {3, 3, 2}};

static int is_swap_bc(int bc) { return bc == BCT_INTERFACE || bc == BCT_PERIODIC; }

/* ------------------------------------------------------------------------------------------
 * Tables: code/set_offset.F:59-263, code/set_loop_bound.F:53-122 (SV_P => start index 2)
 * ------------------------------------------------------------------------------------------ */
int lmgo_init(lmgo *g, int nblk, int nlev_req, const int *nijkp1, const int *bc,
              const int *bl2bl, const int *bl2side)
{
    memset(g, 0, sizeof(*g));
    if (nblk < 1 || nblk > LMGO_MAXB) return -1;
    g->nblk = nblk;
    for (int b = 0; b < nblk; ++b) {
        for (int d = 0; d < 3; ++d) g->nijkp1[0][b][d] = nijkp1[3 * b + d];
        for (int s = 0; s < 6; ++s) {
            g->bc[b][s] = bc[6 * b + s];
            g->bl2bl[b][s] = bl2bl[6 * b + s];
            g->bl2side[b][s] = bl2side[6 * b + s];
        }
    }
    /* level search, set_offset.F:94-132 (maxmgl = LMGO_MAXL) */
    int found = 1;
    for (int lev = 0; lev < LMGO_MAXL - 1; ++lev) {
        int coarser = 1;
        for (int b = 0; b < nblk; ++b) {
            const int *n = g->nijkp1[lev][b];
            if (n[0] % 2 == 0 && n[1] % 2 == 0 && n[2] % 2 == 0 && coarser) {
                int c0 = n[0] / 2 + 1, c1 = n[1] / 2 + 1, c2 = n[2] / 2 + 1;
                if (c0 >= 4 && c1 >= 4 && c2 >= 4) {
                    g->nijkp1[lev + 1][b][0] = c0;
                    g->nijkp1[lev + 1][b][1] = c1;
                    g->nijkp1[lev + 1][b][2] = c2;
                } else {
                    coarser = 0;
                }
            } else {
                coarser = 0;
            }
        }
        if (coarser) found = lev + 2; else break;
    }
    if (nlev_req == 0) nlev_req = found;
    if (nlev_req > found) return -2;
    g->nlev = nlev_req;
    const int L = g->nlev;
    /* offsets: blocks consecutive within a level, levels consecutive, set_offset.F:166-234 */
    long o = 0, o2 = 0, o1 = 0, oxy = 0, oz = 0;
    for (int l = 0; l < L; ++l)
        for (int b = 0; b < nblk; ++b) {
            const int *n = g->nijkp1[l][b];
            g->off[l][b] = o;
            g->off2d[l][b] = o2;
            g->off1d[l][b] = o1;
            g->offxy[l][b] = oxy;
            g->offz[l][b] = oz;
            o += (long)n[0] * n[1] * n[2];
            o2 += (long)n[0] * n[1];
            o1 += n[2];
            oxy += (long)(n[0] + 1) * (n[1] + 1);
            oz += n[2] + 1;
            if (l == 0) g->n1 = o;
        }
    g->ntot = o;
    g->n2d = o2;
    g->n1d = o1;
    g->nxy = oxy;
    g->nz = oz;
    for (int l = 0; l < L; ++l) {
        g->npts[l] = 0.0;
        for (int b = 0; b < nblk; ++b) {
            const int *n = g->nijkp1[l][b];
            int *se = g->nijkse[l][b];
            se[0] = 2; se[1] = n[0] - 1; se[2] = 2; se[3] = n[1] - 1; se[4] = 2; se[5] = n[2] - 1;
            g->npts[l] = g->npts[l] + (double)((long)(n[0] - 2) * (n[1] - 2) * (n[2] - 2));
            if (l + 1 < L)
                for (int d = 0; d < 3; ++d)
                    g->mult[l][b][d] = (n[d] == g->nijkp1[l + 1][b][d]) ? 1 : 2;
        }
    }
    /* face matching, set_interfaces.F:98-184 (the reference only prints a mismatch) */
    for (int b = 0; b < nblk; ++b)
        for (int s = 1; s <= 6; ++s) {
            int nb = g->bl2bl[b][s - 1], ns = g->bl2side[b][s - 1];
            if (nb == 0) continue;
            if (nb < 1 || nb > nblk || ns < 1 || ns > 6) return -3;
            if (s % 2 == ns % 2) return -4;
            for (int l = 0; l < L; ++l) {
                int d1 = (s + 1) / 2, e1 = (ns + 1) / 2;
                int d2 = CYC[0][d1 - 1], d3 = CYC[1][d1 - 1];
                int e2 = CYC[0][e1 - 1], e3 = CYC[1][e1 - 1];
                if (g->nijkp1[l][b][d2 - 1] != g->nijkp1[l][nb - 1][e2 - 1] ||
                    g->nijkp1[l][b][d3 - 1] != g->nijkp1[l][nb - 1][e3 - 1])
                    return -5;
            }
        }
    g->jumble = 0;
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * Halo swap, code/swap_sc0.F:64-151 (pack every face) then :229-283 (unpack in block order,
 * side 1..6), face tables from code/set_interfaces.F:127-188.
 * ------------------------------------------------------------------------------------------ */
typedef struct { long base_pack, base_unpack, jo, ko; int nj, nk; } face_t;

static void face_geom(const lmgo *g, int l, int b, int s, face_t *f)
{
    int d1 = (s + 1) / 2, d2 = CYC[0][d1 - 1], d3 = CYC[1][d1 - 1];
    const int *n = g->nijkp1[l][b];
    long ijko[3] = {1, n[0], (long)n[0] * n[1]};
    long io = ijko[d1 - 1];
    int ni = n[d1 - 1];
    int ipk = (s % 2 == 0) ? ni - 1 : 2;
    int iun = (s % 2 == 0) ? ni : 1;
    f->jo = ijko[d2 - 1];
    f->ko = ijko[d3 - 1];
    f->nj = n[d2 - 1];
    f->nk = n[d3 - 1];
    f->base_pack = g->off[l][b] + io * (ipk - 1);
    f->base_unpack = g->off[l][b] + io * (iun - 1);
}

void lmgo_swap(const lmgo *g, int l, double *var)
{
    /* slot[b][s] holds what block b packed from its side s */
    static double *buf = NULL;
    static long buflen = 0;
    long need = 0;
    for (int b = 0; b < g->nblk; ++b)
        for (int s = 1; s <= 6; ++s)
            if (is_swap_bc(g->bc[b][s - 1])) {
                face_t f;
                face_geom(g, l, b, s, &f);
                need += (long)f.nj * f.nk;
            }
    if (need > buflen) {
        free(buf);
        buf = (double *)malloc(sizeof(double) * (size_t)need);
        buflen = need;
    }
    long slot[LMGO_MAXB][6];
    long pos = 0;
    for (int b = 0; b < g->nblk; ++b)
        for (int s = 1; s <= 6; ++s) {
            slot[b][s - 1] = -1;
            if (!is_swap_bc(g->bc[b][s - 1])) continue;
            face_t f;
            face_geom(g, l, b, s, &f);
            slot[b][s - 1] = pos;
            for (int k = 0; k < f.nk; ++k)
                for (int j = 0; j < f.nj; ++j) buf[pos++] = var[f.base_pack + f.ko * k + f.jo * j];
        }
    for (int b = 0; b < g->nblk; ++b)
        for (int s = 1; s <= 6; ++s) {
            int nb = g->bl2bl[b][s - 1], ns = g->bl2side[b][s - 1];
            if (nb == 0 || !is_swap_bc(g->bc[b][s - 1])) continue;
            face_t f;
            face_geom(g, l, b, s, &f);
            long p = slot[nb - 1][ns - 1];
            if (p < 0) continue;
            for (int k = 0; k < f.nk; ++k)
                for (int j = 0; j < f.nj; ++j) var[f.base_unpack + f.ko * k + f.jo * j] = buf[p++];
        }
}

/* ------------------------------------------------------------------------------------------
 * Residual and norms, code/norms.F:61-124 and code/norms_block.F:36-64
 * ------------------------------------------------------------------------------------------ */
void lmgo_norms(const lmgo *g, int l, const double *const a[7], const double *x, const double *b,
                double *res, double out[4])
{
    double rsum = 0.0, fsum = 0.0, esum = 0.0, vsum = 0.0;
    for (int bl = 0; bl < g->nblk; ++bl) {
        const int *n = g->nijkp1[l][bl];
        const int *se = g->nijkse[l][bl];
        long o = g->off[l][bl];
        const double *ap = a[0] + o, *aw = a[1] + o, *ae = a[2] + o, *as = a[3] + o,
                     *an = a[4] + o, *al = a[5] + o, *ar = a[6] + o;
        const double *xx = x + o, *bb = b + o;
        double *rr = res + o;
        long jo = n[0], ko = (long)n[0] * n[1];
        double rn = 0.0, fn = 0.0, en = 0.0, vn = 0.0;
        for (int k = se[4]; k <= se[5]; ++k)
            for (int j = se[2]; j <= se[3]; ++j)
                for (int i = se[0]; i <= se[1]; ++i) {
                    long ijk = (i - 1) + jo * (j - 1) + ko * (k - 1);
                    double rhs = bb[ijk];
                    double var = xx[ijk];
                    double r = bb[ijk] - ap[ijk] * var;
                    r = r + aw[ijk] * xx[ijk - 1];
                    r = r + ae[ijk] * xx[ijk + 1];
                    r = r + as[ijk] * xx[ijk - jo];
                    r = r + an[ijk] * xx[ijk + jo];
                    r = r + al[ijk] * xx[ijk - ko];
                    r = r + ar[ijk] * xx[ijk + ko];
                    double err = r / ap[ijk];
                    rn = rn + r * r;
                    fn = fn + rhs * rhs;
                    en = en + err * err;
                    vn = vn + var * var;
                    rr[ijk] = r;
                }
        rsum = rsum + rn;
        fsum = fsum + fn;
        esum = esum + en;
        vsum = vsum + vn;
    }
    double rnp = 1.0 / g->npts[l];
    out[0] = sqrt(rsum * rnp);
    out[1] = sqrt(fsum * rnp);
    out[2] = sqrt(esum * rnp);
    out[3] = sqrt(vsum * rnp);
    if (out[3] == 0.0) out[3] = 1.0;
}

/* ------------------------------------------------------------------------------------------
 * Point SOR, code/sporlx.F:105-148 (axis permutation) and :243-273 (the sweep)
 * ------------------------------------------------------------------------------------------ */
static void spor_block(const lmgo *g, int l, int bl, int d1, const double *const a[7], double *f,
                       const double *src, double omega, double *rs, double *es, double *vs)
{
    const int *n = g->nijkp1[l][bl];
    const int *se = g->nijkse[l][bl];
    long o = g->off[l][bl];
    long st[3] = {1, n[0], (long)n[0] * n[1]};
    /* loop axes (inner, middle, outer) and neighbour-coefficient order per d1 */
    int ax_in, ax_mid, ax_out;
    const double *c_m_in, *c_p_in, *c_m_mid, *c_p_mid, *c_m_out, *c_p_out;
    if (d1 == 1) {
        ax_in = 0; ax_mid = 1; ax_out = 2;
        c_m_in = a[1]; c_p_in = a[2]; c_m_mid = a[3]; c_p_mid = a[4]; c_m_out = a[5]; c_p_out = a[6];
    } else if (d1 == 2) {
        ax_in = 1; ax_mid = 0; ax_out = 2;
        c_m_in = a[3]; c_p_in = a[4]; c_m_mid = a[1]; c_p_mid = a[2]; c_m_out = a[5]; c_p_out = a[6];
    } else {
        ax_in = 2; ax_mid = 0; ax_out = 1;
        c_m_in = a[5]; c_p_in = a[6]; c_m_mid = a[1]; c_p_mid = a[2]; c_m_out = a[3]; c_p_out = a[4];
    }
    long io = st[ax_in], jo = st[ax_mid], ko = st[ax_out];
    const double *ap = a[0] + o;
    c_m_in += o; c_p_in += o; c_m_mid += o; c_p_mid += o; c_m_out += o; c_p_out += o;
    double *x = f + o;
    const double *bb = src + o;
    double rn = 0.0, en = 0.0, vn = 0.0;
    for (int k = se[2 * ax_out]; k <= se[2 * ax_out + 1]; ++k)
        for (int j = se[2 * ax_mid]; j <= se[2 * ax_mid + 1]; ++j)
            for (int i = se[2 * ax_in]; i <= se[2 * ax_in + 1]; ++i) {
                long ijk = io * (i - 1) + jo * (j - 1) + ko * (k - 1);
                double f_old = x[ijk];
                double t = bb[ijk] + c_m_in[ijk] * x[ijk - io];
                t = t + c_p_in[ijk] * x[ijk + io];
                t = t + c_m_mid[ijk] * x[ijk - jo];
                t = t + c_p_mid[ijk] * x[ijk + jo];
                t = t + c_m_out[ijk] * x[ijk - ko];
                t = t + c_p_out[ijk] * x[ijk + ko];
                double dx = t / ap[ijk] - f_old;
                x[ijk] = f_old + omega * dx;
                dx = x[ijk] - f_old;
                double r = dx * ap[ijk];
                rn = rn + r * r;
                en = en + dx * dx;
                vn = vn + x[ijk] * x[ijk];
            }
    *rs = rn; *es = en; *vs = vn;
}

/* inout = (res, rhs, err, var) as in the dummy arguments of sporlx */
void lmgo_sporlx(lmgo *g, int l, const double *const a[7], double *f, const double *b,
                 double omega, double inout[4], double soltol)
{
    double rnpts = 1.0 / g->npts[l];
    int jumfor = (g->jumble % 16 >= 8) ? 1 : 2;
    int jumst = (g->jumble % 24) / 8 + 1;
    int d[3];
    d[0] = jumst;
    d[1] = CYC[((jumfor + 1) % 2)][jumst - 1];
    d[2] = CYC[(jumfor % 2)][jumst - 1];
    int isweep = 0, niter = 1;
    double res_norm = inout[0], err_norm = inout[2], var_norm = inout[3];
    while (isweep < niter && err_norm / var_norm > soltol) {
        ++isweep;
        for (int iline = 0; iline < 3; ++iline) {
            res_norm = 0.0; err_norm = 0.0; var_norm = 0.0;
            for (int bl = 0; bl < g->nblk; ++bl) {
                double rs, es, vs;
                spor_block(g, l, bl, d[iline], a, f, b, omega, &rs, &es, &vs);
                res_norm = res_norm + rs;
                err_norm = err_norm + es;
                var_norm = var_norm + vs;
            }
            lmgo_swap(g, l, f);
        }
        err_norm = sqrt(err_norm * rnpts);
        var_norm = sqrt(var_norm * rnpts);
        if (var_norm == 0.0) var_norm = 1.0;
    }
    res_norm = sqrt(res_norm * rnpts);
    inout[0] = res_norm; inout[2] = err_norm; inout[3] = var_norm;
    g->jumble = g->jumble + 13;
    if (g->jumble > 99999) g->jumble = g->jumble % 99999;
}

/* code/sporlx2.F:64-146: per iteration 7 sweeps, each over every block with spor in natural
 * order (io, jo, ko = 1, nip1, nip1*njp1, i.e. the d1 = 1 summation order W,E,S,N,L,R) and a
 * halo swap after it; the norms are those of the 7th sweep.  jumble is left alone. */
void lmgo_sporlx2(lmgo *g, int l, const double *const a[7], double *f, const double *b,
                  double omega, double inout[4], double soltol)
{
    double rnpts = 1.0 / g->npts[l];
    int isweep = 0, niter = 1;
    double res_norm = inout[0], err_norm = inout[2], var_norm = inout[3];
    while (isweep < niter && err_norm / var_norm > soltol) {
        ++isweep;
        for (int iline = 0; iline < 7; ++iline) {
            res_norm = 0.0; err_norm = 0.0; var_norm = 0.0;
            for (int bl = 0; bl < g->nblk; ++bl) {
                double rs, es, vs;
                spor_block(g, l, bl, 1, a, f, b, omega, &rs, &es, &vs);
                res_norm = res_norm + rs;
                err_norm = err_norm + es;
                var_norm = var_norm + vs;
            }
            lmgo_swap(g, l, f);
        }
        err_norm = sqrt(err_norm * rnpts);
        var_norm = sqrt(var_norm * rnpts);
        if (var_norm == 0.0) var_norm = 1.0;
    }
    res_norm = sqrt(res_norm * rnpts);
    inout[0] = res_norm; inout[2] = err_norm; inout[3] = var_norm;
}

/* ------------------------------------------------------------------------------------------
 * Restriction, code/restriction.F:34-65 + code/rest_block.F:61-76 (mult = 2, fac = 1)
 * ------------------------------------------------------------------------------------------ */
void lmgo_restrict(const lmgo *g, int lc, const double *fine, double *coarse)
{
    int lf = lc - 1;
    for (int bl = 0; bl < g->nblk; ++bl) {
        const int *nc = g->nijkp1[lc][bl], *nf = g->nijkp1[lf][bl];
        long joc = nc[0], koc = (long)nc[0] * nc[1], jof = nf[0], kof = (long)nf[0] * nf[1];
        const double *ff = fine + g->off[lf][bl];
        double *fc = coarse + g->off[lc][bl];
        for (int k = 2; k <= nc[2] - 1; ++k)
            for (int j = 2; j <= nc[1] - 1; ++j)
                for (int i = 2; i <= nc[0] - 1; ++i) {
                    long ijkc = (i - 1) + joc * (j - 1) + koc * (k - 1);
                    long ijkf = (2 * i - 2) + jof * (2 * j - 2) + kof * (2 * k - 2);
                    double s = ff[ijkf] + ff[ijkf - 1];
                    s = s + ff[ijkf - jof];
                    s = s + ff[ijkf - 1 - jof];
                    s = s + ff[ijkf - kof];
                    s = s + ff[ijkf - 1 - kof];
                    s = s + ff[ijkf - jof - kof];
                    s = s + ff[ijkf - 1 - jof - kof];
                    fc[ijkc] = 1.0 * s;
                }
    }
}

void lmgo_init_level(const lmgo *g, int l, double *f, double value)
{
    for (int bl = 0; bl < g->nblk; ++bl) {
        const int *n = g->nijkp1[l][bl];
        long np = (long)n[0] * n[1] * n[2];
        double *p = f + g->off[l][bl];
        for (long i = 0; i < np; ++i) p[i] = value;
    }
}

/* ------------------------------------------------------------------------------------------
 * Prolongation + correction, code/prolongation.F:45-71 and code/prol_block.F:103-247
 * ------------------------------------------------------------------------------------------ */
#define C0 0.015625
#define C1I 0.421875
#define C2I 0.140625
#define C3I 0.046875

void lmgo_prolong(const lmgo *g, int lf, double *f)
{
    int lc = lf + 1;
    for (int bl = 0; bl < g->nblk; ++bl) {
        const int *nc = g->nijkp1[lc][bl], *nf = g->nijkp1[lf][bl];
        const int *bc = g->bc[bl];
        double *fc = f + g->off[lc][bl];
        double *ff = f + g->off[lf][bl];
        long ijkoc[3] = {1, nc[0], (long)nc[0] * nc[1]};
        /* zero-gradient coarse halos on every non-interface side (periodic included) */
        for (int s = 1; s <= 6; ++s) {
            if (bc[s - 1] == BCT_INTERFACE) continue;
            int d1 = (s + 1) / 2, d2 = CYC1[0][d1 - 1], d3 = CYC1[1][d1 - 1];
            int ni = nc[d1 - 1], nj = nc[d2 - 1], nk = nc[d3 - 1];
            long io = ijkoc[d1 - 1], jo = ijkoc[d2 - 1], ko = ijkoc[d3 - 1];
            long ioff, inext;
            if (s % 2 == 1) { ioff = 0; inext = io; } else { ioff = io * (ni - 1); inext = -io; }
            for (int kc = 1; kc <= nk; ++kc)
                for (int jc = 1; jc <= nj; ++jc) {
                    long ijkc = ioff + jo * (jc - 1) + ko * (kc - 1);
                    fc[ijkc] = fc[ijkc + inext];
                }
        }
        long np = (long)nf[0] * nf[1] * nf[2];
        double *cr = (double *)calloc((size_t)np, sizeof(double));
        long joc = nc[0], koc = (long)nc[0] * nc[1], jof = nf[0], kof = (long)nf[0] * nf[1];
        long e = 1, nn = joc, r = koc, ne = 1 + joc, nr = joc + koc, er = 1 + koc, enr = 1 + joc + koc;
        long ef = 1, nf_ = jof, rf = kof, nef = 1 + jof, nrf = jof + kof, erf = 1 + kof,
             enrf = 1 + jof + kof;
        for (int kc = 1; kc <= nc[2] - 1; ++kc) {
            int kf = 2 * kc - 1;
            for (int jc = 1; jc <= nc[1] - 1; ++jc) {
                int jf = 2 * jc - 1;
                for (int ic = 1; ic <= nc[0] - 1; ++ic) {
                    int i_f = 2 * ic - 1;
                    long c = (ic - 1) + joc * (jc - 1) + koc * (kc - 1);
                    long q = (i_f - 1) + jof * (jf - 1) + kof * (kf - 1);
                    const double v0 = fc[c], ve = fc[c + e], vn = fc[c + nn], vr = fc[c + r],
                                 vne = fc[c + ne], vnr = fc[c + nr], ver = fc[c + er],
                                 venr = fc[c + enr];
#define P8(dst, w0, w1, w2, w3, w4, w5, w6, w7)                                              \
    do {                                                                                      \
        double t_ = cr[dst] + w0 * v0;                                                        \
        t_ = t_ + w1 * ve; t_ = t_ + w2 * vn; t_ = t_ + w3 * vr; t_ = t_ + w4 * vne;          \
        t_ = t_ + w5 * vnr; t_ = t_ + w6 * ver; t_ = t_ + w7 * venr;                          \
        cr[dst] = t_;                                                                         \
    } while (0)
                    P8(q, C1I, C2I, C2I, C2I, C3I, C3I, C3I, C0);
                    P8(q + ef, C2I, C1I, C3I, C3I, C2I, C0, C2I, C3I);
                    P8(q + nf_, C2I, C3I, C1I, C3I, C2I, C2I, C0, C3I);
                    P8(q + rf, C2I, C3I, C3I, C1I, C0, C2I, C2I, C3I);
                    P8(q + nef, C3I, C2I, C2I, C0, C1I, C3I, C3I, C2I);
                    P8(q + nrf, C3I, C0, C2I, C2I, C3I, C1I, C3I, C2I);
                    P8(q + erf, C3I, C2I, C0, C2I, C3I, C3I, C1I, C2I);
                    P8(q + enrf, C0, C3I, C3I, C3I, C2I, C2I, C2I, C1I);
#undef P8
                }
            }
        }
        int is = bc[0] == BCT_INTERFACE ? 1 : 2, ie = bc[1] == BCT_INTERFACE ? nf[0] : nf[0] - 1;
        int js = bc[2] == BCT_INTERFACE ? 1 : 2, je = bc[3] == BCT_INTERFACE ? nf[1] : nf[1] - 1;
        int ks = bc[4] == BCT_INTERFACE ? 1 : 2, ke = bc[5] == BCT_INTERFACE ? nf[2] : nf[2] - 1;
        for (int k = ks; k <= ke; ++k)
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie; ++i) {
                    long q = (i - 1) + jof * (j - 1) + kof * (k - 1);
                    ff[q] = ff[q] + cr[q];
                }
        free(cr);
    }
    lmgo_swap(g, lf, f);
}

/* ------------------------------------------------------------------------------------------
 * Schedule: code/smooth.F:34-100, code/solversc.F:32-35, code/lmg_solve.F:44-160
 * ------------------------------------------------------------------------------------------ */
static void smooth(lmgo *g, int l, const double *const a[7], double *f, double *b, double *res,
                   int nsmooth, double omega, double soltol, double nrm[4], int meth_smooth)
{
    if (l == 0 || l == g->nlev - 1) {
        lmgo_norms(g, l, a, f, b, res, nrm);
    } else {
        nrm[2] = 1.0;
        nrm[3] = 1.0;
    }
    int ismooth = 0;
    while (nrm[2] / nrm[3] > soltol && ismooth < nsmooth) {
        ++ismooth;
        if (meth_smooth == 2) lmgo_sporlx2(g, l, a, f, b, omega, nrm, soltol);
        else lmgo_sporlx(g, l, a, f, b, omega, nrm, soltol);
        if (ismooth == nsmooth) lmgo_norms(g, l, a, f, b, res, nrm);
    }
}

int lmgo_solve(lmgo *g, const lmgo_params *p, const double *ap, const double *aw,
               const double *ae, const double *as, const double *an, const double *al,
               const double *ar, double *f, double *b, double *res, double *hist)
{
    const double *a[7] = {ap, aw, ae, as, an, al, ar};
    const int L = g->nlev;
    double nrm[LMGO_MAXL][4];
    memset(nrm, 0, sizeof(nrm));
    for (int call = 0; call < p->ncycles; ++call) {
        int vcycle = p->cycle == 1, fcycle = p->cycle == 2;
        int gamma = vcycle ? 1 : 2;
        double tolc = 0.0;
        if (L == 1) tolc = p->tolf;
        lmgo_norms(g, 0, a, f, b, res, nrm[0]);
        double norm_last = (nrm[0][0] <= 1.0e-30) ? 1.0 : nrm[0][0];
        int nummg[LMGO_MAXL];
        nummg[0] = 1;
        int ilev = 0;
        while (nummg[0] > 0 && nrm[0][2] / nrm[0][3] > p->tolf) {
            int goto_fine = (ilev == L - 1) || (nummg[ilev] == 0);
            if (goto_fine) {
                if (ilev == L - 1) {
                    smooth(g, ilev, a, f, b, res, p->numcits, p->omega_c, tolc, nrm[ilev],
                           p->meth_smooth);
                    if (fcycle) gamma = 1;
                }
                if (ilev > 0 && L != 1) {
                    ilev = ilev - 1;
                    lmgo_prolong(g, ilev, f);
                    smooth(g, ilev, a, f, b, res, p->numpost, p->omega, tolc, nrm[ilev], p->meth_smooth);
                    if (ilev == 0 && fcycle) gamma = 2;
                }
                if (ilev == 0) norm_last = nrm[0][0];
                nummg[ilev] = nummg[ilev] - 1;
            } else {
                smooth(g, ilev, a, f, b, res, p->numpre, p->omega, tolc, nrm[ilev], p->meth_smooth);
                ilev = ilev + 1;
                nummg[ilev] = gamma;
                lmgo_init_level(g, ilev, f, 0.0);
                lmgo_restrict(g, ilev, res, b);
            }
        }
        (void)norm_last;
        if (hist)
            for (int l = 0; l < L; ++l)
                for (int q = 0; q < 4; ++q) hist[((long)call * L + l) * 4 + q] = nrm[l][q];
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * After the solve: code/extrap_p.F:30-44 + code/extrap_p_block.F:43-74, code/p_ref_subtract.F
 * ------------------------------------------------------------------------------------------ */
static void zero_grad_halos(const lmgo *g, int l, double *p)
{
    for (int bl = 0; bl < g->nblk; ++bl) {
        const int *n = g->nijkp1[l][bl];
        long ijko[3] = {1, n[0], (long)n[0] * n[1]};
        double *q = p + g->off[l][bl];
        for (int s = 1; s <= 6; ++s) {
            if (is_swap_bc(g->bc[bl][s - 1])) continue;
            int d1 = (s + 1) / 2, d2 = CYC1[0][d1 - 1], d3 = CYC1[1][d1 - 1];
            long io = ijko[d1 - 1], jo = ijko[d2 - 1], ko = ijko[d3 - 1];
            long io1 = (s % 2 == 0) ? -io : io;
            int i = (s % 2 == 0) ? n[d1 - 1] : 1;
            long ioff = io * (i - 1);
            for (int k = 1; k <= n[d3 - 1]; ++k)
                for (int j = 1; j <= n[d2 - 1]; ++j) {
                    long ijk = ioff + jo * (j - 1) + ko * (k - 1);
                    q[ijk] = q[ijk + io1];
                }
        }
    }
}

void lmgo_extrap_p(const lmgo *g, double *p)
{
    zero_grad_halos(g, 0, p);
    lmgo_swap(g, 0, p);
}

void lmgo_p_ref_subtract(const lmgo *g, double *p)
{
    const int *n = g->nijkp1[0][0];
    double pref = p[g->off[0][0] + 1 + (long)n[0] * 1 + (long)n[0] * n[1] * 1];
    for (int bl = 0; bl < g->nblk; ++bl) {
        const int *m = g->nijkp1[0][bl];
        long np = (long)m[0] * m[1] * m[2];
        double *q = p + g->off[0][bl];
        for (long i = 0; i < np; ++i) q[i] = q[i] - pref;
    }
}

/* code/global_sum.F:125-147: sum, max, min over the interior, blocks in order */
void lmgo_global_sum(const lmgo *g, int l, const double *f, double out[3])
{
    double sum = 0.0, mx = -1.0e30, mn = 1.0e30;
    for (int bl = 0; bl < g->nblk; ++bl) {
        const int *n = g->nijkp1[l][bl];
        const double *q = f + g->off[l][bl];
        double s = 0.0;
        for (int k = 2; k <= n[2] - 1; ++k)
            for (int j = 2; j <= n[1] - 1; ++j)
                for (int i = 2; i <= n[0] - 1; ++i) {
                    double v = q[(i - 1) + (long)n[0] * (j - 1) + (long)n[0] * n[1] * (k - 1)];
                    s = s + v;
                    if (v > mx) mx = v;
                    if (v < mn) mn = v;
                }
        sum = sum + s;
    }
    out[0] = sum; out[1] = mx; out[2] = mn;
}

void lmgo_p_mean_subtract(const lmgo *g, double *p, double *mean_out)
{
    double s[3];
    lmgo_global_sum(g, 0, p, s);
    double mean = s[0] / g->npts[0];
    for (long i = 0; i < g->n1; ++i) p[i] = p[i] - mean;
    if (mean_out) *mean_out = mean;
}

/* ------------------------------------------------------------------------------------------
 * Geometry: code/mg_xyz.F:43-132, code/xyz_extrap0.F:30-56, code/swap_xyz.F (INTERFACE sides
 * 1-4 only), code/metric_block2.F:136-253, code/comp_metric.F:68-252 with
 * code/metric_swap1.F / metric_swap2.F (pack plane 2 / n-1, unpack into plane 1 / n).
 * ------------------------------------------------------------------------------------------ */
static void extrap2d(double *x, int nip2, int njp2)
{
    for (int j = 1; j <= njp2; ++j) {
        double *r = x + (long)nip2 * (j - 1);
        r[0] = 2.0 * r[1] - r[2];
        r[nip2 - 1] = 2.0 * r[nip2 - 2] - r[nip2 - 3];
    }
    for (int i = 1; i <= nip2; ++i) {
        double *c = x + (i - 1);
        c[0] = 2.0 * c[nip2] - c[2L * nip2];
        c[(long)nip2 * (njp2 - 1)] = 2.0 * c[(long)nip2 * (njp2 - 2)] - c[(long)nip2 * (njp2 - 3)];
    }
}

static void extrap1d(double *x, int n2)
{
    x[0] = 2.0 * x[1] - x[2];
    x[n2 - 1] = 2.0 * x[n2 - 2] - x[n2 - 3];
}

/* swap_xyz for x/y on sides 1-4 (all packs, then unpacks in block/side order) */
static void swap_xy(const lmgo *g, int l, double *x)
{
    int maxn = 0;
    for (int b = 0; b < g->nblk; ++b)
        for (int d = 0; d < 3; ++d)
            if (g->nijkp1[l][b][d] + 1 > maxn) maxn = g->nijkp1[l][b][d] + 1;
    double *buf = (double *)calloc((size_t)g->nblk * 4 * maxn, sizeof(double));
    for (int pass = 0; pass < 2; ++pass)
        for (int b = 0; b < g->nblk; ++b)
            for (int s = 1; s <= 4; ++s) {
                if (g->bc[b][s - 1] != BCT_INTERFACE || g->bl2bl[b][s - 1] <= 0) continue;
                int nip2 = g->nijkp1[l][b][0] + 1, njp2 = g->nijkp1[l][b][1] + 1;
                double *xb = x + g->offxy[l][b];
                int even = (s % 2 == 0);
                if (pass == 0) {
                    double *w = buf + ((long)b * 4 + (s - 1)) * maxn;
                    if (s <= 2) {
                        int i = even ? nip2 - 2 : 3;
                        for (int j = 1; j <= njp2; ++j) w[j - 1] = xb[(i - 1) + (long)nip2 * (j - 1)];
                    } else {
                        int j = even ? njp2 - 2 : 3;
                        for (int i = 1; i <= nip2; ++i) w[i - 1] = xb[(i - 1) + (long)nip2 * (j - 1)];
                    }
                } else {
                    int nb = g->bl2bl[b][s - 1] - 1, ns = g->bl2side[b][s - 1];
                    const double *w = buf + ((long)nb * 4 + (ns - 1)) * maxn;
                    if (s <= 2) {
                        int i = even ? nip2 : 1;
                        for (int j = 1; j <= njp2; ++j) xb[(i - 1) + (long)nip2 * (j - 1)] = w[j - 1];
                    } else {
                        int j = even ? njp2 : 1;
                        for (int i = 1; i <= nip2; ++i) xb[(i - 1) + (long)nip2 * (j - 1)] = w[i - 1];
                    }
                }
            }
    free(buf);
}

static void metric_block(int nip1, int njp1, int nkp1, const double *x, const double *y,
                         const double *t, double *diu, double *djv, double *rdiv, double *rdju,
                         double *didjp, double *rdidjp, double *rp, double *dkp, double *dkw)
{
    const int nip2 = nip1 + 1;
    const double rad_on = 0.0, rad_off = 1.0 - rad_on;
#define X(i, j) x[((i) - 1) + (long)nip2 * ((j) - 1)]
#define Y(i, j) y[((i) - 1) + (long)nip2 * ((j) - 1)]
    for (int j = 1; j <= njp1; ++j)
        for (int i = 1; i <= nip1; ++i) {
            int ip1 = i + 1 < nip1 ? i + 1 : nip1, im1 = i - 1 > 1 ? i - 1 : 1;
            int jp1 = j + 1 < njp1 ? j + 1 : njp1, jm1 = j - 1 > 1 ? j - 1 : 1;
            double x_ij = X(i, j), x_ip1j = X(ip1, j), x_ijp1 = X(i, jp1), x_ip1jp1 = X(ip1, jp1),
                   x_im1j = X(im1, j), x_im1jp1 = X(im1, jp1), x_ijm1 = X(i, jm1),
                   x_ip1jm1 = X(ip1, jm1);
            double y_ij = Y(i, j), y_ip1j = Y(ip1, j), y_ijp1 = Y(i, jp1), y_ip1jp1 = Y(ip1, jp1),
                   y_im1j = Y(im1, j), y_im1jp1 = Y(im1, jp1), y_ijm1 = Y(i, jm1),
                   y_ip1jm1 = Y(ip1, jm1);
            double xu_ij = 0.5 * (x_ij + x_ijp1), yu_ij = 0.5 * (y_ij + y_ijp1);
            double xu_ip1j = 0.5 * (x_ip1j + x_ip1jp1), yu_ip1j = 0.5 * (y_ip1j + y_ip1jp1);
            double xv_ij = 0.5 * (x_ij + x_ip1j), yv_ij = 0.5 * (y_ij + y_ip1j);
            double xv_ijp1 = 0.5 * (x_ijp1 + x_ip1jp1), yv_ijp1 = 0.5 * (y_ijp1 + y_ip1jp1);
            double xp_ij = 0.25 * (x_ij + x_ip1j + x_ijp1 + x_ip1jp1);
            double yp_ij = 0.25 * (y_ij + y_ip1j + y_ijp1 + y_ip1jp1);
            double xp_ijm1 = 0.25 * (x_ij + x_ip1j + x_ijm1 + x_ip1jm1);
            double yp_ijm1 = 0.25 * (y_ij + y_ip1j + y_ijm1 + y_ip1jm1);
            double xp_im1j = 0.25 * (x_ij + x_im1j + x_ijp1 + x_im1jp1);
            double yp_im1j = 0.25 * (y_ij + y_im1j + y_ijp1 + y_im1jp1);
            double rp_ij = yp_ij * rad_on + rad_off;
            double ru_ij = yu_ij * rad_on + rad_off;
            double rv_ij = yv_ij * rad_on + rad_off;
            long ij = (i - 1) + (long)nip1 * (j - 1);
            rp[ij] = rp_ij;
            double dip = sqrt((xu_ip1j - xu_ij) * (xu_ip1j - xu_ij) + (yu_ip1j - yu_ij) * (yu_ip1j - yu_ij));
            diu[ij] = sqrt((xp_ij - xp_im1j) * (xp_ij - xp_im1j) + (yp_ij - yp_im1j) * (yp_ij - yp_im1j));
            double div = sqrt((x_ip1j - x_ij) * (x_ip1j - x_ij) + (y_ip1j - y_ij) * (y_ip1j - y_ij));
            rdiv[ij] = rv_ij * div;
            double djp = sqrt((xv_ijp1 - xv_ij) * (xv_ijp1 - xv_ij) + (yv_ijp1 - yv_ij) * (yv_ijp1 - yv_ij));
            double dju = sqrt((x_ijp1 - x_ij) * (x_ijp1 - x_ij) + (y_ijp1 - y_ij) * (y_ijp1 - y_ij));
            djv[ij] = sqrt((xp_ij - xp_ijm1) * (xp_ij - xp_ijm1) + (yp_ij - yp_ijm1) * (yp_ij - yp_ijm1));
            rdju[ij] = ru_ij * dju;
            didjp[ij] = dip * djp;
            rdidjp[ij] = rp_ij * didjp[ij];
        }
#undef X
#undef Y
    for (int k = 1; k <= nkp1; ++k) {
        int kp1 = k + 1 < nkp1 ? k + 1 : nkp1, km1 = k - 1 > 1 ? k - 1 : 1;
        dkp[k - 1] = t[kp1 - 1] - t[k - 1];
        dkw[k - 1] = (t[kp1 - 1] - t[km1 - 1]) * 0.5;
    }
}

#define NMET 7
int lmgo_geometry(const lmgo *g, const double *xv1, const double *yv1, const double *zv1,
                  double *x, double *y, double *z, double *met, double *dkp, double *dkw)
{
    for (int b = 0; b < g->nblk; ++b)
        for (int s = 5; s <= 6; ++s)
            if (g->bc[b][s - 1] == BCT_INTERFACE) return -6; /* k-interfaces: unsupported */
    memset(x, 0, sizeof(double) * g->nxy);
    memset(y, 0, sizeof(double) * g->nxy);
    memset(z, 0, sizeof(double) * g->nz);
    long px = 0, pz = 0;
    for (int b = 0; b < g->nblk; ++b) {
        int nip1 = g->nijkp1[0][b][0], njp1 = g->nijkp1[0][b][1], nkp1 = g->nijkp1[0][b][2];
        double *xb = x + g->offxy[0][b], *yb = y + g->offxy[0][b], *zb = z + g->offz[0][b];
        for (int j = 2; j <= njp1; ++j)
            for (int i = 2; i <= nip1; ++i, ++px) {
                xb[(i - 1) + (long)(nip1 + 1) * (j - 1)] = xv1[px];
                yb[(i - 1) + (long)(nip1 + 1) * (j - 1)] = yv1[px];
            }
        for (int k = 2; k <= nkp1; ++k, ++pz) zb[k - 1] = zv1[pz];
        extrap2d(xb, nip1 + 1, njp1 + 1);
        extrap2d(yb, nip1 + 1, njp1 + 1);
        extrap1d(zb, nkp1 + 1);
    }
    swap_xy(g, 0, x);
    swap_xy(g, 0, y);
    for (int l = 1; l < g->nlev; ++l) {
        for (int b = 0; b < g->nblk; ++b) {
            int nip2c = g->nijkp1[l][b][0] + 1, njp2c = g->nijkp1[l][b][1] + 1,
                nkp2c = g->nijkp1[l][b][2] + 1;
            int nip2f = g->nijkp1[l - 1][b][0] + 1;
            for (int pass = 0; pass < 2; ++pass) {
                double *xc = (pass ? y : x) + g->offxy[l][b];
                const double *xf = (pass ? y : x) + g->offxy[l - 1][b];
                for (int jc = 2; jc <= njp2c - 1; ++jc)
                    for (int ic = 2; ic <= nip2c - 1; ++ic)
                        xc[(ic - 1) + (long)nip2c * (jc - 1)] =
                            xf[(2 * ic - 3) + (long)nip2f * (2 * jc - 3)];
                extrap2d(xc, nip2c, njp2c);
            }
            double *zc = z + g->offz[l][b];
            const double *zf = z + g->offz[l - 1][b];
            for (int kc = 2; kc <= nkp2c - 1; ++kc) zc[kc - 1] = zf[2 * kc - 3];
            extrap1d(zc, nkp2c);
        }
        swap_xy(g, l, x);
        swap_xy(g, l, y);
    }
    const long n2d = g->n2d;
    double *m_diu = met, *m_djv = met + n2d, *m_rdiv = met + 2 * n2d, *m_rdju = met + 3 * n2d,
           *m_didjp = met + 4 * n2d, *m_rdidjp = met + 5 * n2d, *m_rp = met + 6 * n2d;
    for (int l = 0; l < g->nlev; ++l) {
        for (int b = 0; b < g->nblk; ++b) {
            const int *n = g->nijkp1[l][b];
            long o2 = g->off2d[l][b], o1 = g->off1d[l][b];
            metric_block(n[0], n[1], n[2], x + g->offxy[l][b], y + g->offxy[l][b],
                         z + g->offz[l][b], m_diu + o2, m_djv + o2, m_rdiv + o2, m_rdju + o2,
                         m_didjp + o2, m_rdidjp + o2, m_rp + o2, dkp + o1, dkw + o1);
        }
        /* halo metrics across periodic/interface sides: all packs, then unpacks */
        int maxn = 0;
        for (int b = 0; b < g->nblk; ++b)
            for (int d = 0; d < 2; ++d)
                if (g->nijkp1[l][b][d] > maxn) maxn = g->nijkp1[l][b][d];
        const int NSW = 6; /* diu djv rdiv rdju didjp rp (rdidjp is not swapped) */
        double *mp[6] = {m_diu, m_djv, m_rdiv, m_rdju, m_didjp, m_rp};
        double *buf = (double *)calloc((size_t)g->nblk * 6 * (NSW * maxn + 2), sizeof(double));
        long slotsz = NSW * maxn + 2;
        for (int pass = 0; pass < 2; ++pass)
            for (int b = 0; b < g->nblk; ++b)
                for (int s = 1; s <= 6; ++s) {
                    int nb = g->bl2bl[b][s - 1];
                    if (nb <= 0 || !is_swap_bc(g->bc[b][s - 1])) continue;
                    const int *n = g->nijkp1[l][b];
                    long o2 = g->off2d[l][b], o1 = g->off1d[l][b];
                    int ns = g->bl2side[b][s - 1];
                    /* pack goes to the slot of the receiving (partner, partner side) */
                    double *w = pass == 0 ? buf + ((long)(nb - 1) * 6 + (ns - 1)) * slotsz
                                          : buf + ((long)b * 6 + (s - 1)) * slotsz;
                    if (s >= 5) {
                        int kk = pass == 0 ? (s == 5 ? 2 : n[2] - 1) : (s == 5 ? 1 : n[2]);
                        if (pass == 0) { w[0] = dkp[o1 + kk - 1]; w[1] = dkw[o1 + kk - 1]; }
                        else { dkp[o1 + kk - 1] = w[0]; dkw[o1 + kk - 1] = w[1]; }
                        continue;
                    }
                    int line, io, jo, je;
                    if (s == 1) { line = pass == 0 ? 2 : 1; io = 1; jo = n[0]; je = n[1]; }
                    else if (s == 2) { line = pass == 0 ? n[0] - 1 : n[0]; io = 1; jo = n[0]; je = n[1]; }
                    else if (s == 3) { line = pass == 0 ? 2 : 1; io = n[0]; jo = 1; je = n[0]; }
                    else { line = pass == 0 ? n[1] - 1 : n[1]; io = n[0]; jo = 1; je = n[0]; }
                    long ioff = (long)io * (line - 1);
                    for (int j = 1; j <= je; ++j) {
                        long ij = o2 + ioff + (long)jo * (j - 1);
                        for (int m = 0; m < NSW; ++m) {
                            if (pass == 0) w[(long)(j - 1) * NSW + m] = mp[m][ij];
                            else mp[m][ij] = w[(long)(j - 1) * NSW + m];
                        }
                    }
                }
        free(buf);
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * Coefficients, code/coeff_p.F:41-75 + code/coeff_p_block.F:69-228.  Extensions (off when
 * rho == NULL and dirichlet == NULL, which reproduces the reference bit for bit):
 *   rho       every face coefficient divided by 0.5*(rho_P + rho_N) (both cells of the face)
 *   dirichlet on a physical side, the face coefficient stays in ap (face value p = 0) and
 *             the neighbour coefficient is zero.
 * ------------------------------------------------------------------------------------------ */
int lmgo_coeff(const lmgo *g, const double *met, const double *dkp_, const double *dkw_,
               const double *rho, const int *dirichlet, double *ap_, double *aw_, double *ae_,
               double *as_, double *an_, double *al_, double *ar_)
{
    const long n2d = g->n2d;
    for (int l = 0; l < g->nlev; ++l)
        for (int b = 0; b < g->nblk; ++b) {
            const int *n = g->nijkp1[l][b];
            const int *bc = g->bc[b];
            const int nip1 = n[0], njp1 = n[1], nkp1 = n[2];
            const int ni = nip1 - 1, nj = njp1 - 1, nk = nkp1 - 1;
            const long jo = nip1, ko = (long)nip1 * njp1;
            const long o2 = g->off2d[l][b], o1 = g->off1d[l][b], o = g->off[l][b];
            const double *diu = met + o2, *djv = met + n2d + o2, *rdiv = met + 2 * n2d + o2,
                         *rdju = met + 3 * n2d + o2, *didjp = met + 4 * n2d + o2,
                         *rp = met + 6 * n2d + o2;
            const double *dkp = dkp_ + o1, *dkw = dkw_ + o1;
            const double *rh = rho ? rho + o : NULL;
            const int *dir = dirichlet ? dirichlet + 6 * b : NULL;
            double *ap = ap_ + o, *aw = aw_ + o, *ae = ae_ + o, *as = as_ + o, *an = an_ + o,
                   *al = al_ + o, *ar = ar_ + o;
#define M2(a, i, j) a[((i) - 1) + (long)nip1 * ((j) - 1)]
#define IJK(i, j, k) (((i) - 1) + jo * ((j) - 1) + ko * ((k) - 1))
#define RF(p, q) (0.5 * (rh[p] + rh[q]))
            for (int k = 2; k <= nk; ++k)
                for (int j = 2; j <= nj; ++j) ae[IJK(1, j, k)] = 0.0;
            for (int k = 2; k <= nk; ++k)
                for (int i = 2; i <= ni; ++i) an[IJK(i, 1, k)] = 0.0;
            for (int j = 2; j <= nj; ++j)
                for (int i = 2; i <= ni; ++i) ar[IJK(i, j, 1)] = 0.0;
            for (int j = 2; j <= nj; ++j)
                for (int i = 2; i <= ni; ++i) {
                    double cef = M2(rdju, i + 1, j) / M2(diu, i + 1, j);
                    double cnf = M2(rdiv, i, j + 1) / M2(djv, i, j + 1);
                    double crf = M2(didjp, i, j) / M2(rp, i, j);
                    for (int k = 2; k <= nk; ++k) {
                        long ijk = IJK(i, j, k);
                        aw[ijk] = ae[ijk - 1];
                        ae[ijk] = cef * dkp[k - 1];
                        as[ijk] = an[ijk - jo];
                        an[ijk] = cnf * dkp[k - 1];
                        al[ijk] = ar[ijk - ko];
                        ar[ijk] = crf / dkw[k];
                        if (rh) {
                            ae[ijk] = ae[ijk] / RF(ijk, ijk + 1);
                            an[ijk] = an[ijk] / RF(ijk, ijk + jo);
                            ar[ijk] = ar[ijk] / RF(ijk, ijk + ko);
                        }
                        ap[ijk] = aw[ijk] + ae[ijk] + as[ijk] + an[ijk] + al[ijk] + ar[ijk];
                    }
                }
            /* WEST */
            if (is_swap_bc(bc[0]) || (dir && dir[0])) {
                for (int k = 2; k <= nk; ++k)
                    for (int j = 2; j <= nj; ++j) {
                        long ijk = IJK(2, j, k);
                        double c = M2(rdju, 2, j) / M2(diu, 2, j) * dkp[k - 1];
                        if (rh) c = c / RF(ijk, ijk - 1);
                        if (is_swap_bc(bc[0])) aw[ijk] = c;
                        ap[ijk] = ap[ijk] + c;
                    }
            }
            /* EAST */
            if (!is_swap_bc(bc[1]))
                for (int k = 2; k <= nk; ++k)
                    for (int j = 2; j <= nj; ++j) {
                        long ijk = IJK(ni, j, k);
                        if (!(dir && dir[1])) ap[ijk] = ap[ijk] - ae[ijk];
                        ae[ijk] = 0.0;
                    }
            /* SOUTH */
            if (is_swap_bc(bc[2]) || (dir && dir[2])) {
                for (int k = 2; k <= nk; ++k)
                    for (int i = 2; i <= ni; ++i) {
                        long ijk = IJK(i, 2, k);
                        double c = M2(rdiv, i, 2) / M2(djv, i, 2) * dkp[k - 1];
                        if (rh) c = c / RF(ijk, ijk - jo);
                        if (is_swap_bc(bc[2])) as[ijk] = c;
                        ap[ijk] = ap[ijk] + c;
                    }
            }
            /* NORTH */
            if (!is_swap_bc(bc[3]))
                for (int k = 2; k <= nk; ++k)
                    for (int i = 2; i <= ni; ++i) {
                        long ijk = IJK(i, nj, k);
                        if (!(dir && dir[3])) ap[ijk] = ap[ijk] - an[ijk];
                        an[ijk] = 0.0;
                    }
            /* LEFT */
            if (is_swap_bc(bc[4]) || (dir && dir[4])) {
                for (int j = 2; j <= nj; ++j)
                    for (int i = 2; i <= ni; ++i) {
                        long ijk = IJK(i, j, 2);
                        double c = M2(didjp, i, j) / (M2(rp, i, j) * dkw[1]);
                        if (rh) c = c / RF(ijk, ijk - ko);
                        if (is_swap_bc(bc[4])) al[ijk] = c;
                        ap[ijk] = ap[ijk] + c;
                    }
            }
            /* RIGHT */
            if (!is_swap_bc(bc[5]))
                for (int j = 2; j <= nj; ++j)
                    for (int i = 2; i <= ni; ++i) {
                        long ijk = IJK(i, j, nk);
                        if (!(dir && dir[5])) ap[ijk] = ap[ijk] - ar[ijk];
                        ar[ijk] = 0.0;
                    }
#undef M2
#undef IJK
#undef RF
        }
    return 0;
}

void lmgo_rho_levels(const lmgo *g, double *rho)
{
    lmgo_swap(g, 0, rho);
    for (int l = 1; l < g->nlev; ++l) {
        for (int bl = 0; bl < g->nblk; ++bl) {
            const int *nc = g->nijkp1[l][bl], *nf = g->nijkp1[l - 1][bl];
            long joc = nc[0], koc = (long)nc[0] * nc[1], jof = nf[0], kof = (long)nf[0] * nf[1];
            const double *ff = rho + g->off[l - 1][bl];
            double *fc = rho + g->off[l][bl];
            for (int k = 2; k <= nc[2] - 1; ++k)
                for (int j = 2; j <= nc[1] - 1; ++j)
                    for (int i = 2; i <= nc[0] - 1; ++i) {
                        long q = (2 * i - 2) + jof * (2 * j - 2) + kof * (2 * k - 2);
                        double s = ff[q] + ff[q - 1];
                        s = s + ff[q - jof];
                        s = s + ff[q - 1 - jof];
                        s = s + ff[q - kof];
                        s = s + ff[q - 1 - kof];
                        s = s + ff[q - jof - kof];
                        s = s + ff[q - 1 - jof - kof];
                        fc[(i - 1) + joc * (j - 1) + koc * (k - 1)] = 0.125 * s;
                    }
        }
        zero_grad_halos(g, l, rho);
        lmgo_swap(g, l, rho);
    }
}

/* ------------------------------------------------------------------------------------------
 * Source, code/sour_p.F:43-66 + code/sour_p_block.F:89-167 (level 1; u,v,w share its layout)
 * ------------------------------------------------------------------------------------------ */
void lmgo_sour(const lmgo *g, const double *met, const double *dkp_, const double *dkw_,
               const double *u_, const double *v_, const double *w_, const double *rho_,
               const double *rhonp1_, double dtime, double *b)
{
    const long n2d = g->n2d;
    const double ddtime = 2.0 / (3.0 * dtime) * 0.5;
    const double cdtime = 2.0 / (3.0 * dtime);
    const double rdtime = 1.0 / dtime;
    for (int bl = 0; bl < g->nblk; ++bl) {
        const int *n = g->nijkp1[0][bl];
        const int nip1 = n[0], njp1 = n[1], ni = n[0] - 1, nj = n[1] - 1, nk = n[2] - 1;
        const long jo = nip1, ko = (long)nip1 * njp1;
        const long o2 = g->off2d[0][bl], o1 = g->off1d[0][bl], o = g->off[0][bl];
        const double *rdiv = met + 2 * n2d + o2, *rdju = met + 3 * n2d + o2,
                     *didjp = met + 4 * n2d + o2, *rdidjp = met + 5 * n2d + o2;
        const double *dkp = dkp_ + o1, *dkw = dkw_ + o1;
        const double *u = u_ + o, *v = v_ + o, *w = w_ + o, *rho = rho_ + o, *r1 = rhonp1_ + o;
        double *s = b + o;
#define M2(a, i, j) a[((i) - 1) + (long)nip1 * ((j) - 1)]
        for (int j = 2; j <= nj; ++j)
            for (int i = 2; i <= ni; ++i)
                for (int k = 2; k <= nk; ++k) {
                    long ijk = (i - 1) + jo * (j - 1) + ko * (k - 1);
                    double gwf = M2(rdju, i, j) * dkp[k - 1] * (r1[ijk] + r1[ijk - 1]) * u[ijk];
                    double gef = M2(rdju, i + 1, j) * dkp[k - 1] * (r1[ijk] + r1[ijk + 1]) * u[ijk + 1];
                    double gsf = M2(rdiv, i, j) * dkp[k - 1] * (r1[ijk] + r1[ijk - jo]) * v[ijk];
                    double gnf = M2(rdiv, i, j + 1) * dkp[k - 1] * (r1[ijk] + r1[ijk + jo]) * v[ijk + jo];
                    double glf = M2(didjp, i, j) * (r1[ijk] + r1[ijk - ko]) * w[ijk];
                    double grf = M2(didjp, i, j) * (r1[ijk] + r1[ijk + ko]) * w[ijk + ko];
                    double sv = -ddtime * (gef - gwf + gnf - gsf + grf - glf);
                    double vol = M2(rdidjp, i, j) * dkw[k - 1];
                    double drho = (r1[ijk] - rho[ijk]) * rdtime;
                    s[ijk] = sv - cdtime * drho * vol;
                }
#undef M2
    }
}

/* ------------------------------------------------------------------------------------------
 * Velocity correction, code/uvw_adv.F:36-62 + code/uvw_adv_block.F:39-102 (level 1):
 *   u = u_int - 1.5 dt (p_P - p_N) / d / (0.5 (rho_P + rho_N)),  N the lower neighbour along the
 * component's axis, d = diu (U), djv (V) or rp*dkw(k) (W) - step.F:418-425 passes djv as the
 * metric of the V call.  Range per component from code/set_loop_bound.F:64-82 (the first face is
 * skipped on a non-periodic, non-interface lower side); the W loop over i starts at jstr as in
 * uvw_adv_block.F:88 (jstr = istr = 2 for W).  Then swap_sc of u (uvw_adv.F:59).
 * ------------------------------------------------------------------------------------------ */
void lmgo_uvw_adv(const lmgo *g, int ivar, const double *met, const double *dkw_, double *u_,
                  const double *uint_, const double *rho_, const double *p_, double dtime)
{
    const long n2d = g->n2d;
    const double ddtime = 1.5 * dtime;
    for (int bl = 0; bl < g->nblk; ++bl) {
        const int *n = g->nijkp1[0][bl];
        const int nip1 = n[0], njp1 = n[1], ni = n[0] - 1, nj = n[1] - 1, nk = n[2] - 1;
        const int *bc = g->bc[bl];
        const int istr = (ivar == 1 && !is_swap_bc(bc[0])) ? 3 : 2;
        const int jstr = (ivar == 2 && !is_swap_bc(bc[2])) ? 3 : 2;
        const int kstr = (ivar == 3 && !is_swap_bc(bc[4])) ? 3 : 2;
        const long jo = nip1, ko = (long)nip1 * njp1;
        const long o2 = g->off2d[0][bl], o = g->off[0][bl];
        const double *d = met + (ivar == 2 ? 1 : 0) * n2d + o2;   /* diu, or djv for V */
        const double *rp = met + 6 * n2d + o2, *dkw = dkw_ + g->off1d[0][bl];
        double *u = u_ + o;
        const double *ui = uint_ + o, *rho = rho_ + o, *p = p_ + o;
        if (ivar != 3) {
            const long io_or_jo = ivar == 1 ? 1 : jo;
            for (int k = kstr; k <= nk; ++k)
                for (int j = jstr; j <= nj; ++j)
                    for (int i = istr; i <= ni; ++i) {
                        const long ijk = (i - 1) + jo * (j - 1) + ko * (k - 1);
                        const long ijkw = ijk - io_or_jo;
                        const long ij = (i - 1) + (long)nip1 * (j - 1);
                        const double dpdi = (p[ijk] - p[ijkw]) / d[ij];
                        const double rhou = 0.5 * (rho[ijk] + rho[ijkw]);
                        u[ijk] = ui[ijk] - ddtime * dpdi / rhou;
                    }
        } else {
            for (int k = kstr; k <= nk; ++k)
                for (int j = jstr; j <= nj; ++j)
                    for (int i = jstr; i <= ni; ++i) {
                        const long ijk = (i - 1) + jo * (j - 1) + ko * (k - 1);
                        const long ijkl = ijk - ko;
                        const long ij = (i - 1) + (long)nip1 * (j - 1);
                        const double dpdi = (p[ijk] - p[ijkl]) / (rp[ij] * dkw[k - 1]);
                        const double rhou = 0.5 * (rho[ijk] + rho[ijkl]);
                        u[ijk] = ui[ijk] - ddtime * dpdi / rhou;
                    }
        }
    }
    lmgo_swap(g, 0, u_);
}

unsigned long lmgo_sizeof(void) { return (unsigned long)sizeof(lmgo); }
