/*
 * FP64 CPU oracle for the NNAP evaluation path of jse (jsex.nnap).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only the tests, the
 * smoke check and the CPU-baseline leg of the benchmark may load it.  The
 * product path (jse_amd/, libjsenn_hip.so) never links, loads or calls it.
 *
 * It restates, in plain C and double precision, the algorithm the reference
 * runs on its CPU path (flt_t = double), each piece citing the reference file
 * it follows (paths relative to the reference repository root):
 *
 *   wrap            src/main/java/jse/atom/IBox.java:145-151 (wrapPBC)
 *   cell grid       src/main/java/jse/atom/NeighborListGetter.java:952-1001
 *   cell binning    NeighborListGetter.java:191-217
 *   27-cell walk    NeighborListGetter.java:265-286 (cell), 330-372, 402-~480
 *   pair filter     NeighborListGetter.java:1018-1054 (dx=(xj+shift)-xi, d2<rc2)
 *   expanded walk   NeighborListGetter.java:704-776, 985-995 (rmax > min box)
 *   Rn, fc          src/nnap/resources/assets/nnap/src/nnap_util.hpp:226-271
 *   Y_lm            basis_SphericalUtil.hpp:12-101 (+ coefficient tables in
 *                   basis_SphericalUtil0.hpp:501-603, recomputed here from
 *                   their closed forms)
 *   dY/dtheta,dphi  basis_SphericalUtil.hpp:103-182
 *   A, fp (L2)      basis_SphericalChebyshev.hpp:264-351, basis_SphericalUtil.hpp:378-394
 *   fp (L3, L4)     basis_SphericalUtil.hpp:395-467 (coefficient tables in
 *                   basis_SphericalUtil0.hpp:10-500, recomputed here from the
 *                   Wigner 3j couplings they hold; column counts
 *                   SphericalChebyshev.java:15-16,39)
 *   wtype mixing W  src/nnap/java/jsex/nnap/basis/WTypeBasis.java:202-304
 *   MLP fwd/bwd     nn_FeedForward.hpp:8-21,131-261; nnap_main.hpp:267-323
 *   fp backward     basis_SphericalChebyshev.hpp:353-525; basis_SphericalUtil.hpp:469-484
 *   accumulation    src/nnap/java/jsex/nnap/NNAP.java:553-594 and
 *                   src/main/java/jse/atom/IPairPotential.java:919-1021
 *
 * Model parameter blocks use the reference's own packing (NNAP.java:226-307):
 *   fp_param[t]   = W[ntypes][P][K]           (WTypeBasis internal rfuse weight)
 *   nn_param[t]   = [hidden W | out W | hidden b | out b]   (FeedForward.java:310-316)
 *   norm_param[t] = [muE+refE, sigmaE, mu[S], sigma[S]]     (NNAP.java:295-306)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <complex.h>
#include <pthread.h>

#define ORC_MAX_LMAX 12
#define ORC_MAX_M ((ORC_MAX_LMAX + 1) * (ORC_MAX_LMAX + 1))
#define ORC_MAX_K 32
#define ORC_MAX_P 128
#define ORC_MAX_LAYERS 8
#define ORC_MAX_H 1024

/* Bases are indexed by basis number b: species t owns bases part0[t] ..
 * part0[t] + nparts[t] - 1 (a merge basis, MergedBasis.java: its fp is the
 * parts' fp concatenated); with part0 / nparts NULL species t owns basis t. */
typedef struct {
    int ntypes;
    const int *nmax;          /* [bases] */
    const int *lmax;          /* [bases] */
    const int *size_np;       /* [bases]  P */
    const double *rcut;       /* [bases] */
    const int *nlayers;       /* [ntypes] */
    const int *hidden;        /* [ntypes][ORC_MAX_LAYERS] */
    const double *const *fp_param;   /* [bases] W[ntypes][P][K] */
    const double *const *nn_param;   /* [ntypes] */
    const double *const *norm_param; /* [ntypes] */
    const int *mirror;        /* [ntypes] mirror basis: the mirrored species, 0 none (NULL: none at all) */
    const int *radial;        /* [bases] 1: radial-only chebyshev basis (lmax 0 sizes), NULL: none */
    const int *part0;         /* [ntypes] first basis of each species (NULL: basis t) */
    const int *nparts;        /* [ntypes] number of bases of each species (NULL: one) */
    const int *l3max;         /* [bases] L3 invariants' lmax (0: none; NULL: none at all) */
    const int *l4max;         /* [bases] L4 invariants' lmax (0: none; NULL: none at all) */
    const int *train_w;       /* [bases] 1: W is a trainable parameter block (fuse, rfuse, exfuse) */
} orc_model;

static int nparts_of(const orc_model *m, int t) { return m->nparts ? m->nparts[t] : 1; }
static int basis_index(const orc_model *m, int t, int k) { return (m->part0 ? m->part0[t] : t) + k; }
/* a species' cutoff: the largest of its bases' (MergedBasis.rcutMax) */
static double species_rcut(const orc_model *m, int t) {
    double r = 0.0;
    for (int k = 0; k < nparts_of(m, t); ++k) {
        const double rk = m->rcut[basis_index(m, t, k)];
        if (rk > r) r = rk;
    }
    return r;
}

/* ------------------------------------------------------------------ */
/* wrap: IBox.wrapPBC for an orthogonal box (IBox.java:145-151).       */
/* s = x/L; if (s<0 || s>=1) s -= floor(s); x' = s*L                   */
/* ------------------------------------------------------------------ */
int orc_wrap(int n, const double *xyz, const double *box, double *out) {
    for (int i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            double s = xyz[3 * i + d] / box[d];
            if (s < 0.0 || s >= 1.0) s -= floor(s);
            out[3 * i + d] = s * box[d];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Cell grid of NeighborListGetter.getProperLinkedCell (:952-1001) for */
/* an orthogonal box and rmax <= min box length.                       */
/* ------------------------------------------------------------------ */
int orc_cell_grid(const double *box, double rmax, int *size) {
    double minbox = box[0];
    if (box[1] < minbox) minbox = box[1];
    if (box[2] < minbox) minbox = box[2];
    if (rmax > minbox) return -1; /* no grid: the expanded walk (neighbors_expanded) */
    int div = (int)floor(minbox / rmax);
    double cell_len = minbox / (double)div;
    for (int d = 0; d < 3; ++d) {
        int s = (int)floor(box[d] / cell_len);
        size[d] = s > div ? s : div;
    }
    return 0;
}

static inline int shift_code(int sx, int sy, int sz) { return (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1); }

/* The fixed 27-image visiting order of LinkedCell.forEachNeighbor
 * (NeighborListGetter.java:345-371) and SingleLinkedCell (:402-~480). */
static const int WALK27[27][3] = {
    {0, 0, 0},  {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
    {1, 1, 0},  {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0}, {0, 1, 1},   {0, 1, -1},  {0, -1, 1},
    {0, -1, -1}, {1, 0, 1},  {-1, 0, 1},  {1, 0, -1},  {-1, 0, -1}, {1, 1, 1},   {1, 1, -1},
    {1, -1, 1}, {1, -1, -1}, {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};

/* emit a candidate: dx = (xj + shift) - xi, kept when dx*dx+dy*dy+dz*dz < rc*rc */
static inline int try_pair(const double *x, const double *box, int i, int j, int sx, int sy, int sz,
                           double rc2, double *d) {
    double xj = x[3 * j], yj = x[3 * j + 1], zj = x[3 * j + 2];
    if (sx) xj = xj + (sx > 0 ? box[0] : -box[0]);
    if (sy) yj = yj + (sy > 0 ? box[1] : -box[1]);
    if (sz) zj = zj + (sz > 0 ? box[2] : -box[2]);
    double dx = xj - x[3 * i];
    double dy = yj - x[3 * i + 1];
    double dz = zj - x[3 * i + 2];
    double a = dx * dx;
    double b = dy * dy;
    double c = dz * dz;
    double d2 = (a + b) + c;
    if (d2 < rc2) {
        d[0] = dx; d[1] = dy; d[2] = dz;
        return 1;
    }
    return 0;
}

/* Image shift past one box length: the key's shift code field holds this
 * marker (the 27 codes of shift_code cover |s| <= 1 only). */
#define EXPAND_CODE 31

/*
 * ExpandLinkedCell (NeighborListGetter.java:704-776; multipliers :985-995),
 * taken when rmax exceeds the smallest box length: multi = ceil(rmax/minbox),
 * mul_d = ceil(multi*minbox / L_d); for atom i, every atom j in index order and
 * every image (a, b, c) in [-mul_x, mul_x] x [-mul_y, mul_y] x [-mul_z, mul_z]
 * (x outermost), the image position x_j + (L_x a, L_y b, L_z c), self skipped
 * only at zero shift, kept when (dx*dx + dy*dy) + dz*dz < rmax^2.
 * key = j*32 + EXPAND_CODE.
 */
static int neighbors_expanded(int n, const double *x, const double *box, double rmax, int cap, int *count,
                              int *key, double *d) {
    double minbox = box[0];
    if (box[1] < minbox) minbox = box[1];
    if (box[2] < minbox) minbox = box[2];
    const int multi = (int)ceil(rmax / minbox);
    const double span = minbox * multi;
    int mul[3];
    for (int q = 0; q < 3; ++q) mul[q] = (int)ceil(span / box[q]);
    const double rc2 = rmax * rmax;
    for (int i = 0; i < n; ++i) {
        int k = 0;
        for (int j = 0; j < n; ++j)
            for (int a = -mul[0]; a <= mul[0]; ++a)
                for (int b = -mul[1]; b <= mul[1]; ++b)
                    for (int c = -mul[2]; c <= mul[2]; ++c) {
                        if (j == i && a == 0 && b == 0 && c == 0) continue;
                        const double px = x[3 * j] + box[0] * a;
                        const double py = x[3 * j + 1] + box[1] * b;
                        const double pz = x[3 * j + 2] + box[2] * c;
                        const double dx = px - x[3 * i], dy = py - x[3 * i + 1], dz = pz - x[3 * i + 2];
                        const double xx = dx * dx, yy = dy * dy, zz = dz * dz;
                        if ((xx + yy) + zz < rc2) {
                            if (k < cap) {
                                key[(size_t)i * cap + k] = j * 32 + EXPAND_CODE;
                                double *o = d + ((size_t)i * cap + k) * 3;
                                o[0] = dx; o[1] = dy; o[2] = dz;
                            }
                            ++k;
                        }
                    }
        count[i] = k;
    }
    return 0;
}

/*
 * Reference-order neighbour walk (cell list).  x must already be wrapped.
 * Outputs per atom i, in the reference's visiting order, up to cap entries:
 *   key[i*cap+k] = j*32 + shift_code, d[(i*cap+k)*3 + c] = dx,dy,dz.
 * count[i] is the true number of neighbours (may exceed cap: caller resizes).
 * rmax above the smallest box length takes the expanded walk above.
 */
int orc_neighbors_celllist(int n, const double *x, const double *box, double rmax, int cap, int *count,
                           int *key, double *d) {
    int size[3];
    if (orc_cell_grid(box, rmax, size) != 0) return neighbors_expanded(n, x, box, rmax, cap, count, key, d);
    const double rc2 = rmax * rmax;
    const int ncell = size[0] * size[1] * size[2];
    int single = (size[0] == 1 && size[1] == 1 && size[2] == 1);
    /* binning in increasing atom index order (NeighborListGetter.java:206-214) */
    int *cnt = (int *)calloc((size_t)ncell + 1, sizeof(int));
    int *cell_of = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *members = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *start = (int *)calloc((size_t)ncell + 1, sizeof(int));
    double cb[3] = {box[0] / size[0], box[1] / size[1], box[2] / size[2]};
    int ci3[3];
    for (int i = 0; i < n; ++i) {
        for (int dd = 0; dd < 3; ++dd) {
            int c = (int)floor(x[3 * i + dd] / cb[dd]);
            /* the reference throws on an out-of-range cell; a wrapped
             * coordinate rounding up to L is clamped into the last cell */
            if (c >= size[dd]) c = size[dd] - 1;
            if (c < 0) c = 0;
            ci3[dd] = c;
        }
        cell_of[i] = ci3[0] + size[0] * ci3[1] + size[0] * size[1] * ci3[2];
        cnt[cell_of[i]]++;
    }
    for (int c = 0; c < ncell; ++c) start[c + 1] = start[c] + cnt[c];
    memset(cnt, 0, sizeof(int) * (size_t)ncell);
    for (int i = 0; i < n; ++i) members[start[cell_of[i]] + cnt[cell_of[i]]++] = i;

    for (int i = 0; i < n; ++i) {
        int k = 0;
        double dd3[3];
        if (single) {
            /* SingleLinkedCell.forEachNeighbor: every atom, 27 images; self only at zero shift skipped */
            for (int j = 0; j < n; ++j) {
                for (int w = 0; w < 27; ++w) {
                    int sx = WALK27[w][0], sy = WALK27[w][1], sz = WALK27[w][2];
                    if (j == i && sx == 0 && sy == 0 && sz == 0) continue;
                    if (try_pair(x, box, i, j, sx, sy, sz, rc2, dd3)) {
                        if (k < cap) {
                            key[(size_t)i * cap + k] = j * 32 + shift_code(sx, sy, sz);
                            memcpy(d + ((size_t)i * cap + k) * 3, dd3, sizeof dd3);
                        }
                        ++k;
                    }
                }
            }
        } else {
            int c = cell_of[i];
            int cx = c % size[0], cy = (c / size[0]) % size[1], cz = c / (size[0] * size[1]);
            for (int w = 0; w < 27; ++w) {
                int ix = cx + WALK27[w][0], iy = cy + WALK27[w][1], iz = cz + WALK27[w][2];
                int sx = 0, sy = 0, sz = 0;
                if (ix >= size[0]) { ix -= size[0]; sx = 1; } else if (ix < 0) { ix += size[0]; sx = -1; }
                if (iy >= size[1]) { iy -= size[1]; sy = 1; } else if (iy < 0) { iy += size[1]; sy = -1; }
                if (iz >= size[2]) { iz -= size[2]; sz = 1; } else if (iz < 0) { iz += size[2]; sz = -1; }
                int cc = ix + size[0] * iy + size[0] * size[1] * iz;
                int mirror = (sx != 0 || sy != 0 || sz != 0);
                for (int q = start[cc]; q < start[cc + 1]; ++q) {
                    int j = members[q];
                    if (!mirror && j == i) continue; /* Cell.forEach skips idx==i; MirrorCell does not */
                    if (try_pair(x, box, i, j, sx, sy, sz, rc2, dd3)) {
                        if (k < cap) {
                            key[(size_t)i * cap + k] = j * 32 + shift_code(sx, sy, sz);
                            memcpy(d + ((size_t)i * cap + k) * 3, dd3, sizeof dd3);
                        }
                        ++k;
                    }
                }
            }
        }
        count[i] = k;
    }
    free(cnt); free(cell_of); free(members); free(start);
    return 0;
}

/* Independent pin of the walk above: every atom, every one of the 27 images,
 * same FP64 expressions, output sorted by key (j*32+code). */
static int cmp_key_d(const void *a, const void *b) {
    const int ka = *(const int *)a, kb = *(const int *)b;
    return (ka > kb) - (ka < kb);
}
int orc_neighbors_bruteforce(int n, const double *x, const double *box, double rmax, int cap, int *count,
                             int *key, double *d) {
    const double rc2 = rmax * rmax;
    /* key + 3 doubles, packed for sorting */
    typedef struct { int key; int pad; double d[3]; } ent;
    ent *buf = (ent *)malloc(sizeof(ent) * (size_t)(n * 27 > 0 ? n * 27 : 1));
    for (int i = 0; i < n; ++i) {
        int k = 0;
        for (int j = 0; j < n; ++j) {
            for (int sz = -1; sz <= 1; ++sz)
                for (int sy = -1; sy <= 1; ++sy)
                    for (int sx = -1; sx <= 1; ++sx) {
                        if (j == i && sx == 0 && sy == 0 && sz == 0) continue;
                        double dd3[3];
                        if (try_pair(x, box, i, j, sx, sy, sz, rc2, dd3)) {
                            buf[k].key = j * 32 + shift_code(sx, sy, sz);
                            memcpy(buf[k].d, dd3, sizeof dd3);
                            ++k;
                        }
                    }
        }
        qsort(buf, (size_t)k, sizeof(ent), cmp_key_d);
        for (int q = 0; q < k && q < cap; ++q) {
            key[(size_t)i * cap + q] = buf[q].key;
            memcpy(d + ((size_t)i * cap + q) * 3, buf[q].d, sizeof(double) * 3);
        }
        count[i] = k;
    }
    free(buf);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Radial basis (nnap_util.hpp:226-271)                                */
/* ------------------------------------------------------------------ */
static void cal_rn(int nmax, double r, double rc, double *rn) {
    double x = 1.0 - 2.0 * (r / rc);
    rn[0] = 1.0;
    if (nmax >= 1) rn[1] = x;
    for (int q = 2; q <= nmax; ++q) rn[q] = 2.0 * x * rn[q - 1] - rn[q - 2];
}
/* -(dRn/dr)/r = 2/(r rc) * n * U_{n-1}(x)  (sign convention of calRnGrad) */
static void cal_rn_grad(int nmax, double r, double rc, double *g) {
    double x = 1.0 - 2.0 * (r / rc);
    double mul = 2.0 / (r * rc);
    double u[ORC_MAX_K];
    u[0] = mul;
    if (nmax >= 2) u[1] = 2.0 * mul * x;
    for (int q = 2; q <= nmax - 1; ++q) u[q] = 2.0 * x * u[q - 1] - u[q - 2];
    g[0] = 0.0;
    for (int q = 1; q <= nmax; ++q) g[q] = u[q - 1] * (double)q;
}
static double cal_fc(double r, double rc) {
    double t = 1.0 - (r / rc) * (r / rc);
    double t2 = t * t;
    return t2 * t2;
}
/* -(dfc/dr)/r = 8 (1-(r/rc)^2)^3 / rc^2 */
static double cal_fc_grad(double r, double rc) {
    double t = 1.0 - (r / rc) * (r / rc);
    return 8.0 * t * t * t / (rc * rc);
}

/* ------------------------------------------------------------------ */
/* Real spherical harmonics (basis_SphericalUtil.hpp:12-101).          */
/* Normalised associated Legendre recurrence with the Condon-Shortley  */
/* phase; coefficients from their closed forms:                        */
/*  A_lm = sqrt((4l^2-1)/(l^2-m^2)),  B_lm = -sqrt(((l-1)^2-m^2)/(4(l-1)^2-1)) */
/*  P_ll = -sqrt((2l+1)/(2l)) sin(t) P_{l-1,l-1};  P_{l,l-1} = sqrt(2l+1) cos(t) P_{l-1,l-1} */
/* Real Y at index l^2+l+m: m>0 sqrt2 P cos(m phi), m<0 sqrt2 P sin(|m| phi). */
/* ------------------------------------------------------------------ */
static void real_sph(int lmax, double ct, double st, double cp, double sp, double *y) {
    double P[ORC_MAX_M];
    /* P indexed l^2+l+m for m>=0 */
    double pll = 0.28209479177387814347403972578039; /* 1/sqrt(4 pi) */
    P[0] = pll;
    for (int l = 1; l <= lmax; ++l) {
        int idx = l * l + l;
        for (int m = 0; m <= l - 2; ++m) {
            double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
            double b = -sqrt((((double)(l - 1) * (l - 1)) - (double)m * m) / (4.0 * (l - 1) * (l - 1) - 1.0));
            P[idx + m] = a * (ct * P[(l - 1) * (l - 1) + (l - 1) + m] + b * P[(l - 2) * (l - 2) + (l - 2) + m]);
        }
        /* m = l-1 and m = l from the diagonal P_{l-1,l-1} */
        P[idx + l - 1] = sqrt(2.0 * l + 1.0) * ct * pll;
        pll *= -sqrt((2.0 * l + 1.0) / (2.0 * l)) * st;
        P[idx + l] = pll;
    }
    double cm = 1.0, sm = 0.0; /* cos(m phi), sin(m phi), m = 0 */
    for (int l = 0; l <= lmax; ++l) y[l * l + l] = P[l * l + l];
    for (int m = 1; m <= lmax; ++m) {
        double c2 = cm * cp - sm * sp;
        double s2 = sm * cp + cm * sp;
        cm = c2; sm = s2;
        for (int l = m; l <= lmax; ++l) {
            int idx = l * l + l;
            y[idx + m] = M_SQRT2 * P[idx + m] * cm;
            y[idx - m] = M_SQRT2 * P[idx + m] * sm;
        }
    }
}

/* angles with the reference's pole convention (calY, calYPthetaPphi) */
typedef struct { double ct, st, cp, sp, dxy; int pole; } angles;
static angles get_angles(double dx, double dy, double dz, double r) {
    angles a;
    a.dxy = hypot(dx, dy);
    a.ct = dz / r;
    a.st = a.dxy / r;
    /* numericEqual(dxy, 0) of nnap_util.hpp:147-152 */
    a.pole = fabs(a.dxy) < 2.2250738585072014E-308 * 8;
    if (a.pole) { a.cp = 1.0; a.sp = 0.0; }
    else { a.cp = dx / a.dxy; a.sp = dy / a.dxy; }
    return a;
}

/* dY/dtheta by the closed form of basis_SphericalUtil.hpp:103-163 */
static void y_dtheta(int lmax, const double *y, double cp, double sp, double *yt) {
    yt[0] = 0.0;
    if (lmax >= 1) {
        yt[1] = -sp * y[2];
        yt[2] = cp * y[3] + sp * y[1];
        yt[3] = -cp * y[2];
    }
    for (int l = 2; l <= lmax; ++l) {
        int idx = l * l + l;
        double mul0 = sqrt((double)l * (l + 1)) * M_SQRT1_2;
        yt[idx] = mul0 * (cp * y[idx + 1] + sp * y[idx - 1]);
        yt[idx + 1] = -mul0 * cp * y[idx];
        yt[idx - 1] = -mul0 * sp * y[idx];
        for (int m = 1; m <= l; ++m) {
            if (m > 1) {
                double t = -0.5 * sqrt((double)(l + m) * (l - m + 1));
                yt[idx + m] = t * (cp * y[idx + m - 1] - sp * y[idx - m + 1]);
                yt[idx - m] = t * (cp * y[idx - m + 1] + sp * y[idx + m - 1]);
            }
            if (m < l) {
                double t = 0.5 * sqrt((double)(l + m + 1) * (l - m));
                yt[idx + m] += t * (cp * y[idx + m + 1] + sp * y[idx - m - 1]);
                yt[idx - m] += t * (cp * y[idx - m - 1] - sp * y[idx + m + 1]);
            }
        }
    }
}

/* exported for the harmonic pins */
int orc_real_sph(int lmax, double dx, double dy, double dz, double *y, double *ytheta, double *yphi) {
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    angles a = get_angles(dx, dy, dz, r);
    real_sph(lmax, a.ct, a.st, a.cp, a.sp, y);
    if (ytheta) y_dtheta(lmax, y, a.cp, a.sp, ytheta);
    if (yphi) {
        for (int l = 0; l <= lmax; ++l)
            for (int m = -l; m <= l; ++m) yphi[l * l + l + m] = a.pole ? 0.0 : -(double)m * y[l * l + l - m];
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* L3 / L4 invariants of one radial channel's A row (index l^2+l+m),   */
/* basis_SphericalUtil.hpp:395-467.  The reference holds them as sparse */
/* coefficient tables; they are the rotation invariants                 */
/*   L3(l1 l2 l3)    = sum_{m1 m2 m3} R(l1 l2 l3)_{m1 m2 m3} A_{l1 m1} A_{l2 m2} A_{l3 m3}  */
/*   L4(l1 l2 l3 l4) = sum_L sum_M u_{LM} v_{LM},                         */
/*     u_{LM} = sum_{m1 m2} R(l1 l2 L)_{m1 m2 M} A_{l1 m1} A_{l2 m2},      */
/*     v_{LM} = sum_{m3 m4} R(l3 l4 L)_{m3 m4 M} A_{l3 m3} A_{l4 m4},      */
/* L over every value both couplings allow, each with weight one, where */
/* R is the Wigner 3j symbol carried to the real harmonics of real_sph: */
/* with Condon-Shortley complex Y^c, Y_{l,m} = sum_k U_{mk} Y^c_{l,k},    */
/*   U_{00} = 1;  m > 0: U_{m,m} = 1/sqrt2, U_{m,-m} = (-1)^m/sqrt2;      */
/*   m < 0 (a = -m): U_{m,-a} = i(-1)^a/sqrt2, U_{m,a} = -i/sqrt2;        */
/*   R(l1 l2 l3)_{m1 m2 m3} = sum_k (l1 l2 l3; k1 k2 k3) conj(U_{m1 k1} U_{m2 k2} U_{m3 k3}). */
/* R is real for l1 + l2 + l3 even and imaginary for odd (then u v is   */
/* real).  The invariants' order within a radial channel's columns,      */
/* after its lmax + 1 L2 values, is the reference's:                     */
/* ------------------------------------------------------------------ */
static const int L3_COLS[7] = {0, 0, 2, 4, 9, 14, 23};
static const int L4_COLS[4] = {0, 1, 3, 9};
static const int L3_LIST[23][3] = {
    {2, 2, 2}, {1, 1, 2},                                                   /* l3max 2 */
    {2, 3, 3}, {1, 2, 3},                                                   /* 3 */
    {4, 4, 4}, {2, 2, 4}, {3, 3, 4}, {2, 4, 4}, {1, 3, 4},                  /* 4 */
    {2, 5, 5}, {4, 5, 5}, {2, 3, 5}, {1, 4, 5}, {3, 4, 5},                  /* 5 */
    {6, 6, 6}, {3, 3, 6}, {4, 4, 6}, {5, 5, 6}, {2, 6, 6}, {4, 6, 6}, {2, 4, 6}, {1, 5, 6}, {3, 5, 6}}; /* 6 */
static const int L4_LIST[9][4] = {
    {1, 1, 1, 1},                                                           /* l4max 1 */
    {2, 2, 2, 2}, {1, 1, 2, 2},                                             /* 2 */
    {3, 3, 3, 3}, {1, 1, 3, 3}, {2, 2, 3, 3}, {1, 1, 1, 3}, {1, 3, 3, 3}, {1, 2, 2, 3}}; /* 3 */

static double FACT[48];
static double w3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0 || j3 < abs(j1 - j2) || j3 > j1 + j2 || abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3)
        return 0.0;
    /* Racah's formula */
    double pre = sqrt(FACT[j1 + j2 - j3] * FACT[j1 - j2 + j3] * FACT[-j1 + j2 + j3] / FACT[j1 + j2 + j3 + 1] *
                      FACT[j1 + m1] * FACT[j1 - m1] * FACT[j2 + m2] * FACT[j2 - m2] * FACT[j3 + m3] * FACT[j3 - m3]);
    int kmin = 0, kmax = j1 + j2 - j3;
    if (j2 - j3 - m1 > kmin) kmin = j2 - j3 - m1;
    if (j1 - j3 + m2 > kmin) kmin = j1 - j3 + m2;
    if (j1 - m1 < kmax) kmax = j1 - m1;
    if (j2 + m2 < kmax) kmax = j2 + m2;
    double s = 0.0;
    for (int k = kmin; k <= kmax; ++k)
        s += ((k & 1) ? -1.0 : 1.0) / (FACT[k] * FACT[j1 + j2 - j3 - k] * FACT[j1 - m1 - k] * FACT[j2 + m2 - k] *
                                        FACT[j3 - j2 + m1 + k] * FACT[j3 - j1 - m2 + k]);
    return (((j1 - j2 - m3) % 2) ? -1.0 : 1.0) * pre * s;
}
/* the (up to two) complex components k of real harmonic m */
static int u_row(int m, int *k, double complex *u) {
    if (m == 0) { k[0] = 0; u[0] = 1.0; return 1; }
    const int a = abs(m);
    const double h = M_SQRT1_2, sg = (a & 1) ? -1.0 : 1.0;
    if (m > 0) { k[0] = a; u[0] = h; k[1] = -a; u[1] = sg * h; }
    else { k[0] = -a; u[0] = I * sg * h; k[1] = a; u[1] = -I * h; }
    return 2;
}
/* R(l1 l2 l3) as [(2l1+1)][(2l2+1)][(2l3+1)], m = index - l */
static void real_coupling(int l1, int l2, int l3, double complex *R) {
    const int n1 = 2 * l1 + 1, n2 = 2 * l2 + 1, n3 = 2 * l3 + 1;
    for (int a = 0; a < n1; ++a)
        for (int b = 0; b < n2; ++b)
            for (int c = 0; c < n3; ++c) {
                int k1[2], k2[2], k3[2];
                double complex u1[2], u2[2], u3[2], s = 0.0;
                const int r1 = u_row(a - l1, k1, u1), r2 = u_row(b - l2, k2, u2), r3 = u_row(c - l3, k3, u3);
                for (int x = 0; x < r1; ++x)
                    for (int y = 0; y < r2; ++y)
                        for (int z = 0; z < r3; ++z)
                            if (k1[x] + k2[y] + k3[z] == 0)
                                s += w3j(l1, l2, l3, k1[x], k2[y], k3[z]) * conj(u1[x]) * conj(u2[y]) * conj(u3[z]);
                R[(a * n2 + b) * n3 + c] = s;
            }
}
/* couplings computed once per process: R3[k] for the L3 list, R4[k][L]
 * (the pair (l1 l2 L) then (l3 l4 L)) for the L4 list */
#define ORC_MAX_L34 6
static double complex *R3[23];
static double complex *R4[9][2 * ORC_MAX_L34 + 1][2];
static pthread_once_t coupling_once = PTHREAD_ONCE_INIT;
static void coupling_init(void) {
    FACT[0] = 1.0;
    for (int k = 1; k < 48; ++k) FACT[k] = FACT[k - 1] * k;
    for (int k = 0; k < 23; ++k) {
        const int *l = L3_LIST[k];
        R3[k] = (double complex *)malloc(sizeof(double complex) * (2 * l[0] + 1) * (2 * l[1] + 1) * (2 * l[2] + 1));
        real_coupling(l[0], l[1], l[2], R3[k]);
    }
    for (int k = 0; k < 9; ++k) {
        const int *l = L4_LIST[k];
        for (int L = 0; L <= 2 * ORC_MAX_L34; ++L) {
            R4[k][L][0] = R4[k][L][1] = NULL;
            if (L < abs(l[0] - l[1]) || L > l[0] + l[1] || L < abs(l[2] - l[3]) || L > l[2] + l[3]) continue;
            for (int h = 0; h < 2; ++h) {
                const int a = l[2 * h], b = l[2 * h + 1];
                R4[k][L][h] = (double complex *)malloc(sizeof(double complex) * (2 * a + 1) * (2 * b + 1) * (2 * L + 1));
                real_coupling(a, b, L, R4[k][L][h]);
            }
        }
    }
}
/* u_{LM} of one pair (a, b) of blocks */
static void pair_coupled(int a, int b, int L, const double complex *R, const double *A, double complex *u) {
    const int na = 2 * a + 1, nb = 2 * b + 1, nL = 2 * L + 1;
    for (int M = 0; M < nL; ++M) u[M] = 0.0;
    for (int x = 0; x < na; ++x)
        for (int y = 0; y < nb; ++y) {
            const double w = A[a * a + x] * A[b * b + y];
            for (int M = 0; M < nL; ++M) u[M] += R[(x * nb + y) * nL + M] * w;
        }
}
/* out[0 .. L3_COLS[l3max] + L4_COLS[l4max] - 1] of one A row */
static void sh_invariants(int l3max, int l4max, const double *A, double *out) {
    pthread_once(&coupling_once, coupling_init);
    int o = 0;
    for (int k = 0; k < L3_COLS[l3max]; ++k) {
        const int *l = L3_LIST[k];
        const int n2 = 2 * l[1] + 1, n3 = 2 * l[2] + 1;
        double s = 0.0;
        for (int a = 0; a < 2 * l[0] + 1; ++a)
            for (int b = 0; b < n2; ++b)
                for (int c = 0; c < n3; ++c)
                    s += creal(R3[k][(a * n2 + b) * n3 + c]) * A[l[0] * l[0] + a] * A[l[1] * l[1] + b] *
                         A[l[2] * l[2] + c];
        out[o++] = s;
    }
    for (int k = 0; k < L4_COLS[l4max]; ++k) {
        const int *l = L4_LIST[k];
        double s = 0.0;
        for (int L = 0; L <= 2 * ORC_MAX_L34; ++L) {
            if (!R4[k][L][0]) continue;
            double complex u[2 * 2 * ORC_MAX_L34 + 1], v[2 * 2 * ORC_MAX_L34 + 1], t = 0.0;
            pair_coupled(l[0], l[1], L, R4[k][L][0], A, u);
            pair_coupled(l[2], l[3], L, R4[k][L][1], A, v);
            for (int M = 0; M < 2 * L + 1; ++M) t += u[M] * v[M];
            s += creal(t);
        }
        out[o++] = s;
    }
}
/* gA += sum_k g[k] d(invariant k)/dA for one A row */
static void sh_invariants_grad(int l3max, int l4max, const double *A, const double *g, double *gA) {
    pthread_once(&coupling_once, coupling_init);
    int o = 0;
    for (int k = 0; k < L3_COLS[l3max]; ++k, ++o) {
        const int *l = L3_LIST[k];
        const int i1 = l[0] * l[0], i2 = l[1] * l[1], i3 = l[2] * l[2];
        const int n2 = 2 * l[1] + 1, n3 = 2 * l[2] + 1;
        for (int a = 0; a < 2 * l[0] + 1; ++a)
            for (int b = 0; b < n2; ++b)
                for (int c = 0; c < n3; ++c) {
                    const double r = g[o] * creal(R3[k][(a * n2 + b) * n3 + c]);
                    gA[i1 + a] += r * A[i2 + b] * A[i3 + c];
                    gA[i2 + b] += r * A[i1 + a] * A[i3 + c];
                    gA[i3 + c] += r * A[i1 + a] * A[i2 + b];
                }
    }
    for (int k = 0; k < L4_COLS[l4max]; ++k, ++o) {
        const int *l = L4_LIST[k];
        for (int L = 0; L <= 2 * ORC_MAX_L34; ++L) {
            if (!R4[k][L][0]) continue;
            double complex uv[2][2 * 2 * ORC_MAX_L34 + 1];
            pair_coupled(l[0], l[1], L, R4[k][L][0], A, uv[0]);
            pair_coupled(l[2], l[3], L, R4[k][L][1], A, uv[1]);
            const int nL = 2 * L + 1;
            /* d(sum_M u_M v_M) = sum_M (du_M v_M + u_M dv_M) */
            for (int h = 0; h < 2; ++h) {
                const int a = l[2 * h], b = l[2 * h + 1], nb = 2 * b + 1;
                const double complex *R = R4[k][L][h], *other = uv[1 - h];
                for (int x = 0; x < 2 * a + 1; ++x)
                    for (int y = 0; y < nb; ++y) {
                        double complex t = 0.0;
                        for (int M = 0; M < nL; ++M) t += R[(x * nb + y) * nL + M] * other[M];
                        const double r = g[o] * creal(t);
                        gA[a * a + x] += r * A[b * b + y];
                        gA[b * b + y] += r * A[a * a + x];
                    }
            }
        }
    }
}
/* exported for the invariant pins: out has L3_COLS[l3max] + L4_COLS[l4max]
 * values; gA (of the A row's size) receives sum_k g[k] d out[k] / dA */
int orc_sh_invariants(int l3max, int l4max, const double *A, double *out) {
    if (l3max < 0 || l3max > 6 || l4max < 0 || l4max > 3) return 1;
    sh_invariants(l3max, l4max, A, out);
    return 0;
}
int orc_sh_invariants_grad(int l3max, int l4max, const double *A, const double *g, double *gA) {
    if (l3max < 0 || l3max > 6 || l4max < 0 || l4max > 3) return 1;
    sh_invariants_grad(l3max, l4max, A, g, gA);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Per-atom model pieces                                               */
/* ------------------------------------------------------------------ */
typedef struct {
    int K, P, M, L1, S, nl, radial;
    int l3, l4, lmm, SL;   /* L3 / L4 lmax, the harmonics' lmax, columns per radial channel */
    double rc;
    const double *W;
    const int *h;
    const double *nn, *norm;
} sp_view;

/* view of basis k of species ctype (S: that basis' size) */
static sp_view basis_view(const orc_model *m, int ctype, int k) {
    sp_view v;
    int t = ctype - 1, b = basis_index(m, t, k);
    v.K = m->nmax[b] + 1;
    v.P = m->size_np[b];
    v.L1 = m->lmax[b] + 1;
    v.l3 = m->l3max ? m->l3max[b] : 0;
    v.l4 = m->l4max ? m->l4max[b] : 0;
    /* A and Y run to max(lmax, l3max, l4max) (SphericalChebyshev.java:36-42) */
    v.lmm = v.L1 - 1;
    if (v.l3 > v.lmm) v.lmm = v.l3;
    if (v.l4 > v.lmm) v.lmm = v.l4;
    v.M = (v.lmm + 1) * (v.lmm + 1);
    v.SL = v.L1 + L3_COLS[v.l3] + L4_COLS[v.l4];
    v.S = v.P * v.SL;
    v.rc = m->rcut[b];
    v.W = m->fp_param[b];
    v.nl = m->nlayers[t];
    v.h = m->hidden + t * ORC_MAX_LAYERS;
    v.nn = m->nn_param[t];
    v.norm = m->norm_param[t];
    v.radial = m->radial ? m->radial[b] : 0;
    return v;
}
/* the species' view for its network: S = the size of all its bases */
static sp_view species(const orc_model *m, int ctype) {
    sp_view v = basis_view(m, ctype, 0);
    for (int k = 1; k < nparts_of(m, ctype - 1); ++k) v.S += basis_view(m, ctype, k).S;
    return v;
}

/* Radial-only chebyshev basis (basis_Chebyshev.hpp:9-33, Chebyshev.java):
 * fp[p] = sum_j fc * Rnp[p], Rnp = W[type_j] Rn. */
static void descriptor_radial(const sp_view *v, const double *dx, const double *dy, const double *dz,
                              const int *ty, int nn, double *fp) {
    double rn[ORC_MAX_K];
    for (int p = 0; p < v->P; ++p) fp[p] = 0.0;
    for (int j = 0; j < nn; ++j) {
        double r = sqrt(dx[j] * dx[j] + dy[j] * dy[j] + dz[j] * dz[j]);
        if (r >= v->rc) continue;
        cal_rn(v->K - 1, r, v->rc, rn);
        double fc = cal_fc(r, v->rc);
        const double *W = v->W + (size_t)(ty[j] - 1) * v->P * v->K;
        for (int p = 0; p < v->P; ++p) {
            double rnp = 0.0;
            for (int q = 0; q < v->K; ++q) rnp += rn[q] * W[p * v->K + q];
            fp[p] += fc * rnp;
        }
    }
}
/* its backward (basis_Chebyshev.hpp:35-80): g_j = -dE_i/d(dx_j) */
static void descriptor_radial_backward(const sp_view *v, const double *dx, const double *dy, const double *dz,
                                       const int *ty, int nn, const double *gfp, double *gx, double *gy,
                                       double *gz) {
    double rn[ORC_MAX_K], rng[ORC_MAX_K], grn[ORC_MAX_K];
    for (int j = 0; j < nn; ++j) {
        gx[j] = gy[j] = gz[j] = 0.0;
        double r = sqrt(dx[j] * dx[j] + dy[j] * dy[j] + dz[j] * dz[j]);
        if (r >= v->rc) continue;
        cal_rn(v->K - 1, r, v->rc, rn);
        double fc = cal_fc(r, v->rc);
        const double *W = v->W + (size_t)(ty[j] - 1) * v->P * v->K;
        double gfc = 0.0;
        for (int q = 0; q < v->K; ++q) grn[q] = 0.0;
        for (int p = 0; p < v->P; ++p) {
            double rnp = 0.0;
            for (int q = 0; q < v->K; ++q) rnp += rn[q] * W[p * v->K + q];
            gfc += gfp[p] * rnp;
            for (int q = 0; q < v->K; ++q) grn[q] += gfp[p] * fc * W[p * v->K + q];
        }
        cal_rn_grad(v->K - 1, r, v->rc, rng);
        double gr = gfc * cal_fc_grad(r, v->rc);
        for (int q = 0; q < v->K; ++q) gr += grn[q] * rng[q];
        gx[j] = gr * dx[j];
        gy[j] = gr * dy[j];
        gz[j] = gr * dz[j];
    }
}

/* A[p][lm] = sum_j fc*Rnp[p]*Y[lm], Rnp = W[type_j] Rn;  fp[p*SL+l] = 4pi/(2l+1) sum_m A^2 (l <= lmax),
 * then the L3 and L4 invariants of A[p] (basis_SphericalChebyshev.hpp:344-350) */
static void descriptor(const sp_view *v, int lmax, const double *dx, const double *dy, const double *dz,
                       const int *ty, int nn, double *A, double *fp) {
    memset(A, 0, sizeof(double) * (size_t)v->P * v->M);
    double rn[ORC_MAX_K], y[ORC_MAX_M];
    for (int j = 0; j < nn; ++j) {
        double r = sqrt(dx[j] * dx[j] + dy[j] * dy[j] + dz[j] * dz[j]);
        if (r >= v->rc) continue;
        angles a = get_angles(dx[j], dy[j], dz[j], r);
        real_sph(v->lmm, a.ct, a.st, a.cp, a.sp, y);
        cal_rn(v->K - 1, r, v->rc, rn);
        double fc = cal_fc(r, v->rc);
        const double *W = v->W + (size_t)(ty[j] - 1) * v->P * v->K;
        for (int p = 0; p < v->P; ++p) {
            double rnp = 0.0;
            for (int q = 0; q < v->K; ++q) rnp += rn[q] * W[p * v->K + q];
            double c = fc * rnp;
            for (int k = 0; k < v->M; ++k) A[p * v->M + k] += c * y[k];
        }
    }
    for (int p = 0; p < v->P; ++p)
        for (int l = 0; l <= lmax; ++l) {
            double s = 0.0;
            for (int m = -l; m <= l; ++m) {
                double q = A[p * v->M + l * l + l + m];
                s += q * q;
            }
            fp[p * v->SL + l] = (4.0 * M_PI / (2.0 * l + 1.0)) * s;
        }
    if (v->SL > lmax + 1)
        for (int p = 0; p < v->P; ++p) sh_invariants(v->l3, v->l4, A + p * v->M, fp + p * v->SL + lmax + 1);
}

static double silu(double x) { return x / (1.0 + exp(-x)); }
static double silu_grad(double x, double *g) {
    double s = 1.0 / (1.0 + exp(-x));
    *g = s * (1.0 + x * (1.0 - s));
    return x * s;
}

/* normalised MLP: x=(fp-mu)/sigma; E = y*sigmaE + (muE+refE); grad wrt fp returned in gfp */
static double mlp(const sp_view *v, const double *fp, double *gfp) {
    double act[ORC_MAX_LAYERS + 1][ORC_MAX_H];
    double der[ORC_MAX_LAYERS][ORC_MAX_H];
    const double *mu = v->norm + 2, *sig = v->norm + 2 + v->S;
    for (int i = 0; i < v->S; ++i) act[0][i] = (fp[i] - mu[i]) / sig[i];
    /* layout [hidden W (per layer [out][in]) | out W | hidden b | out b] */
    size_t hw = 0, hb = 0;
    int in = v->S;
    for (int l = 0; l < v->nl; ++l) { hw += (size_t)in * v->h[l]; hb += v->h[l]; in = v->h[l]; }
    const double *Wo = v->nn + hw;
    const double *B = Wo + v->h[v->nl - 1];
    const double *Bo = B + hb;
    const double *Wl = v->nn;
    const double *Bl = B;
    in = v->S;
    for (int l = 0; l < v->nl; ++l) {
        int out = v->h[l];
        for (int o = 0; o < out; ++o) {
            double z = Bl[o];
            /* dot then + bias (nn_FeedForward.hpp:136) */
            double dsum = 0.0;
            for (int i = 0; i < in; ++i) dsum += act[l][i] * Wl[(size_t)o * in + i];
            z = dsum + Bl[o];
            act[l + 1][o] = silu_grad(z, &der[l][o]);
        }
        Wl += (size_t)in * out;
        Bl += out;
        in = out;
    }
    double y = 0.0;
    for (int i = 0; i < in; ++i) y += act[v->nl][i] * Wo[i];
    y += Bo[0];
    double E = y * v->norm[1] + v->norm[0];
    if (gfp) {
        double g[ORC_MAX_H], gp[ORC_MAX_H];
        double ge = v->norm[1];
        for (int i = 0; i < in; ++i) g[i] = ge * Wo[i];
        const double *Wb = Wo;
        for (int l = v->nl - 1; l >= 0; --l) {
            int out = v->h[l];
            int inl = (l == 0) ? v->S : v->h[l - 1];
            Wb -= (size_t)inl * out;
            for (int i = 0; i < inl; ++i) gp[i] = 0.0;
            for (int o = 0; o < out; ++o) {
                double gz = g[o] * der[l][o];
                for (int i = 0; i < inl; ++i) gp[i] += gz * Wb[(size_t)o * inl + i];
            }
            for (int i = 0; i < inl; ++i) g[i] = gp[i];
        }
        for (int i = 0; i < v->S; ++i) gfp[i] = g[i] / sig[i];
    }
    return E;
}

/* dE/dA of one basis from dE/dfp: the L2 part 2 4pi/(2l+1) gfp A (calGradSphL2,
 * basis_SphericalUtil.hpp:469-484) and the L3 / L4 invariants' derivatives */
static void basis_grad_A(const sp_view *v, int lmax, const double *A, const double *gfp, double *gA) {
    for (int k = 0; k < v->P * v->M; ++k) gA[k] = 0.0;
    for (int p = 0; p < v->P; ++p) {
        for (int l = 0; l <= lmax; ++l) {
            double mul = (2.0 * 4.0 * M_PI / (2.0 * l + 1.0)) * gfp[p * v->SL + l];
            for (int m = -l; m <= l; ++m) gA[p * v->M + l * l + l + m] = mul * A[p * v->M + l * l + l + m];
        }
        if (v->SL > lmax + 1)
            sh_invariants_grad(v->l3, v->l4, A + p * v->M, gfp + p * v->SL + lmax + 1, gA + p * v->M);
    }
}

/* backward of descriptor: g_j = -dE_i/d(dx_j)  (basis_SphericalChebyshev.hpp:353-465) */
static void descriptor_backward(const sp_view *v, int lmax, const double *dx, const double *dy,
                                const double *dz, const int *ty, int nn, const double *A, const double *gfp,
                                double *gx, double *gy, double *gz) {
    double gA[ORC_MAX_P * ORC_MAX_M];
    basis_grad_A(v, lmax, A, gfp, gA);
    double rn[ORC_MAX_K], rng[ORC_MAX_K], y[ORC_MAX_M], yt[ORC_MAX_M];
    double rnp[ORC_MAX_P], grn[ORC_MAX_K], gy_[ORC_MAX_M];
    for (int j = 0; j < nn; ++j) {
        gx[j] = gy[j] = gz[j] = 0.0;
        double r = sqrt(dx[j] * dx[j] + dy[j] * dy[j] + dz[j] * dz[j]);
        if (r >= v->rc) continue;
        angles a = get_angles(dx[j], dy[j], dz[j], r);
        real_sph(v->lmm, a.ct, a.st, a.cp, a.sp, y);
        cal_rn(v->K - 1, r, v->rc, rn);
        double fc = cal_fc(r, v->rc);
        const double *W = v->W + (size_t)(ty[j] - 1) * v->P * v->K;
        for (int p = 0; p < v->P; ++p) {
            double s = 0.0;
            for (int q = 0; q < v->K; ++q) s += rn[q] * W[p * v->K + q];
            rnp[p] = s;
        }
        double gfc = 0.0;
        for (int k = 0; k < v->M; ++k) gy_[k] = 0.0;
        for (int q = 0; q < v->K; ++q) grn[q] = 0.0;
        for (int p = 0; p < v->P; ++p) {
            double fcrnp = fc * rnp[p];
            double gfcrnp = 0.0;
            for (int k = 0; k < v->M; ++k) {
                gy_[k] += gA[p * v->M + k] * fcrnp;
                gfcrnp += gA[p * v->M + k] * y[k];
            }
            gfc += gfcrnp * rnp[p];
            double grnp = gfcrnp * fc;
            for (int q = 0; q < v->K; ++q) grn[q] += grnp * W[p * v->K + q];
        }
        cal_rn_grad(v->K - 1, r, v->rc, rng);
        double fcg = cal_fc_grad(r, v->rc);
        y_dtheta(v->lmm, y, a.cp, a.sp, yt);
        double gr = 0.0;
        for (int q = 0; q < v->K; ++q) gr += grn[q] * rng[q];
        gr += gfc * fcg;
        double gth = 0.0, gph = 0.0;
        for (int l = 0; l <= v->lmm; ++l)
            for (int m = -l; m <= l; ++m) {
                int k = l * l + l + m;
                gth += gy_[k] * yt[k];
                gph += gy_[k] * (-(double)m * y[l * l + l - m]);
            }
        if (a.pole) gph = 0.0;
        double tpx = -a.ct * a.cp / r, tpy = -a.ct * a.sp / r, tpz = a.st / r;
        double ppx = a.pole ? 0.0 : (a.sp / a.dxy), ppy = a.pole ? 0.0 : (-a.cp / a.dxy);
        gx[j] = gr * dx[j] + gth * tpx + gph * ppx;
        gy[j] = gr * dy[j] + gth * tpy + gph * ppy;
        gz[j] = gr * dz[j] + gth * tpz;
    }
}

/* A mirror basis (MirrorBasis.java; nnap_main.hpp:40-47,73-79): for the
 * centres of species ctype, neighbours of the mirrored species and of ctype
 * swap roles while the basis is evaluated.  Returns the types to use (ty
 * itself, or buf holding the swapped copy). */
static const int *role_types(const orc_model *m, int ctype, const int *ty, int nn, int *buf) {
    const int mir = m->mirror ? m->mirror[ctype - 1] : 0;
    if (!mir) return ty;
    for (int k = 0; k < nn; ++k) buf[k] = ty[k] == mir ? ctype : (ty[k] == ctype ? mir : ty[k]);
    return buf;
}

/* fp of every basis of species ctype, concatenated in order; A of basis k
 * (spherical) at A + k * ORC_MAX_P * ORC_MAX_M */
static void fp_all(const orc_model *m, int ctype, const double *dx, const double *dy, const double *dz,
                   const int *ty, int nn, double *A, double *fp) {
    int off = 0;
    for (int k = 0; k < nparts_of(m, ctype - 1); ++k) {
        sp_view v = basis_view(m, ctype, k);
        if (v.radial) descriptor_radial(&v, dx, dy, dz, ty, nn, fp + off);
        else descriptor(&v, v.L1 - 1, dx, dy, dz, ty, nn, A + (size_t)k * ORC_MAX_P * ORC_MAX_M, fp + off);
        off += v.S;
    }
}
/* g_j summed over the bases (each basis' backward with its slice of gfp) */
static void backward_all(const orc_model *m, int ctype, const double *dx, const double *dy, const double *dz,
                         const int *ty, int nn, const double *A, const double *gfp, double *gx, double *gy,
                         double *gz) {
    const int np = nparts_of(m, ctype - 1);
    double *t = np > 1 ? (double *)malloc(sizeof(double) * 3 * (size_t)(nn > 0 ? nn : 1)) : NULL;
    int off = 0;
    for (int k = 0; k < np; ++k) {
        sp_view v = basis_view(m, ctype, k);
        double *ox = k == 0 ? gx : t, *oy = k == 0 ? gy : t + nn, *oz = k == 0 ? gz : t + 2 * nn;
        if (v.radial) descriptor_radial_backward(&v, dx, dy, dz, ty, nn, gfp + off, ox, oy, oz);
        else descriptor_backward(&v, v.L1 - 1, dx, dy, dz, ty, nn, A + (size_t)k * ORC_MAX_P * ORC_MAX_M,
                                 gfp + off, ox, oy, oz);
        if (k > 0)
            for (int j = 0; j < nn; ++j) {
                gx[j] += ox[j];
                gy[j] += oy[j];
                gz[j] += oz[j];
            }
        off += v.S;
    }
    free(t);
}

/* jse_nnap_calFp restated (nnap_interface.cpp:11-24) */
int orc_cal_fp(const orc_model *m, int ctype, const double *dx, const double *dy, const double *dz,
               const int *ty_in, int nn, double *fp) {
    if (ctype < 1 || ctype > m->ntypes) return 1;
    int *buf = (int *)malloc(sizeof(int) * (size_t)(nn > 0 ? nn : 1));
    const int *ty = role_types(m, ctype, ty_in, nn, buf);
    double *A = (double *)malloc(sizeof(double) * (size_t)nparts_of(m, ctype - 1) * ORC_MAX_P * ORC_MAX_M);
    fp_all(m, ctype, dx, dy, dz, ty, nn, A, fp);
    free(A);
    free(buf);
    return 0;
}
/* jse_nnap_calEnergy restated (nnap_interface.cpp:26-46) */
int orc_cal_energy(const orc_model *m, int ctype, const double *dx, const double *dy, const double *dz,
                   const int *ty_in, int nn, double *e) {
    if (ctype < 1 || ctype > m->ntypes) return 1;
    int *buf = (int *)malloc(sizeof(int) * (size_t)(nn > 0 ? nn : 1));
    const int *ty = role_types(m, ctype, ty_in, nn, buf);
    sp_view v = species(m, ctype);
    double *A = (double *)malloc(sizeof(double) * (size_t)nparts_of(m, ctype - 1) * ORC_MAX_P * ORC_MAX_M);
    double *fp = (double *)malloc(sizeof(double) * (size_t)v.S);
    fp_all(m, ctype, dx, dy, dz, ty, nn, A, fp);
    *e = mlp(&v, fp, NULL);
    free(fp);
    free(A);
    free(buf);
    return 0;
}
/* jse_nnap_calEnergyForce restated (nnap_interface.cpp:48-89): g_j = -dE_i/d(r_j - r_i) */
int orc_cal_energy_force(const orc_model *m, int ctype, const double *dx, const double *dy, const double *dz,
                         const int *ty_in, int nn, double *e, double *gx, double *gy, double *gz) {
    if (ctype < 1 || ctype > m->ntypes) return 1;
    int *buf = (int *)malloc(sizeof(int) * (size_t)(nn > 0 ? nn : 1));
    const int *ty = role_types(m, ctype, ty_in, nn, buf);
    sp_view v = species(m, ctype);
    double *A = (double *)malloc(sizeof(double) * (size_t)nparts_of(m, ctype - 1) * ORC_MAX_P * ORC_MAX_M);
    double *fp = (double *)malloc(sizeof(double) * 2 * (size_t)v.S);
    double *gfp = fp + v.S;
    fp_all(m, ctype, dx, dy, dz, ty, nn, A, fp);
    *e = mlp(&v, fp, gfp);
    backward_all(m, ctype, dx, dy, dz, ty, nn, A, gfp, gx, gy, gz);
    free(fp);
    free(A);
    free(buf);
    return 0;
}
/* the MLP alone on a given fp vector (for the MLP pins) */
int orc_mlp(const orc_model *m, int ctype, const double *fp, double *e, double *gfp) {
    if (ctype < 1 || ctype > m->ntypes) return 1;
    sp_view v = species(m, ctype);
    *e = mlp(&v, fp, gfp);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Whole-system E/F/virial, accumulated as the reference does:         */
/* E[i] += E_i; F[i] -= g; F[j] += g; V[j] += (dx*gx, dy*gy, dz*gz,    */
/* dx*gy, dx*gz, dy*gz, dy*gx, dz*gx, dz*gy)  (NNAP.java:575-592,       */
/* IPairPotential.java:919-1006).  The neighbour walk is the cell list */
/* in the reference's order, at the centre species' rcut.              */
/* Virial component order: xx yy zz xy xz yz yx zx zy.                 */
/* ------------------------------------------------------------------ */
int orc_system_efv(const orc_model *m, int n, const double *xyz_in, const int *type, const double *box,
                   int cap, double *E, double *F, double *V, int *nei_count) {
    double rmax = 0.0;
    for (int t = 0; t < m->ntypes; ++t) if (species_rcut(m, t) > rmax) rmax = species_rcut(m, t);
    double *x = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    orc_wrap(n, xyz_in, box, x);
    int *count = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *key = (int *)malloc(sizeof(int) * (size_t)n * cap + 1);
    double *d = (double *)malloc(sizeof(double) * 3 * (size_t)n * cap + 1);
    int rc = orc_neighbors_celllist(n, x, box, rmax, cap, count, key, d);
    if (rc != 0) { free(x); free(count); free(key); free(d); return rc; }
    for (int i = 0; i < n; ++i) if (count[i] > cap) { free(x); free(count); free(key); free(d); return 2; }
    memset(E, 0, sizeof(double) * (size_t)n);
    memset(F, 0, sizeof(double) * 3 * (size_t)n);
    memset(V, 0, sizeof(double) * 9 * (size_t)n);
    double *dx = (double *)malloc(sizeof(double) * (size_t)cap * 6 + 8);
    double *dy = dx + cap, *dz = dy + cap, *gx = dz + cap, *gy = gx + cap, *gz = gy + cap;
    int *ty = (int *)malloc(sizeof(int) * (size_t)cap + 4);
    int *jj = (int *)malloc(sizeof(int) * (size_t)cap + 4);
    for (int i = 0; i < n; ++i) {
        int ct = type[i];
        if (ct < 1 || ct > m->ntypes) { rc = 3; break; }
        double rci = species_rcut(m, ct - 1);
        double rc2 = rci * rci;
        int nn = 0;
        for (int k = 0; k < count[i]; ++k) {
            const double *dd = d + ((size_t)i * cap + k) * 3;
            if (dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2] >= rc2 && rci < rmax) continue;
            int j = key[(size_t)i * cap + k] >> 5;
            dx[nn] = dd[0]; dy[nn] = dd[1]; dz[nn] = dd[2];
            ty[nn] = type[j]; jj[nn] = j;
            if (ty[nn] < 1 || ty[nn] > m->ntypes) { rc = 3; break; }
            ++nn;
        }
        if (rc) break;
        if (nei_count) nei_count[i] = nn;
        double e;
        orc_cal_energy_force(m, ct, dx, dy, dz, ty, nn, &e, gx, gy, gz);
        E[i] += e;
        for (int k = 0; k < nn; ++k) {
            int j = jj[k];
            F[3 * i] -= gx[k]; F[3 * i + 1] -= gy[k]; F[3 * i + 2] -= gz[k];
            F[3 * j] += gx[k]; F[3 * j + 1] += gy[k]; F[3 * j + 2] += gz[k];
            double *vj = V + 9 * (size_t)j;
            vj[0] += dx[k] * gx[k]; vj[1] += dy[k] * gy[k]; vj[2] += dz[k] * gz[k];
            vj[3] += dx[k] * gy[k]; vj[4] += dx[k] * gz[k]; vj[5] += dy[k] * gz[k];
            vj[6] += dy[k] * gx[k]; vj[7] += dz[k] * gx[k]; vj[8] += dz[k] * gy[k];
        }
    }
    free(x); free(count); free(key); free(d); free(dx); free(ty); free(jj);
    return rc;
}

/* descriptors for every atom of a system (used to fix normalisation constants in tests) */
int orc_system_fp(const orc_model *m, int n, const double *xyz_in, const int *type, const double *box, int cap,
                  int smax, double *fp_out) {
    double rmax = 0.0;
    for (int t = 0; t < m->ntypes; ++t) if (species_rcut(m, t) > rmax) rmax = species_rcut(m, t);
    double *x = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    orc_wrap(n, xyz_in, box, x);
    int *count = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *key = (int *)malloc(sizeof(int) * (size_t)n * cap + 1);
    double *d = (double *)malloc(sizeof(double) * 3 * (size_t)n * cap + 1);
    int rc = orc_neighbors_celllist(n, x, box, rmax, cap, count, key, d);
    double *dx = (double *)malloc(sizeof(double) * (size_t)cap * 3 + 8);
    double *dy = dx + cap, *dz = dy + cap;
    int *ty = (int *)malloc(sizeof(int) * (size_t)cap + 4);
    for (int i = 0; rc == 0 && i < n; ++i) {
        if (count[i] > cap) { rc = 2; break; }
        int nn = 0;
        for (int k = 0; k < count[i]; ++k) {
            const double *dd = d + ((size_t)i * cap + k) * 3;
            dx[nn] = dd[0]; dy[nn] = dd[1]; dz[nn] = dd[2];
            ty[nn] = type[key[(size_t)i * cap + k] >> 5];
            ++nn;
        }
        rc = orc_cal_fp(m, type[i], dx, dy, dz, ty, nn, fp_out + (size_t)i * smax);
    }
    free(x); free(count); free(key); free(d); free(dx); free(ty);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Training derivative of the trainable mixing weights: the rGradFpParam */
/* part of jse_nnap_backwardEnergy (nnap_interface.cpp:111-134) summed    */
/* over a system with per-atom weights w[i] (the trainer's aGradEng).     */
/* For a basis whose W is a parameter block (fuse, rfuse, exfuse; the    */
/* others return at once, basis_SphericalChebyshev.hpp:361-364):          */
/*   dE/dW[t][p][n] += w_i sum_{j of species t} Rn_j[n] dE_i/dRnp_j[p],    */
/*   dE_i/dRnp_j[p] = fc_j sum_m dE_i/dA[p][m] Y_j[m]                      */
/* (backwardMplusAnlm then backwardRnp, nnap_util.hpp:282-296); for the  */
/* radial basis dE_i/dRnp_j[p] = fc_j dE_i/dfp[p] (basis_Chebyshev.hpp:    */
/* 135-220).  grad[t]                                                     */
/* holds, per basis of species t in order, ntypes x P x K values (zeros   */
/* for a basis without parameters), with neighbour species as the basis  */
/* sees them (a mirror basis' swapped roles).                             */
/* ------------------------------------------------------------------ */
int orc_system_fp_param_grad(const orc_model *m, int n, const double *xyz_in, const int *type, const double *box,
                             int cap, const double *w, double *const *grad) {
    double rmax = 0.0;
    for (int t = 0; t < m->ntypes; ++t) if (species_rcut(m, t) > rmax) rmax = species_rcut(m, t);
    for (int t = 0; t < m->ntypes; ++t) {
        size_t sz = 0;
        for (int k = 0; k < nparts_of(m, t); ++k) {
            sp_view v = basis_view(m, t + 1, k);
            sz += (size_t)m->ntypes * v.P * v.K;
        }
        memset(grad[t], 0, sizeof(double) * sz);
    }
    double *x = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    orc_wrap(n, xyz_in, box, x);
    int *count = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *key = (int *)malloc(sizeof(int) * (size_t)n * cap + 1);
    double *d = (double *)malloc(sizeof(double) * 3 * (size_t)n * cap + 1);
    int rc = orc_neighbors_celllist(n, x, box, rmax, cap, count, key, d);
    double *dx = (double *)malloc(sizeof(double) * (size_t)cap * 3 + 8);
    double *dy = dx + cap, *dz = dy + cap;
    int *ty = (int *)malloc(sizeof(int) * (size_t)cap + 4), *buf = (int *)malloc(sizeof(int) * (size_t)cap + 4);
    double *A = (double *)malloc(sizeof(double) * ORC_MAX_P * ORC_MAX_M * 8);
    double *gA = (double *)malloc(sizeof(double) * ORC_MAX_P * ORC_MAX_M);
    for (int i = 0; rc == 0 && i < n; ++i) {
        if (count[i] > cap) { rc = 2; break; }
        const int ct = type[i];
        if (ct < 1 || ct > m->ntypes || nparts_of(m, ct - 1) > 8) { rc = 3; break; }
        const double rci = species_rcut(m, ct - 1);
        int nn = 0;
        for (int k = 0; k < count[i]; ++k) {
            const double *dd = d + ((size_t)i * cap + k) * 3;
            if (dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2] >= rci * rci && rci < rmax) continue;
            dx[nn] = dd[0]; dy[nn] = dd[1]; dz[nn] = dd[2];
            ty[nn] = type[key[(size_t)i * cap + k] >> 5];
            ++nn;
        }
        const int *rt = role_types(m, ct, ty, nn, buf);
        sp_view sv = species(m, ct);
        double *fp = (double *)malloc(sizeof(double) * 2 * (size_t)sv.S);
        double *gfp = fp + sv.S;
        fp_all(m, ct, dx, dy, dz, rt, nn, A, fp);
        mlp(&sv, fp, gfp);
        size_t goff = 0;
        int off = 0;
        for (int k = 0; k < nparts_of(m, ct - 1); ++k) {
            sp_view v = basis_view(m, ct, k);
            const int b = basis_index(m, ct - 1, k);
            if (m->train_w && m->train_w[b]) {
                if (!v.radial) basis_grad_A(&v, v.L1 - 1, A + (size_t)k * ORC_MAX_P * ORC_MAX_M, gfp + off, gA);
                double rn[ORC_MAX_K], y[ORC_MAX_M];
                for (int j = 0; j < nn; ++j) {
                    const double r = sqrt(dx[j] * dx[j] + dy[j] * dy[j] + dz[j] * dz[j]);
                    if (r >= v.rc) continue;
                    angles a = get_angles(dx[j], dy[j], dz[j], r);
                    if (!v.radial) real_sph(v.lmm, a.ct, a.st, a.cp, a.sp, y);
                    cal_rn(v.K - 1, r, v.rc, rn);
                    const double fc = cal_fc(r, v.rc);
                    double *g = grad[ct - 1] + goff + (size_t)(rt[j] - 1) * v.P * v.K;
                    for (int p = 0; p < v.P; ++p) {
                        double s = 0.0;
                        if (v.radial) s = gfp[off + p];
                        else
                            for (int q = 0; q < v.M; ++q) s += gA[p * v.M + q] * y[q];
                        const double grnp = fc * s;
                        for (int q = 0; q < v.K; ++q) g[p * v.K + q] += w[i] * grnp * rn[q];
                    }
                }
            }
            goff += (size_t)m->ntypes * v.P * v.K;
            off += v.S;
        }
        free(fp);
    }
    free(x); free(count); free(key); free(d); free(dx); free(ty); free(buf); free(A); free(gA);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Thread-parallel whole-system evaluation for the CPU baseline, shaped */
/* like the reference's parfor over atoms with per-thread partial       */
/* force/virial vectors summed at the end (IPairPotential.java:895-1021,*/
/* ParforThreadPool.java:76-148).  Same arithmetic as orc_system_efv.   */
/* ------------------------------------------------------------------ */
#ifdef _OPENMP
#include <omp.h>
#endif
int orc_system_efv_par(const orc_model *m, int n, const double *xyz_in, const int *type, const double *box,
                       int cap, int nthreads, double *E, double *F, double *V) {
    double rmax = 0.0;
    for (int t = 0; t < m->ntypes; ++t) if (species_rcut(m, t) > rmax) rmax = species_rcut(m, t);
    double *x = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    orc_wrap(n, xyz_in, box, x);
    int *count = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *key = (int *)malloc(sizeof(int) * (size_t)n * cap + 1);
    double *d = (double *)malloc(sizeof(double) * 3 * (size_t)n * cap + 1);
    int rc = orc_neighbors_celllist(n, x, box, rmax, cap, count, key, d);
    for (int i = 0; rc == 0 && i < n; ++i) if (count[i] > cap) rc = 2;
    if (nthreads < 1) nthreads = 1;
    double *Fp = (double *)calloc((size_t)nthreads * 3 * (size_t)(n > 0 ? n : 1), sizeof(double));
    double *Vp = (double *)calloc((size_t)nthreads * 9 * (size_t)(n > 0 ? n : 1), sizeof(double));
    if (rc == 0) {
        memset(E, 0, sizeof(double) * (size_t)n);
#pragma omp parallel num_threads(nthreads)
        {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            double *Ft = Fp + (size_t)tid * 3 * n, *Vt = Vp + (size_t)tid * 9 * n;
            double *dx = (double *)malloc(sizeof(double) * (size_t)cap * 6 + 8);
            double *dy = dx + cap, *dz = dy + cap, *gx = dz + cap, *gy = gx + cap, *gz = gy + cap;
            int *ty = (int *)malloc(sizeof(int) * (size_t)cap + 4);
            int *jj = (int *)malloc(sizeof(int) * (size_t)cap + 4);
#pragma omp for schedule(dynamic, 16)
            for (int i = 0; i < n; ++i) {
                int nn = count[i];
                for (int k = 0; k < nn; ++k) {
                    const double *dd = d + ((size_t)i * cap + k) * 3;
                    int j = key[(size_t)i * cap + k] >> 5;
                    dx[k] = dd[0]; dy[k] = dd[1]; dz[k] = dd[2];
                    ty[k] = type[j]; jj[k] = j;
                }
                double e;
                orc_cal_energy_force(m, type[i], dx, dy, dz, ty, nn, &e, gx, gy, gz);
                E[i] = e;
                for (int k = 0; k < nn; ++k) {
                    int j = jj[k];
                    Ft[3 * i] -= gx[k]; Ft[3 * i + 1] -= gy[k]; Ft[3 * i + 2] -= gz[k];
                    Ft[3 * j] += gx[k]; Ft[3 * j + 1] += gy[k]; Ft[3 * j + 2] += gz[k];
                    double *vj = Vt + 9 * (size_t)j;
                    vj[0] += dx[k] * gx[k]; vj[1] += dy[k] * gy[k]; vj[2] += dz[k] * gz[k];
                    vj[3] += dx[k] * gy[k]; vj[4] += dx[k] * gz[k]; vj[5] += dy[k] * gz[k];
                    vj[6] += dy[k] * gx[k]; vj[7] += dz[k] * gx[k]; vj[8] += dz[k] * gy[k];
                }
            }
            free(dx); free(ty); free(jj);
        }
        for (size_t q = 0; q < 3 * (size_t)n; ++q) {
            double s = 0.0;
            for (int t = 0; t < nthreads; ++t) s += Fp[(size_t)t * 3 * n + q];
            F[q] = s;
        }
        for (size_t q = 0; q < 9 * (size_t)n; ++q) {
            double s = 0.0;
            for (int t = 0; t < nthreads; ++t) s += Vp[(size_t)t * 9 * n + q];
            V[q] = s;
        }
    }
    free(x); free(count); free(key); free(d); free(Fp); free(Vp);
    return rc;
}
