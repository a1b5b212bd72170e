/*
 * cpab_oracle.c -- CPU restatement of the reference's CPA integration kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; the product path
 * (cpabdiffeo_amd) never links, imports or calls it.
 *
 * Every function restates the arithmetic of the reference, operation for
 * operation and in source order, so that a HIP kernel built with the same
 * rounding (no FMA contraction: this file is compiled with -ffp-contract=off)
 * can be compared bit for bit.  Citations are to freifeld/cpabDiffeo.
 *
 *   cell lookups ........ cpab/gpu/calc_1Dto1D.cu:62-65
 *                         cpab/gpu/calc_2Dto2D.cu:38-153
 *                         cpab/gpu/calc_3Dto3D.cu:43-137
 *   affine products ..... cpab/gpu/calc_2Dto2D.cu:24-36, calc_1Dto1D.cu:18-52,
 *                         calc_3Dto3D.cu:25-39
 *   RK2 (solveODE) ...... cpab/gpu/calc_2Dto2D.cu:167-197
 *   calc_T .............. cpab/gpu/calc_2Dto2D.cu:348-408
 *   calc_T_simple ....... cpab/gpu/calc_2Dto2D.cu:412-445
 *   calc_trajectory ..... cpab/gpu/calc_2Dto2D.cu:499-556
 *   calc_v .............. cpab/gpu/calc_2Dto2D.cu:561-602
 *   calc_cell_idx ....... cpab/gpu/calc_2Dto2D.cu:323-344
 *   sensitivity ......... cpab/gpu/calc_2Dto2D.cu:202-319,448-491
 *   expm 2D ............. cpab/gpu/expm/expm_affine_by_series_2D.py:19-237
 *   expm 3D ............. cpab/gpu/expm/expm_affine_by_series_3D.py:19-219
 *   1D closed-form Trel . cpab/cpaNd/_CpaCalcs.py:285-292
 *   sign / dt prep ...... cpab/cpaNd/_CpaCalcs.py:243-256
 *   SSD likelihood ...... cpab/cpa2d/model/transformations/register/gpu/gaussian.py:15-34
 *
 * Bilinear / trilinear remap is not in the reference (it lives in the
 * un-vendored `of` package); the functions here define this project's
 * convention (OpenCV INTER_LINEAR with a zero constant border).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TESS_I 1  /* triangles (2D) / tetrahedra (3D); reference TESS_TYPE 1 */
#define TESS_II 2 /* rectangles / boxes; reference TESS_TYPE 2 */

typedef struct {
    int dim, tess;
    int nC0, nC1, nC2;
    double inc_x, inc_y, inc_z;
} geom_t;

/* calc_2Dto2D.cu:20-22.  Returns a if b >= a (or b is NaN), else round(b). */
static inline int mymin(int a, double b) { return !(b < a) ? a : (int)round(b); }

/* ---------------------------------------------------------------- lookups */

/* calc_1Dto1D.cu:62-65 */
static int cell_1d(const double *p, const geom_t *g) {
    return (int)round(fmin((double)(g->nC0 - 1),
                           fmax(0.0, (p[0] - fmod(p[0], g->inc_x)) / g->inc_x)));
}

/* calc_2Dto2D.cu:38-153 */
static int cell_2d(const double *p, const geom_t *g) {
    const int nC0 = g->nC0, nC1 = g->nC1;
    const double inc_x = g->inc_x, inc_y = g->inc_y;
    if (g->tess == TESS_II) {
        double a = round(fmin((double)(nC0 - 1), fmax(0.0, (p[0] - fmod(p[0], inc_x)) / inc_x)));
        double b = round(fmin((double)(nC1 - 1), fmax(0.0, (p[1] - fmod(p[1], inc_y)) / inc_y)));
        return (int)(a + b * nC0);
    }
    double p0 = fmin((nC0 * inc_x - 0.0000000001), fmax(0.0, p[0]));
    double p1 = fmin((nC1 * inc_y - 0.0000000001), fmax(0.0, p[1]));
    double xmod = fmod(p0, inc_x);
    double ymod = fmod(p1, inc_y);
    double x = xmod / inc_x;
    double y = ymod / inc_y;
    int c = mymin(nC0 - 1, (p0 - xmod) / inc_x) + mymin(nC1 - 1, (p1 - ymod) / inc_y) * nC0;
    c *= 4; /* triangle order in a rectangle: 0 top, 1 right, 2 bottom, 3 left */
    if (p[0] <= 0) {
        if (p[1] <= 0 && p[1] / inc_y < p[0] / inc_x) {
        } else if (p[1] >= nC1 * inc_y && p[1] / inc_y - nC1 > -p[0] / inc_x) {
            c += 2;
        } else {
            c += 3;
        }
        return c;
    }
    if (p[0] >= nC0 * inc_x) {
        if (p[1] <= 0 && -p[1] / inc_y > p[0] / inc_x - nC0) {
        } else if (p[1] >= nC1 * inc_y && p[1] / inc_y - nC1 > p[0] / inc_x - nC0) {
            c += 2;
        } else {
            c += 1;
        }
        return c;
    }
    if (p[1] <= 0) return c;
    if (p[1] >= nC1 * inc_y) return c + 2;
    if (x < y) {
        if (1 - x < y) c += 2;
        else c += 3;
    } else if (1 - x < y) {
        c += 1;
    }
    return c;
}

/* calc_3Dto3D.cu:43-137 (note the z clamp uses inc_x, :59) */
static int cell_3d(const double *p, const geom_t *g) {
    const int nC0 = g->nC0, nC1 = g->nC1, nC2 = g->nC2;
    const double inc_x = g->inc_x, inc_y = g->inc_y, inc_z = g->inc_z;
    if (g->tess == TESS_II) {
        double a = round(fmin((double)(nC0 - 1), fmax(0.0, (p[0] - fmod(p[0], inc_x)) / inc_x)));
        double b = round(fmin((double)(nC1 - 1), fmax(0.0, (p[1] - fmod(p[1], inc_y)) / inc_y)));
        double cc = round(fmin((double)(nC2 - 1), fmax(0.0, (p[2] - fmod(p[2], inc_z)) / inc_z)));
        return (int)(a + b * nC0 + cc * nC1 * nC0);
    }
    double p0 = fmin((nC0 * inc_x - 0.0000000001), fmax(0.0, p[0]));
    double p1 = fmin((nC1 * inc_y - 0.0000000001), fmax(0.0, p[1]));
    double p2 = fmin((nC2 * inc_x - 0.0000000001), fmax(0.0, p[2]));
    double xmod = fmod(p0, inc_x);
    double ymod = fmod(p1, inc_y);
    double zmod = fmod(p2, inc_z);
    int i = mymin(nC0 - 1, ((p0 - xmod) / inc_x));
    int j = mymin(nC1 - 1, ((p1 - ymod) / inc_y));
    int k = mymin(nC2 - 1, ((p2 - zmod) / inc_z));
    int c = i + j * nC0 + k * nC1 * nC0;
    c *= 5;
    double x = xmod / inc_x;
    double y = ymod / inc_y;
    double z = zmod / inc_z;
    int tf = 0;
    if (k % 2 == 0) {
        if ((i % 2 == 0 && j % 2 == 1) || (i % 2 == 1 && j % 2 == 0)) tf = 1;
    } else if ((i % 2 == 0 && j % 2 == 0) || (i % 2 == 1 && j % 2 == 1)) {
        tf = 1;
    }
    if (tf) {
        double tmp = x;
        x = y;
        y = 1 - tmp;
    }
    if (-x - y + z >= 0) c += 1;
    else if (x + y + z - 2 >= 0) c += 2;
    else if (-x + y - z >= 0) c += 3;
    else if (x - y - z >= 0) c += 4;
    return c;
}

static inline int cell_of(const double *p, const geom_t *g) {
    if (g->dim == 1) return cell_1d(p, g);
    if (g->dim == 2) return cell_2d(p, g);
    return cell_3d(p, g);
}

/* ------------------------------------------------- facet distance (report) */
/* Distance from p to the nearest facet of the lookup's decision surfaces
 * (cell walls, triangle diagonals / tetra faces, domain walls).  Used only to
 * report which points ran close enough to a facet for an ulp to flip a cell
 * id.  For out-of-domain points the distance to the domain wall is used. */
static double facet_dist(const double *p, const geom_t *g) {
    double best = INFINITY;
    const double incs[3] = {g->inc_x, g->inc_y, g->inc_z};
    const int ncs[3] = {g->nC0, g->nC1, g->nC2};
    double u[3];
    for (int d = 0; d < g->dim; d++) {
        double m = fmod(p[d], incs[d]);
        if (m < 0) m += incs[d];
        double a = fmin(m, incs[d] - m);
        if (a < best) best = a;
        double lo = fabs(p[d]), hi = fabs(p[d] - ncs[d] * incs[d]);
        if (lo < best) best = lo;
        if (hi < best) best = hi;
        u[d] = m / incs[d];
    }
    if (g->tess == TESS_I && g->dim == 2) {
        double s = fmin(g->inc_x, g->inc_y) / sqrt(2.0);
        double a = fabs(u[0] - u[1]) * s, b = fabs(1 - u[0] - u[1]) * s;
        if (a < best) best = a;
        if (b < best) best = b;
    } else if (g->tess == TESS_I && g->dim == 3) {
        double s = fmin(g->inc_x, fmin(g->inc_y, g->inc_z)) / sqrt(3.0);
        double x = u[0], y = u[1], z = u[2];
        double planes[8] = {-x - y + z, x + y + z - 2, -x + y - z, x - y - z,
                            -y - x + z, y + (1 - x) + z - 2, -y + (1 - x) - z, y - (1 - x) - z};
        for (int q = 0; q < 8; q++) {
            double a = fabs(planes[q]) * s;
            if (a < best) best = a;
        }
    }
    return best;
}

/* ------------------------------------------------------- affine products */

/* calc_2Dto2D.cu:24-29 / calc_1Dto1D.cu:18-34 / calc_3Dto3D.cu:25-31 */
static inline void A_times_b_affine(int dim, double *x, const double *A, const double *b) {
    if (dim == 1) {
        x[0] = A[0] * b[0] + A[1];
    } else if (dim == 2) {
        x[0] = A[0] * b[0] + A[1] * b[1] + A[2];
        x[1] = A[3] * b[0] + A[4] * b[1] + A[5];
    } else {
        x[0] = A[0] * b[0] + A[1] * b[1] + A[2] * b[2] + A[3];
        x[1] = A[4] * b[0] + A[5] * b[1] + A[6] * b[2] + A[7];
        x[2] = A[8] * b[0] + A[9] * b[1] + A[10] * b[2] + A[11];
    }
}

/* calc_2Dto2D.cu:31-36 / calc_1Dto1D.cu:36-52 / calc_3Dto3D.cu:33-39 */
static inline void A_times_b_linear(int dim, double *x, const double *A, const double *b) {
    if (dim == 1) {
        x[0] = A[0] * b[0];
    } else if (dim == 2) {
        x[0] = A[0] * b[0] + A[1] * b[1];
        x[1] = A[3] * b[0] + A[4] * b[1];
    } else {
        x[0] = A[0] * b[0] + A[1] * b[1] + A[2] * b[2];
        x[1] = A[4] * b[0] + A[5] * b[1] + A[6] * b[2];
        x[2] = A[8] * b[0] + A[9] * b[1] + A[10] * b[2];
    }
}

typedef struct {
    int n_lookups;
    double min_dist;
} probe_t;

static inline int cell_probe(const double *p, const geom_t *g, probe_t *pr) {
    if (pr) {
        double d = facet_dist(p, g);
        if (d < pr->min_dist) pr->min_dist = d;
        pr->n_lookups++;
    }
    return cell_of(p, g);
}

/* calc_2Dto2D.cu:167-197 (and the 1D / 3D copies) */
static void solveODE(int dim, double *p, const double *As, double h, int nSteps, const geom_t *g,
                     probe_t *pr) {
    const int ne = dim * (dim + 1);
    double v[3], pMid[3];
    for (int t = 0; t < nSteps; ++t) {
        int c = cell_probe(p, g, pr);
        const double *A = As + (size_t)c * ne;
        A_times_b_affine(dim, v, A, p);
        for (int i = 0; i < dim; ++i) pMid[i] = p[i] + h * v[i] / 2.;
        A_times_b_affine(dim, v, A, pMid);
        for (int i = 0; i < dim; ++i) p[i] += v[i] * h;
    }
}

static void fill_geom(geom_t *g, int dim, int tess, int nC0, int nC1, int nC2, double ix,
                      double iy, double iz) {
    g->dim = dim;
    g->tess = tess;
    g->nC0 = nC0;
    g->nC1 = nC1;
    g->nC2 = nC2;
    g->inc_x = ix;
    g->inc_y = iy;
    g->inc_z = iz;
}

static int pick_threads(int nthreads) {
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/* ------------------------------------------------------------ kernels */

/* calc_2Dto2D.cu:348-408.  n_fallback[i] (optional) counts the steps whose
 * Trel proposal left the cell; min_dist[i] (optional) is the smallest facet
 * distance over every point the integration looked up. */
void oracle_calc_T(int dim, int tess, const double *pos0, double *pos, const double *Trels,
                   const double *As, double dt, int nTimeSteps, int nStepsODE, int64_t nPts,
                   int nC0, int nC1, int nC2, double inc_x, double inc_y, double inc_z,
                   int32_t *n_fallback, double *min_dist, int nthreads) {
    geom_t g;
    fill_geom(&g, dim, tess, nC0, nC1, nC2, inc_x, inc_y, inc_z);
    const int ne = dim * (dim + 1);
    const int want_probe = (min_dist != NULL);
#pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t idx = 0; idx < nPts; idx++) {
        double p[3], pNew[3];
        probe_t pr = {0, INFINITY};
        probe_t *ppr = want_probe ? &pr : NULL;
        for (int i = 0; i < dim; ++i) p[i] = pos0[idx * dim + i];
        double h = dt / (double)nStepsODE;
        int nfb = 0;
        for (int t = 0; t < nTimeSteps; ++t) {
            int c = cell_probe(p, &g, ppr);
            A_times_b_affine(dim, pNew, Trels + (size_t)c * ne, p);
            int cn = cell_probe(pNew, &g, ppr);
            if (cn == c) {
                for (int i = 0; i < dim; ++i) p[i] = pNew[i];
            } else {
                nfb++;
                solveODE(dim, p, As, h, nStepsODE, &g, ppr);
            }
        }
        for (int i = 0; i < dim; ++i) pos[idx * dim + i] = p[i];
        if (n_fallback) n_fallback[idx] = nfb;
        if (min_dist) min_dist[idx] = pr.min_dist;
    }
}

/* calc_2Dto2D.cu:412-445 */
void oracle_calc_T_simple(int dim, int tess, const double *pos0, double *pos, const double *As,
                          double dt, int nTimeSteps, int nStepsODE, int64_t nPts, int nC0,
                          int nC1, int nC2, double inc_x, double inc_y, double inc_z,
                          double *min_dist, int nthreads) {
    geom_t g;
    fill_geom(&g, dim, tess, nC0, nC1, nC2, inc_x, inc_y, inc_z);
#pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t idx = 0; idx < nPts; idx++) {
        double p[3];
        probe_t pr = {0, INFINITY};
        for (int i = 0; i < dim; ++i) p[i] = pos0[idx * dim + i];
        double h = dt / (double)nStepsODE;
        solveODE(dim, p, As, h, nStepsODE * nTimeSteps, &g, min_dist ? &pr : NULL);
        for (int i = 0; i < dim; ++i) pos[idx * dim + i] = p[i];
        if (min_dist) min_dist[idx] = pr.min_dist;
    }
}

/* calc_2Dto2D.cu:499-556.  traj has shape (nTimeSteps, nPts, dim): row t is
 * the position after t+1 steps (the kernel's in-place layout minus row 0's
 * initial copy, which the host glue re-prepends, _CpaCalcs.py:866-868). */
void oracle_calc_trajectory(int dim, int tess, const double *pos0, double *traj,
                            const double *Trels, const double *As, double dt, int nTimeSteps,
                            int nStepsODE, int64_t nPts, int nC0, int nC1, int nC2,
                            double inc_x, double inc_y, double inc_z, int nthreads) {
    geom_t g;
    fill_geom(&g, dim, tess, nC0, nC1, nC2, inc_x, inc_y, inc_z);
    const int ne = dim * (dim + 1);
#pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t idx = 0; idx < nPts; idx++) {
        double p[3], pNew[3];
        for (int i = 0; i < dim; ++i) p[i] = pos0[idx * dim + i];
        double h = dt / (double)nStepsODE;
        for (int t = 0; t < nTimeSteps; ++t) {
            int c = cell_of(p, &g);
            A_times_b_affine(dim, pNew, Trels + (size_t)c * ne, p);
            int cn = cell_of(pNew, &g);
            if (cn == c) {
                for (int i = 0; i < dim; ++i) p[i] = pNew[i];
            } else {
                solveODE(dim, p, As, h, nStepsODE, &g, NULL);
            }
            for (int i = 0; i < dim; ++i) traj[((int64_t)t * nPts + idx) * dim + i] = p[i];
        }
    }
}

/* calc_2Dto2D.cu:561-602 */
void oracle_calc_v(int dim, int tess, const double *pos, double *vel, const double *As,
                   int64_t nPts, int nC0, int nC1, int nC2, double inc_x, double inc_y,
                   double inc_z, int nthreads) {
    geom_t g;
    fill_geom(&g, dim, tess, nC0, nC1, nC2, inc_x, inc_y, inc_z);
    const int ne = dim * (dim + 1);
#pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t idx = 0; idx < nPts; idx++) {
        double p[3], v[3];
        for (int i = 0; i < dim; ++i) p[i] = pos[idx * dim + i];
        int c = cell_of(p, &g);
        A_times_b_affine(dim, v, As + (size_t)c * ne, p);
        for (int i = 0; i < dim; ++i) vel[idx * dim + i] = v[i];
    }
}

/* calc_2Dto2D.cu:323-344 */
void oracle_calc_cell_idx(int dim, int tess, const double *pts, int32_t *cell_idx, int64_t nPts,
                          int nC0, int nC1, int nC2, double inc_x, double inc_y, double inc_z,
                          int nthreads) {
    geom_t g;
    fill_geom(&g, dim, tess, nC0, nC1, nC2, inc_x, inc_y, inc_z);
#pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t idx = 0; idx < nPts; idx++) cell_idx[idx] = cell_of(pts + idx * dim, &g);
}

/* Facet distance per point (reporting helper; no reference counterpart). */
void oracle_facet_dist(int dim, int tess, const double *pts, double *out, int64_t nPts, int nC0,
                       int nC1, int nC2, double inc_x, double inc_y, double inc_z) {
    geom_t g;
    fill_geom(&g, dim, tess, nC0, nC1, nC2, inc_x, inc_y, inc_z);
    for (int64_t idx = 0; idx < nPts; idx++) out[idx] = facet_dist(pts + idx * dim, &g);
}

/* calc_2Dto2D.cu:202-319 + 448-491: forward sensitivity of calc_T_simple.
 * Bs has layout (d, nC, dim, dim+1); grad has layout (nPts, dim, d). */
void oracle_calc_grad_theta(int dim, int tess, const double *pos0, double *pos, const double *As,
                            const double *Bs, double *grad, int d, int nC, double dt,
                            int nTimeSteps, int nStepsODE, int64_t nPts, int nC0, int nC1,
                            int nC2, double inc_x, double inc_y, double inc_z, int nthreads) {
    geom_t g;
    fill_geom(&g, dim, tess, nC0, nC1, nC2, inc_x, inc_y, inc_z);
    const int ne = dim * (dim + 1);
#pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t idx = 0; idx < nPts; idx++) {
        double p[3], v[3], pMid[3], vMid[3], q[3], qMid[3], u[3], uMid[3], BtT[3], AdT[3];
        for (int i = 0; i < dim; ++i) p[i] = pos0[idx * dim + i];
        double h = dt / (double)nStepsODE;
        double *gp = grad + idx * (int64_t)dim * d;
        for (int j = 0; j < d; j++)
            for (int i = 0; i < dim; ++i) gp[i * d + j] = 0;
        const int nSteps = nStepsODE * nTimeSteps;
        for (int t = 0; t < nSteps; ++t) {
            int c = cell_of(p, &g);
            int mi = c * ne;
            A_times_b_affine(dim, v, As + mi, p);
            for (int i = 0; i < dim; ++i) pMid[i] = p[i] + h * v[i] / 2.;
            A_times_b_affine(dim, vMid, As + mi, pMid);
            for (int j = 0; j < d; j++) {
                int64_t bi = (int64_t)j * ne * nC + mi;
                for (int i = 0; i < dim; ++i) q[i] = gp[i * d + j];
                A_times_b_affine(dim, BtT, Bs + bi, p);
                A_times_b_linear(dim, AdT, As + mi, q);
                for (int i = 0; i < dim; ++i) u[i] = BtT[i] + AdT[i];
                for (int i = 0; i < dim; ++i) qMid[i] = q[i] + h * u[i] / 2.;
                A_times_b_affine(dim, BtT, Bs + bi, pMid);
                A_times_b_linear(dim, AdT, As + mi, qMid);
                for (int i = 0; i < dim; ++i) uMid[i] = BtT[i] + AdT[i];
                for (int i = 0; i < dim; ++i) q[i] += uMid[i] * h;
                for (int i = 0; i < dim; ++i) gp[i * d + j] = q[i];
            }
            for (int i = 0; i < dim; ++i) p[i] += vMid[i] * h;
        }
        for (int i = 0; i < dim; ++i) pos[idx * dim + i] = p[i];
    }
}

/* Forward sensitivity of calc_T (calc_2Dto2D.cu:348-408), the map the
 * reference integrates for calc_T_fwd / calc_T_inv: per time step the Trel
 * proposal pNew = Trel_c [p;1], kept when its cell is c, else nStepsODE RK2
 * substeps.  With the decisions of the step held (they change only on a set
 * of measure zero), an accepted step propagates q_j' = E_c q_j + dT_jc [p;1]
 * (E_c the linear part of Trel_c, dT_jc = dTrel_c / dtheta_j, layout
 * (d, nC, dim, dim+1) like Bs), a redone step the RK2 substeps of
 * solveODE2 (calc_2Dto2D.cu:202-319) with Bs = dAs / dtheta.  grad has
 * layout (nPts, dim, d); pos receives calc_T's end points.  No reference
 * counterpart computes this derivative (the reference differentiates only
 * calc_T_simple's RK2). */
void oracle_calc_grad_theta_T(int dim, int tess, const double *pos0, double *pos,
                              const double *Trels, const double *As, const double *Bs,
                              const double *dTs, double *grad, int d, int nC, double dt,
                              int nTimeSteps, int nStepsODE, int64_t nPts, int nC0, int nC1,
                              int nC2, double inc_x, double inc_y, double inc_z, int nthreads) {
    geom_t g;
    fill_geom(&g, dim, tess, nC0, nC1, nC2, inc_x, inc_y, inc_z);
    const int ne = dim * (dim + 1);
#pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t idx = 0; idx < nPts; idx++) {
        double p[3], pNew[3], v[3], pMid[3], vMid[3], q[3], qn[3], qMid[3], u[3], uMid[3], BtT[3],
            AdT[3];
        for (int i = 0; i < dim; ++i) p[i] = pos0[idx * dim + i];
        double h = dt / (double)nStepsODE;
        double *gp = grad + idx * (int64_t)dim * d;
        for (int j = 0; j < d; j++)
            for (int i = 0; i < dim; ++i) gp[i * d + j] = 0;
        for (int t = 0; t < nTimeSteps; ++t) {
            int c = cell_of(p, &g);
            const double *T = Trels + (size_t)c * ne;
            A_times_b_affine(dim, pNew, T, p);
            if (cell_of(pNew, &g) == c) {
                for (int j = 0; j < d; j++) {
                    int64_t bi = (int64_t)j * ne * nC + (int64_t)c * ne;
                    for (int i = 0; i < dim; ++i) q[i] = gp[i * d + j];
                    A_times_b_linear(dim, qn, T, q);
                    A_times_b_affine(dim, BtT, dTs + bi, p);
                    for (int i = 0; i < dim; ++i) gp[i * d + j] = qn[i] + BtT[i];
                }
                for (int i = 0; i < dim; ++i) p[i] = pNew[i];
                continue;
            }
            for (int s = 0; s < nStepsODE; ++s) {
                int cs = cell_of(p, &g);
                int mi = cs * ne;
                A_times_b_affine(dim, v, As + mi, p);
                for (int i = 0; i < dim; ++i) pMid[i] = p[i] + h * v[i] / 2.;
                A_times_b_affine(dim, vMid, As + mi, pMid);
                for (int j = 0; j < d; j++) {
                    int64_t bi = (int64_t)j * ne * nC + mi;
                    for (int i = 0; i < dim; ++i) q[i] = gp[i * d + j];
                    A_times_b_affine(dim, BtT, Bs + bi, p);
                    A_times_b_linear(dim, AdT, As + mi, q);
                    for (int i = 0; i < dim; ++i) u[i] = BtT[i] + AdT[i];
                    for (int i = 0; i < dim; ++i) qMid[i] = q[i] + h * u[i] / 2.;
                    A_times_b_affine(dim, BtT, Bs + bi, pMid);
                    A_times_b_linear(dim, AdT, As + mi, qMid);
                    for (int i = 0; i < dim; ++i) uMid[i] = BtT[i] + AdT[i];
                    for (int i = 0; i < dim; ++i) q[i] += uMid[i] * h;
                    for (int i = 0; i < dim; ++i) gp[i * d + j] = q[i];
                }
                for (int i = 0; i < dim; ++i) p[i] += vMid[i] * h;
            }
        }
        for (int i = 0; i < dim; ++i) pos[idx * dim + i] = p[i];
    }
}

/* --------------------------------------------------------------- expm */

/* expm_affine_by_series_2D.py:28-74 */
static void expm_2x2(const double *A, double *expA) {
    double a = A[0], b = A[1], c = A[2], d = A[3];
    double delta_tmp = (a - d) * (a - d) + 4 * b * c;
    double e = exp((a + d) / 2);
    if (delta_tmp == 0) {
        expA[0] = (1 + (a - d) / 2) * e;
        expA[1] = b * e;
        expA[2] = c * e;
        expA[3] = (1 - (a - d) / 2) * e;
    } else if (delta_tmp > 0) {
        double delta = sqrt(delta_tmp) / 2;
        double ch = cosh(delta), sh = sinh(delta);
        double s = sh / delta;
        expA[0] = (ch + (a - d) / 2 * s) * e;
        expA[1] = b * s * e;
        expA[2] = c * s * e;
        expA[3] = (ch - (a - d) / 2 * s) * e;
    } else {
        double delta = sqrt(-delta_tmp) / 2;
        double co = cos(delta), si = sin(delta);
        double s = si / delta;
        expA[0] = (co + (a - d) / 2 * s) * e;
        expA[1] = b * s * e;
        expA[2] = c * s * e;
        expA[3] = (co - (a - d) / 2 * s) * e;
    }
}

/* expm_affine_by_series_2D.py:85-237, one cell.  A, T: 2x3 row-major. */
static void expm_affine_2d(const double *A, double *T, int p) {
    double A00 = A[0], A01 = A[1], A02 = A[2], A10 = A[3], A11 = A[4], A12 = A[5];
    double det = A00 * A11 - A01 * A10;
    if (det != 0) {
        double A2[4] = {A00, A01, A10, A11}, E[4];
        expm_2x2(A2, E);
        T[0] = E[0];
        T[1] = E[1];
        T[3] = E[2];
        T[4] = E[3];
        double i00 = A11 / det, i11 = A00 / det, i01 = -A01 / det, i10 = -A10 / det;
        double a = i00 * A02 + i01 * A12;
        double b = i10 * A02 + i11 * A12;
        T[2] = (T[0] - 1) * a + (T[1]) * b;
        T[5] = (T[3]) * a + (T[4] - 1) * b;
        return;
    }
    T[0] = 1; T[1] = 0; T[2] = 0; T[3] = 0; T[4] = 1; T[5] = 0;
    double An[6] = {A00, A01, A02, A10, A11, A12}, B[6];
    int fact = 1;
    for (int k = 0; k < 6; k++) T[k] += An[k];
    if (p == 1) return;
    for (int i = 2; i <= p; i++) {
        memcpy(B, An, sizeof B);
        fact *= i;
        An[0] = A00 * B[0] + A01 * B[3];
        An[1] = A00 * B[1] + A01 * B[4];
        An[2] = A00 * B[2] + A01 * B[5];
        An[3] = A10 * B[0] + A11 * B[3];
        An[4] = A10 * B[1] + A11 * B[4];
        An[5] = A10 * B[2] + A11 * B[5];
        for (int k = 0; k < 6; k++) T[k] += An[k] / fact;
    }
}

/* expm_affine_by_series_3D.py:27-219, one cell.  A, T: 3x4 row-major. */
static void expm_affine_3d(const double *A, double *T, int p) {
    static const double I[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    double An[12], B[12];
    for (int k = 0; k < 12; k++) T[k] = I[k];
    for (int k = 0; k < 12; k++) An[k] = A[k];
    int fact = 1;
    for (int k = 0; k < 12; k++) T[k] += An[k];
    if (p == 1) return;
    for (int i = 2; i <= p; i++) {
        memcpy(B, An, sizeof B);
        fact *= i;
        for (int r = 0; r < 3; r++)
            for (int col = 0; col < 4; col++)
                An[r * 4 + col] = A[r * 4 + 0] * B[0 * 4 + col] + A[r * 4 + 1] * B[1 * 4 + col] +
                                  A[r * 4 + 2] * B[2 * 4 + col];
        for (int k = 0; k < 12; k++) T[k] += An[k] / fact;
    }
}

/* _CpaCalcs.py:243-313 for dim 2/3 and :285-292 for dim 1.
 * As: (nC, dim*(dim+1)) unsigned table.  signedAs_out = sign*As;
 * Trels_out = expm(dt*signedAs) (closed form in 1D). */
void oracle_prep(int dim, int nC, const double *As, int sign, double dt, double *signedAs_out,
                 double *Trels_out) {
    const int ne = dim * (dim + 1);
    for (int c = 0; c < nC; c++) {
        double sA[12], sAdt[12];
        for (int k = 0; k < ne; k++) {
            sA[k] = As[c * ne + k];
            if (sign == -1) sA[k] *= -1;
            sAdt[k] = sA[k] * dt;
            signedAs_out[c * ne + k] = sA[k];
        }
        double *T = Trels_out + (size_t)c * ne;
        if (dim == 1) {
            T[0] = exp(sAdt[0]);
            T[1] = sAdt[1];
            if (sA[0] != 0) T[1] = T[1] * (T[0] - 1) / sAdt[0];
        } else if (dim == 2) {
            expm_affine_2d(sAdt, T, 12);
        } else {
            expm_affine_3d(sAdt, T, 12);
        }
    }
}

/* expm on an already scaled table (the reference's gpu_expm(As, Ts, p)). */
void oracle_expm(int dim, int nC, const double *As, double *Ts, int p) {
    const int ne = dim * (dim + 1);
    for (int c = 0; c < nC; c++) {
        if (dim == 2) expm_affine_2d(As + c * ne, Ts + c * ne, p);
        else expm_affine_3d(As + c * ne, Ts + c * ne, p);
    }
}

/* --------------------------------------------------- remap + likelihood */

/* Bilinear sample of img (H x W, float) at (x, y) pixel coordinates, zero
 * outside: OpenCV INTER_LINEAR / BORDER_CONSTANT(0) semantics, computed in
 * float with weights from the double coordinates' fractional parts. */
static inline float sample_bilinear(const float *img, int H, int W, double x, double y) {
    double xf = floor(x), yf = floor(y);
    /* a full pixel or more outside, or not a number: zero */
    if (!(xf >= -1.0 && xf < (double)W && yf >= -1.0 && yf < (double)H)) return 0.f;
    float fx = (float)(x - xf), fy = (float)(y - yf);
    float v00 = 0.f, v01 = 0.f, v10 = 0.f, v11 = 0.f;
    int x0 = (int)xf, y0 = (int)yf;
    int x1 = x0 + 1, y1 = y0 + 1;
    int okx0 = x0 >= 0, okx1 = x1 < W, oky0 = y0 >= 0, oky1 = y1 < H;
    if (oky0 && okx0) v00 = img[(size_t)y0 * W + x0];
    if (oky0 && okx1) v01 = img[(size_t)y0 * W + x1];
    if (oky1 && okx0) v10 = img[(size_t)y1 * W + x0];
    if (oky1 && okx1) v11 = img[(size_t)y1 * W + x1];
    float gx = 1.f - fx, gy = 1.f - fy;
    float a = gx * v00 + fx * v01;
    float b = gx * v10 + fx * v11;
    return gy * a + fy * b;
}

/* out[r, c] = img(pts[r*W + c]) for an H x W image (TransformWrapper.remap_fwd
 * semantics, cpab/cpa2d/TransformWrapper.py:261-317). */
void oracle_remap2d(const double *pts, const float *img, float *out, int H, int W, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t i = 0; i < (int64_t)H * W; i++)
        out[i] = sample_bilinear(img, H, W, pts[2 * i], pts[2 * i + 1]);
}

static inline float sample_trilinear(const float *vol, int D, int H, int W, double x, double y,
                                     double z) {
    double xf = floor(x), yf = floor(y), zf = floor(z);
    if (!(xf >= -1.0 && xf < (double)W && yf >= -1.0 && yf < (double)H && zf >= -1.0 &&
          zf < (double)D))
        return 0.f;
    float fx = (float)(x - xf), fy = (float)(y - yf), fz = (float)(z - zf);
    float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int x0 = (int)xf, y0 = (int)yf, z0 = (int)zf;
    for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++) {
                int xx = x0 + dx, yy = y0 + dy, zz = z0 + dz;
                if (xx >= 0 && xx < W && yy >= 0 && yy < H && zz >= 0 && zz < D)
                    v[dz * 4 + dy * 2 + dx] = vol[((size_t)zz * H + yy) * W + xx];
            }
    float gx = 1.f - fx, gy = 1.f - fy, gz = 1.f - fz;
    float a0 = gx * v[0] + fx * v[1];
    float a1 = gx * v[2] + fx * v[3];
    float a2 = gx * v[4] + fx * v[5];
    float a3 = gx * v[6] + fx * v[7];
    float b0 = gy * a0 + fy * a1;
    float b1 = gy * a2 + fy * a3;
    return gz * b0 + fz * b1;
}

/* out[z, y, x] = vol(pts[(z*H + y)*W + x]) for a D x H x W volume. */
void oracle_remap3d(const double *pts, const float *vol, float *out, int D, int H, int W,
                    int nthreads) {
#pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t i = 0; i < (int64_t)D * H * W; i++)
        out[i] = sample_trilinear(vol, D, H, W, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
}

/* gaussian.py:15-34 + the sum of _ScaleDependentLogLikelihoodGeneral.py:118:
 * sum_i -0.5 * (x_i - y_i)^2 / sigma^2, accumulated in double. */
double oracle_ssd_ll(const float *x, const float *y, int64_t n, double sigma) {
    double acc = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double e = (double)x[i] - (double)y[i];
        double l = e * e;
        l *= -0.5 / (sigma * sigma);
        acc += l;
    }
    return acc;
}

/* Timed CPU baseline of one unit of work: prep (sign, dt, expm) + calc_T. */
void oracle_transform(int dim, int tess, int nC, const double *As, int sign, const double *pos0,
                      double *pos, double dt, int nTimeSteps, int nStepsODE, int64_t nPts,
                      int nC0, int nC1, int nC2, double inc_x, double inc_y, double inc_z,
                      int nthreads) {
    const int ne = dim * (dim + 1);
    double *sA = (double *)malloc(sizeof(double) * nC * ne);
    double *T = (double *)malloc(sizeof(double) * nC * ne);
    oracle_prep(dim, nC, As, sign, dt, sA, T);
    oracle_calc_T(dim, tess, pos0, pos, T, sA, dt, nTimeSteps, nStepsODE, nPts, nC0, nC1, nC2,
                  inc_x, inc_y, inc_z, NULL, NULL, nthreads);
    free(sA);
    free(T);
}

int oracle_max_threads(void) { return pick_threads(0); }
