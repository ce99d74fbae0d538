/*
 * pd_oracle.c -- FP64 CPU restatement of the reference NOSB peridynamics path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker.  The product (simul_amd, libpdx)
 * never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Arithmetic is plain IEEE FP64; build with
 * -ffp-contract=off so no FMA contraction changes the distance tests.
 *
 * Layout: particles in caller order, vectors n*dim, tensors n*dim*dim
 * (row-major per particle), bonds as CSR (row[n+1], col[nb]) with every row
 * sorted by ascending j.  "fixed" is a per-particle DOF bitmask (bit d set =>
 * DOF d is Dirichlet); the reference's all-or-nothing `fixed` flag
 * (peridynamics/core/particles.py:62) is the mask (1<<dim)-1.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t n, dim;
    int32_t model;              /* 0 linear elastic, 1 Neo-Hookean, 2 J2 */
    int32_t formulation;        /* 0 NOSB correspondence + stabilisation, 1 bond-based PMB */
    int32_t per_bond_recompute; /* 1: P_j K_j^-1 recomputed per bond as nosb.py:254-256 */
    int32_t nthreads;
    double horizon;
    double G_s;
    /* particles */
    const double *X;
    double *x;
    const double *vol, *mass;
    const double *bulk, *shear, *cstab;
    const double *sy0, *h_iso, *h_kin;   /* J2 parameters (model 2 only) */
    double *ep, *epbar, *beta;           /* committed J2 state (9, 1, 9 per particle) */
    double *ep_t, *epbar_t, *beta_t;     /* trial J2 state written by orc_stress */
    const uint8_t *fixed;
    const double *fext;                  /* n*dim or NULL */
    /* bonds */
    const int64_t *row;
    const int32_t *col;
    uint8_t *broken;
    double *xi, *xlen, *omega;
    /* per-particle tensors */
    double *K, *Kinv;
    uint8_t *sing;
    double *F, *P, *A;
    /* dynamics */
    double *f, *v, *a;
} orc_pd;

static int nthr(int32_t t) { return t > 0 ? t : 1; }

/* ------------------------------------------------------------------ */
/* small dense helpers                                                  */
/* ------------------------------------------------------------------ */
static double det_d(const double *M, int d)
{
    if (d == 2) return M[0] * M[3] - M[1] * M[2];
    return M[0] * (M[4] * M[8] - M[5] * M[7])
         - M[1] * (M[3] * M[8] - M[5] * M[6])
         + M[2] * (M[3] * M[7] - M[4] * M[6]);
}

/* inverse by adjugate / determinant */
static void inv_d(const double *M, double det, double *R, int d)
{
    if (d == 2) {
        R[0] = M[3] / det; R[1] = -M[1] / det;
        R[2] = -M[2] / det; R[3] = M[0] / det;
        return;
    }
    R[0] = (M[4] * M[8] - M[5] * M[7]) / det;
    R[1] = (M[2] * M[7] - M[1] * M[8]) / det;
    R[2] = (M[1] * M[5] - M[2] * M[4]) / det;
    R[3] = (M[5] * M[6] - M[3] * M[8]) / det;
    R[4] = (M[0] * M[8] - M[2] * M[6]) / det;
    R[5] = (M[2] * M[3] - M[0] * M[5]) / det;
    R[6] = (M[3] * M[7] - M[4] * M[6]) / det;
    R[7] = (M[1] * M[6] - M[0] * M[7]) / det;
    R[8] = (M[0] * M[4] - M[1] * M[3]) / det;
}

static void matmul_d(const double *A, const double *B, double *C, int d)
{
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c) {
            double s = 0.0;
            for (int k = 0; k < d; ++k) s += A[r * d + k] * B[k * d + c];
            C[r * d + c] = s;
        }
}

/* ------------------------------------------------------------------ */
/* a1: neighbour sets  (peridynamics/core/neighbor.py:154-206)          */
/* set semantics: {j != i : |X_j - X_i|^2 < h^2}, FP64, distance summed */
/* as (dx*dx + dy*dy) + dz*dz (neighbor.py:180-182,200-202); no cap.    */
/* ------------------------------------------------------------------ */
static double dist2(const double *X, int dim, int64_t i, int64_t j)
{
    double s = 0.0;
    for (int d = 0; d < dim; ++d) {
        double t = X[j * dim + d] - X[i * dim + d];
        s = s + t * t;
    }
    return s;
}

typedef struct {
    int dim;
    int64_t nc[3];
    double lo[3], h;
    int64_t *start; /* ncells+1 */
    int32_t *idx;   /* particles bucketed by cell, ascending index inside a cell */
} cellgrid;

static int64_t cell_coord(const cellgrid *g, const double *p, int d)
{
    int64_t c = (int64_t)floor((p[d] - g->lo[d]) / g->h);
    if (c < 0) c = 0;
    if (c >= g->nc[d]) c = g->nc[d] - 1;
    return c;
}

static int cellgrid_build(cellgrid *g, int32_t n, int dim, const double *X, double h)
{
    g->dim = dim;
    g->h = h;
    for (int d = 0; d < 3; ++d) { g->nc[d] = 1; g->lo[d] = 0.0; }
    for (int d = 0; d < dim; ++d) {
        double lo = X[d], hi = X[d];
        for (int32_t i = 1; i < n; ++i) {
            double t = X[(int64_t)i * dim + d];
            if (t < lo) lo = t;
            if (t > hi) hi = t;
        }
        g->lo[d] = lo;
        g->nc[d] = (int64_t)floor((hi - lo) / h) + 1;
    }
    int64_t ncell = g->nc[0] * g->nc[1] * g->nc[2];
    g->start = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
    g->idx = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    int64_t *cid = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    if (!g->start || !g->idx || !cid) { free(cid); return -1; }
    for (int32_t i = 0; i < n; ++i) {
        const double *p = X + (int64_t)i * dim;
        int64_t c[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) c[d] = cell_coord(g, p, d);
        cid[i] = (c[0] * g->nc[1] + c[1]) * g->nc[2] + c[2];
        g->start[cid[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) g->start[c + 1] += g->start[c];
    int64_t *fill = (int64_t *)malloc((size_t)ncell * sizeof(int64_t));
    memcpy(fill, g->start, (size_t)ncell * sizeof(int64_t));
    for (int32_t i = 0; i < n; ++i) g->idx[fill[cid[i]]++] = i;
    free(fill);
    free(cid);
    return 0;
}

static void cellgrid_free(cellgrid *g) { free(g->start); free(g->idx); }

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* scans the 27 (9 in 2-D) cells around i; writes sorted j into out (if non-NULL) */
static int64_t scan_one(const cellgrid *g, int32_t n, const double *X, double h2, int32_t i,
                        int32_t *out)
{
    int dim = g->dim;
    const double *p = X + (int64_t)i * dim;
    int64_t c[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) c[d] = cell_coord(g, p, d);
    int64_t cnt = 0;
    int rz = dim == 3 ? 1 : 0;
    for (int64_t a = c[0] - 1; a <= c[0] + 1; ++a) {
        if (a < 0 || a >= g->nc[0]) continue;
        for (int64_t b = c[1] - 1; b <= c[1] + 1; ++b) {
            if (b < 0 || b >= g->nc[1]) continue;
            for (int64_t e = c[2] - rz; e <= c[2] + rz; ++e) {
                if (e < 0 || e >= g->nc[2]) continue;
                int64_t cell = (a * g->nc[1] + b) * g->nc[2] + e;
                for (int64_t k = g->start[cell]; k < g->start[cell + 1]; ++k) {
                    int32_t j = g->idx[k];
                    if (j == i) continue;
                    if (dist2(X, dim, i, j) < h2) {
                        if (out) out[cnt] = j;
                        cnt++;
                    }
                }
            }
        }
    }
    (void)n;
    if (out && cnt > 1) qsort(out, (size_t)cnt, sizeof(int32_t), cmp_i32);
    return cnt;
}

int64_t orc_neighbor_count(int32_t n, int32_t dim, const double *X, double horizon,
                           int32_t *counts, int32_t nthreads)
{
    cellgrid g;
    if (cellgrid_build(&g, n, dim, X, horizon * 1.01) != 0) return -1;
    double h2 = horizon * horizon;
    int64_t total = 0;
#pragma omp parallel for schedule(static) reduction(+ : total) num_threads(nthr(nthreads))
    for (int32_t i = 0; i < n; ++i) {
        counts[i] = (int32_t)scan_one(&g, n, X, h2, i, NULL);
        total += counts[i];
    }
    cellgrid_free(&g);
    return total;
}

void orc_neighbor_fill(int32_t n, int32_t dim, const double *X, double horizon,
                       const int64_t *row, int32_t *col, int32_t nthreads)
{
    cellgrid g;
    if (cellgrid_build(&g, n, dim, X, horizon * 1.01) != 0) return;
    double h2 = horizon * horizon;
#pragma omp parallel for schedule(static) num_threads(nthr(nthreads))
    for (int32_t i = 0; i < n; ++i) scan_one(&g, n, X, h2, i, col + row[i]);
    cellgrid_free(&g);
}

/* O(n^2) brute force, the independent check of the cell-list search */
int64_t orc_neighbor_count_brute(int32_t n, int32_t dim, const double *X, double horizon,
                                 int32_t *counts)
{
    double h2 = horizon * horizon;
    int64_t total = 0;
    for (int32_t i = 0; i < n; ++i) {
        int32_t c = 0;
        for (int32_t j = 0; j < n; ++j)
            if (j != i && dist2(X, dim, i, j) < h2) c++;
        counts[i] = c;
        total += c;
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* a2: bond data  (peridynamics/core/bonds.py:116-137)                  */
/* ------------------------------------------------------------------ */
void orc_bond_data(const orc_pd *p)
{
    int dim = p->dim;
#pragma omp parallel for schedule(static) num_threads(nthr(p->nthreads))
    for (int32_t i = 0; i < p->n; ++i) {
        for (int64_t b = p->row[i]; b < p->row[i + 1]; ++b) {
            int64_t j = p->col[b];
            double s = 0.0;
            for (int d = 0; d < dim; ++d) {
                double t = p->X[j * dim + d] - p->X[(int64_t)i * dim + d];
                p->xi[b * dim + d] = t;
                s = s + t * t;
            }
            double L = sqrt(s);
            p->xlen[b] = L;
            if (L > 0) {
                double w = 1.0 - L / p->horizon;
                p->omega[b] = w > 0.0 ? w : 0.0;
            } else {
                p->omega[b] = 1.0;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* a3: shape tensor and inverse  (peridynamics/core/nosb.py:59-94)      */
/* mode 0: reference test |det K| > tol (absolute, nosb.py:89-90)       */
/* mode 1: scale-free test |det K| > tol * (tr K / dim)^dim             */
/* ------------------------------------------------------------------ */
void orc_shape_tensor(orc_pd *p, int32_t mode, double tol)
{
    int dim = p->dim, dd = dim * dim;
#pragma omp parallel for schedule(static) num_threads(nthr(p->nthreads))
    for (int32_t i = 0; i < p->n; ++i) {
        double K[9] = {0};
        for (int64_t b = p->row[i]; b < p->row[i + 1]; ++b) {
            if (p->broken[b]) continue;
            int64_t j = p->col[b];
            const double *xi = p->xi + b * dim;
            double w = p->omega[b], Vj = p->vol[j];
            for (int m = 0; m < dim; ++m)
                for (int q = 0; q < dim; ++q) K[m * dim + q] += w * xi[m] * xi[q] * Vj;
        }
        double *Ko = p->K + (int64_t)i * dd, *Ki = p->Kinv + (int64_t)i * dd;
        memcpy(Ko, K, sizeof(double) * dd);
        double det = det_d(K, dim);
        double thr = tol;
        if (mode == 1) {
            double tr = 0.0;
            for (int d = 0; d < dim; ++d) tr += K[d * dim + d];
            tr /= dim;
            thr = tol * (dim == 3 ? tr * tr * tr : tr * tr);
        }
        if (fabs(det) > thr && det != 0.0) {
            inv_d(K, det, Ki, dim);
            p->sing[i] = 0;
        } else {
            for (int m = 0; m < dd; ++m) Ki[m] = 0.0;
            for (int d = 0; d < dim; ++d) Ki[d * dim + d] = 1.0;
            p->sing[i] = 1;
        }
    }
}

/* ------------------------------------------------------------------ */
/* a4: deformation gradient  (peridynamics/core/nosb.py:96-126)         */
/* ------------------------------------------------------------------ */
void orc_deformation_gradient(orc_pd *p)
{
    int dim = p->dim, dd = dim * dim;
#pragma omp parallel for schedule(static) num_threads(nthr(p->nthreads))
    for (int32_t i = 0; i < p->n; ++i) {
        double N[9] = {0};
        for (int64_t b = p->row[i]; b < p->row[i + 1]; ++b) {
            if (p->broken[b]) continue;
            int64_t j = p->col[b];
            const double *xi = p->xi + b * dim;
            double eta[3];
            for (int d = 0; d < dim; ++d) eta[d] = p->x[j * dim + d] - p->x[(int64_t)i * dim + d];
            double w = p->omega[b], Vj = p->vol[j];
            for (int m = 0; m < dim; ++m)
                for (int q = 0; q < dim; ++q) N[m * dim + q] += w * eta[m] * xi[q] * Vj;
        }
        matmul_d(N, p->Kinv + (int64_t)i * dd, p->F + (int64_t)i * dd, dim);
    }
}

/* ------------------------------------------------------------------ */
/* a5 / a5' / a5'': first Piola-Kirchhoff stress                       */
/* ------------------------------------------------------------------ */
static double lame_lambda(double Kb, double mu, int dim)
{
    /* nosb.py:222-224 */
    return dim == 3 ? Kb - 2.0 * mu / 3.0 : Kb - mu;
}

/* linear elastic, nosb.py:211-227 */
static void stress_le(const double *F, double Kb, double mu, int dim, double *P)
{
    double eps[9], tr = 0.0;
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c)
            eps[r * dim + c] = 0.5 * (F[r * dim + c] + F[c * dim + r]) - (r == c ? 1.0 : 0.0);
    for (int d = 0; d < dim; ++d) tr += eps[d * dim + d];
    double lam = lame_lambda(Kb, mu, dim);
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c)
            P[r * dim + c] = lam * tr * (r == c ? 1.0 : 0.0) + 2.0 * mu * eps[r * dim + c];
}

/* Neo-Hookean, fem/material/neo_hookean.py:117-127 (sigma) and :176-180 (P = J sigma F^-T) */
static void stress_nh(const double *F, double Kb, double mu, int dim, double *P)
{
    double lam = lame_lambda(Kb, mu, dim);
    double J = det_d(F, dim);
    double Js = J > 1e-8 ? J : 1e-8;
    double lnJ = log(Js);
    double B[9], sig[9], Finv[9];
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) {
            double s = 0.0;
            for (int k = 0; k < dim; ++k) s += F[r * dim + k] * F[c * dim + k];
            B[r * dim + c] = s;
        }
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) {
            double I = (r == c) ? 1.0 : 0.0;
            sig[r * dim + c] = (1.0 / Js) * (mu * (B[r * dim + c] - I) + lam * lnJ * I);
        }
    inv_d(F, J, Finv, dim);
    /* P = Js * sigma @ F^-T */
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) {
            double s = 0.0;
            for (int k = 0; k < dim; ++k) s += sig[r * dim + k] * Finv[c * dim + k];
            P[r * dim + c] = Js * s;
        }
}

/* J2 small-strain radial return, fem/material/j2_plasticity.py:202-249, with an
 * optional linear kinematic term (h_kin = 0 reproduces the reference exactly).
 * Reads the committed state, writes stress and the trial state; never commits. */
static void stress_j2(const double *F, double Kb, double mu, double sy0, double Hi, double Hk,
                      const double *ep, double epbar, const double *beta, double *P, double *ep_t,
                      double *epbar_t, double *beta_t)
{
    double lam = Kb - 2.0 * mu / 3.0;
    double eps[9], ee[9], sig[9], s[9], eta[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            eps[r * 3 + c] = 0.5 * (F[r * 3 + c] + F[c * 3 + r]) - (r == c ? 1.0 : 0.0);
    for (int m = 0; m < 9; ++m) ee[m] = eps[m] - ep[m];
    double tre = ee[0] + ee[4] + ee[8];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            sig[r * 3 + c] = lam * tre * (r == c ? 1.0 : 0.0) + 2.0 * mu * ee[r * 3 + c];
    double pr = (sig[0] + sig[4] + sig[8]) / 3.0;
    for (int m = 0; m < 9; ++m) s[m] = sig[m] - ((m % 4) == 0 ? pr : 0.0);
    for (int m = 0; m < 9; ++m) eta[m] = s[m] - beta[m];
    double nsq = 0.0;
    for (int m = 0; m < 9; ++m) nsq += eta[m] * eta[m];
    double nrm = sqrt(nsq > 1e-30 ? nsq : 1e-30);
    double q = sqrt(1.5) * nrm;
    double fy = q - (sy0 + Hi * epbar);
    if (fy <= 0.0) {
        memcpy(P, sig, sizeof(double) * 9);
        memcpy(ep_t, ep, sizeof(double) * 9);
        *epbar_t = epbar;
        memcpy(beta_t, beta, sizeof(double) * 9);
        return;
    }
    double dg = fy / (3.0 * mu + Hi + Hk);
    double fac = 1.0 - 3.0 * mu * dg / q;
    for (int m = 0; m < 9; ++m) {
        double nflow = eta[m] / nrm;
        ep_t[m] = ep[m] + sqrt(1.5) * dg * nflow;
        beta_t[m] = beta[m] + sqrt(2.0 / 3.0) * Hk * dg * nflow;
    }
    *epbar_t = epbar + dg;
    /* s = s_trial - 2 mu d(eps_p) = beta_n + fac (s_trial - beta_n): the deviator
     * stays 2 mu dev(eps - eps_p) whatever the hardening; the back stress moves
     * the yield surface, not the elastic law */
    for (int m = 0; m < 9; ++m) P[m] = beta[m] + fac * eta[m] + ((m % 4) == 0 ? pr : 0.0);
}

void orc_stress(orc_pd *p)
{
    int dim = p->dim, dd = dim * dim;
#pragma omp parallel for schedule(static) num_threads(nthr(p->nthreads))
    for (int32_t i = 0; i < p->n; ++i) {
        const double *F = p->F + (int64_t)i * dd;
        double *P = p->P + (int64_t)i * dd;
        if (p->model == 1) {
            stress_nh(F, p->bulk[i], p->shear[i], dim, P);
        } else if (p->model == 2) {
            stress_j2(F, p->bulk[i], p->shear[i], p->sy0[i], p->h_iso[i],
                      p->h_kin ? p->h_kin[i] : 0.0, p->ep + (int64_t)i * 9, p->epbar[i],
                      p->beta + (int64_t)i * 9, P, p->ep_t + (int64_t)i * 9, p->epbar_t + i,
                      p->beta_t + (int64_t)i * 9);
        } else {
            stress_le(F, p->bulk[i], p->shear[i], dim, P);
        }
        matmul_d(P, p->Kinv + (int64_t)i * dd, p->A + (int64_t)i * dd, dim);
    }
}

/* J2 commit: trial -> committed (done once per converged load increment) */
void orc_commit(orc_pd *p)
{
    if (p->model != 2) return;
    memcpy(p->ep, p->ep_t, sizeof(double) * 9 * (size_t)p->n);
    memcpy(p->epbar, p->epbar_t, sizeof(double) * (size_t)p->n);
    memcpy(p->beta, p->beta_t, sizeof(double) * 9 * (size_t)p->n);
}

/* ------------------------------------------------------------------ */
/* a6: force gather  (peridynamics/core/nosb.py:229-275)                */
/* a10: bond-based PMB variant (peridynamics/solver/quasi_static.py:119-143) */
/* The stabilisation modulus of a bond is (c_i + c_j)/2; with one       */
/* material this is the reference's scalar bond_constant.               */
/* External force density fext is added afterwards (LoadControl.apply,  */
/* peridynamics/solver/quasi_static.py:372-377).                        */
/* ------------------------------------------------------------------ */
void orc_force(orc_pd *p)
{
    int dim = p->dim, dd = dim * dim;
#pragma omp parallel for schedule(static) num_threads(nthr(p->nthreads))
    for (int32_t i = 0; i < p->n; ++i) {
        double fi[3] = {0, 0, 0};
        const double *Ai = p->A + (int64_t)i * dd;
        double Ai_loc[9];
        if (p->formulation == 0 && p->per_bond_recompute)
            matmul_d(p->P + (int64_t)i * dd, p->Kinv + (int64_t)i * dd, Ai_loc, dim);
        for (int64_t b = p->row[i]; b < p->row[i + 1]; ++b) {
            if (p->broken[b]) continue;
            int64_t j = p->col[b];
            const double *xi = p->xi + b * dim;
            double w = p->omega[b], Vj = p->vol[j], L = p->xlen[b];
            double eta[3], el = 0.0;
            for (int d = 0; d < dim; ++d) {
                eta[d] = p->x[j * dim + d] - p->x[(int64_t)i * dim + d];
                el += eta[d] * eta[d];
            }
            el = sqrt(el);
            double fb[3] = {0, 0, 0};
            double cij = 0.5 * (p->cstab[i] + p->cstab[j]);
            if (p->formulation == 0) {
                const double *Aj = p->A + j * dd;
                const double *Aiu = Ai;
                double Aj_loc[9];
                if (p->per_bond_recompute) {
                    matmul_d(p->P + j * dd, p->Kinv + j * dd, Aj_loc, dim);
                    Aj = Aj_loc;
                    Aiu = Ai_loc;
                }
                for (int r = 0; r < dim; ++r) {
                    double ti = 0.0, tj = 0.0;
                    for (int c = 0; c < dim; ++c) {
                        ti += Aiu[r * dim + c] * xi[c];
                        tj += Aj[r * dim + c] * (-xi[c]);
                    }
                    /* f_corr = (t_ij - t_ji) V_j, nosb.py:250-264 */
                    fb[r] = (w * ti - w * tj) * Vj;
                }
                if (el > 1e-10 && L > 1e-10) {
                    double s = (el - L) / L;
                    for (int r = 0; r < dim; ++r) fb[r] += p->G_s * cij * s * w * (eta[r] / el) * Vj;
                }
            } else {
                if (el > 1e-10) {
                    double s = (el - L) / L;
                    double fm = cij * s * w * Vj;
                    for (int r = 0; r < dim; ++r) fb[r] = fm * (eta[r] / el);
                }
            }
            for (int r = 0; r < dim; ++r) fi[r] += fb[r];
        }
        for (int r = 0; r < dim; ++r)
            p->f[(int64_t)i * dim + r] = fi[r] + (p->fext ? p->fext[(int64_t)i * dim + r] : 0.0);
    }
}

/* NOSBSolver.compute_forces (peridynamics/solver/nosb_solver.py:200-216) */
void orc_compute_forces(orc_pd *p)
{
    if (p->formulation == 0) {
        orc_deformation_gradient(p);
        orc_stress(p);
    }
    orc_force(p);
}

/* _compute_residual_norm (nosb_solver.py:190-198), over free DOFs */
double orc_residual_norm(const orc_pd *p)
{
    int dim = p->dim;
    double s = 0.0;
    for (int32_t i = 0; i < p->n; ++i)
        for (int d = 0; d < dim; ++d)
            if (!((p->fixed[i] >> d) & 1)) {
                double t = p->f[(int64_t)i * dim + d];
                s += t * t;
            }
    return sqrt(s);
}

/* ------------------------------------------------------------------ */
/* a7: stable dt estimate (nosb_solver.py:94-154)                       */
/* ------------------------------------------------------------------ */
double orc_estimate_dt(const orc_pd *p, double bulk, double shear, double rho_mean)
{
    int dim = p->dim, dd = dim * dim;
    double lam = dim == 3 ? bulk - 2.0 * shear / 3.0 : bulk - shear;
    double modulus = lam + 2.0 * shear;
    double lmax = 0.0;
    for (int32_t i = 0; i < p->n; ++i) {
        if (p->mass[i] < 1e-30) continue;
        const double *Ki = p->Kinv + (int64_t)i * dd;
        double ds[3] = {0, 0, 0}, dsq = 0.0;
        for (int64_t b = p->row[i]; b < p->row[i + 1]; ++b) {
            if (p->broken[b]) continue;
            const double *xi = p->xi + b * dim;
            double w = p->omega[b];
            for (int r = 0; r < dim; ++r) {
                double t = 0.0;
                for (int c = 0; c < dim; ++c) t += Ki[r * dim + c] * xi[c];
                t *= w;
                ds[r] += t;
                dsq += t * t;
            }
        }
        double s2 = 0.0;
        for (int r = 0; r < dim; ++r) s2 += ds[r] * ds[r];
        double keff = modulus * p->vol[i] * (s2 + dsq);
        double li = keff / p->mass[i];
        if (li > lmax) lmax = li;
    }
    if (lmax > 0) return 0.5 * (2.0 / sqrt(lmax));
    double avg_vol = 0.0;
    for (int32_t i = 0; i < p->n; ++i) avg_vol += p->vol[i];
    avg_vol /= p->n;
    double c_wave = sqrt((bulk + 4 * shear / 3) / (rho_mean + 1e-20));
    double spacing = dim == 3 ? cbrt(avg_vol) : sqrt(avg_vol);
    return 0.1 * spacing / (c_wave + 1e-20);
}

/* ------------------------------------------------------------------ */
/* a7: kinetic-damping dynamic relaxation, reference-faithful            */
/* NOSBSolver.solve/step (nosb_solver.py:156-188, 218-256, 262-332).     */
/* out[0]=residual out[1]=rel out[2]=ref out[3]=resets; returns iters,   */
/* negated when not converged.                                          */
/* ------------------------------------------------------------------ */
int64_t orc_kd_solve(orc_pd *p, double dt, double visc, int64_t max_it, double tol,
                     double ref_in, double *out)
{
    int dim = p->dim;
    int64_t n = p->n;
    for (int64_t k = 0; k < n * dim; ++k) p->v[k] = 0.0;
    orc_compute_forces(p);
    double ref = ref_in > 0 ? ref_in : orc_residual_norm(p);
    if (ref < 1e-20) ref = 1.0;
    double prev_ke = 0.0;
    int inc = 1;
    int64_t resets = 0, it = 0;
    double res = 0.0, rel = 1.0;
    int conv = 0;
    double damping = 1.0 - visc;
    for (it = 0; it < max_it; ++it) {
#pragma omp parallel for schedule(static) num_threads(nthr(p->nthreads))
        for (int64_t i = 0; i < n; ++i)
            for (int d = 0; d < dim; ++d)
                if (!((p->fixed[i] >> d) & 1))
                    p->x[i * dim + d] += p->v[i * dim + d] * dt + 0.5 * p->a[i * dim + d] * dt * dt;
        orc_compute_forces(p);
        double ke = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            double m = p->mass[i];
            for (int d = 0; d < dim; ++d) {
                int64_t k = i * dim + d;
                if (!((p->fixed[i] >> d) & 1) && m > 1e-20) {
                    double an = p->f[k] / m;
                    p->v[k] += 0.5 * (p->a[k] + an) * dt;
                    p->v[k] *= damping;
                    p->a[k] = an;
                    ke += 0.5 * m * p->v[k] * p->v[k];
                } else {
                    p->v[k] = 0.0;
                    p->a[k] = 0.0;
                }
            }
        }
        if (ke < prev_ke && inc) {
            for (int64_t k = 0; k < n * dim; ++k) p->v[k] = 0.0;
            ke = 0.0;
            inc = 0;
            resets++;
        } else if (ke > prev_ke) {
            inc = 1;
        }
        prev_ke = ke;
        res = orc_residual_norm(p);
        rel = res / ref;
        if (rel < tol) { conv = 1; it++; break; }
    }
    if (out) { out[0] = res; out[1] = rel; out[2] = ref; out[3] = (double)resets; }
    return conv ? it : -it;
}

/* ------------------------------------------------------------------ */
/* Explicit dynamics with a per-step damage schedule:                   */
/* ExplicitSolver.step (peridynamics/solver/explicit.py:62-71, 120-158) */
/* followed by DamageModel.step (core/damage.py:94-152) when s_c > 0.   */
/* Velocity Verlet: a_old = a; x += v dt + a dt^2 / 2 on free DOFs;      */
/* forces; a = f / m (m > 0); v = (1 - damping)(v + (a_old + a) dt / 2); */
/* fixed DOFs v = a = 0.  Starts from the current x, v, a.  Returns the  */
/* bonds broken over the run; out[0] = kinetic energy sum m v^2 / 2     */
/* after the last step (explicit.py:226-233).                           */
/* ------------------------------------------------------------------ */
int64_t orc_update_damage(orc_pd *p, double s_c);

int64_t orc_explicit_run(orc_pd *p, double dt, double damping, int64_t n_steps, double s_c, double *out)
{
    int dim = p->dim;
    int64_t n = p->n, broken = 0;
    double *a_old = (double *)malloc(sizeof(double) * (size_t)(n * dim));
    for (int64_t step = 0; step < n_steps; ++step) {
#pragma omp parallel for schedule(static) num_threads(nthr(p->nthreads))
        for (int64_t i = 0; i < n; ++i)
            for (int d = 0; d < dim; ++d) {
                int64_t k = i * dim + d;
                a_old[k] = p->a[k];
                if (!((p->fixed[i] >> d) & 1)) p->x[k] += p->v[k] * dt + 0.5 * p->a[k] * dt * dt;
            }
        orc_compute_forces(p);
#pragma omp parallel for schedule(static) num_threads(nthr(p->nthreads))
        for (int64_t i = 0; i < n; ++i) {
            double m = p->mass[i];
            for (int d = 0; d < dim; ++d) {
                int64_t k = i * dim + d;
                if (!((p->fixed[i] >> d) & 1)) {
                    if (m > 0.0) p->a[k] = p->f[k] / m;
                    p->v[k] = (1.0 - damping) * (p->v[k] + 0.5 * (a_old[k] + p->a[k]) * dt);
                } else {
                    p->v[k] = 0.0;
                    p->a[k] = 0.0;
                }
            }
        }
        if (s_c > 0.0) broken += orc_update_damage(p, s_c);
    }
    if (out) {
        double ke = 0.0;
        for (int64_t i = 0; i < n; ++i)
            for (int d = 0; d < dim; ++d) ke += 0.5 * p->mass[i] * p->v[i * dim + d] * p->v[i * dim + d];
        out[0] = ke;
    }
    free(a_old);
    return broken;
}

/* ------------------------------------------------------------------ */
/* Adaptive dynamic relaxation (no reference counterpart; Kilic & Madenci */
/* 2010, Theor. Appl. Fract. Mech. 53:194-204) -- the same algorithm the */
/* HIP solver runs, here in FP64.                                       */
/*                                                                      */
/* Diagonal fictitious mass lam_i = beta/4 * (k_corr_i + k_stab_i), a    */
/* Gershgorin bound of the force-density stiffness:                      */
/*   g_ik = w_ik V_k Kinv_i xi_ik, G_i = sum_k g_ik,                     */
/*   S_i = sum_k |g_ik| + |G_i|,  C_i = max(3K_i, 2mu_i)                 */
/*   k_corr_i = sqrt3 (C_i S_i |G_i| + sum_{j in N(i)} w V_j C_j S_j |Kinv_j xi_ji|) */
/*   k_stab_i = 2 sqrt3 G_s sum_j c_ij w V_j / |xi|                      */
/* bond-based formulation: k_corr = 0 and G_s -> 1.                      */
/* ------------------------------------------------------------------ */
void orc_adr_mass(const orc_pd *p, double beta, double *lam)
{
    int dim = p->dim, dd = dim * dim;
    int64_t n = p->n;
    double *S = (double *)calloc((size_t)n, sizeof(double));
    double *CS = (double *)calloc((size_t)n, sizeof(double));
    double *Gn = (double *)calloc((size_t)n, sizeof(double));
    const double r3 = sqrt(3.0);
    if (p->formulation == 0) {
#pragma omp parallel for schedule(static) num_threads(nthr(p->nthreads))
        for (int64_t i = 0; i < n; ++i) {
            const double *Ki = p->Kinv + i * dd;
            double G[3] = {0, 0, 0}, sabs = 0.0;
            for (int64_t b = p->row[i]; b < p->row[i + 1]; ++b) {
                if (p->broken[b]) continue;
                int64_t j = p->col[b];
                const double *xi = p->xi + b * dim;
                double g[3], gn = 0.0;
                for (int r = 0; r < dim; ++r) {
                    double t = 0.0;
                    for (int c = 0; c < dim; ++c) t += Ki[r * dim + c] * xi[c];
                    g[r] = p->omega[b] * p->vol[j] * t;
                    G[r] += g[r];
                    gn += g[r] * g[r];
                }
                sabs += sqrt(gn);
            }
            double gl = 0.0;
            for (int r = 0; r < dim; ++r) gl += G[r] * G[r];
            gl = sqrt(gl);
            double C = 3.0 * p->bulk[i] > 2.0 * p->shear[i] ? 3.0 * p->bulk[i] : 2.0 * p->shear[i];
            S[i] = sabs + gl;
            CS[i] = C * S[i];
            Gn[i] = gl;
        }
    }
    double gs = p->formulation == 0 ? p->G_s : 1.0;
#pragma omp parallel for schedule(static) num_threads(nthr(p->nthreads))
    for (int64_t i = 0; i < n; ++i) {
        double kc = 0.0, ks = 0.0;
        if (p->formulation == 0) kc = CS[i] * Gn[i];
        for (int64_t b = p->row[i]; b < p->row[i + 1]; ++b) {
            if (p->broken[b]) continue;
            int64_t j = p->col[b];
            const double *xi = p->xi + b * dim;
            double w = p->omega[b], Vj = p->vol[j], L = p->xlen[b];
            if (p->formulation == 0) {
                const double *Kj = p->Kinv + j * dd;
                double gn = 0.0;
                for (int r = 0; r < dim; ++r) {
                    double t = 0.0;
                    for (int c = 0; c < dim; ++c) t += Kj[r * dim + c] * (-xi[c]);
                    gn += t * t;
                }
                kc += w * Vj * CS[j] * sqrt(gn);
            }
            if (L > 1e-10) ks += 0.5 * (p->cstab[i] + p->cstab[j]) * w * Vj / L;
        }
        lam[i] = 0.25 * beta * (r3 * kc + 2.0 * r3 * gs * ks);
    }
    free(S);
    free(CS);
    free(Gn);
}

/* ADR iteration.  State: x (u = x - X), v = V^{n-1/2}, a = F^{n-1}.
 * Step n (F^n = f(u^n) + fext already computed):
 *   rel = |F^n|_free / ref ; stop when rel < tol (solution u^n)
 *   n == 0: V^{1/2} = F^0 / (2 lam)
 *   n >= 1: V^{n+1/2} = ((2 - c) V^{n-1/2} + 2 F^n / lam) / (2 + c), with
 *           c the damping computed at step n-1 (lagged one step)
 *   c_next = 2 sqrt(sum u k u / sum u u),  k = -(F^n - F^{n-1}) / (lam V^{n-1/2})
 *   u^{n+1} = u^n + V^{n+1/2}
 * out[0]=residual out[1]=rel out[2]=ref out[3]=last damping */
int64_t orc_adr_solve_mode(orc_pd *p, const double *lam, int64_t max_it, double tol, double ref_in,
                           double cmax, int32_t mode, double *out)
{
    int dim = p->dim;
    int64_t n = p->n;
    orc_compute_forces(p);
    double ref = ref_in > 0 ? ref_in : orc_residual_norm(p);
    if (ref < 1e-20) ref = 1.0;
    double c = 0.0, res = 0.0, rel = 1.0;
    int conv = 0;
    int64_t it;
    for (it = 0; it < max_it; ++it) {
        res = orc_residual_norm(p);
        rel = res / ref;
        if (rel < tol) { conv = 1; break; }
        double num = 0.0, den = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            for (int d = 0; d < dim; ++d) {
                if ((p->fixed[i] >> d) & 1) continue;
                int64_t k = i * dim + d;
                double Fk = p->f[k];
                double u = p->x[k] - p->X[k];
                double vold = p->v[k];
                double vnew;
                if (it == 0) {
                    vnew = Fk / (2.0 * lam[i]);
                } else {
                    if (mode == 0) {
                        if (vold != 0.0) {
                            double kk = -(Fk - p->a[k]) / (lam[i] * vold);
                            num += u * kk * u;
                        }
                        den += u * u;
                    } else {
                        num -= vold * (Fk - p->a[k]);
                        den += lam[i] * vold * vold;
                    }
                    vnew = ((2.0 - c) * vold + 2.0 * Fk / lam[i]) / (2.0 + c);
                }
                p->v[k] = vnew;
                p->a[k] = Fk;
            }
        }
        if (it > 0) {
            double cn = (den > 0.0 && num > 0.0) ? 2.0 * sqrt(num / den) : 0.0;
            c = cn < cmax ? cn : cmax;
        }
#pragma omp parallel for schedule(static) num_threads(nthr(p->nthreads))
        for (int64_t i = 0; i < n; ++i)
            for (int d = 0; d < dim; ++d)
                if (!((p->fixed[i] >> d) & 1)) p->x[i * dim + d] += p->v[i * dim + d];
        orc_compute_forces(p);
    }
    if (out) { out[0] = res; out[1] = rel; out[2] = ref; out[3] = c; }
    return conv ? it : -it;
}

int64_t orc_adr_solve(orc_pd *p, const double *lam, int64_t max_it, double tol, double ref_in,
                      double cmax, double *out)
{
    return orc_adr_solve_mode(p, lam, max_it, tol, ref_in, cmax, 0, out);
}

/* ------------------------------------------------------------------ */
/* a9: damage  (peridynamics/core/damage.py:94-142, bonds.py:161-180)   */
/* ------------------------------------------------------------------ */
int64_t orc_update_damage(orc_pd *p, double s_c)
{
    int dim = p->dim;
    int64_t nb = 0;
#pragma omp parallel for schedule(static) reduction(+ : nb) num_threads(nthr(p->nthreads))
    for (int32_t i = 0; i < p->n; ++i) {
        for (int64_t b = p->row[i]; b < p->row[i + 1]; ++b) {
            if (p->broken[b]) continue;
            int64_t j = p->col[b];
            double s2 = 0.0;
            for (int d = 0; d < dim; ++d) {
                double t = p->x[j * dim + d] - p->x[(int64_t)i * dim + d];
                s2 = s2 + t * t;
            }
            double el = sqrt(s2), L = p->xlen[b];
            double st = 0.0;
            if (L > 1e-10) st = (el - L) / L;
            if (st > s_c) { p->broken[b] = 1; nb++; }
        }
    }
    return nb;
}

void orc_local_damage(const orc_pd *p, const int32_t *initial, double *damage)
{
    for (int32_t i = 0; i < p->n; ++i) {
        if (initial[i] > 0) {
            int32_t c = 0;
            for (int64_t b = p->row[i]; b < p->row[i + 1]; ++b) c += p->broken[b] ? 1 : 0;
            damage[i] = (double)c / (double)initial[i];
        } else {
            damage[i] = 0.0;
        }
    }
}
