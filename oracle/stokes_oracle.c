/*
 * CPU FP64 oracle for the Stokes/Brinkman topology-optimization hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the results tests compare against and as the
 * CPU baseline.  The
 * product path (stokes_to_amd/, include/) never links or calls it.
 *
 * Every operator here is written as the reference composes it: separate DEC
 * primitives (divergence D2, gradient Cod0, face->edge curl Cod1, edge->face curl
 * D1) with explicit masking passes between them, exactly in the order of
 * ST/StokeFlowMapping3D.cu:176-306.  The GPU path instead evaluates one fused
 * stencil; agreement of the two is the parity evidence.
 *
 * Index convention: C-order (i, j, k), k fastest; values outside an array's
 * index range read as zero ("zero extension"), as the reference kernels do with
 * their face_valid / edge_valid guards (CX/grid3D.cu:124-146).
 */
#include "stokes_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OMP_FOR _Pragma("omp parallel for schedule(static)")

typedef struct {
    int dim, nx, ny, nz;
    long nc, nf[3], fo[3], nft, n; /* fo = offsets of face blocks inside the face part */
} G;

static G mkg(int dim, int nx, int ny, int nz)
{
    G g;
    g.dim = dim; g.nx = nx; g.ny = ny; g.nz = (dim == 2) ? 1 : nz;
    nz = g.nz;
    g.nc = (long)nx * ny * nz;
    g.nf[0] = (long)(nx + 1) * ny * nz;
    g.nf[1] = (long)nx * (ny + 1) * nz;
    g.nf[2] = (dim == 3) ? (long)nx * ny * (nz + 1) : 0;
    g.fo[0] = 0; g.fo[1] = g.nf[0]; g.fo[2] = g.nf[0] + g.nf[1];
    g.nft = g.nf[0] + g.nf[1] + g.nf[2];
    g.n = g.nc + g.nft;
    return g;
}

long orc_dof(int dim, int nx, int ny, int nz, long *s)
{
    G g = mkg(dim, nx, ny, nz);
    if (s) { s[0] = g.nc; s[1] = g.nf[0]; s[2] = g.nf[1]; s[3] = g.nf[2]; }
    return g.n;
}

/* dims of face array d and edge array d */
static inline void fdims(const G *g, int d, int *a, int *b, int *c)
{
    *a = g->nx + (d == 0); *b = g->ny + (d == 1); *c = g->nz + (d == 2 && g->dim == 3);
}
static inline void edims(const G *g, int d, int *a, int *b, int *c)
{
    if (g->dim == 2) { *a = g->nx + 1; *b = g->ny + 1; *c = 1; return; } /* nodes = z-edges */
    *a = g->nx + (d != 0); *b = g->ny + (d != 1); *c = g->nz + (d != 2);
}
long orc_edge_count(int dim, int nx, int ny, int nz)
{
    G g = mkg(dim, nx, ny, nz);
    if (dim == 2) return (long)(nx + 1) * (ny + 1);
    long s = 0; int a, b, c;
    for (int d = 0; d < 3; d++) { edims(&g, d, &a, &b, &c); s += (long)a * b * c; }
    return s;
}
static inline long eoff(const G *g, int d)
{
    if (g->dim == 2) return 0;
    long s = 0; int a, b, c;
    for (int q = 0; q < d; q++) { edims(g, q, &a, &b, &c); s += (long)a * b * c; }
    return s;
}

static inline double at(const double *v, int A, int B, int C, int i, int j, int k)
{
    if (i < 0 || j < 0 || k < 0 || i >= A || j >= B || k >= C) return 0.0;
    return v[((long)i * B + j) * C + k];
}
static inline double cget(const G *g, const double *c, int i, int j, int k)
{
    return at(c, g->nx, g->ny, g->nz, i, j, k);
}
static inline double fget(const G *g, const double *f, int d, int i, int j, int k)
{
    if (d >= g->dim) return 0.0;
    int a, b, c; fdims(g, d, &a, &b, &c);
    return at(f + g->fo[d], a, b, c, i, j, k);
}
static inline double eget(const G *g, const double *e, int d, int i, int j, int k)
{
    if (g->dim == 2 && d != 2) return 0.0;
    int a, b, c; edims(g, d, &a, &b, &c);
    return at(e + eoff(g, d), a, b, c, i, j, k);
}

/* ------------------------------------------------------------------ DEC */

/* D2: faces -> cells, c = sum_d f_d(c+e_d) - f_d(c)   (CX/grid3D.cu:330-390) */
static void d2(const G *g, const double *f, double *c)
{
    OMP_FOR
    for (int i = 0; i < g->nx; i++)
        for (int j = 0; j < g->ny; j++)
            for (int k = 0; k < g->nz; k++) {
                double s = fget(g, f, 0, i + 1, j, k) - fget(g, f, 0, i, j, k)
                         + fget(g, f, 1, i, j + 1, k) - fget(g, f, 1, i, j, k);
                if (g->dim == 3) s += fget(g, f, 2, i, j, k + 1) - fget(g, f, 2, i, j, k);
                c[((long)i * g->ny + j) * g->nz + k] = s;
            }
}
/* Cod0: cells -> faces, f_d(i) = c(i) - c(i-e_d)   (CX/grid3D.cu:401-456) */
static void cod0(const G *g, const double *c, double *f)
{
    for (int d = 0; d < g->dim; d++) {
        int A, B, C; fdims(g, d, &A, &B, &C);
        double *o = f + g->fo[d];
        OMP_FOR
        for (int i = 0; i < A; i++)
            for (int j = 0; j < B; j++)
                for (int k = 0; k < C; k++)
                    o[((long)i * B + j) * C + k] =
                        cget(g, c, i, j, k) - cget(g, c, i - (d == 0), j - (d == 1), k - (d == 2));
    }
}
/* Cod1: faces -> edges (CX/grid3D.cu:469-568) */
static void cod1(const G *g, const double *f, double *e)
{
    for (int d = (g->dim == 2 ? 2 : 0); d < 3; d++) {
        int A, B, C; edims(g, d, &A, &B, &C);
        double *o = e + eoff(g, d);
        OMP_FOR
        for (int i = 0; i < A; i++)
            for (int j = 0; j < B; j++)
                for (int k = 0; k < C; k++) {
                    double v;
                    if (d == 0)
                        v = fget(g, f, 1, i, j, k) - fget(g, f, 1, i, j, k - 1)
                          - fget(g, f, 2, i, j, k) + fget(g, f, 2, i, j - 1, k);
                    else if (d == 1)
                        v = fget(g, f, 2, i, j, k) - fget(g, f, 2, i - 1, j, k)
                          - fget(g, f, 0, i, j, k) + fget(g, f, 0, i, j, k - 1);
                    else
                        v = fget(g, f, 0, i, j, k) - fget(g, f, 0, i, j - 1, k)
                          - fget(g, f, 1, i, j, k) + fget(g, f, 1, i - 1, j, k);
                    o[((long)i * B + j) * C + k] = v;
                }
    }
}
/* D1: edges -> faces (CX/grid3D.cu:219-318), D1 = Cod1^T */
static void d1(const G *g, const double *e, double *f)
{
    for (int d = 0; d < g->dim; d++) {
        int A, B, C; fdims(g, d, &A, &B, &C);
        double *o = f + g->fo[d];
        OMP_FOR
        for (int i = 0; i < A; i++)
            for (int j = 0; j < B; j++)
                for (int k = 0; k < C; k++) {
                    double v;
                    if (d == 0)
                        v = eget(g, e, 1, i, j, k + 1) - eget(g, e, 1, i, j, k)
                          + eget(g, e, 2, i, j, k) - eget(g, e, 2, i, j + 1, k);
                    else if (d == 1)
                        v = eget(g, e, 0, i, j, k) - eget(g, e, 0, i, j, k + 1)
                          + eget(g, e, 2, i + 1, j, k) - eget(g, e, 2, i, j, k);
                    else
                        v = eget(g, e, 0, i, j + 1, k) - eget(g, e, 0, i, j, k)
                          + eget(g, e, 1, i, j, k) - eget(g, e, 1, i + 1, j, k);
                    o[((long)i * B + j) * C + k] = v;
                }
    }
}

void orc_d2(int dim, int nx, int ny, int nz, const double *f, double *c)
{ G g = mkg(dim, nx, ny, nz); d2(&g, f, c); }
void orc_cod0(int dim, int nx, int ny, int nz, const double *c, double *f)
{ G g = mkg(dim, nx, ny, nz); cod0(&g, c, f); }
void orc_cod1(int dim, int nx, int ny, int nz, const double *f, double *e)
{ G g = mkg(dim, nx, ny, nz); cod1(&g, f, e); }
void orc_d1(int dim, int nx, int ny, int nz, const double *e, double *f)
{ G g = mkg(dim, nx, ny, nz); d1(&g, e, f); }

/* ------------------------------------------------------------ operators */

static double *dalloc(long n) { double *p = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); return p; }

/* the velocity-block terms shared by the operator and the objective operator:
 * y_f += -(M_f Cod0 M_c D2 V u) + V D1 Cod1 M_f u        (ST/StokeFlowMapping3D.cu:194-258) */
static void velocity_terms(const G *g, const uint8_t *fixed, const double *vol, const double *x,
                           double *y)
{
    const uint8_t *ffix = fixed + g->nc;
    const double *fvol = vol + g->nc;
    const double *xf = x + g->nc;
    double *yf = y + g->nc;
    double *tf = dalloc(g->nft), *tc = dalloc(g->nc);
    double *te = dalloc(orc_edge_count(g->dim, g->nx, g->ny, g->nz));
    /* grad div v */
    OMP_FOR for (long q = 0; q < g->nft; q++) tf[q] = xf[q] * fvol[q];
    d2(g, tf, tc);
    OMP_FOR for (long q = 0; q < g->nc; q++) if (fixed[q]) tc[q] = 0.0;
    cod0(g, tc, tf);
    OMP_FOR for (long q = 0; q < g->nft; q++) { if (ffix[q]) tf[q] = 0.0; yf[q] -= tf[q]; }
    /* - curl curl v */
    OMP_FOR for (long q = 0; q < g->nft; q++) tf[q] = ffix[q] ? 0.0 : xf[q];
    cod1(g, tf, te);
    d1(g, te, tf);
    OMP_FOR for (long q = 0; q < g->nft; q++) yf[q] += tf[q] * fvol[q];
    free(tf); free(tc); free(te);
}

void orc_apply(int dim, int nx, int ny, int nz, const uint8_t *fixed, const double *vol,
               const double *pen, const double *x, double *y)
{
    G g = mkg(dim, nx, ny, nz);
    const double *fvol = vol + g.nc;
    memset(y, 0, sizeof(double) * g.n);
    velocity_terms(&g, fixed, vol, x, y);
    double *tf = dalloc(g.nft), *tc = dalloc(g.nc);
    /* -grad p */
    cod0(&g, x, tf);
    OMP_FOR for (long q = 0; q < g.nft; q++) y[g.nc + q] -= tf[q] * fvol[q];
    /* div v */
    OMP_FOR for (long q = 0; q < g.nft; q++) tf[q] = x[g.nc + q] * fvol[q];
    d2(&g, tf, tc);
    OMP_FOR for (long q = 0; q < g.nc; q++) y[q] += fixed[q] ? 0.0 : tc[q];
    /* penalty, then definite fix (ST/StokeFlowMapping3D.cu:292-305) */
    OMP_FOR for (long q = 0; q < g.n; q++) { y[q] += x[q] * pen[q]; if (fixed[q]) y[q] += x[q]; }
    free(tf); free(tc);
}

void orc_energy_op(int dim, int nx, int ny, int nz, const uint8_t *fixed, const double *vol,
                   const double *pen, const double *x, double *y)
{
    /* the reference's 2D objective is another operator than its 3D one: the face
     * 5-point Laplacian of FT/FluidEnergyGrad.cu:30-61 (Appendix B 5) */
    if (dim == 2) {
        orc_energy_op_ref2d(nx, ny, fixed, pen, x, y);
        return;
    }
    G g = mkg(dim, nx, ny, nz);
    memset(y, 0, sizeof(double) * g.n);
    velocity_terms(&g, fixed, vol, x, y);
    OMP_FOR for (long q = 0; q < g.n; q++) y[q] += x[q] * pen[q];
}

double orc_objective(int dim, int nx, int ny, int nz, const uint8_t *fixed, const double *vol,
                     const double *pen, const double *x)
{
    G g = mkg(dim, nx, ny, nz);
    double *y = dalloc(g.n);
    orc_energy_op(dim, nx, ny, nz, fixed, vol, pen, x, y);
    double s = 0.0;
    for (long q = 0; q < g.n; q++) s += x[q] * y[q]; /* FT/FluidEnergy3D.cpp:106-108 */
    free(y);
    return 0.5 * s;
}

/* ---- literal 2D reference forms (CX/form01mapping.cu) ---- */
static double a2(const double *v, int A, int B, int i, int j)
{
    if (i < 0 || j < 0 || i >= A || j >= B) return 0.0;
    return v[(long)i * B + j];
}
void orc_apply_ref2d(int nx, int ny, const uint8_t *fixed, const double *vol, const double *pen,
                     const double *x, double *y)
{
    long nc = (long)nx * ny, nfx = (long)(nx + 1) * ny, nfy = (long)nx * (ny + 1), n = nc + nfx + nfy;
    const double *xfx = x + nc, *xfy = x + nc + nfx;
    const double *vx = vol + nc, *vy = vol + nc + nfx;
    const uint8_t *fx = fixed + nc, *fy = fixed + nc + nfx;
    double *tx = dalloc(nfx), *ty = dalloc(nfy), *tc = dalloc(nc), *nd = dalloc((long)(nx + 1) * (ny + 1));
    memset(y, 0, sizeof(double) * n);
    double *yfx = y + nc, *yfy = y + nc + nfx;
    /* grad div: t=(u*vol, -v*vol); c=D1(t) (y-flipped div); c*=!fixed; c=-c; f=Cod0(c) */
    for (long q = 0; q < nfx; q++) tx[q] = xfx[q] * vx[q];
    for (long q = 0; q < nfy; q++) ty[q] = -xfy[q] * vy[q];
    for (int i = 0; i < nx; i++) for (int j = 0; j < ny; j++) {
        double dv = a2(tx, nx + 1, ny, i + 1, j) - a2(tx, nx + 1, ny, i, j)
                  + a2(ty, nx, ny + 1, i, j) - a2(ty, nx, ny + 1, i, j + 1);
        tc[(long)i * ny + j] = fixed[(long)i * ny + j] ? 0.0 : -dv;
    }
    for (int i = 0; i <= nx; i++) for (int j = 0; j < ny; j++) {
        long q = (long)i * ny + j;
        double v = a2(tc, nx, ny, i, j) - a2(tc, nx, ny, i - 1, j);
        yfx[q] += fx[q] ? 0.0 : v;
    }
    for (int i = 0; i < nx; i++) for (int j = 0; j <= ny; j++) {
        long q = (long)i * (ny + 1) + j;
        double v = a2(tc, nx, ny, i, j) - a2(tc, nx, ny, i, j - 1);
        yfy[q] += fy[q] ? 0.0 : v;
    }
    /* curl curl: node = Cod1(raw faces); f = D0(node); f_y = -f_y; f *= vol */
    for (int i = 0; i <= nx; i++) for (int j = 0; j <= ny; j++)
        nd[(long)i * (ny + 1) + j] = a2(xfx, nx + 1, ny, i, j - 1) - a2(xfx, nx + 1, ny, i, j)
                                   + a2(xfy, nx, ny + 1, i, j) - a2(xfy, nx, ny + 1, i - 1, j);
    for (int i = 0; i <= nx; i++) for (int j = 0; j < ny; j++) {
        long q = (long)i * ny + j;
        yfx[q] += vx[q] * (a2(nd, nx + 1, ny + 1, i, j + 1) - a2(nd, nx + 1, ny + 1, i, j));
    }
    for (int i = 0; i < nx; i++) for (int j = 0; j <= ny; j++) {
        long q = (long)i * (ny + 1) + j;
        yfy[q] += vy[q] * -(a2(nd, nx + 1, ny + 1, i + 1, j) - a2(nd, nx + 1, ny + 1, i, j));
    }
    /* -grad p */
    for (int i = 0; i <= nx; i++) for (int j = 0; j < ny; j++) {
        long q = (long)i * ny + j;
        yfx[q] -= vx[q] * (a2(x, nx, ny, i, j) - a2(x, nx, ny, i - 1, j));
    }
    for (int i = 0; i < nx; i++) for (int j = 0; j <= ny; j++) {
        long q = (long)i * (ny + 1) + j;
        yfy[q] -= vy[q] * (a2(x, nx, ny, i, j) - a2(x, nx, ny, i, j - 1));
    }
    /* div */
    for (int i = 0; i < nx; i++) for (int j = 0; j < ny; j++) {
        long q = (long)i * ny + j;
        double dv = a2(tx, nx + 1, ny, i + 1, j) - a2(tx, nx + 1, ny, i, j)
                  + a2(ty, nx, ny + 1, i, j) - a2(ty, nx, ny + 1, i, j + 1);
        y[q] += fixed[q] ? 0.0 : dv;
    }
    /* penalty; the 2D reference has no definite fix */
    for (long q = 0; q < n; q++) y[q] += x[q] * pen[q];
    free(tx); free(ty); free(tc); free(nd);
}

void orc_energy_op_ref2d(int nx, int ny, const uint8_t *fixed, const double *pen, const double *x,
                         double *y)
{
    /* face 5-point Laplacian on each velocity component, zero-extended, masked rows */
    long nc = (long)nx * ny, nfx = (long)(nx + 1) * ny, nfy = (long)nx * (ny + 1), n = nc + nfx + nfy;
    const double *ux = x + nc, *uy = x + nc + nfx;
    memset(y, 0, sizeof(double) * n);
    for (int i = 0; i <= nx; i++) for (int j = 0; j < ny; j++) {
        long q = (long)i * ny + j;
        double l = 4 * ux[q] - a2(ux, nx + 1, ny, i - 1, j) - a2(ux, nx + 1, ny, i + 1, j)
                 - a2(ux, nx + 1, ny, i, j - 1) - a2(ux, nx + 1, ny, i, j + 1);
        y[nc + q] = fixed[nc + q] ? 0.0 : l;
    }
    for (int i = 0; i < nx; i++) for (int j = 0; j <= ny; j++) {
        long q = (long)i * (ny + 1) + j;
        double l = 4 * uy[q] - a2(uy, nx, ny + 1, i - 1, j) - a2(uy, nx, ny + 1, i + 1, j)
                 - a2(uy, nx, ny + 1, i, j - 1) - a2(uy, nx, ny + 1, i, j + 1);
        y[nc + nfx + q] = fixed[nc + nfx + q] ? 0.0 : l;
    }
    for (long q = 0; q < n; q++) y[q] += x[q] * pen[q];
}

/* b = lifted right-hand side for Dirichlet values xd on fixed faces:
 * free rows:  -(M_f Cod0^T... ) i.e. minus the unmasked-column operator applied to xd;
 * fixed rows: xd itself. */
void orc_lift_rhs(int dim, int nx, int ny, int nz, const uint8_t *fixed, const double *xd,
                  double *b)
{
    G g = mkg(dim, nx, ny, nz);
    uint8_t *nofix = (uint8_t *)calloc((size_t)g.n, 1);
    double *one = dalloc(g.n), *zero = dalloc(g.n), *y = dalloc(g.n);
    /* columns unmasked: operator with no fixed DOFs and unit volume on every face */
    for (long q = 0; q < g.n; q++) one[q] = 1.0;
    double *xv = dalloc(g.n);
    for (long q = g.nc; q < g.n; q++) xv[q] = fixed[q] ? xd[q] : 0.0;
    /* rows masked afterwards, exactly as the masked operator masks its rows */
    {
        double *tf = dalloc(g.nft), *tc = dalloc(g.nc);
        double *te = dalloc(orc_edge_count(dim, nx, ny, nz));
        d2(&g, xv + g.nc, tc);
        OMP_FOR for (long q = 0; q < g.nc; q++) if (fixed[q]) tc[q] = 0.0;
        /* continuity rows */
        OMP_FOR for (long q = 0; q < g.nc; q++) y[q] = tc[q];
        cod0(&g, tc, tf);
        OMP_FOR for (long q = 0; q < g.nft; q++) y[g.nc + q] = -tf[q];
        cod1(&g, xv + g.nc, te);
        d1(&g, te, tf);
        OMP_FOR for (long q = 0; q < g.nft; q++) y[g.nc + q] += tf[q];
        free(tf); free(tc); free(te);
    }
    for (long q = 0; q < g.n; q++) b[q] = fixed[q] ? xd[q] : -y[q];
    free(nofix); free(one); free(zero); free(y); free(xv);
}

/* ------------------------------------------------------ design / TO */

double orc_alpha(double rho, const double *am)
{
    /* FT/FluidTopoOpt3D.cpp:110-115 (alpha_os=am[0], alpha_us=am[1], q=am[2]) */
    double q = am[2];
    return am[0] + (am[1] - am[0]) * rho * (1.0 + q) / (rho + q);
}
double orc_dalpha(double rho, const double *am)
{
    double q = am[2];
    return (am[1] - am[0]) * ((1.0 + q) / (rho + q) - rho * (1.0 + q) / ((rho + q) * (rho + q)));
}
static double powi(double r, double p) { return (p == 1.0) ? r : pow(r, p); }

/* face alpha = sum over adjacent cells of 0.5*alpha(rho^p)*scale (FT/FluidTopoOpt3D.cpp:145-159) */
void orc_design_to_alpha(int dim, int nx, int ny, int nz, const double *rho, const double *am,
                         const uint8_t *face_fixed, double *fa)
{
    G g = mkg(dim, nx, ny, nz);
    for (int d = 0; d < dim; d++) {
        int A, B, C; fdims(&g, d, &A, &B, &C);
        double *o = fa + g.fo[d];
        const uint8_t *ff = face_fixed + g.fo[d];
        OMP_FOR
        for (int i = 0; i < A; i++)
            for (int j = 0; j < B; j++)
                for (int k = 0; k < C; k++) {
                    long q = ((long)i * B + j) * C + k;
                    double s = 0.0;
                    int ci[2][3] = {{i, j, k}, {i - (d == 0), j - (d == 1), k - (d == 2)}};
                    for (int t = 0; t < 2; t++) {
                        int a = ci[t][0], b = ci[t][1], c = ci[t][2];
                        if (a < 0 || b < 0 || c < 0 || a >= g.nx || b >= g.ny || c >= g.nz) continue;
                        s += 0.5 * am[4] * orc_alpha(powi(rho[((long)a * g.ny + b) * g.nz + c], am[3]), am);
                    }
                    o[q] = ff[q] ? 0.0 : s;
                }
    }
}

/* dJ/drho_c = 0.5*scale*alpha'(rho^p)*p*rho^(p-1) * sum_{free faces of c}(u^2/2 - u*lam)
 * (FT/FluidTopoOpt3D.cpp:170-198; fixed faces excluded since their alpha is zeroed) */
void orc_sensitivity(int dim, int nx, int ny, int nz, const double *rho, const double *am,
                     const uint8_t *face_fixed, const double *u, const double *lam, double *dj)
{
    G g = mkg(dim, nx, ny, nz);
    OMP_FOR
    for (int i = 0; i < g.nx; i++)
        for (int j = 0; j < g.ny; j++)
            for (int k = 0; k < g.nz; k++) {
                long c = ((long)i * g.ny + j) * g.nz + k;
                double s = 0.0;
                for (int d = 0; d < dim; d++) {
                    int A, B, C; fdims(&g, d, &A, &B, &C);
                    for (int t = 0; t < 2; t++) {
                        int a = i + t * (d == 0), b = j + t * (d == 1), cc = k + t * (d == 2);
                        long q = g.fo[d] + ((long)a * B + b) * C + cc;
                        if (face_fixed[q]) continue;
                        s += 0.5 * u[q] * u[q] - u[q] * lam[q];
                    }
                }
                double r = rho[c], p = am[3];
                double chain = (p == 1.0) ? 1.0 : p * pow(r, p - 1.0);
                dj[c] = 0.5 * am[4] * orc_dalpha(powi(r, p), am) * chain * s;
            }
}

/* density filter, hat weights max(0, R - |o|) (Bruns & Tortorelli 2001; Bourdin 2001) */
void orc_filter(int dim, int nx, int ny, int nz, double R, const double *in, double *out, int transpose)
{
    G g = mkg(dim, nx, ny, nz);
    int r = (int)ceil(R) - 1; if (r < 0) r = 0;
    int rz = (dim == 3) ? r : 0;
    double *W = dalloc(g.nc), *tmp = dalloc(g.nc);
    OMP_FOR
    for (int i = 0; i < g.nx; i++)
        for (int j = 0; j < g.ny; j++)
            for (int k = 0; k < g.nz; k++) {
                double w = 0.0;
                for (int a = -r; a <= r; a++) for (int b = -r; b <= r; b++) for (int c = -rz; c <= rz; c++) {
                    int ii = i + a, jj = j + b, kk = k + c;
                    if (ii < 0 || jj < 0 || kk < 0 || ii >= g.nx || jj >= g.ny || kk >= g.nz) continue;
                    double h = R - sqrt((double)(a * a + b * b + c * c));
                    if (h > 0) w += h;
                }
                W[((long)i * g.ny + j) * g.nz + k] = w;
            }
    const double *src = in;
    if (transpose) { for (long q = 0; q < g.nc; q++) tmp[q] = in[q] / W[q]; src = tmp; }
    OMP_FOR
    for (int i = 0; i < g.nx; i++)
        for (int j = 0; j < g.ny; j++)
            for (int k = 0; k < g.nz; k++) {
                double s = 0.0;
                for (int a = -r; a <= r; a++) for (int b = -r; b <= r; b++) for (int c = -rz; c <= rz; c++) {
                    int ii = i + a, jj = j + b, kk = k + c;
                    if (ii < 0 || jj < 0 || kk < 0 || ii >= g.nx || jj >= g.ny || kk >= g.nz) continue;
                    double h = R - sqrt((double)(a * a + b * b + c * c));
                    if (h > 0) s += h * src[((long)ii * g.ny + jj) * g.nz + kk];
                }
                long q = ((long)i * g.ny + j) * g.nz + k;
                out[q] = transpose ? s : s / W[q];
            }
    free(W); free(tmp);
}

/* optimality-criteria update with bisection on the volume multiplier
 * (Andreassen et al. 2011, "Efficient topology optimization in MATLAB using 88
 * lines of code"): volume measured on the filtered density. */
int orc_oc_update(int dim, int nx, int ny, int nz, double R, const double *rho, const double *dc,
                  double volfrac, double move, double eta, double rho_min, double *rn, double *change)
{
    G g = mkg(dim, nx, ny, nz);
    long n = g.nc;
    double *dv = dalloc(n), *ones = dalloc(n);
    for (long q = 0; q < n; q++) ones[q] = 1.0 / (double)n;
    orc_filter(dim, nx, ny, nz, R, ones, dv, 1); /* dV/drho = H^T(1/(n W)) */
    double l1 = 0.0, l2 = 1e9;
    int it = 0;
    while ((l2 - l1) > 1e-13 * (l1 + l2) && it < 200) {
        double lm = 0.5 * (l1 + l2), vol = 0.0;
        for (long q = 0; q < n; q++) {
            double be = -dc[q] / (dv[q] * lm);
            if (be < 0) be = 0;
            double v = rho[q] * pow(be, eta);
            double lo = rho[q] - move, hi = rho[q] + move;
            if (lo < rho_min) lo = rho_min;
            if (hi > 1.0) hi = 1.0;
            v = v < lo ? lo : (v > hi ? hi : v);
            rn[q] = v;
            vol += v * dv[q];
        }
        if (vol > volfrac) l1 = lm; else l2 = lm;
        it++;
    }
    double ch = 0.0;
    for (long q = 0; q < n; q++) { double a = fabs(rn[q] - rho[q]); if (a > ch) ch = a; }
    if (change) *change = ch;
    free(dv); free(ones);
    return it;
}

/* ------------------------------------------------------------ multigrid */

typedef struct { G g; uint8_t *fixed; double *vol; double *pen; } Lvl;
struct orc_mg {
    int L;
    Lvl lv[16];
    int nu_pre, nu_post, tang, closed;
    double omega;
    double *lu; int *piv; long nco;
    double *b[16], *x[16], *r[16], *t[16];
};

static void desc_vol(Lvl *l)
{
    for (long q = 0; q < l->g.n; q++) {
        l->vol[q] = (q >= l->g.nc && !l->fixed[q]) ? 1.0 : 0.0;
        if (l->fixed[q]) l->pen[q] = 0.0;
    }
}

/* coarse flags: a coarse DOF is free if any coincident child is free.
 * coarse face alpha: 4 x (normal-linear weighted mean of the child faces); the
 * factor 4 = (h_c/h_f)^2 keeps alpha in unit-spacing form on every level. */
void orc_coarsen(int dim, int nx, int ny, int nz, const uint8_t *ffixed, const double *fpen,
                 uint8_t *cfixed, double *cpen)
{
    G f = mkg(dim, nx, ny, nz), c = mkg(dim, nx / 2, ny / 2, (dim == 3) ? nz / 2 : 1);
    int kz = (dim == 3) ? 2 : 1;
    for (int I = 0; I < c.nx; I++) for (int J = 0; J < c.ny; J++) for (int K = 0; K < c.nz; K++) {
        long Q = ((long)I * c.ny + J) * c.nz + K;
        int allfix = 1; double s = 0.0; int cnt = 0;
        for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int e = 0; e < kz; e++) {
            long q = ((long)(2 * I + a) * f.ny + 2 * J + b) * f.nz + kz * K + e;
            allfix &= ffixed[q]; s += fpen[q]; cnt++;
        }
        cfixed[Q] = (uint8_t)allfix;
        cpen[Q] = allfix ? 0.0 : s / cnt;
    }
    for (int d = 0; d < dim; d++) {
        int A, B, C, a_, b_, c_;
        fdims(&c, d, &A, &B, &C); fdims(&f, d, &a_, &b_, &c_);
        for (int I = 0; I < A; I++) for (int J = 0; J < B; J++) for (int K = 0; K < C; K++) {
            long Q = c.nc + c.fo[d] + ((long)I * B + J) * C + K;
            int CI[3] = {I, J, K};
            int allfix = 1, zero = 0; double sw = 0.0, sa = 0.0, sh = 0.0;
            /* tangential children: 2 per tangential axis (1 along z in 2D) */
            int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            for (int on = -1; on <= 1; on++) {
                double wn = (on == 0) ? 1.0 : 0.5;
                for (int u = 0; u < 2; u++) for (int v = 0; v < ((t2 == 2 && dim == 2) ? 1 : 2); v++) {
                    int fi[3];
                    fi[d] = 2 * CI[d] + on;
                    fi[t1] = 2 * CI[t1] + u;
                    fi[t2] = (dim == 2 && t2 == 2) ? 0 : 2 * CI[t2] + v;
                    if (fi[0] < 0 || fi[1] < 0 || fi[2] < 0 || fi[0] >= a_ || fi[1] >= b_ || fi[2] >= c_) continue;
                    long q = f.nc + f.fo[d] + ((long)fi[0] * b_ + fi[1]) * c_ + fi[2];
                    if (on == 0) allfix &= ffixed[q];
                    sw += wn; sa += wn * fpen[q];
                    if (fpen[q] > 0.0) sh += wn / fpen[q]; else zero = 1;
                }
            }
            cfixed[Q] = (uint8_t)allfix;
            /* midpoint of the weighted arithmetic and harmonic means (the two Wiener
             * bounds of an effective coefficient), times (h_c/h_f)^2 = 4 */
            const double ar = sa / sw, hm = (zero || sh <= 0.0) ? 0.0 : sw / sh;
            cpen[Q] = allfix ? 0.0 : 4.0 * 0.5 * (ar + hm);
        }
    }
}

/* prolongation weights for a fine face (normal d) from coarse faces: visits the
 * (up to 8) coarse faces with their weights */
typedef struct { int I[3]; double w; } PW;
static int pweights(int dim, int d, int tang, const int fi[3], PW *out)
{
    int n = 0;
    double wn[2]; int In[2]; int nn;
    if ((fi[d] & 1) == 0) { In[0] = fi[d] / 2; wn[0] = 1.0; nn = 1; }
    else { In[0] = (fi[d] - 1) / 2; In[1] = In[0] + 1; wn[0] = wn[1] = 0.5; nn = 2; }
    int ta[2], nt = 0;
    for (int a = 0; a < dim; a++) if (a != d) ta[nt++] = a;
    int Tt[2][2]; double wt[2][2]; int ntt[2];
    for (int s = 0; s < nt; s++) {
        int t = fi[ta[s]];
        if (!tang) { Tt[s][0] = t / 2; wt[s][0] = 1.0; ntt[s] = 1; }
        else {
            Tt[s][0] = t / 2; wt[s][0] = 0.75;
            Tt[s][1] = (t & 1) ? t / 2 + 1 : t / 2 - 1; wt[s][1] = 0.25; ntt[s] = 2;
        }
    }
    for (int a = 0; a < nn; a++)
        for (int b = 0; b < ntt[0]; b++)
            for (int c = 0; c < (nt > 1 ? ntt[1] : 1); c++) {
                PW p;
                p.I[0] = p.I[1] = p.I[2] = 0;
                p.I[d] = In[a];
                p.I[ta[0]] = Tt[0][b];
                if (nt > 1) p.I[ta[1]] = Tt[1][c];
                p.w = wn[a] * wt[0][b] * (nt > 1 ? wt[1][c] : 1.0);
                out[n++] = p;
            }
    return n;
}

/* x_f += M_f P e_c (velocity), x_c += M_c * 0.5 * e_C (pressure, unit-spacing scaling) */
void orc_prolong_add(int dim, int nx, int ny, int nz, const uint8_t *ffixed, int tang,
                     const double *ec, double *xf)
{
    G f = mkg(dim, nx, ny, nz), c = mkg(dim, nx / 2, ny / 2, (dim == 3) ? nz / 2 : 1);
    int kz = (dim == 3) ? 2 : 1;
    OMP_FOR
    for (int i = 0; i < f.nx; i++) for (int j = 0; j < f.ny; j++) for (int k = 0; k < f.nz; k++) {
        long q = ((long)i * f.ny + j) * f.nz + k;
        if (!ffixed[q]) xf[q] += 0.5 * cget(&c, ec, i / 2, j / 2, k / kz);
    }
    for (int d = 0; d < dim; d++) {
        int A, B, C; fdims(&f, d, &A, &B, &C);
        int ca, cb, cc; fdims(&c, d, &ca, &cb, &cc);
        OMP_FOR
        for (int i = 0; i < A; i++) for (int j = 0; j < B; j++) for (int k = 0; k < C; k++) {
            long q = f.nc + f.fo[d] + ((long)i * B + j) * C + k;
            if (ffixed[q]) continue;
            int fi[3] = {i, j, k};
            PW pw[8]; int np = pweights(dim, d, tang, fi, pw);
            double s = 0.0;
            for (int t = 0; t < np; t++)
                s += pw[t].w * at(ec + c.nc + c.fo[d], ca, cb, cc, pw[t].I[0], pw[t].I[1], pw[t].I[2]);
            xf[q] += s;
        }
    }
}

/* b_c = R r_f with R = (4/2^dim) P^T on velocity, (2/2^dim) P_p^T on pressure, masked */
void orc_restrict(int dim, int nx, int ny, int nz, const uint8_t *cfixed, int tang,
                  const double *rf, double *bc)
{
    G f = mkg(dim, nx, ny, nz), c = mkg(dim, nx / 2, ny / 2, (dim == 3) ? nz / 2 : 1);
    int kz = (dim == 3) ? 2 : 1;
    double cu = 4.0 / (double)(1 << dim), cp = 2.0 / (double)(1 << dim);
    memset(bc, 0, sizeof(double) * c.n);
    for (int i = 0; i < f.nx; i++) for (int j = 0; j < f.ny; j++) for (int k = 0; k < f.nz; k++)
        bc[((long)(i / 2) * c.ny + j / 2) * c.nz + k / kz] += cp * rf[((long)i * f.ny + j) * f.nz + k];
    for (int d = 0; d < dim; d++) {
        int A, B, C; fdims(&f, d, &A, &B, &C);
        int ca, cb, cc; fdims(&c, d, &ca, &cb, &cc);
        for (int i = 0; i < A; i++) for (int j = 0; j < B; j++) for (int k = 0; k < C; k++) {
            long q = f.nc + f.fo[d] + ((long)i * B + j) * C + k;
            int fi[3] = {i, j, k};
            PW pw[8]; int np = pweights(dim, d, tang, fi, pw);
            for (int t = 0; t < np; t++) {
                int *I = pw[t].I;
                if (I[0] < 0 || I[1] < 0 || I[2] < 0 || I[0] >= ca || I[1] >= cb || I[2] >= cc) continue;
                bc[c.nc + c.fo[d] + ((long)I[0] * cb + I[1]) * cc + I[2]] += cu * pw[t].w * rf[q];
            }
        }
    }
    for (long Q = 0; Q < c.n; Q++) if (cfixed[Q]) bc[Q] = 0.0;
}

/* one colour step of the box-Vanka smoother: for every cell of the colour, solve the
 * local 7x7 (5x5 in 2D) block [K, b_f; mc b_f^T, 1 - mc] exactly and add
 * omega*correction.  With E = blockdiag_d [[a_lo + (1-mc) m_lo, -m_lo m_hi],
 * [-m_lo m_hi, a_hi + (1-mc) m_hi]] and b_f = (-m_lo, m_hi) per axis, the velocity block
 * is K = E - (1 - mc) b_f b_f^T: in a free cell (mc = 1) the cross-axis couplings of
 * curl-curl and grad-div cancel and K = E; in a fixed cell the curl-curl couplings
 * remain (rank one), and the block is solved by Sherman-Morrison. */
void orc_smooth_color(int dim, int nx, int ny, int nz, const uint8_t *fixed, const double *pen,
                      const double *b, const double *xin, double *xout, int color, double omega)
{
    G g = mkg(dim, nx, ny, nz);
    double *vol = dalloc(g.n), *r = dalloc(g.n);
    for (long q = g.nc; q < g.n; q++) vol[q] = fixed[q] ? 0.0 : 1.0;
    orc_apply(dim, nx, ny, nz, fixed, vol, pen, xin, r);
    for (long q = 0; q < g.n; q++) r[q] = b[q] - r[q];
    memcpy(xout, xin, sizeof(double) * g.n);
    int ne = (dim == 3) ? 4 : 2;
    OMP_FOR
    for (int i = 0; i < g.nx; i++) for (int j = 0; j < g.ny; j++) for (int k = 0; k < g.nz; k++) {
        if (((i + j + k) & 1) != color) continue;
        long c = ((long)i * g.ny + j) * g.nz + k;
        double mc = fixed[c] ? 0.0 : 1.0, s0 = pen[c] + (1.0 - mc);
        int ci[3] = {i, j, k};
        long fq[3][2]; double K[3][3], bfv[3][2], rd[3][2];
        double beta = 0.0, tb = 0.0; /* b^T E^-1 b, b^T E^-1 r_u */
        for (int d = 0; d < dim; d++) {
            int A, B, C; fdims(&g, d, &A, &B, &C);
            double m[2], dg[2];
            for (int s = 0; s < 2; s++) {
                int fi[3] = {ci[0], ci[1], ci[2]}; fi[d] += s;
                long q = g.nc + g.fo[d] + ((long)fi[0] * B + fi[1]) * C + fi[2];
                fq[d][s] = q;
                m[s] = fixed[q] ? 0.0 : 1.0;
                /* cells adjacent to this face: fi - e_d and fi */
                double nadj = 0.0;
                for (int u = 0; u < 2; u++) {
                    int cc[3] = {fi[0], fi[1], fi[2]}; cc[d] -= (1 - u);
                    if (cc[0] < 0 || cc[1] < 0 || cc[2] < 0 || cc[0] >= g.nx || cc[1] >= g.ny || cc[2] >= g.nz) continue;
                    nadj += fixed[((long)cc[0] * g.ny + cc[1]) * g.nz + cc[2]] ? 0.0 : 1.0;
                }
                /* diagonal of A: free-cell count + edge count + penalty (identity on a
                 * fixed face), plus the (1-mc) m term that b b^T takes back in K */
                dg[s] = m[s] * (nadj + ne) + pen[q] + (1.0 - m[s]) + (1.0 - mc) * m[s];
                rd[d][s] = r[q];
            }
            double off = -m[0] * m[1];
            K[d][0] = dg[0]; K[d][1] = off; K[d][2] = dg[1];
            bfv[d][0] = -m[0]; bfv[d][1] = m[1];
            double det = dg[0] * dg[1] - off * off, id = 1.0 / det;
            /* E_d^-1 = id*[[dg1,-off],[-off,dg0]] */
            double kb0 = id * (dg[1] * bfv[d][0] - off * bfv[d][1]);
            double kb1 = id * (-off * bfv[d][0] + dg[0] * bfv[d][1]);
            double kr0 = id * (dg[1] * rd[d][0] - off * rd[d][1]);
            double kr1 = id * (-off * rd[d][0] + dg[0] * rd[d][1]);
            beta += bfv[d][0] * kb0 + bfv[d][1] * kb1;
            tb += bfv[d][0] * kr0 + bfv[d][1] * kr1;
        }
        /* free cell: Schur complement -beta (+ the cell penalty s0 = pen_c);
         * fixed cell: the identity row gives dp = r_p, and the velocity solve of
         * K = E - b b^T by Sherman-Morrison is E^-1 (r_u - b g) with g = (dp - tb)/(1 - beta) */
        double S = mc * (s0 - beta) + (1.0 - mc) * s0;
        double dp = (fabs(S) > 1e-300) ? (r[c] - mc * tb) / S : 0.0;
        double gm = (mc != 0.0) ? dp : (dp - tb) / (1.0 - beta);
        xout[c] += omega * dp;
        for (int d = 0; d < dim; d++) {
            double dg0 = K[d][0], off = K[d][1], dg1 = K[d][2];
            double id = 1.0 / (dg0 * dg1 - off * off);
            double r0 = rd[d][0] - bfv[d][0] * gm, r1 = rd[d][1] - bfv[d][1] * gm;
            xout[fq[d][0]] += omega * id * (dg1 * r0 - off * r1);
            xout[fq[d][1]] += omega * id * (-off * r0 + dg0 * r1);
        }
    }
    free(vol); free(r);
}

static void lu_factor(double *a, int *piv, long n)
{
    for (long k = 0; k < n; k++) {
        long p = k; double mx = fabs(a[k * n + k]);
        for (long i = k + 1; i < n; i++) if (fabs(a[i * n + k]) > mx) { mx = fabs(a[i * n + k]); p = i; }
        piv[k] = (int)p;
        if (p != k) for (long j = 0; j < n; j++) { double t = a[k * n + j]; a[k * n + j] = a[p * n + j]; a[p * n + j] = t; }
        double d = a[k * n + k];
        if (d == 0.0) continue;
        for (long i = k + 1; i < n; i++) {
            double l = a[i * n + k] / d;
            a[i * n + k] = l;
            for (long j = k + 1; j < n; j++) a[i * n + j] -= l * a[k * n + j];
        }
    }
}
static void lu_solve(const double *a, const int *piv, long n, double *x)
{
    for (long k = 0; k < n; k++) { long p = piv[k]; if (p != k) { double t = x[k]; x[k] = x[p]; x[p] = t; } }
    for (long i = 0; i < n; i++) for (long j = 0; j < i; j++) x[i] -= a[i * n + j] * x[j];
    for (long i = n - 1; i >= 0; i--) {
        for (long j = i + 1; j < n; j++) x[i] -= a[i * n + j] * x[j];
        x[i] = (a[i * n + i] != 0.0) ? x[i] / a[i * n + i] : 0.0;
    }
}

static void lvl_apply(const Lvl *l, const double *x, double *y)
{
    orc_apply(l->g.dim, l->g.nx, l->g.ny, l->g.nz, l->fixed, l->vol, l->pen, x, y);
}

orc_mg *orc_mg_create(int dim, int nx, int ny, int nz, const uint8_t *fixed, const double *pen,
                      const double *prm)
{
    orc_mg *mg = (orc_mg *)calloc(1, sizeof(orc_mg));
    mg->nu_pre = (int)prm[0]; mg->nu_post = (int)prm[1]; mg->omega = prm[2];
    mg->tang = (int)prm[3]; mg->closed = (int)prm[4];
    int coarsest = (int)prm[5];
    if (coarsest < 2) coarsest = 2;
    G g = mkg(dim, nx, ny, nz);
    int L = 0;
    for (;;) {
        Lvl *l = &mg->lv[L];
        l->g = g;
        l->fixed = (uint8_t *)calloc((size_t)g.n, 1);
        l->vol = dalloc(g.n); l->pen = dalloc(g.n);
        if (L == 0) { memcpy(l->fixed, fixed, (size_t)g.n); memcpy(l->pen, pen, sizeof(double) * g.n); }
        else {
            Lvl *f = &mg->lv[L - 1];
            orc_coarsen(dim, f->g.nx, f->g.ny, f->g.nz, f->fixed, f->pen, l->fixed, l->pen);
        }
        desc_vol(l);
        mg->b[L] = dalloc(g.n); mg->x[L] = dalloc(g.n); mg->r[L] = dalloc(g.n); mg->t[L] = dalloc(g.n);
        L++;
        int can = (g.nx % 2 == 0) && (g.ny % 2 == 0) && (dim == 2 || g.nz % 2 == 0)
               && g.nx / 2 >= coarsest && g.ny / 2 >= coarsest && (dim == 2 || g.nz / 2 >= coarsest)
               && L < 16;
        if (!can) break;
        g = mkg(dim, g.nx / 2, g.ny / 2, dim == 3 ? g.nz / 2 : 1);
    }
    mg->L = L;
    /* dense coarsest matrix by unit probes, + all-ones pressure gauge when closed
     * (ST/StokeLike.cu:605-608) */
    Lvl *c = &mg->lv[L - 1];
    long n = c->g.n;
    mg->nco = n;
    mg->lu = dalloc(n * n); mg->piv = (int *)calloc((size_t)n, sizeof(int));
    double *e = dalloc(n), *y = dalloc(n);
    for (long j = 0; j < n; j++) {
        memset(e, 0, sizeof(double) * n); e[j] = 1.0;
        lvl_apply(c, e, y);
        for (long i = 0; i < n; i++) mg->lu[i * n + j] = y[i];
    }
    if (mg->closed)
        for (long i = 0; i < c->g.nc; i++) for (long j = 0; j < c->g.nc; j++)
            if (!c->fixed[i] && !c->fixed[j]) mg->lu[i * n + j] += 1.0;
    lu_factor(mg->lu, mg->piv, n);
    free(e); free(y);
    return mg;
}

void orc_mg_free(orc_mg *mg)
{
    if (!mg) return;
    for (int l = 0; l < mg->L; l++) {
        free(mg->lv[l].fixed); free(mg->lv[l].vol); free(mg->lv[l].pen);
        free(mg->b[l]); free(mg->x[l]); free(mg->r[l]); free(mg->t[l]);
    }
    free(mg->lu); free(mg->piv); free(mg);
}
int orc_mg_levels(const orc_mg *mg) { return mg->L; }
void orc_mg_level_dims(const orc_mg *mg, int l, int *d)
{ d[0] = mg->lv[l].g.nx; d[1] = mg->lv[l].g.ny; d[2] = mg->lv[l].g.nz; }
void orc_mg_level_desc(const orc_mg *mg, int l, uint8_t *fixed, double *pen)
{
    memcpy(fixed, mg->lv[l].fixed, (size_t)mg->lv[l].g.n);
    memcpy(pen, mg->lv[l].pen, sizeof(double) * mg->lv[l].g.n);
}
void orc_mg_apply(orc_mg *mg, int l, const double *x, double *y) { lvl_apply(&mg->lv[l], x, y); }

static void smooth(orc_mg *mg, int l, const double *b, double *x, int nu, int first)
{
    Lvl *v = &mg->lv[l];
    double *t = mg->t[l];
    for (int s = 0; s < nu; s++)
        for (int h = 0; h < 2; h++) {
            int color = first ^ h;
            orc_smooth_color(v->g.dim, v->g.nx, v->g.ny, v->g.nz, v->fixed, v->pen, b, x, t, color, mg->omega);
            memcpy(x, t, sizeof(double) * v->g.n);
        }
}

static void vcycle(orc_mg *mg, int l)
{
    Lvl *v = &mg->lv[l];
    long n = v->g.n;
    double *b = mg->b[l], *x = mg->x[l], *r = mg->r[l];
    memset(x, 0, sizeof(double) * n);
    if (l == mg->L - 1) {
        memcpy(x, b, sizeof(double) * n);
        lu_solve(mg->lu, mg->piv, n, x);
        return;
    }
    smooth(mg, l, b, x, mg->nu_pre, 0);
    lvl_apply(v, x, r);
    for (long q = 0; q < n; q++) r[q] = b[q] - r[q];
    Lvl *c = &mg->lv[l + 1];
    orc_restrict(v->g.dim, v->g.nx, v->g.ny, v->g.nz, c->fixed, mg->tang, r, mg->b[l + 1]);
    vcycle(mg, l + 1);
    orc_prolong_add(v->g.dim, v->g.nx, v->g.ny, v->g.nz, v->fixed, mg->tang, mg->x[l + 1], x);
    smooth(mg, l, b, x, mg->nu_post, 1);
}

void orc_mg_vcycle(orc_mg *mg, const double *b, double *x)
{
    memcpy(mg->b[0], b, sizeof(double) * mg->lv[0].g.n);
    vcycle(mg, 0);
    memcpy(x, mg->x[0], sizeof(double) * mg->lv[0].g.n);
}

static double dot(const double *a, const double *b, long n)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long q = 0; q < n; q++) s += a[q] * b[q];
    return s;
}

/* FGMRES(m) right-preconditioned by one V-cycle, classical Gram-Schmidt,
 * stop on the true relative residual ||b - A x|| / ||b|| <= rtol */
int orc_fgmres(orc_mg *mg, const double *b, double *x, double rtol, int m, int maxit,
               double *relres, double *hist)
{
    Lvl *l0 = &mg->lv[0];
    long n = l0->g.n;
    double **V = (double **)calloc((size_t)m + 1, sizeof(double *));
    double **Z = (double **)calloc((size_t)m, sizeof(double *));
    for (int i = 0; i <= m; i++) V[i] = dalloc(n);
    for (int i = 0; i < m; i++) Z[i] = dalloc(n);
    double *H = dalloc((long)(m + 1) * m), *cs = dalloc(m), *sn = dalloc(m), *gv = dalloc(m + 1);
    double *w = dalloc(n), *yv = dalloc(m);
    double bn = sqrt(dot(b, b, n));
    if (bn == 0.0) bn = 1.0;
    int it = 0;
    double rr = 0.0;
    for (;;) {
        lvl_apply(l0, x, w);
        OMP_FOR for (long q = 0; q < n; q++) V[0][q] = b[q] - w[q];
        double beta = sqrt(dot(V[0], V[0], n));
        rr = beta / bn;
        if (hist && it == 0) hist[0] = rr;
        if (rr <= rtol || it >= maxit) break;
        OMP_FOR for (long q = 0; q < n; q++) V[0][q] /= beta;
        memset(gv, 0, sizeof(double) * (m + 1)); gv[0] = beta;
        int j;
        for (j = 0; j < m && it < maxit; j++) {
            orc_mg_vcycle(mg, V[j], Z[j]);
            lvl_apply(l0, Z[j], w);
            for (int i = 0; i <= j; i++) H[i * m + j] = dot(w, V[i], n);
            for (int i = 0; i <= j; i++) { double h = H[i * m + j]; OMP_FOR for (long q = 0; q < n; q++) w[q] -= h * V[i][q]; }
            double hn = sqrt(dot(w, w, n));
            H[(j + 1) * m + j] = hn;
            if (hn > 0) { OMP_FOR for (long q = 0; q < n; q++) V[j + 1][q] = w[q] / hn; }
            for (int i = 0; i < j; i++) {
                double a = H[i * m + j], bb = H[(i + 1) * m + j];
                H[i * m + j] = cs[i] * a + sn[i] * bb;
                H[(i + 1) * m + j] = -sn[i] * a + cs[i] * bb;
            }
            double a = H[j * m + j], bb = H[(j + 1) * m + j], den = sqrt(a * a + bb * bb);
            cs[j] = a / den; sn[j] = bb / den;
            H[j * m + j] = den; H[(j + 1) * m + j] = 0.0;
            gv[j + 1] = -sn[j] * gv[j]; gv[j] = cs[j] * gv[j];
            it++;
            if (hist) hist[it] = fabs(gv[j + 1]) / bn;
            if (fabs(gv[j + 1]) / bn <= rtol) { j++; break; }
        }
        for (int i = j - 1; i >= 0; i--) {
            double s = gv[i];
            for (int k = i + 1; k < j; k++) s -= H[i * m + k] * yv[k];
            yv[i] = s / H[i * m + i];
        }
        for (int i = 0; i < j; i++) { double y = yv[i]; OMP_FOR for (long q = 0; q < n; q++) x[q] += y * Z[i][q]; }
    }
    if (relres) *relres = rr;
    for (int i = 0; i <= m; i++) free(V[i]);
    for (int i = 0; i < m; i++) free(Z[i]);
    free(V); free(Z); free(H); free(cs); free(sn); free(gv); free(w); free(yv);
    return it;
}

/* ------------------------------------------------------- learned-part oracle */
/* 3x3x3 (or 1x1x1) convolution on an R^3 grid, channels-last, zero padding.
 * tap t = (a+1)*9 + (b+1)*3 + (c+1) for offsets (a,b,c) along (i,j,k).
 * out[v][co] = bias[co] + sum_t sum_ci in[v+o_t][ci] W[t][ci][co]  (FP64) */
void orc_conv3d(int R, int ntap, int Cin, int Cout, const double *in, const double *W, const double *bias,
                double *out, int relu)
{
    const long nv = (long)R * R * R;
    const int rr = (ntap == 27) ? 1 : 0;
    OMP_FOR
    for (long v = 0; v < nv; v++) {
        int i = (int)(v / ((long)R * R)), j = (int)((v / R) % R), k = (int)(v % R);
        for (int co = 0; co < Cout; co++) {
            double s = bias ? bias[co] : 0.0;
            for (int a = -rr; a <= rr; a++) for (int b = -rr; b <= rr; b++) for (int c = -rr; c <= rr; c++) {
                int ii = i + a, jj = j + b, kk = k + c;
                if (ii < 0 || jj < 0 || kk < 0 || ii >= R || jj >= R || kk >= R) continue;
                int t = (ntap == 27) ? (a + 1) * 9 + (b + 1) * 3 + (c + 1) : 0;
                const double *x = in + (((long)ii * R + jj) * R + kk) * Cin;
                const double *w = W + (long)t * Cin * Cout + co;
                for (int ci = 0; ci < Cin; ci++) s += x[ci] * w[(long)ci * Cout];
            }
            out[v * Cout + co] = (relu && s < 0) ? 0.0 : s;
        }
    }
}
/* gin[v][ci] = sum_t sum_co gout[v-o_t][co] W[t][ci][co] */
void orc_conv3d_bwd_data(int R, int ntap, int Cin, int Cout, const double *gout, const double *W, double *gin)
{
    const long nv = (long)R * R * R;
    const int rr = (ntap == 27) ? 1 : 0;
    OMP_FOR
    for (long v = 0; v < nv; v++) {
        int i = (int)(v / ((long)R * R)), j = (int)((v / R) % R), k = (int)(v % R);
        for (int ci = 0; ci < Cin; ci++) {
            double s = 0.0;
            for (int a = -rr; a <= rr; a++) for (int b = -rr; b <= rr; b++) for (int c = -rr; c <= rr; c++) {
                int ii = i - a, jj = j - b, kk = k - c;
                if (ii < 0 || jj < 0 || kk < 0 || ii >= R || jj >= R || kk >= R) continue;
                int t = (ntap == 27) ? (a + 1) * 9 + (b + 1) * 3 + (c + 1) : 0;
                const double *g = gout + (((long)ii * R + jj) * R + kk) * Cout;
                const double *w = W + ((long)t * Cin + ci) * Cout;
                for (int co = 0; co < Cout; co++) s += g[co] * w[co];
            }
            gin[v * Cin + ci] = s;
        }
    }
}
/* gW[t][ci][co] = sum_v in[v+o_t][ci] gout[v][co];  gb[co] = sum_v gout[v][co] */
void orc_conv3d_bwd_weight(int R, int ntap, int Cin, int Cout, const double *in, const double *gout, double *gW,
                           double *gb)
{
    const long nv = (long)R * R * R;
    const int rr = (ntap == 27) ? 1 : 0;
    OMP_FOR
    for (int t = 0; t < ntap; t++) {
        int a = (ntap == 27) ? t / 9 - 1 : 0, b = (ntap == 27) ? (t / 3) % 3 - 1 : 0, c = (ntap == 27) ? t % 3 - 1 : 0;
        double *w = gW + (long)t * Cin * Cout;
        for (long q = 0; q < (long)Cin * Cout; q++) w[q] = 0.0;
        for (long v = 0; v < nv; v++) {
            int i = (int)(v / ((long)R * R)), j = (int)((v / R) % R), k = (int)(v % R);
            int ii = i + a, jj = j + b, kk = k + c;
            if (ii < 0 || jj < 0 || kk < 0 || ii >= R || jj >= R || kk >= R) continue;
            const double *x = in + (((long)ii * R + jj) * R + kk) * Cin;
            const double *g = gout + v * Cout;
            for (int ci = 0; ci < Cin; ci++)
                for (int co = 0; co < Cout; co++) w[(long)ci * Cout + co] += x[ci] * g[co];
        }
    }
    (void)rr;
    if (gb) {
        for (int co = 0; co < Cout; co++) {
            double s = 0.0;
            for (long v = 0; v < nv; v++) s += gout[v * Cout + co];
            gb[co] = s;
        }
    }
}

/* ------------------------------------------------------------------ MMA (m = 1)
 * Svanberg's method of moving asymptotes for min f(x) s.t. g(x) <= 0 (one
 * constraint), x in [xmin, xmax], with the subproblem solved through its dual by
 * a primal-dual interior-point Newton iteration on (lambda, mu) -- the variant of
 * EXT/mma/MMASolver.cpp:74-469 (GenSub :360-443, SolveDIP :92-145, XYZofLAMBDA
 * :332-358, DualGrad :322-330, DualHess :224-320, DualLineSearch :196-222,
 * DualResidual :165-194) with its constants (a = 0, c = 1000, d = 0, raa0 = 1e-5,
 * move 0.5, albefa 0.1, asymptotes 0.5 / 0.7 / 1.2, epsimin = 1e-9 sqrt(n + 1)).
 * Sums run serially in index order, like the reference's, so the two agree to
 * the last bit where the operations match. */
struct orc_mma {
    long n;
    int iter;
    double *low, *upp, *xo1, *xo2, *alf, *bet, *p0, *q0, *p1, *q1;
    double lam, mu, y, z, b;
};

orc_mma *orc_mma_create(long n)
{
    orc_mma *s = (orc_mma *)calloc(1, sizeof(orc_mma));
    s->n = n;
    double **arr[10] = {&s->low, &s->upp, &s->xo1, &s->xo2, &s->alf, &s->bet, &s->p0, &s->q0, &s->p1, &s->q1};
    for (int a = 0; a < 10; a++) *arr[a] = (double *)calloc((size_t)n, sizeof(double));
    return s;
}

void orc_mma_free(orc_mma *s)
{
    if (!s) return;
    double *arr[10] = {s->low, s->upp, s->xo1, s->xo2, s->alf, s->bet, s->p0, s->q0, s->p1, s->q1};
    for (int a = 0; a < 10; a++) free(arr[a]);
    free(s);
}

static double mma_max(double a, double b) { return a > b ? a : b; }
static double mma_min(double a, double b) { return a < b ? a : b; }

/* the subproblem's separable coefficients for one constraint */
static void mma_subproblem(orc_mma *s, const double *x, const double *df, double g, const double *dg,
                           const double *xmin, const double *xmax)
{
    const double eps = 1.0e-5, raa0 = 1.0e-5, albefa = 0.1, movel = 0.5;
    s->iter++;
    for (long i = 0; i < s->n; i++) {
        const double w = xmax[i] - xmin[i];
        if (s->iter < 3) {
            s->low[i] = x[i] - 0.5 * w;
            s->upp[i] = x[i] + 0.5 * w;
        } else {
            const double osc = (x[i] - s->xo1[i]) * (s->xo1[i] - s->xo2[i]);
            const double gam = osc < 0.0 ? 0.7 : (osc > 0.0 ? 1.2 : 1.0);
            double L = x[i] - gam * (s->xo1[i] - s->low[i]);
            double U = x[i] + gam * (s->upp[i] - s->xo1[i]);
            const double wm = mma_max(eps, w);
            L = mma_min(mma_max(L, x[i] - 100.0 * wm), x[i] - 1.0e-5 * wm);
            U = mma_min(mma_max(U, x[i] + 1.0e-5 * wm), x[i] + 100.0 * wm);
            const double lo = xmin[i] - 1.0e-6, hi = xmax[i] + 1.0e-6;
            if (x[i] < lo) {
                L = x[i] - (hi - x[i]) / 0.9;
                U = x[i] + (hi - x[i]) / 0.9;
            }
            if (x[i] > hi) {
                L = x[i] - (x[i] - lo) / 0.9;
                U = x[i] + (x[i] - lo) / 0.9;
            }
            s->low[i] = L;
            s->upp[i] = U;
        }
        const double L = s->low[i], U = s->upp[i];
        s->alf[i] = mma_min(mma_max(mma_max(xmin[i], L + albefa * (x[i] - L)), x[i] - movel * w), xmax[i]);
        s->bet[i] = mma_max(mma_min(mma_min(xmax[i], U - albefa * (U - x[i])), x[i] + movel * w), xmin[i]);
        const double winv = 1.0 / mma_max(eps, w);
        const double du = pow(U - x[i], 2.0), dl = pow(x[i] - L, 2.0);
        double pq = 0.001 * fabs(df[i]) + raa0 * winv;
        s->p0[i] = du * (mma_max(0.0, df[i]) + pq);
        s->q0[i] = dl * (mma_max(0.0, -df[i]) + pq);
        pq = 0.001 * fabs(dg[i]) + raa0 * winv;
        s->p1[i] = du * (mma_max(0.0, dg[i]) + pq);
        s->q1[i] = dl * (mma_max(0.0, -dg[i]) + pq);
    }
    double b = -g;
    for (long i = 0; i < s->n; i++) b += s->p1[i] / (s->upp[i] - x[i]) + s->q1[i] / (x[i] - s->low[i]);
    s->b = b;
}

/* primal point of the dual at the current lambda (clamped to [alpha, beta]) */
static void mma_primal(orc_mma *s, double *x)
{
    const double cc = 1000.0;
    if (s->lam < 0.0) s->lam = 0.0;
    s->y = mma_max(0.0, s->lam - cc);
    s->z = mma_max(0.0, 10.0 * (s->lam * 0.0 - 1.0));
    for (long i = 0; i < s->n; i++) {
        const double P = s->p0[i] + s->p1[i] * s->lam, Q = s->q0[i] + s->q1[i] * s->lam;
        double v = (sqrt(P) * s->low[i] + sqrt(Q) * s->upp[i]) / (sqrt(P) + sqrt(Q));
        if (v < s->alf[i]) v = s->alf[i];
        if (v > s->bet[i]) v = s->bet[i];
        x[i] = v;
    }
}

/* dual gradient -b - a z - y (+ extra) + sum(p1/(U-x) + q1/(x-L)); the KKT residual
 * adds mu before the sum, in the reference's order */
static double mma_dual_grad(const orc_mma *s, const double *x, double extra)
{
    double gsum = -s->b - 0.0 * s->z - s->y + extra;
    for (long i = 0; i < s->n; i++) gsum += s->p1[i] / (s->upp[i] - x[i]) + s->q1[i] / (x[i] - s->low[i]);
    return gsum;
}

static double mma_dual_hess(orc_mma *s, const double *x)
{
    const double cc = 1000.0;
    double h = 0.0;
    for (long i = 0; i < s->n; i++) {
        const double P = s->p0[i] + s->p1[i] * s->lam, Q = s->q0[i] + s->q1[i] * s->lam;
        const double du = s->upp[i] - x[i], dl = x[i] - s->low[i];
        const double pq = s->p1[i] / pow(du, 2.0) - s->q1[i] / pow(dl, 2.0);
        double d2 = -1.0 / (2.0 * P / pow(du, 3.0) + 2.0 * Q / pow(dl, 3.0));
        const double xp = (sqrt(P) * s->low[i] + sqrt(Q) * s->upp[i]) / (sqrt(P) + sqrt(Q));
        if (xp < s->alf[i] || xp > s->bet[i]) d2 = 0.0;
        h += pq * d2 * pq;
    }
    if (s->lam < 0.0) s->lam = 0.0;
    if (s->lam > cc) h += -1.0;
    h += -s->mu / s->lam;
    double corr = 1e-4 * h;
    if (-corr < 1.0e-7) corr = -1.0e-7;
    return h + corr;
}

/* one MMA update of x (in place); returns the number of Newton steps */
int orc_mma_update(orc_mma *s, double *x, const double *df, double g, const double *dg, const double *xmin,
                   const double *xmax)
{
    mma_subproblem(s, x, df, g, dg, xmin, xmax);
    memcpy(s->xo2, s->xo1, sizeof(double) * (size_t)s->n);
    memcpy(s->xo1, x, sizeof(double) * (size_t)s->n);
    const double tol = sqrt((double)s->n + 1.0) * 1e-9;
    s->lam = 1000.0 / 2.0;
    s->mu = 1.0;
    double epsi = 1.0, err = 1.0;
    int steps = 0;
    while (epsi > tol) {
        int loop = 0;
        while (err > 0.9 * epsi && loop < 100) {
            loop++;
            steps++;
            mma_primal(s, x);
            const double gr = -mma_dual_grad(s, x, 0.0) - epsi / s->lam;
            const double h = mma_dual_hess(s, x);
            const double s0 = gr / h;
            const double s1 = -s->mu + epsi / s->lam - s0 * s->mu / s->lam;
            double th = 1.005;
            if (th < -1.01 * s0 / s->lam) th = -1.01 * s0 / s->lam;
            if (th < -1.01 * s1 / s->mu) th = -1.01 * s1 / s->mu;
            th = 1.0 / th;
            s->lam += th * s0;
            s->mu += th * s1;
            mma_primal(s, x);
            const double r0 = mma_dual_grad(s, x, s->mu), r1 = s->mu * s->lam - epsi;
            err = fabs(r0) > fabs(r1) ? fabs(r0) : fabs(r1);
        }
        epsi *= 0.1;
    }
    return steps;
}

void orc_mma_state(const orc_mma *s, double *lam_mu, double *low, double *upp)
{
    if (lam_mu) {
        lam_mu[0] = s->lam;
        lam_mu[1] = s->mu;
    }
    if (low) memcpy(low, s->low, sizeof(double) * (size_t)s->n);
    if (upp) memcpy(upp, s->upp, sizeof(double) * (size_t)s->n);
}
