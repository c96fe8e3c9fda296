/*
 * CPU FP64 oracle for the Stokes/Brinkman topology-optimization hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (stokes_to_amd/, include/)
 * links, loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the results tests compare
 * against / the CPU baseline.
 *
 * Layout (the ABI layout, identical to the reference's host Field/FaceField,
 * SX/reservoir/Field.h:67-72): every array is C-order over (i, j, k) with k
 * fastest.  A "dof vector" is [p (nx*ny*nz) | u_x ((nx+1)*ny*nz) |
 * u_y (nx*(ny+1)*nz) | u_z (nx*ny*(nz+1))].  2D uses nz = 1 and has no u_z.
 * A "descriptor" (fixed / vol / pen) uses the same [cells | faces] order, as the
 * reference's StokeFlowDescriptor3D does (ST/StokeFlowDescriptor3D.cu:6-50).
 *
 * Each function is pinned by identities, closed forms, dense solves or finite
 * differences in tests/test_oracle.py.
 */
#ifndef STOKES_ORACLE_H
#define STOKES_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* sizes[0]=cells, sizes[1..3]=faces per axis (sizes[3]=0 in 2D); returns total dofs */
long orc_dof(int dim, int nx, int ny, int nz, long *sizes4);

/* ---- DEC primitives (Appendix-A forms of CX/grid3D.cu, CX/form01mapping.cu) ---- */
void orc_d2(int dim, int nx, int ny, int nz, const double *faces, double *cells);
void orc_cod0(int dim, int nx, int ny, int nz, const double *cells, double *faces);
void orc_cod1(int dim, int nx, int ny, int nz, const double *faces, double *edges);
void orc_d1(int dim, int nx, int ny, int nz, const double *edges, double *faces);
long orc_edge_count(int dim, int nx, int ny, int nz);

/* ---- operators ---- */
/* 3D-form operator, dim 2 or 3 (ST/StokeFlowMapping3D.cu:176-306 composition) */
void orc_apply(int dim, int nx, int ny, int nz, const uint8_t *fixed, const double *vol,
               const double *pen, const double *x, double *y);
/* literal 2D reference operator (ST/StokeFlowMapping.cu:30-139): y-sign flips,
 * unmasked curl input, no definite fix */
void orc_apply_ref2d(int nx, int ny, const uint8_t *fixed, const double *vol,
                     const double *pen, const double *x, double *y);
/* objective operator G: dim 3 FT/FluidEnergyGrad3D.cu:27-134 (grad-div + curl-curl on the
 * masked velocity, plus the penalty); dim 2 the reference's 2D form, orc_energy_op_ref2d */
void orc_energy_op(int dim, int nx, int ny, int nz, const uint8_t *fixed, const double *vol,
                   const double *pen, const double *x, double *y);
/* literal 2D reference objective operator (FT/FluidEnergyGrad.cu:30-61): face 5-point
 * Laplacian 4u - sum of the 4 same-component neighbours (zero outside the face array,
 * CX/form01mapping.cu:429-525) of the unmasked velocity, fixed rows zeroed, lap_coeff 1,
 * plus pen * x on every dof */
void orc_energy_op_ref2d(int nx, int ny, const uint8_t *fixed, const double *pen,
                         const double *x, double *y);
/* right-hand side that lifts Dirichlet values given on fixed faces */
void orc_lift_rhs(int dim, int nx, int ny, int nz, const uint8_t *fixed, const double *xd,
                  double *b);

/* ---- topology-optimization pieces ---- */
/* am = {alpha_max, alpha_min, q, power, scale} */
double orc_alpha(double rho, const double *am);
double orc_dalpha(double rho, const double *am);
void orc_design_to_alpha(int dim, int nx, int ny, int nz, const double *rho, const double *am,
                         const uint8_t *face_fixed, double *face_alpha);
void orc_sensitivity(int dim, int nx, int ny, int nz, const double *rho, const double *am,
                     const uint8_t *face_fixed, const double *u, const double *lam, double *dj);
double orc_objective(int dim, int nx, int ny, int nz, const uint8_t *fixed, const double *vol,
                     const double *pen, const double *x);
void orc_filter(int dim, int nx, int ny, int nz, double radius, const double *in, double *out,
                int transpose);
int orc_oc_update(int dim, int nx, int ny, int nz, double radius, const double *rho,
                  const double *dc, double volfrac, double move, double eta, double rho_min,
                  double *rho_new, double *change);

/* ---- multigrid / Krylov (the same algorithm as the HIP path) ---- */
typedef struct orc_mg orc_mg;
/* prm = {nu_pre, nu_post, omega, tang_interp(0/1), closed(0/1), coarsest(min cells), rtol_unused} */
orc_mg *orc_mg_create(int dim, int nx, int ny, int nz, const uint8_t *fixed, const double *pen,
                      const double *prm);
void orc_mg_free(orc_mg *mg);
int orc_mg_levels(const orc_mg *mg);
void orc_mg_level_dims(const orc_mg *mg, int l, int *dims3);
void orc_mg_level_desc(const orc_mg *mg, int l, uint8_t *fixed, double *pen);
void orc_mg_vcycle(orc_mg *mg, const double *b, double *x);
void orc_mg_apply(orc_mg *mg, int l, const double *x, double *y);
void orc_smooth_color(int dim, int nx, int ny, int nz, const uint8_t *fixed, const double *pen,
                      const double *b, const double *xin, double *xout, int color, double omega);
void orc_restrict(int dim, int nx, int ny, int nz, const uint8_t *cfixed, int tang,
                  const double *rf, double *bc);
void orc_prolong_add(int dim, int nx, int ny, int nz, const uint8_t *ffixed, int tang,
                     const double *ec, double *xf);
void orc_coarsen(int dim, int nx, int ny, int nz, const uint8_t *ffixed, const double *fpen,
                 uint8_t *cfixed, double *cpen);
/* learned part: FP64 3D convolution (channels-last, zero padding) and its adjoints */
void orc_conv3d(int R, int ntap, int Cin, int Cout, const double *in, const double *W, const double *bias,
                double *out, int relu);
void orc_conv3d_bwd_data(int R, int ntap, int Cin, int Cout, const double *gout, const double *W, double *gin);
void orc_conv3d_bwd_weight(int R, int ntap, int Cin, int Cout, const double *in, const double *gout, double *gW,
                           double *gb);
/* MMA with one constraint (EXT/mma/MMASolver.cpp restated for m = 1): state holds
 * the asymptotes and the two previous iterates */
typedef struct orc_mma orc_mma;
orc_mma *orc_mma_create(long n);
void orc_mma_free(orc_mma *s);
int orc_mma_update(orc_mma *s, double *x, const double *df, double g, const double *dg, const double *xmin,
                   const double *xmax);
void orc_mma_state(const orc_mma *s, double *lam_mu, double *low, double *upp);
/* FGMRES(m), right-preconditioned by one V-cycle; returns iterations, relres out */
int orc_fgmres(orc_mg *mg, const double *b, double *x, double rtol, int restart, int maxit,
               double *relres, double *history);

#ifdef __cplusplus
}
#endif
#endif
