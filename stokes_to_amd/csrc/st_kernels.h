// Launchers for the HIP kernels of st_kernels.hip (host-callable).
#pragma once
#include "st_common.h"

namespace st {

// alpha(r) = amax + (amin - amax) r(1+q)/(r+q) on r = rho^power, times scale
// (FT/FluidTopoOpt3D.cpp:110-123)
struct AlphaParams {
    double amax, amin, q, power, scale;
};

// the V-cycle below and including level l0 in one launch (levels small enough for
// one workgroup): pre-smoothing, residual and restriction down to the coarsest
// level, its dense solve, then prolongation and post-smoothing back to l0
constexpr int ST_MAX_LEVELS = 12;
constexpr long ST_TAIL_MAX_CELLS = 512;  // 8^3: one cell per thread of the workgroup
// (kept in device memory: the kernel indexes it by level; b and x of level 0 are
// the caller's vectors when l0 = 0, the level buffers b[l], x[l] otherwise)
constexpr int ST_TAIL_LDS_BYTES = 150 * 1024;  // LDS pool of the LDS-resident variant
struct TailLds {  // byte offsets of the level buffers in the LDS pool
    int fl[ST_MAX_LEVELS], al[ST_MAX_LEVELS], b[ST_MAX_LEVELS], x[ST_MAX_LEVELS], t[ST_MAX_LEVELS],
        r[ST_MAX_LEVELS];
    int bs, bytes, ok;
};
template <typename TS> struct TailArgs {
    int l0, L, nu_pre, nu_post, nd, dim;
    float omega;
    TailLds lds;
    Lvl g[ST_MAX_LEVELS];
    const uint8_t *fl[ST_MAX_LEVELS];
    const TS *al[ST_MAX_LEVELS];
    TS *b[ST_MAX_LEVELS], *x[ST_MAX_LEVELS], *t[ST_MAX_LEVELS], *r[ST_MAX_LEVELS];
    const long *dofs;
    const double *minv;
};
template <typename TS>
void launch_vcycle_tail(const TailArgs<TS> *dev_args, int dim, const TS *b0, TS *x0, int nd, bool lds, hipStream_t);

template <typename TS> void launch_apply(const Lvl &, const uint8_t *, const TS *, const TS *, TS *, hipStream_t);
template <typename TS>
void launch_residual(const Lvl &, const uint8_t *, const TS *, const TS *b, const TS *x, TS *r, double *part,
                     hipStream_t);
unsigned residual_blocks(const Lvl &);
template <typename TS>
void launch_smooth(const Lvl &, const uint8_t *, const TS *, const TS *b, const TS *xin, TS *xout, int color,
                   float omega, bool zero, hipStream_t);
template <typename TS>
void launch_restrict(const Lvl &gf, const Lvl &gc, const uint8_t *flc, const TS *rf, TS *bc, hipStream_t);
template <typename TS>
void launch_prolong(const Lvl &gf, const Lvl &gc, const uint8_t *flf, const TS *ec, TS *xf, hipStream_t);
template <typename TF, typename TC>
void launch_coarsen(const Lvl &gf, const Lvl &gc, const uint8_t *flf, const TF *alf, uint8_t *flc, TC *alc,
                    hipStream_t);
template <typename TS>
void launch_probe(const Lvl &, const uint8_t *, const TS *, const long *dofs, int nd, int closed, double *A,
                  hipStream_t);
void launch_gauss_jordan(int n, double *A, int *perm, int *info, hipStream_t);
template <typename TS>
void launch_coarse_solve(int nd, const long *dofs, const double *Minv, const TS *b, TS *x, hipStream_t);
template <typename TS>
void launch_pack(const Lvl &, const double *abi, TS *pad, int fixed_only, const uint8_t *fl, double scale,
                 hipStream_t);
template <typename TS> void launch_unpack(const Lvl &, const TS *pad, double *abi, hipStream_t);
void launch_flags(const Lvl &, const uint8_t *abi_fixed, uint8_t *fl, hipStream_t);
template <typename TS> void launch_alpha_from_pen(const Lvl &, const double *pen, const uint8_t *fl, TS *al, hipStream_t);
void launch_design_alpha(const Lvl &, const double *rho, const AlphaParams &, const uint8_t *fl, double *al,
                         hipStream_t);
void launch_sensitivity(const Lvl &, const double *rho, const AlphaParams &, const uint8_t *fl, const double *u,
                        const double *lam, double *dj, hipStream_t);
void launch_lift(const Lvl &, const uint8_t *fl, const double *xd, const double *body, double *b, hipStream_t);
void launch_energy(const Lvl &, const uint8_t *fl, const double *al, const double *x, double *y, double *part,
                   hipStream_t);
unsigned energy_blocks(const Lvl &);
// Wi: reciprocal weight sums of the cells, written by launch_filter_w
void launch_filter(const Lvl &, double R, const double *Wi, const double *in, double *out, int transpose, hipStream_t);
void launch_filter_w(const Lvl &, double R, double *Wi, hipStream_t);
constexpr int OC_K = 16;  // multipliers per pass of the OC search
struct OcFactors {
    double f[OC_K];
};
unsigned oc_grid(long n);
void launch_oc_base(long n, const double *dc, const double *dv, double eta, double *base, hipStream_t);
void launch_oc_vols(long n, const double *rho, const double *base, const double *dv, const OcFactors &f, double move,
                    double rho_min, double *part, hipStream_t);
void launch_oc_apply(long n, const double *rho, const double *base, const double *dv, double fl, double move,
                     double rho_min, double *rn, double *part, hipStream_t);
unsigned absmax_blocks(long n);
void launch_absdiff_max(long n, const double *a, const double *b, double *part, hipStream_t);
void launch_sum(int np, int nvec, const double *part, double *out, hipStream_t);
unsigned blas_blocks(long n);
template <typename TA, typename TB> void launch_dot(long n, const TA *a, const TB *b, double *part, hipStream_t);
template <typename TW, typename TV>
void launch_multidot(long n, const TW *w, const TV *V, long ld, int m, double *part, hipStream_t);
template <typename TW, typename TV>
void launch_multiaxpy(long n, TW *w, const TV *V, long ld, int m, const double *h, const double *sc, double *part,
                      hipStream_t);
template <typename TX, typename TY> void launch_axpby(long n, double a, const TX *x, double b, TY *y, hipStream_t);
void launch_lincomb3(long n, double a, const double *x, double b, const double *w, double c, double *y, hipStream_t);
// minimal-residual start: d = float(a - b); Gram of nd <= 3 FP32 vectors y_i (stride ld)
// and their products with r (partials: nd(nd+1)/2 + nd per block, mr_blocks(n) blocks);
// x0 = x1 + sum d_i D_i, r -= sum d_i y_i with the partials of the new ||r||^2
struct MrCoef {
    double d[3];
};
void launch_diff_f32(long n, const double *a, const double *b, float *d, hipStream_t);
unsigned mr_blocks(long n);
void launch_mr_gram(int nd, long n, const float *y, long ld, const float *r, double *part, hipStream_t);
void launch_mr_update(int nd, long n, const double *x1, const float *const *D, const float *y, long ld,
                      const MrCoef &cf, double *x0, float *r, double *part, hipStream_t);
template <typename TS> void launch_fill(long n, TS *y, double v, hipStream_t);

// 2.5D marching kernels (st_march.hip) for 3D levels with ny % 4 == 0, nz % 64 == 0
bool march_ok(const Lvl &);
unsigned march_blocks(const Lvl &);
template <typename TS> void march_apply(const Lvl &, const uint8_t *, const TS *, const TS *, TS *, hipStream_t);
template <typename TS>
void march_residual(const Lvl &, const uint8_t *, const TS *, const TS *b, const TS *x, TS *r, double *part,
                    hipStream_t);
template <typename TS>
void march_smooth(const Lvl &, const uint8_t *, const TS *, const TS *b, const TS *xin, TS *xout, int color,
                  float omega, hipStream_t);
// FP64 residual r = b - A x stored in FP32 (+ FP64 partial sums of |r|^2)
void march_residual_f32(const Lvl &, const uint8_t *, const double *al, const double *b, const double *x, float *r,
                        double *part, hipStream_t);
// FP64 operator applied to an FP32-stored vector (x float, alpha and y double)
void march_apply_mixed(const Lvl &, const uint8_t *, const double *al, const float *x, double *y, hipStream_t);
// one red-black sweep (colour col0, then the other) in one pass (st_sweep.hip);
// FP32 arithmetic, TS = float or bf16 storage, levels with march_ok
template <typename TS>
void march_sweep(const Lvl &, const uint8_t *, const TS *al, const TS *b, const TS *xin, TS *xout, int col0,
                 float omega, hipStream_t);

}  // namespace st
