// Method of moving asymptotes with one constraint on the device.
//
// The reference updates the design with Svanberg's MMA (EXT/mma/MMASolver.cpp,
// driven by SX/optimization/OptimizerMma.h:138-170): each update builds the
// separable convex subproblem (GenSub, :360-443) and solves its one-dimensional
// dual by a primal-dual interior-point Newton iteration on (lambda, mu)
// (SolveDIP :92-145).  Every n-sized step is a data-parallel pass here:
//   k_mma_sub   asymptotes, move bounds alpha/beta and the p/q coefficients of
//               objective and constraint, plus partial sums of the constraint
//               constant b;
//   k_mma_eval  the primal point x(lambda) (clamped to [alpha, beta]) and partial
//               sums of the dual gradient sum(p1/(U-x) + q1/(x-L)) and of the dual
//               Hessian sum(PQ^2 d2) in one read of the eight coefficient arrays.
// The scalar Newton/line-search logic runs on the host on the reduced sums (one
// fixed-order FP64 reduction per pass, so results are deterministic).
#include <cmath>
#include <vector>

#include "st_mma.h"

namespace st {

namespace {
constexpr int MMA_NB = 2048;   // fixed grid: partial sums have a fixed order
constexpr double MMA_C = 1000.0, MMA_EPS = 1.0e-5, MMA_RAA0 = 1.0e-5, MMA_ALBEFA = 0.1, MMA_MOVE = 0.5;

__device__ __forceinline__ double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

__device__ __forceinline__ void block_store2(double a, double b, double *part, int nb)
{
    __shared__ double sh[2][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    a = wsum(a);
    b = wsum(b);
    if (lane == 0) {
        sh[0][w] = a;
        sh[1][w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
        part[nb + blockIdx.x] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
    }
}
}  // namespace

__global__ void __launch_bounds__(256) k_mma_sub(long n, int iter, const double *__restrict__ x,
                                                 const double *__restrict__ df, const double *__restrict__ dg,
                                                 const double *__restrict__ xmin, const double *__restrict__ xmax,
                                                 const double *__restrict__ xo1, const double *__restrict__ xo2,
                                                 double *low, double *upp, double *alf, double *bet, double *p0,
                                                 double *q0, double *p1, double *q1, double *part)
{
    double acc = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double xv = x[i], w = xmax[i] - xmin[i];
        double L, U;
        if (iter < 3) {
            L = xv - 0.5 * w;
            U = xv + 0.5 * w;
        } else {
            const double o1 = xo1[i];
            const double osc = (xv - o1) * (o1 - xo2[i]);
            const double gam = osc < 0.0 ? 0.7 : (osc > 0.0 ? 1.2 : 1.0);
            L = xv - gam * (o1 - low[i]);
            U = xv + gam * (upp[i] - o1);
            const double wm = fmax(MMA_EPS, w);
            L = fmin(fmax(L, xv - 100.0 * wm), xv - 1.0e-5 * wm);
            U = fmin(fmax(U, xv + 1.0e-5 * wm), xv + 100.0 * wm);
            const double lo = xmin[i] - 1.0e-6, hi = xmax[i] + 1.0e-6;
            if (xv < lo) {
                L = xv - (hi - xv) / 0.9;
                U = xv + (hi - xv) / 0.9;
            }
            if (xv > hi) {
                L = xv - (xv - lo) / 0.9;
                U = xv + (xv - lo) / 0.9;
            }
        }
        low[i] = L;
        upp[i] = U;
        alf[i] = fmin(fmax(fmax(xmin[i], L + MMA_ALBEFA * (xv - L)), xv - MMA_MOVE * w), xmax[i]);
        bet[i] = fmax(fmin(fmin(xmax[i], U - MMA_ALBEFA * (U - xv)), xv + MMA_MOVE * w), xmin[i]);
        const double winv = 1.0 / fmax(MMA_EPS, w);
        const double du = (U - xv) * (U - xv), dl = (xv - L) * (xv - L);
        const double f = df[i], gq = dg[i];
        double pq = 0.001 * fabs(f) + MMA_RAA0 * winv;
        p0[i] = du * (fmax(0.0, f) + pq);
        q0[i] = dl * (fmax(0.0, -f) + pq);
        pq = 0.001 * fabs(gq) + MMA_RAA0 * winv;
        const double a = du * (fmax(0.0, gq) + pq), b = dl * (fmax(0.0, -gq) + pq);
        p1[i] = a;
        q1[i] = b;
        acc += a / (U - xv) + b / (xv - L);
    }
    block_store2(acc, 0.0, part, gridDim.x);
}

__global__ void __launch_bounds__(256) k_mma_eval(long n, double lam, const double *__restrict__ low,
                                                  const double *__restrict__ upp, const double *__restrict__ alf,
                                                  const double *__restrict__ bet, const double *__restrict__ p0,
                                                  const double *__restrict__ q0, const double *__restrict__ p1,
                                                  const double *__restrict__ q1, double *x, double *part)
{
    double gs = 0.0, hs = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double L = low[i], U = upp[i], a1 = p1[i], b1 = q1[i];
        const double P = p0[i] + a1 * lam, Q = q0[i] + b1 * lam;
        const double sp = sqrt(P), sq = sqrt(Q);
        const double xp = (sp * L + sq * U) / (sp + sq);
        const double lo = alf[i], hi = bet[i];
        const double xv = xp < lo ? lo : (xp > hi ? hi : xp);
        x[i] = xv;
        const double du = U - xv, dl = xv - L;
        gs += a1 / du + b1 / dl;
        const double pq = a1 / (du * du) - b1 / (dl * dl);
        double d2 = -1.0 / (2.0 * P / (du * du * du) + 2.0 * Q / (dl * dl * dl));
        if (xp < lo || xp > hi) d2 = 0.0;
        hs += pq * d2 * pq;
    }
    block_store2(gs, hs, part, gridDim.x);
}

// outer move limits of OptimizerMma::Optimize_MMA (SX/optimization/OptimizerMma.h:157-164)
__global__ void __launch_bounds__(256) k_mma_box(long n, const double *x, double move, double lb, double ub,
                                                 double *xmin, double *xmax)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    xmin[i] = fmax(lb, x[i] - move);
    xmax[i] = fmin(ub, x[i] + move);
}

void launch_mma_box(long n, const double *x, double move, double lb, double ub, double *xmin, double *xmax,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_mma_box, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, move, lb, ub, xmin, xmax);
}

__global__ void __launch_bounds__(256) k_mma_sum2(int nb, const double *part, double *out)
{
    __shared__ double sh[2][4];
    double a = 0.0, b = 0.0;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) {
        a += part[q];
        b += part[nb + q];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    a = wsum(a);
    b = wsum(b);
    if (lane == 0) {
        sh[0][w] = a;
        sh[1][w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
        out[1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
    }
}

struct MmaDev {
    long n = 0;
    int iter = 0;
    hipStream_t s = nullptr;
    double *low = nullptr, *upp = nullptr, *xo1 = nullptr, *xo2 = nullptr, *alf = nullptr, *bet = nullptr;
    double *p0 = nullptr, *q0 = nullptr, *p1 = nullptr, *q1 = nullptr, *part = nullptr, *red = nullptr;
    std::vector<void *> al;
};

MmaDev *mma_create(long n, hipStream_t s)
{
    MmaDev *m = new MmaDev;
    m->n = n;
    m->s = s;
    double **arr[10] = {&m->low, &m->upp, &m->xo1, &m->xo2, &m->alf, &m->bet, &m->p0, &m->q0, &m->p1, &m->q1};
    for (int a = 0; a < 10; a++) {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(double) * (size_t)n) != hipSuccess) {
            mma_destroy(m);
            return nullptr;
        }
        m->al.push_back(p);
        *arr[a] = (double *)p;
    }
    void *p = nullptr, *q = nullptr;
    if (hipMalloc(&p, sizeof(double) * 2 * MMA_NB) != hipSuccess || hipMalloc(&q, sizeof(double) * 2) != hipSuccess) {
        if (p) hipFree(p);
        mma_destroy(m);
        return nullptr;
    }
    m->al.push_back(p);
    m->al.push_back(q);
    m->part = (double *)p;
    m->red = (double *)q;
    return m;
}

void mma_destroy(MmaDev *m)
{
    if (!m) return;
    for (void *p : m->al) hipFree(p);
    delete m;
}

void mma_reset(MmaDev *m)
{
    if (m) m->iter = 0;
}

// blocks of the two data-parallel kernels: fixed for a given n, so the partial
// sums always have the same order
static int mma_nb(long n)
{
    long b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : (b > MMA_NB ? MMA_NB : b));
}

static int mma_sums(MmaDev *m, double *g, double *h)
{
    hipLaunchKernelGGL(k_mma_sum2, dim3(1), dim3(256), 0, m->s, mma_nb(m->n), m->part, m->red);
    double r[2];
    if (hipMemcpyAsync(r, m->red, sizeof(r), hipMemcpyDeviceToHost, m->s) != hipSuccess) return -1;
    if (hipStreamSynchronize(m->s) != hipSuccess) return -1;
    *g = r[0];
    if (h) *h = r[1];
    return 0;
}

static int mma_eval(MmaDev *m, double lam, double *x, double *g, double *h)
{
    hipLaunchKernelGGL(k_mma_eval, dim3(mma_nb(m->n)), dim3(256), 0, m->s, m->n, lam, m->low, m->upp, m->alf, m->bet,
                       m->p0, m->q0, m->p1, m->q1, x, m->part);
    return mma_sums(m, g, h);
}

int mma_update(MmaDev *m, double *x, const double *df, double g, const double *dg, const double *xmin,
               const double *xmax, int *steps_out)
{
    const long n = m->n;
    m->iter++;
    hipLaunchKernelGGL(k_mma_sub, dim3(mma_nb(n)), dim3(256), 0, m->s, n, m->iter, (const double *)x, df, dg, xmin,
                       xmax, (const double *)m->xo1, (const double *)m->xo2, m->low, m->upp, m->alf, m->bet, m->p0,
                       m->q0, m->p1, m->q1, m->part);
    double bsum = 0.0;
    if (mma_sums(m, &bsum, nullptr) != 0) return -1;
    const double b = -g + bsum;
    hipMemcpyAsync(m->xo2, m->xo1, sizeof(double) * n, hipMemcpyDeviceToDevice, m->s);
    hipMemcpyAsync(m->xo1, x, sizeof(double) * n, hipMemcpyDeviceToDevice, m->s);
    const double tol = std::sqrt((double)n + 1.0) * 1e-9;
    double lam = MMA_C / 2.0, mu = 1.0, epsi = 1.0, err = 1.0;
    double G = 0.0, H = 0.0;
    int steps = 0;
    if (mma_eval(m, lam, x, &G, &H) != 0) return -1;
    while (epsi > tol) {
        int loop = 0;
        while (err > 0.9 * epsi && loop < 100) {
            loop++;
            steps++;
            // Newton step on the perturbed KKT conditions of the dual
            double y = lam - MMA_C > 0.0 ? lam - MMA_C : 0.0;
            const double grad = -((-b - y) + G) - epsi / lam;
            double hess = H;
            if (lam > MMA_C) hess += -1.0;
            hess += -mu / lam;
            double corr = 1e-4 * hess;
            if (-corr < 1.0e-7) corr = -1.0e-7;
            hess += corr;
            const double s0 = grad / hess;
            const double s1 = -mu + epsi / lam - s0 * mu / lam;
            double th = 1.005;
            if (th < -1.01 * s0 / lam) th = -1.01 * s0 / lam;
            if (th < -1.01 * s1 / mu) th = -1.01 * s1 / mu;
            th = 1.0 / th;
            lam += th * s0;
            mu += th * s1;
            if (lam < 0.0) lam = 0.0;
            if (mma_eval(m, lam, x, &G, &H) != 0) return -1;
            y = lam - MMA_C > 0.0 ? lam - MMA_C : 0.0;
            const double r0 = (-b - y + mu) + G, r1 = mu * lam - epsi;
            err = std::fabs(r0) > std::fabs(r1) ? std::fabs(r0) : std::fabs(r1);
            if (!std::isfinite(err)) return -2;
        }
        epsi *= 0.1;
    }
    if (steps_out) *steps_out = steps;
    return hipStreamSynchronize(m->s) == hipSuccess ? 0 : -1;
}

}  // namespace st
