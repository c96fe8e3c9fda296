// HIP kernels for the MAC-grid Stokes/Brinkman hot path (gfx950).
//
// Every kernel is a gather-form stencil over the padded node layout of
// st_common.h: one thread writes only its own outputs, so results are
// deterministic (the reference's scatter kernels use double atomicAdd,
// CX/grid3D.cu:431-455, and are not).  Reductions are two-stage, fixed order.
#include "st_kernels.h"

namespace st {

// ----------------------------------------------------------------- stencil

// Evaluates rows of the operator of ST/StokeFlowMapping3D.cu:176-306 at a
// single DOF, with vol = free mask and pen_c = 0:
//   y_f = m_f [D2^T m_c D2 (m u) + D1 D1^T (m u) + D2^T p] + alpha_f u_f + (1-m_f) u_f
//   y_c = m_c D2 (m u) + (1-m_c) p_c
// D2 = divergence (CX/grid3D.cu:330-390), D1^T = Cod1 face->edge curl
// (CX/grid3D.cu:469-568), D1 edge->face curl (CX/grid3D.cu:219-318).
// MODE 0: load x;  MODE 1: x is the unit vector e_hot (dense probing);
// MODE 2: load x with column masks off (Dirichlet lifting).
template <typename TS, int DIM, int MODE = 0>
struct Sten {
    typedef typename Cvt<TS>::C C;
    Lvl g;
    const uint8_t *__restrict__ fl;
    const TS *__restrict__ x;
    const TS *__restrict__ al;
    long hot;

    __device__ __forceinline__ long id(int i, int j, int k) const { return ((long)i * g.Py + j) * g.Pz + k; }
    __device__ __forceinline__ bool in(int i, int j, int k) const
    {
        if (DIM == 2) return (unsigned)i <= (unsigned)g.nx && (unsigned)j <= (unsigned)g.ny && k == 0;
        return (unsigned)i <= (unsigned)g.nx && (unsigned)j <= (unsigned)g.ny && (unsigned)k <= (unsigned)g.nz;
    }
    // Flag and masked-velocity reads are branch-free: a node outside the box reads
    // node 0 and the value is discarded, so the loads of a stencil do not wait for the
    // flag and range tests and can be in flight at once (the gather kernels of the
    // small levels are latency-bound).
    __device__ __forceinline__ long idc(int i, int j, int k) const { return in(i, j, k) ? id(i, j, k) : 0; }
    __device__ __forceinline__ int F(int i, int j, int k) const
    {
        const int f = (int)fl[idc(i, j, k)];
        return in(i, j, k) ? f : 0;
    }
    __device__ __forceinline__ C ldx(long idx) const
    {
        if (MODE == 1) return (idx == hot) ? C(1) : C(0);
        return Cvt<TS>::ld(x[idx]);
    }
    __device__ __forceinline__ C X(int f, int i, int j, int k) const
    {
        return in(i, j, k) ? ldx((long)f * g.S + id(i, j, k)) : C(0);
    }
    // X with the load issued whatever the range test gives (the restriction's 12-point
    // gathers; the prolongation's two-point ones are faster with the branch)
    __device__ __forceinline__ C Xb(int f, int i, int j, int k) const
    {
        const C v = ldx((long)f * g.S + idc(i, j, k));
        return in(i, j, k) ? v : C(0);
    }
    // masked velocity w = m u
    __device__ __forceinline__ C W(int d, int i, int j, int k) const
    {
        if (DIM == 2 && d == 2) return C(0);
        const long q = idc(i, j, k);
        const C v = ldx((long)(1 + d) * g.S + q);
        const int f = fl[q];  // loaded whatever the tests give: no branch around the loads
        const bool on = in(i, j, k) && (MODE == 2 || ((f >> (1 + d)) & 1));
        return on ? v : C(0);
    }
    __device__ __forceinline__ C dv(int i, int j, int k) const
    {
        C s = W(0, i + 1, j, k) - W(0, i, j, k) + W(1, i, j + 1, k) - W(1, i, j, k);
        if (DIM == 3) s += W(2, i, j, k + 1) - W(2, i, j, k);
        return s;
    }
    __device__ __forceinline__ C gdiv(int i, int j, int k) const
    {
        const C d = dv(i, j, k);
        return (F(i, j, k) & 1) ? d : C(0);
    }
    __device__ __forceinline__ C ex(int i, int j, int k) const
    { return W(1, i, j, k) - W(1, i, j, k - 1) - W(2, i, j, k) + W(2, i, j - 1, k); }
    __device__ __forceinline__ C ey(int i, int j, int k) const
    { return W(2, i, j, k) - W(2, i - 1, j, k) - W(0, i, j, k) + W(0, i, j, k - 1); }
    __device__ __forceinline__ C ez(int i, int j, int k) const
    { return W(0, i, j, k) - W(0, i, j - 1, k) - W(1, i, j, k) + W(1, i - 1, j, k); }

    // (D2^T m_c D2 + D1 D1^T)(m u) at free face d of node (i,j,k)
    template <int d>
    __device__ __forceinline__ C visc(int i, int j, int k) const
    {
        const int di = (d == 0), dj = (d == 1), dk = (d == 2);
        C r = gdiv(i - di, j - dj, k - dk) - gdiv(i, j, k);
        if (DIM == 3) {
            if (d == 0) r += ey(i, j, k + 1) - ey(i, j, k) + ez(i, j, k) - ez(i, j + 1, k);
            if (d == 1) r += ex(i, j, k) - ex(i, j, k + 1) + ez(i + 1, j, k) - ez(i, j, k);
            if (d == 2) r += ex(i, j + 1, k) - ex(i, j, k) + ey(i, j, k) - ey(i + 1, j, k);
        } else {
            if (d == 0) r += ez(i, j, 0) - ez(i, j + 1, 0);
            if (d == 1) r += ez(i + 1, j, 0) - ez(i, j, 0);
        }
        return r;
    }
    // Aface and Acell evaluate both rows (free and fixed) and select by the flag, so
    // that no load waits for the flag
    template <int d>
    __device__ __forceinline__ C Aface(int i, int j, int k) const
    {
        long q = id(i, j, k);
        const C u = ldx((long)(1 + d) * g.S + q);
        const int f = fl[q];
        const int di = (d == 0), dj = (d == 1), dk = (d == 2);
        C r = visc<d>(i, j, k);
        r += Xb(0, i - di, j - dj, k - dk) - ldx(q);
        r += Cvt<TS>::ld(al[(long)d * g.S + q]) * u;
        return ((f >> (1 + d)) & 1) ? r : u;  // definite fix / padding: u
    }
    __device__ __forceinline__ C Acell(int i, int j, int k) const
    {
        long q = id(i, j, k);
        const C dvc = dv(i, j, k);
        const C p = ldx(q);
        return (fl[q] & 1) ? dvc : p;
    }
    // objective operator G of FT/FluidEnergyGrad3D.cu:27-134 (no pressure terms, no fix)
    template <int d>
    __device__ __forceinline__ C Gface(int i, int j, int k) const
    {
        long q = id(i, j, k);
        if (!((fl[q] >> (1 + d)) & 1)) return C(0);
        C u = ldx((long)(1 + d) * g.S + q);
        return visc<d>(i, j, k) + Cvt<TS>::ld(al[(long)d * g.S + q]) * u;
    }
    // 2D objective operator of FT/FluidEnergyGrad.cu:30-61: the face 5-point Laplacian
    // 4u - (4 neighbours of the same component) of the *unmasked* velocity, zero outside
    // the face array (grid2DOperator::faceLapMapping, CX/form01mapping.cu:429-525),
    // rows of fixed faces zeroed, times lap_coeff = 1, plus the penalty alpha_f u_f
    template <int d>
    __device__ __forceinline__ C Gface2(int i, int j) const
    {
        long q = id(i, j, 0);
        if (!((fl[q] >> (1 + d)) & 1)) return C(0);
        const int A = g.nx + (d == 0), B = g.ny + (d == 1);  // face array dims
        auto u = [&](int a, int b) -> C {
            return ((unsigned)a < (unsigned)A && (unsigned)b < (unsigned)B) ? ldx((long)(1 + d) * g.S + id(a, b, 0))
                                                                            : C(0);
        };
        const C uc = u(i, j);
        const C lap = C(4) * uc - u(i - 1, j) - u(i + 1, j) - u(i, j - 1) - u(i, j + 1);
        return lap + Cvt<TS>::ld(al[(long)d * g.S + q]) * uc;
    }
};

__device__ __forceinline__ void node_ijk(const Lvl &g, long n, int &i, int &j, int &k)
{
    k = (int)(n % g.Pz);
    long t = n / g.Pz;
    j = (int)(t % g.Py);
    i = (int)(t / g.Py);
}

// ------------------------------------------------------- block reductions

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
// block of up to 1024 threads; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        int nw = (blockDim.x + 63) >> 6;
        for (int q = 0; q < nw; q++) s += sh[q];
    }
    __syncthreads();
    return s;
}

// ------------------------------------------------------------- operator

template <typename TS, int DIM>
__global__ void __launch_bounds__(256) k_apply(Lvl g, const uint8_t *fl, const TS *al, const TS *x, TS *y)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.S) return;
    int i, j, k;
    node_ijk(g, n, i, j, k);
    Sten<TS, DIM> s{g, fl, x, al, -1};
    y[n] = Cvt<TS>::st(s.Acell(i, j, k));
    y[g.S + n] = Cvt<TS>::st(s.template Aface<0>(i, j, k));
    y[2 * g.S + n] = Cvt<TS>::st(s.template Aface<1>(i, j, k));
    if (DIM == 3) y[3 * g.S + n] = Cvt<TS>::st(s.template Aface<2>(i, j, k));
}

// r = b - A x at node n; returns the node's contribution to |r|^2
template <typename TS, int DIM>
__device__ __forceinline__ double residual_body(const Lvl &g, const uint8_t *fl, const TS *al, const TS *b,
                                                const TS *x, TS *r, long n)
{
    if (n >= g.S) return 0.0;
    int i, j, k;
    node_ijk(g, n, i, j, k);
    Sten<TS, DIM> s{g, fl, x, al, -1};
    typedef typename Cvt<TS>::C C;
    C v0 = Cvt<TS>::ld(b[n]) - s.Acell(i, j, k);
    C v1 = Cvt<TS>::ld(b[g.S + n]) - s.template Aface<0>(i, j, k);
    C v2 = Cvt<TS>::ld(b[2 * g.S + n]) - s.template Aface<1>(i, j, k);
    if (r) {  // without r: the norm only
        r[n] = Cvt<TS>::st(v0);
        r[g.S + n] = Cvt<TS>::st(v1);
        r[2 * g.S + n] = Cvt<TS>::st(v2);
    }
    double acc = (double)v0 * v0 + (double)v1 * v1 + (double)v2 * v2;
    if (DIM == 3) {
        C v3 = Cvt<TS>::ld(b[3 * g.S + n]) - s.template Aface<2>(i, j, k);
        if (r) r[3 * g.S + n] = Cvt<TS>::st(v3);
        acc += (double)v3 * v3;
    }
    return acc;
}

// r = b - A x, optional partial sums of r^2 (FP64) per block
template <typename TS, int DIM>
__global__ void __launch_bounds__(256) k_residual(Lvl g, const uint8_t *fl, const TS *al, const TS *b,
                                                  const TS *x, TS *r, double *part)
{
    __shared__ double sh[16];
    const double acc = residual_body<TS, DIM>(g, fl, al, b, x, r, (long)blockIdx.x * blockDim.x + threadIdx.x);
    if (part) {
        double t = block_sum(acc, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
}

// --------------------------------------------------------------- smoother

// One colour of the red-black box-Vanka smoother (the reference smooths the same
// cell blocks [6 faces + cell] with batched LU, ST/StokeLike.cu:41-372,
// CX/DescentSmoother.cu:31-53).  For a cell c of the colour the local block of A is
//   [K  b_f ; m_c b_f^T  1 - m_c],  K = E - (1 - m_c) b_f b_f^T,
//   E = blockdiag_d [[a_lo + (1-m_c) m_lo, -m_lo m_hi], [-m_lo m_hi, a_hi + (1-m_c) m_hi]],
//   b_f = (-m_lo, m_hi) per axis.
// In a free cell the cross-axis couplings of curl-curl and grad-div cancel (vol = free
// mask) and K = E; the block is solved through the scalar Schur complement.  In a
// fixed cell the curl-curl couplings remain as the rank-one term, p is decoupled
// (dp = r_p) and the velocity solve is Sherman-Morrison on E.  Both are the exact
// block inverse (equal to LU of the probed block).  xin -> xout (ping-pong); DOFs not
// owned by a cell of the colour are copied.  ZERO: xin is identically zero.
template <typename TS, int DIM, bool ZERO>
__device__ __forceinline__ void smooth_cell(const Lvl &g, const uint8_t *fl, const TS *al, const TS *b,
                                            const TS *xin, TS *xout, int color, float omega, int i, int j, int k)
{
    typedef typename Cvt<TS>::C C;
    Sten<TS, DIM> s{g, fl, xin, al, -1};
    const long q = s.id(i, j, k);
    const long S = g.S;
    const long st[3] = {(long)g.Py * g.Pz, (long)g.Pz, 1};
    if (((i + j + k) & 1) != color) {
        xout[q] = ZERO ? Cvt<TS>::st(C(0)) : xin[q];
        if (i == 0) xout[S + q] = ZERO ? Cvt<TS>::st(C(0)) : xin[S + q];
        if (i == g.nx - 1) xout[S + q + st[0]] = ZERO ? Cvt<TS>::st(C(0)) : xin[S + q + st[0]];
        if (j == 0) xout[2 * S + q] = ZERO ? Cvt<TS>::st(C(0)) : xin[2 * S + q];
        if (j == g.ny - 1) xout[2 * S + q + st[1]] = ZERO ? Cvt<TS>::st(C(0)) : xin[2 * S + q + st[1]];
        if (DIM == 3) {
            if (k == 0) xout[3 * S + q] = ZERO ? Cvt<TS>::st(C(0)) : xin[3 * S + q];
            if (k == g.nz - 1) xout[3 * S + q + st[2]] = ZERO ? Cvt<TS>::st(C(0)) : xin[3 * S + q + st[2]];
        }
        return;
    }
    const int fb = fl[q];
    const C mc = (fb & 1) ? C(1) : C(0);
    const C NE = (DIM == 3) ? C(4) : C(2);
    C rp = Cvt<TS>::ld(b[q]);
    if (!ZERO) rp -= s.Acell(i, j, k);
    C beta = C(0), tb = C(0);  // b^T E^-1 b, b^T E^-1 r_u
    C a0[3], a1[3], of[3], r0[3], r1[3], m0[3], m1[3];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const int di = (d == 0), dj = (d == 1), dk = (d == 2);
        const long qh = q + st[d];
        const int fl_lo = fb, fl_hi = fl[qh];
        m0[d] = ((fl_lo >> (1 + d)) & 1) ? C(1) : C(0);
        m1[d] = ((fl_hi >> (1 + d)) & 1) ? C(1) : C(0);
        // free cells adjacent to the two faces: (c - e_d, c) and (c, c + e_d)
        C c_lo = (C)(s.F(i - di, j - dj, k - dk) & 1);
        C c_hi = (C)(s.F(i + di, j + dj, k + dk) & 1);
        C al0 = Cvt<TS>::ld(al[(long)d * S + q]);
        C al1 = Cvt<TS>::ld(al[(long)d * S + qh]);
        a0[d] = m0[d] * (c_lo + mc + NE) + al0 + (C(1) - m0[d]) + (C(1) - mc) * m0[d];
        a1[d] = m1[d] * (mc + c_hi + NE) + al1 + (C(1) - m1[d]) + (C(1) - mc) * m1[d];
        of[d] = -m0[d] * m1[d];
        C rl = Cvt<TS>::ld(b[(long)(1 + d) * S + q]);
        C rh = Cvt<TS>::ld(b[(long)(1 + d) * S + qh]);
        if (!ZERO) {
            if (d == 0) { rl -= s.template Aface<0>(i, j, k); rh -= s.template Aface<0>(i + 1, j, k); }
            if (d == 1) { rl -= s.template Aface<1>(i, j, k); rh -= s.template Aface<1>(i, j + 1, k); }
            if (d == 2) { rl -= s.template Aface<2>(i, j, k); rh -= s.template Aface<2>(i, j, k + 1); }
        }
        r0[d] = rl;
        r1[d] = rh;
        const C id = C(1) / (a0[d] * a1[d] - of[d] * of[d]);
        // b_f = (-m0, m1), b_c = mc * b_f
        const C kb0 = id * (a1[d] * (-m0[d]) - of[d] * m1[d]);
        const C kb1 = id * (-of[d] * (-m0[d]) + a0[d] * m1[d]);
        const C kr0 = id * (a1[d] * rl - of[d] * rh);
        const C kr1 = id * (-of[d] * rl + a0[d] * rh);
        beta += (-m0[d]) * kb0 + m1[d] * kb1;
        tb += (-m0[d]) * kr0 + m1[d] * kr1;
    }
    const C Sc = (mc != C(0)) ? -beta : C(1);
    const C dp = (Sc != C(0)) ? (rp - mc * tb) / Sc : C(0);
    const C gm = (mc != C(0)) ? dp : (dp - tb) / (C(1) - beta);
    const C om = (C)omega;
    const C p0 = ZERO ? C(0) : Cvt<TS>::ld(xin[q]);
    xout[q] = Cvt<TS>::st(p0 + om * dp);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const long qh = q + st[d];
        const C id = C(1) / (a0[d] * a1[d] - of[d] * of[d]);
        const C e0 = r0[d] + m0[d] * gm;   // r - b_f g, b_f = (-m0, m1)
        const C e1 = r1[d] - m1[d] * gm;
        const C u0 = ZERO ? C(0) : Cvt<TS>::ld(xin[(long)(1 + d) * S + q]);
        const C u1 = ZERO ? C(0) : Cvt<TS>::ld(xin[(long)(1 + d) * S + qh]);
        xout[(long)(1 + d) * S + q] = Cvt<TS>::st(u0 + om * id * (a1[d] * e0 - of[d] * e1));
        xout[(long)(1 + d) * S + qh] = Cvt<TS>::st(u1 + om * id * (-of[d] * e0 + a0[d] * e1));
    }
}

// one thread per cell t (k fastest)
template <typename TS, int DIM, bool ZERO>
__device__ __forceinline__ void smooth_body(const Lvl &g, const uint8_t *fl, const TS *al, const TS *b,
                                            const TS *xin, TS *xout, int color, float omega, long t)
{
    const long ncell = (long)g.nx * g.ny * g.nz;
    if (t >= ncell) return;
    const int k = (int)(t % g.nz);
    const long tt = t / g.nz;
    smooth_cell<TS, DIM, ZERO>(g, fl, al, b, xin, xout, color, omega, (int)(tt / g.ny), (int)(tt % g.ny), k);
}

// The pass from x = 0 in 3D with one thread per pair of cells (k, k+1), k even: the
// pair holds one cell of each colour, so every lane runs the block solve of its
// colour cell (one thread per cell leaves half the lanes idle through it)
template <typename TS>
__global__ void __launch_bounds__(256) k_smooth0_pairs(Lvl g, const uint8_t *fl, const TS *al, const TS *b,
                                                       TS *xout, int color, float omega)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int hz = g.nz / 2;
    if (t >= (long)g.nx * g.ny * hz) return;
    const int kp = (int)(t % hz);
    const long tt = t / hz;
    const int j = (int)(tt % g.ny), i = (int)(tt / g.ny);
    const int k0 = 2 * kp;
    const int km = (((i + j + k0) & 1) == color) ? k0 : k0 + 1;
    smooth_cell<TS, 3, true>(g, fl, al, b, (const TS *)nullptr, xout, color, omega, i, j, km);
    smooth_cell<TS, 3, true>(g, fl, al, b, (const TS *)nullptr, xout, color, omega, i, j, 2 * k0 + 1 - km);
}

template <typename TS, int DIM, bool ZERO>
__global__ void __launch_bounds__(256) k_smooth(Lvl g, const uint8_t *fl, const TS *al, const TS *b,
                                                const TS *xin, TS *xout, int color, float omega)
{
    smooth_body<TS, DIM, ZERO>(g, fl, al, b, xin, xout, color, omega, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ------------------------------------------------------------- transfers

// coarse rhs from the fine residual: R = (4/2^dim) P^T on velocity, (2/2^dim)
// P_p^T on pressure (unit-spacing scaling of the residual per level), masked.
// P: velocity linear along the face normal, constant across it; pressure
// constant.  The reference restricts by plain 8-child sums
// (ST/StokeFlowSubSystem3D.cu:32-52, CX/SubsystemUtils3D.cuh:7-185).
template <typename TS, int DIM>
__device__ __forceinline__ void restrict_body(const Lvl &gf, const Lvl &gc, const uint8_t *flc, const TS *rf, TS *bc,
                                              long n)
{
    typedef typename Cvt<TS>::C C;
    if (n >= gc.S) return;
    int I, J, K;
    node_ijk(gc, n, I, J, K);
    const int fb = flc[n];
    Sten<TS, DIM> f{gf, nullptr, rf, nullptr, -1};
    const C cu = (DIM == 3) ? C(0.5) : C(1.0);
    const C cp = (DIM == 3) ? C(0.25) : C(0.5);
    const int kz = (DIM == 3) ? 2 : 1;
    C sp = C(0);
    if (fb & 1) {
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int c = 0; c < kz; c++) sp += f.Xb(0, 2 * I + a, 2 * J + b, kz * K + c);
        sp *= cp;
    }
    bc[n] = Cvt<TS>::st(sp);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        C su = C(0);
        if ((fb >> (1 + d)) & 1) {
            for (int on = -1; on <= 1; on++) {
                const C w = (on == 0) ? C(1) : C(0.5);
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < ((DIM == 3) ? 2 : 1); b++) {
                        int fi, fj, fk;
                        if (d == 0) { fi = 2 * I + on; fj = 2 * J + a; fk = (DIM == 3) ? 2 * K + b : 0; }
                        else if (d == 1) { fi = 2 * I + a; fj = 2 * J + on; fk = (DIM == 3) ? 2 * K + b : 0; }
                        else { fi = 2 * I + a; fj = 2 * J + b; fk = 2 * K + on; }
                        su += w * f.Xb(1 + d, fi, fj, fk);
                    }
            }
            su *= cu;
        }
        bc[(long)(1 + d) * gc.S + n] = Cvt<TS>::st(su);
    }
}

template <typename TS, int DIM>
__global__ void __launch_bounds__(256) k_restrict(Lvl gf, Lvl gc, const uint8_t *flc, const TS *rf, TS *bc)
{
    restrict_body<TS, DIM>(gf, gc, flc, rf, bc, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// xf += P ec on free fine DOFs (pressure correction scaled by 1/2: unit spacing)
template <typename TS, int DIM>
__device__ __forceinline__ void prolong_body(const Lvl &gf, const Lvl &gc, const uint8_t *flf, const TS *ec, TS *xf,
                                             long n)
{
    typedef typename Cvt<TS>::C C;
    if (n >= gf.S) return;
    int i, j, k;
    node_ijk(gf, n, i, j, k);
    const int fb = flf[n];
    if (!fb) return;
    Sten<TS, DIM> c{gc, nullptr, ec, nullptr, -1};
    const int kz = (DIM == 3) ? 2 : 1;
    if (fb & 1) xf[n] = Cvt<TS>::st(Cvt<TS>::ld(xf[n]) + C(0.5) * c.X(0, i >> 1, j >> 1, k / kz));
    int co[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        if (!((fb >> (1 + d)) & 1)) continue;
        int I = i >> 1, J = j >> 1, K = k / kz;
        C e;
        if ((co[d] & 1) == 0) {
            e = c.X(1 + d, I, J, K);
        } else {
            e = c.X(1 + d, I, J, K) + c.X(1 + d, I + (d == 0), J + (d == 1), K + (d == 2));
            e *= C(0.5);
        }
        long o = (long)(1 + d) * gf.S + n;
        xf[o] = Cvt<TS>::st(Cvt<TS>::ld(xf[o]) + e);
    }
}

template <typename TS, int DIM>
__global__ void __launch_bounds__(256) k_prolong(Lvl gf, Lvl gc, const uint8_t *flf, const TS *ec, TS *xf)
{
    prolong_body<TS, DIM>(gf, gc, flf, ec, xf, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// coarse flags (free if any coincident child is free) and coarse face alpha
// (4 x the midpoint of the normal-linear weighted arithmetic and harmonic means of
// the child faces; 4 = (h_c/h_f)^2 keeps the penalty in unit-spacing form).  Reference coarsening:
// ST/StokeFlowSubSystem3D.cu:103-154.
template <typename TF, typename TC, int DIM>
__global__ void __launch_bounds__(256) k_coarsen(Lvl gf, Lvl gc, const uint8_t *flf, const TF *alf, uint8_t *flc,
                                                 TC *alc)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= gc.S) return;
    int I, J, K;
    node_ijk(gc, n, I, J, K);
    Sten<TF, DIM> f{gf, flf, nullptr, alf, -1};
    const int kz = (DIM == 3) ? 2 : 1;
    int out = 0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < kz; c++) out |= f.F(2 * I + a, 2 * J + b, kz * K + c) & 1;
    int fdim[3] = {gf.nx, gf.ny, gf.nz};
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        int fr = 0, zero = 0;
        double sw = 0.0, sa = 0.0, sh = 0.0;
        for (int on = -1; on <= 1; on++) {
            const double w = (on == 0) ? 1.0 : 0.5;
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < ((DIM == 3) ? 2 : 1); b++) {
                    int fi[3];
                    if (d == 0) { fi[0] = 2 * I + on; fi[1] = 2 * J + a; fi[2] = (DIM == 3) ? 2 * K + b : 0; }
                    else if (d == 1) { fi[0] = 2 * I + a; fi[1] = 2 * J + on; fi[2] = (DIM == 3) ? 2 * K + b : 0; }
                    else { fi[0] = 2 * I + a; fi[1] = 2 * J + b; fi[2] = 2 * K + on; }
                    bool ok = true;
                    for (int e = 0; e < DIM; e++) {
                        int hi = fdim[e] + (e == d ? 1 : 0);
                        if (fi[e] < 0 || fi[e] >= hi) ok = false;
                    }
                    if (!ok) continue;
                    long q = f.id(fi[0], fi[1], fi[2]);
                    if (on == 0) fr |= (flf[q] >> (1 + d)) & 1;
                    const double av = (double)Cvt<TF>::ld(alf[(long)d * gf.S + q]);
                    sw += w;
                    sa += w * av;
                    if (av > 0.0) sh += w / av;
                    else zero = 1;
                }
        }
        out |= fr << (1 + d);
        // midpoint of the weighted arithmetic and harmonic child means (the two Wiener
        // bounds of an effective coefficient), times (h_c/h_f)^2 = 4
        const double ar = sw > 0 ? sa / sw : 0.0, hm = (zero || sh <= 0.0) ? 0.0 : sw / sh;
        alc[(long)d * gc.S + n] = Cvt<TC>::st((typename Cvt<TC>::C)((fr && sw > 0) ? 2.0 * (ar + hm) : 0.0));
    }
    flc[n] = (uint8_t)out;
}

// ------------------------------------------------------ coarsest solve

// column `col` of the dense coarsest matrix over the compact DOF list
template <typename TS, int DIM>
__global__ void k_probe(Lvl g, const uint8_t *fl, const TS *al, const long *dofs, int nd, int closed, double *A)
{
    int col = blockIdx.x;
    Sten<TS, DIM, 1> s{g, fl, nullptr, al, dofs[col]};
    for (int row = threadIdx.x; row < nd; row += blockDim.x) {
        long p = dofs[row];
        int f = (int)(p / g.S);
        long n = p - (long)f * g.S;
        int i, j, k;
        node_ijk(g, n, i, j, k);
        double v;
        if (f == 0) v = (double)s.Acell(i, j, k);
        else if (f == 1) v = (double)s.template Aface<0>(i, j, k);
        else if (f == 2) v = (double)s.template Aface<1>(i, j, k);
        else v = (double)s.template Aface<2>(i, j, k);
        if (closed) {
            long pc = dofs[col];
            if (f == 0 && pc < g.S && (fl[n] & 1) && (fl[pc] & 1)) v += 1.0;  // pressure gauge (ST/StokeLike.cu:605-608)
        }
        A[(long)row * nd + col] = v;
    }
}

// in-place Gauss-Jordan inverse with partial pivoting, one workgroup.  info: 0 ok,
// k > 0 zero pivot at step k (singular), -1 a non-finite entry (nothing is inverted)
__global__ void __launch_bounds__(1024) k_gauss_jordan(int n, double *A, int *perm, int *info)
{
    __shared__ double sv[1024];
    __shared__ int si[1024];
    __shared__ int piv_s, bad;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) {
        *info = 0;
        bad = 0;
    }
    __syncthreads();
    for (long q = tid; q < (long)n * n; q += nt)
        if (!isfinite(A[q])) bad = 1;
    __syncthreads();
    if (bad) {  // the same value for every thread: the whole workgroup returns here
        if (tid == 0) *info = -1;
        return;
    }
    for (int k = 0; k < n; k++) {
        double best = -1.0;
        int bi = k;
        for (int r = k + tid; r < n; r += nt) {
            double a = fabs(A[(long)r * n + k]);
            if (a > best) { best = a; bi = r; }
        }
        sv[tid] = best;
        si[tid] = bi;
        __syncthreads();
        for (int o = nt >> 1; o > 0; o >>= 1) {
            if (tid < o && (sv[tid + o] > sv[tid] || (sv[tid + o] == sv[tid] && si[tid + o] < si[tid]))) {
                sv[tid] = sv[tid + o];
                si[tid] = si[tid + o];
            }
            __syncthreads();
        }
        if (tid == 0) {
            piv_s = si[0];
            perm[k] = si[0];
            if (sv[0] == 0.0) *info = k + 1;
        }
        __syncthreads();
        const int p = piv_s;
        if (p != k)
            for (int c = tid; c < n; c += nt) {
                double t = A[(long)k * n + c];
                A[(long)k * n + c] = A[(long)p * n + c];
                A[(long)p * n + c] = t;
            }
        __syncthreads();
        const double pv = A[(long)k * n + k];
        const double pinv = (pv != 0.0) ? 1.0 / pv : 0.0;
        __syncthreads();
        for (int c = tid; c < n; c += nt) {
            double v = (c == k) ? 1.0 : A[(long)k * n + c];
            A[(long)k * n + c] = v * pinv;
        }
        __syncthreads();
        for (int r = tid >> 6; r < n; r += (nt >> 6)) {
            if (r == k) continue;
            const double f = A[(long)r * n + k];
            for (int c = tid & 63; c < n; c += 64) {
                if (c == k) continue;
                A[(long)r * n + c] -= f * A[(long)k * n + c];
            }
        }
        __syncthreads();
        for (int r = tid; r < n; r += nt) {
            if (r == k) continue;
            double f = A[(long)r * n + k];
            A[(long)r * n + k] = -f * pinv;
        }
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; k--) {
        int p = perm[k];
        if (p != k)
            for (int r = tid; r < n; r += nt) {
                double t = A[(long)r * n + k];
                A[(long)r * n + k] = A[(long)r * n + p];
                A[(long)r * n + p] = t;
            }
        __syncthreads();
    }
}

// x = Minv b over the compact list (FP64 arithmetic whatever the storage type)
template <typename TS>
__global__ void k_coarse_solve(int nd, const long *dofs, const double *Minv, const TS *b, TS *x)
{
    extern __shared__ double bs[];
    for (int c = threadIdx.x; c < nd; c += blockDim.x) bs[c] = (double)Cvt<TS>::ld(b[dofs[c]]);
    __syncthreads();
    for (int r = threadIdx.x; r < nd; r += blockDim.x) {
        double s = 0.0;
        const double *row = Minv + (long)r * nd;
        for (int c = 0; c < nd; c++) s += row[c] * bs[c];
        x[dofs[r]] = Cvt<TS>::st((typename Cvt<TS>::C)s);
    }
}

// ------------------------------------------------ fused coarse V-cycle

// One workgroup runs every pass of the coarse levels: each pass is a
// workgroup-strided loop over the level's cells or nodes, passes are separated by
// barriers (the level data, a few hundred KB, stays in L2).  The same bodies as the
// per-level kernels, so the result is identical; it replaces about eleven
// latency-bound launches per level.  b[l0] is the caller's right-hand side; bw[l]
// (l > l0) receive the restricted residuals.
template <typename TS, int DIM>
__global__ void __launch_bounds__(512) k_vcycle_tail(const TailArgs<TS> *__restrict__ a, const TS *b0, TS *x0)
{
    extern __shared__ double bs[];
    const long T = blockDim.x;
    const int l0 = a->l0, L = a->L;
    const float om = a->omega;
    auto bof = [&](int l) -> const TS * { return l == 0 ? b0 : a->b[l]; };
    auto xof = [&](int l) -> TS * { return l == 0 ? x0 : a->x[l]; };
    auto smooth = [&](int l, int nu, int first, bool zero) {
        const Lvl g = a->g[l];
        const uint8_t *fl = a->fl[l];
        const TS *al = a->al[l], *b = bof(l);
        TS *x = xof(l), *t = a->t[l];
        const long nc = (long)g.nx * g.ny * g.nz;
        for (int sw = 0; sw < nu; sw++) {
            if (zero && sw == 0)
                for (long q = threadIdx.x; q < nc; q += T) smooth_body<TS, DIM, true>(g, fl, al, b, x, t, first, om, q);
            else
                for (long q = threadIdx.x; q < nc; q += T) smooth_body<TS, DIM, false>(g, fl, al, b, x, t, first, om, q);
            __syncthreads();
            for (long q = threadIdx.x; q < nc; q += T) smooth_body<TS, DIM, false>(g, fl, al, b, t, x, first ^ 1, om, q);
            __syncthreads();
        }
    };
    for (int l = l0; l + 1 < L; l++) {
        const Lvl g = a->g[l];
        if (a->nu_pre > 0) {
            smooth(l, a->nu_pre, 0, true);
        } else {
            TS *x = xof(l);
            for (long q = threadIdx.x; q < (long)g.nfield * g.S; q += T) x[q] = Cvt<TS>::st((typename Cvt<TS>::C)0);
            __syncthreads();
        }
        for (long n = threadIdx.x; n < g.S; n += T) residual_body<TS, DIM>(g, a->fl[l], a->al[l], bof(l), xof(l), a->r[l], n);
        __syncthreads();
        const Lvl gc = a->g[l + 1];
        for (long n = threadIdx.x; n < gc.S; n += T) restrict_body<TS, DIM>(g, gc, a->fl[l + 1], a->r[l], a->b[l + 1], n);
        __syncthreads();
    }
    const int Lc = L - 1, nd = a->nd;
    const TS *bc = bof(Lc);
    TS *xc = xof(Lc);
    const long *dofs = a->dofs;
    for (int c = threadIdx.x; c < nd; c += T) bs[c] = (double)Cvt<TS>::ld(bc[dofs[c]]);
    __syncthreads();
    for (int r = threadIdx.x; r < nd; r += T) {
        double sacc = 0.0;
        const double *row = a->minv + (long)r * nd;
        for (int c = 0; c < nd; c++) sacc += row[c] * bs[c];
        xc[dofs[r]] = Cvt<TS>::st((typename Cvt<TS>::C)sacc);
    }
    __syncthreads();
    for (int l = L - 2; l >= l0; l--) {
        const Lvl g = a->g[l];
        for (long n = threadIdx.x; n < g.S; n += T) prolong_body<TS, DIM>(g, a->g[l + 1], a->fl[l], a->x[l + 1], xof(l), n);
        __syncthreads();
        if (a->nu_post > 0) smooth(l, a->nu_post, 1, false);
    }
}

// The same cycle with every level buffer staged in LDS (the levels of the tail
// fit: ~136 KB in FP64 from 8^3 down): flags, alpha and the right-hand side are
// loaded once, all passes read and write LDS, x of level l0 is written back.
template <typename TS, int DIM>
__global__ void __launch_bounds__(512) k_vcycle_tail_lds(const TailArgs<TS> *__restrict__ a, const TS *b0g, TS *x0g)
{
    __shared__ __align__(16) unsigned char pool[ST_TAIL_LDS_BYTES];
    const long T = blockDim.x;
    const int l0 = a->l0, L = a->L;
    const float om = a->omega;
    // level buffers in the pool, formed where they are used (so the compiler sees LDS
    // pointers and emits LDS instructions, not flat ones)
    auto Pfl = [&](int l) { return (uint8_t *)(pool + a->lds.fl[l]); };
    auto Pal = [&](int l) { return (TS *)(pool + a->lds.al[l]); };
    auto Pb = [&](int l) { return (TS *)(pool + a->lds.b[l]); };
    auto Px = [&](int l) { return (TS *)(pool + a->lds.x[l]); };
    auto Pt = [&](int l) { return (TS *)(pool + a->lds.t[l]); };
    auto Pr = [&](int l) { return (TS *)(pool + a->lds.r[l]); };
    for (long q = threadIdx.x; q < a->lds.bytes / 4; q += T) ((unsigned *)pool)[q] = 0u;  // padding = 0
    __syncthreads();
    for (int l = l0; l < L; l++) {
        const Lvl g = a->g[l];
        uint8_t *fl = (uint8_t *)Pfl(l);
        TS *al = (TS *)Pal(l);
        for (long q = threadIdx.x; q < g.S; q += T) fl[q] = a->fl[l][q];
        for (long q = threadIdx.x; q < (long)g.dim * g.S; q += T) al[q] = a->al[l][q];
    }
    {
        const Lvl g = a->g[l0];
        const TS *bsrc = (l0 == 0) ? b0g : a->b[l0];
        for (long q = threadIdx.x; q < (long)g.nfield * g.S; q += T) Pb(l0)[q] = bsrc[q];
    }
    __syncthreads();
    auto smooth = [&](int l, int nu, int first, bool zero) {
        const Lvl g = a->g[l];
        const uint8_t *fl = Pfl(l);
        const TS *al = Pal(l), *b = Pb(l);
        TS *x = Px(l), *t = Pt(l);
        const long nc = (long)g.nx * g.ny * g.nz;
        for (int sw = 0; sw < nu; sw++) {
            if (zero && sw == 0)
                for (long q = threadIdx.x; q < nc; q += T) smooth_body<TS, DIM, true>(g, fl, al, b, x, t, first, om, q);
            else
                for (long q = threadIdx.x; q < nc; q += T) smooth_body<TS, DIM, false>(g, fl, al, b, x, t, first, om, q);
            __syncthreads();
            for (long q = threadIdx.x; q < nc; q += T) smooth_body<TS, DIM, false>(g, fl, al, b, t, x, first ^ 1, om, q);
            __syncthreads();
        }
    };
    for (int l = l0; l + 1 < L; l++) {
        const Lvl g = a->g[l];
        if (a->nu_pre > 0) {
            smooth(l, a->nu_pre, 0, true);
        } else {
            for (long q = threadIdx.x; q < (long)g.nfield * g.S; q += T) Px(l)[q] = Cvt<TS>::st((typename Cvt<TS>::C)0);
            __syncthreads();
        }
        for (long n = threadIdx.x; n < g.S; n += T) residual_body<TS, DIM>(g, Pfl(l), Pal(l), Pb(l), Px(l), Pr(l), n);
        __syncthreads();
        const Lvl gc = a->g[l + 1];
        for (long n = threadIdx.x; n < gc.S; n += T) restrict_body<TS, DIM>(g, gc, Pfl(l + 1), Pr(l), Pb(l + 1), n);
        __syncthreads();
    }
    const int Lc = L - 1, nd = a->nd;
    double *bs = (double *)(pool + a->lds.bs);
    const long *dofs = a->dofs;
    for (int c = threadIdx.x; c < nd; c += T) bs[c] = (double)Cvt<TS>::ld(Pb(Lc)[dofs[c]]);
    __syncthreads();
    for (int r = threadIdx.x; r < nd; r += T) {
        double sacc = 0.0;
        const double *row = a->minv + (long)r * nd;
        for (int c = 0; c < nd; c++) sacc += row[c] * bs[c];
        Px(Lc)[dofs[r]] = Cvt<TS>::st((typename Cvt<TS>::C)sacc);
    }
    __syncthreads();
    for (int l = L - 2; l >= l0; l--) {
        const Lvl g = a->g[l];
        for (long n = threadIdx.x; n < g.S; n += T) prolong_body<TS, DIM>(g, a->g[l + 1], Pfl(l), Px(l + 1), Px(l), n);
        __syncthreads();
        if (a->nu_post > 0) smooth(l, a->nu_post, 1, false);
    }
    {
        const Lvl g = a->g[l0];
        TS *xdst = (l0 == 0) ? x0g : a->x[l0];
        for (long q = threadIdx.x; q < (long)g.nfield * g.S; q += T) xdst[q] = Px(l0)[q];
    }
}

// --------------------------------------------------- ABI <-> padded layout

// field f of the ABI dof vector: dims (nx+(f==1), ny+(f==2), nz+(f==3)),
// C-order, fields concatenated [p | u_x | u_y | u_z]
__device__ __forceinline__ bool abi_index(const Lvl &g, int f, int i, int j, int k, long &idx)
{
    const int A = g.nx + (f == 1), B = g.ny + (f == 2), Cc = g.nz + (f == 3);
    if (i >= A || j >= B || k >= Cc) return false;
    long off = 0;
    long nc = (long)g.nx * g.ny * g.nz;
    long nfx = (long)(g.nx + 1) * g.ny * g.nz, nfy = (long)g.nx * (g.ny + 1) * g.nz;
    if (f >= 1) off += nc;
    if (f >= 2) off += nfx;
    if (f >= 3) off += nfy;
    idx = off + ((long)i * B + j) * Cc + k;
    return true;
}

template <typename TS>
__global__ void __launch_bounds__(256) k_pack(Lvl g, const double *abi, TS *pad, int fixed_only, const uint8_t *fl,
                                              double scale)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.S) return;
    int i, j, k;
    node_ijk(g, n, i, j, k);
    for (int f = 0; f < g.nfield; f++) {
        long idx;
        double v = 0.0;
        if (abi_index(g, f, i, j, k, idx)) v = abi[idx] * scale;
        if (fixed_only && ((fl[n] >> f) & 1)) v = 0.0;
        pad[(long)f * g.S + n] = Cvt<TS>::st((typename Cvt<TS>::C)v);
    }
}

template <typename TS>
__global__ void __launch_bounds__(256) k_unpack(Lvl g, const TS *pad, double *abi)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.S) return;
    int i, j, k;
    node_ijk(g, n, i, j, k);
    for (int f = 0; f < g.nfield; f++) {
        long idx;
        if (abi_index(g, f, i, j, k, idx)) abi[idx] = (double)Cvt<TS>::ld(pad[(long)f * g.S + n]);
    }
}

// flags from ABI fixed flags ([cells | faces], 1 = fixed); padding -> 0
__global__ void k_flags(Lvl g, const uint8_t *abi_fixed, uint8_t *fl)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.S) return;
    int i, j, k;
    node_ijk(g, n, i, j, k);
    int out = 0;
    for (int f = 0; f < g.nfield; f++) {
        long idx;
        if (abi_index(g, f, i, j, k, idx) && !abi_fixed[idx]) out |= 1 << f;
    }
    fl[n] = (uint8_t)out;
}

// alpha from an ABI penalty array ([cells | faces]); zero on non-free faces
template <typename TS>
__global__ void k_alpha_from_pen(Lvl g, const double *pen, const uint8_t *fl, TS *al)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.S) return;
    int i, j, k;
    node_ijk(g, n, i, j, k);
    for (int d = 0; d < g.dim; d++) {
        long idx;
        double v = 0.0;
        if (abi_index(g, 1 + d, i, j, k, idx) && ((fl[n] >> (1 + d)) & 1)) v = pen[idx];
        al[(long)d * g.S + n] = Cvt<TS>::st((typename Cvt<TS>::C)v);
    }
}

// ------------------------------------------------------------- design map

__device__ __forceinline__ double alpha_of(double r, const AlphaParams &a)
{
    const double rp = (a.power == 1.0) ? r : pow(r, a.power);
    return a.amax + (a.amin - a.amax) * rp * (1.0 + a.q) / (rp + a.q);
}
__device__ __forceinline__ double dalpha_of(double r, const AlphaParams &a)
{
    const double rp = (a.power == 1.0) ? r : pow(r, a.power);
    const double dr = (a.amin - a.amax) * ((1.0 + a.q) / (rp + a.q) - rp * (1.0 + a.q) / ((rp + a.q) * (rp + a.q)));
    const double chain = (a.power == 1.0) ? 1.0 : a.power * pow(r, a.power - 1.0);
    return dr * chain;
}

// face alpha = sum over the (<=2) adjacent cells of 0.5*scale*alpha(rho^p)
// (FT/FluidTopoOpt3D.cpp:145-159), zero on fixed faces (ST/StokeFlowDescriptor3D.cu:37-50)
__global__ void __launch_bounds__(256) k_design_alpha(Lvl g, const double *rho, AlphaParams a, const uint8_t *fl,
                                                      double *al)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.S) return;
    int i, j, k;
    node_ijk(g, n, i, j, k);
    const int fb = fl[n];
    for (int d = 0; d < g.dim; d++) {
        double v = 0.0;
        if ((fb >> (1 + d)) & 1) {
            for (int t = 0; t < 2; t++) {
                int a0 = i - t * (d == 0), b0 = j - t * (d == 1), c0 = k - t * (d == 2);
                if (a0 < 0 || b0 < 0 || c0 < 0 || a0 >= g.nx || b0 >= g.ny || c0 >= g.nz) continue;
                v += 0.5 * a.scale * alpha_of(rho[((long)a0 * g.ny + b0) * g.nz + c0], a);
            }
        }
        al[(long)d * g.S + n] = v;
    }
}

// dJ/drho_c = 0.5*scale*alpha'(rho) * sum_{free faces of c}(u^2/2 - u*lam)
// (FT/FluidTopoOpt3D.cpp:170-198), ABI cell order output
__global__ void __launch_bounds__(256) k_sensitivity(Lvl g, const double *rho, AlphaParams a, const uint8_t *fl,
                                                     const double *u, const double *lam, double *dj)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long ncell = (long)g.nx * g.ny * g.nz;
    if (t >= ncell) return;
    int k = (int)(t % g.nz);
    long tt = t / g.nz;
    int j = (int)(tt % g.ny);
    int i = (int)(tt / g.ny);
    const long q = ((long)i * g.Py + j) * g.Pz + k;
    const long st[3] = {(long)g.Py * g.Pz, (long)g.Pz, 1};
    double s = 0.0;
    for (int d = 0; d < g.dim; d++)
        for (int h = 0; h < 2; h++) {
            long qq = q + h * st[d];
            if (!((fl[qq] >> (1 + d)) & 1)) continue;
            double uu = u[(long)(1 + d) * g.S + qq], ll = lam[(long)(1 + d) * g.S + qq];
            s += 0.5 * uu * uu - uu * ll;
        }
    dj[t] = 0.5 * a.scale * dalpha_of(rho[t], a) * s;
}

// ------------------------------------------------ Dirichlet lifting, G x

// b = body (free rows) - A_unmasked-columns(xd); fixed rows take xd
template <int DIM>
__global__ void __launch_bounds__(256) k_lift(Lvl g, const uint8_t *fl, const double *xd, const double *body,
                                              double *b)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.S) return;
    int i, j, k;
    node_ijk(g, n, i, j, k);
    Sten<double, DIM, 2> s{g, fl, xd, nullptr, -1};
    const int fb = fl[n];
    b[n] = (fb & 1) ? body[n] - s.dv(i, j, k) : xd[n];
    b[g.S + n] = ((fb >> 1) & 1) ? body[g.S + n] - s.template visc<0>(i, j, k) : xd[g.S + n];
    b[2 * g.S + n] = ((fb >> 2) & 1) ? body[2 * g.S + n] - s.template visc<1>(i, j, k) : xd[2 * g.S + n];
    if (DIM == 3) b[3 * g.S + n] = ((fb >> 3) & 1) ? body[3 * g.S + n] - s.template visc<2>(i, j, k) : xd[3 * g.S + n];
}

// y = G x (optional) and partial sums of x.Gx: 3D FT/FluidEnergyGrad3D.cu:27-134,
// 2D FT/FluidEnergyGrad.cu:30-61 (J = x.Gx / 2, FT/FluidEnergy{,3D}.cpp:95-109)
template <int DIM>
__global__ void __launch_bounds__(256) k_energy(Lvl g, const uint8_t *fl, const double *al, const double *x, double *y,
                                                double *part)
{
    __shared__ double sh[16];
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (n < g.S) {
        int i, j, k;
        node_ijk(g, n, i, j, k);
        Sten<double, DIM> s{g, fl, x, al, -1};
        double v1, v2, v3 = 0.0;
        if (DIM == 3) {
            v1 = s.template Gface<0>(i, j, k);
            v2 = s.template Gface<1>(i, j, k);
            v3 = s.template Gface<2>(i, j, k);
        } else {  // the reference's 2D objective differs from its 3D one (face Laplacian)
            v1 = s.template Gface2<0>(i, j);
            v2 = s.template Gface2<1>(i, j);
        }
        acc = x[g.S + n] * v1 + x[2 * g.S + n] * v2 + (DIM == 3 ? x[3 * g.S + n] * v3 : 0.0);
        if (y) {
            y[n] = 0.0;
            y[g.S + n] = v1;
            y[2 * g.S + n] = v2;
            if (DIM == 3) y[3 * g.S + n] = v3;
        }
    }
    if (part) {
        double t = block_sum(acc, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
}

// --------------------------------------------------------------- filter

// hat filter weights max(0, R - |o|), neighbours inside the domain only:
// W_i = sum_j w_ij (row weight), forward out_i = sum_j w_ij in_j / W_i,
// transpose out_j = sum_i w_ij in_i / W_i (H symmetric)
// Density filter (the reference's cone filter: weight R - |o| for the neighbours o
// with |o| < R, normalised by the weight sum W of the cell).  Wi = 1 / W per cell is
// computed once per context; the forward filter is (sum h in) Wi, its transpose
// sum h in Wi over the neighbours.
__global__ void __launch_bounds__(256) k_filter_w(int dim, int nx, int ny, int nz, int r, double R, double *Wi)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nc = (long)nx * ny * nz;
    if (t >= nc) return;
    int k = (int)(t % nz);
    long tt = t / nz;
    int j = (int)(tt % ny);
    int i = (int)(tt / ny);
    const int rz = (dim == 3) ? r : 0;
    double w = 0.0;
    for (int a = -r; a <= r; a++)
        for (int b = -r; b <= r; b++)
            for (int c = -rz; c <= rz; c++) {
                int ii = i + a, jj = j + b, kk = k + c;
                if (ii < 0 || jj < 0 || kk < 0 || ii >= nx || jj >= ny || kk >= nz) continue;
                double h = R - sqrt((double)(a * a + b * b + c * c));
                if (h > 0) w += h;
            }
    Wi[t] = 1.0 / w;
}

// neighbour weights by squared offset length, up to radius FILT_RT.  RR > 0: the
// filter's half-width r = RR and the dimension D are compile-time (the taps unroll and
// their loads are issued together; cell index in 32 bits, nc < 2^32); RR = 0: runtime
// r and dimension.
constexpr int FILT_RT = 6;
template <int RR, int D, bool TR>
__global__ void __launch_bounds__(256) k_filter(int dim, int nx, int ny, int nz, int r, double R,
                                                const double *__restrict__ Wi, const double *__restrict__ in,
                                                double *__restrict__ out)
{
    __shared__ double hw[3 * FILT_RT * FILT_RT + 1];
    if (RR > 0) {
        r = RR;
        dim = D;
    }
    const int d2max = 3 * r * r;
    if (r <= FILT_RT)
        for (int q = threadIdx.x; q <= d2max; q += blockDim.x) {
            const double h = R - sqrt((double)q);
            hw[q] = (h > 0) ? h : 0.0;
        }
    __syncthreads();
    // the grid is a multiple of 8 workgroups; workgroup L runs on XCD L % 8 and takes
    // block (L % 8) (grid / 8) + L / 8, so each XCD sweeps a contiguous range of cells
    // and the neighbour planes it re-reads stay in its L2
    const long per = gridDim.x >> 3;
    const long blk = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const long t = blk * blockDim.x + threadIdx.x;
    const long nc = (long)nx * ny * nz;
    if (t >= nc) return;
    int i, j, k;
    if (RR > 0) {
        const unsigned tu = (unsigned)t, tt = tu / (unsigned)nz;
        k = (int)(tu - tt * (unsigned)nz);
        i = (int)(tt / (unsigned)ny);
        j = (int)(tt - (unsigned)i * (unsigned)ny);
    } else {
        k = (int)(t % nz);
        const long tt = t / nz;
        j = (int)(tt % ny);
        i = (int)(tt / ny);
    }
    const int rz = (dim == 3) ? r : 0;
    double s = 0.0;
    if (RR == 1) {
        // per-axis offsets and validity, combined per tap: a neighbour outside the grid
        // reads the cell itself with weight 0 (branch-free, 32-bit cell offsets)
        const int A = ny * nz, B = nz;
        const int oi[3] = {i > 0 ? -A : 0, 0, i < nx - 1 ? A : 0};
        const int oj[3] = {j > 0 ? -B : 0, 0, j < ny - 1 ? B : 0};
        const int ok[3] = {k > 0 ? -1 : 0, 0, k < nz - 1 ? 1 : 0};
        const bool vi[3] = {i > 0, true, i < nx - 1}, vj[3] = {j > 0, true, j < ny - 1},
                   vk[3] = {k > 0, true, k < nz - 1};
        const unsigned tu = (unsigned)t;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++)
#pragma unroll
                for (int c = (D == 3 ? 0 : 1); c < (D == 3 ? 3 : 2); c++) {
                    const int d2 = (a - 1) * (a - 1) + (b - 1) * (b - 1) + (c - 1) * (c - 1);
                    const unsigned q = tu + (unsigned)(oi[a] + oj[b] + ok[c]);
                    const double hv = (vi[a] && vj[b] && vk[c]) ? hw[d2] : 0.0;
                    s += TR ? hv * in[q] * Wi[q] : hv * in[q];
                }
    } else {
#pragma unroll
        for (int a = -r; a <= r; a++)
#pragma unroll
            for (int b = -r; b <= r; b++)
#pragma unroll
                for (int c = -rz; c <= rz; c++) {
                    const int d2 = a * a + b * b + c * c;
                    double h;
                    if (r <= FILT_RT) h = hw[d2];
                    else {
                        h = R - sqrt((double)d2);
                        if (h < 0) h = 0.0;
                    }
                    const int ii = i + a, jj = j + b, kk = k + c;
                    if (h == 0.0 || ii < 0 || jj < 0 || kk < 0 || ii >= nx || jj >= ny || kk >= nz) continue;
                    const long q = ((long)ii * ny + jj) * nz + kk;
                    s += TR ? h * in[q] * Wi[q] : h * in[q];
                }
    }
    out[t] = TR ? s : s * Wi[t];
}

// OC in factored form: at multiplier l the update is rho * base * l^-eta clamped to the
// move limits and bounds, base = (-dc/dv)^eta where -dc/dv > 0 (else 0), so the search
// over l needs no pow per cell and step.  k_oc_vols evaluates the volume sum(v dv) for
// OC_K multipliers in one pass (factors f_k = l_k^-eta); k_oc_apply writes the update
// for one and its volume.
__global__ void __launch_bounds__(256) k_oc_base(long n, const double *dc, const double *dv, double eta, double *base)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const double b0 = -dc[t] / dv[t];
        base[t] = (b0 > 0.0) ? pow(b0, eta) : 0.0;
    }
}

__global__ void __launch_bounds__(256) k_oc_vols(long n, const double *rho, const double *base, const double *dv,
                                                 OcFactors f, double move, double rho_min, double *part)
{
    __shared__ double sh[16];
    double acc[OC_K];
#pragma unroll
    for (int k = 0; k < OC_K; k++) acc[k] = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const double r = rho[t], rb = r * base[t], w = dv[t];
        const double lo = fmax(r - move, rho_min), hi = fmin(r + move, 1.0);
#pragma unroll
        for (int k = 0; k < OC_K; k++) {
            const double v = rb * f.f[k];
            acc[k] += (v < lo ? lo : (v > hi ? hi : v)) * w;
        }
    }
    for (int k = 0; k < OC_K; k++) {
        const double s = block_sum(acc[k], sh);
        if (threadIdx.x == 0) part[(long)k * gridDim.x + blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(256) k_oc_apply(long n, const double *rho, const double *base, const double *dv,
                                                  double fl, double move, double rho_min, double *rn, double *part)
{
    __shared__ double sh[16];
    double acc = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const double r = rho[t];
        const double lo = fmax(r - move, rho_min), hi = fmin(r + move, 1.0);
        const double v0 = r * base[t] * fl;
        const double v = v0 < lo ? lo : (v0 > hi ? hi : v0);
        rn[t] = v;
        acc += v * dv[t];
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_absdiff_max(long n, const double *a, const double *b, double *part)
{
    __shared__ double sh[256];
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    for (long q = t; q < n; q += (long)gridDim.x * blockDim.x) m = fmax(m, fabs(a[q] - b[q]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = blockDim.x >> 1; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ----------------------------------------------------------------- BLAS-1

// sum of partials in fixed order (one block)
// out[v] = sum of the np partials of vector v; one workgroup per vector (grid nvec)
__global__ void __launch_bounds__(1024) k_sum(int np, const double *part, double *out)
{
    __shared__ double sh[16];
    const int v = blockIdx.x;
    double a = 0.0;
    for (int q = threadIdx.x; q < np; q += blockDim.x) a += part[(long)v * np + q];
    const double s = block_sum(a, sh);
    if (threadIdx.x == 0) out[v] = s;
}

template <typename TA, typename TB>
__global__ void __launch_bounds__(256) k_dot(long n, const TA *a, const TB *b, double *part)
{
    __shared__ double sh[16];
    double acc = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        acc += (double)Cvt<TA>::ld(a[q]) * (double)Cvt<TB>::ld(b[q]);
    double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// dots of w with the m vectors V[0..m), V contiguous with stride ld
// partial sums of w . V[v] for v < m (FP64 accumulation; V may be stored in a
// lower precision than w)
template <typename TW, typename TV, int M>
__global__ void __launch_bounds__(256) k_multidot(long n, const TW *w, const TV *V, long ld, int m, double *part)
{
    __shared__ double sh[16];
    double acc[M];
#pragma unroll
    for (int v = 0; v < M; v++) acc[v] = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const double wv = (double)Cvt<TW>::ld(w[q]);
#pragma unroll
        for (int v = 0; v < M; v++)
            if (v < m) acc[v] += wv * (double)Cvt<TV>::ld(V[(long)v * ld + q]);
    }
    for (int v = 0; v < m; v++) {
        double s = block_sum(acc[v], sh);
        if (threadIdx.x == 0) part[(long)v * gridDim.x + blockIdx.x] = s;
    }
}

// w -= sum_v h[v] sc[v] V[v] (sc optional); partial sums of |w|^2
template <typename TW, typename TV, int M>
__global__ void __launch_bounds__(256) k_multiaxpy(long n, TW *w, const TV *V, long ld, int m, const double *h,
                                                   const double *sc, double *part)
{
    __shared__ double sh[16];
    __shared__ double hs[64];
    if (threadIdx.x < (unsigned)m) hs[threadIdx.x] = sc ? h[threadIdx.x] * sc[threadIdx.x] : h[threadIdx.x];
    __syncthreads();
    double acc = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        double wv = (double)Cvt<TW>::ld(w[q]);
#pragma unroll
        for (int v = 0; v < M; v++)
            if (v < m) wv -= hs[v] * (double)Cvt<TV>::ld(V[(long)v * ld + q]);
        w[q] = Cvt<TW>::st((typename Cvt<TW>::C)wv);
        acc += wv * wv;
    }
    if (part) {
        double s = block_sum(acc, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = s;
    }
}

// Four consecutive elements per lane in one 16-byte (float, double: two) or 8-byte
// (bf16) load or store, as FP64 values.
template <typename T> struct Quad;
template <> struct Quad<double> {
    __device__ static void ld(const double *p, long i, double *o)
    {
        const double2 a = reinterpret_cast<const double2 *>(p)[2 * i], b = reinterpret_cast<const double2 *>(p)[2 * i + 1];
        o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
    }
    __device__ static void st(double *p, long i, const double *o)
    {
        reinterpret_cast<double2 *>(p)[2 * i] = make_double2(o[0], o[1]);
        reinterpret_cast<double2 *>(p)[2 * i + 1] = make_double2(o[2], o[3]);
    }
};
template <> struct Quad<float> {
    __device__ static void ld(const float *p, long i, double *o)
    {
        const float4 a = reinterpret_cast<const float4 *>(p)[i];
        o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    }
    __device__ static void st(float *p, long i, const double *o)
    {
        reinterpret_cast<float4 *>(p)[i] = make_float4((float)o[0], (float)o[1], (float)o[2], (float)o[3]);
    }
};
template <> struct Quad<bf16> {
    __device__ static void ld(const bf16 *p, long i, double *o)
    {
        const uint2 a = reinterpret_cast<const uint2 *>(p)[i];
        bf16 h[4] = {{(uint16_t)(a.x & 0xffff)}, {(uint16_t)(a.x >> 16)}, {(uint16_t)(a.y & 0xffff)}, {(uint16_t)(a.y >> 16)}};
        for (int e = 0; e < 4; e++) o[e] = to_f(h[e]);
    }
    __device__ static void st(bf16 *p, long i, const double *o)
    {
        uint16_t h[4];
        for (int e = 0; e < 4; e++) h[e] = to_bf16((float)o[e]).v;
        reinterpret_cast<uint2 *>(p)[i] = make_uint2(h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16));
    }
};

// k_multidot / k_multiaxpy on four elements per lane (n and ld multiples of 4, the
// vectors 16-byte aligned): one wide load per vector and group instead of four
// narrow ones; n4 = n / 4 groups
template <typename TW, typename TV, int M>
__global__ void __launch_bounds__(256) k_multidot4(long n4, const TW *w, const TV *V, long ld, int m, double *part)
{
    __shared__ double sh[16];
    double acc[M];
#pragma unroll
    for (int v = 0; v < M; v++) acc[v] = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += (long)gridDim.x * blockDim.x) {
        double a[4];
        Quad<TW>::ld(w, q, a);
#pragma unroll
        for (int v = 0; v < M; v++)
            if (v < m) {
                double b[4];
                Quad<TV>::ld(V + (long)v * ld, q, b);
                acc[v] += (a[0] * b[0] + a[1] * b[1]) + (a[2] * b[2] + a[3] * b[3]);
            }
    }
    for (int v = 0; v < m; v++) {
        double s = block_sum(acc[v], sh);
        if (threadIdx.x == 0) part[(long)v * gridDim.x + blockIdx.x] = s;
    }
}

template <typename TW, typename TV, int M>
__global__ void __launch_bounds__(256) k_multiaxpy4(long n4, TW *w, const TV *V, long ld, int m, const double *h,
                                                    const double *sc, double *part)
{
    __shared__ double sh[16];
    __shared__ double hs[64];
    if (threadIdx.x < (unsigned)m) hs[threadIdx.x] = sc ? h[threadIdx.x] * sc[threadIdx.x] : h[threadIdx.x];
    __syncthreads();
    double acc = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += (long)gridDim.x * blockDim.x) {
        double a[4];
        Quad<TW>::ld(w, q, a);
#pragma unroll
        for (int v = 0; v < M; v++)
            if (v < m) {
                double b[4];
                Quad<TV>::ld(V + (long)v * ld, q, b);
                const double hv = hs[v];
#pragma unroll
                for (int e = 0; e < 4; e++) a[e] -= hv * b[e];
            }
        Quad<TW>::st(w, q, a);
        acc += (a[0] * a[0] + a[1] * a[1]) + (a[2] * a[2] + a[3] * a[3]);
    }
    if (part) {
        double s = block_sum(acc, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = s;
    }
}

// y = a*x + b*y  (types may differ: conversions between precisions)
template <typename TX, typename TY>
__global__ void __launch_bounds__(256) k_axpby(long n, double a, const TX *x, double b, TY *y)
{
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        double v = a * (double)Cvt<TX>::ld(x[q]);
        if (b != 0.0) v += b * (double)Cvt<TY>::ld(y[q]);
        y[q] = Cvt<TY>::st((typename Cvt<TY>::C)v);
    }
}

// y = a x + b w + c y (FP64; c = 0 does not read y): the extrapolated start of a solve
__global__ void __launch_bounds__(256) k_lincomb3(long n, double a, const double *x, double b, const double *w,
                                                  double c, double *y)
{
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        double v = a * x[q] + b * w[q];
        if (c != 0.0) v += c * y[q];
        y[q] = v;
    }
}

// Minimal-residual start (see st_solver.hip): d = float(a - b); the Gram matrix of the
// images y_i = A D_i and their products with the residual r in one pass; the update
// x0 = x1 + sum d_i D_i, r <- r - sum d_i y_i with the new ||r||^2.
__global__ void __launch_bounds__(256) k_diff_f32(long n, const double *a, const double *b, float *d)
{
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        d[q] = (float)(a[q] - b[q]);
}

// part[k][block], k over G_ij (i <= j < ND, row by row) then g_i = y_i . r
template <int ND>
__global__ void __launch_bounds__(256) k_mr_gram(long n, const float *y, long ld, const float *r, double *part)
{
    __shared__ double sh[16];
    constexpr int NQ = ND * (ND + 1) / 2 + ND;
    double acc[NQ];
#pragma unroll
    for (int k = 0; k < NQ; k++) acc[k] = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        double v[ND];
#pragma unroll
        for (int i = 0; i < ND; i++) v[i] = (double)y[i * ld + q];
        const double rq = (double)r[q];
        int k = 0;
#pragma unroll
        for (int i = 0; i < ND; i++)
#pragma unroll
            for (int j = i; j < ND; j++) acc[k++] += v[i] * v[j];
#pragma unroll
        for (int i = 0; i < ND; i++) acc[k++] += v[i] * rq;
    }
#pragma unroll
    for (int k = 0; k < NQ; k++) {
        const double t = block_sum(acc[k], sh);
        if (threadIdx.x == 0) part[(long)k * gridDim.x + blockIdx.x] = t;
    }
}

template <int ND>
__global__ void __launch_bounds__(256) k_mr_update(long n, const double *x1, const float *D0, const float *D1,
                                                   const float *D2, const float *y, long ld, MrCoef cf, double *x0,
                                                   float *r, double *part)
{
    __shared__ double sh[16];
    const float *D[3] = {D0, D1, D2};
    double acc = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        double xv = x1[q], rv = (double)r[q];
#pragma unroll
        for (int i = 0; i < ND; i++) {
            xv += cf.d[i] * (double)D[i][q];
            rv -= cf.d[i] * (double)y[i * ld + q];
        }
        x0[q] = xv;
        const float rf = (float)rv;
        r[q] = rf;
        acc += (double)rf * (double)rf;
    }
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <typename TS>
__global__ void __launch_bounds__(256) k_fill(long n, TS *y, double v)
{
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        y[q] = Cvt<TS>::st((typename Cvt<TS>::C)v);
}

// ============================================================ launchers

static inline unsigned nb(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }
static inline unsigned nb_grid(long n)
{
    long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (unsigned)b;
}

#define DISPATCH_DIM(dim, F, ...) \
    do { if ((dim) == 3) F<3>(__VA_ARGS__); else F<2>(__VA_ARGS__); } while (0)

template <typename TS>
void launch_apply(const Lvl &g, const uint8_t *fl, const TS *al, const TS *x, TS *y, hipStream_t s)
{
    if (march_ok(g)) { march_apply<TS>(g, fl, al, x, y, s); return; }
    if (g.dim == 3) hipLaunchKernelGGL((k_apply<TS, 3>), dim3(nb(g.S)), dim3(256), 0, s, g, fl, al, x, y);
    else hipLaunchKernelGGL((k_apply<TS, 2>), dim3(nb(g.S)), dim3(256), 0, s, g, fl, al, x, y);
}
template <typename TS>
void launch_residual(const Lvl &g, const uint8_t *fl, const TS *al, const TS *b, const TS *x, TS *r, double *part,
                     hipStream_t s)
{
    if (march_ok(g)) { march_residual<TS>(g, fl, al, b, x, r, part, s); return; }
    if (g.dim == 3) hipLaunchKernelGGL((k_residual<TS, 3>), dim3(nb(g.S)), dim3(256), 0, s, g, fl, al, b, x, r, part);
    else hipLaunchKernelGGL((k_residual<TS, 2>), dim3(nb(g.S)), dim3(256), 0, s, g, fl, al, b, x, r, part);
}
unsigned residual_blocks(const Lvl &g) { return march_ok(g) ? march_blocks(g) : nb(g.S); }

template <typename TS>
void launch_smooth(const Lvl &g, const uint8_t *fl, const TS *al, const TS *b, const TS *xin, TS *xout, int color,
                   float omega, bool zero, hipStream_t s)
{
    long nc = (long)g.nx * g.ny * g.nz;
    if (!zero && march_ok(g)) { march_smooth<TS>(g, fl, al, b, xin, xout, color, omega, s); return; }
    if (zero && g.dim == 3 && g.nz % 2 == 0) {
        hipLaunchKernelGGL((k_smooth0_pairs<TS>), dim3(nb(nc / 2)), dim3(256), 0, s, g, fl, al, b, xout, color, omega);
        return;
    }
    if (g.dim == 3) {
        if (zero) hipLaunchKernelGGL((k_smooth<TS, 3, true>), dim3(nb(nc)), dim3(256), 0, s, g, fl, al, b, xin, xout, color, omega);
        else hipLaunchKernelGGL((k_smooth<TS, 3, false>), dim3(nb(nc)), dim3(256), 0, s, g, fl, al, b, xin, xout, color, omega);
    } else {
        if (zero) hipLaunchKernelGGL((k_smooth<TS, 2, true>), dim3(nb(nc)), dim3(256), 0, s, g, fl, al, b, xin, xout, color, omega);
        else hipLaunchKernelGGL((k_smooth<TS, 2, false>), dim3(nb(nc)), dim3(256), 0, s, g, fl, al, b, xin, xout, color, omega);
    }
}
template <typename TS>
void launch_restrict(const Lvl &gf, const Lvl &gc, const uint8_t *flc, const TS *rf, TS *bc, hipStream_t s)
{
    if (gf.dim == 3) hipLaunchKernelGGL((k_restrict<TS, 3>), dim3(nb(gc.S)), dim3(256), 0, s, gf, gc, flc, rf, bc);
    else hipLaunchKernelGGL((k_restrict<TS, 2>), dim3(nb(gc.S)), dim3(256), 0, s, gf, gc, flc, rf, bc);
}
template <typename TS>
void launch_prolong(const Lvl &gf, const Lvl &gc, const uint8_t *flf, const TS *ec, TS *xf, hipStream_t s)
{
    if (gf.dim == 3) hipLaunchKernelGGL((k_prolong<TS, 3>), dim3(nb(gf.S)), dim3(256), 0, s, gf, gc, flf, ec, xf);
    else hipLaunchKernelGGL((k_prolong<TS, 2>), dim3(nb(gf.S)), dim3(256), 0, s, gf, gc, flf, ec, xf);
}
template <typename TF, typename TC>
void launch_coarsen(const Lvl &gf, const Lvl &gc, const uint8_t *flf, const TF *alf, uint8_t *flc, TC *alc,
                    hipStream_t s)
{
    if (gf.dim == 3) hipLaunchKernelGGL((k_coarsen<TF, TC, 3>), dim3(nb(gc.S)), dim3(256), 0, s, gf, gc, flf, alf, flc, alc);
    else hipLaunchKernelGGL((k_coarsen<TF, TC, 2>), dim3(nb(gc.S)), dim3(256), 0, s, gf, gc, flf, alf, flc, alc);
}
template <typename TS>
void launch_probe(const Lvl &g, const uint8_t *fl, const TS *al, const long *dofs, int nd, int closed, double *A,
                  hipStream_t s)
{
    if (g.dim == 3) hipLaunchKernelGGL((k_probe<TS, 3>), dim3(nd), dim3(256), 0, s, g, fl, al, dofs, nd, closed, A);
    else hipLaunchKernelGGL((k_probe<TS, 2>), dim3(nd), dim3(256), 0, s, g, fl, al, dofs, nd, closed, A);
}
void launch_gauss_jordan(int n, double *A, int *perm, int *info, hipStream_t s)
{
    hipLaunchKernelGGL(k_gauss_jordan, dim3(1), dim3(1024), 0, s, n, A, perm, info);
}
template <typename TS>
void launch_coarse_solve(int nd, const long *dofs, const double *Minv, const TS *b, TS *x, hipStream_t s)
{
    hipLaunchKernelGGL((k_coarse_solve<TS>), dim3(1), dim3(512), nd * sizeof(double), s, nd, dofs, Minv, b, x);
}
template <typename TS>
void launch_pack(const Lvl &g, const double *abi, TS *pad, int fixed_only, const uint8_t *fl, double scale,
                 hipStream_t s)
{
    hipLaunchKernelGGL((k_pack<TS>), dim3(nb(g.S)), dim3(256), 0, s, g, abi, pad, fixed_only, fl, scale);
}
template <typename TS>
void launch_unpack(const Lvl &g, const TS *pad, double *abi, hipStream_t s)
{
    hipLaunchKernelGGL((k_unpack<TS>), dim3(nb(g.S)), dim3(256), 0, s, g, pad, abi);
}
void launch_flags(const Lvl &g, const uint8_t *abi_fixed, uint8_t *fl, hipStream_t s)
{
    hipLaunchKernelGGL(k_flags, dim3(nb(g.S)), dim3(256), 0, s, g, abi_fixed, fl);
}
template <typename TS>
void launch_alpha_from_pen(const Lvl &g, const double *pen, const uint8_t *fl, TS *al, hipStream_t s)
{
    hipLaunchKernelGGL((k_alpha_from_pen<TS>), dim3(nb(g.S)), dim3(256), 0, s, g, pen, fl, al);
}
void launch_design_alpha(const Lvl &g, const double *rho, const AlphaParams &a, const uint8_t *fl, double *al,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_design_alpha, dim3(nb(g.S)), dim3(256), 0, s, g, rho, a, fl, al);
}
void launch_sensitivity(const Lvl &g, const double *rho, const AlphaParams &a, const uint8_t *fl, const double *u,
                        const double *lam, double *dj, hipStream_t s)
{
    long nc = (long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_sensitivity, dim3(nb(nc)), dim3(256), 0, s, g, rho, a, fl, u, lam, dj);
}
void launch_lift(const Lvl &g, const uint8_t *fl, const double *xd, const double *body, double *b, hipStream_t s)
{
    if (g.dim == 3) hipLaunchKernelGGL((k_lift<3>), dim3(nb(g.S)), dim3(256), 0, s, g, fl, xd, body, b);
    else hipLaunchKernelGGL((k_lift<2>), dim3(nb(g.S)), dim3(256), 0, s, g, fl, xd, body, b);
}
void launch_energy(const Lvl &g, const uint8_t *fl, const double *al, const double *x, double *y, double *part,
                   hipStream_t s)
{
    if (g.dim == 3) hipLaunchKernelGGL((k_energy<3>), dim3(nb(g.S)), dim3(256), 0, s, g, fl, al, x, y, part);
    else hipLaunchKernelGGL((k_energy<2>), dim3(nb(g.S)), dim3(256), 0, s, g, fl, al, x, y, part);
}
unsigned energy_blocks(const Lvl &g) { return nb(g.S); }
void launch_filter(const Lvl &g, double R, const double *Wi, const double *in, double *out, int transpose,
                   hipStream_t s)
{
    long nc = (long)g.nx * g.ny * g.nz;
    int r = (int)ceil(R) - 1;
    if (r < 0) r = 0;
    const dim3 gr((nb(nc) + 7) / 8 * 8), bl(256);
#define ST_FILT(RR, D)                                                                                        \
    do {                                                                                                      \
        if (transpose) hipLaunchKernelGGL((k_filter<RR, D, true>), gr, bl, 0, s, g.dim, g.nx, g.ny, g.nz, r, R, Wi, in, out); \
        else hipLaunchKernelGGL((k_filter<RR, D, false>), gr, bl, 0, s, g.dim, g.nx, g.ny, g.nz, r, R, Wi, in, out);        \
    } while (0)
    if (r == 1 && nc < (1L << 32)) {
        if (g.dim == 3) ST_FILT(1, 3);
        else ST_FILT(1, 2);
    } else ST_FILT(0, 0);
#undef ST_FILT
}
void launch_filter_w(const Lvl &g, double R, double *Wi, hipStream_t s)
{
    long nc = (long)g.nx * g.ny * g.nz;
    int r = (int)ceil(R) - 1;
    if (r < 0) r = 0;
    hipLaunchKernelGGL(k_filter_w, dim3(nb(nc)), dim3(256), 0, s, g.dim, g.nx, g.ny, g.nz, r, R, Wi);
}
unsigned oc_grid(long n) { return nb_grid(n); }
void launch_oc_base(long n, const double *dc, const double *dv, double eta, double *base, hipStream_t s)
{
    hipLaunchKernelGGL(k_oc_base, dim3(nb_grid(n)), dim3(256), 0, s, n, dc, dv, eta, base);
}
void launch_oc_vols(long n, const double *rho, const double *base, const double *dv, const OcFactors &f, double move,
                    double rho_min, double *part, hipStream_t s)
{
    hipLaunchKernelGGL(k_oc_vols, dim3(nb_grid(n)), dim3(256), 0, s, n, rho, base, dv, f, move, rho_min, part);
}
void launch_oc_apply(long n, const double *rho, const double *base, const double *dv, double fl, double move,
                     double rho_min, double *rn, double *part, hipStream_t s)
{
    hipLaunchKernelGGL(k_oc_apply, dim3(nb_grid(n)), dim3(256), 0, s, n, rho, base, dv, fl, move, rho_min, rn, part);
}
unsigned absmax_blocks(long n) { return nb_grid(n); }
void launch_absdiff_max(long n, const double *a, const double *b, double *part, hipStream_t s)
{
    hipLaunchKernelGGL(k_absdiff_max, dim3(nb_grid(n)), dim3(256), 0, s, n, a, b, part);
}
void launch_sum(int np, int nvec, const double *part, double *out, hipStream_t s)
{
    if (nvec < 1) return;
    hipLaunchKernelGGL(k_sum, dim3((unsigned)nvec), dim3(1024), 0, s, np, part, out);
}
unsigned blas_blocks(long n) { return nb_grid(n); }
template <typename TA, typename TB>
void launch_dot(long n, const TA *a, const TB *b, double *part, hipStream_t s)
{
    hipLaunchKernelGGL((k_dot<TA, TB>), dim3(nb_grid(n)), dim3(256), 0, s, n, a, b, part);
}
// the four-element kernels apply when every vector starts on a 16-byte boundary
template <typename TW, typename TV>
static bool quad_ok(long n, const TW *w, const TV *V, long ld)
{
    return n % 4 == 0 && ld % 4 == 0 && ((uintptr_t)w % 16) == 0 && ((uintptr_t)V % 16) == 0 &&
           (ld * (long)sizeof(TV)) % 16 == 0;
}

template <typename TW, typename TV>
void launch_multidot(long n, const TW *w, const TV *V, long ld, int m, double *part, hipStream_t s)
{
    if (quad_ok(n, w, V, ld)) {
        const dim3 gr(nb_grid(n));
        if (m <= 8) hipLaunchKernelGGL((k_multidot4<TW, TV, 8>), gr, dim3(256), 0, s, n / 4, w, V, ld, m, part);
        else if (m <= 16) hipLaunchKernelGGL((k_multidot4<TW, TV, 16>), gr, dim3(256), 0, s, n / 4, w, V, ld, m, part);
        else hipLaunchKernelGGL((k_multidot4<TW, TV, 32>), gr, dim3(256), 0, s, n / 4, w, V, ld, m, part);
        return;
    }
    if (m <= 8) hipLaunchKernelGGL((k_multidot<TW, TV, 8>), dim3(nb_grid(n)), dim3(256), 0, s, n, w, V, ld, m, part);
    else if (m <= 16) hipLaunchKernelGGL((k_multidot<TW, TV, 16>), dim3(nb_grid(n)), dim3(256), 0, s, n, w, V, ld, m, part);
    else hipLaunchKernelGGL((k_multidot<TW, TV, 32>), dim3(nb_grid(n)), dim3(256), 0, s, n, w, V, ld, m, part);
}
template <typename TW, typename TV>
void launch_multiaxpy(long n, TW *w, const TV *V, long ld, int m, const double *h, const double *sc, double *part,
                      hipStream_t s)
{
    if (quad_ok(n, (const TW *)w, V, ld)) {
        const dim3 gr(nb_grid(n));
        if (m <= 8) hipLaunchKernelGGL((k_multiaxpy4<TW, TV, 8>), gr, dim3(256), 0, s, n / 4, w, V, ld, m, h, sc, part);
        else if (m <= 16) hipLaunchKernelGGL((k_multiaxpy4<TW, TV, 16>), gr, dim3(256), 0, s, n / 4, w, V, ld, m, h, sc, part);
        else hipLaunchKernelGGL((k_multiaxpy4<TW, TV, 32>), gr, dim3(256), 0, s, n / 4, w, V, ld, m, h, sc, part);
        return;
    }
    if (m <= 8) hipLaunchKernelGGL((k_multiaxpy<TW, TV, 8>), dim3(nb_grid(n)), dim3(256), 0, s, n, w, V, ld, m, h, sc, part);
    else if (m <= 16) hipLaunchKernelGGL((k_multiaxpy<TW, TV, 16>), dim3(nb_grid(n)), dim3(256), 0, s, n, w, V, ld, m, h, sc, part);
    else hipLaunchKernelGGL((k_multiaxpy<TW, TV, 32>), dim3(nb_grid(n)), dim3(256), 0, s, n, w, V, ld, m, h, sc, part);
}
void launch_lincomb3(long n, double a, const double *x, double b, const double *w, double c, double *y, hipStream_t s)
{
    hipLaunchKernelGGL(k_lincomb3, dim3(nb_grid(n)), dim3(256), 0, s, n, a, x, b, w, c, y);
}
void launch_diff_f32(long n, const double *a, const double *b, float *d, hipStream_t s)
{
    hipLaunchKernelGGL(k_diff_f32, dim3(nb_grid(n)), dim3(256), 0, s, n, a, b, d);
}
unsigned mr_blocks(long n) { return nb_grid(n); }
void launch_mr_gram(int nd, long n, const float *y, long ld, const float *r, double *part, hipStream_t s)
{
    const dim3 g(nb_grid(n)), b(256);
    if (nd == 1) hipLaunchKernelGGL((k_mr_gram<1>), g, b, 0, s, n, y, ld, r, part);
    else if (nd == 2) hipLaunchKernelGGL((k_mr_gram<2>), g, b, 0, s, n, y, ld, r, part);
    else hipLaunchKernelGGL((k_mr_gram<3>), g, b, 0, s, n, y, ld, r, part);
}
void launch_mr_update(int nd, long n, const double *x1, const float *const *D, const float *y, long ld,
                      const MrCoef &cf, double *x0, float *r, double *part, hipStream_t s)
{
    const dim3 g(nb_grid(n)), b(256);
    const float *D0 = D[0], *D1 = nd > 1 ? D[1] : D[0], *D2 = nd > 2 ? D[2] : D[0];
    if (nd == 1)
        hipLaunchKernelGGL((k_mr_update<1>), g, b, 0, s, n, x1, D0, D1, D2, y, ld, cf, x0, r, part);
    else if (nd == 2)
        hipLaunchKernelGGL((k_mr_update<2>), g, b, 0, s, n, x1, D0, D1, D2, y, ld, cf, x0, r, part);
    else hipLaunchKernelGGL((k_mr_update<3>), g, b, 0, s, n, x1, D0, D1, D2, y, ld, cf, x0, r, part);
}
template <typename TX, typename TY>
void launch_axpby(long n, double a, const TX *x, double b, TY *y, hipStream_t s)
{
    hipLaunchKernelGGL((k_axpby<TX, TY>), dim3(nb_grid(n)), dim3(256), 0, s, n, a, x, b, y);
}
template <typename TS>
void launch_fill(long n, TS *y, double v, hipStream_t s)
{
    hipLaunchKernelGGL((k_fill<TS>), dim3(nb_grid(n)), dim3(256), 0, s, n, y, v);
}

template <typename TS>
void launch_vcycle_tail(const TailArgs<TS> *a, int dim, const TS *b0, TS *x0, int nd, bool lds, hipStream_t s)
{
    if (lds) {
        if (dim == 3) hipLaunchKernelGGL((k_vcycle_tail_lds<TS, 3>), dim3(1), dim3(512), 0, s, a, b0, x0);
        else hipLaunchKernelGGL((k_vcycle_tail_lds<TS, 2>), dim3(1), dim3(512), 0, s, a, b0, x0);
        return;
    }
    const size_t sh = sizeof(double) * (size_t)(nd > 0 ? nd : 1);
    if (dim == 3) hipLaunchKernelGGL((k_vcycle_tail<TS, 3>), dim3(1), dim3(512), sh, s, a, b0, x0);
    else hipLaunchKernelGGL((k_vcycle_tail<TS, 2>), dim3(1), dim3(512), sh, s, a, b0, x0);
}

// explicit instantiations
#define INST_TS(T)                                                                                          \
    template void launch_apply<T>(const Lvl &, const uint8_t *, const T *, const T *, T *, hipStream_t);     \
    template void launch_residual<T>(const Lvl &, const uint8_t *, const T *, const T *, const T *, T *,     \
                                     double *, hipStream_t);                                               \
    template void launch_smooth<T>(const Lvl &, const uint8_t *, const T *, const T *, const T *, T *, int,  \
                                   float, bool, hipStream_t);                                              \
    template void launch_restrict<T>(const Lvl &, const Lvl &, const uint8_t *, const T *, T *, hipStream_t); \
    template void launch_prolong<T>(const Lvl &, const Lvl &, const uint8_t *, const T *, T *, hipStream_t);  \
    template void launch_probe<T>(const Lvl &, const uint8_t *, const T *, const long *, int, int, double *,   \
                                  hipStream_t);                                                            \
    template void launch_coarse_solve<T>(int, const long *, const double *, const T *, T *, hipStream_t);    \
    template void launch_pack<T>(const Lvl &, const double *, T *, int, const uint8_t *, double, hipStream_t); \
    template void launch_unpack<T>(const Lvl &, const T *, double *, hipStream_t);                          \
    template void launch_alpha_from_pen<T>(const Lvl &, const double *, const uint8_t *, T *, hipStream_t);  \
    template void launch_fill<T>(long, T *, double, hipStream_t);                                           \
    template void launch_vcycle_tail<T>(const TailArgs<T> *, int, const T *, T *, int, bool, hipStream_t);

INST_TS(double)
INST_TS(float)
INST_TS(bf16)

#define INST_PAIR(A, B)                                                                   \
    template void launch_dot<A, B>(long, const A *, const B *, double *, hipStream_t);   \
    template void launch_axpby<A, B>(long, double, const A *, double, B *, hipStream_t); \
    template void launch_coarsen<A, B>(const Lvl &, const Lvl &, const uint8_t *, const A *, uint8_t *, B *, hipStream_t);

INST_PAIR(double, double)
INST_PAIR(double, float)
INST_PAIR(float, double)
INST_PAIR(float, float)
INST_PAIR(double, bf16)
INST_PAIR(bf16, double)
INST_PAIR(bf16, bf16)
INST_PAIR(float, bf16)

#define INST_KRYLOV(W, V)                                                                                   \
    template void launch_multidot<W, V>(long, const W *, const V *, long, int, double *, hipStream_t);        \
    template void launch_multiaxpy<W, V>(long, W *, const V *, long, int, const double *, const double *, double *, \
                                         hipStream_t);
INST_KRYLOV(double, double)
INST_KRYLOV(double, float)
INST_KRYLOV(float, float)
INST_PAIR(bf16, float)

}  // namespace st
