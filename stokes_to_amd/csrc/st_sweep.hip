// One whole red-black box-Vanka sweep (first colour, then the second) of a fine 3D
// level in one pass over HBM, for the mixed-precision smoother (FP32 arithmetic;
// FP32 or BF16 storage).
//
// The two-pass form (st_march.hip, M_SMOOTH) reads x, b, alpha and writes x once per
// colour, i.e. twice per sweep.  Here a workgroup owns a (j, k) tile of BJ x 64 cells
// and marches i through a chunk of planes, running both colours in a pipeline:
//   step t:  stage x of plane t+4 in LDS; edge curls E and m_c div G of plane t+3;
//            first colour on the cells of plane t+2 -> x' (LDS, raw values);
//            E', G' of x' for plane t+1; second colour on the tile cells of plane t
//            from x', written to HBM.
// The second colour of a tile needs x' one cell around the tile, which needs the
// first colour on a ring of two cells around it, which reads x three cells around
// it: the first colour is recomputed on that ring by every tile that needs it (and on
// two planes before and after the chunk), so no workgroup waits for another.  Every
// block update uses the arithmetic of the two-pass kernel (the exact closed-form block
// solve of st_kernels.hip smooth_cell), so the result equals two colour passes; the
// intermediate x' stays in FP32 in LDS (with BF16 storage it is not rounded to BF16
// between the colours).
//
// LDS (FP32): x 4 planes x 4 fields, E 2 x 3, G 3 planes on the (BJ+7) x 71 staging
// region; x' 4 planes x 4 fields, E' 2 x 3, G' 2 on the (BJ+3) x 67 tile region; flags
// 8 planes: 126 KB, one workgroup of 512 threads per CU.
#include "st_kernels.h"

namespace st {

namespace {
constexpr int SB_J = 4, SB_K = 64;            // tile of cells
constexpr int S3_H = SB_J + 7, S3_W = SB_K + 7;  // x staging region: j0-3 .. j0+BJ+3, k0-3 .. k0+BK+3
constexpr int S3_N = S3_H * S3_W;
constexpr int S1_H = SB_J + 3, S1_W = SB_K + 3;  // x' region: j0-1 .. j0+BJ+1, k0-1 .. k0+BK+1
constexpr int S1_N = S1_H * S1_W;
constexpr int SW_NT = 512;
constexpr int C0_ROW = (SB_K + 4) / 2;          // first-colour cells per row of the ring region
constexpr int C0_N = (SB_J + 4) * C0_ROW;        // first-colour cells per plane (rows j0-2 .. j0+BJ+1)
static_assert(2 * SW_NT >= S3_N, "a plane is staged two nodes per thread");
static_assert(SW_NT >= C0_N, "one first-colour cell per thread");

__device__ __forceinline__ int m4(int p) { return p & 3; }       // planes >= -8 (offset below)
__device__ __forceinline__ int m8(int p) { return p & 7; }
__device__ __forceinline__ int m2(int p) { return p & 1; }
__device__ __forceinline__ int m3(int p) { return (p + 9) % 3; }  // p >= -9
__device__ __forceinline__ float rcpf_(float x) { return __builtin_amdgcn_rcpf(x); }

// closed-form block solve of one cell (st_kernels.hip smooth_cell, st_march.hip), in
// the operation order of the two-pass kernel: inputs the 7 residuals, masks and
// diagonal entries; outputs dp and the two face corrections per axis, times omega
struct Blk {
    float dp, d0[3], d1[3];
};
__device__ __forceinline__ Blk block_solve(float mc, float rp, const float *r0, const float *r1, const float *m0,
                                           const float *m1, const float *a0, const float *a1, float om)
{
    Blk o;
    float beta = 0.f, tb = 0.f, of[3], id[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        of[d] = -m0[d] * m1[d];
        id[d] = rcpf_(a0[d] * a1[d] - of[d] * of[d]);
        const float kb0 = id[d] * (a1[d] * (-m0[d]) - of[d] * m1[d]);
        const float kb1 = id[d] * (-of[d] * (-m0[d]) + a0[d] * m1[d]);
        const float kr0 = id[d] * (a1[d] * r0[d] - of[d] * r1[d]);
        const float kr1 = id[d] * (-of[d] * r0[d] + a0[d] * r1[d]);
        beta += (-m0[d]) * kb0 + m1[d] * kb1;
        tb += (-m0[d]) * kr0 + m1[d] * kr1;
    }
    const float Sc = (mc != 0.f) ? -beta : 1.f;
    const float dp = (Sc != 0.f) ? (rp - mc * tb) * rcpf_(Sc) : 0.f;
    const float gm = (mc != 0.f) ? dp : (dp - tb) * rcpf_(1.f - beta);
    o.dp = dp;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const float e0 = r0[d] + m0[d] * gm, e1 = r1[d] - m1[d] * gm;
        o.d0[d] = om * id[d] * (a1[d] * e0 - of[d] * e1);
        o.d1[d] = om * id[d] * (-of[d] * e0 + a0[d] * e1);
    }
    return o;
}
}  // namespace

template <typename TS>
__global__ void __launch_bounds__(SW_NT) k_sweep(Lvl g, const uint8_t *__restrict__ fl, const TS *__restrict__ al,
                                                 const TS *__restrict__ b, const TS *__restrict__ x,
                                                 TS *__restrict__ y, int chunk, int col0, float omega)
{
    typedef float C;
    __shared__ C sX[4][4][S3_N];   // raw x: p, ux, uy, uz of planes t+1 .. t+4
    __shared__ uint8_t sF[8][S3_N];
    __shared__ C sE[2][3][S3_N];   // edge curls of x, planes t+2, t+3
    __shared__ C sG[3][S3_N];      // m_c div w of x, planes t+1 .. t+3
    __shared__ C sXp[4][4][S1_N];  // raw x' (after the first colour), planes t .. t+3
    __shared__ C sEp[2][3][S1_N];
    __shared__ C sGp[2][S1_N];

    const int tid = threadIdx.x;
    const int nbx = gridDim.x, nby = gridDim.y, nbz = gridDim.z;
    const int nbt = nbx * nby * nbz;
    const int hw = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const int lid = (nbt % 8 == 0) ? (hw % 8) * (nbt / 8) + hw / 8 : hw;  // XCD-aware, as k_march
    const int bx = lid % nbx, by = (lid / nbx) % nby, bz = lid / (nbx * nby);
    const int k0 = bx * SB_K, j0 = by * SB_J;
    const int ic0 = bz * chunk;
    int ic1 = (bz == nbz - 1) ? g.nx : ic0 + chunk;
    if (ic1 > g.nx) ic1 = g.nx;
    const int col1 = col0 ^ 1;
    const long S = g.S;
    const long sI = (long)g.Py * g.Pz, sJ = g.Pz;
    const C om = (C)omega;

    // masked velocity of x (staging region) and of x' (tile region)
    auto Wx = [&](int pl, int d, int c) -> C {
        return ((sF[m8(pl)][c] >> (1 + d)) & 1) ? sX[m4(pl)][1 + d][c] : C(0);
    };
    auto c3of1 = [](int c1) { return (c1 / S1_W + 2) * S3_W + (c1 % S1_W + 2); };
    auto Wp = [&](int pl, int d, int c1) -> C {
        return ((sF[m8(pl)][c3of1(c1)] >> (1 + d)) & 1) ? sXp[m4(pl)][1 + d][c1] : C(0);
    };

    // ---- staging of x (global -> registers -> LDS), two nodes per thread
    C rw[2][4];
    uint8_t rf[2];
    auto fetch = [&](int s) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int t = tid + e * SW_NT;
            rf[e] = 0;
            rw[e][0] = rw[e][1] = rw[e][2] = rw[e][3] = C(0);
            const int jj = j0 - 3 + t / S3_W, kk = k0 - 3 + t % S3_W;
            if (t < S3_N && s >= 0 && s <= g.nx && jj >= 0 && jj <= g.ny && kk >= 0 && kk <= g.nz) {
                const long q = (long)s * sI + (long)jj * sJ + kk;
                rf[e] = fl[q];
                rw[e][0] = Cvt<TS>::ld(x[q]);
                rw[e][1] = Cvt<TS>::ld(x[S + q]);
                rw[e][2] = Cvt<TS>::ld(x[2 * S + q]);
                rw[e][3] = Cvt<TS>::ld(x[3 * S + q]);
            }
        }
    };
    auto store = [&](int s) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int t = tid + e * SW_NT;
            if (t < S3_N) {
                sF[m8(s)][t] = rf[e];
#pragma unroll
                for (int f = 0; f < 4; f++) sX[m4(s)][f][t] = rw[e][f];
            }
        }
    };
    // E and G of x for plane s: E on rows 1 .. BJ+5, columns 1 .. BK+5; G on rows
    // 0 .. BJ+5, columns 0 .. BK+5 (what the first colour of the ring cells reads)
    auto edges_gdiv = [&](int s) {
        const int es = m2(s);
        for (int t = tid; t < (SB_J + 5) * (SB_K + 5); t += SW_NT) {
            const int c = (t / (SB_K + 5) + 1) * S3_W + (t % (SB_K + 5) + 1);
            sE[es][0][c] = Wx(s, 1, c) - Wx(s, 1, c - 1) - Wx(s, 2, c) + Wx(s, 2, c - S3_W);
            sE[es][1][c] = Wx(s, 2, c) - Wx(s - 1, 2, c) - Wx(s, 0, c) + Wx(s, 0, c - 1);
            sE[es][2][c] = Wx(s, 0, c) - Wx(s, 0, c - S3_W) - Wx(s, 1, c) + Wx(s - 1, 1, c);
        }
        const int gs = m3(s);
        for (int t = tid; t < (SB_J + 6) * (SB_K + 6); t += SW_NT) {
            const int c = (t / (SB_K + 6)) * S3_W + (t % (SB_K + 6));
            const C dv = Wx(s + 1, 0, c) - Wx(s, 0, c) + Wx(s, 1, c + S3_W) - Wx(s, 1, c) + Wx(s, 2, c + 1) - Wx(s, 2, c);
            sG[gs][c] = (sF[m8(s)][c] & 1) ? dv : C(0);
        }
    };
    // x' of plane s starts as x (the first colour then overwrites its DOFs)
    auto init_xp = [&](int s) {
        for (int t = tid; t < S1_N; t += SW_NT) {
            const int c3 = c3of1(t);
#pragma unroll
            for (int f = 0; f < 4; f++) sXp[m4(s)][f][t] = sX[m4(s)][f][c3];
        }
    };
    // E' and G' of x' for plane s on the tile region (as st_march.hip edges / gdiv)
    auto edges_gdiv_p = [&](int s) {
        const int es = m2(s);
        for (int t = tid; t < (SB_J + 1) * (SB_K + 1); t += SW_NT) {
            const int c = (t / (SB_K + 1) + 1) * S1_W + (t % (SB_K + 1) + 1);
            sEp[es][0][c] = Wp(s, 1, c) - Wp(s, 1, c - 1) - Wp(s, 2, c) + Wp(s, 2, c - S1_W);
            sEp[es][1][c] = Wp(s, 2, c) - Wp(s - 1, 2, c) - Wp(s, 0, c) + Wp(s, 0, c - 1);
            sEp[es][2][c] = Wp(s, 0, c) - Wp(s, 0, c - S1_W) - Wp(s, 1, c) + Wp(s - 1, 1, c);
        }
        for (int t = tid; t < (SB_J + 2) * (SB_K + 2); t += SW_NT) {
            const int c = (t / (SB_K + 2)) * S1_W + (t % (SB_K + 2));
            const C dv = Wp(s + 1, 0, c) - Wp(s, 0, c) + Wp(s, 1, c + S1_W) - Wp(s, 1, c) + Wp(s, 2, c + 1) - Wp(s, 2, c);
            sGp[es][c] = (sF[m8(s)][c3of1(c)] & 1) ? dv : C(0);
        }
    };

    // ---- first-colour cell of this thread on plane s (ring region), its b and alpha
    C cb[7], ca[6];
    auto c0_cell = [&](int s, int &jj, int &kk) -> bool {
        if (tid >= C0_N) return false;
        const int row = tid / C0_ROW, pr = tid % C0_ROW;
        jj = j0 - 2 + row;
        kk = k0 - 2 + 2 * pr + ((col0 + s + jj) & 1);  // k0 - 2 is even
        return s >= 0 && s < g.nx && jj >= 0 && jj < g.ny && kk >= 0 && kk < g.nz;
    };
    auto c0_load = [&](int s) {
        int jj, kk;
        if (!c0_cell(s, jj, kk)) return;
        const long q = (long)s * sI + (long)jj * sJ + kk;
        cb[0] = Cvt<TS>::ld(b[q]);
        cb[1] = Cvt<TS>::ld(b[S + q]);
        cb[2] = Cvt<TS>::ld(b[2 * S + q]);
        cb[3] = Cvt<TS>::ld(b[3 * S + q]);
        cb[4] = Cvt<TS>::ld(b[S + q + sI]);
        cb[5] = Cvt<TS>::ld(b[2 * S + q + sJ]);
        cb[6] = Cvt<TS>::ld(b[3 * S + q + 1]);
        ca[0] = Cvt<TS>::ld(al[q]);
        ca[1] = Cvt<TS>::ld(al[S + q]);
        ca[2] = Cvt<TS>::ld(al[2 * S + q]);
        ca[3] = Cvt<TS>::ld(al[q + sI]);
        ca[4] = Cvt<TS>::ld(al[S + q + sJ]);
        ca[5] = Cvt<TS>::ld(al[2 * S + q + 1]);
    };
    // the first colour on plane s: residuals from x (LDS), block solve, x' updated
    // where it lies inside the tile region
    auto c0_update = [&](int s) {
        int jj, kk;
        if (!c0_cell(s, jj, kk)) return;
        const int c = (jj - (j0 - 3)) * S3_W + (kk - (k0 - 3));
        const int ps = m4(s), pn = m4(s + 1), pm = m4(s - 1);
        const int fs = m8(s), fn = m8(s + 1), fm = m8(s - 1);
        const int es = m2(s), en = m2(s + 1);
        const int gm = m3(s - 1), gs = m3(s), gn = m3(s + 1);
        // operator rows at plane s (node c) / plane s+1, raw value xr for a fixed face
        // (the high face groups its last terms as the two-pass kernel does)
        auto yx = [&](int gA, int gB, int eA, int pA, int pB, int fB, int cc, C av, C xr, bool hi) -> C {
            C r = sG[gA][cc] - sG[gB][cc] + sE[eA][1][cc + 1] - sE[eA][1][cc] + sE[eA][2][cc] - sE[eA][2][cc + S3_W];
            const C w = (sF[fB][cc] & 2) ? sX[pB][1][cc] : C(0);
            if (hi) r += sX[pA][0][cc] - sX[pB][0][cc] + av * w;
            else r = r + sX[pA][0][cc] - sX[pB][0][cc] + av * w;
            return (sF[fB][cc] & 2) ? r : xr;
        };
        auto yy = [&](int cc, C av, C xr) -> C {
            C r = sG[gs][cc - S3_W] - sG[gs][cc] + sE[es][0][cc] - sE[es][0][cc + 1] + sE[en][2][cc] - sE[es][2][cc];
            r = r + sX[ps][0][cc - S3_W] - sX[ps][0][cc] + av * ((sF[fs][cc] & 4) ? sX[ps][2][cc] : C(0));
            return (sF[fs][cc] & 4) ? r : xr;
        };
        auto yz = [&](int cc, C av, C xr) -> C {
            C r = sG[gs][cc - 1] - sG[gs][cc] + sE[es][0][cc + S3_W] - sE[es][0][cc] + sE[es][1][cc] - sE[en][1][cc];
            r = r + sX[ps][0][cc - 1] - sX[ps][0][cc] + av * ((sF[fs][cc] & 8) ? sX[ps][3][cc] : C(0));
            return (sF[fs][cc] & 8) ? r : xr;
        };
        const int fb = sF[fs][c];
        const C mc = (fb & 1) ? C(1) : C(0);
        const C rp = cb[0] - ((fb & 1) ? sG[gs][c] : sX[ps][0][c]);
        C r0[3], r1[3], m0[3], m1[3], a0[3], a1[3], u0[3], u1[3];
        u0[0] = sX[ps][1][c];
        u0[1] = sX[ps][2][c];
        u0[2] = sX[ps][3][c];
        u1[0] = sX[pn][1][c];
        u1[1] = sX[ps][2][c + S3_W];
        u1[2] = sX[ps][3][c + 1];
        r0[0] = cb[1] - yx(gm, gs, es, pm, ps, fs, c, ca[0], u0[0], false);
        r0[1] = cb[2] - yy(c, ca[1], u0[1]);
        r0[2] = cb[3] - yz(c, ca[2], u0[2]);
        r1[0] = cb[4] - yx(gs, gn, en, ps, pn, fn, c, ca[3], u1[0], true);
        r1[1] = cb[5] - yy(c + S3_W, ca[4], u1[1]);
        r1[2] = cb[6] - yz(c + 1, ca[5], u1[2]);
        const int fl1[3] = {sF[fn][c], sF[fs][c + S3_W], sF[fs][c + 1]};
        const C clo[3] = {(C)(sF[fm][c] & 1), (C)(sF[fs][c - S3_W] & 1), (C)(sF[fs][c - 1] & 1)};
        const C chi[3] = {(C)(sF[fn][c] & 1), (C)(sF[fs][c + S3_W] & 1), (C)(sF[fs][c + 1] & 1)};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            m0[d] = ((fb >> (1 + d)) & 1) ? C(1) : C(0);
            m1[d] = ((fl1[d] >> (1 + d)) & 1) ? C(1) : C(0);
            a0[d] = m0[d] * (clo[d] + mc + C(4)) + ca[d] + (C(1) - m0[d]) + (C(1) - mc) * m0[d];
            a1[d] = m1[d] * (mc + chi[d] + C(4)) + ca[3 + d] + (C(1) - m1[d]) + (C(1) - mc) * m1[d];
        }
        const Blk o = block_solve(mc, rp, r0, r1, m0, m1, a0, a1, om);
        // x' updates of the DOFs that lie in the tile region (rows j0-1 .. j0+BJ+1,
        // columns k0-1 .. k0+BK+1)
        const int r1i = jj - (j0 - 1), q1i = kk - (k0 - 1);
        auto put = [&](int pl, int f, int rr, int qq, C v) {
            if (rr >= 0 && rr < S1_H && qq >= 0 && qq < S1_W) sXp[m4(pl)][f][rr * S1_W + qq] = v;
        };
        put(s, 0, r1i, q1i, sX[ps][0][c] + om * o.dp);
        put(s, 1, r1i, q1i, u0[0] + o.d0[0]);
        put(s, 2, r1i, q1i, u0[1] + o.d0[1]);
        put(s, 3, r1i, q1i, u0[2] + o.d0[2]);
        put(s + 1, 1, r1i, q1i, u1[0] + o.d1[0]);
        put(s, 2, r1i + 1, q1i, u1[1] + o.d1[1]);
        put(s, 3, r1i, q1i + 1, u1[2] + o.d1[2]);
    };

    // ---- second colour: the tile cell of this thread (threads 0 .. 255)
    const bool own = tid < SB_J * SB_K;
    const int tj = tid / SB_K, tk = tid % SB_K;
    const int j = j0 + tj, k = k0 + tk;
    const int cown = (tj + 1) * S1_W + (tk + 1);       // tile region index
    const int cown3 = (tj + 3) * S3_W + (tk + 3);      // staging region index (flags)
    C pb[7], pa[6];
    auto c1_load = [&](int i) {
        if (!own) return;
        const long q = (long)i * sI + (long)j * sJ + k;
        pb[0] = Cvt<TS>::ld(b[q]);
        pb[1] = Cvt<TS>::ld(b[S + q]);
        pb[2] = Cvt<TS>::ld(b[2 * S + q]);
        pb[3] = Cvt<TS>::ld(b[3 * S + q]);
        pb[4] = Cvt<TS>::ld(b[S + q + sI]);
        pb[5] = Cvt<TS>::ld(b[2 * S + q + sJ]);
        pb[6] = Cvt<TS>::ld(b[3 * S + q + 1]);
        pa[0] = Cvt<TS>::ld(al[q]);
        pa[1] = Cvt<TS>::ld(al[S + q]);
        pa[2] = Cvt<TS>::ld(al[2 * S + q]);
        pa[3] = Cvt<TS>::ld(al[q + sI]);
        pa[4] = Cvt<TS>::ld(al[S + q + sJ]);
        pa[5] = Cvt<TS>::ld(al[2 * S + q + 1]);
    };
    C gm1 = C(0), pm1 = C(0);
    int fm1 = 0;
    auto c1_update = [&](int i) {
        if (!own) return;
        const long q = (long)i * sI + (long)j * sJ + k;
        const int ps = m4(i), pn = m4(i + 1);
        const int fs = m8(i), fn = m8(i + 1);
        const int es = m2(i), en = m2(i + 1);
        const int c = cown, c3 = cown3;
        auto F1 = [&](int f8, int cc) -> int { return sF[f8][c3of1(cc)]; };
        const int fb = sF[fs][c3];
        const bool mine = ((i + j + k) & 1) == col1;
        if (!mine) {
            // a first-colour cell: p' and its domain-boundary faces come from x'
            y[q] = Cvt<TS>::st(sXp[ps][0][c]);
            if (i == 0) y[S + q] = Cvt<TS>::st(sXp[ps][1][c]);
            if (i == g.nx - 1) y[S + q + sI] = Cvt<TS>::st(sXp[pn][1][c]);
            if (j == 0) y[2 * S + q] = Cvt<TS>::st(sXp[ps][2][c]);
            if (j == g.ny - 1) y[2 * S + q + sJ] = Cvt<TS>::st(sXp[ps][2][c + S1_W]);
            if (k == 0) y[3 * S + q] = Cvt<TS>::st(sXp[ps][3][c]);
            if (k == g.nz - 1) y[3 * S + q + 1] = Cvt<TS>::st(sXp[ps][3][c + 1]);
            return;
        }
        auto yx = [&](int cc, C gA, C gB, int eA, C pA, C pB, int fB8, int pB4, C av, C xr, bool hi) -> C {
            C r = gA - gB + sEp[eA][1][cc + 1] - sEp[eA][1][cc] + sEp[eA][2][cc] - sEp[eA][2][cc + S1_W];
            const int fB = F1(fB8, cc);
            const C w = (fB & 2) ? sXp[pB4][1][cc] : C(0);
            if (hi) r += pA - pB + av * w;
            else r = r + pA - pB + av * w;
            return (fB & 2) ? r : xr;
        };
        auto yy = [&](int cc, C av, C xr) -> C {
            C r = sGp[es][cc - S1_W] - sGp[es][cc] + sEp[es][0][cc] - sEp[es][0][cc + 1] + sEp[en][2][cc] - sEp[es][2][cc];
            const int fB = F1(fs, cc);
            r = r + sXp[ps][0][cc - S1_W] - sXp[ps][0][cc] + av * ((fB & 4) ? sXp[ps][2][cc] : C(0));
            return (fB & 4) ? r : xr;
        };
        auto yz = [&](int cc, C av, C xr) -> C {
            C r = sGp[es][cc - 1] - sGp[es][cc] + sEp[es][0][cc + S1_W] - sEp[es][0][cc] + sEp[es][1][cc] - sEp[en][1][cc];
            const int fB = F1(fs, cc);
            r = r + sXp[ps][0][cc - 1] - sXp[ps][0][cc] + av * ((fB & 8) ? sXp[ps][3][cc] : C(0));
            return (fB & 8) ? r : xr;
        };
        const C mc = (fb & 1) ? C(1) : C(0);
        const C rp = pb[0] - ((fb & 1) ? sGp[es][c] : sXp[ps][0][c]);
        C r0[3], r1[3], m0[3], m1[3], a0[3], a1[3], u0[3], u1[3];
        u0[0] = sXp[ps][1][c];
        u0[1] = sXp[ps][2][c];
        u0[2] = sXp[ps][3][c];
        u1[0] = sXp[pn][1][c];
        u1[1] = sXp[ps][2][c + S1_W];
        u1[2] = sXp[ps][3][c + 1];
        r0[0] = pb[1] - yx(c, gm1, sGp[es][c], es, pm1, sXp[ps][0][c], fs, ps, pa[0], u0[0], false);
        r0[1] = pb[2] - yy(c, pa[1], u0[1]);
        r0[2] = pb[3] - yz(c, pa[2], u0[2]);
        r1[0] = pb[4] - yx(c, sGp[es][c], sGp[en][c], en, sXp[ps][0][c], sXp[pn][0][c], fn, pn, pa[3], u1[0], true);
        r1[1] = pb[5] - yy(c + S1_W, pa[4], u1[1]);
        r1[2] = pb[6] - yz(c + 1, pa[5], u1[2]);
        const int fl1[3] = {sF[fn][c3], sF[fs][c3 + S3_W], sF[fs][c3 + 1]};
        const C clo[3] = {(C)(fm1 & 1), (C)(sF[fs][c3 - S3_W] & 1), (C)(sF[fs][c3 - 1] & 1)};
        const C chi[3] = {(C)(sF[fn][c3] & 1), (C)(sF[fs][c3 + S3_W] & 1), (C)(sF[fs][c3 + 1] & 1)};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            m0[d] = ((fb >> (1 + d)) & 1) ? C(1) : C(0);
            m1[d] = ((fl1[d] >> (1 + d)) & 1) ? C(1) : C(0);
            a0[d] = m0[d] * (clo[d] + mc + C(4)) + pa[d] + (C(1) - m0[d]) + (C(1) - mc) * m0[d];
            a1[d] = m1[d] * (mc + chi[d] + C(4)) + pa[3 + d] + (C(1) - m1[d]) + (C(1) - mc) * m1[d];
        }
        const Blk o = block_solve(mc, rp, r0, r1, m0, m1, a0, a1, om);
        const long qh[3] = {q + sI, q + sJ, q + 1};
        y[q] = Cvt<TS>::st(sXp[ps][0][c] + om * o.dp);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            y[(long)(1 + d) * S + q] = Cvt<TS>::st(u0[d] + o.d0[d]);
            y[(long)(1 + d) * S + qh[d]] = Cvt<TS>::st(u1[d] + o.d1[d]);
        }
    };

    // ---- pipeline: first colour on planes ic0-2 .. ic1+1, second colour on ic0 .. ic1-1
    const int t0 = ic0 - 4;
    for (int s = t0 + 1; s <= t0 + 3; s++) {
        fetch(s);
        store(s);
    }
    fetch(t0 + 4);
    __syncthreads();
    edges_gdiv(t0 + 2);  // E(t0+2) and G(t0+2); G(t0+1) below
    {
        const int s = t0 + 1, gs = m3(s);
        for (int t = tid; t < (SB_J + 6) * (SB_K + 6); t += SW_NT) {
            const int c = (t / (SB_K + 6)) * S3_W + (t % (SB_K + 6));
            const C dv = Wx(s + 1, 0, c) - Wx(s, 0, c) + Wx(s, 1, c + S3_W) - Wx(s, 1, c) + Wx(s, 2, c + 1) - Wx(s, 2, c);
            sG[gs][c] = (sF[m8(s)][c] & 1) ? dv : C(0);
        }
    }
    init_xp(t0 + 2);
    init_xp(t0 + 3);
    for (int t = t0; t < ic1; t++) {
        // A: plane t+4 into LDS, plane t+5 into registers; global operands of both colours
        store(t + 4);
        fetch(t + 5);
        c0_load(t + 2);
        if (t >= ic0) c1_load(t);
        __syncthreads();
        // B: E, G of plane t+3; x' of plane t+3 starts as x
        edges_gdiv(t + 3);
        init_xp(t + 3);
        __syncthreads();
        // C: first colour on plane t+2
        c0_update(t + 2);
        __syncthreads();
        // D: E', G' of plane t+1 (x' of planes t .. t+2 are final)
        if (t >= ic0 - 2) edges_gdiv_p(t + 1);
        __syncthreads();
        // E: second colour on plane t
        if (t >= ic0) c1_update(t);
        if (t >= ic0 - 1 && own) {
            gm1 = sGp[m2(t)][cown];
            pm1 = sXp[m4(t)][0][cown];
            fm1 = sF[m8(t)][cown3];
        }
    }
}

template <typename TS>
void march_sweep(const Lvl &g, const uint8_t *fl, const TS *al, const TS *b, const TS *xin, TS *xout, int col0,
                 float omega, hipStream_t s)
{
    // tiles of 4 x 64 cells; chunks of at least 16 planes (the first colour is
    // recomputed on 4 extra planes per chunk), about 1024 workgroups at 256^3
    const int tiles = (g.nz / SB_K) * (g.ny / SB_J);
    int nch = (1024 + tiles - 1) / tiles;
    const int maxch = g.nx / 16;
    if (nch > maxch) nch = maxch;
    if (nch < 1) nch = 1;
    int chunk = (g.nx + nch - 1) / nch;
    nch = (g.nx + chunk - 1) / chunk;
    hipLaunchKernelGGL((k_sweep<TS>), dim3(g.nz / SB_K, g.ny / SB_J, nch), dim3(SW_NT), 0, s, g, fl, al, b, xin, xout,
                       chunk, col0, omega);
}

template void march_sweep<float>(const Lvl &, const uint8_t *, const float *, const float *, const float *, float *,
                                 int, float, hipStream_t);
template void march_sweep<bf16>(const Lvl &, const uint8_t *, const bf16 *, const bf16 *, const bf16 *, bf16 *, int,
                                float, hipStream_t);

}  // namespace st
