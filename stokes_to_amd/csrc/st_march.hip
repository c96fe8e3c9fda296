// 2.5D "marching" kernels for the fine 3D levels (gfx950).
//
// A workgroup of BJ x 64 threads owns a (j, k) tile of node columns and walks a
// chunk of i-planes.  For each plane it stages, once, in LDS:
//   w = m u (masked face velocities), p, flags          3 rolling planes
//   edge curls e = Cod1 w   (CX/grid3D.cu:469-568)       2 rolling planes
//   g = m_c D2 w            (CX/grid3D.cu:330-390)       2 rolling planes
// so every stencil term of ST/StokeFlowMapping3D.cu:176-306 is formed once per
// node and read from LDS by its consumers instead of being re-derived from
// global memory (what the simple gather kernels of st_kernels.hip do).
// The k direction (contiguous in memory) maps to the 64 lanes of a wave, so all
// global loads and stores are coalesced.
//
// Modes: APPLY y = A x;  RESID r = b - A x (+ FP64 partial sums of r^2);
//        SMOOTH one colour of the box-Vanka smoother (xin -> xout), with the
//        7 local residuals of each colour cell read from the staged planes.
#include "st_kernels.h"

namespace st {

namespace {
constexpr int BK = 64, LK = BK + 3;
enum { M_APPLY = 0, M_RESID = 1, M_SMOOTH = 2 };

__device__ __forceinline__ int slot3(int s) { return (s + 3) % 3; }
__device__ __forceinline__ int slot2(int s) { return s & 1; }

__device__ __forceinline__ double wsum64(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// 1/x and a/b of the block solve: IEEE division in FP64; in FP32 (the
// mixed-precision smoother) the hardware reciprocal, 1 ulp
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ double quo(double a, double b) { return a / b; }
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float quo(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
}  // namespace

// Global access at a node offset: with O32 (every field under 4 GiB) the offset is
// a 32-bit byte offset from the field's base pointer, which the loads and stores
// take in their SGPR-base + VGPR-offset form (no 64-bit address arithmetic per
// field); otherwise a 64-bit element index.
template <bool O32> struct Ix { typedef long T; };
template <> struct Ix<true> { typedef unsigned T; };
template <bool O32, typename T>
__device__ __forceinline__ T gld(const T *p, typename Ix<O32>::T i)
{
    if constexpr (O32)
        return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + i * (unsigned)sizeof(T));
    else
        return p[i];
}
template <bool O32, typename T>
__device__ __forceinline__ void gst(T *p, typename Ix<O32>::T i, T v)
{
    if constexpr (O32)
        *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + i * (unsigned)sizeof(T)) = v;
    else
        p[i] = v;
}

// TS: storage of alpha and b (and the compute type through Cvt<TS>::C); TX: storage
// of x (a lower-precision x is widened on load: the FP64 operator on an FP32 vector);
// BJ: tile rows; TY: storage of y (an FP64 residual stored in FP32); O32: 32-bit
// offsets.
//
// SMOOTH: half of the workgroup's waves (BJ x 32 lanes) take the BJ x 32 cells of the
// colour in the tile's plane, one per lane, the other half the other cells
// (copy-through of p and of the domain-boundary faces).  A lane per colour cell, not
// per node, so no wave runs the block solve with half of its lanes idle.
template <typename TS, int MODE, typename TX = TS, int BJ = 4, typename TY = TS, bool O32 = false>
__global__ void __launch_bounds__(BJ * 64) k_march(Lvl g, const uint8_t *__restrict__ fl, const TS *__restrict__ al,
                                                   const TS *__restrict__ b, const TX *__restrict__ x,
                                                   TY *__restrict__ y, double *part, int chunk, int color,
                                                   float omega)
{
    typedef typename Cvt<TS>::C C;
    typedef typename Cvt<TY>::C CY;
    typedef typename Ix<O32>::T I;
    constexpr int LJ = BJ + 3, NL = LJ * LK, NT = BJ * BK;
    static_assert(NL <= 2 * NT, "a plane's staged nodes are fetched two per thread");
    constexpr bool SM = (MODE == M_SMOOTH);
    __shared__ C sW[3][3][NL];
    __shared__ C sP[3][NL];
    __shared__ uint8_t sF[3][NL];
    __shared__ C sE[2][3][NL];
    __shared__ C sG[2][NL];
    __shared__ double sred[NT / 64];
    const int tid = threadIdx.x;
    const int tk = tid & (BK - 1), tj = tid / BK;
    // XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs, so map
    // hardware id h to logical tile (h % 8) * (nb / 8) + h / 8 -- neighbouring tiles
    // (which share halo planes) then run on the same XCD and meet in its L2
    const int nbx = gridDim.x, nby = gridDim.y, nbz = gridDim.z;
    const int nbt = nbx * nby * nbz;
    const int hw = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const int lid = (nbt % 8 == 0) ? (hw % 8) * (nbt / 8) + hw / 8 : hw;
    const int bx = lid % nbx, by = (lid / nbx) % nby, bz = lid / (nbx * nby);
    const int k0 = bx * BK, j0 = by * BJ;
    const int j = j0 + tj, k = k0 + tk;
    const int ilast = SM ? g.nx : g.nx + 1;  // output planes [0, ilast)
    const int ic0 = bz * chunk;
    int ic1 = (bz == nbz - 1) ? ilast : ic0 + chunk;
    if (ic1 > ilast) ic1 = ilast;
    const I sI = (I)g.Py * (I)g.Pz, sJ = (I)g.Pz;
    const int cown = (tj + 1) * LK + (tk + 1);
    // field base pointers (uniform)
    const long S = g.S;
    const TX *xf[4] = {x, x + S, x + 2 * S, x + 3 * S};
    const TS *af[3] = {al, al + S, al + 2 * S};
    const TS *bf[4] = {b, b, b, b};
    if (MODE != M_APPLY) {
        bf[1] = b + S;
        bf[2] = b + 2 * S;
        bf[3] = b + 3 * S;
    }
    TY *yf[4] = {y, y, y, y};  // y is null for a residual norm without the residual
    if (y) {
        yf[1] = y + S;
        yf[2] = y + 2 * S;
        yf[3] = y + 3 * S;
    }
    auto ldx = [&](int f, I q) -> C { return (C)Cvt<TX>::ld(gld<O32>(xf[f], q)); };
    auto lda = [&](int f, I q) -> C { return (C)Cvt<TS>::ld(gld<O32>(af[f], q)); };
    auto ldb = [&](int f, I q) -> C { return (C)Cvt<TS>::ld(gld<O32>(bf[f], q)); };
    auto sty = [&](int f, I q, C v) { gst<O32>(yf[f], q, Cvt<TY>::st((CY)v)); };

    // plane staging: global -> registers (fetch, issued one plane ahead) and
    // registers -> LDS (store); each thread owns LDS entries tid and tid + NT.  The
    // loads are unconditional (an entry outside the level or the tile reads node 0 and
    // is zeroed when stored).
    C rw[2][4];
    int rf[2];
    bool rok[2];
    auto fetch = [&](int s) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int t = tid + e * NT;
            const int jj = j0 - 1 + t / LK, kk = k0 - 1 + t % LK;
            rok[e] = t < NL && s >= 0 && s <= g.nx && jj >= 0 && jj <= g.ny && kk >= 0 && kk <= g.nz;
            const I q = rok[e] ? (I)s * sI + (I)jj * sJ + (I)kk : (I)0;
            rf[e] = gld<O32>(fl, q);
#pragma unroll
            for (int f = 0; f < 4; f++) rw[e][f] = ldx(f, q);
        }
    };
    auto store = [&](int s) {
        const int sl = slot3(s);
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int t = tid + e * NT;
            if (t < NL) {
                const int f = rok[e] ? rf[e] : 0;
                sP[sl][t] = rok[e] ? rw[e][0] : C(0);
                sW[sl][0][t] = (f & 2) ? rw[e][1] : C(0);
                sW[sl][1][t] = (f & 4) ? rw[e][2] : C(0);
                sW[sl][2][t] = (f & 8) ? rw[e][3] : C(0);
                sF[sl][t] = (uint8_t)f;
            }
        }
    };
    auto load_plane = [&](int s) {
        fetch(s);
        store(s);
    };
    // edge curls of plane s on [j0, j0+BJ] x [k0, k0+BK]
    auto edges = [&](int s) {
        const int ws = slot3(s), wm = slot3(s - 1), es = slot2(s);
        for (int t = tid; t < (BJ + 1) * (BK + 1); t += NT) {
            const int c = (t / (BK + 1) + 1) * LK + (t % (BK + 1) + 1);
            sE[es][0][c] = sW[ws][1][c] - sW[ws][1][c - 1] - sW[ws][2][c] + sW[ws][2][c - LK];
            sE[es][1][c] = sW[ws][2][c] - sW[wm][2][c] - sW[ws][0][c] + sW[ws][0][c - 1];
            sE[es][2][c] = sW[ws][0][c] - sW[ws][0][c - LK] - sW[ws][1][c] + sW[wm][1][c];
        }
    };
    // g = m_c div(w) of plane s on [j0-1, j0+BJ] x [k0-1, k0+BK]
    auto gdiv = [&](int s) {
        const int ws = slot3(s), wn = slot3(s + 1), gs = slot2(s);
        for (int t = tid; t < (BJ + 2) * (BK + 2); t += NT) {
            const int c = (t / (BK + 2)) * LK + (t % (BK + 2));
            const C dv = sW[wn][0][c] - sW[ws][0][c] + sW[ws][1][c + LK] - sW[ws][1][c] + sW[ws][2][c + 1] -
                         sW[ws][2][c];
            sG[gs][c] = (sF[ws][c] & 1) ? dv : C(0);
        }
    };

    // SMOOTH lane roles: the waves alternate between the colour cells and the copy
    // role plane by plane (even waves take the colour cells on even planes, odd waves
    // on odd planes), so the block solves of a tile spread over all four SIMDs.  A lane
    // keeps its row sj of the tile and its pair of columns k0 + 2 sm, k0 + 2 sm + 1 (LDS
    // index cE of the even one) in both roles.
    static_assert(!SM || BJ % 2 == 0, "the smoother pairs the waves of a workgroup");
    const int wv = tid >> 6;
    const int hc = ((wv >> 1) << 6) | (tid & 63);
    const int sj = hc >> 5, sm = hc & 31;
    const int cE = (sj + 1) * LK + 2 * sm + 1;

    // ---- preamble for the chunk
    load_plane(ic0 - 1);
    load_plane(ic0);
    load_plane(ic0 + 1);
    __syncthreads();
    edges(ic0);
    gdiv(ic0);
    // g, p and flags of the plane below (i - 1) at a column: the low x face's row
    // needs them, and their slots are reused before the plane is processed.  Kept for
    // the own column (APPLY, RESID) or for both columns of the lane's pair (SMOOTH).
    auto below = [&](int c, C &gv, C &pv, int &fv) {
        const int a = slot3(ic0 - 1), bq = slot3(ic0);
        const C dv = sW[bq][0][c] - sW[a][0][c] + sW[a][1][c + LK] - sW[a][1][c] + sW[a][2][c + 1] - sW[a][2][c];
        fv = sF[a][c];
        gv = (fv & 1) ? dv : C(0);
        pv = sP[a][c];
    };
    C gm1[2], pm1[2];
    int fm1[2];
    if (SM) {
        below(cE, gm1[0], pm1[0], fm1[0]);
        below(cE + 1, gm1[1], pm1[1], fm1[1]);
    } else {
        below(cown, gm1[0], pm1[0], fm1[0]);
    }
    fetch(ic0 + 2);
    __syncthreads();

    // global operands of the node q (LDS index c) and of its three high faces, loaded
    // before the two LDS phases of the iteration so that their latency overlaps them:
    // pb = b (p, x, y, z rows; then the x/y/z rows of the high faces), pa = alpha, px =
    // the raw x of the own faces then of the high faces where the face is fixed (a free
    // face's value is the staged w; 0 elsewhere).  Planes i and i+1 are staged.
    C pb[7], px[6], pa[6];
    auto preload = [&](int i, I q, int c) {
        const int so = slot3(i), sn = slot3(i + 1);
        const int fo = sF[so][c];
        auto raw = [&](int fixed, int f, I qq) -> C { return fixed ? ldx(f, qq) : C(0); };
        if (MODE != M_APPLY) {
#pragma unroll
            for (int f = 0; f < 4; f++) pb[f] = ldb(f, q);
        }
#pragma unroll
        for (int f = 0; f < 3; f++) pa[f] = lda(f, q);
        if (SM) {  // a cell: i < nx, j < ny, k < nz, so the high faces exist
            pb[4] = ldb(1, q + sI);
            pb[5] = ldb(2, q + sJ);
            pb[6] = ldb(3, q + 1);
            pa[3] = lda(0, q + sI);
            pa[4] = lda(1, q + sJ);
            pa[5] = lda(2, q + 1);
        }
        px[0] = raw(!(fo & 2), 1, q);
        px[1] = raw(!(fo & 4), 2, q);
        px[2] = raw(!(fo & 8), 3, q);
        if (SM) {
            px[3] = raw(!(sF[sn][c] & 2), 1, q + sI);
            px[4] = raw(!(sF[so][c + LK] & 4), 2, q + sJ);
            px[5] = raw(!(sF[so][c + 1] & 8), 3, q + 1);
        }
    };

    double acc = 0.0;
    for (int i = ic0; i < ic1; i++) {
        store(i + 2);
        if (i + 1 < ic1) fetch(i + 3);
        // the node (or cell) of this lane in plane i
        int par = 0, c = cown, jn = j, kn = k;
        const bool cw = ((wv + i) & 1) == 0;
        if (SM) {
            const int pc = (color + i + j0 + sj) & 1;  // parity of the colour's columns in this row
            par = cw ? pc : 1 - pc;
            c = cE + par;
            jn = j0 + sj;
            kn = k0 + 2 * sm + par;
        }
        const I q = (I)i * sI + (I)jn * sJ + (I)kn;
        if (!SM || cw) preload(i, q, c);
        __syncthreads();
        edges(i + 1);
        gdiv(i + 1);
        __syncthreads();
        const int ps = slot3(i), pn = slot3(i + 1), es = slot2(i), en = slot2(i + 1), gs = slot2(i), gn = slot2(i + 1);
        const C gb = gm1[par], pbl = pm1[par];
        const int fbl = fm1[par];

        // operator row values at plane i, LDS position c; xr / av are the raw x and
        // the alpha of that face (used for a fixed face / the Brinkman term)
        auto yp = [&](int c) -> C { return (sF[ps][c] & 1) ? sG[gs][c] : sP[ps][c]; };
        auto yx = [&](int c, C gim1, C pim1, C xr, C av) -> C {
            C r = gim1 - sG[gs][c] + sE[es][1][c + 1] - sE[es][1][c] + sE[es][2][c] - sE[es][2][c + LK];
            r = r + pim1 - sP[ps][c] + av * sW[ps][0][c];
            return (sF[ps][c] & 2) ? r : xr;
        };
        auto yy = [&](int c, C xr, C av) -> C {
            C r = sG[gs][c - LK] - sG[gs][c] + sE[es][0][c] - sE[es][0][c + 1] + sE[en][2][c] - sE[es][2][c];
            r = r + sP[ps][c - LK] - sP[ps][c] + av * sW[ps][1][c];
            return (sF[ps][c] & 4) ? r : xr;
        };
        auto yz = [&](int c, C xr, C av) -> C {
            C r = sG[gs][c - 1] - sG[gs][c] + sE[es][0][c + LK] - sE[es][0][c] + sE[es][1][c] - sE[en][1][c];
            r = r + sP[ps][c - 1] - sP[ps][c] + av * sW[ps][2][c];
            return (sF[ps][c] & 8) ? r : xr;
        };

        if (!SM) {
            auto emit = [&](int c, I qq, C gim1, C pim1, const C *bb, const C *xr, const C *av) {
                C v0 = yp(c), v1 = yx(c, gim1, pim1, xr[0], av[0]), v2 = yy(c, xr[1], av[1]), v3 = yz(c, xr[2], av[2]);
                if (MODE == M_RESID) {
                    v0 = bb[0] - v0;
                    v1 = bb[1] - v1;
                    v2 = bb[2] - v2;
                    v3 = bb[3] - v3;
                    acc += (double)v0 * v0 + (double)v1 * v1 + (double)v2 * v2 + (double)v3 * v3;
                }
                if (MODE != M_RESID || y) {  // a residual without y: its norm only
                    sty(0, qq, v0);
                    sty(1, qq, v1);
                    sty(2, qq, v2);
                    sty(3, qq, v3);
                }
            };
            emit(cown, q, gb, pbl, pb, px, pa);
            // the padded node column k = nz and row j = ny belong to the last tiles
            const bool lastk = (k == g.nz - 1), lastj = (j == g.ny - 1);
            auto emit_edge = [&](int c, I qq) {
                C bb[4], xr[3], av[3];
#pragma unroll
                for (int f = 0; f < 4; f++) bb[f] = (MODE == M_RESID) ? ldb(f, qq) : C(0);
#pragma unroll
                for (int f = 0; f < 3; f++) {
                    xr[f] = ldx(f + 1, qq);
                    av[f] = lda(f, qq);
                }
                emit(c, qq, C(0), C(0), bb, xr, av);
            };
            if (lastk) emit_edge(cown + 1, q + 1);
            if (lastj) emit_edge(cown + LK, q + sJ);
            if (lastk && lastj) emit_edge(cown + LK + 1, q + sJ + 1);
            // the row padding k in (nz, Pz) is written too (zeros), so that y covers
            // every entry of the vector: callers reuse buffers across precisions
            if (y && k0 + BK == g.nz && g.nz + 1 + tk < g.Pz) {
                const I qp = (I)i * sI + (I)j * sJ + (I)(g.nz + 1 + tk);
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    sty(f, qp, C(0));
                    if (lastj) sty(f, qp + sJ, C(0));
                }
            }
        } else if (!cw) {
            // ---- copy-through of the cells of the other colour: p, and the
            // domain-boundary faces (no cell of the colour owns them)
            const int fb = sF[ps][c];
            auto cp = [&](int f, I qq, int free, C w) { sty(f, qq, free ? w : ldx(f, qq)); };
            sty(0, q, sP[ps][c]);
            if (i == 0) cp(1, q, fb & 2, sW[ps][0][c]);
            if (i == g.nx - 1) cp(1, q + sI, sF[pn][c] & 2, sW[pn][0][c]);
            if (jn == 0) cp(2, q, fb & 4, sW[ps][1][c]);
            if (jn == g.ny - 1) cp(2, q + sJ, sF[ps][c + LK] & 4, sW[ps][1][c + LK]);
            if (kn == 0) cp(3, q, fb & 8, sW[ps][2][c]);
            if (kn == g.nz - 1) cp(3, q + 1, sF[ps][c + 1] & 8, sW[ps][2][c + 1]);
        } else {
            // ---- block solve of a colour cell
            const int fb = sF[ps][c];
            const C mc = (fb & 1) ? C(1) : C(0);
            const I qh[3] = {q + sI, q + sJ, q + 1};
            // residuals at the 7 DOFs
            const C rp = pb[0] - yp(c);
            C r0[3], r1[3], u0[3], u1[3], m0[3], m1[3], a0[3], a1[3];
            // low faces (node i,j,k)
            r0[0] = pb[1] - yx(c, gb, pbl, px[0], pa[0]);
            r0[1] = pb[2] - yy(c, px[1], pa[1]);
            r0[2] = pb[3] - yz(c, px[2], pa[2]);
            // high x face: node (i+1, j, k), from plane i+1 (slots pn, en, gn)
            {
                const int f1 = sF[pn][c];
                C v = sG[gs][c] - sG[gn][c] + sE[en][1][c + 1] - sE[en][1][c] + sE[en][2][c] - sE[en][2][c + LK];
                v += sP[ps][c] - sP[pn][c] + pa[3] * sW[pn][0][c];
                r1[0] = pb[4] - ((f1 & 2) ? v : px[3]);
            }
            r1[1] = pb[5] - yy(c + LK, px[4], pa[4]);
            r1[2] = pb[6] - yz(c + 1, px[5], pa[5]);
            // masks, raw values, diagonals
            const int fl0[3] = {fb, fb, fb};
            const int fl1[3] = {sF[pn][c], sF[ps][c + LK], sF[ps][c + 1]};
            const C clo[3] = {(C)(fbl & 1), (C)(sF[ps][c - LK] & 1), (C)(sF[ps][c - 1] & 1)};
            const C chi[3] = {(C)(sF[pn][c] & 1), (C)(sF[ps][c + LK] & 1), (C)(sF[ps][c + 1] & 1)};
            const int cw1[3] = {c, c + LK, c + 1};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                m0[d] = ((fl0[d] >> (1 + d)) & 1) ? C(1) : C(0);
                m1[d] = ((fl1[d] >> (1 + d)) & 1) ? C(1) : C(0);
                u0[d] = (m0[d] != C(0)) ? sW[ps][d][c] : px[d];
                const int slot1 = (d == 0) ? pn : ps;
                u1[d] = (m1[d] != C(0)) ? sW[slot1][d][cw1[d]] : px[3 + d];
                const C al0 = pa[d], al1 = pa[3 + d];
                a0[d] = m0[d] * (clo[d] + mc + C(4)) + al0 + (C(1) - m0[d]) + (C(1) - mc) * m0[d];
                a1[d] = m1[d] * (mc + chi[d] + C(4)) + al1 + (C(1) - m1[d]) + (C(1) - mc) * m1[d];
            }
            // block solve of st_kernels.hip smooth_cell: Schur complement in a free
            // cell, Sherman-Morrison (K = E - b b^T) in a fixed cell
            C beta = C(0), tb = C(0);
            C of[3], id[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                of[d] = -m0[d] * m1[d];
                id[d] = rcp(a0[d] * a1[d] - of[d] * of[d]);
                const C kb0 = id[d] * (a1[d] * (-m0[d]) - of[d] * m1[d]);
                const C kb1 = id[d] * (-of[d] * (-m0[d]) + a0[d] * m1[d]);
                const C kr0 = id[d] * (a1[d] * r0[d] - of[d] * r1[d]);
                const C kr1 = id[d] * (-of[d] * r0[d] + a0[d] * r1[d]);
                beta += (-m0[d]) * kb0 + m1[d] * kb1;
                tb += (-m0[d]) * kr0 + m1[d] * kr1;
            }
            const C Sc = (mc != C(0)) ? -beta : C(1);
            const C dp = (Sc != C(0)) ? quo(rp - mc * tb, Sc) : C(0);
            const C gm = (mc != C(0)) ? dp : quo(dp - tb, C(1) - beta);
            const C om = (C)omega;
            sty(0, q, sP[ps][c] + om * dp);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const C e0 = r0[d] + m0[d] * gm, e1 = r1[d] - m1[d] * gm;
                sty(1 + d, q, u0[d] + om * id[d] * (a1[d] * e0 - of[d] * e1));
                sty(1 + d, qh[d], u1[d] + om * id[d] * (-of[d] * e0 + a0[d] * e1));
            }
        }
        if (SM) {
            gm1[0] = sG[gs][cE];
            gm1[1] = sG[gs][cE + 1];
            pm1[0] = sP[ps][cE];
            pm1[1] = sP[ps][cE + 1];
            fm1[0] = sF[ps][cE];
            fm1[1] = sF[ps][cE + 1];
        } else {
            gm1[0] = sG[gs][cown];
            pm1[0] = sP[ps][cown];
            fm1[0] = sF[ps][cown];
        }
        __syncthreads();
    }
    if (MODE == M_RESID && part) {
        double v = wsum64(acc);
        if ((tid & 63) == 0) sred[tid >> 6] = v;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < NT / 64; w++) s += sred[w];
            part[hw] = s;
        }
    }
}

bool march_ok(const Lvl &g)
{
    return g.tiled && g.dim == 3 && g.nz >= 64 && g.nz % BK == 0 && g.ny % 4 == 0 && g.nx >= 8;
}

// grid of BJ-row tiles x i-chunks: at least `target` workgroups, at least 8 planes
// per chunk
static dim3 march_grid(const Lvl &g, int bj, int target, int &chunk)
{
    const int tiles = (g.nz / BK) * (g.ny / bj);
    int nch = (target + tiles - 1) / tiles;
    int maxch = g.nx / 8;
    if (nch > maxch) nch = maxch;
    if (nch < 1) nch = 1;
    chunk = (g.nx + nch - 1) / nch;
    nch = (g.nx + chunk - 1) / chunk;
    return dim3(g.nz / BK, g.ny / bj, nch);
}

unsigned march_blocks(const Lvl &g)
{
    int ch;
    dim3 d = march_grid(g, 4, 2048, ch);
    return d.x * d.y * d.z;
}

// every field of the level addressable with 32-bit byte offsets (the largest offset
// used is a high face one plane up: S + sI nodes, at most 8 bytes each)
static bool off32(const Lvl &g) { return (unsigned long)(g.S + (long)g.Py * g.Pz) * 8UL < (1UL << 32); }

template <typename TS, int MODE, typename TX, typename TY>
static void launch_march(const Lvl &g, const uint8_t *fl, const TS *al, const TS *b, const TX *x, TY *y, double *part,
                         int color, float omega, hipStream_t s)
{
    int ch;
    dim3 grid = march_grid(g, 4, 2048, ch);
    if (off32(g))
        hipLaunchKernelGGL((k_march<TS, MODE, TX, 4, TY, true>), grid, dim3(256), 0, s, g, fl, al, b, x, y, part, ch,
                           color, omega);
    else
        hipLaunchKernelGGL((k_march<TS, MODE, TX, 4, TY, false>), grid, dim3(256), 0, s, g, fl, al, b, x, y, part, ch,
                           color, omega);
}

template <typename TS>
void march_apply(const Lvl &g, const uint8_t *fl, const TS *al, const TS *x, TS *y, hipStream_t s)
{
    launch_march<TS, M_APPLY>(g, fl, al, (const TS *)nullptr, x, y, (double *)nullptr, 0, 0.f, s);
}
template <typename TS>
void march_residual(const Lvl &g, const uint8_t *fl, const TS *al, const TS *b, const TS *x, TS *r, double *part,
                    hipStream_t s)
{
    launch_march<TS, M_RESID>(g, fl, al, b, x, r, part, 0, 0.f, s);
}
template <typename TS>
void march_smooth(const Lvl &g, const uint8_t *fl, const TS *al, const TS *b, const TS *xin, TS *xout, int color,
                  float omega, hipStream_t s)
{
    launch_march<TS, M_SMOOTH>(g, fl, al, b, xin, xout, (double *)nullptr, color, omega, s);
}

void march_residual_f32(const Lvl &g, const uint8_t *fl, const double *al, const double *b, const double *x, float *r,
                        double *part, hipStream_t s)
{
    launch_march<double, M_RESID>(g, fl, al, b, x, r, part, 0, 0.f, s);
}

void march_apply_mixed(const Lvl &g, const uint8_t *fl, const double *al, const float *x, double *y, hipStream_t s)
{
    launch_march<double, M_APPLY>(g, fl, al, (const double *)nullptr, x, y, (double *)nullptr, 0, 0.f, s);
}

#define INST_M(T)                                                                                              \
    template void march_apply<T>(const Lvl &, const uint8_t *, const T *, const T *, T *, hipStream_t);       \
    template void march_residual<T>(const Lvl &, const uint8_t *, const T *, const T *, const T *, T *,       \
                                    double *, hipStream_t);                                                   \
    template void march_smooth<T>(const Lvl &, const uint8_t *, const T *, const T *, const T *, T *, int,    \
                                  float, hipStream_t);
INST_M(double)
INST_M(float)
INST_M(bf16)

}  // namespace st
