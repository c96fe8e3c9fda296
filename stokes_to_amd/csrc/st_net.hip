// Learned initial guess: a 3-level 3D convolutional U-Net (16/32/64 channels,
// 3x3x3 kernels) with BF16 weights and activations and FP32 accumulation,
// trained and evaluated on the device by the kernels below.  All convolutions
// are implicit GEMMs on the CDNA matrix cores (v_mfma_f32_16x16x16_bf16).
// There is no counterpart in the reference (its GMRES starts from zero,
// CX/GMRES.cpp:29-40).
//
// The network runs on multigrid level lc (n/4 cells per axis): inputs are the
// block-averaged filtered density plus static boundary channels, outputs the
// coarse state (p, u_x, u_y, u_z at the level's nodes).  In the optimisation
// loop the prediction enters as a coarse-grid correction of the previous
// solution, x0 = x_prev + P P (net(rho_new) - net(rho_prev)).
#include "st_net.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace st {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct alignas(8) bf4 {
    uint16_t v[4];
};

__device__ __forceinline__ s4 as_s4(const bf4 &x)
{
    s4 r;
    r[0] = (short)x.v[0];
    r[1] = (short)x.v[1];
    r[2] = (short)x.v[2];
    r[3] = (short)x.v[3];
    return r;
}

__device__ __forceinline__ f4 mfma16(const s4 &a, const s4 &b, const f4 &c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void tap_off(int t, int ntap, int &a, int &b, int &c)
{
    if (ntap == 1) { a = b = c = 0; return; }
    a = t / 9 - 1;
    b = (t / 3) % 3 - 1;
    c = t % 3 - 1;
}

// ------------------------------------------------------------- convolution

// out[v][co] (+)= bias[co] + sum_t sum_ci in[v + o_t][ci] W[t][ci][co]
// D = A B with A = W^T (16 out channels x 16 in channels per tap/block, packed in
// the MFMA A-operand lane order) and B = 16 input channels x 16 voxels; each wave
// produces 16 out channels x 16 voxels.  OUTM: 0 bf16 store, 1 f32 store,
// 2 f32 accumulate.
template <int NTAP, int OUTM, bool RELU>
__global__ void __launch_bounds__(256) k_conv(int R, const bf16 *__restrict__ in, int csi, int cio, int ncib,
                                              const bf4 *__restrict__ Wp, const float *__restrict__ bias, void *out,
                                              int cso, int coo, int ncob)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long nv = (long)R * R * R;
    const long v0 = ((long)blockIdx.x * 4 + w) * 16;
    const int ob = blockIdx.y;
    const long v = v0 + (lane & 15);
    const int kq = lane >> 4;
    const int vi = (int)(v / ((long)R * R)), vj = (int)((v / R) % R), vk = (int)(v % R);
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < NTAP; t++) {
        int a, b, c;
        tap_off(t, NTAP, a, b, c);
        const int ii = vi + a, jj = vj + b, kk = vk + c;
        const bool ok = v < nv && ii >= 0 && jj >= 0 && kk >= 0 && ii < R && jj < R && kk < R;
        const long nb = (((long)ii * R + jj) * R + kk);
        for (int cb = 0; cb < ncib; cb++) {
            bf4 xb = {{0, 0, 0, 0}};
            if (ok) xb = *(const bf4 *)(in + nb * csi + cio + cb * 16 + kq * 4);
            const bf4 wa = Wp[(((long)t * ncib + cb) * ncob + ob) * 64 + lane];
            acc = mfma16(as_s4(wa), as_s4(xb), acc);
        }
    }
    if (v >= nv) return;
    const int co = ob * 16 + kq * 4;
    if (OUTM == 0) {
        bf4 o;
        for (int q = 0; q < 4; q++) {
            float s = acc[q] + (bias ? bias[co + q] : 0.f);
            if (RELU && s < 0.f) s = 0.f;
            o.v[q] = to_bf16(s).v;
        }
        *(bf4 *)((bf16 *)out + v * cso + coo + co) = o;
    } else {
        float *op = (float *)out + v * cso + coo + co;
        for (int q = 0; q < 4; q++) {
            float s = acc[q] + (bias ? bias[co + q] : 0.f);
            if (RELU && s < 0.f) s = 0.f;
            if (OUTM == 2) op[q] += s;
            else op[q] = s;
        }
    }
}

// weight gradient partials: P[s][t][ci][co] = sum_{v in chunk s} x[v + o_t][ci] gz[v][co]
// one wave per (tap, ci block, co block, chunk); 16-voxel tiles are transposed
// through LDS into the MFMA operand order (K = voxels).
template <int NTAP>
__global__ void __launch_bounds__(64) k_conv_wgrad(int R, const bf16 *__restrict__ x, int csx, int cix, int ncib,
                                                   const bf16 *__restrict__ gz, int ncob, long chunk,
                                                   float *__restrict__ part)
{
    __shared__ uint16_t xs[16][17], gs[16][17];
    const int lane = threadIdx.x;
    const int id = blockIdx.x;
    const int ob = id % ncob, cb = (id / ncob) % ncib, t = id / (ncob * ncib);
    const int sidx = blockIdx.y;
    const long nv = (long)R * R * R;
    const int Cout = ncob * 16, Cin = ncib * 16;
    int a, b, c;
    tap_off(t, NTAP, a, b, c);
    const int kq = lane >> 4, r16 = lane & 15;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    const long vs = (long)sidx * chunk, ve = (vs + chunk < nv) ? vs + chunk : nv;
    for (long v0 = vs; v0 < ve; v0 += 16) {
        const long v = v0 + r16;
        bf4 xb = {{0, 0, 0, 0}}, gb = {{0, 0, 0, 0}};
        if (v < ve) {
            const int vi = (int)(v / ((long)R * R)), vj = (int)((v / R) % R), vk = (int)(v % R);
            const int ii = vi + a, jj = vj + b, kk = vk + c;
            if (ii >= 0 && jj >= 0 && kk >= 0 && ii < R && jj < R && kk < R)
                xb = *(const bf4 *)(x + (((long)ii * R + jj) * R + kk) * csx + cix + cb * 16 + kq * 4);
            gb = *(const bf4 *)(gz + v * Cout + ob * 16 + kq * 4);
        }
        for (int q = 0; q < 4; q++) {
            xs[r16][kq * 4 + q] = xb.v[q];
            gs[r16][kq * 4 + q] = gb.v[q];
        }
        __syncthreads();
        bf4 A, Bm;
        for (int q = 0; q < 4; q++) {
            A.v[q] = xs[kq * 4 + q][r16];   // A[m = ci][k = voxel]
            Bm.v[q] = gs[kq * 4 + q][r16];  // B[k = voxel][n = co]
        }
        acc = mfma16(as_s4(A), as_s4(Bm), acc);
        __syncthreads();
    }
    float *o = part + (((long)sidx * NTAP + t) * Cin + cb * 16 + kq * 4) * Cout + ob * 16 + r16;
    for (int q = 0; q < 4; q++) o[(long)q * Cout] = acc[q];
}

// Weight gradient from staged operands (R % 4 == 0).  The MFMA wants four consecutive
// voxels of one channel per lane (K = voxels, M = 16 channels); the activations and
// gradients are channel-last, so both are staged once per layer with the 16 channels
// of a block interleaved per quad of voxels:
//   xs[cb][s][(i+1) P + (j+1)][k/4][ci][k%4] = x(i, j, k + s - 1) of channel cb*16 + ci,
//     zero outside the grid (P = R + 2: a zero row and plane all round; s: the three
//     k shifts)
//   gs[ob][v/4][co%16][v%4] = gz(v, ob*16 + co%16)
// so the operand of lane l = 16 qd + m for the 16 voxels from v0 (one k row) is the
// 8 bytes at 64 (v0/4 + qd) + 4 m: a wave's load is 512 contiguous bytes.  One wave
// accumulates all taps of one (ci block, co block) over its voxel groups and writes
// its partial.
// one (i+1, j+1) row of xs for one ci block per workgroup (grid: P*P rows x ncib): the
// row's R voxels x 16 channels are read channel-contiguous into LDS with a zero at
// each end, then written as the three k-shifted interleaved rows; pad rows are zero
__global__ void __launch_bounds__(256) k_wg_stage_x(int R, const bf16 *__restrict__ x, int csx, int cix, int ncib,
                                                    bf16 *__restrict__ xs)
{
    __shared__ uint16_t sm[16][130];  // R <= 128
    const int P = R + 2;
    const int row = blockIdx.x, cb = blockIdx.y;
    const int i = row / P - 1, j = row % P - 1;
    const bool pad = i < 0 || i >= R || j < 0 || j >= R;
    if (!pad) {
        for (int t = threadIdx.x; t < 16 * R; t += blockDim.x) {
            const int ci = t & 15, k = t >> 4;
            sm[ci][k + 1] = x[(((long)i * R + j) * R + k) * csx + cix + cb * 16 + ci].v;
        }
        if (threadIdx.x < 16) {
            sm[threadIdx.x][0] = 0;
            sm[threadIdx.x][R + 1] = 0;
        }
    }
    __syncthreads();
    const long rowl = 16L * R, plane = (long)P * P * rowl;
    bf16 *o = xs + (long)cb * 3 * plane + (long)row * rowl;
    for (int t = threadIdx.x; t < 3 * 16 * R; t += blockDim.x) {
        const int u = t % (16 * R), sh = t / (16 * R);
        const int k = ((u >> 6) << 2) | (u & 3), ci = (u >> 2) & 15;
        bf16 v;
        v.v = pad ? (uint16_t)0 : sm[ci][k + sh];
        o[(long)sh * plane + u] = v;
    }
}

// position of gz(v, co) in the staged gradient (C % 16 == 0, nv % 4 == 0)
__device__ __forceinline__ long gs_at(long nv, long v, int co)
{
    return (((long)(co >> 4) * (nv >> 2) + (v >> 2)) * 16 + (co & 15)) * 4 + (v & 3);
}

__global__ void __launch_bounds__(256) k_wg_stage_g(long nv, int C, const bf16 *__restrict__ gz, bf16 *__restrict__ gs)
{
    const long tot = nv * C;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (long)gridDim.x * blockDim.x) {
        // q = ((ob (nv/4) + v/4) 16 + co%16) 4 + v%4
        const long v = ((q >> 6) % (nv >> 2)) * 4 + (q & 3);
        const int co = (int)(q / (16 * nv)) * 16 + (int)((q >> 2) & 15);
        gs[q] = gz[v * C + co];
    }
}

// 1D grid of nch chunks x (ci blocks x co blocks) pairs, a multiple of 8 workgroups,
// 4 waves per block, gpw voxel groups of 16 per wave; part[split][t][ci][co], split =
// chunk * 4 + wave (every entry of every split is written, zeros for a wave past the
// last group).  Workgroup L runs on XCD L % 8; it takes logical block
// (L % 8) * (nblocks / 8) + L / 8, so each XCD sweeps a contiguous run of chunks and
// the rows that neighbouring chunks share (taps one row or plane away) meet in its L2.
template <int NTAP>
__global__ void __launch_bounds__(256) k_wg_staged(int R, const bf16 *__restrict__ xs, const bf16 *__restrict__ gs,
                                                   int ncib, int ncob, int nch, long gpw, float *__restrict__ part)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int L = blockIdx.x, per = gridDim.x >> 3;
    const int lb = (L & 7) * per + (L >> 3);
    const int chunk = lb % nch, pair = lb / nch;
    const int cb = pair / ncob, ob = pair % ncob;
    const long nv = (long)R * R * R, ngroups = nv / 16;
    const int P = R + 2;
    const int m = lane & 15, qd = lane >> 4;
    const long rowl = 16L * R, plane = (long)P * P * rowl;
    const bf16 *xb = xs + (long)cb * 3 * plane + 4 * m;
    const bf16 *gb = gs + (long)ob * 16 * nv + 4 * m;
    f4 acc[NTAP];
#pragma unroll
    for (int t = 0; t < NTAP; t++) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    const long split = (long)chunk * 4 + w;
    const long g0 = split * gpw;
    const long g1 = (g0 + gpw < ngroups) ? g0 + gpw : ngroups;
    for (long g = g0; g < g1; g++) {
        const long v = g * 16 + 4 * qd;  // this lane's quad of voxels (one k row: R % 4 == 0)
        const int k = (int)(v % R), j = (int)((v / R) % R), i = (int)(v / ((long)R * R));
        const bf4 bq = *(const bf4 *)(gb + 16 * v);  // B[k = voxel][n = co]
#pragma unroll
        for (int t = 0; t < NTAP; t++) {
            int a, b, c;
            tap_off(t, NTAP, a, b, c);
            const long row = (long)(i + a + 1) * P + (j + b + 1);
            const bf4 aq = *(const bf4 *)(xb + (c + 1) * plane + row * rowl + 16 * k);
            acc[t] = mfma16(as_s4(aq), as_s4(bq), acc[t]);  // A[m = ci][k = voxel]
        }
    }
    const int Cin = ncib * 16, Cout = ncob * 16;
#pragma unroll
    for (int t = 0; t < NTAP; t++)
        for (int q = 0; q < 4; q++)
            part[((split * NTAP + t) * Cin + cb * 16 + 4 * qd + q) * Cout + ob * 16 + m] = acc[t][q];
}

// first stage of the reduction of many partials: group g of the nsplit partials of
// every entry summed into part2[g][q] (grid: entry blocks x groups)
__global__ void __launch_bounds__(256) k_reduce_parts_grouped(long n, int nsplit, const float *part, float *part2)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int G = gridDim.y, g = blockIdx.y;
    const int p0 = (int)((long)nsplit * g / G), p1 = (int)((long)nsplit * (g + 1) / G);
    float s = 0.f;
    for (int p = p0; p < p1; p++) s += part[(long)p * n + q];
    part2[(long)g * n + q] = s;
}

// bias gradient partials from the staged gradient: block (g, co) sums its segment of
// the voxels of channel co; part[g][co]
__global__ void __launch_bounds__(256) k_rowsum_parts(long nv, const bf16 *__restrict__ gs, float *part)
{
    __shared__ float sh[256];
    const int G = gridDim.x, g = blockIdx.x, co = blockIdx.y, C = gridDim.y;
    const long v0 = nv * g / G, v1 = nv * (g + 1) / G;
    float s = 0.f;
    for (long v = v0 + threadIdx.x; v < v1; v += blockDim.x) s += to_f(gs[gs_at(nv, v, co)]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x >> 1; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)g * C + co] = sh[0];
}

__global__ void k_reduce_parts(long n, int nsplit, const float *part, float *out)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    float s = 0.f;
    for (int p = 0; p < nsplit; p++) s += part[(long)p * n + q];
    out[q] = s;
}

// bias gradient: gb[c] = sum_v gz[v][c]; one block per channel, fixed order
__global__ void __launch_bounds__(256) k_bias_grad(long nv, int C, const bf16 *gz, float *gb)
{
    __shared__ float sh[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (long v = threadIdx.x; v < nv; v += blockDim.x) s += to_f(gz[v * C + c]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x >> 1; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) gb[c] = sh[0];
}

// gz[v][c] = bf16(gy[v][c] * (y[v][c] > 0))   (relu) or bf16(gy) (linear)
__global__ void k_act_bwd(long nv, int C, const bf16 *y, int csy, int cy, const float *gy, int csg, int cg,
                          bf16 *gz, int relu)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nv * C) return;
    long v = q / C;
    int c = (int)(q % C);
    float g = gy[v * csg + cg + c];
    if (relu && !(to_f(y[v * csy + cy + c]) > 0.f)) g = 0.f;
    gz[q] = to_bf16(g);
}

// 2x average pooling (bf16 -> bf16) and its adjoint (f32 accumulate)
__global__ void k_pool(int Rc, int C, const bf16 *in, int csi, int cio, bf16 *out)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nvc = (long)Rc * Rc * Rc;
    if (q >= nvc * C) return;
    long vc = q / C;
    int c = (int)(q % C);
    int I = (int)(vc / ((long)Rc * Rc)), J = (int)((vc / Rc) % Rc), K = (int)(vc % Rc);
    int Rf = 2 * Rc;
    float s = 0.f;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++)
                s += to_f(in[(((long)(2 * I + a) * Rf + 2 * J + b) * Rf + 2 * K + d) * csi + cio + c]);
    out[q] = to_bf16(0.125f * s);
}
__global__ void k_pool_bwd(int Rc, int C, const float *gout, float *gin, int csg, int cig)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int Rf = 2 * Rc;
    long nvf = (long)Rf * Rf * Rf;
    if (q >= nvf * C) return;
    long vf = q / C;
    int c = (int)(q % C);
    int i = (int)(vf / ((long)Rf * Rf)), j = (int)((vf / Rf) % Rf), k = (int)(vf % Rf);
    long vc = ((long)(i / 2) * Rc + j / 2) * Rc + k / 2;
    gin[vf * csg + cig + c] += 0.125f * gout[vc * C + c];
}
// nearest 2x upsampling (bf16 -> bf16 slice) and its adjoint (f32 sum of children)
__global__ void k_up(int Rf, int C, const bf16 *in, bf16 *out, int cso, int coo)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nvf = (long)Rf * Rf * Rf;
    if (q >= nvf * C) return;
    long vf = q / C;
    int c = (int)(q % C);
    int Rc = Rf / 2;
    int i = (int)(vf / ((long)Rf * Rf)), j = (int)((vf / Rf) % Rf), k = (int)(vf % Rf);
    long vc = ((long)(i / 2) * Rc + j / 2) * Rc + k / 2;
    out[vf * cso + coo + c] = in[vc * C + c];
}
__global__ void k_up_bwd(int Rf, int C, const float *gout, int csg, int cig, float *gin)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int Rc = Rf / 2;
    long nvc = (long)Rc * Rc * Rc;
    if (q >= nvc * C) return;
    long vc = q / C;
    int c = (int)(q % C);
    int I = (int)(vc / ((long)Rc * Rc)), J = (int)((vc / Rc) % Rc), K = (int)(vc % Rc);
    float s = 0.f;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++) s += gout[(((long)(2 * I + a) * Rf + 2 * J + b) * Rf + 2 * K + d) * csg + cig + c];
    gin[q] = s;
}

// loss 0.5 * mean (y - t)^2 over voxels and the 4 used channels; gradient into gout
__global__ void __launch_bounds__(256) k_loss(long nv, const float *y, const float *tg, float *gout, double *part)
{
    __shared__ double sh[256];
    long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    const float inv = 1.0f / (float)(nv * NET_NOUT);
    if (v < nv) {
        for (int c = 0; c < 16; c++) {
            float g = 0.f;
            if (c < NET_NOUT) {
                float d = y[v * 16 + c] - tg[v * NET_NOUT + c];
                acc += 0.5 * (double)d * d;
                g = d * inv;
            }
            gout[v * 16 + c] = g;
        }
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x >> 1; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] / (double)(nv * NET_NOUT);
}

__global__ void k_adam(long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    float gg = g[q];
    float mm = b1 * m[q] + (1.f - b1) * gg;
    float vv = b2 * v[q] + (1.f - b2) * gg * gg;
    m[q] = mm;
    v[q] = vv;
    p[q] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// pack W[t][ci][co] (f32) into the MFMA A-operand lane order:
// forward  A[m = co][k = ci] per (t, ci block, co block)
// backward A[m = ci][k = co] of the flipped kernel W'[t][co][ci] = W[T-1-t][ci][co]
__global__ void k_pack_w(int ntap, int ncib, int ncob, const float *W, bf4 *Wp, int flip)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)ntap * ncib * ncob * 64;
    if (q >= n) return;
    int lane = (int)(q % 64);
    long r = q / 64;
    int Cin = ncib * 16, Cout = ncob * 16;
    bf4 o;
    if (!flip) {
        int ob = (int)(r % ncob), cb = (int)((r / ncob) % ncib), t = (int)(r / (ncob * ncib));
        for (int j = 0; j < 4; j++) {
            int ci = cb * 16 + 4 * (lane >> 4) + j, co = ob * 16 + (lane & 15);
            o.v[j] = to_bf16(W[((long)t * Cin + ci) * Cout + co]).v;
        }
    } else {
        // conv over gz (in channels = co, blocks ncob) producing gx (out channels = ci)
        int cb = (int)(r % ncib), ob = (int)((r / ncib) % ncob), t = (int)(r / (ncib * ncob));
        for (int j = 0; j < 4; j++) {
            int ci = cb * 16 + (lane & 15), co = ob * 16 + 4 * (lane >> 4) + j;
            o.v[j] = to_bf16(W[((long)(ntap - 1 - t) * Cin + ci) * Cout + co]).v;
        }
    }
    Wp[q] = o;
}

template <typename TS>
__global__ void k_cvt_bf16(long n, const TS *in, bf16 *out)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = to_bf16((float)Cvt<TS>::ld(in[q]));
}

// ------------------------------------------------ solver <-> network glue

// state restriction (fine solution -> coarse state): velocity = mean of the
// coincident child faces, pressure = 2 x mean of the children (unit spacing)
template <int DIM>
__global__ void k_state_restrict(Lvl gf, Lvl gc, const uint8_t *flc, const double *xf, double *xc)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= gc.S) return;
    int I = (int)(n / ((long)gc.Py * gc.Pz)), J = (int)((n / gc.Pz) % gc.Py), K = (int)(n % gc.Pz);
    const int kz = (DIM == 3) ? 2 : 1;
    auto fid = [&](int i, int j, int k) { return ((long)i * gf.Py + j) * gf.Pz + k; };
    const int fb = flc[n];
    double sp = 0.0;
    if (I < gc.nx && J < gc.ny && K < gc.nz) {
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int c = 0; c < kz; c++) sp += xf[fid(2 * I + a, 2 * J + b, kz * K + c)];
        sp *= 2.0 / (double)(4 * kz);
    }
    xc[n] = (fb & 1) ? sp : 0.0;
    for (int d = 0; d < DIM; d++) {
        double s = 0.0;
        int cnt = 0;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < ((DIM == 3) ? 2 : 1); b++) {
                int fi, fj, fk;
                if (d == 0) { fi = 2 * I; fj = 2 * J + a; fk = (DIM == 3) ? 2 * K + b : 0; }
                else if (d == 1) { fi = 2 * I + a; fj = 2 * J; fk = (DIM == 3) ? 2 * K + b : 0; }
                else { fi = 2 * I + a; fj = 2 * J + b; fk = 2 * K; }
                if (fi > gf.nx || fj > gf.ny || fk > ((DIM == 3) ? gf.nz : 0)) continue;
                s += xf[(long)(1 + d) * gf.S + fid(fi, fj, fk)];
                cnt++;
            }
        xc[(long)(1 + d) * gc.S + n] = (((fb >> (1 + d)) & 1) && cnt) ? s / cnt : 0.0;
    }
}

// features on the network grid R^3 (R = level-lc cells): block-mean density and
// three static boundary channels (fixed-face fraction, free boundary faces,
// Dirichlet x-velocity at the low x face), zero-padded to 16 channels
__global__ void k_features(Lvl gc, int dim, int nf, const double *rho, int ratio, const uint8_t *flc,
                           const double *xdc, bf16 *X0, float *feat)
{
    long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int R = gc.nx;
    long nv = (dim == 3) ? (long)R * R * R : (long)R * R;
    if (v >= nv) return;
    int I, J, K;
    if (dim == 3) { I = (int)(v / ((long)R * R)); J = (int)((v / R) % R); K = (int)(v % R); }
    else { I = (int)(v / R); J = (int)(v % R); K = 0; }
    double s = 0.0;
    const int kz = (dim == 3) ? ratio : 1;
    for (int a = 0; a < ratio; a++)
        for (int b = 0; b < ratio; b++)
            for (int c = 0; c < kz; c++) {
                long fi = I * ratio + a, fj = J * ratio + b, fk = K * kz + c;
                s += rho[(dim == 3) ? (fi * nf + fj) * nf + fk : fi * nf + fj];
            }
    s /= (double)(ratio * ratio * kz);
    const long q = ((long)I * gc.Py + J) * gc.Pz + K;
    const long sI = (long)gc.Py * gc.Pz, sJ = gc.Pz;
    int nfix = 0, nfree_b = 0;
    for (int d = 0; d < dim; d++)
        for (int h = 0; h < 2; h++) {
            long qq = q + h * (d == 0 ? sI : (d == 1 ? sJ : 1));
            int coord = (d == 0 ? I : (d == 1 ? J : K)) + h;
            bool fr = (flc[qq] >> (1 + d)) & 1;
            if (!fr) nfix++;
            if (fr && (coord == 0 || coord == R)) nfree_b++;
        }
    float f[4] = {(float)s, (float)nfix / (2.f * dim), (float)nfree_b, (float)xdc[gc.S + q]};
    for (int c = 0; c < 16; c++) X0[v * 16 + c] = to_bf16(c < 4 ? f[c] : 0.f);
    if (feat)
        for (int c = 0; c < 4; c++) feat[v * 4 + c] = f[c];
}

// targets: coarse state at the R^3 nodes, scaled by 1/scale[c]
__global__ void k_targets(Lvl gc, int dim, const double *xc, const float *scale, float *tg)
{
    long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int R = gc.nx;
    long nv = (dim == 3) ? (long)R * R * R : (long)R * R;
    if (v >= nv) return;
    int I, J, K;
    if (dim == 3) { I = (int)(v / ((long)R * R)); J = (int)((v / R) % R); K = (int)(v % R); }
    else { I = (int)(v / R); J = (int)(v % R); K = 0; }
    const long q = ((long)I * gc.Py + J) * gc.Pz + K;
    for (int c = 0; c < NET_NOUT; c++) {
        double val = (c <= dim) ? xc[(long)c * gc.S + q] : 0.0;
        tg[v * NET_NOUT + c] = (float)(val / scale[c]);
    }
}

// network output -> coarse state (padded FP64 vector of level lc), masked by flags
__global__ void k_out_state(Lvl gc, int dim, const float *y, const float *scale, const uint8_t *flc, double *xc)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= gc.S) return;
    int I = (int)(n / ((long)gc.Py * gc.Pz)), J = (int)((n / gc.Pz) % gc.Py), K = (int)(n % gc.Pz);
    const int R = gc.nx;
    const bool inside = I < R && J < R && (dim == 2 || K < R);
    const long v = (dim == 3) ? ((long)I * R + J) * R + K : (long)I * R + J;
    for (int c = 0; c <= dim; c++) {
        double val = 0.0;
        if (inside && ((flc[n] >> c) & 1)) val = (double)y[v * 16 + c] * scale[c];
        xc[(long)c * gc.S + n] = val;
    }
}

}  // namespace st

namespace st {
__global__ void k_feat_pad(long nv, const float *feat, bf16 *X0)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nv * 16) return;
    long v = q / 16;
    int c = (int)(q % 16);
    X0[q] = to_bf16(c < 4 ? feat[v * 4 + c] : 0.f);
}

__global__ void k_scale_targets(long n, const float *raw, const float *scale, float *out)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = raw[q] / scale[q % NET_NOUT];
}
}  // namespace st

// ================================================================= host side

namespace st {

namespace {
inline unsigned nblk(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }
inline unsigned nblk_cap(long n) { const unsigned b = nblk(n); return b > 8192 ? 8192 : (b < 1 ? 1 : b); }

// staged weight gradient (R % 4 == 0): about 2048 waves per layer, each over gpw groups
// of 16 voxels of one (ci block, co block)
struct WgPlan {
    long nchunks, gpw, nsplit;
};
WgPlan wg_plan(int R, int cin, int cout)
{
    const long ngroups = (long)R * R * R / 16, npairs = (long)(cin / 16) * (cout / 16);
    long nch = (512 + npairs - 1) / npairs;
    long gpw = (ngroups + 4 * nch - 1) / (4 * nch);
    if (gpw < 1) gpw = 1;
    nch = (ngroups + 4 * gpw - 1) / (4 * gpw);
    while ((nch * npairs) % 8) nch++;  // whole XCD runs (chunks past the last group write zeros)
    return {nch, gpw, nch * 4};
}
long wg_xs_elems(int R, int cin)
{
    const long P = R + 2;
    return (long)(cin / 16) * 3 * 16 * P * P * R;
}
bool wg_staged_ok(int R) { return R >= 4 && R % 4 == 0 && R <= 128; }
constexpr int WG_G = 32;  // groups of the first reduction stage (weights and bias)
// gW = sum over voxels of x(v + o_t) gz(v) for every tap, gb = sum over voxels of gz(v)
// (xs, gs: staging buffers of wg_xs_elems(R, cin) and R^3 cout elements; part:
// wg_plan(...).nsplit * ntap cin cout; part2: WG_G * max(ntap cin cout, cout))
void wgrad_staged(int R, int ntap, int cin, int cout, const bf16 *x, int csx, int cix, const bf16 *gz, bf16 *xs,
                  bf16 *gs, float *part, float *part2, float *gW, float *gb, hipStream_t s)
{
    const WgPlan pl = wg_plan(R, cin, cout);
    const long nv = (long)R * R * R, nw = (long)ntap * cin * cout;
    hipLaunchKernelGGL(k_wg_stage_x, dim3((unsigned)((R + 2) * (R + 2)), (unsigned)(cin / 16)), dim3(256), 0, s, R, x, csx,
                       cix, cin / 16, xs);
    hipLaunchKernelGGL(k_wg_stage_g, dim3(nblk_cap(nv * cout)), dim3(256), 0, s, nv, cout, gz, gs);
    const dim3 grid((unsigned)(pl.nchunks * (cin / 16) * (cout / 16)));
    const int nch = (int)pl.nchunks;
    if (ntap == 27)
        hipLaunchKernelGGL((k_wg_staged<27>), grid, dim3(256), 0, s, R, xs, gs, cin / 16, cout / 16, nch, pl.gpw, part);
    else
        hipLaunchKernelGGL((k_wg_staged<1>), grid, dim3(256), 0, s, R, xs, gs, cin / 16, cout / 16, nch, pl.gpw, part);
    hipLaunchKernelGGL(k_reduce_parts_grouped, dim3(nblk(nw), WG_G), dim3(256), 0, s, nw, (int)pl.nsplit, part, part2);
    hipLaunchKernelGGL(k_reduce_parts, dim3(nblk(nw)), dim3(256), 0, s, nw, WG_G, part2, gW);
    hipLaunchKernelGGL(k_rowsum_parts, dim3(WG_G, cout), dim3(256), 0, s, nv, gs, part2);
    hipLaunchKernelGGL(k_reduce_parts, dim3(nblk(cout)), dim3(256), 0, s, (long)cout, WG_G, part2, gb);
}

struct Layer {
    NetLayerSpec sp;
    long nw = 0;
    float *W = nullptr, *b = nullptr, *mW = nullptr, *vW = nullptr, *mb = nullptr, *vb = nullptr, *gW = nullptr,
          *gb = nullptr;
    bf4 *Wp = nullptr, *Wq = nullptr;
};
}  // namespace

struct NetImpl {
    int R = 0, R2 = 0, R4 = 0, maxs = 0, nsamp = 0, step = 0;
    long nv0 = 0, nv1 = 0, nv2 = 0;
    hipStream_t s = nullptr;
    std::vector<void *> allocs;
    Layer L[NET_NLAYER];
    bf16 *X0, *E1a, *D1in, *P1, *E2a, *D2in, *P2, *E3a, *E3b, *D2a, *D2b, *D1a, *D1b, *gz;
    float *OUT, *gOUT, *gD1b, *gD1a, *gD1in, *gD2b, *gD2a, *gD2in, *gE3b, *gE3a, *gP2, *gE2a, *gP1, *gE1a;
    float *part, *feat, *tgt_raw, *tgt, *scale;
    bf16 *wxs = nullptr, *wgs = nullptr;  // staging of the weight gradient
    float *part2 = nullptr;               // its first reduction stage
    double *lpart;
    bool scaled = false;
    float hscale[NET_NOUT] = {1, 1, 1, 1};
    int nsplit = 64;

    template <typename T> bool alloc(T **p, long n)
    {
        void *q = nullptr;
        if (hipMalloc(&q, (size_t)(n > 0 ? n : 1) * sizeof(T)) != hipSuccess) return false;
        hipMemsetAsync(q, 0, (size_t)(n > 0 ? n : 1) * sizeof(T), s);
        allocs.push_back(q);
        *p = (T *)q;
        return true;
    }
    long nv(int lvl) const { return lvl == 0 ? nv0 : (lvl == 1 ? nv1 : nv2); }
    int res(int lvl) const { return lvl == 0 ? R : (lvl == 1 ? R2 : R4); }

    void conv_fwd(int l, const bf16 *in, int csi, int cio, void *out, int cso, int coo, int outm)
    {
        const NetLayerSpec &p = L[l].sp;
        const int Rl = res(p.lvl);
        dim3 grid(nblk(nv(p.lvl), 64), p.cout / 16);
        const int ncib = p.cin / 16, ncob = p.cout / 16;
        if (p.ntap == 27) {
            if (outm == 0 && p.relu)
                hipLaunchKernelGGL((k_conv<27, 0, true>), grid, dim3(256), 0, s, Rl, in, csi, cio, ncib, L[l].Wp, L[l].b, out, cso, coo, ncob);
            else
                hipLaunchKernelGGL((k_conv<27, 1, false>), grid, dim3(256), 0, s, Rl, in, csi, cio, ncib, L[l].Wp, L[l].b, out, cso, coo, ncob);
        } else {
            hipLaunchKernelGGL((k_conv<1, 1, false>), grid, dim3(256), 0, s, Rl, in, csi, cio, ncib, L[l].Wp, L[l].b, out, cso, coo, ncob);
        }
    }
    // gx = conv^T(gz) written (f32) into a channel slice
    void conv_bwd_data(int l, const bf16 *gzin, float *gx, int csg, int cig)
    {
        const NetLayerSpec &p = L[l].sp;
        const int Rl = res(p.lvl);
        dim3 grid(nblk(nv(p.lvl), 64), p.cin / 16);
        if (p.ntap == 27)
            hipLaunchKernelGGL((k_conv<27, 1, false>), grid, dim3(256), 0, s, Rl, gzin, p.cout, 0, p.cout / 16, L[l].Wq, (const float *)nullptr, (void *)gx, csg, cig, p.cin / 16);
        else
            hipLaunchKernelGGL((k_conv<1, 1, false>), grid, dim3(256), 0, s, Rl, gzin, p.cout, 0, p.cout / 16, L[l].Wq, (const float *)nullptr, (void *)gx, csg, cig, p.cin / 16);
    }
    void conv_wgrad(int l, const bf16 *x, int csx, int cix, const bf16 *gzin)
    {
        const NetLayerSpec &p = L[l].sp;
        const int Rl = res(p.lvl);
        const long n = nv(p.lvl);
        if (wg_staged_ok(Rl)) {
            wgrad_staged(Rl, p.ntap, p.cin, p.cout, x, csx, cix, gzin, wxs, wgs, part, part2, L[l].gW, L[l].gb, s);
            return;
        }
        int ns = nsplit;
        long chunk = (n + ns - 1) / ns;
        chunk = (chunk + 15) / 16 * 16;
        ns = (int)((n + chunk - 1) / chunk);
        dim3 grid(p.ntap * (p.cin / 16) * (p.cout / 16), ns);
        if (p.ntap == 27)
            hipLaunchKernelGGL((k_conv_wgrad<27>), grid, dim3(64), 0, s, Rl, x, csx, cix, p.cin / 16, gzin, p.cout / 16, chunk, part);
        else
            hipLaunchKernelGGL((k_conv_wgrad<1>), grid, dim3(64), 0, s, Rl, x, csx, cix, p.cin / 16, gzin, p.cout / 16, chunk, part);
        hipLaunchKernelGGL(k_reduce_parts, dim3(nblk(L[l].nw)), dim3(256), 0, s, L[l].nw, ns, part, L[l].gW);
        hipLaunchKernelGGL(k_bias_grad, dim3(p.cout), dim3(256), 0, s, n, p.cout, gzin, L[l].gb);
    }
    void act_bwd(int l, const bf16 *y, int csy, int cy, const float *gy, int csg, int cg)
    {
        const NetLayerSpec &p = L[l].sp;
        const long n = nv(p.lvl);
        hipLaunchKernelGGL(k_act_bwd, dim3(nblk(n * p.cout)), dim3(256), 0, s, n, p.cout, y, csy, cy, gy, csg, cg, gz, p.relu);
    }
    void pack()
    {
        for (int l = 0; l < NET_NLAYER; l++) {
            const NetLayerSpec &p = L[l].sp;
            long n = (long)p.ntap * (p.cin / 16) * (p.cout / 16) * 64;
            hipLaunchKernelGGL(k_pack_w, dim3(nblk(n)), dim3(256), 0, s, p.ntap, p.cin / 16, p.cout / 16, L[l].W, L[l].Wp, 0);
            hipLaunchKernelGGL(k_pack_w, dim3(nblk(n)), dim3(256), 0, s, p.ntap, p.cin / 16, p.cout / 16, L[l].W, L[l].Wq, 1);
        }
    }
    void forward()
    {
        conv_fwd(0, X0, 16, 0, E1a, 16, 0, 0);
        conv_fwd(1, E1a, 16, 0, D1in, 48, 32, 0);
        hipLaunchKernelGGL(k_pool, dim3(nblk(nv1 * 16)), dim3(256), 0, s, R2, 16, D1in, 48, 32, P1);
        conv_fwd(2, P1, 16, 0, E2a, 32, 0, 0);
        conv_fwd(3, E2a, 32, 0, D2in, 96, 64, 0);
        hipLaunchKernelGGL(k_pool, dim3(nblk(nv2 * 32)), dim3(256), 0, s, R4, 32, D2in, 96, 64, P2);
        conv_fwd(4, P2, 32, 0, E3a, 64, 0, 0);
        conv_fwd(5, E3a, 64, 0, E3b, 64, 0, 0);
        hipLaunchKernelGGL(k_up, dim3(nblk(nv1 * 64)), dim3(256), 0, s, R2, 64, E3b, D2in, 96, 0);
        conv_fwd(6, D2in, 96, 0, D2a, 32, 0, 0);
        conv_fwd(7, D2a, 32, 0, D2b, 32, 0, 0);
        hipLaunchKernelGGL(k_up, dim3(nblk(nv0 * 32)), dim3(256), 0, s, R, 32, D2b, D1in, 48, 0);
        conv_fwd(8, D1in, 48, 0, D1a, 16, 0, 0);
        conv_fwd(9, D1a, 16, 0, D1b, 16, 0, 0);
        conv_fwd(10, D1b, 16, 0, OUT, 16, 0, 1);
    }
    // gradients of all parameters from gOUT (dL/dOUT)
    void backward()
    {
        act_bwd(10, nullptr, 16, 0, gOUT, 16, 0);
        conv_wgrad(10, D1b, 16, 0, gz);
        conv_bwd_data(10, gz, gD1b, 16, 0);
        act_bwd(9, D1b, 16, 0, gD1b, 16, 0);
        conv_wgrad(9, D1a, 16, 0, gz);
        conv_bwd_data(9, gz, gD1a, 16, 0);
        act_bwd(8, D1a, 16, 0, gD1a, 16, 0);
        conv_wgrad(8, D1in, 48, 0, gz);
        conv_bwd_data(8, gz, gD1in, 48, 0);
        hipLaunchKernelGGL(k_up_bwd, dim3(nblk(nv1 * 32)), dim3(256), 0, s, R, 32, gD1in, 48, 0, gD2b);
        act_bwd(7, D2b, 32, 0, gD2b, 32, 0);
        conv_wgrad(7, D2a, 32, 0, gz);
        conv_bwd_data(7, gz, gD2a, 32, 0);
        act_bwd(6, D2a, 32, 0, gD2a, 32, 0);
        conv_wgrad(6, D2in, 96, 0, gz);
        conv_bwd_data(6, gz, gD2in, 96, 0);
        hipLaunchKernelGGL(k_up_bwd, dim3(nblk(nv2 * 64)), dim3(256), 0, s, R2, 64, gD2in, 96, 0, gE3b);
        act_bwd(5, E3b, 64, 0, gE3b, 64, 0);
        conv_wgrad(5, E3a, 64, 0, gz);
        conv_bwd_data(5, gz, gE3a, 64, 0);
        act_bwd(4, E3a, 64, 0, gE3a, 64, 0);
        conv_wgrad(4, P2, 32, 0, gz);
        conv_bwd_data(4, gz, gP2, 32, 0);
        hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(nv1 * 32)), dim3(256), 0, s, R4, 32, gP2, gD2in, 96, 64);
        act_bwd(3, D2in, 96, 64, gD2in, 96, 64);
        conv_wgrad(3, E2a, 32, 0, gz);
        conv_bwd_data(3, gz, gE2a, 32, 0);
        act_bwd(2, E2a, 32, 0, gE2a, 32, 0);
        conv_wgrad(2, P1, 16, 0, gz);
        conv_bwd_data(2, gz, gP1, 16, 0);
        hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(nv0 * 16)), dim3(256), 0, s, R2, 16, gP1, gD1in, 48, 32);
        act_bwd(1, D1in, 48, 32, gD1in, 48, 32);
        conv_wgrad(1, E1a, 16, 0, gz);
        conv_bwd_data(1, gz, gE1a, 16, 0);
        act_bwd(0, E1a, 16, 0, gE1a, 16, 0);
        conv_wgrad(0, X0, 16, 0, gz);
    }
    double loss_of(const float *tg)
    {
        const unsigned nb = nblk(nv0);
        hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, nv0, OUT, tg, gOUT, lpart);
        std::vector<double> h(nb);
        hipMemcpyAsync(h.data(), lpart, sizeof(double) * nb, hipMemcpyDeviceToHost, s);
        hipStreamSynchronize(s);
        double t = 0.0;
        for (double v : h) t += v;
        return t;
    }
    void features_from_host(const float *feat_dev)
    {
        // feat (R^3 x 4, f32) -> X0 (R^3 x 16, bf16, zero-padded)
        hipLaunchKernelGGL(k_feat_pad, dim3(nblk(nv0 * 16)), dim3(256), 0, s, nv0, feat_dev, X0);
    }
};

}  // namespace st

namespace st {

static unsigned long long g_rng = 0;
static double urand()
{
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return (double)(g_rng >> 11) * (1.0 / 9007199254740992.0);
}

NetImpl *net_create(int R, int max_samples, unsigned seed, hipStream_t s)
{
    if (R < 4 || R % 4 != 0) return nullptr;
    NetImpl *n = new NetImpl();
    n->R = R;
    n->R2 = R / 2;
    n->R4 = R / 4;
    n->nv0 = (long)R * R * R;
    n->nv1 = n->nv0 / 8;
    n->nv2 = n->nv1 / 8;
    n->maxs = max_samples;
    n->s = s;
    bool ok = true;
    long maxparts = 0, wgparts = 0, wgx = 0, wgg = 0;
    for (int l = 0; l < NET_NLAYER && ok; l++) {
        Layer &L = n->L[l];
        L.sp = NET_SPEC[l];
        L.nw = (long)L.sp.ntap * L.sp.cin * L.sp.cout;
        long np = (long)L.sp.ntap * (L.sp.cin / 16) * (L.sp.cout / 16) * 64;
        ok = ok && n->alloc(&L.W, L.nw) && n->alloc(&L.b, L.sp.cout) && n->alloc(&L.mW, L.nw) && n->alloc(&L.vW, L.nw) &&
             n->alloc(&L.mb, L.sp.cout) && n->alloc(&L.vb, L.sp.cout) && n->alloc(&L.gW, L.nw) &&
             n->alloc(&L.gb, L.sp.cout) && n->alloc(&L.Wp, np) && n->alloc(&L.Wq, np);
        if (L.nw > maxparts) maxparts = L.nw;
        const int Rl = R >> L.sp.lvl;
        if (wg_staged_ok(Rl)) {
            wgparts = std::max(wgparts, wg_plan(Rl, L.sp.cin, L.sp.cout).nsplit * L.nw);
            wgx = std::max(wgx, wg_xs_elems(Rl, L.sp.cin));
            wgg = std::max(wgg, (long)Rl * Rl * Rl * L.sp.cout);
        }
    }
    const long v0 = n->nv0, v1 = n->nv1, v2 = n->nv2;
    ok = ok && n->alloc(&n->X0, v0 * 16) && n->alloc(&n->E1a, v0 * 16) && n->alloc(&n->D1in, v0 * 48) &&
         n->alloc(&n->P1, v1 * 16) && n->alloc(&n->E2a, v1 * 32) && n->alloc(&n->D2in, v1 * 96) &&
         n->alloc(&n->P2, v2 * 32) && n->alloc(&n->E3a, v2 * 64) && n->alloc(&n->E3b, v2 * 64) &&
         n->alloc(&n->D2a, v1 * 32) && n->alloc(&n->D2b, v1 * 32) && n->alloc(&n->D1a, v0 * 16) &&
         n->alloc(&n->D1b, v0 * 16) && n->alloc(&n->gz, v0 * 64) && n->alloc(&n->OUT, v0 * 16);
    ok = ok && n->alloc(&n->gOUT, v0 * 16) && n->alloc(&n->gD1b, v0 * 16) && n->alloc(&n->gD1a, v0 * 16) &&
         n->alloc(&n->gD1in, v0 * 48) && n->alloc(&n->gD2b, v1 * 32) && n->alloc(&n->gD2a, v1 * 32) &&
         n->alloc(&n->gD2in, v1 * 96) && n->alloc(&n->gE3b, v2 * 64) && n->alloc(&n->gE3a, v2 * 64) &&
         n->alloc(&n->gP2, v2 * 32) && n->alloc(&n->gE2a, v1 * 32) && n->alloc(&n->gP1, v1 * 16) &&
         n->alloc(&n->gE1a, v0 * 16);
    ok = ok && n->alloc(&n->part, std::max(maxparts * n->nsplit, wgparts)) && n->alloc(&n->lpart, nblk(v0) + 1) &&
         n->alloc(&n->wxs, wgx) && n->alloc(&n->wgs, wgg) && n->alloc(&n->part2, (long)WG_G * maxparts) &&
         n->alloc(&n->feat, (long)(max_samples + 1) * v0 * 4) && n->alloc(&n->tgt_raw, (long)(max_samples + 1) * v0 * 4) &&
         n->alloc(&n->tgt, v0 * 4) && n->alloc(&n->scale, NET_NOUT);
    if (!ok) {
        net_destroy(n);
        return nullptr;
    }
    // deterministic He-uniform initialisation (fan-in = taps x used input channels)
    g_rng = 0x9E3779B97F4A7C15ull ^ (unsigned long long)seed;
    for (int l = 0; l < NET_NLAYER; l++) {
        Layer &L = n->L[l];
        const int cin_used = (l == 0) ? 4 : L.sp.cin;
        const double bound = std::sqrt(6.0 / (double)(L.sp.ntap * cin_used)) * (l == NET_NLAYER - 1 ? 0.1 : 1.0);
        std::vector<float> w(L.nw, 0.f);
        for (int t = 0; t < L.sp.ntap; t++)
            for (int ci = 0; ci < L.sp.cin; ci++)
                for (int co = 0; co < L.sp.cout; co++)
                    w[((long)t * L.sp.cin + ci) * L.sp.cout + co] = (ci < cin_used) ? (float)((2.0 * urand() - 1.0) * bound) : 0.f;
        hipMemcpyAsync(L.W, w.data(), sizeof(float) * L.nw, hipMemcpyHostToDevice, s);
    }
    std::vector<float> one(NET_NOUT, 1.f);
    hipMemcpyAsync(n->scale, one.data(), sizeof(float) * NET_NOUT, hipMemcpyHostToDevice, s);
    n->pack();
    hipStreamSynchronize(s);
    return n;
}

void net_destroy(NetImpl *n)
{
    if (!n) return;
    hipStreamSynchronize(n->s);
    for (void *q : n->allocs) hipFree(q);
    delete n;
}

long net_num_params(const NetImpl *n)
{
    long t = 0;
    for (int l = 0; l < NET_NLAYER; l++) t += n->L[l].nw + n->L[l].sp.cout;
    return t;
}

int net_get_params(NetImpl *n, float *h)
{
    long o = 0;
    for (int l = 0; l < NET_NLAYER; l++) {
        hipMemcpyAsync(h + o, n->L[l].W, sizeof(float) * n->L[l].nw, hipMemcpyDeviceToHost, n->s);
        o += n->L[l].nw;
        hipMemcpyAsync(h + o, n->L[l].b, sizeof(float) * n->L[l].sp.cout, hipMemcpyDeviceToHost, n->s);
        o += n->L[l].sp.cout;
    }
    return hipStreamSynchronize(n->s) == hipSuccess ? 0 : -1;
}

int net_set_params(NetImpl *n, const float *h)
{
    long o = 0;
    for (int l = 0; l < NET_NLAYER; l++) {
        hipMemcpyAsync(n->L[l].W, h + o, sizeof(float) * n->L[l].nw, hipMemcpyHostToDevice, n->s);
        o += n->L[l].nw;
        hipMemcpyAsync(n->L[l].b, h + o, sizeof(float) * n->L[l].sp.cout, hipMemcpyHostToDevice, n->s);
        o += n->L[l].sp.cout;
    }
    n->pack();
    return hipStreamSynchronize(n->s) == hipSuccess ? 0 : -1;
}

static void features(NetImpl *n, const Lvl &gc, int nf, int ratio, const double *rho, const uint8_t *flc,
                     const double *xdc, float *feat_out)
{
    hipLaunchKernelGGL(k_features, dim3(nblk(n->nv0)), dim3(256), 0, n->s, gc, 3, nf, rho, ratio, flc, xdc, n->X0,
                       feat_out);
}

int net_add_sample(NetImpl *n, const Lvl &gc, int nf, int ratio, const double *rho, const uint8_t *flc,
                   const double *xdc, const double *xc)
{
    if (gc.nx != n->R || gc.dim != 3) return -1;
    const int slot = n->nsamp % n->maxs;
    features(n, gc, nf, ratio, rho, flc, xdc, n->feat + (long)slot * n->nv0 * 4);
    std::vector<float> one(NET_NOUT, 1.f);
    float *ones = n->tgt;  // temporary unit scale
    hipMemcpyAsync(ones, one.data(), sizeof(float) * NET_NOUT, hipMemcpyHostToDevice, n->s);
    hipLaunchKernelGGL(k_targets, dim3(nblk(n->nv0)), dim3(256), 0, n->s, gc, 3, xc, ones,
                       n->tgt_raw + (long)slot * n->nv0 * 4);
    n->nsamp++;
    return hipStreamSynchronize(n->s) == hipSuccess ? 0 : -1;
}

int net_num_samples(const NetImpl *n) { return n->nsamp < n->maxs ? n->nsamp : n->maxs; }

int net_eval(NetImpl *n, const Lvl &gc, int nf, int ratio, const double *rho, const uint8_t *flc, const double *xdc,
             const double *xc, double *loss, double *loss_zero)
{
    if (gc.nx != n->R || gc.dim != 3 || !n->scaled) return -1;
    float *feat = n->feat + (long)n->maxs * n->nv0 * 4, *raw = n->tgt_raw + (long)n->maxs * n->nv0 * 4;
    features(n, gc, nf, ratio, rho, flc, xdc, feat);
    std::vector<float> one(NET_NOUT, 1.f);
    hipMemcpyAsync(n->tgt, one.data(), sizeof(float) * NET_NOUT, hipMemcpyHostToDevice, n->s);
    hipLaunchKernelGGL(k_targets, dim3(nblk(n->nv0)), dim3(256), 0, n->s, gc, 3, xc, n->tgt, raw);
    hipLaunchKernelGGL(k_scale_targets, dim3(nblk(n->nv0 * 4)), dim3(256), 0, n->s, n->nv0 * 4, raw, n->scale, n->tgt);
    n->forward();
    *loss = n->loss_of(n->tgt);
    std::vector<float> t((size_t)n->nv0 * 4);
    hipMemcpyAsync(t.data(), n->tgt, sizeof(float) * t.size(), hipMemcpyDeviceToHost, n->s);
    if (hipStreamSynchronize(n->s) != hipSuccess) return -1;
    double z = 0.0;
    for (float v : t) z += (double)v * v;
    *loss_zero = 0.5 * z / (double)t.size();
    return 0;
}

static void set_scales(NetImpl *n)
{
    const int ns = net_num_samples(n);
    std::vector<float> h((long)ns * n->nv0 * 4);
    hipMemcpyAsync(h.data(), n->tgt_raw, sizeof(float) * h.size(), hipMemcpyDeviceToHost, n->s);
    hipStreamSynchronize(n->s);
    double ss[NET_NOUT] = {0, 0, 0, 0};
    for (size_t q = 0; q < h.size(); q++) ss[q % NET_NOUT] += (double)h[q] * h[q];
    for (int c = 0; c < NET_NOUT; c++) {
        double rms = std::sqrt(ss[c] / (double)(h.size() / NET_NOUT));
        n->hscale[c] = (float)(rms > 0 ? rms : 1.0);
    }
    hipMemcpyAsync(n->scale, n->hscale, sizeof(float) * NET_NOUT, hipMemcpyHostToDevice, n->s);
    n->scaled = true;
}

static void adam_step(NetImpl *n, float lr)
{
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
    for (int l = 0; l < NET_NLAYER; l++) {
        Layer &L = n->L[l];
        hipLaunchKernelGGL(k_adam, dim3(nblk(L.nw)), dim3(256), 0, n->s, L.nw, L.W, L.gW, L.mW, L.vW, lr, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(k_adam, dim3(nblk(L.sp.cout)), dim3(256), 0, n->s, (long)L.sp.cout, L.b, L.gb, L.mb, L.vb, lr, b1, b2, eps, c1, c2);
    }
    n->pack();
}

int net_train(NetImpl *n, int epochs, float lr, double *loss)
{
    const int ns = net_num_samples(n);
    if (ns < 1) return -1;
    if (!n->scaled) set_scales(n);
    double last = 0.0;
    for (int e = 0; e < epochs; e++) {
        double acc = 0.0;
        for (int q = 0; q < ns; q++) {
            hipLaunchKernelGGL(k_feat_pad, dim3(nblk(n->nv0 * 16)), dim3(256), 0, n->s, n->nv0, n->feat + (long)q * n->nv0 * 4, n->X0);
            hipLaunchKernelGGL(k_scale_targets, dim3(nblk(n->nv0 * 4)), dim3(256), 0, n->s, n->nv0 * 4, n->tgt_raw + (long)q * n->nv0 * 4, n->scale, n->tgt);
            n->forward();
            acc += n->loss_of(n->tgt);
            n->backward();
            adam_step(n, lr);
        }
        last = acc / ns;
    }
    if (loss) *loss = last;
    return hipStreamSynchronize(n->s) == hipSuccess ? 0 : -1;
}

int net_predict(NetImpl *n, const Lvl &gc, int nf, int ratio, const double *rho, const uint8_t *flc,
                const double *xdc, double *xc)
{
    if (gc.nx != n->R || gc.dim != 3) return -1;
    features(n, gc, nf, ratio, rho, flc, xdc, nullptr);
    n->forward();
    hipLaunchKernelGGL(k_out_state, dim3(nblk(gc.S)), dim3(256), 0, n->s, gc, 3, n->OUT, n->scale, flc, xc);
    return 0;
}

int net_grad_host(NetImpl *n, const float *feat, const float *tgt, double *loss, float *grad, float *out)
{
    float *dfeat = n->feat + (long)n->maxs * n->nv0 * 4, *dt = n->tgt;  // staging slot after the samples
    hipMemcpyAsync(dfeat, feat, sizeof(float) * n->nv0 * 4, hipMemcpyHostToDevice, n->s);
    hipMemcpyAsync(dt, tgt, sizeof(float) * n->nv0 * 4, hipMemcpyHostToDevice, n->s);
    hipLaunchKernelGGL(k_feat_pad, dim3(nblk(n->nv0 * 16)), dim3(256), 0, n->s, n->nv0, dfeat, n->X0);
    n->forward();
    double L = n->loss_of(dt);
    n->backward();
    if (loss) *loss = L;
    if (out) hipMemcpyAsync(out, n->OUT, sizeof(float) * n->nv0 * 16, hipMemcpyDeviceToHost, n->s);
    long o = 0;
    for (int l = 0; l < NET_NLAYER; l++) {
        hipMemcpyAsync(grad + o, n->L[l].gW, sizeof(float) * n->L[l].nw, hipMemcpyDeviceToHost, n->s);
        o += n->L[l].nw;
        hipMemcpyAsync(grad + o, n->L[l].gb, sizeof(float) * n->L[l].sp.cout, hipMemcpyDeviceToHost, n->s);
        o += n->L[l].sp.cout;
    }
    return hipStreamSynchronize(n->s) == hipSuccess ? 0 : -1;
}

void launch_state_restrict(const Lvl &gf, const Lvl &gc, const uint8_t *flc, const double *xf, double *xc,
                           hipStream_t s)
{
    if (gf.dim == 3) hipLaunchKernelGGL((k_state_restrict<3>), dim3(nblk(gc.S)), dim3(256), 0, s, gf, gc, flc, xf, xc);
    else hipLaunchKernelGGL((k_state_restrict<2>), dim3(nblk(gc.S)), dim3(256), 0, s, gf, gc, flc, xf, xc);
}

int net_conv_test(int R, int ntap, int cin, int cout, int relu, int mode, const float *in, const float *w,
                  const float *bias, const float *gout, float *out, float *gw, float *gb)
{
    if (cin % 16 || cout % 16 || (ntap != 1 && ntap != 27) || R < 1) return -1;
    const long nv = (long)R * R * R, nw = (long)ntap * cin * cout;
    const long np = (long)ntap * (cin / 16) * (cout / 16) * 64;
    hipStream_t s = nullptr;
    std::vector<void *> al;
    auto A = [&](long bytes) { void *p = nullptr; hipMalloc(&p, bytes > 0 ? bytes : 1); hipMemset(p, 0, bytes > 0 ? bytes : 1); al.push_back(p); return p; };
    float *din = (float *)A(nv * cin * 4), *dw = (float *)A(nw * 4), *db = (float *)A(cout * 4), *dgo = (float *)A(nv * cout * 4);
    float *dout = (float *)A(nv * (cin > cout ? cin : cout) * 4), *dgw = (float *)A(nw * 4), *dgb = (float *)A(cout * 4);
    bf16 *bin = (bf16 *)A(nv * cin * 2), *bgo = (bf16 *)A(nv * cout * 2);
    bf4 *Wp = (bf4 *)A(np * 8), *Wq = (bf4 *)A(np * 8);
    const int nsplit = 16;
    float *part = (float *)A(nw * nsplit * 4);
    hipMemcpy(dw, w, nw * 4, hipMemcpyHostToDevice);
    if (bias) hipMemcpy(db, bias, cout * 4, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_pack_w, dim3(nblk(np)), dim3(256), 0, s, ntap, cin / 16, cout / 16, dw, Wp, 0);
    hipLaunchKernelGGL(k_pack_w, dim3(nblk(np)), dim3(256), 0, s, ntap, cin / 16, cout / 16, dw, Wq, 1);
    if (mode == 0 || mode == 2) {
        hipMemcpy(din, in, nv * cin * 4, hipMemcpyHostToDevice);
        hipLaunchKernelGGL((k_cvt_bf16<float>), dim3(nblk(nv * cin)), dim3(256), 0, s, nv * cin, (const float *)din, bin);
    }
    if (mode == 1 || mode == 2) {
        hipMemcpy(dgo, gout, nv * cout * 4, hipMemcpyHostToDevice);
        hipLaunchKernelGGL((k_cvt_bf16<float>), dim3(nblk(nv * cout)), dim3(256), 0, s, nv * cout, (const float *)dgo, bgo);
    }
    if (mode == 0) {
        dim3 grid(nblk(nv, 64), cout / 16);
        if (ntap == 27 && relu) hipLaunchKernelGGL((k_conv<27, 1, true>), grid, dim3(256), 0, s, R, bin, cin, 0, cin / 16, Wp, db, (void *)dout, cout, 0, cout / 16);
        else if (ntap == 27) hipLaunchKernelGGL((k_conv<27, 1, false>), grid, dim3(256), 0, s, R, bin, cin, 0, cin / 16, Wp, db, (void *)dout, cout, 0, cout / 16);
        else hipLaunchKernelGGL((k_conv<1, 1, false>), grid, dim3(256), 0, s, R, bin, cin, 0, cin / 16, Wp, db, (void *)dout, cout, 0, cout / 16);
        hipMemcpy(out, dout, nv * cout * 4, hipMemcpyDeviceToHost);
    } else if (mode == 1) {
        dim3 grid(nblk(nv, 64), cin / 16);
        if (ntap == 27) hipLaunchKernelGGL((k_conv<27, 1, false>), grid, dim3(256), 0, s, R, bgo, cout, 0, cout / 16, Wq, (const float *)nullptr, (void *)dout, cin, 0, cin / 16);
        else hipLaunchKernelGGL((k_conv<1, 1, false>), grid, dim3(256), 0, s, R, bgo, cout, 0, cout / 16, Wq, (const float *)nullptr, (void *)dout, cin, 0, cin / 16);
        hipMemcpy(out, dout, nv * cin * 4, hipMemcpyDeviceToHost);
    } else if (wg_staged_ok(R)) {
        bf16 *xs = (bf16 *)A(wg_xs_elems(R, cin) * 2), *gs = (bf16 *)A(nv * cout * 2);
        float *wpart = (float *)A(wg_plan(R, cin, cout).nsplit * nw * 4);
        float *wpart2 = (float *)A((long)WG_G * (nw > cout ? nw : cout) * 4);
        wgrad_staged(R, ntap, cin, cout, bin, cin, 0, bgo, xs, gs, wpart, wpart2, dgw, dgb, s);
        hipMemcpy(gw, dgw, nw * 4, hipMemcpyDeviceToHost);
        hipMemcpy(gb, dgb, cout * 4, hipMemcpyDeviceToHost);
    } else {
        long chunk = (nv + nsplit - 1) / nsplit;
        chunk = (chunk + 15) / 16 * 16;
        int ns = (int)((nv + chunk - 1) / chunk);
        dim3 grid(ntap * (cin / 16) * (cout / 16), ns);
        if (ntap == 27) hipLaunchKernelGGL((k_conv_wgrad<27>), grid, dim3(64), 0, s, R, bin, cin, 0, cin / 16, bgo, cout / 16, chunk, part);
        else hipLaunchKernelGGL((k_conv_wgrad<1>), grid, dim3(64), 0, s, R, bin, cin, 0, cin / 16, bgo, cout / 16, chunk, part);
        hipLaunchKernelGGL(k_reduce_parts, dim3(nblk(nw)), dim3(256), 0, s, nw, ns, part, dgw);
        hipLaunchKernelGGL(k_bias_grad, dim3(cout), dim3(256), 0, s, nv, cout, bgo, dgb);
        hipMemcpy(gw, dgw, nw * 4, hipMemcpyDeviceToHost);
        hipMemcpy(gb, dgb, cout * 4, hipMemcpyDeviceToHost);
    }
    hipError_t e = hipDeviceSynchronize();
    for (void *p : al) hipFree(p);
    return e == hipSuccess ? 0 : -2;
}

}  // namespace st
