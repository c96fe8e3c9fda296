// st_objective_noalloc: the objective J = x.Gx / 2 of st_objective for callers that
// cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).  No context: the descriptor and the state come in as device arrays in
// the ABI layout, J goes out to a device scalar, and the only scratch -- one FP64
// partial sum per workgroup -- lives in the file-scope device array `scratch` below.
//
// G is the reference's objective operator: 3D FT/FluidEnergyGrad3D.cu:27-134
// (grad-div + curl-curl of the masked velocity plus the penalty), 2D
// FT/FluidEnergyGrad.cu:30-61 (face 5-point Laplacian of the unmasked velocity, fixed
// rows zeroed, plus the penalty); J = x.Gx / 2 (FT/FluidEnergy3D.cpp:94-109).  The
// stencil reads the ABI layout directly (no packing into the padded layout), so no
// vector-sized buffer is needed.
#include "../../include/stokes_to.h"
#include "st_common.h"

#include <cstdio>
#include <mutex>

namespace {

constexpr int NA_BLOCKS = 2048;  // fixed grid: partial sums in `scratch`
constexpr int NA_THREADS = 256;
constexpr int NA_MAX_N = 1024;   // largest n accepted (64-bit index math up to here)
constexpr int NA_MAX_DEV = 64;

// one partial per workgroup plus one slot of slack: 16 KiB on every device the
// library runs on, held until the process ends
__device__ __align__(256) double scratch[NA_BLOCKS + 1];

// device address of `scratch` per device (hipGetSymbolAddress at the first call on it)
double *g_scratch[NA_MAX_DEV];
std::mutex g_mu;
// message of the last failed call on this thread (st_last_error(NULL)); a fixed
// buffer, so that reporting an error allocates nothing either
thread_local char g_err[192];

int fail_na(int code, const char *msg)
{
    std::snprintf(g_err, sizeof(g_err), "st_objective_noalloc: %s", msg);
    return code;
}

// ABI layout: [cells | faces_x | faces_y | faces_z], each C-order with k fastest;
// field f has dims (n + (f==1), n + (f==2), n + (f==3)) (2D: nz = 1, no z faces)
struct AbiG {
    int dim, n, nz;
    long off[4];
    const uint8_t *fixed;
    const double *pen, *x;

    __device__ bool in(int f, int i, int j, int k) const
    {
        const int A = n + (f == 1), B = n + (f == 2), C = nz + (f == 3);
        return (unsigned)i < (unsigned)A && (unsigned)j < (unsigned)B && (unsigned)k < (unsigned)C;
    }
    __device__ long at(int f, int i, int j, int k) const
    {
        const int B = n + (f == 2), C = nz + (f == 3);
        return off[f] + ((long)i * B + j) * C + k;
    }
    // masked velocity w = m u (zero outside the face array and on fixed faces)
    __device__ double W(int d, int i, int j, int k) const
    {
        if (d >= dim || !in(1 + d, i, j, k)) return 0.0;
        const long q = at(1 + d, i, j, k);
        return fixed[q] ? 0.0 : x[q];
    }
    __device__ double U(int d, int i, int j, int k) const  // unmasked, zero outside
    {
        return in(1 + d, i, j, k) ? x[at(1 + d, i, j, k)] : 0.0;
    }
    __device__ double gdiv(int i, int j, int k) const  // m_c D2 w (CX/grid3D.cu:330-390)
    {
        if (!in(0, i, j, k) || fixed[at(0, i, j, k)]) return 0.0;
        double s = W(0, i + 1, j, k) - W(0, i, j, k) + W(1, i, j + 1, k) - W(1, i, j, k);
        if (dim == 3) s += W(2, i, j, k + 1) - W(2, i, j, k);
        return s;
    }
    // edge curls Cod1 w (CX/grid3D.cu:469-568)
    __device__ double ex(int i, int j, int k) const
    { return W(1, i, j, k) - W(1, i, j, k - 1) - W(2, i, j, k) + W(2, i, j - 1, k); }
    __device__ double ey(int i, int j, int k) const
    { return W(2, i, j, k) - W(2, i - 1, j, k) - W(0, i, j, k) + W(0, i, j, k - 1); }
    __device__ double ez(int i, int j, int k) const
    { return W(0, i, j, k) - W(0, i, j - 1, k) - W(1, i, j, k) + W(1, i - 1, j, k); }

    // row d of G x at face (i,j,k); the caller has checked that the face is free
    __device__ double Grow(int d, int i, int j, int k) const
    {
        const double u = x[at(1 + d, i, j, k)];
        const double a = pen[at(1 + d, i, j, k)];
        if (dim == 2) {
            const double lap = 4.0 * u - U(d, i - 1, j, 0) - U(d, i + 1, j, 0) - U(d, i, j - 1, 0) - U(d, i, j + 1, 0);
            return lap + a * u;
        }
        const int di = (d == 0), dj = (d == 1), dk = (d == 2);
        double r = gdiv(i - di, j - dj, k - dk) - gdiv(i, j, k);
        if (d == 0) r += ey(i, j, k + 1) - ey(i, j, k) + ez(i, j, k) - ez(i, j + 1, k);
        if (d == 1) r += ex(i, j, k) - ex(i, j, k + 1) + ez(i + 1, j, k) - ez(i, j, k);
        if (d == 2) r += ex(i, j + 1, k) - ex(i, j, k) + ey(i, j, k) - ey(i + 1, j, k);
        return r + a * u;
    }
};

__global__ void __launch_bounds__(NA_THREADS) k_objective_abi(AbiG a, double *part)
{
    __shared__ double sh[NA_THREADS / 64];
    const int P = a.n + 1, PZ = a.nz + (a.dim == 3 ? 1 : 0);
    const long nodes = (long)P * P * PZ;
    double acc = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nodes; t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t % PZ);
        const long tt = t / PZ;
        const int j = (int)(tt % P), i = (int)(tt / P);
        for (int d = 0; d < a.dim; d++) {
            if (!a.in(1 + d, i, j, k)) continue;
            const long q = a.at(1 + d, i, j, k);
            if (a.fixed[q]) continue;  // fixed rows of G are zero
            acc += a.x[q] * a.Grow(d, i, j, k);
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < NA_THREADS / 64; w++) s += sh[w];
        part[blockIdx.x] = s;
    }
}

// fixed-order sum of the partials, J = sum / 2
__global__ void __launch_bounds__(NA_THREADS) k_objective_sum(const double *part, int np, double *J)
{
    __shared__ double sh[NA_THREADS / 64];
    double acc = 0.0;
    for (int q = threadIdx.x; q < np; q += blockDim.x) acc += part[q];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < NA_THREADS / 64; w++) s += sh[w];
        *J = 0.5 * s;
    }
}

// device address of `scratch` on the current device, looked up once per device
double *scratch_on(int dev)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch[dev] = (double *)p;
    }
    return g_scratch[dev];
}

}  // namespace

namespace st {
const char *noctx_error() { return g_err[0] ? g_err : "null context"; }
}  // namespace st

extern "C" int st_objective_noalloc(int dim, int n, const uint8_t *d_fixed, const double *d_pen, const double *d_x,
                                    double *d_J, void *stream)
{
    if (dim != 2 && dim != 3) return fail_na(ST_EINVAL, "dim must be 2 or 3");
    if (n < 1 || n > NA_MAX_N) return fail_na(ST_EINVAL, "n must lie in [1, 1024]");
    if (!d_fixed || !d_pen || !d_x || !d_J) return fail_na(ST_EINVAL, "null array");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEV) return fail_na(ST_EHIP, "no usable device");
    double *part = scratch_on(dev);
    if (!part) return fail_na(ST_EHIP, "hipGetSymbolAddress(scratch) failed");
    AbiG a;
    a.dim = dim;
    a.n = n;
    a.nz = (dim == 3) ? n : 1;
    const long nc = (long)n * n * a.nz;
    const long nfx = (long)(n + 1) * n * a.nz, nfy = (long)n * (n + 1) * a.nz;
    a.off[0] = 0;
    a.off[1] = nc;
    a.off[2] = nc + nfx;
    a.off[3] = nc + nfx + nfy;
    a.fixed = d_fixed;
    a.pen = d_pen;
    a.x = d_x;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_objective_abi, dim3(NA_BLOCKS), dim3(NA_THREADS), 0, s, a, part);
    hipLaunchKernelGGL(k_objective_sum, dim3(1), dim3(NA_THREADS), 0, s, (const double *)part, NA_BLOCKS, d_J);
    if (hipGetLastError() != hipSuccess) return fail_na(ST_EHIP, "kernel launch failed");
    return ST_OK;
}
