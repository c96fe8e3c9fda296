// Shared definitions for the HIP Stokes/Brinkman kernels (gfx950).
//
// Device layout of one multigrid level ("padded node layout"): a level with
// nx*ny*nz cells stores every field on the padded node box
// [0,nx] x [0,ny] x [0,nz] (nz=0 box depth 1 in 2D), C-order with k fastest:
//     node(i,j,k) = (i*Py + j)*Pz + k,  Py = ny+1,  Pz >= nz+1 (1 in 2D),
// so cell (i,j,k), its three low faces u_x(i,j,k), u_y(i,j,k), u_z(i,j,k) and
// its three low edges share one index.  A vector is [p | u_x | u_y | u_z], each
// block S = (nx+1)*Py*Pz long.  Entries outside a field's valid range (e.g.
// p(nx,j,k), u_y(nx,j,k), and the columns k > nz) are padding and always hold 0,
// which is the zero extension the reference kernels implement with their
// validity guards (CX/grid3D.cu:124-146).
// On the levels the tiled kernels run (3D, nz >= 64) the row pitch Pz is nz+1
// rounded up to ST_ROW_ALIGN values, so that every row of every field starts on
// a 128-byte line (FP64) and a wave's 64-value row segment k0..k0+63 covers
// whole lines: with Pz = 257 the same streams run about 20% slower at 256^3.
//
// Per-node flag byte: bit0 = cell free, bit1..3 = face x/y/z free.  Padding and
// fixed DOFs have the bit clear.  The face "vol" of the reference descriptor
// equals the free mask on every level here (ST/StokeFlowDescriptor3D.cu:37-50).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef ST_ROW_ALIGN
#define ST_ROW_ALIGN 16
#endif

namespace st {

struct Lvl {
    int dim, nx, ny, nz;   // cells (nz = 1 in 2D)
    int Px, Py, Pz;        // padded node box
    long S;                // nodes per field
    int nfield;            // 1 + dim
    int tiled;             // may run the 2.5D tiled kernels (when its shape allows them)
};

static inline Lvl make_lvl(int dim, int nx, int ny, int nz)
{
    Lvl g;
    g.dim = dim; g.nx = nx; g.ny = ny; g.nz = (dim == 3) ? nz : 1;
    g.Px = nx + 1; g.Py = ny + 1; g.Pz = (dim == 3) ? nz + 1 : 1;
    if (dim == 3 && nz >= 64) g.Pz = (g.Pz + ST_ROW_ALIGN - 1) / ST_ROW_ALIGN * ST_ROW_ALIGN;
    g.S = (long)g.Px * g.Py * g.Pz;
    g.nfield = 1 + dim;
    g.tiled = 1;
    return g;
}

// bfloat16 storage type (round-to-nearest-even), computed in float
struct bf16 {
    uint16_t v;
};
__host__ __device__ inline float to_f(bf16 h)
{
    union { uint32_t u; float f; } c;
    c.u = ((uint32_t)h.v) << 16;
    return c.f;
}
__host__ __device__ inline bf16 to_bf16(float f)
{
    union { uint32_t u; float f; } c;
    c.f = f;
    uint32_t u = c.u;
    uint32_t r = ((u >> 16) & 1u) + 0x7fffu;
    bf16 h;
    h.v = (uint16_t)((u + r) >> 16);
    return h;
}

template <typename TS> struct Cvt;
template <> struct Cvt<double> {
    typedef double C;
    __host__ __device__ static double ld(double v) { return v; }
    __host__ __device__ static double st(double v) { return v; }
};
template <> struct Cvt<float> {
    typedef float C;
    __host__ __device__ static float ld(float v) { return v; }
    __host__ __device__ static float st(float v) { return v; }
};
template <> struct Cvt<bf16> {
    typedef float C;
    __host__ __device__ static float ld(bf16 v) { return to_f(v); }
    __host__ __device__ static bf16 st(float v) { return to_bf16(v); }
};

}  // namespace st
