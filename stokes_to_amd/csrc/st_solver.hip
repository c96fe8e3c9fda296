// Host orchestration of the HIP Stokes/Brinkman solver and the C ABI of
// libstokes_hip.so (declared in include/stokes_to.h).
//
// Structure (MI355X-first, one device, one stream, all state resident in HBM):
//  - level hierarchy n, n/2, ..., 4 in the padded node layout (st_common.h);
//    the FP64 level data (flags, face alpha) is always kept, the multigrid
//    preconditioner runs in FP64, FP32 or BF16 storage;
//  - FGMRES(m) in FP64 on the device, right-preconditioned by one V-cycle, with
//    classical Gram-Schmidt as two fused passes (multi-dot, multi-axpy + norm),
//    stopped on the true FP64 residual.  Only the (m+1) Hessenberg scalars per
//    iteration cross PCIe (the reference copies two full vectors per iteration
//    and orthogonalises on the host, CX/GMRES.cpp:46-95).
#include "../../include/stokes_to.h"
#include "st_kernels.h"
#include "st_mma.h"
#include "st_net.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

using namespace st;

namespace st {
const char *noctx_error();  // last failure of a context-free entry point (st_noalloc.hip)
}

namespace {

template <typename TS> struct Hier {
    std::vector<TS *> al, b, x, t, r;
    TailArgs<TS> *tail = nullptr;  // device copy of the fused coarse V-cycle's arguments
    bool tail_lds = false;         // its levels fit the LDS pool
};

struct Net;  // learned predictor (st_net.hip)

}  // namespace

struct st_ctx {
    st_params prm;
    int dim = 3, n = 0, L = 0;
    std::vector<Lvl> lv;
    std::vector<uint8_t *> fl;
    Hier<double> h64;
    Hier<float> h32;
    Hier<bf16> hbf;
    // coarsest dense inverse
    long *d_dofs = nullptr;
    int nd = 0;
    double *d_minv = nullptr;
    int *d_perm = nullptr, *d_info = nullptr;
    // finest FP64 state (padded)
    long ndof = 0, ncell = 0, nvec = 0;  // ABI sizes and padded vector length
    double *X = nullptr, *B = nullptr, *XD = nullptr, *BODY = nullptr, *LAM = nullptr, *BA = nullptr, *ZERO = nullptr;
    double *Wv = nullptr, *V = nullptr, *Zb = nullptr;
    double *part = nullptr, *red = nullptr, *h_dev = nullptr, *sig2_dev = nullptr;
    int npart = 0;
    // design
    double *rho = nullptr, *rhof = nullptr, *dj = nullptr, *dc = nullptr, *rn = nullptr, *fW = nullptr, *dv = nullptr;
    double *tmpc = nullptr;  // cell scratch of the OC update
    MmaDev *mma = nullptr;   // MMA state (created on first use)
    double *mlo = nullptr, *mhi = nullptr, *mdg = nullptr;
    // staging
    double *abi_a = nullptr, *abi_b = nullptr;
    uint8_t *abi_u8 = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    bool geom = false, alpha = false, rhs = false, dirty = true, have_x = false;
    std::string err;
    std::vector<void *> allocs;
    double last_setup_ms = 0.0;
    NetImpl *net = nullptr;     // learned predictor (created on first use)
    int lc = -1;                // network level (n/4)
    const double *rho_used = nullptr;  // density the current alpha was computed from
    double *xdc = nullptr, *xcp = nullptr, *xcn = nullptr;  // level-lc Dirichlet state, two predictions
    double *rhop = nullptr;     // design of the solution held in X (for the learned correction)
    bool have_rhop = false, have_lam = false, have_grad = false;
    int learned_rejects = 0;  // consecutive rejections of the network's candidate
    long learned_calls = 0;
    // previous solutions of the current sequence of forward (XH) and adjoint (LH) solves,
    // for the extrapolated start: [0] = the one before the held solution, [1] = the one
    // before that; nxh / nlh of them valid
    double *XH[2] = {nullptr, nullptr}, *LH[2] = {nullptr, nullptr};
    int nxh = 0, nlh = 0;
    // the FP32 path's first residual of the next solve, already in u_0 (set by the
    // extrapolated start when it computed it; consumed or dropped by the next solve)
    bool r0_ready = false;
    double r0_beta = 0.0;
    // differences of consecutive solutions of the forward (DX) and adjoint (DL)
    // sequences in FP32, newest first, for the minimal-residual start (warm_order 3)
    float *DX[3] = {nullptr, nullptr, nullptr}, *DL[3] = {nullptr, nullptr, nullptr};
    // multiplier of the last OC update (0: none), where the next search starts
    double oc_lambda = 0.0;
    // dv = dV/drho = H^T(1/(n W)) depends on the grid and the filter only: formed once
    bool dv_ready = false;
    // ||B||^2 of the forward right-hand side (it changes only with st_set_rhs / geometry)
    double bn2 = 0.0;
    bool bn2_ok = false;
    // kernel timing: event pairs around finest-level launches
    bool prof = false;
    std::vector<hipEvent_t> evpool[ST_PROF_NKIND];
    size_t evused[ST_PROF_NKIND] = {};
    double prof_ms[ST_PROF_NKIND] = {};
    long prof_n[ST_PROF_NKIND] = {};
};

static void prof_mark(st_ctx *c, int kind)
{
    if (!c->prof) return;
    auto &pool = c->evpool[kind];
    if (c->evused[kind] >= pool.size()) {
        hipEvent_t e;
        hipEventCreate(&e);
        pool.push_back(e);
    }
    hipEventRecord(pool[c->evused[kind]++], c->s);
}

// called after the stream is synchronised: fold recorded pairs into totals
static void prof_collect(st_ctx *c)
{
    for (int k = 0; k < ST_PROF_NKIND; k++) {
        for (size_t i = 0; i + 1 < c->evused[k]; i += 2) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, c->evpool[k][i], c->evpool[k][i + 1]);
            c->prof_ms[k] += ms;
            c->prof_n[k] += 1;
        }
        c->evused[k] = 0;
    }
}

#define CK(call)                                                                     \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            c->err = std::string(#call) + ": " + hipGetErrorString(e_);             \
            return ST_EHIP;                                                          \
        }                                                                            \
    } while (0)
#define CKR(expr)                 \
    do {                          \
        int r_ = (expr);          \
        if (r_ != ST_OK) return r_; \
    } while (0)

static int fail(st_ctx *c, int code, const std::string &m)
{
    if (c) c->err = m;
    return code;
}

template <typename T> static int dalloc(st_ctx *c, T **p, long n)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, (size_t)(n > 0 ? n : 1) * sizeof(T));
    if (e != hipSuccess) {
        c->err = std::string("hipMalloc(") + std::to_string((long long)n * sizeof(T)) + " B): " + hipGetErrorString(e);
        return ST_EHIP;
    }
    // padding entries (outside a field's valid range) start at 0 and no kernel writes
    // anything else there
    e = hipMemsetAsync(q, 0, (size_t)(n > 0 ? n : 1) * sizeof(T), c->s);
    if (e != hipSuccess) {
        hipFree(q);
        c->err = std::string("hipMemsetAsync: ") + hipGetErrorString(e);
        return ST_EHIP;
    }
    c->allocs.push_back(q);
    *p = (T *)q;
    return ST_OK;
}

static long abi_dofs(int dim, int n)
{
    if (dim == 3) return (long)n * n * n + 3L * (n + 1) * n * n;
    return (long)n * n + 2L * (n + 1) * n;
}

// ---------------------------------------------------------------- helpers

static int sum_partials(st_ctx *c, int np, int nvec, double *out_dev)
{
    launch_sum(np, nvec, c->part, out_dev, c->s);
    return ST_OK;
}

static int dot64(st_ctx *c, const double *a, const double *b, long n, double *res)
{
    launch_dot<double, double>(n, a, b, c->part, c->s);
    launch_sum((int)blas_blocks(n), 1, c->part, c->red, c->s);
    CK(hipMemcpyAsync(res, c->red, sizeof(double), hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    return ST_OK;
}

// -------------------------------------------------------------- hierarchy

template <typename TS> static Hier<TS> &hier(st_ctx *c);
template <> Hier<double> &hier<double>(st_ctx *c) { return c->h64; }
template <> Hier<float> &hier<float>(st_ctx *c) { return c->h32; }
template <> Hier<bf16> &hier<bf16>(st_ctx *c) { return c->hbf; }

template <typename TS> static int alloc_hier(st_ctx *c, bool level0_vectors)
{
    Hier<TS> &h = hier<TS>(c);
    h.al.assign(c->L, nullptr);
    h.b.assign(c->L, nullptr);
    h.x.assign(c->L, nullptr);
    h.t.assign(c->L, nullptr);
    h.r.assign(c->L, nullptr);
    for (int l = 0; l < c->L; l++) {
        const Lvl &g = c->lv[l];
        long nv = (long)g.nfield * g.S;
        if (l > 0 || !std::is_same<TS, double>::value) CKR(dalloc(c, &h.al[l], (long)g.dim * g.S));
        if (l > 0 || level0_vectors) {
            CKR(dalloc(c, &h.b[l], nv));
            CKR(dalloc(c, &h.x[l], nv));
        }
        CKR(dalloc(c, &h.t[l], nv));
        CKR(dalloc(c, &h.r[l], nv));
    }
    return ST_OK;
}

template <typename TS> static void convert_alpha(st_ctx *c)
{
    Hier<TS> &h = hier<TS>(c);
    for (int l = 0; l < c->L; l++) {
        const Lvl &g = c->lv[l];
        launch_axpby<double, TS>((long)g.dim * g.S, 1.0, c->h64.al[l], 0.0, h.al[l], c->s);
    }
}

static int setup_hierarchy(st_ctx *c)
{
    if (!c->geom) return fail(c, ST_ESTATE, "geometry not set (st_set_geometry)");
    if (!c->alpha) return fail(c, ST_ESTATE, "penalty/design not set (st_set_penalty or st_set_design)");
    CK(hipEventRecord(c->e2, c->s));
    for (int l = 0; l + 1 < c->L; l++)
        launch_coarsen<double, double>(c->lv[l], c->lv[l + 1], c->fl[l], c->h64.al[l], c->fl[l + 1], c->h64.al[l + 1],
                                       c->s);
    if (c->prm.precision == ST_FP32) convert_alpha<float>(c);
    if (c->prm.precision == ST_BF16) convert_alpha<bf16>(c);
    const int Lc = c->L - 1;
    launch_probe<double>(c->lv[Lc], c->fl[Lc], c->h64.al[Lc], c->d_dofs, c->nd, c->prm.closed, c->d_minv, c->s);
    launch_gauss_jordan(c->nd, c->d_minv, c->d_perm, c->d_info, c->s);
    int info = 0;
    CK(hipMemcpyAsync(&info, c->d_info, sizeof(int), hipMemcpyDeviceToHost, c->s));
    hipEvent_t e3;
    CK(hipEventCreate(&e3));
    CK(hipEventRecord(e3, c->s));
    CK(hipStreamSynchronize(c->s));
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->e2, e3);
    hipEventDestroy(e3);
    c->last_setup_ms = ms;
    if (info < 0)
        return fail(c, ST_ENAN, "non-finite value in the coarsest-level matrix (a NaN or Inf in the design or penalty)");
    if (info != 0)
        return fail(c, ST_ESINGULAR,
                    "coarsest-level matrix is singular at pivot " + std::to_string(info) +
                        " (closed domain? set params.closed=1)");
    c->dirty = false;
    return ST_OK;
}

// ---------------------------------------------------------------- V-cycle

template <typename TS> static void smooth(st_ctx *c, int l, const TS *b, TS *x, int nu, int first, bool zero)
{
    Hier<TS> &h = hier<TS>(c);
    const Lvl &g = c->lv[l];
    TS *t = h.t[l];
    const float om = (float)c->prm.omega;
    const bool pr = (l == 0);
    if constexpr (!std::is_same<TS, double>::value) {
        if (c->prm.fused_sweep && march_ok(g)) {
            // FP32 arithmetic on a tiled level: whole sweeps in one pass (st_sweep.hip),
            // ping-pong x -> t -> x; a V-cycle's first sweep starts from x = 0 with the
            // per-cell zero-start pass and one colour pass (it reads no x)
            TS *src = x, *dst = t;
            int sw = 0;
            if (zero) {
                if (pr) prof_mark(c, ST_PROF_SMOOTH0);
                launch_smooth<TS>(g, c->fl[l], h.al[l], b, x, x, first, om, true, c->s);
                if (pr) { prof_mark(c, ST_PROF_SMOOTH0); prof_mark(c, ST_PROF_SMOOTH); }
                launch_smooth<TS>(g, c->fl[l], h.al[l], b, x, t, first ^ 1, om, false, c->s);
                if (pr) prof_mark(c, ST_PROF_SMOOTH);
                src = t;
                dst = x;
                sw = 1;
            }
            for (; sw < nu; sw++) {
                if (pr) prof_mark(c, ST_PROF_SWEEP);
                march_sweep<TS>(g, c->fl[l], h.al[l], b, src, dst, first, om, c->s);
                if (pr) prof_mark(c, ST_PROF_SWEEP);
                std::swap(src, dst);
            }
            if (src != x) launch_axpby<TS, TS>((long)g.nfield * g.S, 1.0, src, 0.0, x, c->s);
            return;
        }
    }
    for (int sw = 0; sw < nu; sw++) {
        // the pass from x = 0 is another kernel (k_smooth0_pairs) with other bytes
        const int k0 = (zero && sw == 0) ? ST_PROF_SMOOTH0 : ST_PROF_SMOOTH;
        if (pr) prof_mark(c, k0);
        launch_smooth<TS>(g, c->fl[l], h.al[l], b, x, t, first, om, zero && sw == 0, c->s);
        if (pr) { prof_mark(c, k0); prof_mark(c, ST_PROF_SMOOTH); }
        launch_smooth<TS>(g, c->fl[l], h.al[l], b, t, x, first ^ 1, om, false, c->s);
        if (pr) prof_mark(c, ST_PROF_SMOOTH);
    }
}

// first level handled by the fused coarse V-cycle: the largest level with at most
// ST_TAIL_MAX_CELLS cells (one workgroup), or L when there is none
static int tail_level(const st_ctx *c)
{
    for (int l = 0; l < c->L; l++) {
        const Lvl &g = c->lv[l];
        if ((long)g.nx * g.ny * g.nz <= ST_TAIL_MAX_CELLS) return l;
    }
    return c->L;
}

template <typename TS> static void vcycle(st_ctx *c, const TS *b0, TS *x0)
{
    Hier<TS> &h = hier<TS>(c);
    const int L = c->L;
    const int lt = tail_level(c);
    for (int l = 0; l + 1 < L && l < lt; l++) {
        const Lvl &g = c->lv[l];
        const TS *b = (l == 0) ? b0 : h.b[l];
        TS *x = (l == 0) ? x0 : h.x[l];
        if (c->prm.nu_pre > 0) smooth<TS>(c, l, b, x, c->prm.nu_pre, 0, true);
        else launch_fill<TS>((long)g.nfield * g.S, x, 0.0, c->s);
        if (l == 0) prof_mark(c, ST_PROF_RESIDUAL);
        launch_residual<TS>(g, c->fl[l], h.al[l], b, x, h.r[l], nullptr, c->s);
        if (l == 0) prof_mark(c, ST_PROF_RESIDUAL);
        launch_restrict<TS>(g, c->lv[l + 1], c->fl[l + 1], h.r[l], h.b[l + 1], c->s);
    }
    if (lt < L) {
        if (!h.tail) {  // built once: the level buffers do not move
            TailArgs<TS> a;
            std::memset(&a, 0, sizeof(a));
            a.l0 = lt;
            a.L = L;
            a.nu_pre = c->prm.nu_pre;
            a.nu_post = c->prm.nu_post;
            a.nd = c->nd;
            a.dim = c->dim;
            a.omega = (float)c->prm.omega;
            for (int l = lt; l < L; l++) {
                a.g[l] = c->lv[l];
                a.fl[l] = c->fl[l];
                a.al[l] = h.al[l];
                a.b[l] = h.b[l];
                a.x[l] = h.x[l];
                a.t[l] = h.t[l];
                a.r[l] = h.r[l];
            }
            a.dofs = c->d_dofs;
            a.minv = c->d_minv;
            // LDS layout of the staged variant (16-byte aligned buffers)
            {
                long off = 0;
                auto take = [&](long bytes) {
                    const long o = off;
                    off += (bytes + 15) / 16 * 16;
                    return (int)o;
                };
                for (int l = lt; l < L; l++) {
                    const Lvl &g = c->lv[l];
                    const long nv = (long)g.nfield * g.S;
                    a.lds.fl[l] = take(g.S);
                    a.lds.al[l] = take((long)g.dim * g.S * sizeof(TS));
                    a.lds.b[l] = take(nv * sizeof(TS));
                    a.lds.x[l] = take(nv * sizeof(TS));
                    a.lds.t[l] = take(nv * sizeof(TS));
                    a.lds.r[l] = take(nv * sizeof(TS));
                }
                a.lds.bs = take((long)c->nd * sizeof(double));
                a.lds.bytes = (int)off;
                a.lds.ok = off <= ST_TAIL_LDS_BYTES;
            }
            h.tail_lds = a.lds.ok != 0;
            void *p = nullptr;
            if (hipMalloc(&p, sizeof(a)) == hipSuccess && hipMemcpy(p, &a, sizeof(a), hipMemcpyHostToDevice) == hipSuccess) {
                c->allocs.push_back(p);
                h.tail = (TailArgs<TS> *)p;
            }
        }
        if (h.tail) {
            launch_vcycle_tail<TS>(h.tail, c->dim, lt == 0 ? b0 : h.b[lt], lt == 0 ? x0 : h.x[lt], c->nd, h.tail_lds,
                                   c->s);
        } else {  // allocation failed: per-level kernels (same arithmetic)
            for (int l = lt; l + 1 < L; l++) {
                const Lvl &g = c->lv[l];
                const TS *b = (l == 0) ? b0 : h.b[l];
                TS *x = (l == 0) ? x0 : h.x[l];
                if (c->prm.nu_pre > 0) smooth<TS>(c, l, b, x, c->prm.nu_pre, 0, true);
                else launch_fill<TS>((long)g.nfield * g.S, x, 0.0, c->s);
                launch_residual<TS>(g, c->fl[l], h.al[l], b, x, h.r[l], nullptr, c->s);
                launch_restrict<TS>(g, c->lv[l + 1], c->fl[l + 1], h.r[l], h.b[l + 1], c->s);
            }
            launch_coarse_solve<TS>(c->nd, c->d_dofs, c->d_minv, (L - 1 == 0) ? b0 : h.b[L - 1],
                                    (L - 1 == 0) ? x0 : h.x[L - 1], c->s);
            for (int l = L - 2; l >= lt; l--) {
                const TS *b = (l == 0) ? b0 : h.b[l];
                TS *x = (l == 0) ? x0 : h.x[l];
                launch_prolong<TS>(c->lv[l], c->lv[l + 1], c->fl[l], h.x[l + 1], x, c->s);
                if (c->prm.nu_post > 0) smooth<TS>(c, l, b, x, c->prm.nu_post, 1, false);
            }
        }
    } else {
        launch_coarse_solve<TS>(c->nd, c->d_dofs, c->d_minv, h.b[L - 1], h.x[L - 1], c->s);
    }
    for (int l = std::min(lt, L - 1) - 1; l >= 0; l--) {
        const TS *b = (l == 0) ? b0 : h.b[l];
        TS *x = (l == 0) ? x0 : h.x[l];
        launch_prolong<TS>(c->lv[l], c->lv[l + 1], c->fl[l], h.x[l + 1], x, c->s);
        if (c->prm.nu_post > 0) smooth<TS>(c, l, b, x, c->prm.nu_post, 1, false);
    }
}

// z = M v.  The FP64 path keeps FP64 Krylov vectors; the mixed-precision paths
// keep them in FP32 (the V-cycle works on them directly in FP32, through a BF16
// copy in the BF16 path)
static void precondition(st_ctx *c, const double *v, double *z) { vcycle<double>(c, v, z); }
static void precondition(st_ctx *c, const float *v, float *z)
{
    if (c->prm.precision == ST_FP32) {
        vcycle<float>(c, v, z);
    } else {
        const long nv = c->nvec;
        launch_axpby<float, bf16>(nv, 1.0, v, 0.0, c->hbf.b[0], c->s);
        vcycle<bf16>(c, c->hbf.b[0], c->hbf.x[0]);
        launch_axpby<bf16, float>(nv, 1.0, c->hbf.x[0], 0.0, z, c->s);
    }
}

// w = A z inside FGMRES: the FP32 operator for the FP32 path
static void inner_apply(st_ctx *c, const float *z, float *w)
{
    launch_apply<float>(c->lv[0], c->fl[0], c->h32.al[0], z, w, c->s);
}

// w = A z with the FP64 operator
static void apply64(st_ctx *c, const double *z, double *w)
{
    launch_apply<double>(c->lv[0], c->fl[0], c->h64.al[0], z, w, c->s);
}
static void apply64(st_ctx *c, const float *z, double *w)
{
    const Lvl &g = c->lv[0];
    if (march_ok(g)) {
        march_apply_mixed(g, c->fl[0], c->h64.al[0], z, w, c->s);
    } else {
        launch_axpby<float, double>(c->nvec, 1.0, z, 0.0, c->h64.t[0], c->s);
        launch_apply<double>(g, c->fl[0], c->h64.al[0], c->h64.t[0], w, c->s);
    }
}

static void inner_apply(st_ctx *c, const double *z, double *w) { apply64(c, z, w); }
static void inner_apply(st_ctx *c, const float *z, double *w) { apply64(c, z, w); }

// ----------------------------------------------------------------- FGMRES

// solves A x = b (padded FP64, device), x holds the initial guess on entry.
// TV: storage of the Krylov vectors V and Z; TW: precision of the inner operator
// w = A z and of w.  (double, double): the plain FP64 method.  (float, float): the
// FP32 path, an FP32 inner FGMRES -- operator, basis and preconditioner in FP32, inner
// products accumulated in FP64 -- inside FP64 iterative refinement: x, its update and
// the residual that starts every restart cycle and decides convergence are FP64.
// (float, double): the BF16 path (FP64 operator on FP32 Krylov vectors).  With an
// FP32 operator the Arnoldi estimate of the residual is only good to ~1e-6 of the
// cycle's starting residual, so a cycle also ends there and the FP64 residual of the
// restart takes over.
template <typename TV, typename TW> static int fgmres(st_ctx *c, const double *b, double *x, st_stats *st, bool use_r0)
{
    // LAZY (Krylov vectors and operator in one precision): the basis is stored
    // unnormalised, v_i = sig_i u_i, and A M u_j is written straight into the slot of
    // u_{j+1}, so no pass scales and copies the Arnoldi vector.  With c_i = u_i . A M u_j:
    // h_ij = sig_i sig_j c_i, u_{j+1} = A M u_j - sum sig_i^2 c_i u_i,
    // h_{j+1,j} = sig_j |u_{j+1}|, sig_{j+1} = 1 / |u_{j+1}|; z_j = M u_j enters the
    // solution update with the weight sig_j.
    constexpr bool LAZY = std::is_same<TV, TW>::value;
    TV *const Vb = reinterpret_cast<TV *>(c->V);
    TV *const Zs = reinterpret_cast<TV *>(c->Zb);
    TW *const Wk = reinterpret_cast<TW *>(c->Wv);
    const double floor_rel = std::is_same<TW, float>::value ? 1e-6 : 0.0;
    const Lvl &g = c->lv[0];
    const long nv = c->nvec;
    const int m = c->prm.restart;
    const double rtol = c->prm.rtol;
    double bn2 = 0.0;
    if (b == c->B && c->bn2_ok) {
        bn2 = c->bn2;
    } else {
        CKR(dot64(c, b, b, nv, &bn2));
        if (b == c->B) {
            c->bn2 = bn2;
            c->bn2_ok = true;
        }
    }
    double bn = std::sqrt(bn2);
    if (!std::isfinite(bn)) return fail(c, ST_ENAN, "non-finite right-hand side");
    if (bn == 0.0) bn = 1.0;
    std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), gv(m + 1), y(m), hh(m + 2);
    std::vector<double> sig(m + 1, 1.0), sig2(m + 1, 1.0);
    int it = 0;
    double rr = 0.0, rr0 = -1.0;
    const int rb = (int)residual_blocks(g);
    for (;;) {
        double *r0 = std::is_same<TV, double>::value ? c->V : c->Wv;  // FP64 residual
        // FP32 path on a tiled level: the FP64 residual goes straight into u_0 in FP32
        const bool r_f32 = LAZY && std::is_same<TV, float>::value && march_ok(g);
        double beta;
        if (r_f32 && use_r0 && it == 0) {  // the start's residual is already in u_0
            beta = c->r0_beta;
        } else {
            if (r_f32) march_residual_f32(g, c->fl[0], c->h64.al[0], b, x, reinterpret_cast<float *>(Vb), c->part, c->s);
            else launch_residual<double>(g, c->fl[0], c->h64.al[0], b, x, r0, c->part, c->s);
            launch_sum(rb, 1, c->part, c->red, c->s);
            double r2 = 0.0;
            CK(hipMemcpyAsync(&r2, c->red, sizeof(double), hipMemcpyDeviceToHost, c->s));
            CK(hipStreamSynchronize(c->s));
            beta = std::sqrt(r2);
        }
        if (!std::isfinite(beta)) return fail(c, ST_ENAN, "non-finite residual");
        rr = beta / bn;
        if (rr0 < 0) rr0 = rr;
        if (rr <= rtol || it >= c->prm.max_iter) break;
        if (LAZY) {
            sig[0] = 1.0 / beta;
            sig2[0] = sig[0] * sig[0];
            CK(hipMemcpyAsync(c->sig2_dev, sig2.data(), sizeof(double), hipMemcpyHostToDevice, c->s));
            if (!std::is_same<TV, double>::value && !r_f32) launch_axpby<double, TV>(nv, 1.0, r0, 0.0, Vb, c->s);
        } else {
            launch_axpby<double, TV>(nv, 1.0 / beta, r0, 0.0, Vb, c->s);
        }
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = beta;
        const double cycle_tol = std::max(rtol * bn, floor_rel * beta);
        int j = 0;
        for (; j < m && it < c->prm.max_iter; j++) {
            TV *vj = Vb + (long)j * nv, *zj = Zs + (long)j * nv;
            TW *w = LAZY ? reinterpret_cast<TW *>(Vb + (long)(j + 1) * nv) : Wk;
            precondition(c, vj, zj);
            prof_mark(c, ST_PROF_APPLY);
            inner_apply(c, zj, w);
            prof_mark(c, ST_PROF_APPLY);
            const unsigned nbk = blas_blocks(nv);
            launch_multidot<TW, TV>(nv, w, Vb, nv, j + 1, c->part, c->s);
            launch_sum((int)nbk, j + 1, c->part, c->h_dev, c->s);
            launch_multiaxpy<TW, TV>(nv, w, Vb, nv, j + 1, c->h_dev, LAZY ? c->sig2_dev : nullptr, c->part, c->s);
            launch_sum((int)nbk, 1, c->part, c->h_dev + (j + 1), c->s);
            CK(hipMemcpyAsync(hh.data(), c->h_dev, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, c->s));
            CK(hipStreamSynchronize(c->s));
            const double un = std::sqrt(hh[j + 1]);  // |u_{j+1}| (LAZY) or |w|
            if (!std::isfinite(un)) return fail(c, ST_ENAN, "non-finite Arnoldi vector");
            const double hn = LAZY ? sig[j] * un : un;
            for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = LAZY ? sig[i] * sig[j] * hh[i] : hh[i];
            H[(size_t)(j + 1) * m + j] = hn;
            if (LAZY) {
                sig[j + 1] = un > 0 ? 1.0 / un : 0.0;
                if (un > 1e12 || (un > 0 && un < 1e-12)) {  // keep the stored vectors in range
                    launch_axpby<TW, TV>(nv, 1.0 / un, w, 0.0, reinterpret_cast<TV *>(w), c->s);
                    sig[j + 1] = 1.0;
                }
                sig2[j + 1] = sig[j + 1] * sig[j + 1];
                CK(hipMemcpyAsync(c->sig2_dev + (j + 1), &sig2[j + 1], sizeof(double), hipMemcpyHostToDevice, c->s));
            } else if (hn > 0) {
                launch_axpby<TW, TV>(nv, 1.0 / hn, Wk, 0.0, Vb + (long)(j + 1) * nv, c->s);
            }
            for (int i = 0; i < j; i++) {
                double a = H[(size_t)i * m + j], bb = H[(size_t)(i + 1) * m + j];
                H[(size_t)i * m + j] = cs[i] * a + sn[i] * bb;
                H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * bb;
            }
            double a = H[(size_t)j * m + j], bb = H[(size_t)(j + 1) * m + j];
            double den = std::sqrt(a * a + bb * bb);
            cs[j] = a / den;
            sn[j] = bb / den;
            H[(size_t)j * m + j] = den;
            H[(size_t)(j + 1) * m + j] = 0.0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            it++;
            if (std::fabs(gv[j + 1]) <= cycle_tol || hn == 0.0) {
                j++;
                break;
            }
        }
        for (int i = j - 1; i >= 0; i--) {
            double sacc = gv[i];
            for (int k = i + 1; k < j; k++) sacc -= H[(size_t)i * m + k] * y[k];
            y[i] = sacc / H[(size_t)i * m + i];
        }
        for (int i = 0; i < j; i++) y[i] = LAZY ? -y[i] * sig[i] : -y[i];
        CK(hipMemcpyAsync(c->h_dev, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, c->s));
        launch_multiaxpy<double, TV>(nv, x, Zs, nv, j, c->h_dev, nullptr, nullptr, c->s);
    }
    if (st) {
        st->iters = it;
        st->learned_used = 0;
        st->relres = rr;
        st->relres0 = rr0;
        st->converged = rr <= rtol;
    }
    return rr <= rtol ? ST_OK : ST_ENOCONV;
}

// ------------------------------------------------------------- ABI core

static int stage_in(st_ctx *c, const double *h, long n)
{
    CK(hipMemcpyAsync(c->abi_a, h, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
    return ST_OK;
}
static int stage_out(st_ctx *c, double *h, long n)
{
    CK(hipMemcpyAsync(h, c->abi_b, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    return ST_OK;
}

static int check_finite(st_ctx *c, const double *h, long n, const char *what)
{
    for (long q = 0; q < n; q++)
        if (!std::isfinite(h[q])) return fail(c, ST_ENAN, std::string("non-finite value in ") + what);
    return ST_OK;
}

// one smoother colour step in FP32 / BF16 storage (test entry st_mg_smooth_mp)
template <typename TS> static int mg_smooth_lowp(st_ctx *c, int l, int color, const double *b, const double *xin,
                                                 double *xout)
{
    const Lvl &g = c->lv[l];
    const long nd = abi_dofs(c->dim, g.nx);
    // scratch only: the held forward / adjoint solutions (X, LAM) are never touched
    TS *bb = reinterpret_cast<TS *>(c->BA), *xi = reinterpret_cast<TS *>(c->Wv), *xo = reinterpret_cast<TS *>(c->V);
    CKR(stage_in(c, b, nd));
    launch_pack<TS>(g, c->abi_a, bb, 0, c->fl[l], 1.0, c->s);
    CKR(stage_in(c, xin, nd));
    launch_pack<TS>(g, c->abi_a, xi, 0, c->fl[l], 1.0, c->s);
    launch_fill<TS>((long)g.nfield * g.S, xo, 0.0, c->s);
    launch_smooth<TS>(g, c->fl[l], hier<TS>(c).al[l], bb, xi, xo, color, (float)c->prm.omega, false, c->s);
    launch_unpack<TS>(g, xo, c->abi_b, c->s);
    return stage_out(c, xout, nd);
}

extern "C" {

void st_default_params(st_params *p, int dim, int n)
{
    std::memset(p, 0, sizeof(*p));
    p->dim = dim;
    p->n = n;
    p->alpha_max = 1e4;
    p->alpha_min = 0.0;
    p->q = 0.1;
    p->power = 1.0;
    p->alpha_scale = (n > 0) ? 1.0 / ((double)n * n) : 1.0;
    p->precision = ST_FP64;
    p->rtol = 1e-8;
    p->max_iter = 200;
    p->restart = 30;
    p->nu_pre = 1;
    p->nu_post = 2;
    p->omega = 1.0;
    p->closed = 0;
    p->filter_radius = 1.5;
    p->optimizer = ST_OPT_OC;
    p->fused_sweep = 0;
    p->warm_order = 3;
}

int st_create(st_ctx **out, const st_params *p)
{
    if (!out || !p) return ST_EINVAL;
    *out = nullptr;
    if (p->dim != 2 && p->dim != 3) return ST_EINVAL;
    int n = p->n;
    if (n < 8 || (n % 4) != 0) return ST_EINVAL;
    {
        int t = n / 4;
        while (t % 2 == 0) t /= 2;
        if (t != 1) return ST_EINVAL;  // n = 4 * 2^k
    }
    if (p->restart < 1 || p->restart > 32 || p->max_iter < 1 || !(p->rtol > 0)) return ST_EINVAL;
    if (p->precision < ST_FP64 || p->precision > ST_BF16) return ST_EINVAL;
    if (p->warm_order < 0 || p->warm_order > 3) return ST_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return ST_EHIP;
    st_ctx *c = new st_ctx();
    c->prm = *p;
    c->dim = p->dim;
    c->n = n;
    if (hipStreamCreate(&c->s) != hipSuccess) {
        delete c;
        return ST_EHIP;
    }
    hipEventCreate(&c->e0);
    hipEventCreate(&c->e1);
    hipEventCreate(&c->e2);
    int m = n;
    for (;;) {
        c->lv.push_back(make_lvl(c->dim, m, m, m));
        if (m % 2 != 0 || m / 2 < 2) break;  // coarsest level 2^dim cells
        m /= 2;
    }
    // a coarse level below 128 cells a side runs the per-cell gather kernels: with
    // few planes per tile the marching kernels are latency-bound and slower there
    for (size_t l = 1; l < c->lv.size(); l++)
        if (c->lv[l].nz < 128) c->lv[l].tiled = 0;
    c->L = (int)c->lv.size();
    if (c->L > ST_MAX_LEVELS) {
        st_destroy(c);
        return ST_EINVAL;
    }
    int rc = ST_OK;
    c->fl.assign(c->L, nullptr);
    for (int l = 0; l < c->L && rc == ST_OK; l++) rc = dalloc(c, &c->fl[l], c->lv[l].S);
    const Lvl &g0 = c->lv[0];
    c->nvec = (long)g0.nfield * g0.S;
    c->ndof = abi_dofs(c->dim, n);
    c->ncell = (c->dim == 3) ? (long)n * n * n : (long)n * n;
    const int mr = p->restart;
    if (rc == ST_OK) rc = dalloc(c, &c->h64.al.emplace_back(), (long)c->dim * g0.S);
    if (rc == ST_OK) {
        double *al0 = c->h64.al[0];
        c->h64.al.clear();
        rc = alloc_hier<double>(c, false);
        if (rc == ST_OK) c->h64.al[0] = al0;
    }
    if (rc == ST_OK && p->precision == ST_FP32) rc = alloc_hier<float>(c, false);
    if (rc == ST_OK && p->precision == ST_BF16) rc = alloc_hier<bf16>(c, true);
    if (rc == ST_OK) rc = dalloc(c, &c->X, c->nvec);
    if (rc == ST_OK) rc = dalloc(c, &c->B, c->nvec);
    if (rc == ST_OK) rc = dalloc(c, &c->XD, c->nvec);
    if (rc == ST_OK) rc = dalloc(c, &c->BODY, c->nvec);
    if (rc == ST_OK) rc = dalloc(c, &c->LAM, c->nvec);
    if (rc == ST_OK) rc = dalloc(c, &c->BA, c->nvec);
    if (rc == ST_OK) rc = dalloc(c, &c->ZERO, c->nvec);
    if (rc == ST_OK) rc = dalloc(c, &c->Wv, c->nvec);
    if (rc == ST_OK) rc = dalloc(c, &c->V, c->nvec * (mr + 1));
    if (rc == ST_OK) rc = dalloc(c, &c->Zb, c->nvec * mr);
    // the held solutions of the warm starts (and the FP32 differences of the
    // minimal-residual start) with the context's other vectors, not during a run
    if (p->warm_order > 0)
        for (double **pp : {&c->XH[0], &c->XH[1], &c->LH[0], &c->LH[1]})
            if (rc == ST_OK) rc = dalloc(c, pp, c->nvec);
    if (p->warm_order == 3 && p->precision == ST_FP32 && march_ok(c->lv[0]))
        for (float **pp : {&c->DX[0], &c->DX[1], &c->DX[2], &c->DL[0], &c->DL[1], &c->DL[2]}) {
            double *t = nullptr;
            if (rc == ST_OK) rc = dalloc(c, &t, (c->nvec + 1) / 2);
            *pp = reinterpret_cast<float *>(t);
        }
    c->npart = 8192 * 33;
    if (rc == ST_OK) rc = dalloc(c, &c->part, c->npart);
    if (rc == ST_OK) rc = dalloc(c, &c->red, 64);
    if (rc == ST_OK) rc = dalloc(c, &c->h_dev, 64);
    if (rc == ST_OK) rc = dalloc(c, &c->sig2_dev, 64);
    for (double **pp : {&c->rho, &c->rhof, &c->dj, &c->dc, &c->rn, &c->fW, &c->dv, &c->tmpc})
        if (rc == ST_OK) rc = dalloc(c, pp, c->ncell);
    if (rc == ST_OK) rc = dalloc(c, &c->abi_a, c->ndof);
    if (rc == ST_OK) rc = dalloc(c, &c->abi_b, c->ndof);
    if (rc == ST_OK) rc = dalloc(c, &c->abi_u8, c->ndof);
    // coarsest compact DOF list (all valid DOFs of the padded coarsest level)
    if (rc == ST_OK) {
        const Lvl &gc = c->lv[c->L - 1];
        std::vector<long> dofs;
        for (int f = 0; f < gc.nfield; f++)
            for (int i = 0; i < gc.Px; i++)
                for (int j = 0; j < gc.Py; j++)
                    for (int k = 0; k < gc.Pz; k++) {
                        int A = gc.nx + (f == 1), Bv = gc.ny + (f == 2), Cv = gc.nz + (f == 3);
                        if (i < A && j < Bv && k < Cv) dofs.push_back((long)f * gc.S + ((long)i * gc.Py + j) * gc.Pz + k);
                    }
        c->nd = (int)dofs.size();
        rc = dalloc(c, &c->d_dofs, c->nd);
        if (rc == ST_OK) rc = dalloc(c, &c->d_minv, (long)c->nd * c->nd);
        if (rc == ST_OK) rc = dalloc(c, &c->d_perm, c->nd);
        if (rc == ST_OK) rc = dalloc(c, &c->d_info, 4);
        if (rc == ST_OK && hipMemcpy(c->d_dofs, dofs.data(), sizeof(long) * c->nd, hipMemcpyHostToDevice) != hipSuccess)
            rc = ST_EHIP;
    }
    if (rc == ST_OK) {
        launch_filter_w(c->lv[0], p->filter_radius > 1.0 ? p->filter_radius : 1.0, c->fW, c->s);
        if (hipStreamSynchronize(c->s) != hipSuccess) rc = ST_EHIP;
    }
    if (rc != ST_OK) {
        std::fprintf(stderr, "st_create: %s\n", c->err.c_str());
        st_destroy(c);
        return rc;
    }
    *out = c;
    return ST_OK;
}

void st_destroy(st_ctx *c)
{
    if (!c) return;
    if (c->s) hipStreamSynchronize(c->s);
    if (c->net) net_destroy(c->net);
    if (c->mma) mma_destroy(c->mma);
    for (void *q : c->allocs) hipFree(q);
    if (c->e0) hipEventDestroy(c->e0);
    if (c->e1) hipEventDestroy(c->e1);
    if (c->e2) hipEventDestroy(c->e2);
    if (c->s) hipStreamDestroy(c->s);
    delete c;
}

const char *st_last_error(const st_ctx *c) { return c ? c->err.c_str() : st::noctx_error(); }

int st_get_params(const st_ctx *c, st_params *p)
{
    if (!c || !p) return ST_EINVAL;
    *p = c->prm;
    return ST_OK;
}
long st_num_dofs(const st_ctx *c) { return c ? c->ndof : -1; }
long st_num_cells(const st_ctx *c) { return c ? c->ncell : -1; }

static void design_to_alpha(st_ctx *c, const double *rho_dev);

int st_set_geometry(st_ctx *c, const uint8_t *fixed, int closed)
{
    if (!c || !fixed || closed < -1 || closed > 1) return ST_EINVAL;
    CK(hipMemcpyAsync(c->abi_u8, fixed, c->ndof, hipMemcpyHostToDevice, c->s));
    launch_flags(c->lv[0], c->abi_u8, c->fl[0], c->s);
    CK(hipStreamSynchronize(c->s));
    // the pressure gauge of the coarsest matrix belongs to the geometry, as in
    // FluidSteadyState3D::init_boundary(..., bool _closed) (ST/FluidSteadyState3D.cpp:92-114);
    // the coarsest inverse is rebuilt on the next solve (dirty)
    if (closed >= 0) c->prm.closed = closed;
    c->geom = true;
    c->dirty = true;
    c->bn2_ok = false;
    c->nxh = c->nlh = 0;
    // a penalty from a design follows the new free faces
    if (c->rho_used) {
        design_to_alpha(c, c->rho_used);
        CK(hipStreamSynchronize(c->s));
    }
    return ST_OK;
}

int st_set_penalty(st_ctx *c, const double *pen)
{
    if (!c || !pen) return ST_EINVAL;
    if (!c->geom) return fail(c, ST_ESTATE, "st_set_geometry must precede st_set_penalty");
    CKR(check_finite(c, pen, c->ndof, "penalty"));
    for (long q = 0; q < c->ncell; q++)
        if (pen[q] != 0.0) return fail(c, ST_EINVAL, "nonzero cell penalty is not supported");
    CKR(stage_in(c, pen, c->ndof));
    launch_alpha_from_pen<double>(c->lv[0], c->abi_a, c->fl[0], c->h64.al[0], c->s);
    CK(hipStreamSynchronize(c->s));
    c->rho_used = nullptr;
    c->alpha = true;
    c->dirty = true;
    return ST_OK;
}

static AlphaParams alpha_params(const st_ctx *c)
{
    AlphaParams a;
    a.amax = c->prm.alpha_max;
    a.amin = c->prm.alpha_min;
    a.q = c->prm.q;
    a.power = c->prm.power;
    a.scale = c->prm.alpha_scale;
    return a;
}

static void design_to_alpha(st_ctx *c, const double *rho_dev)
{
    c->rho_used = rho_dev;
    launch_design_alpha(c->lv[0], rho_dev, alpha_params(c), c->fl[0], c->h64.al[0], c->s);
    c->alpha = true;
    c->dirty = true;
}

int st_set_design(st_ctx *c, const double *rho)
{
    if (!c || !rho) return ST_EINVAL;
    if (!c->geom) return fail(c, ST_ESTATE, "st_set_geometry must precede st_set_design");
    CKR(check_finite(c, rho, c->ncell, "design"));
    CK(hipMemcpyAsync(c->rho, rho, sizeof(double) * c->ncell, hipMemcpyHostToDevice, c->s));
    design_to_alpha(c, c->rho);
    CK(hipStreamSynchronize(c->s));
    return ST_OK;
}

int st_set_design_device(st_ctx *c, const double *d_rho)
{
    if (!c || !d_rho) return ST_EINVAL;
    if (!c->geom) return fail(c, ST_ESTATE, "st_set_geometry must precede st_set_design_device");
    CK(hipMemcpyAsync(c->rho, d_rho, sizeof(double) * c->ncell, hipMemcpyDeviceToDevice, c->s));
    design_to_alpha(c, c->rho);
    CK(hipStreamSynchronize(c->s));
    return ST_OK;
}

int st_profile(st_ctx *c, int on)
{
    if (!c) return ST_EINVAL;
    CK(hipStreamSynchronize(c->s));
    prof_collect(c);
    if (on) {
        for (int k = 0; k < ST_PROF_NKIND; k++) { c->prof_ms[k] = 0.0; c->prof_n[k] = 0; }
    }
    c->prof = on != 0;
    return ST_OK;
}

int st_profile_read(st_ctx *c, int kind, double *total_ms, long *launches, double *bytes)
{
    if (!c || kind < 0 || kind >= ST_PROF_NKIND) return ST_EINVAL;
    if (total_ms) *total_ms = c->prof_ms[kind];
    if (launches) *launches = c->prof_n[kind];
    if (bytes) {
        // algorithmic HBM bytes of one finest-level launch over the valid DOFs, with the
        // masks counted as bit fields (n_f/8 bytes):
        //   operator apply      (s_x + s_y)(n_c + n_f) + s_a n_f + n_f/8  (x, alpha read, y write)
        //   residual            s(3 n_c + 4 n_f) + n_f/8   (x, b, alpha read, r write)
        //   smoother colour     half of a two-colour sweep s(3 n_c + 4 n_f) + n_f/8
        //                       (x, b, alpha read once, x written once per sweep)
        //   zero-start pass     s(2 n_c + 3 n_f) + n_f/8  (b, alpha read, x written; x = 0
        //                       is not read)
        //   one-pass sweep      s(3 n_c + 4 n_f) + n_f/8  (both colours)
        // s: storage of the level-0 multigrid vectors; the Krylov operator of the FP32
        // path is FP32, that of the BF16 path FP64 on FP32 vectors
        const double nc = (double)c->ncell, nf = (double)(c->ndof - c->ncell);
        const int pr = c->prm.precision;
        const double sv = (pr == ST_FP32) ? 4.0 : (pr == ST_BF16 ? 2.0 : 8.0);
        if (kind == ST_PROF_SMOOTH) *bytes = 0.5 * (sv * (3.0 * nc + 4.0 * nf) + nf / 8.0);
        else if (kind == ST_PROF_SWEEP) *bytes = sv * (3.0 * nc + 4.0 * nf) + nf / 8.0;
        else if (kind == ST_PROF_SMOOTH0) *bytes = sv * (2.0 * nc + 3.0 * nf) + nf / 8.0;
        else if (kind == ST_PROF_APPLY) {
            const double sx = (pr == ST_FP64) ? 8.0 : 4.0, sa = (pr == ST_FP32) ? 4.0 : 8.0, sy = sa;
            *bytes = (sx + sy) * (nc + nf) + sa * nf + nf / 8.0;
        } else *bytes = sv * (3.0 * nc + 4.0 * nf) + nf / 8.0;
    }
    return ST_OK;
}

int st_get_design(st_ctx *c, double *rho)
{
    if (!c || !rho) return ST_EINVAL;
    CK(hipMemcpyAsync(rho, c->rho, sizeof(double) * c->ncell, hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    return ST_OK;
}

int st_set_rhs(st_ctx *c, const double *body, const double *dirichlet)
{
    if (!c) return ST_EINVAL;
    if (!c->geom) return fail(c, ST_ESTATE, "st_set_geometry must precede st_set_rhs");
    const Lvl &g = c->lv[0];
    if (dirichlet) {
        CKR(check_finite(c, dirichlet, c->ndof, "dirichlet"));
        CKR(stage_in(c, dirichlet, c->ndof));
        launch_pack<double>(g, c->abi_a, c->XD, 1, c->fl[0], 1.0, c->s);
    } else {
        launch_fill<double>(c->nvec, c->XD, 0.0, c->s);
    }
    if (body) {
        CKR(check_finite(c, body, c->ndof, "body"));
        CKR(stage_in(c, body, c->ndof));
        launch_pack<double>(g, c->abi_a, c->BODY, 0, c->fl[0], 1.0, c->s);
    } else {
        launch_fill<double>(c->nvec, c->BODY, 0.0, c->s);
    }
    launch_lift(g, c->fl[0], c->XD, c->BODY, c->B, c->s);
    CK(hipStreamSynchronize(c->s));
    c->bn2_ok = false;
    c->rhs = true;
    c->nxh = c->nlh = 0;
    return ST_OK;
}

static int solve_core_(st_ctx *c, const double *b, double *x, const double *xd, int init, st_stats *st);

// A failed solve (other than the iteration cap) leaves x undefined: the sequence of
// held solutions it belonged to starts again
static int solve_core(st_ctx *c, const double *b, double *x, const double *xd, int init, st_stats *st)
{
    const int rc = solve_core_(c, b, x, xd, init, st);
    if (rc != ST_OK && rc != ST_ENOCONV) {
        if (x == c->X) c->nxh = 0;
        if (x == c->LAM) c->nlh = 0;
    }
    return rc;
}

static int solve_core_(st_ctx *c, const double *b, double *x, const double *xd, int init, st_stats *st)
{
    if (!c->geom || !c->alpha) return fail(c, ST_ESTATE, "geometry and penalty/design must be set before solving");
    st_stats local;
    std::memset(&local, 0, sizeof(local));
    CK(hipEventRecord(c->e0, c->s));
    if (c->dirty) CKR(setup_hierarchy(c));
    local.setup_ms = c->last_setup_ms;
    const bool use_r0 = c->r0_ready && init == ST_INIT_KEEP;
    c->r0_ready = false;
    if (init == ST_INIT_ZERO) launch_axpby<double, double>(c->nvec, 1.0, xd, 0.0, x, c->s);
    int rc = (c->prm.precision == ST_FP64)   ? fgmres<double, double>(c, b, x, &local, false)
             : (c->prm.precision == ST_FP32) ? fgmres<float, float>(c, b, x, &local, use_r0)
                                              : fgmres<float, double>(c, b, x, &local, false);
    CK(hipEventRecord(c->e1, c->s));
    CK(hipEventSynchronize(c->e1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->e0, c->e1);
    local.ms = ms;
    prof_collect(c);
    if (st) *st = local;
    return rc;
}

// ------------------------------------------------- extrapolated start

static int resid_norm(st_ctx *c, const double *b, const double *x, double *nrm);
static int resid_norm_u0(st_ctx *c, const double *b, const double *x, double *nrm);

// Minimal-residual start (warm_order 3, FP32 path on a tiled finest level): with x1 the
// held solution and D_i the differences of the last solutions of the sequence (up to
// 3, newest first, D_1 = x1 - x_2 formed here), the start is
//   x0 = x1 + sum d_i D_i,  d = argmin || r1 - sum d_i A D_i ||,  r1 = b - A x1,
// taken when its residual is below ||r1||.  The span holds x1 itself and both
// polynomial extrapolations (2 x1 - x2 = x1 + D_1, 3 x1 - 3 x2 + x3 = x1 + 2 D_1 - D_2),
// so the start is never worse than any of them.  r1 is the FP64 residual written in FP32
// into u_0; A D_i (the FP32 operator of the solve) go into the next three basis slots
// (free until the solve); d from the normal equations of their Gram matrix (one pass);
// the update writes x0 and r0 = r1 - sum d_i A D_i into u_0, which the solve takes as its
// first residual.  The corrections are small next to x1, so the FP32 differences and
// products change x0 by ~1e-7 of the correction only.
static int warm_mr(st_ctx *c, const double *b, double **xp, double **h, int &nh, float **dh, float *ms)
{
    const long nv = c->nvec;
    const Lvl &g = c->lv[0];
    if (!h[0]) {
        CKR(dalloc(c, &h[0], nv));
        CKR(dalloc(c, &h[1], nv));
    }
    if (!dh[0])
        for (int i = 0; i < 3; i++) {
            double *t = nullptr;
            CKR(dalloc(c, &t, (nv + 1) / 2));
            dh[i] = reinterpret_cast<float *>(t);
        }
    hipEvent_t a, e;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&e));
    CK(hipEventRecord(a, c->s));
    // the products below use the FP32 operator of the new design: set up the hierarchy
    // the solve then uses (its time counts in the start's)
    if (c->dirty) {
        const int rc = setup_hierarchy(c);
        if (rc != ST_OK) {
            hipEventDestroy(a);
            hipEventDestroy(e);
            nh = 0;
            return rc;
        }
    }
    double *x1 = *xp, *f = h[1];
    if (nh >= 1) {  // newest difference x1 - x_2 into the oldest slot, then rotate
        float *t = dh[2];
        dh[2] = dh[1];
        dh[1] = dh[0];
        dh[0] = t;
        launch_diff_f32(nv, x1, h[0], dh[0], c->s);
    }
    const int nd = std::min(nh, 3);
    float *u0 = reinterpret_cast<float *>(c->V);
    double r1 = 0.0;
    CKR(resid_norm_u0(c, b, x1, &r1));
    double r0 = r1;
    bool take = false, updated = false;  // updated: u_0 holds r0, no longer r1
    if (nd >= 1 && std::isfinite(r1) && r1 > 0.0) {
        float *y = u0 + nv;  // basis slots 1..nd
        for (int i = 0; i < nd; i++) march_apply<float>(g, c->fl[0], c->h32.al[0], dh[i], y + i * nv, c->s);
        const int nq = nd * (nd + 1) / 2 + nd;
        const int nbk = (int)mr_blocks(nv);
        launch_mr_gram(nd, nv, y, nv, u0, c->part, c->s);
        launch_sum(nbk, nq, c->part, c->red, c->s);
        double q[9];
        CK(hipMemcpyAsync(q, c->red, sizeof(double) * nq, hipMemcpyDeviceToHost, c->s));
        CK(hipStreamSynchronize(c->s));
        // G d = g (G symmetric positive semi-definite) by elimination in order; a pivot
        // below 1e-12 of the largest diagonal marks a (near-)dependent direction, which
        // gets a zero coefficient
        double G[3][3] = {}, rhs[3] = {};
        int k = 0;
        for (int i = 0; i < nd; i++)
            for (int j = i; j < nd; j++) G[i][j] = G[j][i] = q[k++];
        for (int i = 0; i < nd; i++) rhs[i] = q[k++];
        double scale = 0.0;
        for (int i = 0; i < nd; i++) scale = std::max(scale, G[i][i]);
        bool ok = scale > 0.0 && std::isfinite(scale), drop[3] = {false, false, false};
        for (int p = 0; ok && p < nd; p++) {
            if (!(G[p][p] > 1e-12 * scale)) {
                drop[p] = true;
                continue;
            }
            for (int i = p + 1; i < nd; i++) {
                const double m = G[i][p] / G[p][p];
                for (int j = p; j < nd; j++) G[i][j] -= m * G[p][j];
                rhs[i] -= m * rhs[p];
            }
        }
        MrCoef cf{{0.0, 0.0, 0.0}};
        for (int p = nd - 1; ok && p >= 0; p--) {
            if (drop[p]) continue;
            double t = rhs[p];
            for (int j = p + 1; j < nd; j++) t -= G[p][j] * cf.d[j];
            cf.d[p] = t / G[p][p];
            if (!std::isfinite(cf.d[p])) ok = false;
        }
        double gd = 0.0;
        for (int i = 0; ok && i < nd; i++) gd += cf.d[i] * q[nd * (nd + 1) / 2 + i];
        // predicted ||r0||^2 = ||r1||^2 - g.d (d solves the normal equations)
        if (ok && gd > 1e-6 * r1 * r1) {
            launch_mr_update(nd, nv, x1, dh, y, nv, cf, f, u0, c->part, c->s);
            updated = true;
            launch_sum(nbk, 1, c->part, c->red, c->s);
            double rr = 0.0;
            CK(hipMemcpyAsync(&rr, c->red, sizeof(double), hipMemcpyDeviceToHost, c->s));
            CK(hipStreamSynchronize(c->s));
            r0 = std::sqrt(rr);
            take = std::isfinite(r0) && r0 < r1;
        }
    }
    if (take) {
        *xp = f;
        h[1] = h[0];
        h[0] = x1;
        c->r0_ready = true;
        c->r0_beta = r0;
    } else {
        // x1 stays the start; u_0 holds r1 unless the update overwrote it
        launch_axpby<double, double>(nv, 1.0, x1, 0.0, f, c->s);
        h[1] = h[0];
        h[0] = f;
        c->r0_ready = !updated;
        c->r0_beta = r1;
    }
    nh = std::min(nh + 1, 3);
    CK(hipEventRecord(e, c->s));
    CK(hipEventSynchronize(e));
    hipEventElapsedTime(ms, a, e);
    hipEventDestroy(a);
    hipEventDestroy(e);
    return ST_OK;
}

// Start of a solve that continues a sequence (ST_INIT_KEEP, ST_INIT_LEARNED; order 3 on
// the FP32 path is warm_mr above): *xp holds the previous solution x1, h[0] = x2 and
// h[1] = x3 the ones before it (nh valid).  The candidate is the polynomial
// extrapolation of order min(nh, warm_order, 2) through them,
// 2 x1 - x2 or 3 x1 - 3 x2 + x3 (the designs of an optimisation change smoothly, and so
// do their solutions), kept when its true FP64 residual is below that of x1.  After the
// call *xp is the start and h[0] = x1, h[1] = x2: the three buffers are rotated; x1 is
// copied only when the candidate loses (or when there is nothing to extrapolate yet).
static int warm_extrapolate(st_ctx *c, const double *b, double **xp, double **h, int &nh, float **dh, float *ms)
{
    *ms = 0.f;
    c->r0_ready = false;
    if (c->prm.warm_order == 0) return ST_OK;
    if (c->prm.warm_order == 3 && c->prm.precision == ST_FP32 && march_ok(c->lv[0]))
        return warm_mr(c, b, xp, h, nh, dh, ms);
    if (!h[0]) {
        CKR(dalloc(c, &h[0], c->nvec));
        CKR(dalloc(c, &h[1], c->nvec));
    }
    hipEvent_t a, e;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&e));
    CK(hipEventRecord(a, c->s));
    double *x1 = *xp, *x2 = h[0], *f = h[1];  // f: x3, or free
    nh = std::min(nh, 2);
    const int ord = std::min(nh, std::min(c->prm.warm_order, 2));
    bool take = false;
    if (ord >= 1) {
        if (ord == 2) launch_lincomb3(c->nvec, 3.0, x1, -3.0, x2, 1.0, f, c->s);  // 3 x1 - 3 x2 + x3
        else launch_lincomb3(c->nvec, 2.0, x1, -1.0, x2, 0.0, f, c->s);            // 2 x1 - x2
        double r_prev = 0.0, r_cand = 0.0;
        CKR(resid_norm(c, b, x1, &r_prev));
        // on the FP32 path the candidate's residual goes where the solve's first
        // residual goes (u_0, in FP32), so the solve need not compute it again
        const bool f32 = c->prm.precision == ST_FP32 && march_ok(c->lv[0]);
        if (f32) CKR(resid_norm_u0(c, b, f, &r_cand));
        else CKR(resid_norm(c, b, f, &r_cand));
        take = std::isfinite(r_cand) && r_cand < r_prev;
        c->r0_ready = take && f32;
        c->r0_beta = r_cand;
    }
    if (take) {
        *xp = f;
        h[0] = x1;
        h[1] = x2;
    } else {
        launch_axpby<double, double>(c->nvec, 1.0, x1, 0.0, f, c->s);
        h[0] = f;
        h[1] = x2;
    }
    nh = std::min(nh + 1, 2);
    CK(hipEventRecord(e, c->s));
    CK(hipEventSynchronize(e));
    hipEventElapsedTime(ms, a, e);
    hipEventDestroy(a);
    hipEventDestroy(e);
    return ST_OK;
}

// ------------------------------------------------------ learned predictor

static int net_ready(st_ctx *c)
{
    if (c->dim != 3) return fail(c, ST_EINVAL, "the learned predictor is 3D only");
    if (c->net) return ST_OK;
    if (c->L < 3 || c->lv[2].nx < 4 || c->lv[2].nx % 4) return fail(c, ST_EINVAL, "grid too small for the predictor");
    c->lc = 2;
    const Lvl &gc = c->lv[c->lc];
    const long nv = (long)gc.nfield * gc.S;
    CKR(dalloc(c, &c->xdc, nv));
    CKR(dalloc(c, &c->xcp, nv));
    CKR(dalloc(c, &c->xcn, nv));
    CKR(dalloc(c, &c->rhop, c->ncell));
    c->net = net_create(gc.nx, 32, 12345u, c->s);
    if (!c->net) return fail(c, ST_EHIP, "cannot allocate the predictor");
    return ST_OK;
}

// fine FP64 padded vector -> level-lc state (two state restrictions)
static void restrict_state(st_ctx *c, const double *xf, double *xc)
{
    launch_state_restrict(c->lv[0], c->lv[1], c->fl[1], xf, c->h64.t[1], c->s);
    launch_state_restrict(c->lv[1], c->lv[2], c->fl[2], c->h64.t[1], xc, c->s);
}

// x += P P xc  (prolongation to the finest level, free DOFs only)
static void prolong_state_add(st_ctx *c, const double *xc, double *x)
{
    const Lvl &g1 = c->lv[1];
    launch_fill<double>((long)g1.nfield * g1.S, c->h64.t[1], 0.0, c->s);
    launch_prolong<double>(c->lv[1], c->lv[2], c->fl[1], xc, c->h64.t[1], c->s);
    launch_prolong<double>(c->lv[0], c->lv[1], c->fl[0], c->h64.t[1], x, c->s);
}

static int predict(st_ctx *c, const double *rho_dev, double *xc)
{
    if (!rho_dev) return fail(c, ST_ESTATE, "the predictor needs a design (st_set_design / st_to_step)");
    restrict_state(c, c->XD, c->xdc);
    if (net_predict(c->net, c->lv[c->lc], c->n, c->n / c->lv[c->lc].nx, rho_dev, c->fl[c->lc], c->xdc, xc) != 0)
        return fail(c, ST_EINVAL, "predictor/level mismatch");
    return ST_OK;
}

// the same with the residual stored in FP32 as the first Krylov vector u_0 (the
// FP32 path on a tiled finest level)
static int resid_norm_u0(st_ctx *c, const double *b, const double *x, double *nrm)
{
    const Lvl &g = c->lv[0];
    march_residual_f32(g, c->fl[0], c->h64.al[0], b, x, reinterpret_cast<float *>(c->V), c->part, c->s);
    launch_sum((int)residual_blocks(g), 1, c->part, c->red, c->s);
    double r2 = 0.0;
    CK(hipMemcpyAsync(&r2, c->red, sizeof(double), hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    *nrm = std::sqrt(r2);
    return ST_OK;
}

// true FP64 residual norm ||b - A x|| (the residual itself is not stored)
static int resid_norm(st_ctx *c, const double *b, const double *x, double *nrm)
{
    const Lvl &g = c->lv[0];
    launch_residual<double>(g, c->fl[0], c->h64.al[0], b, x, (double *)nullptr, c->part, c->s);
    launch_sum((int)residual_blocks(g), 1, c->part, c->red, c->s);
    double r2 = 0.0;
    CK(hipMemcpyAsync(&r2, c->red, sizeof(double), hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    *nrm = std::sqrt(r2);
    return ST_OK;
}

// after a forward solve: remember the design its solution belongs to
// back-off of the learned candidate: after 3 consecutive rejections the network is tried
// every 4th call, the period doubling with every further rejection up to every 64th
static bool learned_backs_off(const st_ctx *c, long call)
{
    const int per = 1 << std::min(std::max(c->learned_rejects - 1, 0), 6);
    return c->learned_rejects >= 3 && (call % per) != 0;
}

static void remember_design(st_ctx *c)
{
    // kept only when the next learned start will try the network (it needs the design
    // the held solution belongs to)
    if (c->net && c->rhop && c->rho_used && !learned_backs_off(c, c->learned_calls + 1)) {
        launch_axpby<double, double>(c->ncell, 1.0, c->rho_used, 0.0, c->rhop, c->s);
        c->have_rhop = true;
    } else {
        c->have_rhop = false;
    }
}

// Learned initial guess in X; returns its device time in ms and whether the
// network's candidate was taken.  Candidates (both network evaluations use the
// current weights):
//   with the solution of a previous design in X:
//     X + P P (net(rho) - net(rho_prev))  against  X (the previous solution);
//   without one:
//     x_D + P P net(rho)                   against  x_D (the zero initial guess).
// The candidate with the smaller true FP64 residual ||b - A x|| is kept, so the
// start without the network wins every tie.  After 3 rejections in a row
// the candidate is built and checked only on every 4th call (the others take the
// start without the network), so a predictor that keeps losing costs little.
static int learned_guess(st_ctx *c, float *ms, int *used)
{
    CKR(net_ready(c));
    if (net_num_samples(c->net) < 1) return fail(c, ST_ESTATE, "the predictor is not trained");
    if (c->dirty) CKR(setup_hierarchy(c));
    const Lvl &gc = c->lv[c->lc];
    const long nvc = (long)gc.nfield * gc.S;
    c->learned_calls++;
    if (c->have_x && (!c->have_rhop || learned_backs_off(c, c->learned_calls))) {
        // no design for the held solution, or backing off: the start without the network
        if (ms) *ms = 0.f;
        if (used) *used = 0;
        return ST_OK;
    }
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    CK(hipEventRecord(a, c->s));
    const bool corr = c->have_x && c->have_rhop;
    if (!c->have_x) launch_axpby<double, double>(c->nvec, 1.0, c->XD, 0.0, c->X, c->s);
    int take = 0;
    {
        CKR(predict(c, c->rho_used, c->xcn));
        if (corr) {
            CKR(predict(c, c->rhop, c->xcp));
            launch_axpby<double, double>(nvc, -1.0, c->xcp, 1.0, c->xcn, c->s);  // xcn = net(rho) - net(rho_prev)
        }
        // the residual of the start without the network: known when the warm start left
        // it in u_0, else computed
        double r_base = 0.0, r_cand = 0.0;
        if (c->r0_ready) r_base = c->r0_beta;
        else CKR(resid_norm(c, c->B, c->X, &r_base));
        launch_axpby<double, double>(c->nvec, 1.0, c->X, 0.0, c->Wv, c->s);  // keep the base
        prolong_state_add(c, c->xcn, c->X);
        CKR(resid_norm(c, c->B, c->X, &r_cand));
        take = (std::isfinite(r_cand) && r_cand < r_base) ? 1 : 0;
        if (take) c->r0_ready = false;  // the start changed
        if (!take) launch_axpby<double, double>(c->nvec, 1.0, c->Wv, 0.0, c->X, c->s);
        c->learned_rejects = take ? 0 : c->learned_rejects + 1;
    }
    CK(hipEventRecord(b, c->s));
    CK(hipEventSynchronize(b));
    float ims = 0.f;
    hipEventElapsedTime(&ims, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (ms) *ms = ims;
    if (used) *used = take;
    return ST_OK;
}

int st_solve(st_ctx *c, int init, const double *x0, double *x_out, st_stats *stats)
{
    if (!c) return ST_EINVAL;
    if (!c->rhs) return fail(c, ST_ESTATE, "st_set_rhs must precede st_solve");
    float wms = 0.f;  // extrapolated start
    if (init == ST_INIT_HOST) {
        if (!x0) return ST_EINVAL;
        CKR(check_finite(c, x0, c->ndof, "x0"));
        CKR(stage_in(c, x0, c->ndof));
        launch_pack<double>(c->lv[0], c->abi_a, c->X, 0, c->fl[0], 1.0, c->s);
        c->nxh = 0;  // a new sequence
    } else if (init == ST_INIT_KEEP) {
        if (!c->have_x) init = ST_INIT_ZERO;
        else CKR(warm_extrapolate(c, c->B, &c->X, c->XH, c->nxh, c->DX, &wms));
    } else if (init == ST_INIT_LEARNED) {
        float ims = 0.f;
        int used = 0;
        if (c->have_x) CKR(warm_extrapolate(c, c->B, &c->X, c->XH, c->nxh, c->DX, &wms));
        else c->nxh = 0;
        CKR(learned_guess(c, &ims, &used));
        int rc = solve_core(c, c->B, c->X, c->XD, ST_INIT_KEEP, stats);
        if (stats) {
            stats->init_ms = wms + ims;
            stats->ms += wms + ims;
            stats->learned_used = used;
        }
        if (rc == ST_OK || rc == ST_ENOCONV) {
            c->have_x = true;
            remember_design(c);
        }
        if (x_out && (rc == ST_OK || rc == ST_ENOCONV)) {
            launch_unpack<double>(c->lv[0], c->X, c->abi_b, c->s);
            int r2 = stage_out(c, x_out, c->ndof);
            if (r2 != ST_OK) return r2;
        }
        return rc;
    } else if (init != ST_INIT_ZERO) {
        return ST_EINVAL;
    }
    if (init == ST_INIT_ZERO) c->nxh = 0;
    int rc = solve_core(c, c->B, c->X, c->XD, init == ST_INIT_HOST ? ST_INIT_KEEP : init, stats);
    if (stats && wms > 0.f) {
        stats->init_ms = wms;
        stats->ms += wms;
    }
    if (rc == ST_OK || rc == ST_ENOCONV) {
        c->have_x = true;
        remember_design(c);
    }
    if (x_out && (rc == ST_OK || rc == ST_ENOCONV)) {
        launch_unpack<double>(c->lv[0], c->X, c->abi_b, c->s);
        int r2 = stage_out(c, x_out, c->ndof);
        if (r2 != ST_OK) return r2;
    }
    return rc;
}

int st_get_gradient(st_ctx *c, double *dc)
{
    if (!c || !dc) return ST_EINVAL;
    if (!c->have_grad) return fail(c, ST_ESTATE, "no design gradient yet (st_to_step)");
    CK(hipMemcpyAsync(dc, c->dc, sizeof(double) * c->ncell, hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    return ST_OK;
}

int st_reset_solution(st_ctx *c)
{
    if (!c) return ST_EINVAL;
    c->have_x = c->have_lam = c->have_rhop = false;
    c->nxh = c->nlh = 0;
    return ST_OK;
}

int st_get_solution(st_ctx *c, double *x)
{
    if (!c || !x) return ST_EINVAL;
    launch_unpack<double>(c->lv[0], c->X, c->abi_b, c->s);
    return stage_out(c, x, c->ndof);
}

int st_apply(st_ctx *c, const double *x, double *y)
{
    if (!c || !x || !y) return ST_EINVAL;
    if (!c->geom || !c->alpha) return fail(c, ST_ESTATE, "geometry and penalty/design must be set");
    CKR(stage_in(c, x, c->ndof));
    launch_pack<double>(c->lv[0], c->abi_a, c->Wv, 0, c->fl[0], 1.0, c->s);
    launch_apply<double>(c->lv[0], c->fl[0], c->h64.al[0], c->Wv, c->BA, c->s);
    launch_unpack<double>(c->lv[0], c->BA, c->abi_b, c->s);
    return stage_out(c, y, c->ndof);
}

int st_apply_device(st_ctx *c, const double *d_x, double *d_y)
{
    if (!c || !d_x || !d_y) return ST_EINVAL;
    if (!c->geom || !c->alpha) return fail(c, ST_ESTATE, "geometry and penalty/design must be set");
    launch_pack<double>(c->lv[0], d_x, c->Wv, 0, c->fl[0], 1.0, c->s);
    launch_apply<double>(c->lv[0], c->fl[0], c->h64.al[0], c->Wv, c->BA, c->s);
    launch_unpack<double>(c->lv[0], c->BA, d_y, c->s);
    CK(hipStreamSynchronize(c->s));
    return ST_OK;
}

static int objective_dev(st_ctx *c, const double *x, double *y, double *J)
{
    const Lvl &g = c->lv[0];
    launch_energy(g, c->fl[0], c->h64.al[0], x, y, c->part, c->s);
    launch_sum((int)energy_blocks(g), 1, c->part, c->red, c->s);
    double s = 0.0;
    CK(hipMemcpyAsync(&s, c->red, sizeof(double), hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    if (J) *J = 0.5 * s;
    return ST_OK;
}

int st_objective(st_ctx *c, const double *x, double *J)
{
    if (!c || !x || !J) return ST_EINVAL;
    if (!c->geom || !c->alpha) return fail(c, ST_ESTATE, "geometry and penalty/design must be set");
    CKR(stage_in(c, x, c->ndof));
    launch_pack<double>(c->lv[0], c->abi_a, c->Wv, 0, c->fl[0], 1.0, c->s);
    return objective_dev(c, c->Wv, nullptr, J);
}

int st_energy_grad(st_ctx *c, const double *x, double *gx)
{
    if (!c || !x || !gx) return ST_EINVAL;
    if (!c->geom || !c->alpha) return fail(c, ST_ESTATE, "geometry and penalty/design must be set");
    CKR(stage_in(c, x, c->ndof));
    launch_pack<double>(c->lv[0], c->abi_a, c->Wv, 0, c->fl[0], 1.0, c->s);
    CKR(objective_dev(c, c->Wv, c->BA, nullptr));
    launch_unpack<double>(c->lv[0], c->BA, c->abi_b, c->s);
    return stage_out(c, gx, c->ndof);
}

int st_sensitivity(st_ctx *c, const double *rho, const double *x, const double *lam, double *dj)
{
    if (!c || !rho || !x || !lam || !dj) return ST_EINVAL;
    if (!c->geom) return fail(c, ST_ESTATE, "geometry must be set");
    CK(hipMemcpyAsync(c->rn, rho, sizeof(double) * c->ncell, hipMemcpyHostToDevice, c->s));
    CKR(stage_in(c, x, c->ndof));
    launch_pack<double>(c->lv[0], c->abi_a, c->Wv, 0, c->fl[0], 1.0, c->s);
    CKR(stage_in(c, lam, c->ndof));
    launch_pack<double>(c->lv[0], c->abi_a, c->BA, 0, c->fl[0], 1.0, c->s);
    launch_sensitivity(c->lv[0], c->rn, alpha_params(c), c->fl[0], c->Wv, c->BA, c->dj, c->s);
    CK(hipMemcpyAsync(dj, c->dj, sizeof(double) * c->ncell, hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    return ST_OK;
}

static void filter_dev(st_ctx *c, const double *in, double *out, int transpose)
{
    if (c->prm.filter_radius > 1.0)
        launch_filter(c->lv[0], c->prm.filter_radius, c->fW, in, out, transpose, c->s);
    else
        launch_axpby<double, double>(c->ncell, 1.0, in, 0.0, out, c->s);
}

int st_filter(st_ctx *c, const double *in, double *out, int transpose)
{
    if (!c || !in || !out) return ST_EINVAL;
    c->have_grad = false;  // c->dc is reused below
    CK(hipMemcpyAsync(c->rn, in, sizeof(double) * c->ncell, hipMemcpyHostToDevice, c->s));
    filter_dev(c, c->rn, c->dc, transpose);
    CK(hipMemcpyAsync(out, c->dc, sizeof(double) * c->ncell, hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    return ST_OK;
}

// OC with bisection on the device: rho (dev), dc (dev) -> rn (dev)
// c->dv = dV/drho = H^T(1/(n W)), formed on first use
static void volume_gradient(st_ctx *c)
{
    if (c->dv_ready) return;
    launch_fill<double>(c->ncell, c->tmpc, 1.0 / (double)c->ncell, c->s);
    filter_dev(c, c->tmpc, c->dv, 1);
    c->dv_ready = true;
}

static int oc_dev(st_ctx *c, const double *rho, const double *dc, double volfrac, double move, double eta,
                  double rho_min, double *rn, double *change, double *vol_out)
{
    const long n = c->ncell;
    volume_gradient(c);
    // search for the multiplier: the volume falls with it; each pass evaluates OC_K
    // multipliers and keeps the sub-bracket where the volume crosses volfrac, down to
    // the bisection's tolerance.  The first pass spreads them geometrically over a
    // factor 2^(+-3.75) around the previous update's multiplier (the multiplier moves
    // little between design steps; if it moved further, the bracket becomes the side
    // it lies on).  Once both ends of a bracket narrower than half its upper end have
    // known volumes, a pass spreads them over a window of 1/32 of the bracket around
    // the linear interpolation of the volume (544-fold narrowing when the crossing lies
    // in the window; after a pass that missed it, the next is even); the others evenly
    // over the bracket (17-fold narrowing per pass).
    double *base = c->tmpc;  // free once dv is formed
    launch_oc_base(n, dc, c->dv, eta, base, c->s);
    const int gb = (int)oc_grid(n);
    double l1 = 0.0, l2 = 1e9, vol = 0.0;
    double v1 = 0.0, v2 = 0.0;  // volumes at l1, l2 when known
    bool v1k = false, v2k = false, missed = false;
    for (int pass = 0; (l2 - l1) > 1e-13 * (l1 + l2) && pass < 60; pass++) {
        OcFactors f;
        double lk[OC_K], vk[OC_K];
        const bool around = pass == 0 && c->oc_lambda > 0.0 && c->oc_lambda < 1e8;
        double w1 = l1, w2 = l2;  // the window the multipliers are spread over
        bool win = false;
        if (!around && !missed && v1k && v2k && v1 > v2 && (l2 - l1) < 0.5 * l2) {
            const double est = l1 + (v1 - volfrac) / (v1 - v2) * (l2 - l1), hw = (l2 - l1) / 64.0;
            w1 = std::max(l1, est - hw);
            w2 = std::min(l2, w1 + 2.0 * hw);
            w1 = std::max(l1, w2 - 2.0 * hw);
            win = true;
        }
        for (int k = 0; k < OC_K; k++) {
            lk[k] = around ? c->oc_lambda * std::exp2(0.5 * ((double)k - 0.5 * (OC_K - 1)))
                           : w1 + (w2 - w1) * (double)(k + 1) / (double)(OC_K + 1);
            f.f[k] = std::pow(lk[k], -eta);
        }
        launch_oc_vols(n, rho, base, c->dv, f, move, rho_min, c->part, c->s);
        launch_sum(gb, OC_K, c->part, c->red, c->s);
        CK(hipMemcpyAsync(vk, c->red, sizeof(double) * OC_K, hipMemcpyDeviceToHost, c->s));
        CK(hipStreamSynchronize(c->s));
        int k = 0;
        while (k < OC_K && vk[k] > volfrac) k++;
        if (k > 0) {
            l1 = lk[k - 1];
            v1 = vk[k - 1];
            v1k = true;
        }
        if (k < OC_K) {
            l2 = lk[k];
            v2 = vk[k];
            v2k = true;
        }
        missed = win && (k == 0 || k == OC_K);
    }
    c->oc_lambda = 0.5 * (l1 + l2);
    launch_oc_apply(n, rho, base, c->dv, std::pow(0.5 * (l1 + l2), -eta), move, rho_min, rn, c->part, c->s);
    launch_sum(gb, 1, c->part, c->red, c->s);
    CK(hipMemcpyAsync(&vol, c->red, sizeof(double), hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    launch_absdiff_max(n, rn, rho, c->part, c->s);
    std::vector<double> pm(absmax_blocks(n));
    CK(hipMemcpyAsync(pm.data(), c->part, sizeof(double) * pm.size(), hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    double ch = 0.0;
    for (double v : pm) ch = v > ch ? v : ch;
    if (change) *change = ch;
    if (vol_out) *vol_out = vol;
    return ST_OK;
}

int st_oc_update(st_ctx *c, const double *rho, const double *dcp, double volfrac, double move, double eta,
                 double rho_min, double *rho_new, double *change)
{
    if (!c || !rho || !dcp || !rho_new) return ST_EINVAL;
    c->have_grad = false;  // c->dc is reused below
    CK(hipMemcpyAsync(c->rn, rho, sizeof(double) * c->ncell, hipMemcpyHostToDevice, c->s));
    CK(hipMemcpyAsync(c->dc, dcp, sizeof(double) * c->ncell, hipMemcpyHostToDevice, c->s));
    CKR(oc_dev(c, c->rn, c->dc, volfrac, move, eta, rho_min, c->dj, change, nullptr));
    CK(hipMemcpyAsync(rho_new, c->dj, sizeof(double) * c->ncell, hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    return ST_OK;
}

static int mma_ready(st_ctx *c)
{
    if (c->mma) return ST_OK;
    CKR(dalloc(c, &c->mlo, c->ncell));
    CKR(dalloc(c, &c->mhi, c->ncell));
    CKR(dalloc(c, &c->mdg, c->ncell));
    c->mma = mma_create(c->ncell, c->s);
    if (!c->mma) return fail(c, ST_EHIP, "cannot allocate the MMA state");
    return ST_OK;
}

int st_mma_update(st_ctx *c, double *x, const double *dfdx, double g, const double *dgdx, const double *xmin,
                  const double *xmax, int *newton_steps)
{
    if (!c || !x || !dfdx || !dgdx || !xmin || !xmax) return ST_EINVAL;
    if (!std::isfinite(g)) return fail(c, ST_ENAN, "non-finite constraint value");
    CKR(check_finite(c, x, c->ncell, "x"));
    CKR(check_finite(c, dfdx, c->ncell, "dfdx"));
    CKR(mma_ready(c));
    c->have_grad = false;  // c->dc is reused below
    const size_t nb = sizeof(double) * c->ncell;
    CK(hipMemcpyAsync(c->rn, x, nb, hipMemcpyHostToDevice, c->s));
    CK(hipMemcpyAsync(c->dc, dfdx, nb, hipMemcpyHostToDevice, c->s));
    CK(hipMemcpyAsync(c->mdg, dgdx, nb, hipMemcpyHostToDevice, c->s));
    CK(hipMemcpyAsync(c->mlo, xmin, nb, hipMemcpyHostToDevice, c->s));
    CK(hipMemcpyAsync(c->mhi, xmax, nb, hipMemcpyHostToDevice, c->s));
    int r = mma_update(c->mma, c->rn, c->dc, g, c->mdg, c->mlo, c->mhi, newton_steps);
    if (r == -2) return fail(c, ST_ENAN, "MMA dual iteration produced a non-finite value");
    if (r != 0) return fail(c, ST_EHIP, "MMA update failed");
    CK(hipMemcpyAsync(x, c->rn, nb, hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    return ST_OK;
}

int st_mma_reset(st_ctx *c)
{
    if (!c) return ST_EINVAL;
    if (c->mma) mma_reset(c->mma);
    return ST_OK;
}

// MMA design update of the device design (one volume constraint on the filtered
// density, scaled like the reference's, FT/FluidTopoOpt3D.cpp:19-39):
// g = 1e4 (mean(H rho) - volfrac), dg = 1e4 H^T(1/n); box max(rho_min, rho - move) ..
// min(1, rho + move) (SX/optimization/OptimizerMma.h:157-164)
static int mma_dev(st_ctx *c, const double *rho, const double *dc, double volfrac, double move, double rho_min,
                   double *rn, double *change, double *vol_out)
{
    const long n = c->ncell;
    CKR(mma_ready(c));
    volume_gradient(c);
    double vol = 0.0;
    launch_dot<double, double>(n, rho, c->dv, c->part, c->s);
    launch_sum((int)blas_blocks(n), 1, c->part, c->red, c->s);
    CK(hipMemcpyAsync(&vol, c->red, sizeof(double), hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    const double coef = 1e4;
    launch_axpby<double, double>(n, coef, c->dv, 0.0, c->mdg, c->s);
    launch_mma_box(n, rho, move, rho_min, 1.0, c->mlo, c->mhi, c->s);
    launch_axpby<double, double>(n, 1.0, rho, 0.0, rn, c->s);
    int r = mma_update(c->mma, rn, dc, coef * (vol - volfrac), c->mdg, c->mlo, c->mhi, nullptr);
    if (r == -2) return fail(c, ST_ENAN, "MMA dual iteration produced a non-finite value");
    if (r != 0) return fail(c, ST_EHIP, "MMA update failed");
    launch_absdiff_max(n, rn, rho, c->part, c->s);
    std::vector<double> pm(absmax_blocks(n));
    CK(hipMemcpyAsync(pm.data(), c->part, sizeof(double) * pm.size(), hipMemcpyDeviceToHost, c->s));
    launch_dot<double, double>(n, rn, c->dv, c->part + pm.size(), c->s);
    launch_sum((int)blas_blocks(n), 1, c->part + pm.size(), c->red, c->s);
    double vn = 0.0;
    CK(hipMemcpyAsync(&vn, c->red, sizeof(double), hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    double ch = 0.0;
    for (double v : pm) ch = v > ch ? v : ch;
    if (change) *change = ch;
    if (vol_out) *vol_out = vn;
    return ST_OK;
}

int st_to_step(st_ctx *c, int init, double volfrac, double move, double eta, double rho_min, double *J,
               double *vol, double *change, st_stats *fwd, st_stats *adj)
{
    if (!c) return ST_EINVAL;
    if (!c->geom || !c->rhs) return fail(c, ST_ESTATE, "geometry and rhs must be set before st_to_step");
    if (init != ST_INIT_ZERO && init != ST_INIT_KEEP && init != ST_INIT_LEARNED) return ST_EINVAL;
    // filtered physical density -> face alpha
    filter_dev(c, c->rho, c->rhof, 0);
    design_to_alpha(c, c->rhof);
    int fmode = (init == ST_INIT_ZERO || !c->have_x) ? ST_INIT_ZERO : ST_INIT_KEEP;
    float ims = 0.f, wms = 0.f;
    int used = 0;
    if (fmode == ST_INIT_KEEP) CKR(warm_extrapolate(c, c->B, &c->X, c->XH, c->nxh, c->DX, &wms));
    else c->nxh = 0;
    if (init == ST_INIT_LEARNED && c->have_x && c->net && net_num_samples(c->net) > 0)
        CKR(learned_guess(c, &ims, &used));
    int rc = solve_core(c, c->B, c->X, c->XD, fmode, fwd);
    if (fwd) {
        fwd->init_ms = wms + ims;
        fwd->ms += wms + ims;
        fwd->learned_used = used;
    }
    if (rc != ST_OK) return rc;
    c->have_x = true;
    remember_design(c);
    CKR(objective_dev(c, c->X, c->BA, J));
    // adjoint: A lam = G x (A symmetric), homogeneous Dirichlet data
    int amode = (init == ST_INIT_ZERO || !c->have_lam) ? ST_INIT_ZERO : ST_INIT_KEEP;
    float lms = 0.f;
    if (amode == ST_INIT_KEEP) CKR(warm_extrapolate(c, c->BA, &c->LAM, c->LH, c->nlh, c->DL, &lms));
    else c->nlh = 0;
    rc = solve_core(c, c->BA, c->LAM, c->ZERO, amode, adj);
    if (adj) {
        adj->init_ms = lms;
        adj->ms += lms;
    }
    if (rc != ST_OK) return rc;
    c->have_lam = true;
    launch_sensitivity(c->lv[0], c->rhof, alpha_params(c), c->fl[0], c->X, c->LAM, c->dj, c->s);
    filter_dev(c, c->dj, c->dc, 1);
    c->have_grad = true;
    if (c->prm.optimizer == ST_OPT_MMA)
        CKR(mma_dev(c, c->rho, c->dc, volfrac, move, rho_min, c->rn, change, vol));
    else
        CKR(oc_dev(c, c->rho, c->dc, volfrac, move, eta, rho_min, c->rn, change, vol));
    launch_axpby<double, double>(c->ncell, 1.0, c->rn, 0.0, c->rho, c->s);
    CK(hipStreamSynchronize(c->s));
    return ST_OK;
}

int st_record_sample(st_ctx *c)
{
    if (!c) return ST_EINVAL;
    CKR(net_ready(c));
    if (!c->have_x || !c->rho_used) return fail(c, ST_ESTATE, "record a sample after a solve of a design");
    restrict_state(c, c->XD, c->xdc);
    restrict_state(c, c->X, c->xcn);
    if (net_add_sample(c->net, c->lv[c->lc], c->n, c->n / c->lv[c->lc].nx, c->rho_used, c->fl[c->lc], c->xdc,
                       c->xcn) != 0)
        return fail(c, ST_EHIP, "recording a sample failed");
    return ST_OK;
}

int st_net_eval(st_ctx *c, double *loss, double *loss_zero)
{
    if (!c || !loss || !loss_zero) return ST_EINVAL;
    CKR(net_ready(c));
    if (!c->have_x || !c->rho_used) return fail(c, ST_ESTATE, "evaluate after a solve of a design");
    restrict_state(c, c->XD, c->xdc);
    restrict_state(c, c->X, c->xcn);
    if (net_eval(c->net, c->lv[c->lc], c->n, c->n / c->lv[c->lc].nx, c->rho_used, c->fl[c->lc], c->xdc, c->xcn, loss,
                 loss_zero) != 0)
        return fail(c, ST_ESTATE, "the predictor is not trained");
    return ST_OK;
}

int st_num_samples(const st_ctx *c) { return (c && c->net) ? net_num_samples(c->net) : 0; }

int st_net_train(st_ctx *c, int epochs, double lr, double *loss)
{
    if (!c || epochs < 0) return ST_EINVAL;
    CKR(net_ready(c));
    if (net_num_samples(c->net) < 1) return fail(c, ST_ESTATE, "no recorded samples");
    if (net_train(c->net, epochs, (float)lr, loss) != 0) return fail(c, ST_EHIP, "training failed");
    return ST_OK;
}

long st_net_num_params(st_ctx *c)
{
    if (!c || net_ready(c) != ST_OK) return -1;
    return net_num_params(c->net);
}
int st_net_get_params(st_ctx *c, float *p)
{
    if (!c || !p) return ST_EINVAL;
    CKR(net_ready(c));
    return net_get_params(c->net, p) == 0 ? ST_OK : ST_EHIP;
}
int st_net_set_params(st_ctx *c, const float *p)
{
    if (!c || !p) return ST_EINVAL;
    CKR(net_ready(c));
    return net_set_params(c->net, p) == 0 ? ST_OK : ST_EHIP;
}
int st_net_grad(st_ctx *c, const float *feat, const float *tgt, double *loss, float *grad, float *out)
{
    if (!c || !feat || !tgt || !grad) return ST_EINVAL;
    CKR(net_ready(c));
    return net_grad_host(c->net, feat, tgt, loss, grad, out) == 0 ? ST_OK : ST_EHIP;
}
int st_net_conv_test(int R, int ntap, int cin, int cout, int relu, int mode, const float *in, const float *w,
                     const float *bias, const float *gout, float *out, float *gw, float *gb)
{
    int r = net_conv_test(R, ntap, cin, cout, relu, mode, in, w, bias, gout, out, gw, gb);
    return r == 0 ? ST_OK : (r == -1 ? ST_EINVAL : ST_EHIP);
}

// ------------------------------------------------------ MG test surface

int st_mg_levels(const st_ctx *c) { return c ? c->L : -1; }
int st_mg_level_n(const st_ctx *c, int l) { return (c && l >= 0 && l < c->L) ? c->lv[l].nx : -1; }

int st_mg_vcycle(st_ctx *c, const double *b, double *z)
{
    if (!c || !b || !z) return ST_EINVAL;
    if (c->dirty) CKR(setup_hierarchy(c));
    CKR(stage_in(c, b, c->ndof));
    launch_pack<double>(c->lv[0], c->abi_a, c->BA, 0, c->fl[0], 1.0, c->s);
    if (c->prm.precision == ST_FP64) {
        precondition(c, c->BA, c->Wv);
    } else {  // through the FP32 Krylov storage, as inside FGMRES
        float *v = reinterpret_cast<float *>(c->V), *z = reinterpret_cast<float *>(c->Zb);
        launch_axpby<double, float>(c->nvec, 1.0, c->BA, 0.0, v, c->s);
        precondition(c, v, z);
        launch_axpby<float, double>(c->nvec, 1.0, z, 0.0, c->Wv, c->s);
    }
    launch_unpack<double>(c->lv[0], c->Wv, c->abi_b, c->s);
    return stage_out(c, z, c->ndof);
}

static long level_dofs(const st_ctx *c, int l) { return abi_dofs(c->dim, c->lv[l].nx); }

int st_mg_level_desc(st_ctx *c, int l, uint8_t *fixed, double *alpha_faces)
{
    if (!c || l < 0 || l >= c->L) return ST_EINVAL;
    if (c->dirty) CKR(setup_hierarchy(c));
    const Lvl &g = c->lv[l];
    // flags -> ABI fixed flags via a float round trip of a mask vector
    std::vector<uint8_t> hfl(g.S);
    CK(hipMemcpyAsync(hfl.data(), c->fl[l], g.S, hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    long nd = level_dofs(c, l);
    long o = 0;
    for (int f = 0; f < g.nfield; f++) {
        int A = g.nx + (f == 1), Bv = g.ny + (f == 2), Cv = g.nz + (f == 3);
        for (int i = 0; i < A; i++)
            for (int j = 0; j < Bv; j++)
                for (int k = 0; k < Cv; k++) {
                    long nn = ((long)i * g.Py + j) * g.Pz + k;
                    if (fixed) fixed[o] = ((hfl[nn] >> f) & 1) ? 0 : 1;
                    o++;
                }
    }
    if (o != nd) return fail(c, ST_EINVAL, "internal: level dof count mismatch");
    if (alpha_faces) {
        // pack alpha faces as the face part of an ABI vector (cells zero)
        launch_fill<double>((long)g.nfield * g.S, c->Wv, 0.0, c->s);
        CK(hipMemcpyAsync(c->Wv + g.S, c->h64.al[l], sizeof(double) * g.dim * g.S, hipMemcpyDeviceToDevice, c->s));
        launch_unpack<double>(g, c->Wv, c->abi_b, c->s);
        std::vector<double> tmp(nd);
        CKR(stage_out(c, tmp.data(), nd));
        long nc = (long)g.nx * g.ny * g.nz;
        std::memcpy(alpha_faces, tmp.data() + nc, sizeof(double) * (nd - nc));
    }
    return ST_OK;
}

int st_mg_smooth(st_ctx *c, int l, int color, const double *b, const double *xin, double *xout)
{
    if (!c || l < 0 || l >= c->L || !b || !xin || !xout) return ST_EINVAL;
    if (c->dirty) CKR(setup_hierarchy(c));
    const Lvl &g = c->lv[l];
    long nd = level_dofs(c, l);
    CKR(stage_in(c, b, nd));
    launch_pack<double>(g, c->abi_a, c->BA, 0, c->fl[l], 1.0, c->s);
    CKR(stage_in(c, xin, nd));
    launch_pack<double>(g, c->abi_a, c->Wv, 0, c->fl[l], 1.0, c->s);
    launch_fill<double>((long)g.nfield * g.S, c->V, 0.0, c->s);
    launch_smooth<double>(g, c->fl[l], c->h64.al[l], c->BA, c->Wv, c->V, color, (float)c->prm.omega, false, c->s);
    launch_unpack<double>(g, c->V, c->abi_b, c->s);
    return stage_out(c, xout, nd);
}

int st_mg_smooth_mp(st_ctx *c, int l, int color, const double *b, const double *xin, double *xout)
{
    if (!c || l < 0 || l >= c->L || !b || !xin || !xout) return ST_EINVAL;
    if (c->dirty) CKR(setup_hierarchy(c));
    if (c->prm.precision == ST_FP32) return mg_smooth_lowp<float>(c, l, color, b, xin, xout);
    if (c->prm.precision == ST_BF16) return mg_smooth_lowp<bf16>(c, l, color, b, xin, xout);
    return st_mg_smooth(c, l, color, b, xin, xout);
}

int st_mg_restrict(st_ctx *c, int l, const double *rf, double *bc)
{
    if (!c || l < 0 || l + 1 >= c->L || !rf || !bc) return ST_EINVAL;
    if (c->dirty) CKR(setup_hierarchy(c));
    const Lvl &g = c->lv[l], &gc = c->lv[l + 1];
    CKR(stage_in(c, rf, level_dofs(c, l)));
    launch_pack<double>(g, c->abi_a, c->BA, 0, c->fl[l], 1.0, c->s);
    launch_restrict<double>(g, gc, c->fl[l + 1], c->BA, c->Wv, c->s);
    launch_unpack<double>(gc, c->Wv, c->abi_b, c->s);
    return stage_out(c, bc, level_dofs(c, l + 1));
}

int st_mg_prolong(st_ctx *c, int l, const double *ec, const double *xf, double *xf_out)
{
    if (!c || l < 0 || l + 1 >= c->L || !ec || !xf || !xf_out) return ST_EINVAL;
    if (c->dirty) CKR(setup_hierarchy(c));
    const Lvl &g = c->lv[l], &gc = c->lv[l + 1];
    CKR(stage_in(c, ec, level_dofs(c, l + 1)));
    launch_pack<double>(gc, c->abi_a, c->BA, 0, c->fl[l + 1], 1.0, c->s);
    CKR(stage_in(c, xf, level_dofs(c, l)));
    launch_pack<double>(g, c->abi_a, c->Wv, 0, c->fl[l], 1.0, c->s);
    launch_prolong<double>(g, gc, c->fl[l], c->BA, c->Wv, c->s);
    launch_unpack<double>(g, c->Wv, c->abi_b, c->s);
    return stage_out(c, xf_out, level_dofs(c, l));
}

}  // extern "C"
