// fluid_topo: command-line topology-optimization driver on the C ABI of
// libstokes_hip.so, with the reference's interface (FT/main.cpp:13-55):
//
//   fluid_topo -driver 0|1 -test 0 -s N -o DIR -frac F -lf LAST
//
// -driver 0 is the 2D driver (FT/FluidTopoOptDriver.cpp), 1 the 3D one
// (FT/FluidTopoOptDriver3D.cpp); -test 0 is the reference's test case (closed box
// with source/sink terms, FT/FluidTopoOptDriver3D.cpp:99-167,
// FT/FluidTopoOptDriver.cpp:104-166); -s the cells per axis; -frac the volume
// fraction; -lf the iteration cap.  The loop is OptimizerMma::Optimize_MMA
// (SX/optimization/OptimizerMma.h:138-170): forward solve, adjoint and sensitivity,
// frame output, MMA update with move 0.2 in [0.01, 1], stop when the largest design
// change is below 1e-3 or after LAST iterations.  Every design iteration runs on the
// GPU (st_to_step); the host only writes the frames.
//
// Output (the reference's viewer format, SX/common/Driver.h:57-70 and the
// Write_To_File_3d writers of SX/reservoir): DIR/<frame>/{x, velocity, psi_D,
// psi_N}, DIR/0/grid and DIR/0/last_frame.txt.  Binary, little-endian:
//   grid      int32 counts[3], double dx, double domain_min[3]
//   x         int32 counts[3], double value per cell (i-major, k fastest)
//   velocity  int32 counts[3], double[3] per cell: the mean of the two faces per axis
//   psi_D/N   int32 n, then n positions (double[3]) of the fixed cells / face centres,
//             followed by the velocity, force (double[3] each), mass, colour (double)
//             and index (int32) attributes of each point, all zero
// 2D writes counts (n, n, 1) (grid: (n, n, 0)) and zero third components.
//
// Extra options: -opt mma|oc (default mma, the reference's optimizer), -filter R
// (density filter radius in cells; default 1 = none, as the reference), -mp
// fp64|fp32|bf16 (multigrid precision, default fp64), -init zero|keep (initial
// guesses of the solves, default keep), -rtol T (default 1e-8), -every K (write
// every K-th frame, default 1), -dump FILE (write the problem's fixed flags and
// source terms, uint8[tot] then double[tot], and exit without touching the GPU).
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/stokes_to.h"

namespace {

struct Args {
    int driver = 1, test = 2, s = 64, lf = 200, every = 1;
    double frac = 200.0, filter = 1.0, rtol = 1e-8;
    std::string out = "output", opt = "mma", mp = "fp64", init = "keep", dump;
};

[[noreturn]] void die(const std::string &m)
{
    std::fprintf(stderr, "fluid_topo: %s\n", m.c_str());
    std::exit(1);
}

Args parse(int argc, char **argv)
{
    Args a;
    for (int i = 1; i < argc; i++) {
        std::string k = argv[i];
        if (i + 1 >= argc) die("missing value for " + k);
        std::string v = argv[++i];
        if (k == "-driver") a.driver = std::atoi(v.c_str());
        else if (k == "-test") a.test = std::atoi(v.c_str());
        else if (k == "-s") a.s = std::atoi(v.c_str());
        else if (k == "-o") a.out = v;
        else if (k == "-frac") a.frac = std::atof(v.c_str());
        else if (k == "-lf") a.lf = std::atoi(v.c_str());
        else if (k == "-opt") a.opt = v;
        else if (k == "-filter") a.filter = std::atof(v.c_str());
        else if (k == "-mp") a.mp = v;
        else if (k == "-init") a.init = v;
        else if (k == "-rtol") a.rtol = std::atof(v.c_str());
        else if (k == "-every") a.every = std::max(1, std::atoi(v.c_str()));
        else if (k == "-dump") a.dump = v;
        else die("unknown option " + k);
    }
    return a;
}

// ---- problem: the reference's test 0 in the ABI layout [cells | faces x | y | z]
struct Problem {
    int dim = 3, n = 0;
    long nc = 0, nfx = 0, nfy = 0, nfz = 0, tot = 0;
    std::vector<uint8_t> fixed;
    std::vector<double> body;
    double alpha_os = 0, alpha_us = 1e-3, q = 0.01;
    long cell(int i, int j, int k) const { return ((long)i * n + j) * (dim == 3 ? n : 1) + k; }
    long fxi(int i, int j, int k) const { return nc + ((long)i * n + j) * (dim == 3 ? n : 1) + k; }
    long fyi(int i, int j, int k) const { return nc + nfx + ((long)i * (n + 1) + j) * (dim == 3 ? n : 1) + k; }
    long fzi(int i, int j, int k) const { return nc + nfx + nfy + ((long)i * n + j) * (n + 1) + k; }
};

Problem test0(int dim, int n)
{
    Problem P;
    P.dim = dim;
    P.n = n;
    const int nz = dim == 3 ? n : 1;
    P.nc = (long)n * n * nz;
    P.nfx = (long)(n + 1) * n * nz;
    P.nfy = (long)n * (n + 1) * nz;
    P.nfz = dim == 3 ? (long)n * n * (n + 1) : 0;
    P.tot = P.nc + P.nfx + P.nfy + P.nfz;
    P.fixed.assign(P.tot, 0);
    P.body.assign(P.tot, 0.0);
    for (int a = 0; a < n; a++)
        for (int c = 0; c < nz; c++) {
            P.fixed[P.fxi(0, a, c)] = 1;
            P.fixed[P.fxi(n, a, c)] = 1;
            P.fixed[P.fyi(a, 0, c)] = 1;
            P.fixed[P.fyi(a, n, c)] = 1;
        }
    P.alpha_os = 500.0 / (n / 32.0);
    if (dim == 3) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                P.fixed[P.fzi(i, j, 0)] = 1;
                P.fixed[P.fzi(i, j, n)] = 1;
            }
        // absolute indices of the reference (valid from n = 42); scaled below 64
        auto sc = [&](int v) { return n >= 64 ? v : v * n / 64; };
        for (int j = sc(28); j < sc(36); j++)
            for (int k = sc(28); k < sc(36); k++) {
                P.body[P.cell(0, j, k)] = 600.0;
                P.body[P.fxi(1, j, k)] = 600.0;
            }
        for (int j = sc(10); j < sc(18); j++)
            for (int k = sc(34); k < sc(42); k++) {
                P.body[P.cell(n - 1, j, k)] = -200.0;
                P.body[P.fxi(n - 1, j, k)] = 200.0;
            }
        for (int i = sc(28); i < sc(36); i++)
            for (int j = sc(22); j < sc(30); j++) {
                P.body[P.cell(i, j, n - 1)] = -200.0;
                P.body[P.fzi(i, j, n - 1)] = 200.0;
            }
        for (int i = sc(22); i < sc(30); i++)
            for (int j = sc(28); j < sc(36); j++) {
                P.body[P.cell(i, j, 0)] = -200.0;
                P.body[P.fzi(i, j, 1)] = -200.0;
            }
        P.q = 0.01;
    } else {
        const int O = 16, hin = n / 32, hout = n / (4 * O);
        const double e = 60.0, r = e * (2 * hin + 1) / (2 * hout + 1) / O;
        for (int i = 0; i < n; i++) {
            if (std::abs(i - n / 2) > hin) continue;
            P.body[P.cell(0, i, 0)] = e;
            P.body[P.fxi(1, i, 0)] = e;
        }
        for (int o = 0; o < O; o++)
            for (int i = 0; i < n; i++) {
                if (std::abs(i - n / (2 * O) * (2 * o + 1)) > hout) continue;
                P.body[P.cell(n - 1, i, 0)] = -r;
                P.body[P.fxi(n - 1, i, 0)] = r;
            }
        P.q = 0.014;
    }
    return P;
}

// ---- output in the reference's binary formats
void mkdir_p(const std::string &d)
{
    if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) die("cannot create " + d);
}

struct Out {
    FILE *f;
    explicit Out(const std::string &path) : f(std::fopen(path.c_str(), "wb"))
    {
        if (!f) die("cannot write " + path);
    }
    ~Out() { std::fclose(f); }
    template <typename T> void put(const T &v) { std::fwrite(&v, sizeof(T), 1, f); }
    template <typename T> void put(const T *v, size_t n) { std::fwrite(v, sizeof(T), n, f); }
};

void write_grid(const std::string &path, const Problem &P)
{
    Out o(path);
    int32_t c[3] = {P.n, P.n, P.dim == 3 ? P.n : 0};
    double dx = 1.0, mn[3] = {0.0, 0.0, 0.0};
    o.put(c, 3);
    o.put(dx);
    o.put(mn, 3);
}

void write_field(const std::string &path, const Problem &P, const double *v)
{
    Out o(path);
    int32_t c[3] = {P.n, P.n, P.dim == 3 ? P.n : 1};
    o.put(c, 3);
    o.put(v, (size_t)P.nc);
}

void write_velocity(const std::string &path, const Problem &P, const double *x)
{
    const int n = P.n, nz = P.dim == 3 ? n : 1;
    std::vector<double> v((size_t)P.nc * 3, 0.0);
    const double *ux = x + P.nc, *uy = ux + P.nfx, *uz = uy + P.nfy;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < nz; k++) {
                const long c = P.cell(i, j, k);
                v[3 * c + 0] = 0.5 * (ux[P.fxi(i, j, k) - P.nc] + ux[P.fxi(i + 1, j, k) - P.nc]);
                v[3 * c + 1] = 0.5 * (uy[P.fyi(i, j, k) - P.nc - P.nfx] + uy[P.fyi(i, j + 1, k) - P.nc - P.nfx]);
                if (P.dim == 3)
                    v[3 * c + 2] = 0.5 * (uz[P.fzi(i, j, k) - P.nc - P.nfx - P.nfy] +
                                          uz[P.fzi(i, j, k + 1) - P.nc - P.nfx - P.nfy]);
            }
    Out o(path);
    int32_t c[3] = {n, n, nz};
    o.put(c, 3);
    o.put(v.data(), v.size());
}

void write_points(const std::string &path, const std::vector<double> &X)
{
    Out o(path);
    const int32_t n = (int32_t)(X.size() / 3);
    o.put(n);
    if (n > 0) {
        o.put(X.data(), X.size());
        std::vector<double> z3(X.size(), 0.0), z1((size_t)n, 0.0);
        std::vector<int32_t> zi((size_t)n, 0);
        o.put(z3.data(), z3.size());  // velocity
        o.put(z3.data(), z3.size());  // force
        o.put(z1.data(), z1.size());  // mass
        o.put(z1.data(), z1.size());  // colour
        o.put(zi.data(), zi.size());  // index
    }
}

void write_boundary(const std::string &dir, const Problem &P)
{
    const int n = P.n, nz = P.dim == 3 ? n : 1;
    std::vector<double> D, N;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < nz; k++)
                if (P.fixed[P.cell(i, j, k)]) D.insert(D.end(), {i + 0.5, j + 0.5, P.dim == 3 ? k + 0.5 : 0.0});
    // x faces, then y faces (2D: j outer as the reference), then z faces
    for (int i = 0; i <= n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < nz; k++)
                if (P.fixed[P.fxi(i, j, k)]) N.insert(N.end(), {(double)i, j + 0.5, P.dim == 3 ? k + 0.5 : 0.0});
    if (P.dim == 3) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j <= n; j++)
                for (int k = 0; k < n; k++)
                    if (P.fixed[P.fyi(i, j, k)]) N.insert(N.end(), {i + 0.5, (double)j, k + 0.5});
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
                for (int k = 0; k <= n; k++)
                    if (P.fixed[P.fzi(i, j, k)]) N.insert(N.end(), {i + 0.5, j + 0.5, (double)k});
    } else {
        for (int j = 0; j <= n; j++)
            for (int i = 0; i < n; i++)
                if (P.fixed[P.fyi(i, j, 0)]) N.insert(N.end(), {i + 0.5, (double)j, 0.0});
    }
    write_points(dir + "/psi_D", D);
    write_points(dir + "/psi_N", N);
}

void check(st_ctx *c, int rc, const char *what)
{
    if (rc != ST_OK) die(std::string(what) + " failed (" + std::to_string(rc) + "): " + st_last_error(c));
}

}  // namespace

int main(int argc, char **argv)
{
    Args a = parse(argc, argv);
    if (a.driver != 0 && a.driver != 1) die("-driver must be 0 (2D) or 1 (3D)");
    if (a.test != 0) die("only -test 0 (the reference's test case) is defined");
    const int dim = a.driver == 0 ? 2 : 3;
    if (a.s < 8 || (a.s & (a.s - 1))) die("-s must be a power of two >= 8");
    Problem P = test0(dim, a.s);
    if (!a.dump.empty()) {
        Out o(a.dump);
        o.put(P.fixed.data(), P.fixed.size());
        o.put(P.body.data(), P.body.size());
        return 0;
    }
    if (a.frac <= 0.0 || a.frac >= 1.0) die("-frac must lie in (0, 1)");

    st_params prm;
    st_default_params(&prm, dim, a.s);
    prm.alpha_max = P.alpha_os;
    prm.alpha_min = P.alpha_us;
    prm.q = P.q;
    prm.power = 3.0;
    prm.alpha_scale = 1.0;
    prm.filter_radius = a.filter;
    prm.rtol = a.rtol;
    prm.optimizer = a.opt == "oc" ? ST_OPT_OC : ST_OPT_MMA;
    prm.precision = a.mp == "fp32" ? ST_FP32 : (a.mp == "bf16" ? ST_BF16 : ST_FP64);
    const int init = a.init == "zero" ? ST_INIT_ZERO : ST_INIT_KEEP;

    st_ctx *c = nullptr;
    int rc = st_create(&c, &prm);
    if (rc != ST_OK) die("st_create failed (" + std::to_string(rc) + "); is a GPU visible and -s = 4*2^k?");
    // test 0 is a closed box: the reference passes closed = true to init_boundary
    // (FT/FluidTopoOptDriver3D.cpp:165)
    check(c, st_set_geometry(c, P.fixed.data(), 1), "st_set_geometry");
    check(c, st_set_rhs(c, P.body.data(), nullptr), "st_set_rhs");
    std::vector<double> rho((size_t)P.nc, a.frac), x((size_t)P.tot);
    check(c, st_set_design(c, rho.data()), "st_set_design");

    mkdir_p(a.out);
    double ch = 1.0;
    int it = 0;
    while (ch > 1e-3 && it < a.lf) {
        check(c, st_get_design(c, rho.data()), "st_get_design");
        double J = 0, vol = 0;
        st_stats f, g;
        check(c, st_to_step(c, init, a.frac, 0.2, 0.5, 0.01, &J, &vol, &ch, &f, &g), "st_to_step");
        std::printf("iteration %d: objective %.10g  volume %.6f  change %.3e  forward %d it (%.1f ms)  adjoint %d it"
                    " (%.1f ms)\n", it, J, vol, ch, f.iters, f.ms, g.iters, g.ms);
        std::fflush(stdout);
        if (it % a.every == 0) {
            const std::string fd = a.out + "/" + std::to_string(it);
            mkdir_p(fd);
            mkdir_p(a.out + "/0");
            if (it == 0) write_grid(fd + "/grid", P);
            check(c, st_get_solution(c, x.data()), "st_get_solution");
            write_field(fd + "/x", P, rho.data());
            write_velocity(fd + "/velocity", P, x.data());
            write_boundary(fd, P);
            FILE *lf = std::fopen((a.out + "/0/last_frame.txt").c_str(), "w");
            if (!lf) die("cannot write last_frame.txt");
            std::fprintf(lf, "%d", it);
            std::fclose(lf);
        }
        it++;
    }
    st_destroy(c);
    return 0;
}
