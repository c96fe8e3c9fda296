// Runs the reference's call sequence through the HipSteadyState3D binding
// (hip_steady_state3d.h) on a problem read from raw files, and writes the solution:
//
//   binding_test0 N DIR
//     DIR/cell_fixed, DIR/face_fixed  int32, [cells] and [faces_x | faces_y | faces_z]
//     DIR/cell_b, DIR/face_b          float64 source terms
//     DIR/cell_pen, DIR/face_pen      float64 penalties
//     DIR/closed                      text "0" or "1"
//   -> DIR/p, DIR/vel (float64) and one line "iters relres" on stdout.
//
// The calls are those of FluidTopoOpt3D (FT/FluidTopoOpt3D.cpp:88-90, :145-168):
// init -> init_boundary(..., closed) -> update_penalty -> update_b -> solve.  The two
// structs below hold only the members of the reference's Field / FaceField that the
// binding reads (SX/reservoir/Field.h:16, SX/reservoir/FaceField.h:16).
#include "hip_steady_state3d.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

template <class T, int d> struct Field {
    std::vector<T> array;
};
template <class T, int d> struct FaceField {
    Field<T, d> face_fields[d];
};

template <class T> static std::vector<T> read_raw(const std::string &path, size_t n)
{
    std::vector<T> v(n);
    std::ifstream f(path, std::ios::binary);
    if (!f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(T))))
        throw std::runtime_error("cannot read " + path);
    return v;
}

template <class T> static void split(const std::vector<T> &all, int n, FaceField<T, 3> &f)
{
    const size_t m = (size_t)(n + 1) * n * n;
    for (int d = 0; d < 3; d++) f.face_fields[d].array.assign(all.begin() + d * m, all.begin() + (d + 1) * m);
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: binding_test0 N DIR\n");
        return 2;
    }
    try {
        const int n = std::atoi(argv[1]);
        const std::string dir = argv[2];
        const size_t nc = (size_t)n * n * n, nf = 3 * (size_t)(n + 1) * n * n;
        Field<int, 3> cf;
        Field<double, 3> cv, cb, cp;
        FaceField<int, 3> ff;
        FaceField<double, 3> fv, fb, fp;
        cf.array = read_raw<int>(dir + "/cell_fixed", nc);
        split(read_raw<int>(dir + "/face_fixed", nf), n, ff);
        cb.array = read_raw<double>(dir + "/cell_b", nc);
        split(read_raw<double>(dir + "/face_b", nf), n, fb);
        cp.array = read_raw<double>(dir + "/cell_pen", nc);
        split(read_raw<double>(dir + "/face_pen", nf), n, fp);
        int closed = 0;
        std::ifstream(dir + "/closed") >> closed;
        cv.array.assign(nc, 0.0);  // vol: the reference passes it; the ABI derives it from fixed
        for (int d = 0; d < 3; d++) {
            auto &a = fv.face_fields[d].array;
            a.resize(ff.face_fields[d].array.size());
            for (size_t q = 0; q < a.size(); q++) a[q] = ff.face_fields[d].array[q] ? 0.0 : 1.0;
        }

        HipSteadyState3D<Field, FaceField> steady;
        steady.init(n, 1e-10);
        steady.init_boundary(cv, cf, fv, ff, closed != 0);
        steady.update_penalty(cp, fp);
        steady.update_b(cb, fb);
        steady.solve();

        std::ofstream(dir + "/p", std::ios::binary)
            .write(reinterpret_cast<const char *>(steady.p.array.data()), (std::streamsize)(nc * sizeof(double)));
        std::ofstream v(dir + "/vel", std::ios::binary);
        for (int d = 0; d < 3; d++)
            v.write(reinterpret_cast<const char *>(steady.vel.face_fields[d].array.data()),
                    (std::streamsize)(steady.vel.face_fields[d].array.size() * sizeof(double)));
        std::printf("%d %.3e\n", steady.stats.iters, steady.stats.relres);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "binding_test0: %s\n", e.what());
        return 1;
    }
    return 0;
}
