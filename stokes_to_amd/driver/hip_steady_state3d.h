// HipSteadyState3D -- the drop-in a maintainer of the reference adds in place of
// FluidSteadyState3D (ST/FluidSteadyState3D.h:11-39): the same five calls, in the
// same order, backed by libstokes_hip.so through the C ABI of include/stokes_to.h.
//
// The class is written against the members of the reference's host fields that it
// uses: Field<T,3>::array (x-major values, SX/reservoir/Field.h:16) and
// FaceField<T,3>::face_fields[d] (SX/reservoir/FaceField.h:16).  Those layouts are the
// ABI layout, so the binding only concatenates [cells | faces_x | faces_y | faces_z].
// The owner sizes vel and p (the reference's init calls Resize on them); solve() fills
// their arrays.
//
//   reference (FT/FluidTopoOpt3D.cpp:88-90, :145-168)    this binding
//   steady.init(N)                                        st_create
//   steady.init_boundary(cv, cf, fv, ff, closed)          st_set_geometry(fixed, closed)
//   steady.update_penalty(cp, fp)                         st_set_penalty
//   steady.update_b(cb, fb)                               st_set_rhs(body, no Dirichlet)
//   steady.solve(); steady.vel, steady.p                  st_solve(ST_INIT_ZERO) + unpack
#pragma once
#include "stokes_to.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

template <template <class, int> class FieldT, template <class, int> class FaceFieldT, class Scalar = double>
class HipSteadyState3D {
public:
    FaceFieldT<Scalar, 3> vel;
    FieldT<Scalar, 3> p;
    st_stats stats{};  // iterations, true FP64 residual, time of the last solve

    // FluidSteadyState3D::init (ST/FluidSteadyState3D.cpp:12-90).  The reference's
    // penalties arrive pre-scaled (unit spacing), so alpha_scale stays 1.
    void init(int n, double rtol = 1e-8)
    {
        st_params prm;
        st_default_params(&prm, 3, n);
        prm.alpha_scale = 1.0;
        prm.rtol = rtol;
        if (ctx) st_destroy(ctx);
        ctx = nullptr;
        int rc = st_create(&ctx, &prm);
        if (rc != ST_OK) throw std::runtime_error("st_create failed: " + std::to_string(rc));
        N = n;
        ndof = st_num_dofs(ctx);
    }
    // FluidSteadyState3D::init_boundary (ST/FluidSteadyState3D.cpp:92-114).  vol is the
    // free mask on every face (StokeFlowDescriptor3D::finish); closed sets the pressure
    // gauge of the coarsest matrix here, where the reference assigns it.
    void init_boundary(const FieldT<Scalar, 3> &, const FieldT<int, 3> &cell_fixed, const FaceFieldT<Scalar, 3> &,
                       const FaceFieldT<int, 3> &face_fixed, bool closed)
    {
        std::vector<uint8_t> fixed;
        fixed.reserve((size_t)ndof);
        for (int v : cell_fixed.array) fixed.push_back(v != 0);
        for (int d = 0; d < 3; d++)
            for (int v : face_fixed.face_fields[d].array) fixed.push_back(v != 0);
        check_size(fixed.size());
        check(st_set_geometry(ctx, fixed.data(), closed ? 1 : 0));
    }
    // FluidSteadyState3D::update_penalty (ST/FluidSteadyState3D.cpp:130-150)
    void update_penalty(const FieldT<Scalar, 3> &cell_penalty, const FaceFieldT<Scalar, 3> &face_penalty)
    {
        const std::vector<double> pen = flatten(cell_penalty, face_penalty);
        check(st_set_penalty(ctx, pen.data()));
    }
    // FluidSteadyState3D::update_b (ST/FluidSteadyState3D.cpp:116-127)
    void update_b(const FieldT<Scalar, 3> &cell_b, const FaceFieldT<Scalar, 3> &face_b)
    {
        const std::vector<double> b = flatten(cell_b, face_b);
        check(st_set_rhs(ctx, b.data(), nullptr));
    }
    // FluidSteadyState3D::solve (ST/FluidSteadyState3D.cpp:152-170): from zero like the
    // reference's GMRES; the result goes to vel and p
    void solve()
    {
        std::vector<double> x((size_t)ndof);
        check(st_solve(ctx, ST_INIT_ZERO, nullptr, x.data(), &stats));
        size_t o = 0;
        p.array.resize((size_t)N * N * N);
        for (auto &v : p.array) v = (Scalar)x[o++];
        for (int d = 0; d < 3; d++) {
            auto &a = vel.face_fields[d].array;
            a.resize((size_t)(N + 1) * N * N);
            for (auto &v : a) v = (Scalar)x[o++];
        }
    }
    ~HipSteadyState3D()
    {
        if (ctx) st_destroy(ctx);
    }

private:
    st_ctx *ctx = nullptr;
    int N = 0;
    long ndof = 0;

    std::vector<double> flatten(const FieldT<Scalar, 3> &c, const FaceFieldT<Scalar, 3> &f) const
    {
        std::vector<double> out;
        out.reserve((size_t)ndof);
        out.insert(out.end(), c.array.begin(), c.array.end());
        for (int d = 0; d < 3; d++) out.insert(out.end(), f.face_fields[d].array.begin(), f.face_fields[d].array.end());
        check_size(out.size());
        return out;
    }
    void check_size(size_t m) const
    {
        if ((long)m != ndof) throw std::runtime_error("field sizes do not match the grid of init()");
    }
    void check(int rc) const
    {
        if (rc != ST_OK) throw std::runtime_error(std::string("libstokes_hip: ") + st_last_error(ctx));
    }
};
