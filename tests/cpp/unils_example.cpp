// examples/LevelSet/UniLS.cpp (ls(), the 2-D case) compiled against the engine's
// front-end: a circle of radius 0.15 advected by a swirling velocity field with
// a third-order Runge-Kutta step of upwind-selected WENO53 derivatives, then
// four iterations of the WENO reinitialisation with its volume correction (a
// 3 x 3 convolution of smoothDelta terms). Written as the reference writes it,
// with three changes: the fields carry three ghost layers (setExt(3)), which
// WENO53 reads at the boundary cells; the smoothed delta is
// Math::smoothDelta(h, e) on an expression (the device has no way to run the
// reference's host functor through UniOpAdaptor / NamedFunctor); and the
// Tecplot output is replaced by writing the final storage block of p to
// argv[2] (raw doubles).
//
// usage: unils_example [steps=128] [out.bin]
#include <OpFlow>

#include <cstdio>
#include <cstdlib>

using namespace OpFlow;

template <std::size_t d>
using DU = D1WENO53Upwind<d>;
template <std::size_t d>
using DD = D1WENO53Downwind<d>;

int main(int argc, char** argv) {
    using Mesh = CartesianMesh<Meta::int_<2>>;
    using Field = CartesianField<Real, Mesh>;

    constexpr auto n = 65;
    auto m = MeshBuilder<Mesh>().newMesh(n, n).setMeshOfDim(0, 0., 1.).setMeshOfDim(1, 0., 1.).build();
    auto u = ExprBuilder<Field>()
                     .setName("u")
                     .setMesh(m)
                     .setBC(0, DimPos::start, BCType::Dirc, 0.)
                     .setBC(0, DimPos::end, BCType::Dirc, 0.)
                     .setBC(1, DimPos::start, BCType::Dirc, 0.)
                     .setBC(1, DimPos::end, BCType::Dirc, 0.)
                     .setLoc({LocOnMesh::Center, LocOnMesh::Center})
                     .setExt(3)
                     .build();
    auto v = u;
    v.name = "v";
    auto p = ExprBuilder<Field>()
                     .setName("p")
                     .setMesh(m)
                     .setBC(0, DimPos::start, BCType::Neum, 0.)
                     .setBC(0, DimPos::end, BCType::Neum, 0.)
                     .setBC(1, DimPos::start, BCType::Neum, 0.)
                     .setBC(1, DimPos::end, BCType::Neum, 0.)
                     .setLoc({LocOnMesh::Center, LocOnMesh::Center})
                     .setExt(3)
                     .build();

    p.initBy([](auto&& x) { return std::sqrt(Math::pow2(x[0] - 0.5) + Math::pow2(x[1] - 0.75)) - 0.15; });
    auto p1 = p, p2 = p, p3 = p, p0 = p;

    u.initBy([](auto&& x) {
        return 2 * std::sin(PI * x[1]) * std::cos(PI * x[1]) * Math::pow2(std::sin(PI * x[0]));
    });
    v.initBy([](auto&& x) {
        return -2 * std::sin(PI * x[0]) * std::cos(PI * x[0]) * Math::pow2(std::sin(PI * x[1]));
    });

    constexpr auto h = 1. / (n - 1);
    auto _eps = 1e-6;
    auto _1 = [&](auto&& _p, auto&& _pp) {
        return _p / OpFlow::sqrt(_p * _p + _eps)
               * (OpFlow::sqrt(
                          OpFlow::pow(
                                  OpFlow::max(-OpFlow::min(0., dx<DU>(_pp)), OpFlow::max(dx<DD>(_pp), 0.)), 2)
                          + OpFlow::pow(
                                  OpFlow::max(-OpFlow::min(0., dy<DU>(_pp)), OpFlow::max(dy<DD>(_pp), 0.)),
                                  2))
                  - 1.);
    };
    auto _2 = [&](auto&& _p, auto&& _pp) {
        return _p / OpFlow::sqrt(_p * _p + _eps)
               * (OpFlow::sqrt(
                          OpFlow::pow(
                                  OpFlow::max(OpFlow::max(0., dx<DU>(_pp)), -OpFlow::min(dx<DD>(_pp), 0.)), 2)
                          + OpFlow::pow(
                                  OpFlow::max(OpFlow::max(0., dy<DU>(_pp)), -OpFlow::min(dy<DD>(_pp), 0.)),
                                  2))
                  - 1.);
    };

    const int steps = argc > 1 ? std::atoi(argv[1]) : 128;
    auto dt = 1. / (n - 1);
    for (auto i = 0; i < 2. / dt && i < steps; ++i) {
        p1 = p
             - dt
                       * (u * conditional(u > 0., dx<DD>(p), dx<DU>(p))
                          + v * conditional(v > 0., dy<DD>(p), dy<DU>(p)));
        p2 = (p1
              - dt
                        * (u * conditional(u > 0., dx<DD>(p1), dx<DU>(p1))
                           + v * conditional(v > 0., dy<DD>(p), dy<DU>(p))))
                     / 4.
             + 0.75 * p;
        p3 = (p2
              - dt
                        * (u * conditional(u > 0., dx<DD>(p2), dx<DU>(p2))
                           + v * conditional(v > 0., dy<DD>(p2), dy<DU>(p2))))
                     * 2. / 3.
             + p / 3.;
        p = p3;
        p0 = p;
        // reinit
        for (auto _ = 0; _ < 4; ++_) {
            auto h1 = conditional(p0 > 0, _1(p0, p), _2(p0, p));
            p1 = p - dt * h1;
            auto h2 = conditional(p0 > 0, _1(p0, p1), _2(p0, p1));
            p2 = p1 - dt / 4. * (-3 * h1 + h2);
            auto h3 = conditional(p0 > 0, _1(p0, p2), _2(p0, p2));
            p3 = p2 - dt / 12. * (-h1 - h2 + 8 * h3);
            constexpr auto _c = 16. / 24., _o = 1. / 24.;
            constexpr DS::FixedSizeTensor<double, 3, 3> conv_ker {_o, _o, _o, _o, _c, _o, _o, _o, _o};
            auto delta_op = [=](auto&& e) { return Math::smoothDelta(h, e); };
            auto int_op = [=](auto&& e) { return h * h * conv(e, conv_ker); };

            auto lambda = -int_op(delta_op(p0) * (p3 - p0) / (_ + 1)) / int_op(pow(delta_op(p0), 2) + 1e-14);
            p = p3 + lambda * (_ + 1) * delta_op(p0);
        }
    }
    double pmin = rangeReduce(p.assignableRange, MinOp{}, p), pmax = rangeReduce(p.assignableRange, MaxOp{}, p);
    std::printf("{\"steps\": %d, \"p_min\": %.17g, \"p_max\": %.17g}\n", steps, pmin, pmax);
    if (argc > 2 && argv[2][0]) {
        const auto& hp = p.hostData();
        FILE* f = std::fopen(argv[2], "wb");
        std::fwrite(hp.data(), sizeof(double), hp.size(), f);
        std::fclose(f);
    }
    return 0;
}
