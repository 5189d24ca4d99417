// examples/CONV1D/CONV1D.cpp compiled against the engine's front-end: 1-D linear
// advection of a square pulse, corner-located field on [0, 1] (101 nodes),
// Dirichlet 0, dt = 0.5e-2, c = 1, 100 steps of scheme 1 (first-order upwind)
//   u = u - dt * c * dx<D1FirstOrderBiasedDownwind>(u);
// and, on a second field, the same transport in conservative form with a
// flux-limiter face interpolation (D1FluxLimiter.hpp; Superbee):
//   w = w - dt * dx<D1FirstOrderCentered>(c * d1IntpCornerToCenter<0, D1Superbee>(c, w));
// The reference records u with a Tecplot stream; this program writes the final
// storage blocks of u and w to argv[1] (raw doubles) and prints one JSON line.
//
// usage: conv1d_example [out.bin] [steps=100]
#include <OpFlow>

#include <cstdio>

using namespace OpFlow;

int main(int argc, char** argv) {
    using Mesh = CartesianMesh<Meta::int_<1>>;
    using Field = CartesianField<Real, Mesh>;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 100;

    auto mesh = MeshBuilder<Mesh>().newMesh(101).setMeshOfDim(0, 0., 1.).build();
    auto u = ExprBuilder<Field>()
                     .setMesh(mesh)
                     .setName("u")
                     .setBC(0, DimPos::start, BCType::Dirc, 0.)
                     .setBC(0, DimPos::end, BCType::Dirc, 0.)
                     .setLoc(LocOnMesh::Corner)
                     .build();
    u.initBy([](auto&& i) { return 0.2 <= i[0] && i[0] <= 0.4 ? 1.0 : 0.0; });
    auto w = ExprBuilder<Field>()
                     .setMesh(mesh)
                     .setName("w")
                     .setBC(0, DimPos::start, BCType::Dirc, 0.)
                     .setBC(0, DimPos::end, BCType::Dirc, 0.)
                     .setLoc(LocOnMesh::Corner)
                     .setExt(2)
                     .build();
    w.initBy([](auto&& i) { return 0.2 <= i[0] && i[0] <= 0.4 ? 1.0 : 0.0; });

    const Real dt = 0.5e-2;
    const Real c = 1.0;
    for (auto i = 0; i < steps; ++i) {
        // Scheme 1: 1st order upwind scheme
        u = u - dt * c * dx<D1FirstOrderBiasedDownwind>(u);
        // limited, conservative: face values by the Superbee-limited upwind interpolation
        w = w - dt * dx<D1FirstOrderCentered>(c * d1IntpCornerToCenter<0, D1Superbee>(c, w));
    }
    const double umin = rangeReduce(u.assignableRange, MinOp{}, u), umax = rangeReduce(u.assignableRange, MaxOp{}, u);
    const double wmin = rangeReduce(w.assignableRange, MinOp{}, w), wmax = rangeReduce(w.assignableRange, MaxOp{}, w);
    const double wsum = rangeReduce(w.assignableRange, SumOp{}, w);
    std::printf("{\"steps\": %d, \"u_min\": %.17g, \"u_max\": %.17g, \"w_min\": %.17g, \"w_max\": %.17g, "
                "\"w_sum\": %.17g}\n",
                steps, umin, umax, wmin, wmax, wsum);
    if (argc > 1 && argv[1][0]) {
        const auto& hu = u.hostData();
        const auto& hw = w.hostData();
        FILE* f = std::fopen(argv[1], "wb");
        std::fwrite(hu.data(), sizeof(double), hu.size(), f);
        std::fwrite(hw.data(), sizeof(double), hw.size(), f);
        std::fclose(f);
    }
    return 0;
}
