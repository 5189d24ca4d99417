// examples/FTCS2D/FTCS-MPI.cpp-style program compiled against the engine's
// OpFlow front-end: cell-centred 256^2 field, Dirichlet 0, dt = 0.2 h^2, 1000
// steps of `u = u + dt * alpha * (d2x(u) + d2y(u))`. Writes the final storage
// block to argv[1] for the parity test and, given argv[4], the final field as
// `<dir of argv[4]>/u_0.cart`.
#include <OpFlow>

#include <chrono>
#include <cstdio>
#include <vector>

using namespace OpFlow;

int main(int argc, char** argv) {
    using Mesh = CartesianMesh<Meta::int_<2>>;
    using Field = CartesianField<Real, Mesh>;
    const int n = argc > 2 ? std::atoi(argv[2]) : 256;
    const int steps = argc > 3 ? std::atoi(argv[3]) : 1000;
    auto mesh = MeshBuilder<Mesh>().newMesh(n + 1, n + 1).setMeshOfDim(0, 0., 1.).setMeshOfDim(1, 0., 1.).build();
    auto u = ExprBuilder<Field>()
                     .setName("u")
                     .setMesh(mesh)
                     .setBC(0, DimPos::start, BCType::Dirc, 0.)
                     .setBC(0, DimPos::end, BCType::Dirc, 0.)
                     .setBC(1, DimPos::start, BCType::Dirc, 0.)
                     .setBC(1, DimPos::end, BCType::Dirc, 0.)
                     .setLoc(std::array{LocOnMesh::Center, LocOnMesh::Center})
                     .setExt(1)
                     .setPadding(1)
                     .build();
    // initial condition: the seeded bumps (4 Gaussians) on the cell centres
    std::vector<double> init(u.storageSize());
    check(opf_gen_bumps(mesh.h.get(), &u.desc, 0xC0FFEEull + 1, 4, 0.0, init.data()), "bumps");
    u.upload(init);
    u.updatePadding();
    const Real dt = 0.2 / ((double)n * n), alpha = 1.0;
    auto step = [&] { u = u + dt * alpha * (d2x<D2SecondOrderCentered>(u) + d2y<D2SecondOrderCentered>(u)); };
    // the first assignment also allocates the field's second time level and loads the
    // kernels (first use in the process): timed on its own
    auto t0 = std::chrono::steady_clock::now();
    if (steps > 0) step();
    deviceSync();
    auto t1 = std::chrono::steady_clock::now();
    for (int i = 2; i <= steps; ++i) step();
    const auto& h = u.hostData();
    auto t2 = std::chrono::steady_clock::now();
    std::printf("Elapsed time: %lld ms\n",
                (long long)std::chrono::duration_cast<std::chrono::milliseconds>(t2 - t0).count());
    std::printf("{\"n\": %d, \"steps\": %d, \"s_loop_and_read\": %.6e, \"s_first_step\": %.6e, "
                "\"s_steps_2_to_n_and_read\": %.6e}\n",
                n, steps, std::chrono::duration<double>(t2 - t0).count(),
                std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count());
    std::printf("Center val: %.17g\n", u.evalAt(DS::MDIndex<2>{n / 2, n / 2}));
    // the same update through the generic interpreter path (no pattern match): u2 = u2 + (d2x + d2y) * c
    auto u2 = ExprBuilder<Field>().setName("u2").setMesh(mesh)
                      .setBC(0, DimPos::start, BCType::Dirc, 0.).setBC(0, DimPos::end, BCType::Dirc, 0.)
                      .setBC(1, DimPos::start, BCType::Dirc, 0.).setBC(1, DimPos::end, BCType::Dirc, 0.)
                      .setLoc(LocOnMesh::Center).setExt(1).build();
    u2.upload(init);
    u2.updatePadding();
    for (int i = 1; i <= 5; ++i) u2 += dt * alpha * (d2x<D2SecondOrderCentered>(u2) + d2y<D2SecondOrderCentered>(u2));
    if (argc > 4) {  // the final field as the reference's examples record it (Utils::RawBinaryOStream)
        Utils::RawBinaryOStream os(argv[4]);
        os << Utils::TimeStamp(steps * dt) << u;
    }
    if (argc > 1 && argv[1][0]) {
        FILE* f = std::fopen(argv[1], "wb");
        std::fwrite(h.data(), sizeof(double), h.size(), f);
        const auto& h2 = u2.hostData();
        std::fwrite(h2.data(), sizeof(double), h2.size(), f);
        std::fwrite(init.data(), sizeof(double), init.size(), f);
        std::fclose(f);
    }
    return 0;
}
