// The 3-D FTCS heat update written as an OpFlow program (the time loop of
// examples/FTCS2D/FTCS-OMP.cpp:23-31 with a third dimension), compiled against
// the engine's front-end: n^3 cell-centred periodic field, i.i.d. U[-1,1] from a
// SplitMix64 seed, dt = 0.2 h^2, alpha = 1, `steps` assignments
//   u = u + dt * alpha * (d2x(u) + d2y(u) + d2z(u));
// Prints one JSON line with the loop's time (steps 2 .. steps: from the second
// assignment to the device having finished the last one) and, given argv[4], writes the final
// storage block there for the parity test (an empty argv[4] writes nothing). argv[5] adds that many
// untimed assignments after the first one, before the timed loop, so that the GPU clocks have
// settled when the timing starts; the field then holds steps + settle updates.
//
// usage: ftcs3d_example [n=512] [steps=200] [f64|f32] [out.bin] [settle=0]
#include <OpFlow>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace OpFlow;

template <typename D>
int run(int n, int steps, const char* out, int settle) {
    using Mesh = CartesianMesh<Meta::int_<3>>;
    using Field = CartesianField<D, Mesh>;
    auto mesh = MeshBuilder<Mesh>().newMesh(n + 1, n + 1, n + 1)
                        .setMeshOfDim(0, 0., 1.).setMeshOfDim(1, 0., 1.).setMeshOfDim(2, 0., 1.).build();
    auto b = ExprBuilder<Field>().setName("u").setMesh(mesh).setLoc(LocOnMesh::Center).setExt(1);
    for (int d = 0; d < 3; ++d) b.setBC(d, DimPos::start, BCType::Periodic).setBC(d, DimPos::end, BCType::Periodic);
    auto u = b.build();
    std::vector<double> init(u.storageSize());
    check(opf_gen_uniform(0xC0FFEEull + 2, init.size(), -1.0, 1.0, init.data()), "uniform");
    std::vector<D> v(init.begin(), init.end());
    u.upload(v);
    u.updatePadding();
    const Real h = 1.0 / n, dt = 0.2 * h * h, alpha = 1.0;
    auto step = [&] {
        u = u + dt * alpha * (d2x<D2SecondOrderCentered>(u) + d2y<D2SecondOrderCentered>(u) +
                              d2z<D2SecondOrderCentered>(u));
    };
    // the first assignment also allocates the field's second time level: untimed
    if (steps > 0) step();
    for (int i = 0; i < settle; ++i) step();
    deviceSync();
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 2; i <= steps; ++i) step();
    deviceSync();
    auto t1 = std::chrono::steady_clock::now();
    const double s = std::chrono::duration<double>(t1 - t0).count();
    const int timed = steps > 1 ? steps - 1 : 1;
    std::printf("{\"n\": %d, \"steps\": %d, \"settle\": %d, \"timed_steps\": %d, \"dtype\": \"%s\", "
                "\"s_loop\": %.6e, \"Mcell_updates_per_s\": %.6e}\n",
                n, steps, settle, timed, sizeof(D) == 8 ? "f64" : "f32", s, (double)n * n * n * timed / s / 1e6);
    if (out) {
        const auto& hd = u.hostData();
        FILE* f = std::fopen(out, "wb");
        std::fwrite(hd.data(), sizeof(D), hd.size(), f);
        std::fclose(f);
    }
    return 0;
}

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 512;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 200;
    const bool f32 = argc > 3 && std::strcmp(argv[3], "f32") == 0;
    const char* out = argc > 4 && argv[4][0] ? argv[4] : nullptr;
    const int settle = argc > 5 ? std::atoi(argv[5]) : 0;
    return f32 ? run<float>(n, steps, out, settle) : run<double>(n, steps, out, settle);
}
