// Utils::IOGroup over the raw binary streams (IOGroupTest.cpp's pattern, with
// RawBinaryOStream / RawBinaryIStream in place of H5Stream): two fields dumped at two
// time stamps, read back in order into fields of the same names, compared exactly.
//
// usage: iogroup_example <dir/>
#include <OpFlow>

#include <cstdio>

using namespace OpFlow;

int main(int argc, char** argv) {
    using Mesh = CartesianMesh<Meta::int_<2>>;
    using Field = CartesianField<Real, Mesh>;
    const std::string root = argc > 1 ? argv[1] : "./";
    auto m = MeshBuilder<Mesh>().newMesh(17, 9).setMeshOfDim(0, 0., 1.).setMeshOfDim(1, 0., 0.5).build();
    auto make = [&](const char* name) {
        return ExprBuilder<Field>()
                .setName(name)
                .setMesh(m)
                .setBC(0, DimPos::start, BCType::Dirc, 0.)
                .setBC(0, DimPos::end, BCType::Dirc, 0.)
                .setBC(1, DimPos::start, BCType::Neum, 0.)
                .setBC(1, DimPos::end, BCType::Neum, 0.)
                .setLoc({LocOnMesh::Center, LocOnMesh::Center})
                .build();
    };
    auto u = make("u"), v = make("v");
    u.initBy([](auto&& x) { return std::sin(3 * x[0]) + x[1]; });
    v.initBy([](auto&& x) { return x[0] * x[1] - 0.25; });
    auto group = Utils::makeIOGroup<Utils::RawBinaryOStream>(root, u, v);
    group.dump(Utils::TimeStamp(0.5));
    const auto u0 = u.hostData(), v0 = v.hostData();
    u = u * 2. + v;
    v = v - 1.;
    group.dump(Utils::TimeStamp(1.0));
    const auto u1 = u.hostData(), v1 = v.hostData();

    auto a = make("u"), b = make("v");
    auto rg = Utils::makeIOGroup<Utils::RawBinaryIStream>(root, StreamIn, a, b);
    bool ok = true;
    rg.read(Utils::TimeStamp(0.5));
    ok &= a.hostData() == u0 && b.hostData() == v0;
    rg.read(Utils::TimeStamp(1.0));
    ok &= a.hostData() == u1 && b.hostData() == v1;
    std::printf("{\"iogroup_roundtrip\": %s}\n", ok ? "true" : "false");
    return ok ? 0 : 1;
}
