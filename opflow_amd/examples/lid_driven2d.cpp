// Lid-driven cavity with the OpFlow front-end on the MI355X engine: staggered
// (MAC) grid, explicit central-difference advection and diffusion, then a
// pressure Poisson projection with Neumann BCs and a pinned first unknown.
// Operator vocabulary follows examples/LidDriven/LidDriven2D.cpp:29-41 and its
// pressure handler (:67-74); the momentum step is explicit here.
//
// usage: lid_driven2d [n=1024] [steps=500] [mixed=1] [out.bin] [dump_dir dump_steps]
//   dump_steps: comma-separated step indices; at each, the storage blocks of
//   u, v before the step (<dir>/s<i>_u_n.bin, _v_n), after the momentum update
//   (_u_star, _v_star), p after the solve (_p) and u, v after the projection
//   (_u_np1, _v_np1) are written as raw doubles, for checks outside the engine.
#include <OpFlow>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <string>

using namespace OpFlow;

int main(int argc, char** argv) {
    using Mesh = CartesianMesh<Meta::int_<2>>;
    using Field = CartesianField<Real, Mesh>;
    const int n = argc > 1 ? std::atoi(argv[1]) : 1024;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 500;
    const bool mixed = argc > 3 ? std::atoi(argv[3]) != 0 : true;
    const char* out = argc > 4 && argv[4][0] ? argv[4] : nullptr;
    const std::string dump_dir = argc > 5 ? argv[5] : "";
    std::set<int> dump_steps;
    if (argc > 6)
        for (const char* c = argv[6]; *c;) {
            char* e;
            dump_steps.insert((int)std::strtol(c, &e, 10));
            c = *e ? e + 1 : e;
        }
    auto dump = [&](int step, const char* tag, const auto& f) {
        const auto& h = f.hostData();
        std::string path = dump_dir + "/s" + std::to_string(step) + "_" + tag + ".bin";
        FILE* fp = std::fopen(path.c_str(), "wb");
        if (!fp || std::fwrite(h.data(), sizeof(h[0]), h.size(), fp) != h.size()) {
            std::fprintf(stderr, "cannot write %s\n", path.c_str());
            std::exit(1);
        }
        std::fclose(fp);
    };

    auto mesh = MeshBuilder<Mesh>().newMesh(n + 1, n + 1).setMeshOfDim(0, 0., 1.).setMeshOfDim(1, 0., 1.).build();
    auto builder = ExprBuilder<Field>().setMesh(mesh).setExt(1)
                           .setBC(0, DimPos::start, BCType::Dirc, 0.).setBC(0, DimPos::end, BCType::Dirc, 0.)
                           .setBC(1, DimPos::start, BCType::Dirc, 0.).setBC(1, DimPos::end, BCType::Dirc, 0.);
    auto u = builder.setName("u").setBC(1, DimPos::end, BCType::Dirc, 1.)
                     .setLoc({LocOnMesh::Corner, LocOnMesh::Center}).build();
    auto du = u;
    du.name = "du";
    auto v = builder.setName("v").setBC(1, DimPos::end, BCType::Dirc, 0.)
                     .setLoc({LocOnMesh::Center, LocOnMesh::Corner}).build();
    auto dv = v;
    dv.name = "dv";
    auto p = builder.setName("p")
                     .setBC(0, DimPos::start, BCType::Neum, 0.).setBC(0, DimPos::end, BCType::Neum, 0.)
                     .setBC(1, DimPos::start, BCType::Neum, 0.).setBC(1, DimPos::end, BCType::Neum, 0.)
                     .setLoc({LocOnMesh::Center, LocOnMesh::Center}).build();
    u = 0; du = 0; v = 0; dv = 0; p = 0;

    // composite operators (LidDriven2D.cpp:29-41)
    auto conv_xx = [&](auto&& _1, auto&& _2) {
        return dx<D1FirstOrderCentered>(d1IntpCornerToCenter<0>(_1) * d1IntpCornerToCenter<0>(_2));
    };
    auto conv_xy = [&](auto&& _1, auto&& _2) {
        return dy<D1FirstOrderCentered>(d1IntpCenterToCorner<1>(_1) * d1IntpCenterToCorner<0>(_2));
    };
    auto conv_yx = [&](auto&& _1, auto&& _2) {
        return dx<D1FirstOrderCentered>(d1IntpCenterToCorner<1>(_1) * d1IntpCenterToCorner<0>(_2));
    };
    auto conv_yy = [&](auto&& _1, auto&& _2) {
        return dy<D1FirstOrderCentered>(d1IntpCornerToCenter<1>(_1) * d1IntpCornerToCenter<1>(_2));
    };

    const Real h = 1.0 / n, nu = 1.0e-3, U = 1.0;
    const Real dt = std::min(0.25 * h * h / nu, 0.5 * h / U);
    StructSolverParams<StructSolverType::PCG> pp;
    pp.tol = 1e-8;
    pp.staticMat = true;
    pp.pinValue = true;
    pp.mixedPrecision = mixed;
    auto p_solver = PrecondStructSolver<StructSolverType::PCG, StructSolverType::PFMG>(pp, {});
    auto p_handler = makeEqnSolveHandler(
            [&](auto&& e) {
                return d2x<D2SecondOrderCentered>(e) + d2y<D2SecondOrderCentered>(e)
                       == (dx<D1FirstOrderCentered>(u) + dy<D1FirstOrderCentered>(v)) / dt;
            },
            p, p_solver);

    long long iters = 0;
    double worst = 0;
    // the first step also builds the pressure hierarchy and compiles the expression kernels
    // (once per process / machine): it is timed on its own, the rest as the steady rate
    auto t0 = std::chrono::steady_clock::now();
    auto t_first = t0;
    for (int i = 0; i < steps; ++i) {
        if (i == 1) {
            opf_ctx_sync(globalContext()->h);
            t_first = std::chrono::steady_clock::now();
        }
        const bool dumping = !dump_dir.empty() && dump_steps.count(i);
        if (dumping) dump(i, "u_n", u), dump(i, "v_n", v);
        du = dt * (nu * (d2x<D2SecondOrderCentered>(u) + d2y<D2SecondOrderCentered>(u)) - (conv_xx(u, u) + conv_xy(u, v)));
        dv = dt * (nu * (d2x<D2SecondOrderCentered>(v) + d2y<D2SecondOrderCentered>(v)) - (conv_yx(u, v) + conv_yy(v, v)));
        u = u + du;
        v = v + dv;
        if (dumping) dump(i, "u_star", u), dump(i, "v_star", v);
        auto st = p_handler->solve();
        iters += st.niter;
        worst = std::max(worst, st.relerr);
        if (dumping) {
            dump(i, "p", p);
            std::printf("{\"dump_step\": %d, \"niter\": %d, \"relres\": %.6e}\n", i, st.niter, st.relerr);
        }
        u = u - dt * dx<D1FirstOrderCentered>(p);
        v = v - dt * dy<D1FirstOrderCentered>(p);
        if (dumping) dump(i, "u_np1", u), dump(i, "v_np1", v);
    }
    const auto& hu = u.hostData();
    const auto& hv = v.hostData();
    auto t1 = std::chrono::steady_clock::now();
    const double first = std::chrono::duration<double>(t_first - t0).count();
    const double rest = std::chrono::duration<double>(t1 - t_first).count();
    const double per_step = steps > 1 ? rest / (steps - 1) : first;
    auto div = ExprBuilder<Field>().setMesh(mesh).setName("div").setLoc(LocOnMesh::Center).setExt(1)
                       .setBC(0, DimPos::start, BCType::Neum, 0.).setBC(0, DimPos::end, BCType::Neum, 0.)
                       .setBC(1, DimPos::start, BCType::Neum, 0.).setBC(1, DimPos::end, BCType::Neum, 0.).build();
    div = dx<D1FirstOrderCentered>(u) + dy<D1FirstOrderCentered>(v);
    double divmax = std::max(std::fabs(rangeReduce(div.assignableRange, MaxOp{}, div)),
                             std::fabs(rangeReduce(div.assignableRange, MinOp{}, div)));
    std::printf("{\"n\": %d, \"steps\": %d, \"mixed\": %d, \"dt\": %.6e, \"s_per_step\": %.6e, \"s_first_step\": %.6e, "
                "\"poisson_iters\": %lld, \"worst_relres\": %.3e, \"max_div\": %.3e, \"u_center\": %.17g}\n",
                n, steps, (int)mixed, dt, per_step, first, iters, worst, divmax, u.evalAt(DS::MDIndex<2>{n / 2, n / 2}));
    if (out) {
        FILE* f = std::fopen(out, "wb");
        std::fwrite(hu.data(), sizeof(double), hu.size(), f);
        std::fwrite(hv.data(), sizeof(double), hv.size(), f);
        std::fclose(f);
    }
    return 0;
}
