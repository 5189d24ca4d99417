// Expression-tree lowering and the stack-program sweep kernel behind
// opf_sweep (Expr::operator= -> FieldAssigner::assign -> rangeFor(evalAt),
// Loops/FieldAssigner.hpp:26-86).
//
// Lowering inlines every operator at the shifted indices it reads, exactly as
// the reference's lazy evalAt recursion does (Expr/Expression.hpp:51-65), and
// spells out each operator's mesh arithmetic in the reference's operation
// order, so FP64 results match OpFlow's evaluation bit for bit. The program
// is a straight line of stack instructions; every lane runs the same
// instruction stream (uniform control flow) and keeps the stack in LDS.
#include <algorithm>
#include <cmath>

#include "opf_internal.hpp"
#include "sweep.hpp"

using namespace opf;

namespace opf {

// OpFlow's PI (Core/Constants.hpp:37)
constexpr double kPi = 3.141592653589793;

struct Lowerer {
    const opf_expr_node* nodes;
    int nnodes;
    opf_field* const* fields;
    int nfields;
    const opf_mesh* mesh;
    const int* unknown_slots = nullptr;  // field slot of each OPF_N_UNKNOWN index (its layout)
    int nunknown = 0;
    bool unknown_as_field = false;  // read OPF_N_UNKNOWN as an ordinary field (bias evaluation)
    std::vector<Insn> code;
    int depth = 0, maxdepth = 0;
    // per field, per dim min/max read offset relative to the evaluation index
    int omin[SWEEP_MAX_FIELDS][3], omax[SWEEP_MAX_FIELDS][3];
    int tmin[3], tmax[3];  // mesh table shifts
    int err = OPF_OK;
    std::vector<std::array<int, 3>> locs;

    void init() {
        for (int f = 0; f < SWEEP_MAX_FIELDS; ++f)
            for (int d = 0; d < 3; ++d) {
                omin[f][d] = 1 << 20;
                omax[f][d] = -(1 << 20);
            }
        for (int d = 0; d < 3; ++d) {
            tmin[d] = 1 << 20;
            tmax[d] = -(1 << 20);
        }
    }
    void push(int dd) {
        depth += dd;
        maxdepth = std::max(maxdepth, depth);
    }
    void op(int o) {
        code.push_back({o, 0, 0, 0, 0.0});
        push(o == I_NEG || o == I_ABS ? 0 : o == I_WENO_UP || o == I_WENO_DOWN ? -6 : o == I_SEL ? -2 : -1);
    }
    void cst(double v) {
        code.push_back({I_CONST, 0, 0, 0, v});
        push(1);
    }
    void fn1(int f) { code.push_back({I_FN1, f, 0, 0, 0.0}); }
    // a ScalarExpr node: one launch-time slot per node (the node may be inlined at
    // several shifts); past SWEEP_MAX_CONSTS distinct nodes the value is baked in
    std::vector<int> cslot;
    std::vector<double> consts;
    void ucst(int n, double v) {
        if (cslot.empty()) cslot.assign(nnodes, -1);
        if (cslot[n] < 0 && (int)consts.size() < SWEEP_MAX_CONSTS) {
            cslot[n] = (int)consts.size();
            consts.push_back(v);
        }
        if (cslot[n] < 0) return cst(v);
        code.push_back({I_CONST, 1, cslot[n], 0, v});
        push(1);
    }
    void tab(int kind, int d, int shift) {
        code.push_back({kind, d, shift, 0, 0.0});
        tmin[d] = std::min(tmin[d], shift);
        tmax[d] = std::max(tmax[d], shift);
        push(1);
    }
    void dx(int d, int s) { tab(I_DX, d, s); }
    void x(int d, int s) { tab(I_X, d, s); }

    int compute_locs() {
        locs.assign(nnodes, {1, 1, 1});
        for (int n = 0; n < nnodes; ++n) {
            const opf_expr_node& e = nodes[n];
            auto kid = [&](int c) -> bool { return c >= 0 && c < n; };
            switch (e.kind) {
                case OPF_N_FIELD:
                    if (e.a < 0 || e.a >= nfields) return OPF_ERR_ARG;
                    for (int d = 0; d < 3; ++d) locs[n][d] = fields[e.a]->d.loc[d];
                    break;
                case OPF_N_UNKNOWN:
                    if (e.a < 0 || e.a >= nunknown || unknown_slots[e.a] < 0 || unknown_slots[e.a] >= nfields)
                        return OPF_ERR_ARG;
                    for (int d = 0; d < 3; ++d) locs[n][d] = fields[unknown_slots[e.a]]->d.loc[d];
                    break;
                case OPF_N_CONST: locs[n] = {-1, -1, -1}; break;
                case OPF_N_ADD:
                case OPF_N_SUB:
                case OPF_N_MUL:
                case OPF_N_DIV:
                case OPF_N_MIN:
                case OPF_N_MAX:
                case OPF_N_GT:
                case OPF_N_LT:
                case OPF_N_GE:
                case OPF_N_LE:
                    if (!kid(e.a) || !kid(e.b)) return OPF_ERR_ARG;
                    locs[n] = locs[e.a][0] >= 0 ? locs[e.a] : locs[e.b];
                    break;
                case OPF_N_COND:  // the location of the value it selects (Conditional.hpp: initPropsFrom(arg2))
                    if (!kid(e.a) || !kid(e.b) || !kid(e.dim)) return OPF_ERR_ARG;
                    locs[n] = locs[e.b][0] >= 0 ? locs[e.b] : locs[e.dim];
                    break;
                case OPF_N_ABS:
                    if (!kid(e.a)) return OPF_ERR_ARG;
                    locs[n] = locs[e.a];
                    break;
                case OPF_N_FUNC1:
                    if (!kid(e.a) || e.dim < 0 || e.dim > OPF_FN_SMOOTH_HEAVISIDE) return OPF_ERR_ARG;
                    if (e.dim >= OPF_FN_SMOOTH_DELTA && !(e.value > 0)) return OPF_ERR_ARG;
                    locs[n] = locs[e.a];
                    break;
                case OPF_N_FUNC2:
                    if (!kid(e.a) || !kid(e.b) || e.dim < 0 || e.dim > OPF_FN2_NEXTAFTER) return OPF_ERR_ARG;
                    locs[n] = locs[e.a][0] >= 0 ? locs[e.a] : locs[e.b];
                    break;
                case OPF_N_CONV: {  // Convolution::prepare: the argument's location
                    const int k[3] = {e.dim & 255, (e.dim >> 8) & 255, (e.dim >> 16) & 255};
                    const int K = k[0] * k[1] * k[2];
                    if (!kid(e.a) || K < 1 || e.b < 0 || e.b + K > n) return OPF_ERR_ARG;
                    for (int d = 0; d < 3; ++d)
                        if (k[d] % 2 == 0 || (d >= mesh->dim && k[d] != 1)) return OPF_ERR_ARG;
                    for (int q = 0; q < K; ++q)
                        if (nodes[e.b + q].kind != OPF_N_CONST) return OPF_ERR_ARG;
                    locs[n] = locs[e.a];
                } break;
                case OPF_N_FLUXLIM: {
                    // the interpolated argument b decides the direction (D1FluxLimiter.hpp prepare)
                    if (!kid(e.a) || !kid(e.b) || e.dim < 0 || e.dim >= mesh->dim) return OPF_ERR_ARG;
                    const int code = (int)e.value, lim = code % 32, dir = code / 32;
                    if ((double)code != e.value || code < 0 || lim > OPF_LIM_BSOU || dir > 2) return OPF_ERR_ARG;
                    // Cen2Cor needs a centre-located argument, Cor2Cen a corner-located one
                    if (dir != 0 && locs[e.b][e.dim] >= 0 &&
                        locs[e.b][e.dim] != (dir == 1 ? OPF_LOC_CENTER : OPF_LOC_CORNER))
                        return OPF_ERR_ARG;
                    locs[n] = locs[e.b];
                    if (locs[n][e.dim] >= 0)
                        locs[n][e.dim] = locs[n][e.dim] == OPF_LOC_CENTER ? OPF_LOC_CORNER : OPF_LOC_CENTER;
                } break;
                default:
                    if (!kid(e.a)) return OPF_ERR_ARG;
                    if (e.kind != OPF_N_NEG && (e.dim < 0 || e.dim >= mesh->dim)) return OPF_ERR_ARG;
                    locs[n] = locs[e.a];
                    if (e.kind == OPF_N_D1C && locs[n][e.dim] >= 0)
                        locs[n][e.dim] = locs[n][e.dim] == OPF_LOC_CENTER ? OPF_LOC_CORNER : OPF_LOC_CENTER;
                    if (e.kind == OPF_N_INTP_C2N) locs[n][e.dim] = OPF_LOC_CORNER;
                    if (e.kind == OPF_N_INTP_N2C) locs[n][e.dim] = OPF_LOC_CENTER;
                    break;
            }
        }
        return OPF_OK;
    }

    // D1FluxLimiterImpl<Kernel, d, dir>::eval (Interpolator/D1FluxLimiter.hpp:43-202) at shift sh:
    //   u[i] > 0 ? upwind : downwind, each  y2 +- dx * 0.5 * K, in the reference's operation order
    void emit_fluxlim(const opf_expr_node& e, const int sh[3]) {
        const int d = e.dim, lim = (int)e.value % 32;
        const bool c2n = locs[e.b][d] == OPF_LOC_CENTER;  // Cen2Cor: b centre-located in d
        const int i = sh[d];
        auto val = [&](int k) {  // b at index i + k along d
            int t[3] = {sh[0], sh[1], sh[2]};
            t[d] += k;
            emit(e.b, t[0], t[1], t[2]);
        };
        // node coordinate (Cor2Cen) or cell centre x + 0.5 dx (Cen2Cor) at i + k
        auto xp = [&](int k) {
            x(d, i + k);
            if (c2n) { cst(0.5); dx(d, i + k); op(I_MUL); op(I_ADD); }
        };
        // (yb - ya) / (xb - xa) for points i + ka, i + kb
        auto slope = [&](int ka, int kb) {
            val(kb); val(ka); op(I_SUB);
            xp(kb); xp(ka); op(I_SUB);
            op(I_DIV);
        };
        // one branch: points p1 < p2 < p3 (offsets), slop_u / slop_f pairs, the dx index and sign
        auto branch = [&](bool up) {
            int k1, k2, k3, kd;
            if (c2n) { k1 = up ? -2 : -1; kd = up ? -1 : 0; }
            else { k1 = up ? -1 : 0; kd = 0; }
            k2 = k1 + 1;
            k3 = k1 + 2;
            // upwind: slop_u = (y2 - y1)/(x2 - x1), slop_f = (y3 - y2)/(x3 - x2); downwind swaps them
            auto su = [&] { up ? slope(k1, k2) : slope(k2, k3); };
            auto sf = [&] { up ? slope(k2, k3) : slope(k1, k2); };
            val(k2);                            // y2
            dx(d, i + kd); cst(0.5); op(I_MUL);  // dx * 0.5
            if (lim <= OPF_LIM_LUI) {
                // KappaKernel::eval(slop_u, slop_f) = (1 + kappa) / 2. * slop_f + (1 - kappa) / 2. * slop_u
                const double kappa = lim == OPF_LIM_CDS ? 1.0 : lim == OPF_LIM_QUICK ? 0.5
                                     : lim == OPF_LIM_CUI ? 1. / 3. : lim == OPF_LIM_FROMM ? 0.0 : -1.0;
                cst((1 + kappa) / 2.); sf(); op(I_MUL);
                cst((1 - kappa) / 2.); su(); op(I_MUL);
                op(I_ADD);
                op(I_MUL);
            } else {
                // r = slop_f / (slop_u + 1e-16); (dx * 0.5 * K(r)) * slop_u
                auto r = [&] { sf(); su(); cst(1e-16); op(I_ADD); op(I_DIV); };
                limiter(lim, r);
                op(I_MUL);
                su();
                op(I_MUL);
            }
            op(up ? I_ADD : I_SUB);
        };
        emit(e.a, sh[0], sh[1], sh[2]);
        cst(0.0);
        op(I_GT);
        branch(true);
        branch(false);
        op(I_SEL);
    }
    // K(r) of the ratio-form kernels (FluxLimiterKernels.hpp); std::min / std::max and their
    // initializer-list forms fold left with the strict comparisons of I_MIN / I_MAX
    template <typename R>
    void limiter(int lim, R&& r) {
        auto mr = [&](double m) { cst(m); r(); op(I_MUL); };
        // 0.5 * (1 + kappa) * r + 0.5 * (1 - kappa)  and its mirror
        auto lin = [&](double kappa, bool mirror) {
            const double c1 = 0.5 * (1 + kappa), c2 = 0.5 * (1 - kappa);
            cst(mirror ? c2 : c1); r(); op(I_MUL); cst(mirror ? c1 : c2); op(I_ADD);
        };
        switch (lim) {
            case OPF_LIM_MINMOD:  // max(0., min(r, 1.))
                cst(0.0); r(); cst(1.0); op(I_MIN); op(I_MAX);
                break;
            case OPF_LIM_SUPERBEE:  // max({0., min(2. * r, 1.), min(r, 2.)})
                cst(0.0); mr(2.0); cst(1.0); op(I_MIN); op(I_MAX); r(); cst(2.0); op(I_MIN); op(I_MAX);
                break;
            case OPF_LIM_MUSCL:  // max(0., min({2 * r, (r + 1) / 2., 2.}))
                cst(0.0); mr(2.0); r(); cst(1.0); op(I_ADD); cst(2.0); op(I_DIV); op(I_MIN); cst(2.0); op(I_MIN);
                op(I_MAX);
                break;
            case OPF_LIM_HARMONIC:  // (r + fabs(r)) / (r + 1)
                r(); r(); op(I_ABS); op(I_ADD); r(); cst(1.0); op(I_ADD); op(I_DIV);
                break;
            case OPF_LIM_VANALBADA:  // r * (r + 1) / (r * r + 1)
                r(); r(); cst(1.0); op(I_ADD); op(I_MUL); r(); r(); op(I_MUL); cst(1.0); op(I_ADD); op(I_DIV);
                break;
            default: {
                // SPL / GPL kernels; m, alpha and kappa are std::ratio num / den as integer divisions
                // in the reference (FluxLimiterKernels.hpp:72-143), so the ratios < 1 count as 0
                double m = 2, alpha = 0, kappa = 0;
                if (lim == OPF_LIM_SMART) m = 4;
                if (lim == OPF_LIM_BSOU) m = 1, kappa = -1;
                cst(0.0);
                if (lim == OPF_LIM_UMIST || lim == OPF_LIM_SPL13) {
                    mr(m); lin(kappa, false); op(I_MIN); lin(kappa, true); op(I_MIN); cst(m); op(I_MIN);
                } else if (lim == OPF_LIM_SPLMAX12 || lim == OPF_LIM_SPLMAX13) {
                    mr(m); lin(kappa, false); lin(kappa, true); op(I_MAX); op(I_MIN); cst(m); op(I_MIN);
                } else {  // GPLKappaKernel: Koren, SMART, BSOU
                    mr(2 + alpha); lin(kappa, false); op(I_MIN); cst(m); op(I_MIN);
                }
                op(I_MAX);
            }
        }
    }

    void emit(int n, int s0, int s1, int s2) {
        if (err) return;
        const opf_expr_node& e = nodes[n];
        int sh[3] = {s0, s1, s2};
        auto at = [&](int c, int d, int k) {
            int t[3] = {s0, s1, s2};
            t[d] += k;
            emit(c, t[0], t[1], t[2]);
        };
        switch (e.kind) {
            case OPF_N_FIELD:
            case OPF_N_UNKNOWN: {
                int f = e.kind == OPF_N_FIELD ? e.a : unknown_slots[e.a];
                const opf_field_desc& fd = fields[f]->d;
                long long lin = sh[0] + (long long)fd.extent[0] * (sh[1] + (long long)fd.extent[1] * sh[2]);
                const bool ldu = e.kind == OPF_N_UNKNOWN && !unknown_as_field;
                // pad mode: a = the unknown's index; otherwise the field slot read
                code.push_back({ldu ? I_LDU : I_LD, ldu ? e.a : f, (int)lin, 0, 0.0});
                for (int d = 0; d < 3; ++d) {
                    omin[f][d] = std::min(omin[f][d], sh[d]);
                    omax[f][d] = std::max(omax[f][d], sh[d]);
                }
                if (ldu) {
                    // remember the offset in the instruction for pad mode slot lookup
                    code.back().c = (sh[0] + 8) | ((sh[1] + 8) << 5) | ((sh[2] + 8) << 10);
                } else {
                    // and for the compiled sweep's tile addressing (offsets within +-63)
                    if (sh[0] < -63 || sh[0] > 63 || sh[1] < -63 || sh[1] > 63 || sh[2] < -63 || sh[2] > 63) {
                        set_error("lowering: stencil offset beyond 63 cells");
                        err = OPF_ERR_UNSUPPORTED;
                        return;
                    }
                    code.back().c = (sh[0] + 64) | ((sh[1] + 64) << 8) | ((sh[2] + 64) << 16);
                }
                push(1);
            } break;
            case OPF_N_CONST: ucst(n, e.value); break;
            case OPF_N_MIN: emit(e.a, s0, s1, s2); emit(e.b, s0, s1, s2); op(I_MIN); break;
            case OPF_N_MAX: emit(e.a, s0, s1, s2); emit(e.b, s0, s1, s2); op(I_MAX); break;
            case OPF_N_GT: emit(e.a, s0, s1, s2); emit(e.b, s0, s1, s2); op(I_GT); break;
            case OPF_N_LT: emit(e.a, s0, s1, s2); emit(e.b, s0, s1, s2); op(I_LT); break;
            case OPF_N_GE: emit(e.a, s0, s1, s2); emit(e.b, s0, s1, s2); op(I_GE); break;
            case OPF_N_LE: emit(e.a, s0, s1, s2); emit(e.b, s0, s1, s2); op(I_LE); break;
            case OPF_N_ABS: emit(e.a, s0, s1, s2); op(I_ABS); break;
            case OPF_N_FUNC1:
                if (e.dim == OPF_FN_SMOOTH_DELTA) {
                    // Math::smoothDelta (Numeric.hpp:94-96):
                    //   std::max(d, -d) > eps ? 0. : 1. / 2. / eps * (1 + std::cos(d * PI / eps))
                    const double eps = e.value;
                    emit(e.a, s0, s1, s2); emit(e.a, s0, s1, s2); op(I_NEG); op(I_MAX); cst(eps); op(I_GT);
                    cst(0.0);
                    cst(1. / 2. / eps);
                    cst(1.0); emit(e.a, s0, s1, s2); cst(kPi); op(I_MUL); cst(eps); op(I_DIV); fn1(OPF_FN_COS);
                    op(I_ADD); op(I_MUL);
                    op(I_SEL);
                } else if (e.dim == OPF_FN_SMOOTH_HEAVISIDE) {
                    // Math::smoothHeviside (Numeric.hpp:86-92): d < -eps ? 0 : d > eps ? 1 :
                    //   0.5 * (1 + d / eps + 1 / PI * std::sin(d * PI / eps))
                    const double eps = e.value;
                    emit(e.a, s0, s1, s2); cst(-eps); op(I_LT);
                    cst(0.0);
                    emit(e.a, s0, s1, s2); cst(eps); op(I_GT);
                    cst(1.0);
                    cst(0.5);
                    cst(1.0); emit(e.a, s0, s1, s2); cst(eps); op(I_DIV); op(I_ADD);
                    cst(1 / kPi); emit(e.a, s0, s1, s2); cst(kPi); op(I_MUL); cst(eps); op(I_DIV); fn1(OPF_FN_SIN);
                    op(I_MUL); op(I_ADD); op(I_MUL);
                    op(I_SEL);
                    op(I_SEL);
                } else {
                    emit(e.a, s0, s1, s2);
                    fn1(e.dim);
                }
                break;
            case OPF_N_FUNC2:
                emit(e.a, s0, s1, s2);
                emit(e.b, s0, s1, s2);
                code.push_back({I_FN2, e.dim, 0, 0, 0.0});
                push(-1);
                break;
            case OPF_N_CONV: {
                // Convolution::eval (Convolution.hpp:52-68): rangeReduce_s over the window
                // x-fastest of kernel[idx - i + n/2] * e(idx): the first term, then += each
                const int k[3] = {e.dim & 255, (e.dim >> 8) & 255, (e.dim >> 16) & 255};
                int q = 0;
                for (int c = 0; c < k[2]; ++c)
                    for (int b = 0; b < k[1]; ++b)
                        for (int a = 0; a < k[0]; ++a, ++q) {
                            ucst(e.b + q, nodes[e.b + q].value);
                            emit(e.a, s0 + a - k[0] / 2, s1 + b - k[1] / 2, s2 + c - k[2] / 2);
                            op(I_MUL);
                            if (q > 0) op(I_ADD);
                        }
            } break;
            case OPF_N_COND:
                emit(e.a, s0, s1, s2);
                emit(e.b, s0, s1, s2);
                emit(e.dim, s0, s1, s2);
                op(I_SEL);
                break;
            case OPF_N_FLUXLIM: emit_fluxlim(e, sh); break;
            case OPF_N_ADD: emit(e.a, s0, s1, s2); emit(e.b, s0, s1, s2); op(I_ADD); break;
            case OPF_N_SUB: emit(e.a, s0, s1, s2); emit(e.b, s0, s1, s2); op(I_SUB); break;
            case OPF_N_MUL: emit(e.a, s0, s1, s2); emit(e.b, s0, s1, s2); op(I_MUL); break;
            case OPF_N_DIV: emit(e.a, s0, s1, s2); emit(e.b, s0, s1, s2); op(I_DIV); break;
            case OPF_N_NEG: emit(e.a, s0, s1, s2); op(I_NEG); break;
            case OPF_N_D2C: {
                // D2SecondOrderCentered::eval (D2SecondOrderCentered.hpp:160-171)
                int d = e.dim, i = sh[d];
                bool corner = locs[e.a][d] == OPF_LOC_CORNER;
                auto dxl = [&] {
                    if (corner) dx(d, i - 1);
                    else { dx(d, i - 1); dx(d, i); op(I_ADD); cst(0.5); op(I_MUL); }
                };
                auto dxr = [&] {
                    if (corner) dx(d, i);
                    else { dx(d, i); dx(d, i + 1); op(I_ADD); cst(0.5); op(I_MUL); }
                };
                at(e.a, d, +1); at(e.a, d, 0); op(I_SUB); dxr(); op(I_DIV);
                at(e.a, d, 0); at(e.a, d, -1); op(I_SUB); dxl(); op(I_DIV);
                op(I_SUB);
                dxl(); dxr(); op(I_ADD); cst(0.5); op(I_MUL);
                op(I_DIV);
            } break;
            case OPF_N_D1C: {
                // D1FirstOrderCentered::eval (D1FirstOrderCentered.hpp:30-37)
                int d = e.dim, i = sh[d];
                if (locs[e.a][d] == OPF_LOC_CENTER) {
                    at(e.a, d, 0); at(e.a, d, -1); op(I_SUB);
                    dx(d, i - 1); dx(d, i); op(I_ADD); op(I_DIV);
                    cst(2.0); op(I_MUL);
                } else {
                    at(e.a, d, +1); at(e.a, d, 0); op(I_SUB); dx(d, i); op(I_DIV);
                }
            } break;
            case OPF_N_D1UP: {
                // D1FirstOrderBiasedUpwind::eval (D1FirstOrderBiasedUpwind.hpp:53-58)
                int d = e.dim, i = sh[d];
                at(e.a, d, +1); at(e.a, d, 0); op(I_SUB);
                if (locs[e.a][d] == OPF_LOC_CORNER) dx(d, i);
                else { dx(d, i); dx(d, i + 1); op(I_ADD); cst(0.5); op(I_MUL); }
                op(I_DIV);
            } break;
            case OPF_N_D1DOWN: {
                // D1FirstOrderBiasedDownwind::eval (D1FirstOrderBiasedDownwind.hpp:52-57)
                int d = e.dim, i = sh[d];
                at(e.a, d, 0); at(e.a, d, -1); op(I_SUB);
                if (locs[e.a][d] == OPF_LOC_CORNER) dx(d, i - 1);
                else { dx(d, i - 1); dx(d, i); op(I_ADD); cst(0.5); op(I_MUL); }
                op(I_DIV);
            } break;
            case OPF_N_WENO_UP:
            case OPF_N_WENO_DOWN: {
                // D1WENO53Upwind/Downwind::eval: six values in index order, then h = dx(d, i)
                int d = e.dim, i = sh[d];
                int k0 = e.kind == OPF_N_WENO_UP ? -2 : -3;
                for (int k = 0; k < 6; ++k) at(e.a, d, k0 + k);
                dx(d, i);
                op(e.kind == OPF_N_WENO_UP ? I_WENO_UP : I_WENO_DOWN);
            } break;
            case OPF_N_INTP_C2N: {
                // D1Linear<d, Cen2Cor>::eval -> Interpolator1D::intp(x1, y1, x2, y2, x)
                // = ((x1-x)*y2 - (x2-x)*y1) / ((x1-x) - (x2-x))   (Interpolator.hpp:21-40)
                int d = e.dim, i = sh[d];
                auto x1mx = [&] { x(d, i - 1); cst(0.5); dx(d, i - 1); op(I_MUL); op(I_ADD); x(d, i); op(I_SUB); };
                auto x2mx = [&] { x(d, i); cst(0.5); dx(d, i); op(I_MUL); op(I_ADD); x(d, i); op(I_SUB); };
                x1mx(); at(e.a, d, 0); op(I_MUL);
                x2mx(); at(e.a, d, -1); op(I_MUL);
                op(I_SUB);
                x1mx(); x2mx(); op(I_SUB);
                op(I_DIV);
            } break;
            case OPF_N_INTP_N2C: {
                // D1Linear<d, Cor2Cen>::eval -> Math::mid(a, b) = (a + b) * 0.5 (Interpolator.hpp:83)
                int d = e.dim;
                at(e.a, d, 0); at(e.a, d, +1); op(I_ADD); cst(0.5); op(I_MUL);
            } break;
            default:
                set_error("lowering: unknown node kind %d", e.kind);
                err = OPF_ERR_ARG;
        }
    }
};

int lower_tree(const opf_expr_node* nodes, int nnodes, int root, opf_field* const* fields, int nfields,
               const opf_mesh* mesh, int unknown_slot, LoweredProgram* out, bool unknown_as_field) {
    return lower_tree_set(nodes, nnodes, root, fields, nfields, mesh, &unknown_slot, unknown_slot >= 0 ? 1 : 0, out,
                          unknown_as_field);
}

int lower_tree_set(const opf_expr_node* nodes, int nnodes, int root, opf_field* const* fields, int nfields,
                   const opf_mesh* mesh, const int* unknown_slots, int nunknown, LoweredProgram* out,
                   bool unknown_as_field) {
    OPF_CHECK(nodes && nnodes > 0 && root >= 0 && root < nnodes, OPF_ERR_ARG, "lowering: bad tree");
    OPF_CHECK(nfields >= 0 && nfields <= SWEEP_MAX_FIELDS, OPF_ERR_UNSUPPORTED, "at most %d fields per sweep",
              SWEEP_MAX_FIELDS);
    Lowerer L;
    L.nodes = nodes;
    L.nnodes = nnodes;
    L.fields = fields;
    L.nfields = nfields;
    L.mesh = mesh;
    L.unknown_slots = unknown_slots;
    L.nunknown = nunknown;
    L.unknown_as_field = unknown_as_field;
    L.init();
    int rc = L.compute_locs();
    OPF_CHECK(rc == OPF_OK, rc, "lowering: malformed tree (children must precede parents, valid field slots)");
    L.emit(root, 0, 0, 0);
    if (L.err) return L.err;
    OPF_CHECK(L.maxdepth <= SWEEP_MAX_DEPTH, OPF_ERR_UNSUPPORTED, "expression needs stack depth %d > %d",
              L.maxdepth, SWEEP_MAX_DEPTH);
    OPF_CHECK(L.code.size() < 100000, OPF_ERR_UNSUPPORTED, "expression too large after inlining");
    out->code = L.code;
    out->consts = L.consts;
    out->maxdepth = L.maxdepth;
    for (int f = 0; f < SWEEP_MAX_FIELDS; ++f)
        for (int d = 0; d < 3; ++d) {
            out->omin[f][d] = L.omin[f][d];
            out->omax[f][d] = L.omax[f][d];
        }
    for (int d = 0; d < 3; ++d) {
        out->tmin[d] = L.tmin[d];
        out->tmax[d] = L.tmax[d];
    }
    for (int d = 0; d < 3; ++d) out->loc[d] = L.locs[root][d];
    return OPF_OK;
}

// Host-side bounds check: every read of every field over the evaluation box
// [s, s+n) stays inside that field's storage, and mesh-table reads inside the
// mesh arrays.
int check_bounds(const LoweredProgram& P, opf_field* const* fields, int nfields, const opf_mesh* mesh, const int s[3],
                 const int n[3]) {
    for (int f = 0; f < nfields; ++f)
        for (int d = 0; d < 3; ++d) {
            if (P.omin[f][d] > P.omax[f][d]) continue;
            const opf_field_desc& fd = fields[f]->d;
            int lo = s[d] + P.omin[f][d], hi = s[d] + n[d] - 1 + P.omax[f][d];
            OPF_CHECK(lo >= fd.offset[d] && hi < fd.offset[d] + fd.extent[d], OPF_ERR_RANGE,
                      "sweep reads field slot %d at [%d,%d] in dim %d outside storage [%d,%d)", f, lo, hi, d,
                      fd.offset[d], fd.offset[d] + fd.extent[d]);
        }
    for (int d = 0; d < mesh->dim; ++d) {
        if (P.tmin[d] > P.tmax[d]) continue;
        int lo = s[d] + P.tmin[d], hi = s[d] + n[d] - 1 + P.tmax[d];
        OPF_CHECK(mesh->has_dx(d, lo) && mesh->has_dx(d, hi), OPF_ERR_RANGE,
                  "sweep reads mesh spacing of dim %d at [%d,%d] outside the mesh padding", d, lo, hi);
    }
    return OPF_OK;
}

// -------------------------------------------------------------- device side
#define OPF_SHARED_SRC(...) __VA_ARGS__
#include "weno_device.inc"
#include "math_device.inc"
#undef OPF_SHARED_SRC

// the interpreter calls the math functions out of line: inlined, every one of them
// would sit in the interpreter loop and raise its register count for every program
template <typename T>
__device__ __noinline__ T interp_fn1(int f, T x) { return opf_fn1<T>(f, x); }
template <typename T>
__device__ __noinline__ T interp_fn2(int f, T x, T y) { return opf_fn2<T>(f, x, y); }

template <typename T>
__global__ __launch_bounds__(256) void sweep_kernel(SweepArgs a) {
    __shared__ T stk[SWEEP_MAX_DEPTH][256];
    const int tid = threadIdx.x + 64 * threadIdx.y;
    const int i = a.s[0] + blockIdx.x * 64 + threadIdx.x;
    const int j = a.s[1] + blockIdx.y * 4 + threadIdx.y;
    const int k = a.s[2] + blockIdx.z;
    const bool live = i < a.s[0] + a.n[0] && j < a.s[1] + a.n[1];
    // dead lanes still walk the program (uniform control flow) but read cell s
    const int ci = live ? i : a.s[0], cj = live ? j : a.s[1];
    const Insn* __restrict__ code = (const Insn*)a.code;
    T tos = 0;
    int sp = 0;
    for (int pc = 0; pc < a.ncode; ++pc) {
        const Insn in = code[pc];
        switch (in.op) {
            case I_LD: {
                const FieldArg& F = a.f[in.a];
                long long base = (ci - F.off0) + (long long)F.e0 * ((cj - F.off1) + (long long)F.e1 * (k - F.off2));
                stk[sp++][tid] = tos;
                tos = ((const T*)F.ptr)[base + in.b];
            } break;
            case I_CONST: stk[sp++][tid] = tos; tos = in.a ? (T)a.cst[in.b] : (T)in.imm; break;
            case I_DX: {
                int c = in.a == 0 ? ci : in.a == 1 ? cj : k;
                stk[sp++][tid] = tos;
                tos = ((const T*)a.dx[in.a])[c + in.b - a.morg[in.a]];
            } break;
            case I_X: {
                int c = in.a == 0 ? ci : in.a == 1 ? cj : k;
                stk[sp++][tid] = tos;
                tos = ((const T*)a.x[in.a])[c + in.b - a.morg[in.a]];
            } break;
            case I_ADD: tos = stk[--sp][tid] + tos; break;
            case I_SUB: tos = stk[--sp][tid] - tos; break;
            case I_MUL: tos = stk[--sp][tid] * tos; break;
            case I_DIV: tos = stk[--sp][tid] / tos; break;
            case I_NEG: tos = -tos; break;
            case I_MIN: { const T a_ = stk[--sp][tid]; tos = tos < a_ ? tos : a_; } break;
            case I_MAX: { const T a_ = stk[--sp][tid]; tos = a_ < tos ? tos : a_; } break;
            case I_ABS: tos = fabs(tos); break;
            case I_GT: tos = stk[--sp][tid] > tos ? (T)1 : (T)0; break;
            case I_LT: tos = stk[--sp][tid] < tos ? (T)1 : (T)0; break;
            case I_GE: tos = stk[--sp][tid] >= tos ? (T)1 : (T)0; break;
            case I_LE: tos = stk[--sp][tid] <= tos ? (T)1 : (T)0; break;
            case I_SEL: {
                sp -= 2;
                tos = stk[sp][tid] != (T)0 ? stk[sp + 1][tid] : tos;
            } break;
            case I_FN1: tos = interp_fn1<T>(in.a, tos); break;
            case I_FN2: tos = interp_fn2<T>(in.a, stk[--sp][tid], tos); break;
            case I_WENO_UP:
            case I_WENO_DOWN: {
                sp -= 6;
                T v0 = stk[sp][tid], v1 = stk[sp + 1][tid], v2 = stk[sp + 2][tid], v3 = stk[sp + 3][tid],
                  v4 = stk[sp + 4][tid], v5 = stk[sp + 5][tid];
                // stack: [.. X v0 v1 v2 v3 v4 v5 | tos = h]; after sp -= 6, v0 = stk[sp]
                tos = weno_kernel<T>(v0, v1, v2, v3, v4, v5, tos, in.op == I_WENO_UP);
            } break;
            default: break;
        }
    }
    if (!live) return;
    long long lo = (i - a.oo[0]) + (long long)a.oe0 * ((j - a.oo[1]) + (long long)a.oe1 * (k - a.oo[2]));
    T v = tos;
    if (a.op != OPF_OP_EQ) {
        T old = ((const T*)a.old)[lo];
        switch (a.op) {
            case OPF_OP_ADD: v = old + v; break;
            case OPF_OP_MINUS: v = old - v; break;
            case OPF_OP_MUL: v = old * v; break;
            case OPF_OP_DIV: v = old / v; break;
        }
    }
    ((T*)a.out)[lo] = v;
}

}  // namespace opf

namespace opf {

std::vector<Insn> structural_code(const LoweredProgram& P) {
    std::vector<Insn> c = P.code;
    for (Insn& in : c)
        if (in.op == I_CONST && in.a) in.imm = 0.0;
    return c;
}

void uniform_spacing(const opf_mesh* m, int dtype, SweepArgs& a) {
    a.umask = 0;
    for (int d = 0; d < m->dim; ++d) {
        a.hdx[d] = a.rdx[d] = 0.0;
        int e;
        const double h = m->h[d];
        // a power of two whose reciprocal is a normal number of the sweep's precision
        if (!m->uniform[d] || !(h > 0) || std::frexp(h, &e) != 0.5) continue;
        const int lim = dtype == OPF_F64 ? 1016 : 120;  // room for multiples 1/4 .. 4 of h
        if (e - 1 > lim || e - 1 < -lim) continue;
        a.umask |= 1 << d;
        a.hdx[d] = h;
        a.rdx[d] = 1.0 / h;
    }
}

int launch_sweep(opf_ctx* ctx, int dtype, const SweepArgs& a0, const LoweredProgram& P) {
    hipStream_t st = ctx->stream;
    SweepArgs a = a0;
    OPF_CHECK(P.consts.size() <= (size_t)SWEEP_MAX_CONSTS, OPF_ERR_ARG, "sweep: %zu constant slots",
              P.consts.size());
    for (size_t q = 0; q < P.consts.size(); ++q) a.cst[q] = P.consts[q];
    const long long cells = (long long)a.n[0] * a.n[1] * a.n[2];
    const bool jit = (ctx->jit_mode >= 2 || (ctx->jit_mode == 1 && cells >= (1LL << 18))) &&
                     P.code.size() <= (size_t)SWEEP_JIT_MAX_INSNS;
    if (jit) {
        const int rc = jit_launch_sweep(st, P, dtype, a, ctx->jit_mode >= 3 ? ctx->jit_mode - 1 : 0);
        if (rc == OPF_OK) return OPF_OK;
        // the compiled engine is an optimisation: a failed compile or module load
        // runs the same program on the interpreter (same bits); argument errors stand
        if (rc != OPF_ERR_HIP) return rc;
        (void)hipGetLastError();
    }
    dim3 grid((a.n[0] + 63) / 64, (a.n[1] + 3) / 4, a.n[2]);
    if (dtype == OPF_F64) hipLaunchKernelGGL(sweep_kernel<double>, grid, dim3(64, 4), 0, st, a);
    else hipLaunchKernelGGL(sweep_kernel<float>, grid, dim3(64, 4), 0, st, a);
    OPF_HIP(hipGetLastError());
    return OPF_OK;
}

int mesh_arrays(const opf_mesh* mc, int dtype, const void* x[3], const void* dx[3]) {
    opf_mesh* m = const_cast<opf_mesh*>(mc);
    for (int d = 0; d < 3; ++d) {
        x[d] = dx[d] = nullptr;
        if (d >= m->dim) continue;
        if (dtype == OPF_F64) {
            x[d] = m->x_dev[d];
            dx[d] = m->dx_dev[d];
            continue;
        }
        if (!m->x_f32[d]) {
            std::vector<float> hx(m->hx[d].begin(), m->hx[d].end()), hdx(m->hdx[d].begin(), m->hdx[d].end());
            OPF_HIP(hipMalloc(&m->x_f32[d], hx.size() * sizeof(float)));
            OPF_HIP(hipMalloc(&m->dx_f32[d], hdx.size() * sizeof(float)));
            OPF_HIP(hipMemcpy(m->x_f32[d], hx.data(), hx.size() * sizeof(float), hipMemcpyHostToDevice));
            OPF_HIP(hipMemcpy(m->dx_f32[d], hdx.data(), hdx.size() * sizeof(float), hipMemcpyHostToDevice));
        }
        x[d] = m->x_f32[d];
        dx[d] = m->dx_f32[d];
    }
    return OPF_OK;
}

}  // namespace opf

extern "C" int opf_cache_stats(opf_ctx* ctx, long long* stats) {
    OPF_CHECK(ctx && stats, OPF_ERR_ARG, "opf_cache_stats: null arg");
    jit_stats(&stats[0], &stats[1], &stats[2]);
    stats[3] = ctx->prog_uploads;
    stats[4] = (long long)ctx->progs.size();
    return OPF_OK;
}

extern "C" int opf_sweep(const opf_expr_node* nodes, int nnodes, int root, opf_field* const* fields, int nfields,
                         opf_field* out, int op) {
    OPF_CHECK(out && op >= OPF_OP_EQ && op <= OPF_OP_DIV, OPF_ERR_ARG, "opf_sweep: bad out/op");
    OPF_CHECK(nfields == 0 || fields, OPF_ERR_ARG, "opf_sweep: null fields");
    const opf_mesh* mesh = out->mesh;
    bool inplace = false;
    for (int f = 0; f < nfields; ++f) {
        OPF_CHECK(fields[f], OPF_ERR_ARG, "opf_sweep: null field slot %d", f);
        OPF_CHECK(fields[f]->mesh == mesh, OPF_ERR_ARG, "opf_sweep: all fields must share the output's mesh");
        OPF_CHECK(fields[f]->d.dtype == out->d.dtype, OPF_ERR_UNSUPPORTED, "opf_sweep: mixed dtypes");
        inplace |= fields[f] == out;
    }
    LoweredProgram P;
    OPF_TRY(lower_tree(nodes, nnodes, root, fields, nfields, mesh, -1, &P));
    const opf_field_desc& d = out->d;
    int s[3], n[3];
    for (int q = 0; q < 3; ++q) {
        s[q] = std::max(d.assign_start[q], d.local_start[q]);
        n[q] = std::min(d.assign_end[q], d.local_end[q]) - s[q];
    }
    opf_ctx* ctx = out->ctx;
    hipStream_t st = ctx->stream;
    if (n[0] > 0 && n[1] > 0 && n[2] > 0) {
        OPF_TRY(check_bounds(P, fields, nfields, mesh, s, n));
        SweepArgs a{};
        for (int f = 0; f < nfields; ++f) {
            const opf_field_desc& fd = fields[f]->d;
            a.f[f] = {fields[f]->data, fd.offset[0], fd.offset[1], fd.offset[2], fd.extent[0], fd.extent[1]};
        }
        OPF_TRY(mesh_arrays(mesh, d.dtype, a.x, a.dx));
        uniform_spacing(mesh, d.dtype, a);
        for (int q = 0; q < 3; ++q) {
            a.morg[q] = mesh->start[q] - mesh->pad;
            a.s[q] = s[q];
            a.n[q] = n[q];
            a.oo[q] = d.offset[q];
        }
        a.oe0 = d.extent[0];
        a.oe1 = d.extent[1];
        a.op = op;
        void* target = out->data;
        if (inplace) {
            if (!out->spare) {
                OPF_HIP(hipMalloc(&out->spare, out->nbytes));
                out->spare_valid = false;
            }
            if (!out->spare_valid) {
                OPF_HIP(hipMemcpyAsync(out->spare, out->data, out->nbytes, hipMemcpyDeviceToDevice, st));
                out->spare_valid = true;
            }
            target = out->spare;
        }
        a.out = target;
        a.old = out->data;
        void* dev = nullptr;
        const std::vector<Insn> sc = structural_code(P);
        OPF_TRY(upload_program(ctx, sc.data(), sc.size() * sizeof(Insn), &dev, true));
        a.code = dev;
        a.ncode = (int)P.code.size();
        OPF_TRY(launch_sweep(ctx, d.dtype, a, P));
        if (inplace) std::swap(out->data, out->spare);
    }
    return launch_ghost_fill(out, out->data);
}
