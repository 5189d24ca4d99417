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
        push(o == I_NEG ? 0 : o == I_WENO_UP || o == I_WENO_DOWN ? -6 : -1);
    }
    void cst(double v) {
        code.push_back({I_CONST, 0, 0, 0, v});
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
                    if (!kid(e.a) || !kid(e.b)) return OPF_ERR_ARG;
                    locs[n] = locs[e.a][0] >= 0 ? locs[e.a] : locs[e.b];
                    break;
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
                }
                push(1);
            } break;
            case OPF_N_CONST: cst(e.value); break;
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
#undef OPF_SHARED_SRC

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
            case I_CONST: stk[sp++][tid] = tos; tos = (T)in.imm; break;
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

int launch_sweep(opf_ctx* ctx, int dtype, const SweepArgs& a, const LoweredProgram& P) {
    hipStream_t st = ctx->stream;
    const long long cells = (long long)a.n[0] * a.n[1] * a.n[2];
    if (ctx->jit_mode == 2 || (ctx->jit_mode == 1 && cells >= (1LL << 18))) return jit_launch_sweep(st, P, dtype, a);
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
        OPF_TRY(upload_program(ctx, P.code.data(), P.code.size() * sizeof(Insn), &dev));
        a.code = dev;
        a.ncode = (int)P.code.size();
        OPF_TRY(launch_sweep(ctx, d.dtype, a, P));
        if (inplace) std::swap(out->data, out->spare);
    }
    return launch_ghost_fill(out, out->data);
}
