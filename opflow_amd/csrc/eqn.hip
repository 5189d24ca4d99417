// Implicit equation handler: the device replacement of
// HYPREEqnSolveHandler::{init, generateAb, generateb, solve, returnValues}
// (Equation/HYPREEqnSolveHandler.hpp:66-209) for centre-located targets.
//
// The equation `lhs == rhs` is lowered to the sweep's stack program with
// OPF_N_UNKNOWN standing for the target's StencilField. A "pad mode" kernel
// runs that program on linear forms -- stack entries are StencilPads over the
// 1 + 2*dim axis slots of the cell plus a bias, exactly the algebra of
// StencilPad.hpp:136-298 -- so every cell's row and right-hand side are
// produced on the device. Ghost unknowns are kept as their own slots and the
// ghost rule's affine constant goes into the bias; the multigrid operator folds
// the ghost slot into the diagonal with the rule's factor, which is the
// folding StencilField::evalAtImpl_final performs (StencilField.hpp:123-350).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>

#include "gmg.hpp"
#include "sweep.hpp"

#define OPF_SHARED_SRC(...) __VA_ARGS__
#include "math_device.inc"
#undef OPF_SHARED_SRC
// out of line in the pad interpreter (see sweep.hip)
__device__ __noinline__ double pad_fn1(int f, double x) { return opf_fn1<double>(f, x); }
__device__ __noinline__ double pad_fn2(int f, double x, double y) { return opf_fn2<double>(f, x, y); }

namespace opf {

// PadArgs: the row box s, n (target interior), the target's local range [tl, te) (every
// unknown has this layout), per unknown the periodic dims and the ghost rule's affine
// constants (PadBC), `live` (the unknown whose coefficients are kept; -1: every unknown,
// in shared slots), the program, coef [PAD_K][N], bias [N], error flags, and the
// launch-time constants of the program
#define OPF_SHARED_SRC(...) __VA_ARGS__
#include "pad_args.inc"
#undef OPF_SHARED_SRC
static_assert(sizeof(((PadArgs*)nullptr)->f) / sizeof(FieldArg) == SWEEP_MAX_FIELDS, "field slots");
constexpr int PAD_W = PAD_K + 2;  // coefficients, bias, linear flag
constexpr int PAD_BLOCK = 32;     // 24-deep stack of 9 doubles per lane: 54 KB of LDS

__device__ __forceinline__ int slot_of(int ox, int oy, int oz) {
    if (ox == 0 && oy == 0 && oz == 0) return 0;
    if (oy == 0 && oz == 0) return ox == -1 ? 1 : ox == 1 ? 2 : -1;
    if (ox == 0 && oz == 0) return oy == -1 ? 3 : oy == 1 ? 4 : -1;
    if (ox == 0 && oy == 0) return oz == -1 ? 5 : oz == 1 ? 6 : -1;
    return -1;
}

__global__ __launch_bounds__(PAD_BLOCK) void pad_kernel(PadArgs a) {
    __shared__ double stk[SWEEP_MAX_DEPTH][PAD_W][PAD_BLOCK];
    const int tid = threadIdx.x;
    long long N = (long long)a.n[0] * a.n[1] * a.n[2];
    long long row = (long long)blockIdx.x * PAD_BLOCK + tid;
    const bool live = row < N;
    long long rr = live ? row : 0;
    const int ci = a.s[0] + (int)(rr % a.n[0]);
    const int cj = a.s[1] + (int)((rr / a.n[0]) % a.n[1]);
    const int ck = a.s[2] + (int)(rr / ((long long)a.n[0] * a.n[1]));
    const Insn* __restrict__ code = (const Insn*)a.code;
    double t[PAD_W];
    for (int q = 0; q < PAD_W; ++q) t[q] = 0;
    int sp = 0;
    int bad = 0;
    auto push = [&]() {
        for (int q = 0; q < PAD_W; ++q) stk[sp][q][tid] = t[q];
        ++sp;
        for (int q = 0; q < PAD_W; ++q) t[q] = 0;
    };
    for (int pc = 0; pc < a.ncode; ++pc) {
        const Insn in = code[pc];
        switch (in.op) {
            case I_LD: {
                const FieldArg& F = a.f[in.a];
                long long base = (ci - F.off0) + (long long)F.e0 * ((cj - F.off1) + (long long)F.e1 * (ck - F.off2));
                push();
                t[PAD_K] = ((const double*)F.ptr)[base + in.b];
            } break;
            case I_LDU: {
                // another unknown than the live one still counts as linear (so products of
                // unknowns are caught) and still brings its boundary constant to the bias
                const int u = in.a;
                int ox = (in.c & 31) - 8, oy = ((in.c >> 5) & 31) - 8, oz = ((in.c >> 10) & 31) - 8;
                push();
                int sl = slot_of(ox, oy, oz);
                if (sl < 0) { bad |= 1; break; }
                if (a.live < 0 || u == a.live) t[sl] = 1.0;
                t[PAD_K + 1] = 1.0;
                int p[3] = {ci + ox, cj + oy, ck + oz};
                for (int d = 0; d < 3; ++d) {
                    if (p[d] >= a.tl[d] && p[d] < a.te[d]) continue;
                    int side = p[d] < a.tl[d] ? 0 : 1;
                    if (!a.bc[u].per[d]) t[PAD_K] = a.bc[u].beta[d][side];
                }
            } break;
            case I_CONST: push(); t[PAD_K] = in.imm; break;
            case I_DX: {
                int c = in.a == 0 ? ci : in.a == 1 ? cj : ck;
                push();
                t[PAD_K] = a.dx[in.a][c + in.b - a.morg[in.a]];
            } break;
            case I_X: {
                int c = in.a == 0 ? ci : in.a == 1 ? cj : ck;
                push();
                t[PAD_K] = a.x[in.a][c + in.b - a.morg[in.a]];
            } break;
            case I_ADD:
            case I_SUB: {
                --sp;
                double sg = in.op == I_ADD ? 1.0 : -1.0;
                for (int q = 0; q <= PAD_K; ++q) {
                    double l = stk[sp][q][tid];
                    t[q] = in.op == I_ADD ? l + t[q] : l - t[q];
                }
                (void)sg;
                t[PAD_K + 1] = (stk[sp][PAD_K + 1][tid] != 0.0 || t[PAD_K + 1] != 0.0) ? 1.0 : 0.0;
            } break;
            case I_MUL: {
                --sp;
                bool ll = stk[sp][PAD_K + 1][tid] != 0.0, rl = t[PAD_K + 1] != 0.0;
                if (ll && rl) { bad |= 2; break; }
                if (rl) {  // const * pad
                    double c = stk[sp][PAD_K][tid];
                    for (int q = 0; q <= PAD_K; ++q) t[q] = c * t[q];
                } else {   // pad * const (or const * const)
                    double c = t[PAD_K];
                    for (int q = 0; q <= PAD_K; ++q) t[q] = stk[sp][q][tid] * c;
                    t[PAD_K + 1] = ll ? 1.0 : 0.0;
                }
            } break;
            case I_DIV: {
                --sp;
                if (t[PAD_K + 1] != 0.0) { bad |= 2; break; }
                double c = t[PAD_K];
                for (int q = 0; q <= PAD_K; ++q) t[q] = stk[sp][q][tid] / c;
                t[PAD_K + 1] = stk[sp][PAD_K + 1][tid];
            } break;
            case I_NEG:
                for (int q = 0; q <= PAD_K; ++q) t[q] = -t[q];
                break;
            case I_WENO_UP:
            case I_WENO_DOWN:  // WENO on linear forms is nonlinear
                bad |= 2;
                sp -= 6;
                break;
            // the nonlinear operations apply to values that do not read the unknown
            // (coefficients such as sqrt(r) or max(r, 0)); on a linear form they fail
            case I_SEL: {
                sp -= 2;
                if (stk[sp][PAD_K + 1][tid] != 0.0 || stk[sp + 1][PAD_K + 1][tid] != 0.0 || t[PAD_K + 1] != 0.0) {
                    bad |= 2;
                    break;
                }
                t[PAD_K] = stk[sp][PAD_K][tid] != 0.0 ? stk[sp + 1][PAD_K][tid] : t[PAD_K];
            } break;
            case I_ABS:
            case I_FN1:
                if (t[PAD_K + 1] != 0.0) { bad |= 2; break; }
                t[PAD_K] = in.op == I_ABS ? fabs(t[PAD_K]) : pad_fn1(in.a, t[PAD_K]);
                break;
            default: {  // I_MIN, I_MAX, I_FN2 and the comparisons: binary
                --sp;
                if (stk[sp][PAD_K + 1][tid] != 0.0 || t[PAD_K + 1] != 0.0) { bad |= 2; break; }
                const double x = stk[sp][PAD_K][tid], y = t[PAD_K];
                double v = 0;
                switch (in.op) {
                    case I_MIN: v = y < x ? y : x; break;
                    case I_MAX: v = x < y ? y : x; break;
                    case I_GT: v = x > y ? 1.0 : 0.0; break;
                    case I_LT: v = x < y ? 1.0 : 0.0; break;
                    case I_GE: v = x >= y ? 1.0 : 0.0; break;
                    case I_LE: v = x <= y ? 1.0 : 0.0; break;
                    case I_FN2: v = pad_fn2(in.a, x, y); break;
                    default: bad |= 2;
                }
                t[PAD_K] = v;
            } break;
        }
    }
    if (!live) return;
    if (bad) atomicOr(a.err, bad);
    for (int q = 0; q < PAD_K; ++q) a.coef[(long long)q * N + row] = t[q];
    a.bias[row] = t[PAD_K];
}

// the pads of program P over the rows of a: the compiled kernel (jit.hip) where the
// context compiles sweeps (every size in modes 2-4, >= 2^16 rows by default) and the
// program is a linear form it covers; else (and on a failed compile) the interpreter
static int launch_pads(opf_ctx* ctx, hipStream_t st, const LoweredProgram& P, PadArgs& a, long long N) {
    for (size_t q = 0; q < P.consts.size() && q < 32; ++q) a.cst[q] = P.consts[q];
    if (ctx->jit_mode >= 2 || (ctx->jit_mode == 1 && N >= (1LL << 16))) {
        const int rc = jit_launch_pad(st, P, a);
        if (rc == OPF_OK) return OPF_OK;
        if (rc != OPF_ERR_UNSUPPORTED && rc != OPF_ERR_HIP) return rc;
        (void)hipGetLastError();
    }
    hipLaunchKernelGGL(pad_kernel, dim3((unsigned)((N + PAD_BLOCK - 1) / PAD_BLOCK)), dim3(PAD_BLOCK), 0, st, a);
    OPF_HIP(hipGetLastError());
    return OPF_OK;
}

// b = -s * bias, gathered over rows
__global__ void k_rhs(long long N, const double* __restrict__ bias, double s, double* __restrict__ b) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) b[t] = -s * bias[t];
}

// field interior <-> dense vector
__global__ void k_gather(long long N, int n0, int n1, int s0, int s1, int s2, const double* __restrict__ f, int o0,
                         int o1, int o2, int e0, int e1, double* __restrict__ x) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    int i = s0 + (int)(t % n0), j = s1 + (int)((t / n0) % n1), k = s2 + (int)(t / ((long long)n0 * n1));
    x[t] = f[(i - o0) + (long long)e0 * ((j - o1) + (long long)e1 * (k - o2))];
}
__global__ void k_scatter(long long N, int n0, int n1, int s0, int s1, int s2, double* __restrict__ f, int o0, int o1,
                          int o2, int e0, int e1, const double* __restrict__ x, const double* __restrict__ shift) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    int i = s0 + (int)(t % n0), j = s1 + (int)((t / n0) % n1), k = s2 + (int)(t / ((long long)n0 * n1));
    f[(i - o0) + (long long)e0 * ((j - o1) + (long long)e1 * (k - o2))] = x[t] - (shift ? shift[0] : 0.0);
}
// b[pin] = -(sum of all other b) so the singular system is consistent
__global__ void k_pin_rhs(long long N, double* __restrict__ b, const double* __restrict__ part, int nb, long long pin) {
    __shared__ double sh[256];
    double acc = 0;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) acc += part[t];
    sh[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int q = 0; q < (int)blockDim.x; ++q) s += sh[q];
        b[pin] = -(s - b[pin]);
    }
}
__global__ void k_sum_rows(long long N, const double* __restrict__ b, double* __restrict__ part) {
    __shared__ double sh[256];
    double a = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < N; t += (long long)gridDim.x * blockDim.x)
        a += b[t];
    sh[threadIdx.x] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int q = 0; q < (int)blockDim.x; ++q) s += sh[q];
        part[blockIdx.x] = s;
    }
}
__global__ void k_copy1(const double* __restrict__ x, long long i, double* __restrict__ out) { out[0] = x[i]; }

// x_ghost = alpha * x_mirror + beta per face of a centre-located target (the
// folding StencilField::evalAtImpl_final applies, StencilField.hpp:123-350)
static int target_bc(const opf_field* t, int per[3], double alpha[3][2], double beta[3][2]) {
    GhostInfo gi;
    OPF_TRY(build_ghost_info(t, &gi));
    for (int d = 0; d < 3; ++d) {
        per[d] = 0;
        for (int s = 0; s < 2; ++s) {
            alpha[d][s] = 1.0;
            beta[d][s] = 0.0;
        }
    }
    for (int d = 0; d < t->d.dim; ++d) {
        per[d] = gi.per[d];
        for (int s = 0; s < 2; ++s) {
            const GhostRule& r = gi.rule[d][s][0];
            switch (r.type) {
                case 1: alpha[d][s] = r.c1 / r.c3; beta[d][s] = (0.0 * r.c1 - r.c2) / r.c3; break;
                case 2: alpha[d][s] = 1.0; beta[d][s] = r.c1; break;
                case 3: alpha[d][s] = 1.0; beta[d][s] = 0.0; break;
                case 4: alpha[d][s] = -1.0; beta[d][s] = 0.0; break;
                case 5: break;
                default: set_error("implicit solve: every boundary needs a BC"); return OPF_ERR_UNSUPPORTED;
            }
        }
    }
    return OPF_OK;
}

}  // namespace opf

using namespace opf;

struct opf_eqn {
    opf_ctx* ctx = nullptr;
    opf_field* target = nullptr;
    std::vector<opf_field*> fields;
    std::vector<opf_expr_node> all;  // lhs - rhs
    LoweredProgram P;
    void* prog = nullptr;
    // right-hand side only (static operators after the first solve): the tree with the
    // unknown read from a zero field whose ghosts hold the boundary constants
    LoweredProgram Pb;
    void* progb = nullptr;
    void* zbuf = nullptr;
    opf_solve_params params{};
    int s[3], n[3];
    long long N = 0;
    double* coef = nullptr;
    double* bias = nullptr;
    double* b = nullptr;
    double* x = nullptr;
    double* shift = nullptr;
    int* err = nullptr;
    double sign = 1.0;
    bool symmetric = true;  // false: BiCGStab with the symmetric part's multigrid as preconditioner
    bool built = false;
    bool first = true;
    DevLevel lvl;  // the operator's fine level (device arrays, reused by every rebuild)
    Hierarchy H;
    int per[3] = {0, 0, 0};
    double alpha[3][2] = {{1, 1}, {1, 1}, {1, 1}};
    double beta[3][2] = {{0, 0}, {0, 0}, {0, 0}};
};

static int eqn_eval_pads(opf_eqn* e) {
    opf_ctx* ctx = e->ctx;
    PadArgs a{};
    for (size_t f = 0; f < e->fields.size(); ++f) {
        const opf_field_desc& fd = e->fields[f]->d;
        a.f[f] = {e->fields[f]->data, fd.offset[0], fd.offset[1], fd.offset[2], fd.extent[0], fd.extent[1]};
    }
    const void* xs[3];
    const void* dxs[3];
    OPF_TRY(mesh_arrays(e->target->mesh, OPF_F64, xs, dxs));
    const opf_field_desc& td = e->target->d;
    for (int q = 0; q < 3; ++q) {
        a.x[q] = (const double*)xs[q];
        a.dx[q] = (const double*)dxs[q];
        a.morg[q] = e->target->mesh->start[q] - e->target->mesh->pad;
        a.s[q] = e->s[q];
        a.n[q] = e->n[q];
        a.tl[q] = td.local_start[q];
        a.te[q] = td.local_end[q];
        a.bc[0].per[q] = e->per[q];
        a.bc[0].beta[q][0] = e->beta[q][0];
        a.bc[0].beta[q][1] = e->beta[q][1];
    }
    a.live = -1;
    a.code = e->prog;
    a.ncode = (int)e->P.code.size();
    a.coef = e->coef;
    a.bias = e->bias;
    a.err = e->err;
    OPF_HIP(hipMemsetAsync(e->err, 0, sizeof(int), ctx->stream));
    OPF_TRY(launch_pads(ctx, ctx->stream, e->P, a, e->N));
    int herr = 0;
    OPF_HIP(hipMemcpyAsync(&herr, e->err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    OPF_HIP(hipStreamSynchronize(ctx->stream));
    OPF_CHECK(!(herr & 1), OPF_ERR_UNSUPPORTED,
              "equation reads the unknown off the 5/7-point axis stencil (not a Laplacian-type operator)");
    OPF_CHECK(!(herr & 2), OPF_ERR_UNSUPPORTED, "equation is nonlinear in the unknown");
    return OPF_OK;
}

// generateb (HYPREEqnSolveHandler.hpp:168-179): the bias of every row without the
// coefficients. The pad of lhs - rhs is linear in the unknown, so its constant
// part is the tree evaluated with the unknown replaced by a field that is zero
// inside and whose ghosts follow the target's boundary rules (the affine
// constants the pad evaluation folds in). One sweep, no host round trip.
static int eqn_eval_bias(opf_eqn* e) {
    opf_ctx* ctx = e->ctx;
    hipStream_t st = ctx->stream;
    opf_field* t = e->target;
    const int nf = (int)e->fields.size();
    if (!e->progb) {
        OPF_CHECK(nf + 1 <= SWEEP_MAX_FIELDS, OPF_ERR_UNSUPPORTED, "equation uses too many fields");
        OPF_HIP(hipMalloc(&e->zbuf, t->nbytes));
        OPF_HIP(hipMemsetAsync(e->zbuf, 0, t->nbytes, st));
        std::vector<opf_field*> fl(e->fields.begin(), e->fields.end());
        fl.push_back(t);  // slot nf: the zero field (the target's layout)
        OPF_TRY(lower_tree(e->all.data(), (int)e->all.size(), (int)e->all.size() - 1, fl.data(), nf + 1, t->mesh, nf,
                           &e->Pb, true));
        OPF_TRY(check_bounds(e->Pb, fl.data(), nf + 1, t->mesh, e->s, e->n));
        OPF_TRY(upload_program(ctx, e->Pb.code.data(), e->Pb.code.size() * sizeof(Insn), &e->progb));
    }
    OPF_TRY(launch_ghost_fill(t, e->zbuf));  // boundary constants (values may change between solves)
    SweepArgs a{};
    for (int f = 0; f < nf; ++f) {
        const opf_field_desc& fd = e->fields[f]->d;
        a.f[f] = {e->fields[f]->data, fd.offset[0], fd.offset[1], fd.offset[2], fd.extent[0], fd.extent[1]};
    }
    const opf_field_desc& td = t->d;
    a.f[nf] = {e->zbuf, td.offset[0], td.offset[1], td.offset[2], td.extent[0], td.extent[1]};
    OPF_TRY(mesh_arrays(t->mesh, OPF_F64, a.x, a.dx));
    uniform_spacing(t->mesh, OPF_F64, a);
    for (int q = 0; q < 3; ++q) {
        a.morg[q] = t->mesh->start[q] - t->mesh->pad;
        a.s[q] = e->s[q];
        a.n[q] = e->n[q];
        a.oo[q] = e->s[q];  // dense rows, x-fastest over the row box
    }
    a.oe0 = e->n[0];
    a.oe1 = e->n[1];
    a.op = OPF_OP_EQ;
    a.out = e->bias;
    a.old = e->bias;
    a.code = e->progb;
    a.ncode = (int)e->Pb.code.size();
    (void)st;
    return launch_sweep(ctx, OPF_F64, a, e->Pb);
}

// rows -> multigrid hierarchy of a single equation, on the device (the rows are
// never copied to the host); a rebuild over an operator of the same layout keeps
// the hierarchy's buffers and recomputes only the coefficients
static int eqn_build_operator(opf_eqn* e) {
    OPF_TRY(rows_to_level_dev(e->ctx->stream, e->coef, e->N, e->n, e->target->d.dim, e->per, &e->sign, e->lvl,
                              &e->symmetric));
    if (e->params.pre_smooth > 0) e->H.pre = e->params.pre_smooth;
    if (e->params.post_smooth > 0) e->H.post = e->params.post_smooth;
    const bool mixed = e->params.precision == OPF_PREC_MIXED && e->symmetric;
    e->built = true;  // gmg_free is safe on a partly built hierarchy
    return gmg_build_dev(e->ctx, e->H, e->lvl, e->per, e->alpha, mixed ? OPF_PREC_MIXED : OPF_PREC_FP64);
}

// non-symmetric operators: BiCGStab on the pinned system (defined with the set solvers)
static int eqn_solve_nonsym(opf_eqn* e, opf_solve_state* st, bool pin);

extern "C" int opf_eqn_create(const opf_expr_node* nodes, int nnodes, int lhs_root, int rhs_root,
                              opf_field* const* fields, int nfields, opf_field* target,
                              const opf_solve_params* params, opf_eqn** out) {
    OPF_CHECK(nodes && target && out && nfields >= 1 && fields, OPF_ERR_ARG, "opf_eqn_create: null arg");
    OPF_CHECK(target->d.dtype == OPF_F64, OPF_ERR_UNSUPPORTED, "implicit solve needs an FP64 target");
    int tslot = -1;
    for (int f = 0; f < nfields; ++f) {
        OPF_CHECK(fields[f] && fields[f]->mesh == target->mesh && fields[f]->d.dtype == OPF_F64, OPF_ERR_ARG,
                  "equation fields must share the target's mesh and be FP64");
        if (fields[f] == target) tslot = f;
    }
    OPF_CHECK(tslot >= 0, OPF_ERR_ARG, "the target must be one of the fields (slot of the StencilField)");
    const opf_field_desc& td = target->d;
    for (int d = 0; d < td.dim; ++d)
        OPF_CHECK(td.loc[d] == OPF_LOC_CENTER, OPF_ERR_UNSUPPORTED, "implicit solve supports centre-located targets");
    auto* e = new opf_eqn();
    e->ctx = target->ctx;
    e->target = target;
    e->fields.assign(fields, fields + nfields);
    e->params = params ? *params : opf_solve_params{};
    // lhs - rhs as one tree: append a SUB node
    e->all.assign(nodes, nodes + nnodes);
    e->all.push_back({OPF_N_SUB, lhs_root, rhs_root, 0, 0.0});
    int rc = lower_tree(e->all.data(), (int)e->all.size(), (int)e->all.size() - 1, fields, nfields, target->mesh, tslot,
                        &e->P);
    if (rc != OPF_OK) {
        delete e;
        return rc;
    }
    for (int q = 0; q < 3; ++q) {
        e->s[q] = std::max(td.assign_start[q], td.local_start[q]);
        e->n[q] = std::min(td.assign_end[q], td.local_end[q]) - e->s[q];
    }
    e->N = (long long)e->n[0] * e->n[1] * e->n[2];
    rc = check_bounds(e->P, fields, nfields, target->mesh, e->s, e->n);
    if (rc != OPF_OK) {
        delete e;
        return rc;
    }
    rc = target_bc(target, e->per, e->alpha, e->beta);
    if (rc != OPF_OK) {
        delete e;
        return rc;
    }
    opf_ctx* ctx = e->ctx;
    rc = upload_program(ctx, e->P.code.data(), e->P.code.size() * sizeof(Insn), &e->prog);
    if (rc == OPF_OK) {
        hipError_t h1 = hipMalloc(&e->coef, sizeof(double) * PAD_K * e->N);
        hipError_t h2 = hipMalloc(&e->bias, sizeof(double) * e->N);
        hipError_t h3 = hipMalloc(&e->b, sizeof(double) * e->N);
        hipError_t h4 = hipMalloc(&e->x, sizeof(double) * e->N);
        hipError_t h5 = hipMalloc(&e->shift, sizeof(double) * 4);
        hipError_t h6 = hipMalloc(&e->err, sizeof(int));
        if (h1 || h2 || h3 || h4 || h5 || h6) {
            set_error("opf_eqn_create: device allocation failed");
            rc = OPF_ERR_NOMEM;
        }
    }
    if (rc != OPF_OK) {
        opf_eqn_destroy(e);
        return rc;
    }
    *out = e;
    return OPF_OK;
}

extern "C" int opf_eqn_solve(opf_eqn* e, opf_solve_state* st) {
    OPF_CHECK(e && st, OPF_ERR_ARG, "opf_eqn_solve: null arg");
    opf_ctx* ctx = e->ctx;
    hipStream_t s = ctx->stream;
    // generateAb / generateb (HYPREEqnSolveHandler.hpp:125-179, solve :190-209)
    if (e->first || !e->params.static_mat) {
        OPF_TRY(eqn_eval_pads(e));
        OPF_TRY(eqn_build_operator(e));
    } else {
        OPF_TRY(eqn_eval_bias(e));
    }
    e->first = false;
    const long long N = e->N;
    const unsigned nb = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(k_rhs, dim3(nb), dim3(256), 0, s, N, e->bias, e->sign, e->b);
    // pinned singular systems: the first unknown is fixed at 0 (HYPREEqnSolveHandler.hpp:145-163); the
    // equivalent consistent singular system replaces that row's rhs by minus the sum of the others, and
    // the solution is shifted so that x[first] = 0.
    const bool pin = e->params.pin_value != 0;
    if (!e->symmetric) return eqn_solve_nonsym(e, st, pin);
    OPF_CHECK(!pin || e->H.singular, OPF_ERR_UNSUPPORTED,
              "pinValue on a non-singular operator is not supported by the multigrid solver");
    if (pin) {
        const int G = (int)std::min<long long>(256, (N + 255) / 256);
        hipLaunchKernelGGL(k_sum_rows, dim3(G), dim3(256), 0, s, N, (const double*)e->b, e->H.part);
        hipLaunchKernelGGL(k_pin_rhs, dim3(1), dim3(256), 0, s, N, e->b, (const double*)e->H.part, G, 0LL);
    }
    // initx: start from the target's current values (HYPREEqnSolveHandler.hpp:119-123)
    const opf_field_desc& td = e->target->d;
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, s, N, e->n[0], e->n[1], e->s[0], e->s[1], e->s[2],
                       (const double*)e->target->data, td.offset[0], td.offset[1], td.offset[2], td.extent[0],
                       td.extent[1], e->x);
    double tol = e->params.tol > 0 ? e->params.tol : 1e-6;
    int maxit = e->params.max_iter > 0 ? e->params.max_iter : 200;
    OPF_TRY(gmg_solve(ctx, e->H, e->b, e->x, tol, maxit, st, e->params.verbose));
    if (pin) hipLaunchKernelGGL(k_copy1, dim3(1), dim3(1), 0, s, (const double*)e->x, 0LL, e->shift);
    // returnValues: write back, updatePadding (HYPREEqnSolveHandler.hpp:181-188)
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, s, N, e->n[0], e->n[1], e->s[0], e->s[1], e->s[2],
                       (double*)e->target->data, td.offset[0], td.offset[1], td.offset[2], td.extent[0], td.extent[1],
                       (const double*)e->x, pin ? (const double*)e->shift : nullptr);
    OPF_HIP(hipGetLastError());
    e->target->spare_valid = false;
    OPF_TRY(launch_ghost_fill(e->target, e->target->data));
    return OPF_OK;
}

extern "C" int opf_eqn_csr(opf_eqn* e, long long* nrows, long long* nnz, long long* ptr, long long* col, double* val,
                           double* rhs) {
    // CSRMatrixGenerator::generate / generate_rhs (CSRMatrixGenerator.hpp:55-220) for one equation:
    // row r = the StencilPad of lhs - rhs at unknown r (x-fastest over the target's
    // assignable range), columns ascending, rhs = -bias; with pinValue the last row
    // becomes x_last = 0.
    OPF_CHECK(e && nrows && nnz, OPF_ERR_ARG, "opf_eqn_csr: null arg");
    OPF_TRY(eqn_eval_pads(e));
    const long long N = e->N;
    const int dim = e->target->d.dim;
    std::vector<double> c((size_t)PAD_K * N), bias((size_t)N);
    OPF_HIP(hipMemcpyAsync(c.data(), e->coef, sizeof(double) * c.size(), hipMemcpyDeviceToHost, e->ctx->stream));
    OPF_HIP(hipMemcpyAsync(bias.data(), e->bias, sizeof(double) * N, hipMemcpyDeviceToHost, e->ctx->stream));
    OPF_HIP(hipStreamSynchronize(e->ctx->stream));
    const bool pin = e->params.pin_value != 0;
    const long long st[3] = {1, e->n[0], (long long)e->n[0] * e->n[1]};
    long long cnt = 0;
    if (ptr) ptr[0] = 0;
    for (long long r = 0; r < N; ++r) {
        std::pair<long long, double> ent[7];
        int ne = 0;
        if (pin && r == N - 1) {
            ent[ne++] = {r, 1.0};
        } else {
            ent[ne++] = {r, c[r]};
            long long rem = r;
            int p[3] = {0, 0, 0};
            for (int d = 0; d < 3; ++d) {
                p[d] = (int)(rem % e->n[d]);
                rem /= e->n[d];
            }
            for (int d = 0; d < dim; ++d)
                for (int sd = 0; sd < 2; ++sd) {
                    const double a = c[(size_t)(1 + 2 * d + sd) * N + r];
                    const bool edge = sd == 0 ? p[d] == 0 : p[d] == e->n[d] - 1;
                    if (edge && !e->per[d]) {
                        // the slot holds the ghost, alpha * (this cell) + beta; beta is in the bias already
                        ent[0].second += e->alpha[d][sd] * a;
                        continue;
                    }
                    long long q = r + (sd == 0 ? -st[d] : st[d]);
                    if (edge) q = r + (sd == 0 ? 1 : -1) * (long long)(e->n[d] - 1) * st[d];
                    bool merged = false;  // a periodic extent of 1 or 2 meets the same cell twice
                    for (int m = 0; m < ne; ++m)
                        if (ent[m].first == q) {
                            ent[m].second += a;
                            merged = true;
                        }
                    if (!merged) ent[ne++] = {q, a};
                }
            std::sort(ent, ent + ne);
        }
        for (int m = 0; m < ne; ++m, ++cnt) {
            if (col) col[cnt] = ent[m].first;
            if (val) val[cnt] = ent[m].second;
        }
        if (ptr) ptr[r + 1] = cnt;
        if (rhs) rhs[r] = (pin && r == N - 1) ? 0.0 : -bias[r];
    }
    *nrows = N;
    *nnz = cnt;
    return OPF_OK;
}

extern "C" int opf_eqn_destroy(opf_eqn* e) {
    if (!e) return OPF_OK;
    hipStreamSynchronize(e->ctx->stream);
    if (e->built) gmg_free(e->H);
    dev_level_free(e->lvl);
    void* p[] = {e->coef, e->bias, e->b, e->x, e->shift, e->err, e->zbuf};
    for (void* q : p)
        if (q) hipFree(q);
    delete e;
    return OPF_OK;
}

// ------------------------------------------------------------------ equation sets
// SolveEqns (UnifiedSolve.hpp:69-100) with EqnHolder / StencilHolder and the
// multi-equation CSRMatrixGenerator::generate (CSRMatrixGenerator.hpp:26-41,
// 55-193): equation i is lhs_i - rhs_i over target i's rows; OPF_N_UNKNOWN's
// `a` names the target whose StencilField it stands for, so an equation may read
// several unknowns. Rows and columns follow ColoredMDRangeMapper
// (ColoredMDRangeMapper.hpp:20-41): block i holds target i's cells x-fastest,
// column = block offset of the unknown + its cell rank.
namespace opf {

struct EqsBC {
    int per[3];
    double alpha[3][2];
};

// y (+)= s * (block row of coefficients c applied to x): slot q of row r multiplies
// the unknown at the slot's neighbour; a neighbour outside the box is the ghost,
// alpha * (this cell) (its beta is in the bias), or the periodic image
__global__ void k_pad_apply(long long N, int n0, int n1, int n2, int dim, const double* __restrict__ c, EqsBC bc,
                            const double* __restrict__ x, double* __restrict__ y, double s, int acc) {
    long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const int n[3] = {n0, n1, n2};
    int p[3] = {(int)(r % n0), (int)((r / n0) % n1), (int)(r / ((long long)n0 * n1))};
    const long long st[3] = {1, n0, (long long)n0 * n1};
    const double xr = x[r];
    double v = c[r] * xr;
    for (int d = 0; d < dim; ++d)
        for (int sd = 0; sd < 2; ++sd) {
            const double a = c[(long long)(1 + 2 * d + sd) * N + r];
            const bool edge = sd == 0 ? p[d] == 0 : p[d] == n[d] - 1;
            double xn;
            if (!edge) xn = x[r + (sd == 0 ? -st[d] : st[d])];
            else if (bc.per[d]) xn = x[r + (sd == 0 ? 1 : -1) * (long long)(n[d] - 1) * st[d]];
            else xn = bc.alpha[d][sd] * xr;
            v += a * xn;
        }
    y[r] = (acc ? y[r] : 0.0) + s * v;
}
__global__ void k_pin_row(double* __restrict__ y, const double* __restrict__ x, long long i) { y[i] = x[i]; }
__global__ void k_sub_scalar(long long N, double* __restrict__ x, const double* __restrict__ s) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) x[t] -= s[0];
}
// partials of (a, b) and, when c is given, (c, d): one per block, fixed order
__global__ void k_dot2(long long M, const double* __restrict__ a, const double* __restrict__ b,
                       const double* __restrict__ c, const double* __restrict__ d, double* __restrict__ part) {
    __shared__ double s1[256], s2[256];
    double u = 0, w = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < M; t += (long long)gridDim.x * blockDim.x) {
        u += a[t] * b[t];
        if (c) w += c[t] * d[t];
    }
    s1[threadIdx.x] = u;
    s2[threadIdx.x] = w;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s1[0];
        part[gridDim.x + blockIdx.x] = s2[0];
    }
}
// p = r + beta (p - omega v)
__global__ void k_bicg_p(long long M, double* __restrict__ p, const double* __restrict__ r, const double* __restrict__ v,
                         double beta, double omega) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < M) p[t] = r[t] + beta * (p[t] - omega * v[t]);
}
// y = x + a z
__global__ void k_axpy(long long M, double* __restrict__ y, const double* __restrict__ x, double a,
                       const double* __restrict__ z) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < M) y[t] = x[t] + a * z[t];
}
// x += a p + w s
__global__ void k_bicg_x(long long M, double* __restrict__ x, double a, const double* __restrict__ p, double w,
                         const double* __restrict__ s) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < M) x[t] += a * p[t] + w * s[t];
}

struct EqsSystem {
    opf_ctx* ctx = nullptr;
    int neq = 0, dim = 2;
    std::vector<opf_field*> tg;
    int s[3], n[3];
    long long N = 0;
    int per[EQS_MAX][3];
    double alpha[EQS_MAX][3][2], beta[EQS_MAX][3][2];
    std::vector<std::vector<opf_field*>> fl;  // per equation: its fields, then the unknowns' layout
    std::vector<LoweredProgram> P;
    std::vector<void*> prog;
    std::vector<std::vector<int>> refs;  // unknowns each equation reads, ascending
    std::vector<int> blk;                // i * neq + j -> coefficient block, -1 if eq i does not read unknown j
    std::vector<unsigned> touched;       // per block: pad slots the equation's program reads
    int nblk = 0;
    double* coef = nullptr;  // [nblk][PAD_K][N]
    double* bias = nullptr;  // [neq][N]
    int* err = nullptr;
    ~EqsSystem() {
        if (ctx) hipStreamSynchronize(ctx->stream);
        if (coef) hipFree(coef);
        if (bias) hipFree(bias);
        if (err) hipFree(err);
    }
};

static int eqs_setup(int neq, const opf_eqn_desc* eqs, opf_field* const* targets, EqsSystem& S) {
    OPF_CHECK(neq >= 1 && neq <= EQS_MAX && eqs && targets, OPF_ERR_ARG, "equation set: 1..%d equations", EQS_MAX);
    const opf_field* t0 = targets[0];
    OPF_CHECK(t0, OPF_ERR_ARG, "equation set: null target");
    const opf_field_desc& d0 = t0->d;
    for (int i = 0; i < neq; ++i) {
        const opf_field* t = targets[i];
        OPF_CHECK(t && t->mesh == t0->mesh && t->d.dtype == OPF_F64, OPF_ERR_ARG,
                  "equation set: targets must be FP64 fields on one mesh");
        const opf_field_desc& d = t->d;
        bool same = d.dim == d0.dim;
        for (int q = 0; q < 3; ++q)
            same &= d.extent[q] == d0.extent[q] && d.offset[q] == d0.offset[q] &&
                    d.local_start[q] == d0.local_start[q] && d.local_end[q] == d0.local_end[q] &&
                    d.assign_start[q] == d0.assign_start[q] && d.assign_end[q] == d0.assign_end[q];
        OPF_CHECK(same, OPF_ERR_UNSUPPORTED, "equation set: targets must share one layout and range");
        for (int q = 0; q < d.dim; ++q)
            OPF_CHECK(d.loc[q] == OPF_LOC_CENTER, OPF_ERR_UNSUPPORTED, "implicit solve supports centre-located targets");
        OPF_TRY(target_bc(t, S.per[i], S.alpha[i], S.beta[i]));
    }
    S.ctx = t0->ctx;
    S.neq = neq;
    S.dim = d0.dim;
    S.tg.assign(targets, targets + neq);
    for (int q = 0; q < 3; ++q) {
        S.s[q] = std::max(d0.assign_start[q], d0.local_start[q]);
        S.n[q] = std::min(d0.assign_end[q], d0.local_end[q]) - S.s[q];
    }
    S.N = (long long)S.n[0] * S.n[1] * S.n[2];
    OPF_CHECK(S.N > 0, OPF_ERR_ARG, "equation set: empty range");
    S.fl.resize(neq);
    S.P.resize(neq);
    S.prog.assign(neq, nullptr);
    S.refs.resize(neq);
    S.blk.assign((size_t)neq * neq, -1);
    for (int i = 0; i < neq; ++i) {
        const opf_eqn_desc& E = eqs[i];
        OPF_CHECK(E.nodes && E.nnodes > 0 && E.nfields >= 0 && (E.fields || E.nfields == 0), OPF_ERR_ARG,
                  "equation set: equation %d is empty", i);
        OPF_CHECK(E.nfields + 1 <= SWEEP_MAX_FIELDS, OPF_ERR_UNSUPPORTED, "equation %d reads too many fields", i);
        for (int f = 0; f < E.nfields; ++f) {
            OPF_CHECK(E.fields[f] && E.fields[f]->mesh == t0->mesh && E.fields[f]->d.dtype == OPF_F64, OPF_ERR_ARG,
                      "equation fields must share the targets' mesh and be FP64");
            S.fl[i].push_back(E.fields[f]);
        }
        S.fl[i].push_back(targets[0]);  // the layout every unknown has
        std::vector<opf_expr_node> all(E.nodes, E.nodes + E.nnodes);
        all.push_back({OPF_N_SUB, E.lhs_root, E.rhs_root, 0, 0.0});
        std::vector<int> uslot(neq, E.nfields);
        OPF_TRY(lower_tree_set(all.data(), (int)all.size(), (int)all.size() - 1, S.fl[i].data(), (int)S.fl[i].size(),
                               t0->mesh, uslot.data(), neq, &S.P[i]));
        OPF_TRY(check_bounds(S.P[i], S.fl[i].data(), (int)S.fl[i].size(), t0->mesh, S.s, S.n));
        OPF_TRY(upload_program(S.ctx, S.P[i].code.data(), S.P[i].code.size() * sizeof(Insn), &S.prog[i]));
        std::vector<unsigned> mask(neq, 0u);
        for (const Insn& in : S.P[i].code) {
            if (in.op != I_LDU) continue;
            int ox = (in.c & 31) - 8, oy = ((in.c >> 5) & 31) - 8, oz = ((in.c >> 10) & 31) - 8;
            int sl = -1;
            if (ox == 0 && oy == 0 && oz == 0) sl = 0;
            else if (oy == 0 && oz == 0 && (ox == -1 || ox == 1)) sl = ox == -1 ? 1 : 2;
            else if (ox == 0 && oz == 0 && (oy == -1 || oy == 1)) sl = oy == -1 ? 3 : 4;
            else if (ox == 0 && oy == 0 && (oz == -1 || oz == 1)) sl = oz == -1 ? 5 : 6;
            OPF_CHECK(sl >= 0, OPF_ERR_UNSUPPORTED,
                      "equation %d reads an unknown off the 5/7-point axis stencil (not a Laplacian-type operator)", i);
            mask[in.a] |= 1u << sl;
        }
        for (int j = 0; j < neq; ++j)
            if (mask[j]) {
                S.refs[i].push_back(j);
                S.blk[(size_t)i * neq + j] = S.nblk++;
                S.touched.push_back(mask[j]);
            }
        OPF_CHECK(!S.refs[i].empty(), OPF_ERR_ARG, "equation %d reads no unknown", i);
    }
    OPF_HIP(hipMalloc(&S.coef, sizeof(double) * PAD_K * S.N * S.nblk));
    OPF_HIP(hipMalloc(&S.bias, sizeof(double) * S.N * neq));
    OPF_HIP(hipMalloc(&S.err, sizeof(int)));
    return OPF_OK;
}

// every block's coefficients and every row's bias, on the device
static int eqs_eval(EqsSystem& S) {
    hipStream_t st = S.ctx->stream;
    OPF_HIP(hipMemsetAsync(S.err, 0, sizeof(int), st));
    const opf_field_desc& td = S.tg[0]->d;
    const void* xs[3];
    const void* dxs[3];
    OPF_TRY(mesh_arrays(S.tg[0]->mesh, OPF_F64, xs, dxs));
    for (int i = 0; i < S.neq; ++i) {
        PadArgs a{};
        for (size_t f = 0; f < S.fl[i].size(); ++f) {
            const opf_field_desc& fd = S.fl[i][f]->d;
            a.f[f] = {S.fl[i][f]->data, fd.offset[0], fd.offset[1], fd.offset[2], fd.extent[0], fd.extent[1]};
        }
        for (int q = 0; q < 3; ++q) {
            a.x[q] = (const double*)xs[q];
            a.dx[q] = (const double*)dxs[q];
            a.morg[q] = S.tg[0]->mesh->start[q] - S.tg[0]->mesh->pad;
            a.s[q] = S.s[q];
            a.n[q] = S.n[q];
            a.tl[q] = td.local_start[q];
            a.te[q] = td.local_end[q];
        }
        for (int u = 0; u < S.neq; ++u)
            for (int q = 0; q < 3; ++q) {
                a.bc[u].per[q] = S.per[u][q];
                a.bc[u].beta[q][0] = S.beta[u][q][0];
                a.bc[u].beta[q][1] = S.beta[u][q][1];
            }
        a.code = S.prog[i];
        a.ncode = (int)S.P[i].code.size();
        a.bias = S.bias + (size_t)i * S.N;  // every pass writes the same bias
        a.err = S.err;
        for (int j : S.refs[i]) {
            a.live = j;
            a.coef = S.coef + (size_t)S.blk[(size_t)i * S.neq + j] * PAD_K * S.N;
            OPF_TRY(launch_pads(S.ctx, st, S.P[i], a, S.N));
        }
    }
    OPF_HIP(hipGetLastError());
    int herr = 0;
    OPF_HIP(hipMemcpyAsync(&herr, S.err, sizeof(int), hipMemcpyDeviceToHost, st));
    OPF_HIP(hipStreamSynchronize(st));
    OPF_CHECK(!(herr & 1), OPF_ERR_UNSUPPORTED,
              "equation reads an unknown off the 5/7-point axis stencil (not a Laplacian-type operator)");
    OPF_CHECK(!(herr & 2), OPF_ERR_UNSUPPORTED, "equation set is nonlinear in the unknowns");
    return OPF_OK;
}

// equations i and j read only their own unknown with the same program (up to which
// unknown), the same fields and the same boundary folding: their diagonal blocks
// hold the same coefficients (only the biases can differ)
static bool same_diagonal(const EqsSystem& S, int i, int j) {
    if (S.refs[i].size() != 1 || S.refs[i][0] != i || S.refs[j].size() != 1 || S.refs[j][0] != j) return false;
    if (S.fl[i] != S.fl[j] || S.P[i].code.size() != S.P[j].code.size()) return false;
    for (size_t q = 0; q < S.P[i].code.size(); ++q) {
        const Insn &x = S.P[i].code[q], &y = S.P[j].code[q];
        const bool same = x.op == y.op && (x.op == I_LDU || x.a == y.a) && x.b == y.b && x.c == y.c &&
                          std::memcmp(&x.imm, &y.imm, sizeof(double)) == 0;
        if (!same) return false;
    }
    for (int d = 0; d < 3; ++d)
        if (S.per[i][d] != S.per[j][d] || S.alpha[i][d][0] != S.alpha[j][d][0] || S.alpha[i][d][1] != S.alpha[j][d][1])
            return false;
    return true;
}

}  // namespace opf

extern "C" int opf_eqns_csr(int neq, const opf_eqn_desc* eqs, opf_field* const* targets, const int* pin,
                            long long* nrows, long long* nnz, long long* ptr, long long* col, double* val,
                            double* rhs) {
    // CSRMatrixGenerator::generate_s per equation, appended (CSRMatrixGenerator.hpp:26-41, 143-193):
    // block i's rows are target i's cells; a row lists the (unknown, cell) entries its
    // pad holds in ascending global column; with pin[i] the block's LAST row is x = 0
    OPF_CHECK(nrows && nnz, OPF_ERR_ARG, "opf_eqns_csr: null arg");
    EqsSystem S;
    OPF_TRY(eqs_setup(neq, eqs, targets, S));
    OPF_TRY(eqs_eval(S));
    const long long N = S.N;
    std::vector<double> c((size_t)PAD_K * N * S.nblk), bias((size_t)N * neq);
    OPF_HIP(hipMemcpyAsync(c.data(), S.coef, sizeof(double) * c.size(), hipMemcpyDeviceToHost, S.ctx->stream));
    OPF_HIP(hipMemcpyAsync(bias.data(), S.bias, sizeof(double) * bias.size(), hipMemcpyDeviceToHost, S.ctx->stream));
    OPF_HIP(hipStreamSynchronize(S.ctx->stream));
    const long long st[3] = {1, S.n[0], (long long)S.n[0] * S.n[1]};
    long long cnt = 0;
    if (ptr) ptr[0] = 0;
    std::vector<std::pair<long long, double>> ent;
    for (int i = 0; i < neq; ++i)
        for (long long r = 0; r < N; ++r) {
            const long long row = (long long)i * N + r;
            ent.clear();
            const bool pinned = pin && pin[i] && r == N - 1;
            if (pinned) {
                ent.push_back({row, 1.0});
            } else {
                int p[3];
                long long rem = r;
                for (int d = 0; d < 3; ++d) {
                    p[d] = (int)(rem % S.n[d]);
                    rem /= S.n[d];
                }
                auto add = [&](long long q, double v) {
                    for (auto& e : ent)
                        if (e.first == q) {
                            e.second += v;
                            return;
                        }
                    ent.push_back({q, v});
                };
                for (int j : S.refs[i]) {
                    const int b = S.blk[(size_t)i * neq + j];
                    const double* cb = c.data() + (size_t)b * PAD_K * N;
                    const unsigned tm = S.touched[b];
                    const long long base = (long long)j * N;
                    if (tm & 1u) add(base + r, cb[r]);
                    for (int d = 0; d < S.dim; ++d)
                        for (int sd = 0; sd < 2; ++sd) {
                            const int q = 1 + 2 * d + sd;
                            if (!(tm & (1u << q))) continue;
                            const double a = cb[(size_t)q * N + r];
                            const bool edge = sd == 0 ? p[d] == 0 : p[d] == S.n[d] - 1;
                            if (edge && !S.per[j][d]) {
                                add(base + r, S.alpha[j][d][sd] * a);  // the ghost, folded onto the mirror
                                continue;
                            }
                            long long nb = r + (sd == 0 ? -st[d] : st[d]);
                            if (edge) nb = r + (sd == 0 ? 1 : -1) * (long long)(S.n[d] - 1) * st[d];
                            add(base + nb, a);
                        }
                }
                std::sort(ent.begin(), ent.end());
            }
            for (const auto& e : ent) {
                if (col) col[cnt] = e.first;
                if (val) val[cnt] = e.second;
                ++cnt;
            }
            if (ptr) ptr[row + 1] = cnt;
            if (rhs) rhs[row] = pinned ? 0.0 : -bias[(size_t)i * N + r];
        }
    *nrows = N * neq;
    *nnz = cnt;
    return OPF_OK;
}

namespace opf {

using ApplyFn = std::function<void(hipStream_t, const double*, double*)>;
using PrecFn = std::function<int(const double*, double*)>;

static double host_sum(const std::vector<double>& v, size_t from, size_t n) {
    double s = 0;
    for (size_t q = 0; q < n; ++q) s += v[from + q];
    return s;
}

// (a, b) and (c, d) over M entries; deterministic (fixed grid, fixed order)
static int dots(hipStream_t st, long long M, const double* a, const double* b, const double* c, const double* d,
                double* part, std::vector<double>& hp, double* o1, double* o2) {
    const int G = (int)std::min<long long>(512, (M + 255) / 256);
    hipLaunchKernelGGL(k_dot2, dim3(G), dim3(256), 0, st, M, a, b, c, d, part);
    OPF_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double) * 2 * G, hipMemcpyDeviceToHost, st));
    OPF_HIP(hipStreamSynchronize(st));
    *o1 = host_sum(hp, 0, G);
    if (o2) *o2 = host_sum(hp, G, G);
    return OPF_OK;
}

// Coupled sets: BiCGStab (the solver of SolveEqns' AMGCL backend, AMGCLBackend.hpp:36-74)
// on the whole system, right-preconditioned by one multigrid V-cycle per diagonal
// block. Row block i is scaled by sign[i] (the orientation of its diagonal block,
// which its hierarchy approximates); pinned rows are x = 0.
struct EqsCoupled {
    EqsSystem& S;
    std::vector<Hierarchy>& H;
    const double* sign;
    const std::vector<int>& pin;
    void apply(hipStream_t st, const double* x, double* y) const {
        const long long N = S.N;
        const unsigned nb = (unsigned)((N + 255) / 256);
        for (int i = 0; i < S.neq; ++i) {
            bool first = true;
            for (int j : S.refs[i]) {
                EqsBC bc;
                for (int d = 0; d < 3; ++d) {
                    bc.per[d] = S.per[j][d];
                    bc.alpha[d][0] = S.alpha[j][d][0];
                    bc.alpha[d][1] = S.alpha[j][d][1];
                }
                const double* cb = S.coef + (size_t)S.blk[(size_t)i * S.neq + j] * PAD_K * N;
                hipLaunchKernelGGL(k_pad_apply, dim3(nb), dim3(256), 0, st, N, S.n[0], S.n[1], S.n[2], S.dim, cb, bc,
                                   x + (size_t)j * N, y + (size_t)i * N, sign[i], first ? 0 : 1);
                first = false;
            }
            if (pin[i]) hipLaunchKernelGGL(k_pin_row, dim3(1), dim3(1), 0, st, y + (size_t)i * N, x + (size_t)i * N, N - 1);
        }
    }
    int precond(opf_ctx* ctx, const double* r, double* z) const {
        for (int i = 0; i < S.neq; ++i) OPF_TRY(gmg_vcycle64(ctx, H[i], r + (size_t)i * S.N, z + (size_t)i * S.N));
        return OPF_OK;
    }
};

// Right-preconditioned BiCGStab on M unknowns from the given x: y = A x by `apply`,
// z = M^-1 r by `precond`; fixed-order dot products; stops at ||b - A x|| <= tol ||b||
// (recurrence residual; the final true residual is reported).
static int bicgstab(opf_ctx* ctx, long long M, const ApplyFn& apply, const PrecFn& precond, double* x, const double* b,
                    double tol, int maxit, opf_solve_state* out, int verbose) {
    hipStream_t st = ctx->stream;
    const unsigned nb = (unsigned)((M + 255) / 256);
    double* w = nullptr;
    OPF_HIP(hipMalloc(&w, sizeof(double) * (8 * M + 1024)));
    double *r = w, *rh = w + M, *p = w + 2 * M, *v = w + 3 * M, *sv = w + 4 * M, *t = w + 5 * M, *ph = w + 6 * M,
           *sh = w + 7 * M, *part = w + 8 * M;
    std::vector<double> hp(1024);
    int rc = OPF_OK;
    double bb = 0, rr = 0, rho = 1, alpha = 1, omega = 1;
    int it = 0;
    auto fail = [&](int code) {
        rc = code;
        return code;
    };
    // the recurrence residual can drift below the true one: a restart from the true
    // residual continues until the true residual meets the tolerance (at most 4 times)
    for (int restart = 0; restart < 5 && it < maxit; ++restart) {
        rho = alpha = omega = 1.0;
        // r = b - A x; r_hat = r; p = v = 0
        apply(st, x, t);
        hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(256), 0, st, M, r, b, -1.0, (const double*)t);
        if (fail(hipMemcpyAsync(rh, r, sizeof(double) * M, hipMemcpyDeviceToDevice, st) ? OPF_ERR_HIP : OPF_OK)) break;
        if (fail(hipMemsetAsync(p, 0, sizeof(double) * 2 * M, st) ? OPF_ERR_HIP : OPF_OK)) break;  // p, v
        if (fail(dots(st, M, b, b, r, r, part, hp, &bb, &rr))) break;
        const double tgt = tol * tol * bb;
        while (rr > tgt && it < maxit) {
            double rho1 = 0;
            if (fail(dots(st, M, rh, r, nullptr, nullptr, part, hp, &rho1, nullptr))) break;
            if (rho1 == 0.0) break;  // breakdown
            const double beta = (rho1 / rho) * (alpha / omega);
            rho = rho1;
            hipLaunchKernelGGL(k_bicg_p, dim3(nb), dim3(256), 0, st, M, p, (const double*)r, (const double*)v, beta,
                               omega);
            if (fail(precond(p, ph))) break;
            apply(st, ph, v);
            double rv = 0;
            if (fail(dots(st, M, rh, v, nullptr, nullptr, part, hp, &rv, nullptr))) break;
            if (rv == 0.0) break;
            alpha = rho / rv;
            hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(256), 0, st, M, sv, (const double*)r, -alpha, (const double*)v);
            if (fail(precond(sv, sh))) break;
            apply(st, sh, t);
            double ts = 0, tt = 0;
            if (fail(dots(st, M, t, sv, t, t, part, hp, &ts, &tt))) break;
            omega = tt > 0 ? ts / tt : 0.0;
            hipLaunchKernelGGL(k_bicg_x, dim3(nb), dim3(256), 0, st, M, x, alpha, (const double*)ph, omega,
                               (const double*)sh);
            hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(256), 0, st, M, r, (const double*)sv, -omega, (const double*)t);
            if (fail(dots(st, M, r, r, nullptr, nullptr, part, hp, &rr, nullptr))) break;
            ++it;
            if (verbose) fprintf(stderr, "[bicgstab] it %d relres %.3e\n", it, bb > 0 ? std::sqrt(rr / bb) : 0.0);
            if (omega == 0.0) break;
        }
        if (rc) break;
        // the true residual of the final x
        apply(st, x, t);
        hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(256), 0, st, M, r, b, -1.0, (const double*)t);
        if (fail(dots(st, M, r, r, nullptr, nullptr, part, hp, &rr, nullptr))) break;
        if (rr <= tgt * (1 + 1e-12) * (1 + 1e-12) || bb == 0) break;
    }
    hipStreamSynchronize(st);
    hipFree(w);
    if (rc) return rc;
    out->niter = it;
    out->nrefine = 0;
    out->abserr = std::sqrt(rr);
    out->relerr = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    out->converged = out->relerr <= tol * (1 + 1e-12) || bb == 0;
    return OPF_OK;
}

}  // namespace opf

extern "C" int opf_eqns_solve(int neq, const opf_eqn_desc* eqs, opf_field* const* targets,
                              const opf_solve_params* params, opf_solve_state* state) {
    // SolveEqns (UnifiedSolve.hpp:69-100): assemble, solve from x = 0 with params[0]'s
    // tolerance (relative residual of the whole system), write every target back and
    // updatePadding
    OPF_CHECK(state && params, OPF_ERR_ARG, "opf_eqns_solve: null arg");
    EqsSystem S;
    OPF_TRY(eqs_setup(neq, eqs, targets, S));
    OPF_TRY(eqs_eval(S));
    opf_ctx* ctx = S.ctx;
    hipStream_t st = ctx->stream;
    const long long N = S.N;
    const unsigned nb = (unsigned)((N + 255) / 256);
    const double tol = params[0].tol > 0 ? params[0].tol : 1e-8;  // amgcl solver defaults
    const int maxit = params[0].max_iter > 0 ? params[0].max_iter : 100;
    bool coupled = false;
    for (int i = 0; i < neq; ++i) coupled |= !(S.refs[i].size() == 1 && S.refs[i][0] == i);
    std::vector<int> pin(neq);
    for (int i = 0; i < neq; ++i) pin[i] = params[i].pin_value != 0;
    // diagonal blocks -> multigrid levels, built on the device; a twin (an earlier
    // equation whose diagonal block has the same program) shares its level's arrays
    struct Levels {
        std::vector<DevLevel> L;
        std::vector<char> own;
        ~Levels() {
            for (size_t i = 0; i < L.size(); ++i)
                if (own[i]) dev_level_free(L[i]);
        }
    } lv{std::vector<DevLevel>(neq), std::vector<char>(neq, 0)};
    std::vector<DevLevel>& L = lv.L;
    std::vector<double> sign(neq, 1.0);
    std::vector<char> symm(neq, 1);
    for (int i = 0; i < neq; ++i) {
        const int b = S.blk[(size_t)i * neq + i];
        OPF_CHECK(b >= 0, OPF_ERR_UNSUPPORTED, "equation %d does not read its own target (no diagonal block)", i);
        int twin = -1;
        for (int j = 0; j < i && twin < 0; ++j)
            if (same_diagonal(S, i, j)) twin = j;
        if (twin >= 0) {
            L[i] = L[twin];
            sign[i] = sign[twin];
            symm[i] = symm[twin];
            continue;
        }
        lv.own[i] = 1;
        bool sy = true;
        OPF_TRY(rows_to_level_dev(st, S.coef + (size_t)b * PAD_K * N, N, S.n, S.dim, S.per[i], &sign[i], L[i], &sy));
        symm[i] = sy;
    }
    // a non-symmetric diagonal block is solved by the Krylov method of coupled sets,
    // its multigrid level (the symmetric part) as preconditioner
    for (int i = 0; i < neq; ++i) coupled |= !symm[i];
    auto hier = [&](Hierarchy& H, int i, int slots, int precision) {
        if (params[0].pre_smooth > 0) H.pre = params[0].pre_smooth;
        if (params[0].post_smooth > 0) H.post = params[0].post_smooth;
        return gmg_build_dev(ctx, H, L[i], S.per[i], S.alpha[i], precision, slots);
    };
    double* vec = nullptr;  // x, b over the whole system (block i at i * N)
    OPF_HIP(hipMalloc(&vec, sizeof(double) * 2 * N * neq + 64));
    double* x = vec;
    double* b = vec + N * neq;
    double* shift = vec + 2 * N * neq;  // [neq]
    int rc = OPF_OK;
    std::vector<Hierarchy> H;
    do {
        if ((rc = hipMemsetAsync(x, 0, sizeof(double) * N * neq, st) ? OPF_ERR_HIP : OPF_OK)) break;
        for (int i = 0; i < neq; ++i) {
            hipLaunchKernelGGL(k_rhs, dim3(nb), dim3(256), 0, st, N, (const double*)(S.bias + (size_t)i * N), sign[i],
                               b + (size_t)i * N);
            if (pin[i] && (rc = hipMemsetAsync(b + (size_t)i * N + (N - 1), 0, sizeof(double), st) ? OPF_ERR_HIP : 0))
                break;  // the pinned row reads x = 0
        }
        if (rc) break;
        opf_solve_state total{};
        if (!coupled) {
            // block-diagonal: one multigrid solve per block, blocks with the same operator
            // batched as the slots of one hierarchy; a pinned block solves the consistent
            // singular system (its last row's rhs = minus the sum of the others, which the
            // pinned system implies) and is shifted so that its last unknown is 0
            std::vector<int> grp(neq, -1);
            for (int i = 0; i < neq && !rc; ++i) {
                if (grp[i] >= 0) continue;
                std::vector<int> mem{i};
                for (int j = i + 1; j < neq && !rc; ++j) {
                    bool same = false;
                    if (grp[j] < 0) rc = dev_levels_equal(st, L[i], L[j], S.dim, &same);
                    for (int d = 0; d < 3 && same; ++d)
                        same &= S.per[i][d] == S.per[j][d] && S.alpha[i][d][0] == S.alpha[j][d][0] &&
                                S.alpha[i][d][1] == S.alpha[j][d][1];
                    if (same) mem.push_back(j);
                }
                if (rc) break;
                for (int m : mem) grp[m] = i;
                H.emplace_back();
                Hierarchy& G = H.back();
                const int B = (int)mem.size();
                if ((rc = hier(G, i, B, params[0].precision == OPF_PREC_MIXED ? OPF_PREC_MIXED : OPF_PREC_FP64)))
                    break;
                double* xg = nullptr;
                if ((rc = hipMalloc(&xg, sizeof(double) * 2 * N * B) ? OPF_ERR_NOMEM : OPF_OK)) break;
                double* bg = xg + N * B;
                for (int m = 0; m < B; ++m) {
                    const int e = mem[m];
                    hipMemcpyAsync(bg + (size_t)m * N, b + (size_t)e * N, sizeof(double) * N, hipMemcpyDeviceToDevice,
                                   st);
                    if (pin[e]) {
                        if (!G.singular) {
                            set_error("pinValue on a non-singular operator is not supported by the multigrid solver");
                            rc = OPF_ERR_UNSUPPORTED;
                            break;
                        }
                        const int P = (int)std::min<long long>(256, (N + 255) / 256);
                        hipLaunchKernelGGL(k_sum_rows, dim3(P), dim3(256), 0, st, N, (const double*)(bg + (size_t)m * N),
                                           G.part);
                        hipLaunchKernelGGL(k_pin_rhs, dim3(1), dim3(256), 0, st, N, bg + (size_t)m * N,
                                           (const double*)G.part, P, N - 1);
                    }
                }
                if (!rc) rc = hipMemsetAsync(xg, 0, sizeof(double) * N * B, st) ? OPF_ERR_HIP : OPF_OK;
                std::vector<opf_solve_state> ss(B);
                if (!rc) rc = gmg_solve_slots(ctx, G, bg, xg, tol, maxit, ss.data(), params[0].verbose);
                if (!rc) {
                    for (int m = 0; m < B; ++m) {
                        const int e = mem[m];
                        double* xe = x + (size_t)e * N;
                        hipMemcpyAsync(xe, xg + (size_t)m * N, sizeof(double) * N, hipMemcpyDeviceToDevice, st);
                        if (pin[e]) {
                            hipLaunchKernelGGL(k_copy1, dim3(1), dim3(1), 0, st, (const double*)xe, N - 1, shift + e);
                            hipLaunchKernelGGL(k_sub_scalar, dim3(nb), dim3(256), 0, st, N, xe,
                                               (const double*)(shift + e));
                        }
                        total.niter = std::max(total.niter, ss[m].niter);
                        total.nrefine = std::max(total.nrefine, ss[m].nrefine);
                    }
                }
                hipStreamSynchronize(st);
                hipFree(xg);
            }
            if (rc) break;
        } else {
            for (int i = 0; i < neq && !rc; ++i) {
                H.emplace_back();
                rc = hier(H.back(), i, 1, OPF_PREC_FP64);
            }
            if (rc) break;
            EqsCoupled A{S, H, sign.data(), pin};
            rc = bicgstab(
                    ctx, N * neq, [&](hipStream_t q, const double* u, double* y) { A.apply(q, u, y); },
                    [&](const double* r, double* z) { return A.precond(ctx, r, z); }, x, b, tol, maxit, &total,
                    params[0].verbose);
            if (rc) break;
        }
        // the state reports the whole system's true residual (pinned rows included)
        {
            EqsCoupled A{S, H, sign.data(), pin};
            const long long M = N * neq;
            double* rv = nullptr;
            if ((rc = hipMalloc(&rv, sizeof(double) * (2 * M + 1024)) ? OPF_ERR_NOMEM : OPF_OK)) break;
            A.apply(st, x, rv + M);
            hipLaunchKernelGGL(k_axpy, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st, M, rv, (const double*)b,
                               -1.0, (const double*)(rv + M));
            std::vector<double> hp(1024);
            double rr = 0, bb = 0;
            rc = dots(st, M, rv, rv, b, b, rv + 2 * M, hp, &rr, &bb);
            hipFree(rv);
            if (rc) break;
            total.abserr = std::sqrt(rr);
            total.relerr = bb > 0 ? std::sqrt(rr / bb) : 0.0;
            total.converged = total.relerr <= tol * (1 + 1e-12) || bb == 0;
        }
        // returnValues: write back, updatePadding
        for (int i = 0; i < neq; ++i) {
            const opf_field_desc& td = S.tg[i]->d;
            hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, st, N, S.n[0], S.n[1], S.s[0], S.s[1], S.s[2],
                               (double*)S.tg[i]->data, td.offset[0], td.offset[1], td.offset[2], td.extent[0],
                               td.extent[1], (const double*)(x + (size_t)i * N), (const double*)nullptr);
        }
        if ((rc = hipGetLastError() ? OPF_ERR_HIP : OPF_OK)) break;
        for (int i = 0; i < neq && !rc; ++i) {
            S.tg[i]->spare_valid = false;
            rc = launch_ghost_fill(S.tg[i], S.tg[i]->data);
        }
        if (rc) break;
        *state = total;
    } while (false);
    hipStreamSynchronize(st);
    for (Hierarchy& G : H) gmg_free(G);
    hipFree(vec);
    return rc;
}

// HYPREEqnSolveHandler::solve for an operator whose StencilPads are not symmetric
// (advection terms): the pinned row (pinValue: the first unknown is 0, as
// HYPREEqnSolveHandler.hpp:145-163 sets it) is kept literally and the system is
// solved by BiCGStab from the target's values, preconditioned by one V-cycle of
// the operator's symmetric part; then the usual write-back.
static int eqn_solve_nonsym(opf_eqn* e, opf_solve_state* st, bool pin) {
    opf_ctx* ctx = e->ctx;
    hipStream_t s = ctx->stream;
    const long long N = e->N;
    const unsigned nb = (unsigned)((N + 255) / 256);
    if (pin) OPF_HIP(hipMemsetAsync(e->b, 0, sizeof(double), s));
    const opf_field_desc& td = e->target->d;
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, s, N, e->n[0], e->n[1], e->s[0], e->s[1], e->s[2],
                       (const double*)e->target->data, td.offset[0], td.offset[1], td.offset[2], td.extent[0],
                       td.extent[1], e->x);
    EqsBC bc;
    for (int d = 0; d < 3; ++d) {
        bc.per[d] = e->per[d];
        bc.alpha[d][0] = e->alpha[d][0];
        bc.alpha[d][1] = e->alpha[d][1];
    }
    const int dim = td.dim;
    auto apply = [&](hipStream_t q, const double* u, double* y) {
        hipLaunchKernelGGL(k_pad_apply, dim3(nb), dim3(256), 0, q, N, e->n[0], e->n[1], e->n[2], dim,
                           (const double*)e->coef, bc, u, y, e->sign, 0);
        if (pin) hipLaunchKernelGGL(k_pin_row, dim3(1), dim3(1), 0, q, y, u, 0LL);
    };
    auto precond = [&](const double* r, double* z) { return gmg_vcycle64(ctx, e->H, r, z); };
    const double tol = e->params.tol > 0 ? e->params.tol : 1e-6;
    const int maxit = e->params.max_iter > 0 ? e->params.max_iter : 200;
    OPF_TRY(bicgstab(ctx, N, apply, precond, e->x, e->b, tol, maxit, st, e->params.verbose));
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, s, N, e->n[0], e->n[1], e->s[0], e->s[1], e->s[2],
                       (double*)e->target->data, td.offset[0], td.offset[1], td.offset[2], td.extent[0], td.extent[1],
                       (const double*)e->x, (const double*)nullptr);
    OPF_HIP(hipGetLastError());
    e->target->spare_valid = false;
    return launch_ghost_fill(e->target, e->target->data);
}
