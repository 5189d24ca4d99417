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

#include "gmg.hpp"
#include "sweep.hpp"

namespace opf {

constexpr int PAD_K = 7;       // centre + 2 per dim
constexpr int PAD_W = PAD_K + 2;  // coefficients, bias, linear flag
constexpr int PAD_BLOCK = 32;  // 16-deep stack of 9 doubles per lane: 36 KB of LDS

struct PadArgs {
    FieldArg f[SWEEP_MAX_FIELDS];
    const double* x[3];
    const double* dx[3];
    int morg[3];
    int s[3], n[3];      // row box (target interior)
    int tl[3], te[3];    // target local range [tl, te)
    int per[3];
    double beta[3][2];   // ghost rule affine constants
    const void* code;
    int ncode;
    double* coef;        // [PAD_K][N]
    double* bias;        // [N]
    int* err;            // error flags
};

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
                int ox = (in.c & 31) - 8, oy = ((in.c >> 5) & 31) - 8, oz = ((in.c >> 10) & 31) - 8;
                push();
                int sl = slot_of(ox, oy, oz);
                if (sl < 0) { bad |= 1; break; }
                t[sl] = 1.0;
                t[PAD_K + 1] = 1.0;
                int p[3] = {ci + ox, cj + oy, ck + oz};
                for (int d = 0; d < 3; ++d) {
                    if (p[d] >= a.tl[d] && p[d] < a.te[d]) continue;
                    int side = p[d] < a.tl[d] ? 0 : 1;
                    if (!a.per[d]) t[PAD_K] = a.beta[d][side];
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
            default:  // WENO on linear forms is nonlinear
                bad |= 2;
                sp -= 6;
                break;
        }
    }
    if (!live) return;
    if (bad) atomicOr(a.err, bad);
    for (int q = 0; q < PAD_K; ++q) a.coef[(long long)q * N + row] = t[q];
    a.bias[row] = t[PAD_K];
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
    bool built = false;
    bool first = true;
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
        a.per[q] = e->per[q];
        a.beta[q][0] = e->beta[q][0];
        a.beta[q][1] = e->beta[q][1];
    }
    a.code = e->prog;
    a.ncode = (int)e->P.code.size();
    a.coef = e->coef;
    a.bias = e->bias;
    a.err = e->err;
    OPF_HIP(hipMemsetAsync(e->err, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(pad_kernel, dim3((unsigned)((e->N + PAD_BLOCK - 1) / PAD_BLOCK)), dim3(PAD_BLOCK), 0,
                       ctx->stream, a);
    OPF_HIP(hipGetLastError());
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

// rows -> multigrid level (face coefficients, sigma); checks symmetry
static int eqn_build_operator(opf_eqn* e) {
    const long long N = e->N;
    const int dim = e->target->d.dim;
    std::vector<double> c((size_t)PAD_K * N);
    OPF_HIP(hipMemcpyAsync(c.data(), e->coef, sizeof(double) * c.size(), hipMemcpyDeviceToHost, e->ctx->stream));
    OPF_HIP(hipStreamSynchronize(e->ctx->stream));
    // orientation: diagonal must be positive after scaling by sign
    double dsum = 0;
    for (long long r = 0; r < N; ++r) dsum += c[r];
    OPF_CHECK(dsum != 0.0, OPF_ERR_UNSUPPORTED, "equation has no diagonal");
    e->sign = dsum > 0 ? 1.0 : -1.0;
    const double s = e->sign;
    GLevelHost L;
    L.var = true;
    for (int q = 0; q < 3; ++q) L.n[q] = e->n[q];
    L.N = N;
    auto cidx = [&](int i, int j, int k) { return i + (long long)e->n[0] * (j + (long long)e->n[1] * k); };
    double maxa = 0;
    for (int d = 0; d < dim; ++d) {
        int fn[3] = {e->n[0] + (d == 0), e->n[1] + (d == 1), e->n[2] + (d == 2)};
        L.af[d].assign((size_t)fn[0] * fn[1] * fn[2], 0.0);
        for (int k = 0; k < fn[2]; ++k)
            for (int j = 0; j < fn[1]; ++j)
                for (int i = 0; i < fn[0]; ++i) {
                    int p[3] = {i, j, k};
                    double alo = NAN, ahi = NAN;  // from the row above (its lo face) and below (its hi face)
                    if (p[d] < e->n[d]) alo = -s * c[(size_t)(1 + 2 * d) * N + cidx(i, j, k)];
                    if (p[d] > 0) {
                        int q[3] = {i, j, k};
                        q[d] -= 1;
                        ahi = -s * c[(size_t)(2 + 2 * d) * N + cidx(q[0], q[1], q[2])];
                    }
                    double a = std::isnan(alo) ? ahi : alo;
                    if (!std::isnan(alo) && !std::isnan(ahi)) {
                        OPF_CHECK(std::fabs(alo - ahi) <= 1e-12 * std::max(std::fabs(alo), std::fabs(ahi)),
                                  OPF_ERR_UNSUPPORTED, "operator is not symmetric (face %d of dim %d): %g vs %g", i,
                                  d, alo, ahi);
                    }
                    L.af[d][(size_t)i + (size_t)fn[0] * (j + (size_t)fn[1] * k)] = a;
                    maxa = std::max(maxa, std::fabs(a));
                }
        if (e->per[d]) {  // face 0 and face n are the same physical face
            for (int k = 0; k < fn[2]; ++k)
                for (int j = 0; j < fn[1]; ++j)
                    for (int i = 0; i < fn[0]; ++i) {
                        int p[3] = {i, j, k};
                        if (p[d] != 0) continue;
                        int q[3] = {i, j, k};
                        q[d] = e->n[d];
                        double& f0 = L.af[d][(size_t)p[0] + (size_t)fn[0] * (p[1] + (size_t)fn[1] * p[2])];
                        double fnv = L.af[d][(size_t)q[0] + (size_t)fn[0] * (q[1] + (size_t)fn[1] * q[2])];
                        OPF_CHECK(std::fabs(f0 - fnv) <= 1e-12 * std::max(std::fabs(f0), std::fabs(fnv)),
                                  OPF_ERR_UNSUPPORTED, "periodic faces disagree in dim %d", d);
                    }
        }
    }
    L.sg.assign(N, 0.0);
    for (int k = 0; k < e->n[2]; ++k)
        for (int j = 0; j < e->n[1]; ++j)
            for (int i = 0; i < e->n[0]; ++i) {
                long long r = cidx(i, j, k);
                double sum = 0;
                int p[3] = {i, j, k};
                for (int d = 0; d < dim; ++d) {
                    int fn0 = e->n[0] + (d == 0), fn1 = e->n[1] + (d == 1);
                    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
                    hi[d] += 1;
                    sum += L.af[d][(size_t)lo[0] + (size_t)fn0 * (lo[1] + (size_t)fn1 * lo[2])];
                    sum += L.af[d][(size_t)hi[0] + (size_t)fn0 * (hi[1] + (size_t)fn1 * hi[2])];
                    (void)p;
                }
                double sg = s * c[r] - sum;
                if (std::fabs(sg) <= 1e-12 * std::max(1.0, sum)) sg = 0.0;
                L.sg[r] = sg;
            }
    // constant-coefficient detection (uniform operator => matrix-free constant mode)
    bool cst = true;
    double a0[3] = {0, 0, 0};
    for (int d = 0; d < dim && cst; ++d) {
        a0[d] = L.af[d][0];
        for (double v : L.af[d]) cst &= v == a0[d];
    }
    for (double v : L.sg) cst &= v == L.sg[0];
    if (cst) {
        GLevelHost C;
        C.var = false;
        for (int q = 0; q < 3; ++q) {
            C.n[q] = L.n[q];
            C.a[q] = a0[q];
        }
        C.N = N;
        C.sigma = L.sg[0];
        L = C;
    }
    if (e->built) gmg_free(e->H);
    e->H = Hierarchy{};
    if (e->params.pre_smooth > 0) e->H.pre = e->params.pre_smooth;
    if (e->params.post_smooth > 0) e->H.post = e->params.post_smooth;
    int rc = gmg_build(e->ctx, e->H, L, e->per, e->alpha,
                       e->params.precision == OPF_PREC_MIXED ? OPF_PREC_MIXED : OPF_PREC_FP64);
    if (rc != OPF_OK) return rc;
    e->built = true;
    return OPF_OK;
}

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
    GhostInfo gi;
    rc = build_ghost_info(target, &gi);
    if (rc != OPF_OK) {
        delete e;
        return rc;
    }
    for (int d = 0; d < td.dim; ++d) {
        e->per[d] = gi.per[d];
        for (int s = 0; s < 2; ++s) {
            const GhostRule& r = gi.rule[d][s][0];
            switch (r.type) {
                case 1: e->alpha[d][s] = r.c1 / r.c3; e->beta[d][s] = (0.0 * r.c1 - r.c2) / r.c3; break;
                case 2: e->alpha[d][s] = 1.0; e->beta[d][s] = r.c1; break;
                case 3: e->alpha[d][s] = 1.0; e->beta[d][s] = 0.0; break;
                case 4: e->alpha[d][s] = -1.0; e->beta[d][s] = 0.0; break;
                case 5: break;
                default:
                    delete e;
                    set_error("implicit solve: every boundary needs a BC");
                    return OPF_ERR_UNSUPPORTED;
            }
        }
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
    void* p[] = {e->coef, e->bias, e->b, e->x, e->shift, e->err, e->zbuf};
    for (void* q : p)
        if (q) hipFree(q);
    delete e;
    return OPF_OK;
}
