// Multigrid levels built on the device from the StencilPad rows of one diagonal
// block: the device form of the matrix setup behind generateAb
// (HYPREEqnSolveHandler.hpp:125-166, where HYPRE receives the rows and builds its
// own hierarchy). The rows are PAD_K coefficient arrays of N unknowns (slot 0: the
// cell itself, slot 1 + 2d + s: its neighbour on side s of dim d, x-fastest over
// the target's assignable range). Nothing of them is copied to the host: the face
// coefficients, the per-cell shift, the symmetry and uniformity tests and every
// coarse level are kernels; the host reads back one small record per level build
// and the coarsest level (at most 1024 cells) for its dense inverse.
#include <algorithm>
#include <cmath>

#include "gmg.hpp"

namespace opf {
namespace {

constexpr int kNT = 256;
constexpr int kDiagBlocks = 1024;
constexpr int kFlagBlocks = 2048;

// results of one rows -> level build that the host needs
struct LevelMeta {
    double dsum;   // sum of the diagonal slots (orients the operator)
    double sign;   // +1 / -1, 0 when the rows have no diagonal
    int asym;      // a face's two rows disagree: the level holds the symmetric part
    int nonconst;  // a coefficient differs from the first of its array
    int shift;     // a cell has a nonzero sigma
    int pad;
    double a0[3];  // first face coefficient per dim
    double sg0;    // first sigma
};

__device__ double blk_sum(double v) {
    __shared__ double sh[kNT / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    for (int q = 0; q < kNT / 64; ++q) r += sh[q];
    return r;
}

// sets *flag when pred holds on any thread of the block; every thread of the
// block calls it (one atomic per block, not per thread: a flag that most
// threads raise would otherwise serialise the whole grid on one address)
__device__ __forceinline__ void block_flag(bool pred, int* flag) {
    if (__syncthreads_or(pred) && threadIdx.x == 0) atomicOr(flag, 1);
}

__device__ __forceinline__ bool differ(double u, double v) { return fabs(u - v) > 1e-12 * fmax(fabs(u), fabs(v)); }

__device__ __forceinline__ long long cell_lin(const int p[3], const int n[3]) {
    return p[0] + (long long)n[0] * (p[1] + (long long)n[1] * p[2]);
}

struct Ext {
    int n[3];
};

__global__ __launch_bounds__(kNT) void k_diag_part(long long N, const double* __restrict__ c, double* __restrict__ part) {
    double s = 0;
    for (long long r = (long long)blockIdx.x * kNT + threadIdx.x; r < N; r += (long long)gridDim.x * kNT) s += c[r];
    s = blk_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kNT) void k_diag_final(int G, const double* __restrict__ part, LevelMeta* m) {
    double s = 0;
    for (int q = threadIdx.x; q < G; q += kNT) s += part[q];
    s = blk_sum(s);
    if (threadIdx.x == 0) {
        m->dsum = s;
        m->sign = s > 0 ? 1.0 : s < 0 ? -1.0 : 0.0;
        m->asym = 0;
        m->nonconst = 0;
        m->shift = 0;
    }
}

// face coefficient of dim d: minus the oriented neighbour slot of the row above
// the face (its lo side), else of the row below (its hi side); two rows that
// disagree give their mean
__global__ __launch_bounds__(kNT) void k_faces(int d, long long N, const double* __restrict__ c, Ext e,
                                               LevelMeta* m, double* __restrict__ af) {
    const int fn[3] = {e.n[0] + (d == 0), e.n[1] + (d == 1), e.n[2] + (d == 2)};
    const long long nf = (long long)fn[0] * fn[1] * fn[2];
    const long long f = (long long)blockIdx.x * kNT + threadIdx.x;
    bool asym = false;
    if (f < nf) {
        int p[3];
        p[0] = (int)(f % fn[0]);
        p[1] = (int)((f / fn[0]) % fn[1]);
        p[2] = (int)(f / ((long long)fn[0] * fn[1]));
        const double s = m->sign;
        const bool has_lo = p[d] < e.n[d], has_hi = p[d] > 0;
        double alo = 0, ahi = 0;
        if (has_lo) alo = -s * c[(size_t)(1 + 2 * d) * N + cell_lin(p, e.n)];
        if (has_hi) {
            int q[3] = {p[0], p[1], p[2]};
            q[d] -= 1;
            ahi = -s * c[(size_t)(2 + 2 * d) * N + cell_lin(q, e.n)];
        }
        double a = has_lo ? alo : ahi;
        if (has_lo && has_hi && differ(alo, ahi)) {
            a = 0.5 * (alo + ahi);
            asym = true;
        }
        af[f] = a;
    }
    block_flag(asym, &m->asym);
}

// a periodic dim's faces 0 and n are the same physical face
__global__ __launch_bounds__(kNT) void k_faces_periodic(int d, Ext e, LevelMeta* m, double* __restrict__ af) {
    const int fn[3] = {e.n[0] + (d == 0), e.n[1] + (d == 1), e.n[2] + (d == 2)};
    int pn[3] = {fn[0], fn[1], fn[2]};
    pn[d] = 1;
    const long long t = (long long)blockIdx.x * kNT + threadIdx.x;
    bool asym = false;
    if (t < (long long)pn[0] * pn[1] * pn[2]) {
        int p[3];
        p[0] = (int)(t % pn[0]);
        p[1] = (int)((t / pn[0]) % pn[1]);
        p[2] = (int)(t / ((long long)pn[0] * pn[1]));
        int q[3] = {p[0], p[1], p[2]};
        q[d] = e.n[d];
        const long long i0 = cell_lin(p, fn), i1 = cell_lin(q, fn);
        const double f0 = af[i0], f1 = af[i1];
        if (differ(f0, f1)) {
            af[i0] = af[i1] = 0.5 * (f0 + f1);
            asym = true;
        }
    }
    block_flag(asym, &m->asym);
}

// sigma = oriented diagonal minus the cell's face coefficients (dims in order, lo
// face then hi face), zeroed when it is rounding noise of that sum
__global__ __launch_bounds__(kNT) void k_sigma(int dim, long long N, const double* __restrict__ c, Ext e,
                                               const LevelMeta* m, const double* __restrict__ af0,
                                               const double* __restrict__ af1, const double* __restrict__ af2,
                                               double* __restrict__ sg) {
    const long long r = (long long)blockIdx.x * kNT + threadIdx.x;
    if (r >= N) return;
    const int i = (int)(r % e.n[0]), j = (int)((r / e.n[0]) % e.n[1]), k = (int)(r / ((long long)e.n[0] * e.n[1]));
    const double* af[3] = {af0, af1, af2};
    double sum = 0;
    for (int d = 0; d < dim; ++d) {
        int hi[3] = {i, j, k};
        hi[d] += 1;
        sum += af[d][face_index(d, i, j, k, e.n[0], e.n[1])];
        sum += af[d][face_index(d, hi[0], hi[1], hi[2], e.n[0], e.n[1])];
    }
    double v = m->sign * c[r] - sum;
    if (fabs(v) <= 1e-12 * fmax(1.0, sum)) v = 0.0;
    sg[r] = v;
}

// uniformity (every entry equals the first) and, for sigma, any nonzero entry;
// grid-stride over at most kFlagBlocks blocks
__global__ __launch_bounds__(kNT) void k_flags(long long n, const double* __restrict__ v, LevelMeta* m, int is_sg) {
    const double v0 = v[0];
    bool nonconst = false, shift = false;
    for (long long t = (long long)blockIdx.x * kNT + threadIdx.x; t < n; t += (long long)gridDim.x * kNT) {
        const double x = v[t];
        nonconst |= x != v0;
        shift |= x != 0.0;
    }
    block_flag(nonconst, &m->nonconst);
    if (is_sg) block_flag(shift, &m->shift);
}

__global__ void k_firsts(const double* __restrict__ af0, const double* __restrict__ af1,
                         const double* __restrict__ af2, const double* __restrict__ sg, LevelMeta* m) {
    m->a0[0] = af0 ? af0[0] : 0.0;
    m->a0[1] = af1 ? af1[0] : 0.0;
    m->a0[2] = af2 ? af2[0] : 0.0;
    m->sg0 = sg[0];
}

// coarse face = mean of the fine faces it covers, scaled by 1/4 (cell-centred
// coarsening by 2 of a divergence-form operator: face area x2^(dim-1), spacing x2)
__global__ __launch_bounds__(kNT) void k_coarsen_faces(int d, int dim, Ext F, Ext C, const double* __restrict__ ff,
                                                       double* __restrict__ cf) {
    const int e0 = C.n[0] + (d == 0), e1 = C.n[1] + (d == 1), e2 = C.n[2] + (d == 2);
    const long long t = (long long)blockIdx.x * kNT + threadIdx.x;
    if (t >= (long long)e0 * e1 * e2) return;
    const int I = (int)(t % e0), J = (int)((t / e0) % e1), K = (int)(t / ((long long)e0 * e1));
    const int nz = (dim == 3 && d != 2) ? 2 : 1, ny = d != 1 ? 2 : 1, nx = d != 0 ? 2 : 1;
    const int tw = dim == 3 ? 4 : 2;
    double s = 0;
    for (int dk = 0; dk < nz; ++dk)
        for (int dj = 0; dj < ny; ++dj)
            for (int di = 0; di < nx; ++di) {
                const int fi = 2 * I + di, fj = 2 * J + dj, fk = dim == 3 ? 2 * K + dk : 0;
                s += ff[face_index(d, fi, fj, fk, F.n[0], F.n[1])];
            }
    cf[t] = s / tw * 0.25;
}

// coarse sigma = mean of the children's
__global__ __launch_bounds__(kNT) void k_coarsen_sigma(int dim, Ext F, Ext C, const double* __restrict__ fs,
                                                       double* __restrict__ cs) {
    const long long c = (long long)blockIdx.x * kNT + threadIdx.x;
    const long long NC = (long long)C.n[0] * C.n[1] * C.n[2];
    if (c >= NC) return;
    const int I = (int)(c % C.n[0]), J = (int)((c / C.n[0]) % C.n[1]), K = (int)(c / ((long long)C.n[0] * C.n[1]));
    const int nz = dim == 3 ? 2 : 1;
    double s = 0;
    for (int dk = 0; dk < nz; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di) {
                const long long f = (2 * I + di) +
                                    (long long)F.n[0] * ((2 * J + dj) + (long long)F.n[1] * (dim == 3 ? 2 * K + dk : 0));
                s += fs[f];
            }
    cs[c] = s / (1 << dim);
}

__global__ __launch_bounds__(kNT) void k_to_f32(long long n, const double* __restrict__ s, float* __restrict__ d) {
    const long long t = (long long)blockIdx.x * kNT + threadIdx.x;
    if (t < n) d[t] = (float)s[t];
}

__global__ __launch_bounds__(kNT) void k_neq(long long n, const double* __restrict__ a, const double* __restrict__ b,
                                             int* flag) {
    bool ne = false;
    for (long long t = (long long)blockIdx.x * kNT + threadIdx.x; t < n; t += (long long)gridDim.x * kNT)
        ne |= a[t] != b[t];
    block_flag(ne, flag);
}

inline unsigned blocks(long long n) { return (unsigned)((n + kNT - 1) / kNT); }
// grid of the grid-stride flag kernels (k_flags, k_neq)
inline unsigned flag_blocks(long long n) { return std::min(blocks(n), (unsigned)kFlagBlocks); }

inline Ext ext_of(const int n[3]) { return Ext{{n[0], n[1], n[2]}}; }

}  // namespace

long long level_face_count(const int n[3], int d) {
    return (long long)(n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2));
}

void dev_level_free(DevLevel& L) {
    for (int d = 0; d < 3; ++d) {
        if (L.af[d]) hipFree(L.af[d]);
        L.af[d] = nullptr;
    }
    if (L.sg) hipFree(L.sg);
    L.sg = nullptr;
}

int rows_to_level_dev(hipStream_t st, const double* coef, long long N, const int nn[3], int dim, const int per[3],
                      double* sign, DevLevel& L, bool* symmetric) {
    OPF_CHECK(N > 0, OPF_ERR_ARG, "rows_to_level_dev: empty block");
    for (int q = 0; q < 3; ++q) L.n[q] = nn[q];
    L.N = N;
    // device scratch: the record and the diagonal partials (freed before return)
    char* scratch = nullptr;
    const size_t meta_off = 0, part_off = 256;
    OPF_HIP(hipMalloc((void**)&scratch, part_off + sizeof(double) * kDiagBlocks));
    LevelMeta* m = (LevelMeta*)(scratch + meta_off);
    double* part = (double*)(scratch + part_off);
    int rc = OPF_OK;
    for (int d = 0; d < dim && rc == OPF_OK; ++d)
        if (!L.af[d] && hipMalloc((void**)&L.af[d], sizeof(double) * level_face_count(nn, d)) != hipSuccess)
            rc = OPF_ERR_NOMEM;
    if (rc == OPF_OK && !L.sg && hipMalloc((void**)&L.sg, sizeof(double) * N) != hipSuccess) rc = OPF_ERR_NOMEM;
    LevelMeta hm{};
    if (rc == OPF_OK) {
        const Ext e = ext_of(nn);
        const int G = (int)std::min<long long>(kDiagBlocks, (N + kNT - 1) / kNT);
        hipLaunchKernelGGL(k_diag_part, dim3(G), dim3(kNT), 0, st, N, coef, part);
        hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kNT), 0, st, G, (const double*)part, m);
        for (int d = 0; d < dim; ++d) {
            hipLaunchKernelGGL(k_faces, dim3(blocks(level_face_count(nn, d))), dim3(kNT), 0, st, d, N, coef, e, m,
                               L.af[d]);
            if (per[d]) {
                long long np = level_face_count(nn, d) / (nn[d] + 1);
                hipLaunchKernelGGL(k_faces_periodic, dim3(blocks(np)), dim3(kNT), 0, st, d, e, m, L.af[d]);
            }
        }
        hipLaunchKernelGGL(k_sigma, dim3(blocks(N)), dim3(kNT), 0, st, dim, N, coef, e, (const LevelMeta*)m,
                           (const double*)L.af[0], (const double*)L.af[1], (const double*)L.af[2], L.sg);
        for (int d = 0; d < dim; ++d)
            hipLaunchKernelGGL(k_flags, dim3(flag_blocks(level_face_count(nn, d))), dim3(kNT), 0, st,
                               level_face_count(nn, d), (const double*)L.af[d], m, 0);
        hipLaunchKernelGGL(k_flags, dim3(flag_blocks(N)), dim3(kNT), 0, st, N, (const double*)L.sg, m, 1);
        hipLaunchKernelGGL(k_firsts, dim3(1), dim3(1), 0, st, (const double*)L.af[0], (const double*)L.af[1],
                           (const double*)L.af[2], (const double*)L.sg, m);
        if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&hm, m, sizeof(hm), hipMemcpyDeviceToHost, st) ||
            hipStreamSynchronize(st))
            rc = OPF_ERR_HIP;
    }
    hipFree(scratch);
    if (rc == OPF_ERR_NOMEM) set_error("rows_to_level_dev: device allocation failed");
    if (rc == OPF_ERR_HIP) set_error("rows_to_level_dev: HIP error");
    if (rc != OPF_OK) return rc;
    OPF_CHECK(hm.sign != 0.0, OPF_ERR_UNSUPPORTED, "equation has no diagonal");
    *sign = hm.sign;
    if (symmetric) *symmetric = hm.asym == 0;
    L.shift = hm.shift != 0;
    // a uniform operator runs matrix-free (constant mode); its arrays stay allocated
    // for the caller to reuse or free
    L.var = hm.nonconst != 0;
    for (int d = 0; d < 3; ++d) L.a[d] = !L.var && d < dim ? hm.a0[d] : 0.0;
    L.sigma = L.var ? 0.0 : hm.sg0;
    return OPF_OK;
}

int dev_levels_equal(hipStream_t st, const DevLevel& a, const DevLevel& b, int dim, bool* eq) {
    *eq = false;
    if (a.var != b.var || a.N != b.N) return OPF_OK;
    for (int q = 0; q < 3; ++q)
        if (a.n[q] != b.n[q]) return OPF_OK;
    if (!a.var) {
        *eq = a.a[0] == b.a[0] && a.a[1] == b.a[1] && a.a[2] == b.a[2] && a.sigma == b.sigma;
        return OPF_OK;
    }
    bool alias = a.sg == b.sg;
    for (int d = 0; d < dim; ++d) alias &= a.af[d] == b.af[d];
    if (alias) {
        *eq = true;
        return OPF_OK;
    }
    int* flag = nullptr;
    OPF_HIP(hipMalloc((void**)&flag, sizeof(int)));
    int h = 0;
    int rc = OPF_OK;
    if (hipMemcpyAsync(flag, &h, sizeof(int), hipMemcpyHostToDevice, st)) rc = OPF_ERR_HIP;
    for (int d = 0; d < dim && !rc; ++d) {
        const long long nf = level_face_count(a.n, d);
        hipLaunchKernelGGL(k_neq, dim3(flag_blocks(nf)), dim3(kNT), 0, st, nf, (const double*)a.af[d],
                           (const double*)b.af[d], flag);
    }
    if (!rc) hipLaunchKernelGGL(k_neq, dim3(flag_blocks(a.N)), dim3(kNT), 0, st, a.N, (const double*)a.sg,
                                (const double*)b.sg, flag);
    if (!rc && (hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st) || hipStreamSynchronize(st)))
        rc = OPF_ERR_HIP;
    hipFree(flag);
    if (rc) return (set_error("dev_levels_equal: HIP error"), rc);
    *eq = h == 0;
    return OPF_OK;
}

void coarsen_dev(hipStream_t st, int dim, const int fn[3], const double* const faf[3], const double* fsg,
                 const int cn[3], double* const caf[3], double* csg) {
    const Ext F = ext_of(fn), C = ext_of(cn);
    for (int d = 0; d < dim; ++d)
        hipLaunchKernelGGL(k_coarsen_faces, dim3(blocks(level_face_count(cn, d))), dim3(kNT), 0, st, d, dim, F, C,
                           faf[d], caf[d]);
    const long long NC = (long long)cn[0] * cn[1] * cn[2];
    hipLaunchKernelGGL(k_coarsen_sigma, dim3(blocks(NC)), dim3(kNT), 0, st, dim, F, C, fsg, csg);
}

void to_f32_dev(hipStream_t st, long long n, const double* src, float* dst) {
    hipLaunchKernelGGL(k_to_f32, dim3(blocks(n)), dim3(kNT), 0, st, n, src, dst);
}

}  // namespace opf
