// Matrix-free geometric multigrid preconditioned flexible CG on gfx950.
//
// Levels are cell-centred with coarsening by 2 in every dimension;
// restriction averages the 2^d children's residuals, prolongation is
// (bi/tri)linear with BC-consistent coarse ghosts, the smoother is weighted
// Jacobi, the coarsest level is solved with a dense FP64 inverse. The Krylov
// method is flexible CG (Polak-Ribiere beta), so the slightly non-symmetric
// V-cycle (average restriction, linear prolongation) is harmless. All
// reductions are two-pass and fixed-order (deterministic); the CG scalars
// live in device memory so an iteration needs no host round trip.
//
// Mixed precision: the V-cycle and the inner CG run in FP32 with FP32
// storage, FP64 is kept for the outer residual b - A x, the solution update
// and the coarsest solve (iterative refinement).
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "gmg.hpp"

namespace opf {

// ------------------------------------------------------------------ helpers
template <typename T>
__device__ __forceinline__ T face_coef(const GOp<T>& o, bool var, int d, int i, int j, int k) {
    return var ? o.af[d][face_index(d, i, j, k, o.n0, o.n1)] : o.a[d];
}

// (A x)_c and the diagonal D_c at cell (i, j, k) = linear index t;
// xv(l) returns x at linear cell index l.
template <typename T, bool VAR, int DIM, typename XV>
__device__ __forceinline__ T apply_op(const GOp<T>& o, int i, int j, int k, uint32_t t, T xc, XV xv, T& D) {
    T acc = 0;
    D = VAR ? o.sg[t] : o.sigma;
    const int c[3] = {i, j, k};
    const int n[3] = {o.n0, o.n1, o.n2};
    const uint32_t st[3] = {1u, o.s1, o.s2};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        int hi[3] = {i, j, k};
        hi[d] += 1;
        T alo = face_coef<T>(o, VAR, d, i, j, k);
        T ahi = face_coef<T>(o, VAR, d, hi[0], hi[1], hi[2]);
        if (c[d] > 0) {
            D += alo;
            acc += alo * xv(t - st[d]);
        } else if (o.per[d]) {
            D += alo;
            acc += alo * xv(t + (uint32_t)(n[d] - 1) * st[d]);
        } else {
            D += alo * o.bfac[d][0];
        }
        if (c[d] < n[d] - 1) {
            D += ahi;
            acc += ahi * xv(t + st[d]);
        } else if (o.per[d]) {
            D += ahi;
            acc += ahi * xv(t - (uint32_t)(n[d] - 1) * st[d]);
        } else {
            D += ahi * o.bfac[d][1];
        }
    }
    return D * xc - acc;
}

// cell of this thread in a one-cell-per-thread launch (XCD-contiguous blocks)
__device__ __forceinline__ uint32_t cell_id() { return xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; }

// contiguous range of cells of this block in a fixed-grid reduction launch
// (ranges are multiples of the block size; the partial-sum order depends only
// on N and the grid, so results are reproducible)
__device__ __forceinline__ void block_range(uint32_t N, uint32_t& beg, uint32_t& end) {
    const uint32_t G = gridDim.x;
    const uint32_t per = (((N + G - 1) / G) + blockDim.x - 1) / blockDim.x * blockDim.x;
    const uint32_t lb = xcd_block(blockIdx.x, G);
    beg = min(N, lb * per);
    end = min(N, beg + per);
}

// linear cell index -> (i, j, k) for extents n0, n1 (f0, f1 divide by them)
__device__ __forceinline__ void cell_of(uint32_t t, const FastDiv& f0, const FastDiv& f1, int& i, int& j, int& k) {
    const uint32_t r = f0.div(t);
    i = (int)(t - r * f0.d);
    const uint32_t q = f1.div(r);
    j = (int)(r - q * f1.d);
    k = (int)q;
}

// Batched execution: blockIdx.y selects a request slot; slot vectors are
// stored back to back (stride vs), scalars and reduction partials likewise.
// mask (optional, device) switches slots off without changing the launch.
struct Bat {
    const int* mask;
    long long vs;
    int ss;
    int ps;
};
#define BAT_SLOT(b)                                  \
    const int slot = blockIdx.y;                     \
    if ((b).mask && !(b).mask[slot]) return;         \
    const long long vo = (long long)slot * (b).vs;   \
    (void)vo;

__device__ double gmg_block_sum(double v) {
    __shared__ double sh[16];  // blocks of up to 1024 threads
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) {
        int nw = (blockDim.x + 63) >> 6;
        for (int q = 0; q < nw; ++q) r += sh[q];
        sh[0] = r;
    }
    __syncthreads();
    r = sh[0];
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------ kernels
// r = b - A x  (x may be null: zero), optional partial sum of r^2 (grid-stride,
// fixed grid => deterministic). TO: output precision.
template <typename T, typename TO, bool VAR, int DIM>
__global__ __launch_bounds__(256) void k_resid(GOp<T> o, const T* __restrict__ x, const T* __restrict__ b,
                                               TO* __restrict__ r, double* __restrict__ part, Bat bt) {
    BAT_SLOT(bt);
    const uint32_t N = o.s2 * (uint32_t)o.n2;
    if (x) x += vo;
    b += vo;
    if (r) r += vo;
    if (part) part += (long long)slot * bt.ps;
    double acc = 0;
    uint32_t tb, te;
    block_range(N, tb, te);
    for (uint32_t t = tb + threadIdx.x; t < te; t += blockDim.x) {
        T v = b[t];
        if (x) {
            int i, j, k;
            cell_of(t, o.f0, o.f1, i, j, k);
            auto xv = [&](uint32_t l) { return x[l]; };
            T D;
            v = v - apply_op<T, VAR, DIM>(o, i, j, k, t, x[t], xv, D);
        }
        if (r) r[t] = (TO)v;
        acc += (double)v * (double)v;
    }
    if (part) {
        acc = gmg_block_sum(acc);
        if (threadIdx.x == 0) part[blockIdx.x] = acc;
    }
}

// y = A x
template <typename T, bool VAR, int DIM>
__global__ __launch_bounds__(256) void k_apply(GOp<T> o, const T* __restrict__ x, T* __restrict__ y) {
    const uint32_t N = o.s2 * (uint32_t)o.n2;
    const uint32_t t = cell_id();
    if (t >= N) return;
    int i, j, k;
    cell_of(t, o.f0, o.f1, i, j, k);
    auto xv = [&](uint32_t l) { return x[l]; };
    T D;
    y[t] = apply_op<T, VAR, DIM>(o, i, j, k, t, x[t], xv, D);
}

// weighted Jacobi: xo = xi + w (b - A xi)/D ; ZERO: xo = w b / D
template <typename T, bool VAR, int DIM, bool ZERO>
__global__ __launch_bounds__(256) void k_jacobi(GOp<T> o, const T* __restrict__ xi, const T* __restrict__ b,
                                                T* __restrict__ xo, T w, Bat bt) {
    BAT_SLOT(bt);
    const uint32_t N = o.s2 * (uint32_t)o.n2;
    if (xi) xi += vo;
    b += vo;
    xo += vo;
    const uint32_t t = cell_id();
    if (t >= N) return;
    int i, j, k;
    cell_of(t, o.f0, o.f1, i, j, k);
    T D;
    if (ZERO) {
        auto zv = [&](uint32_t) { return (T)0; };
        apply_op<T, VAR, DIM>(o, i, j, k, t, (T)0, zv, D);
        xo[t] = w * b[t] / D;
    } else {
        auto xv = [&](uint32_t l) { return xi[l]; };
        T xc = xi[t];
        T ax = apply_op<T, VAR, DIM>(o, i, j, k, t, xc, xv, D);
        xo[t] = xc + w * (b[t] - ax) / D;
    }
}

// Diagonal of the constant-coefficient operator at cell (i, j, k), accumulated
// in exactly apply_op's order (so it has the same bits).
template <typename T, int DIM>
__device__ __forceinline__ T diag_const(const GOp<T>& o, int i, int j, int k) {
    T D = o.sigma;
    const int c[3] = {i, j, k};
    const int n[3] = {o.n0, o.n1, o.n2};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        if (c[d] > 0 || o.per[d]) D += o.a[d];
        else D += o.a[d] * o.bfac[d][0];
        if (c[d] < n[d] - 1 || o.per[d]) D += o.a[d];
        else D += o.a[d] * o.bfac[d][1];
    }
    return D;
}

// The first two pre-smoothing sweeps from x = 0 in one pass (constant
// coefficients): x1 = w b / D at every cell, x2 = x1 + w (b - A x1) / D. x1 at
// the neighbours is recomputed from b there instead of going through memory;
// the arithmetic is that of k_jacobi<ZERO> followed by k_jacobi, so x2 has the
// same bits as the two-launch sequence.
template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_jacobi2z(GOp<T> o, const T* __restrict__ b, T* __restrict__ xo, T w, Bat bt) {
    BAT_SLOT(bt);
    const uint32_t N = o.s2 * (uint32_t)o.n2;
    b += vo;
    xo += vo;
    const uint32_t t = cell_id();
    if (t >= N) return;
    int i, j, k;
    cell_of(t, o.f0, o.f1, i, j, k);
    const int c[3] = {i, j, k};
    const int n[3] = {o.n0, o.n1, o.n2};
    const uint32_t st[3] = {1u, o.s1, o.s2};
    auto x1 = [&](uint32_t l, int d, int cd) {
        int q[3] = {i, j, k};
        q[d] = cd;
        return w * b[l] / diag_const<T, DIM>(o, q[0], q[1], q[2]);
    };
    const T bc = b[t];
    T D = o.sigma, acc = 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        if (c[d] > 0) {
            D += o.a[d];
            acc += o.a[d] * x1(t - st[d], d, c[d] - 1);
        } else if (o.per[d]) {
            D += o.a[d];
            acc += o.a[d] * x1(t + (uint32_t)(n[d] - 1) * st[d], d, n[d] - 1);
        } else {
            D += o.a[d] * o.bfac[d][0];
        }
        if (c[d] < n[d] - 1) {
            D += o.a[d];
            acc += o.a[d] * x1(t + st[d], d, c[d] + 1);
        } else if (o.per[d]) {
            D += o.a[d];
            acc += o.a[d] * x1(t - (uint32_t)(n[d] - 1) * st[d], d, 0);
        } else {
            D += o.a[d] * o.bfac[d][1];
        }
    }
    const T xc = w * bc / D;
    const T ax = D * xc - acc;
    xo[t] = xc + w * (bc - ax) / D;
}

// last post-smoothing sweep of the finest level fused with the dot products
// the flexible CG needs next: out = xi + w (b - A xi)/D, then partials of
// (b, out), (out, q) and sum(out) (b is the residual r the V-cycle smooths).
// Same block ranges and per-thread order as k_dot3, so the sums are identical.
template <typename T, bool VAR, int DIM>
__global__ __launch_bounds__(256) void k_jacobi_dot3(GOp<T> o, const T* __restrict__ xi, const T* __restrict__ b,
                                                     T* __restrict__ xo, T w, const T* __restrict__ q,
                                                     double* __restrict__ part, Bat bt) {
    BAT_SLOT(bt);
    const uint32_t N = o.s2 * (uint32_t)o.n2;
    xi += vo;
    b += vo;
    xo += vo;
    q += vo;
    part += (long long)slot * bt.ps;
    double sa = 0, sb = 0, sc = 0;
    uint32_t tb, te;
    block_range(N, tb, te);
    for (uint32_t t = tb + threadIdx.x; t < te; t += blockDim.x) {
        int i, j, k;
        cell_of(t, o.f0, o.f1, i, j, k);
        auto xv = [&](uint32_t l) { return xi[l]; };
        T D;
        const T xc = xi[t];
        const T bv = b[t];
        const T ax = apply_op<T, VAR, DIM>(o, i, j, k, t, xc, xv, D);
        const T zv = xc + w * (bv - ax) / D;
        xo[t] = zv;
        const double zd = (double)zv;
        sa += (double)bv * zd;
        sb += zd * (double)q[t];
        sc += zd;
    }
    sa = gmg_block_sum(sa);
    sb = gmg_block_sum(sb);
    sc = gmg_block_sum(sc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sa;
        part[GMG_MAXB + blockIdx.x] = sb;
        part[2 * GMG_MAXB + blockIdx.x] = sc;
    }
}

// ---- constant-coefficient kernels over 64 x 4 tiles (3-D: marching in z)
// A block of 64 x 4 threads owns a 64 x 4 column of the x-y plane and (3-D) walks ZC
// planes in z, keeping the stencilled vector's values at k-1, k, k+1 in
// registers (the z neighbours come for free). Cells with no boundary face take
// a straight-line path (interior diagonal, omega / D precomputed); the others
// take apply_op's general path. Launches are a fixed grid striding over
// (tile, z chunk) items, so the partial sums of reductions are reproducible.
struct ZTile {
    int tx, ty, nz, zc;  // tiles in x, y, z chunks, planes per chunk
    int items;
};

__device__ __forceinline__ void ztile_item(const ZTile& z, int item, int& i, int& j, int& k0, int& k1, int n0, int n1,
                                           int n2) {
    const int x = item % z.tx;
    const int r = item / z.tx;
    const int y = r % z.ty;
    const int c = r / z.ty;
    i = x * 64 + (int)(threadIdx.x & 63);
    j = y * 4 + (int)(threadIdx.x >> 6);
    k0 = c * z.zc;
    k1 = min(k0 + z.zc, n2);
    if (i >= n0 || j >= n1) k1 = k0;  // lanes past the grid edge do no cells
}

template <typename T, int DIM>
__device__ __forceinline__ bool interior3(const GOp<T>& o, int i, int j, int k) {
    return i > 0 && i < o.n0 - 1 && j > 0 && j < o.n1 - 1 && (DIM == 2 || (k > 0 && k < o.n2 - 1));
}

// (A v)_c for an interior cell: x/y neighbours through the accessor, z neighbours from the window
template <typename T, int DIM, typename V>
__device__ __forceinline__ T apply_in3(const GOp<T>& o, uint32_t t, T vc, T vm, T vp, V v) {
    T acc = o.a[0] * v(t - 1);
    acc += o.a[0] * v(t + 1);
    acc += o.a[1] * v(t - o.s1);
    acc += o.a[1] * v(t + o.s1);
    if (DIM == 3) {
        acc += o.a[2] * vm;
        acc += o.a[2] * vp;
    }
    return o.dint * vc - acc;
}

// weighted Jacobi sweep xo = xi + w (b - A xi) / D
template <typename T, int DIM>
__global__ __launch_bounds__(256) void zk_jacobi(GOp<T> o, const T* __restrict__ xi, const T* __restrict__ b,
                                                 T* __restrict__ xo, T w, ZTile zt, Bat bt) {
    BAT_SLOT(bt);
    xi += vo;
    b += vo;
    xo += vo;
    auto xv = [&](uint32_t l) { return xi[l]; };
    for (int item = (int)xcd_block(blockIdx.x, gridDim.x); item < zt.items; item += gridDim.x) {
        int i, j, k0, k1;
        ztile_item(zt, item, i, j, k0, k1, o.n0, o.n1, o.n2);
        if (k0 >= k1) continue;
        uint32_t t = (uint32_t)i + o.s1 * (uint32_t)j + o.s2 * (uint32_t)k0;
        T vm = (DIM == 3 && k0 > 0) ? xi[t - o.s2] : (T)0, vc = xi[t];
        for (int k = k0; k < k1; ++k, t += o.s2) {
            const T vp = (DIM == 3 && k + 1 < o.n2) ? xi[t + o.s2] : (T)0;
            const T bv = b[t];
            if (interior3<T, DIM>(o, i, j, k)) {
                xo[t] = vc + o.wdi * (bv - apply_in3<T, DIM>(o, t, vc, vm, vp, xv));
            } else {
                T D;
                const T ax = apply_op<T, false, DIM>(o, i, j, k, t, vc, xv, D);
                xo[t] = vc + w * (bv - ax) / D;
            }
            vm = vc;
            vc = vp;
        }
    }
}

// last sweep of the finest level + partials of (b, out), (out, q), sum(out)
template <typename T, int DIM>
__global__ __launch_bounds__(256) void zk_jacobi_dot3(GOp<T> o, const T* __restrict__ xi, const T* __restrict__ b,
                                                      T* __restrict__ xo, T w, const T* __restrict__ q,
                                                      double* __restrict__ part, ZTile zt, Bat bt) {
    BAT_SLOT(bt);
    xi += vo;
    b += vo;
    xo += vo;
    q += vo;
    part += (long long)slot * bt.ps;
    auto xv = [&](uint32_t l) { return xi[l]; };
    double sa = 0, sb = 0, sc = 0;
    for (int item = (int)xcd_block(blockIdx.x, gridDim.x); item < zt.items; item += gridDim.x) {
        int i, j, k0, k1;
        ztile_item(zt, item, i, j, k0, k1, o.n0, o.n1, o.n2);
        if (k0 >= k1) continue;
        uint32_t t = (uint32_t)i + o.s1 * (uint32_t)j + o.s2 * (uint32_t)k0;
        T vm = (DIM == 3 && k0 > 0) ? xi[t - o.s2] : (T)0, vc = xi[t];
        for (int k = k0; k < k1; ++k, t += o.s2) {
            const T vp = (DIM == 3 && k + 1 < o.n2) ? xi[t + o.s2] : (T)0;
            const T bv = b[t];
            T zv;
            if (interior3<T, DIM>(o, i, j, k)) {
                zv = vc + o.wdi * (bv - apply_in3<T, DIM>(o, t, vc, vm, vp, xv));
            } else {
                T D;
                const T ax = apply_op<T, false, DIM>(o, i, j, k, t, vc, xv, D);
                zv = vc + w * (bv - ax) / D;
            }
            xo[t] = zv;
            const double zd = (double)zv;
            sa += (double)bv * zd;
            sb += zd * (double)q[t];
            sc += zd;
            vm = vc;
            vc = vp;
        }
    }
    sa = gmg_block_sum(sa);
    sb = gmg_block_sum(sb);
    sc = gmg_block_sum(sc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sa;
        part[GMG_MAXB + blockIdx.x] = sb;
        part[2 * GMG_MAXB + blockIdx.x] = sc;
    }
}

// two sweeps from x = 0 (see k_jacobi2z); cells two or more cells away from
// every boundary take the straight-line path with x1 = (omega / D) b
template <typename T, int DIM>
__global__ __launch_bounds__(256) void zk_jacobi2z(GOp<T> o, const T* __restrict__ b, T* __restrict__ xo, T w,
                                                   ZTile zt, Bat bt) {
    BAT_SLOT(bt);
    b += vo;
    xo += vo;
    auto x1v = [&](uint32_t l) { return o.wdi * b[l]; };
    for (int item = (int)xcd_block(blockIdx.x, gridDim.x); item < zt.items; item += gridDim.x) {
        int i, j, k0, k1;
        ztile_item(zt, item, i, j, k0, k1, o.n0, o.n1, o.n2);
        if (k0 >= k1) continue;
        uint32_t t = (uint32_t)i + o.s1 * (uint32_t)j + o.s2 * (uint32_t)k0;
        T bm = (DIM == 3 && k0 > 0) ? b[t - o.s2] : (T)0, bc = b[t];
        const bool in2xy = i > 1 && i < o.n0 - 2 && j > 1 && j < o.n1 - 2;
        for (int k = k0; k < k1; ++k, t += o.s2) {
            const T bp = (DIM == 3 && k + 1 < o.n2) ? b[t + o.s2] : (T)0;
            if (in2xy && (DIM == 2 || (k > 1 && k < o.n2 - 2))) {
                const T xc = o.wdi * bc;
                const T ax = apply_in3<T, DIM>(o, t, xc, o.wdi * bm, o.wdi * bp, x1v);
                xo[t] = xc + o.wdi * (bc - ax);
            } else {
                const int c[3] = {i, j, k};
                const int n[3] = {o.n0, o.n1, o.n2};
                const uint32_t st[3] = {1u, o.s1, o.s2};
                auto x1 = [&](uint32_t l, int d, int cd) {
                    int qq[3] = {i, j, k};
                    qq[d] = cd;
                    return w * b[l] / diag_const<T, DIM>(o, qq[0], qq[1], qq[2]);
                };
                T D = o.sigma, acc = 0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    if (c[d] > 0) {
                        D += o.a[d];
                        acc += o.a[d] * x1(t - st[d], d, c[d] - 1);
                    } else if (o.per[d]) {
                        D += o.a[d];
                        acc += o.a[d] * x1(t + (uint32_t)(n[d] - 1) * st[d], d, n[d] - 1);
                    } else {
                        D += o.a[d] * o.bfac[d][0];
                    }
                    if (c[d] < n[d] - 1) {
                        D += o.a[d];
                        acc += o.a[d] * x1(t + st[d], d, c[d] + 1);
                    } else if (o.per[d]) {
                        D += o.a[d];
                        acc += o.a[d] * x1(t - (uint32_t)(n[d] - 1) * st[d], d, 0);
                    } else {
                        D += o.a[d] * o.bfac[d][1];
                    }
                }
                const T xc = w * bc / D;
                xo[t] = xc + w * (bc - (D * xc - acc)) / D;
            }
            bm = bc;
            bc = bp;
        }
    }
}

// p_new = (z - zmean) + beta p_old ; q = A p_new ; partial (p_new, q)
template <typename T, int DIM>
__global__ __launch_bounds__(256) void zk_spmv_p(GOp<T> o, const T* __restrict__ z, const T* __restrict__ pold,
                                                 T* __restrict__ pnew, T* __restrict__ q,
                                                 const double* __restrict__ s, double* __restrict__ part, ZTile zt,
                                                 Bat bt) {
    BAT_SLOT(bt);
    z += vo;
    pold += vo;
    pnew += vo;
    q += vo;
    s += (long long)slot * bt.ss;
    part += (long long)slot * bt.ps;
    const T beta = (T)s[S_BETA];
    const T zm = (T)s[S_ZMEAN];
    auto pv = [&](uint32_t l) { return (z[l] - zm) + beta * pold[l]; };
    double acc = 0;
    for (int item = (int)xcd_block(blockIdx.x, gridDim.x); item < zt.items; item += gridDim.x) {
        int i, j, k0, k1;
        ztile_item(zt, item, i, j, k0, k1, o.n0, o.n1, o.n2);
        if (k0 >= k1) continue;
        uint32_t t = (uint32_t)i + o.s1 * (uint32_t)j + o.s2 * (uint32_t)k0;
        T vm = (DIM == 3 && k0 > 0) ? pv(t - o.s2) : (T)0, vc = pv(t);
        for (int k = k0; k < k1; ++k, t += o.s2) {
            const T vp = (DIM == 3 && k + 1 < o.n2) ? pv(t + o.s2) : (T)0;
            T qc;
            if (interior3<T, DIM>(o, i, j, k)) {
                qc = apply_in3<T, DIM>(o, t, vc, vm, vp, pv);
            } else {
                T D;
                qc = apply_op<T, false, DIM>(o, i, j, k, t, vc, pv, D);
            }
            pnew[t] = vc;
            q[t] = qc;
            acc += (double)vc * (double)qc;
            vm = vc;
            vc = vp;
        }
    }
    acc = gmg_block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// r = b - A x (x may be null), partials of r^2
template <typename T, typename TO, int DIM>
__global__ __launch_bounds__(256) void zk_resid(GOp<T> o, const T* __restrict__ x, const T* __restrict__ b,
                                                TO* __restrict__ r, double* __restrict__ part, ZTile zt, Bat bt) {
    BAT_SLOT(bt);
    if (x) x += vo;
    b += vo;
    if (r) r += vo;
    if (part) part += (long long)slot * bt.ps;
    auto xv = [&](uint32_t l) { return x[l]; };
    double acc = 0;
    for (int item = (int)xcd_block(blockIdx.x, gridDim.x); item < zt.items; item += gridDim.x) {
        int i, j, k0, k1;
        ztile_item(zt, item, i, j, k0, k1, o.n0, o.n1, o.n2);
        if (k0 >= k1) continue;
        uint32_t t = (uint32_t)i + o.s1 * (uint32_t)j + o.s2 * (uint32_t)k0;
        T vm = (DIM == 3 && x && k0 > 0) ? x[t - o.s2] : (T)0, vc = x ? x[t] : (T)0;
        for (int k = k0; k < k1; ++k, t += o.s2) {
            const T vp = (DIM == 3 && x && k + 1 < o.n2) ? x[t + o.s2] : (T)0;
            T v = b[t];
            if (x) {
                if (interior3<T, DIM>(o, i, j, k)) {
                    v = v - apply_in3<T, DIM>(o, t, vc, vm, vp, xv);
                } else {
                    T D;
                    v = v - apply_op<T, false, DIM>(o, i, j, k, t, vc, xv, D);
                }
            }
            if (r) r[t] = (TO)v;
            acc += (double)v * (double)v;
            vm = vc;
            vc = vp;
        }
    }
    if (part) {
        acc = gmg_block_sum(acc);
        if (threadIdx.x == 0) part[blockIdx.x] = acc;
    }
}

#include "gmg_zp.hpp"

// coarse rhs = average over the 2^d children of the fine residual b - A x
template <typename T, bool VAR, int DIM>
__global__ __launch_bounds__(256) void k_restrict(GOp<T> o, const T* __restrict__ x, const T* __restrict__ b,
                                                  T* __restrict__ bc, int c0, int c1, int c2, FastDiv fc0,
                                                  FastDiv fc1, Bat bt) {
    BAT_SLOT(bt);
    const uint32_t NC = (uint32_t)c0 * c1 * c2;
    x += vo;
    b += vo;
    bc += (long long)slot * NC;
    const uint32_t t = cell_id();
    if (t >= NC) return;
    int I, J, K;
    cell_of(t, fc0, fc1, I, J, K);
    auto xv = [&](uint32_t l) { return x[l]; };
    T s = 0;
    const int nz = DIM == 3 ? 2 : 1;
    if (!VAR && I > 0 && I < c0 - 1 && J > 0 && J < c1 - 1 && (DIM == 2 || (K > 0 && K < c2 - 1))) {
        // every child has six (four) neighbours inside the fine grid: straight-line residuals
        for (int dk = 0; dk < nz; ++dk)
            for (int dj = 0; dj < 2; ++dj)
                for (int di = 0; di < 2; ++di) {
                    const uint32_t f = (uint32_t)(2 * I + di) + o.s1 * (uint32_t)(2 * J + dj) +
                                       o.s2 * (uint32_t)(DIM == 3 ? 2 * K + dk : 0);
                    T acc = o.a[0] * x[f - 1];
                    acc += o.a[0] * x[f + 1];
                    acc += o.a[1] * x[f - o.s1];
                    acc += o.a[1] * x[f + o.s1];
                    if (DIM == 3) {
                        acc += o.a[2] * x[f - o.s2];
                        acc += o.a[2] * x[f + o.s2];
                    }
                    s += b[f] - (o.dint * x[f] - acc);
                }
        bc[t] = s * (T)(DIM == 3 ? 0.125 : 0.25);
        return;
    }
    for (int dk = 0; dk < nz; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di) {
                const int i = 2 * I + di, j = 2 * J + dj, k = DIM == 3 ? 2 * K + dk : 0;
                const uint32_t f = (uint32_t)i + o.s1 * j + o.s2 * k;
                T D;
                s += b[f] - apply_op<T, VAR, DIM>(o, i, j, k, f, x[f], xv, D);
            }
    bc[t] = s * (T)(DIM == 3 ? 0.125 : 0.25);
}

// The same restriction with one thread per fine child (2^DIM consecutive lanes per
// coarse cell), the children's residuals summed through LDS in k_restrict's order. For
// small levels: there one thread per coarse cell leaves most CUs idle and runs its
// 2^DIM residuals in series (32^3 -> 16^3: 13 us).
template <typename T, bool VAR, int DIM>
__global__ __launch_bounds__(256) void k_restrict_ch(GOp<T> o, const T* __restrict__ x, const T* __restrict__ b,
                                                     T* __restrict__ bc, int c0, int c1, int c2, FastDiv fc0,
                                                     FastDiv fc1, Bat bt) {
    constexpr int NCH = 1 << DIM;
    __shared__ T sh[256];
    BAT_SLOT(bt);
    const uint32_t NC = (uint32_t)c0 * c1 * c2;
    x += vo;
    b += vo;
    bc += (long long)slot * NC;
    const uint32_t t = cell_id();
    const uint32_t c = t / NCH;
    const int ch = (int)(t % NCH);
    T r = 0;
    if (c < NC) {
        int I, J, K;
        cell_of(c, fc0, fc1, I, J, K);
        const int i = 2 * I + (ch & 1), j = 2 * J + ((ch >> 1) & 1), k = DIM == 3 ? 2 * K + (ch >> 2) : 0;
        const uint32_t f = (uint32_t)i + o.s1 * j + o.s2 * k;
        auto xv = [&](uint32_t l) { return x[l]; };
        T D;
        r = b[f] - apply_op<T, VAR, DIM>(o, i, j, k, f, x[f], xv, D);
    }
    sh[threadIdx.x] = r;
    __syncthreads();
    if (ch == 0 && c < NC) {
        T s = 0;
        for (int q = 0; q < NCH; ++q) s += sh[threadIdx.x + q];
        bc[c] = s * (T)(DIM == 3 ? 0.125 : 0.25);
    }
}

// x_f += P x_c (linear interpolation, coarse ghosts: periodic wrap or alpha * boundary value);
// o is the fine level's operator
template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_prolong(T* __restrict__ xf, const T* __restrict__ xc, int c0, int c1, int c2,
                                                 GOp<T> o, Bat bt) {
    BAT_SLOT(bt);
    const uint32_t N = o.s2 * (uint32_t)o.n2;
    xf += vo;
    xc += (long long)slot * ((long long)c0 * c1 * c2);
    const uint32_t t = cell_id();
    if (t >= N) return;
    int i, j, k;
    cell_of(t, o.f0, o.f1, i, j, k);
    const int f[3] = {i, j, k}, cn[3] = {c0, c1, c2};
    int par[3], oth[3];
    for (int d = 0; d < 3; ++d) {
        par[d] = f[d] >> 1;
        oth[d] = (f[d] & 1) ? par[d] + 1 : par[d] - 1;
    }
    const bool inner = i > 0 && i < o.n0 - 1 && j > 0 && j < o.n1 - 1 && (DIM == 2 || (k > 0 && k < o.n2 - 1));
    if (inner) {
        // every coarse cell used is inside the coarse grid: fixed (bi/tri)linear weights
        const uint32_t cs1 = (uint32_t)c0, cs2 = (uint32_t)c0 * (uint32_t)c1;
        const uint32_t p0 = (uint32_t)par[0] + cs1 * (uint32_t)par[1] + cs2 * (uint32_t)par[2];
        const int o0 = oth[0] - par[0], o1 = (oth[1] - par[1]) * (int)cs1;
        if (DIM == 2) {
            const T v = (T)0.5625 * xc[p0] + (T)0.1875 * (xc[p0 + o0] + xc[p0 + o1]) + (T)0.0625 * xc[p0 + o0 + o1];
            xf[t] += v;
        } else {
            const int o2 = (oth[2] - par[2]) * (int)cs2;
            const T v = (T)0.421875 * xc[p0] + (T)0.140625 * (xc[p0 + o0] + xc[p0 + o1] + xc[p0 + o2]) +
                        (T)0.046875 * (xc[p0 + o0 + o1] + xc[p0 + o0 + o2] + xc[p0 + o1 + o2]) +
                        (T)0.015625 * xc[p0 + o0 + o1 + o2];
            xf[t] += v;
        }
        return;
    }
    T s = 0;
    const int ncomb = 1 << DIM;
    for (int m = 0; m < ncomb; ++m) {
        T w = 1;
        T fac = 1;
        int q[3] = {0, 0, 0};
        for (int d = 0; d < DIM; ++d) {
            bool use_oth = (m >> d) & 1;
            w *= use_oth ? (T)0.25 : (T)0.75;
            int c = use_oth ? oth[d] : par[d];
            if (c < 0 || c >= cn[d]) {
                if (o.per[d]) c = c < 0 ? c + cn[d] : c - cn[d];
                else {
                    // ghost = alpha * boundary value; alpha = 1 - bfac
                    fac *= (T)1 - o.bfac[d][c < 0 ? 0 : 1];
                    c = par[d];
                }
            }
            q[d] = c;
        }
        s += w * fac * xc[(uint32_t)q[0] + (uint32_t)c0 * ((uint32_t)q[1] + (uint32_t)c1 * q[2])];
    }
    xf[t] += s;
}

// (P x_c) at fine cell (i, j, k): the (bi/tri)linear weights of k_prolong's
// general path, boundary ghosts folded (alpha = 1 - bfac), periodic wrap
template <typename T, int DIM>
__device__ __forceinline__ T prolong_at(const GOp<T>& o, const T* __restrict__ xc, const int cn[3], int i, int j,
                                        int k) {
    const int f[3] = {i, j, k};
    int par[3], oth[3];
    for (int d = 0; d < 3; ++d) {
        par[d] = f[d] >> 1;
        oth[d] = (f[d] & 1) ? par[d] + 1 : par[d] - 1;
    }
    T s = 0;
    for (int m = 0; m < (1 << DIM); ++m) {
        T wgt = 1, fac = 1;
        int q[3] = {0, 0, 0};
        for (int d = 0; d < DIM; ++d) {
            const bool use_oth = (m >> d) & 1;
            wgt *= use_oth ? (T)0.25 : (T)0.75;
            int c = use_oth ? oth[d] : par[d];
            if (c < 0 || c >= cn[d]) {
                if (o.per[d]) c = c < 0 ? c + cn[d] : c - cn[d];
                else {
                    fac *= (T)1 - o.bfac[d][c < 0 ? 0 : 1];
                    c = par[d];
                }
            }
            q[d] = c;
        }
        s += wgt * fac * xc[(uint32_t)q[0] + (uint32_t)cn[0] * ((uint32_t)q[1] + (uint32_t)cn[1] * q[2])];
    }
    return s;
}

// first post-smoothing sweep fused with the prolongation: with y = x + P x_c (never
// stored), xo = y + w (b - A y) / D; the stencil's neighbours of y are formed on
// the fly. Replaces k_prolong + one Jacobi launch on 2-D levels.
template <typename T, bool VAR, int DIM>
__global__ __launch_bounds__(256) void k_pjacobi(GOp<T> o, const T* __restrict__ xi, const T* __restrict__ xc, int c0,
                                                 int c1, int c2, const T* __restrict__ b, T* __restrict__ xo, T w,
                                                 Bat bt) {
    BAT_SLOT(bt);
    const uint32_t N = o.s2 * (uint32_t)o.n2;
    xi += vo;
    b += vo;
    xo += vo;
    xc += (long long)slot * ((long long)c0 * c1 * c2);
    const uint32_t t = cell_id();
    if (t >= N) return;
    int i, j, k;
    cell_of(t, o.f0, o.f1, i, j, k);
    const int cn[3] = {c0, c1, c2};
    auto yv = [&](uint32_t l) {
        int a, e, c;
        cell_of(l, o.f0, o.f1, a, e, c);
        return xi[l] + prolong_at<T, DIM>(o, xc, cn, a, e, c);
    };
    const T yc = xi[t] + prolong_at<T, DIM>(o, xc, cn, i, j, k);
    T D;
    const T ay = apply_op<T, VAR, DIM>(o, i, j, k, t, yc, yv, D);
    xo[t] = yc + w * (b[t] - ay) / D;
}

// x = Cinv b in FP64 for the coarsest level: 16 lanes per row, each summing every 16th
// column, the partial sums combined in a fixed order, so a row's reads are coalesced
// and every thread of the block works (bv(c): b at column c)
template <typename T, typename BV>
__device__ __forceinline__ void dense_apply(const double* __restrict__ cinv, int n, BV bv, T* x) {
    const int lane = threadIdx.x & 15, groups = blockDim.x >> 4;
    for (int row = threadIdx.x >> 4; row < n; row += groups) {  // the 16 lanes of a group share the row
        const double* a = cinv + (long long)row * n;
        double s = 0;
        for (int c = lane; c < n; c += 16) s += a[c] * bv(c);
        for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
        if (lane == 0) x[row] = (T)s;
    }
}

// coarsest level: x = Cinv b (FP64 inverse), one block per slot
template <typename T>
__global__ __launch_bounds__(256) void k_dense(const double* __restrict__ cinv, int n, const T* __restrict__ b,
                                               T* __restrict__ x, Bat bt) {
    __shared__ double sb[1024];
    BAT_SLOT(bt);
    b += (long long)slot * n;
    x += (long long)slot * n;
    for (int q = threadIdx.x; q < n; q += blockDim.x) sb[q] = (double)b[q];
    __syncthreads();
    dense_apply<T>(cinv, n, [&](int c) { return sb[c]; }, x);
}


// ---- 2-D levels with constant coefficients, V(2,2): a level in two launches.
// down: x1 = w b / D, x2 = x1 + w (b - A x1) / D, r = b - A x2, coarse b = the mean of r
//       over each 2 x 2 block; x2 is stored.
// up:   x3 = x2 + P x_c, x4 = x3 + w (b - A x3) / D, x5 = x4 + w (b - A x4) / D -> out,
//       on the finest level with the partials of (b, x5), (x5, q), sum(x5).
// A workgroup owns a TS x TS tile (32 x 32 with 256 threads on levels of >= 2^21
// cells, 16 x 16 with 256 threads down to 2^16, 8 x 8 with one wave below: the small
// levels are latency-bound, so fewer cells per thread and more workgroups pay); the
// intermediate vectors live in LDS on the tile plus a halo of 2 (x1, x3), 1 (x2, x4)
// and are recomputed by neighbouring tiles instead of going through memory, so a
// level costs two launches instead of four or five and each fine vector crosses HBM
// once per phase. Halo positions of a periodic dim hold the wrapped cells; those
// outside a non-periodic boundary are never read (the boundary cell folds its ghost
// into D, as apply_op does). Blocks stride over the tiles (a fixed grid: the partial
// sums are reproducible).

struct F2Tile {
    int tx, ty, tiles;
};

// global coordinate g of a position, wrapped into [0, n) on a periodic dim; false when
// it lies outside a non-periodic dim
__device__ __forceinline__ bool f2_map(int g, int n, int per, int& w) {
    if (g >= 0 && g < n) {
        w = g;
        return true;
    }
    if (!per) return false;
    w = ((g % n) + n) % n;
    return true;
}

// D and (A v)_c at cell (i, j) of a constant-coefficient 2-D level, the neighbours of v
// read at LDS offsets +-1 / +-W from c (apply_op's order and boundary folding)
template <typename T>
__device__ __forceinline__ T f2_apply(const GOp<T>& o, int i, int j, const T* v, int c, int W, T& D) {
    T acc = 0;
    D = o.sigma;
    const int cc[2] = {i, j}, nn[2] = {o.n0, o.n1}, st[2] = {1, W};
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        if (cc[d] > 0 || o.per[d]) {
            D += o.a[d];
            acc += o.a[d] * v[c - st[d]];
        } else {
            D += o.a[d] * o.bfac[d][0];
        }
        if (cc[d] < nn[d] - 1 || o.per[d]) {
            D += o.a[d];
            acc += o.a[d] * v[c + st[d]];
        } else {
            D += o.a[d] * o.bfac[d][1];
        }
    }
    return D * v[c] - acc;
}

// w r / D at cell (i, j): off the boundary as o.wdi r (the interior scaling of the
// per-level kernels: no division on the cells of a wave that touches no boundary)
template <typename T>
__device__ __forceinline__ T f2_corr(const GOp<T>& o, int i, int j, T r, T D, T w) {
    if (i > 0 && i < o.n0 - 1 && j > 0 && j < o.n1 - 1) return o.wdi * r;
    return w * r / D;
}

// A tile whose region with halo 2 lies inside the level and whose cells are all off
// the boundary: no wrapped or missing neighbour, D = dint everywhere. Such tiles take
// the straight-line path (f2_apply_in, prolong_in, o.wdi), which gives the general
// path's bits for those cells; tiles along the boundary take the general path.
template <typename T, int TS>
__device__ __forceinline__ bool f2_inner(const GOp<T>& o, int x0, int y0) {
    return x0 >= 3 && y0 >= 3 && x0 + TS + 3 <= o.n0 && y0 + TS + 3 <= o.n1;
}

// (A v)_c of an interior cell (f2_apply's order, D = dint)
template <typename T>
__device__ __forceinline__ T f2_apply_in(const GOp<T>& o, const T* v, int c, int W) {
    T acc = 0;
    acc += o.a[0] * v[c - 1];
    acc += o.a[0] * v[c + 1];
    acc += o.a[1] * v[c - W];
    acc += o.a[1] * v[c + W];
    return o.dint * v[c] - acc;
}

// (P x_c) at fine cell (i, j) whose four coarse cells lie inside the coarse level
// (prolong_at's weights and order, no boundary factor)
template <typename T>
__device__ __forceinline__ T prolong_in(const T* __restrict__ xc, int c0, int i, int j) {
    const int pi = i >> 1, pj = j >> 1, oi = (i & 1) ? pi + 1 : pi - 1, oj = (j & 1) ? pj + 1 : pj - 1;
    T s = 0;
    s += (T)0.5625 * xc[(uint32_t)pi + (uint32_t)c0 * (uint32_t)pj];
    s += (T)0.1875 * xc[(uint32_t)oi + (uint32_t)c0 * (uint32_t)pj];
    s += (T)0.1875 * xc[(uint32_t)pi + (uint32_t)c0 * (uint32_t)oj];
    s += (T)0.0625 * xc[(uint32_t)oi + (uint32_t)c0 * (uint32_t)oj];
    return s;
}

// one TS x TS tile of k2_down (IN: f2_inner holds for the tile)
template <typename T, int TS, bool IN>
__device__ __forceinline__ void k2_down_tile(const GOp<T>& o, const T* __restrict__ b, T* __restrict__ x2o,
                                             T* __restrict__ bc, int c0, int c1, T w, int x0, int y0, T* sb, T* s1,
                                             T* s2) {
    constexpr int W2 = TS + 4, W1 = TS + 2;
    // b and x1 = w b / D on the tile + 2
    for (int p = threadIdx.x; p < W2 * W2; p += blockDim.x) {
        int i, j;
        bool ok = true;
        if (IN) {
            i = x0 - 2 + p % W2;
            j = y0 - 2 + p / W2;
        } else {
            ok = f2_map(x0 - 2 + p % W2, o.n0, o.per[0], i) && f2_map(y0 - 2 + p / W2, o.n1, o.per[1], j);
        }
        T bv = 0, x1 = 0;
        if (ok) {
            bv = b[(uint32_t)i + o.s1 * (uint32_t)j];
            x1 = IN ? o.wdi * bv : f2_corr<T>(o, i, j, bv, diag_const<T, 2>(o, i, j, 0), w);
        }
        sb[p] = bv;
        s1[p] = x1;
    }
    __syncthreads();
    // x2 on the tile + 1
    for (int p = threadIdx.x; p < W1 * W1; p += blockDim.x) {
        const int li = p % W1, lj = p / W1, c = (lj + 1) * W2 + li + 1;
        T x2 = 0;
        if (IN) {
            x2 = s1[c] + o.wdi * (sb[c] - f2_apply_in<T>(o, s1, c, W2));
        } else {
            int i, j;
            if (f2_map(x0 - 1 + li, o.n0, o.per[0], i) && f2_map(y0 - 1 + lj, o.n1, o.per[1], j)) {
                T D;
                const T ax = f2_apply<T>(o, i, j, s1, c, W2, D);
                x2 = s1[c] + f2_corr<T>(o, i, j, sb[c] - ax, D, w);
            }
        }
        s2[p] = x2;
    }
    __syncthreads();
    // thread q: coarse cell q of the tile; r = b - A x2 on its four children (x2
    // stored), coarse b = the children's r summed as k_restrict sums them
    if (threadIdx.x < (TS / 2) * (TS / 2)) {
        const int q = threadIdx.x, ci = q % (TS / 2), cj = q / (TS / 2);
        T s = 0;
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const int li = 2 * ci + di, lj = 2 * cj + dj, i = x0 + li, j = y0 + lj;
                if (IN || (i < o.n0 && j < o.n1)) {
                    const int c = (lj + 1) * W1 + li + 1;
                    T ax;
                    if (IN) {
                        ax = f2_apply_in<T>(o, s2, c, W1);
                    } else {
                        T D;
                        ax = f2_apply<T>(o, i, j, s2, c, W1, D);
                    }
                    s += sb[(lj + 2) * W2 + li + 2] - ax;
                    x2o[(uint32_t)i + o.s1 * (uint32_t)j] = s2[c];
                } else {
                    s += (T)0;
                }
            }
        const int I = x0 / 2 + ci, J = y0 / 2 + cj;
        if (IN || (I < c0 && J < c1)) bc[(uint32_t)I + (uint32_t)c0 * (uint32_t)J] = s * (T)0.25;
    }
    __syncthreads();  // the LDS tiles are reused by the next tile
}

template <typename T, int TS, int NT>
__global__ __launch_bounds__(NT) void k2_down(GOp<T> o, const T* __restrict__ b, T* __restrict__ x2o,
                                              T* __restrict__ bc, int c0, int c1, T w, F2Tile ft, Bat bt) {
    static_assert((TS / 2) * (TS / 2) <= NT, "one coarse cell of the tile per thread");
    constexpr int W2 = TS + 4, W1 = TS + 2;
    __shared__ T sb[W2 * W2], s1[W2 * W2], s2[W1 * W1];
    BAT_SLOT(bt);
    b += vo;
    x2o += vo;
    bc += (long long)slot * ((long long)c0 * c1);
    for (int tile = (int)xcd_block(blockIdx.x, gridDim.x); tile < ft.tiles; tile += gridDim.x) {
        const int x0 = (tile % ft.tx) * TS, y0 = (tile / ft.tx) * TS;
        if (f2_inner<T, TS>(o, x0, y0)) k2_down_tile<T, TS, true>(o, b, x2o, bc, c0, c1, w, x0, y0, sb, s1, s2);
        else k2_down_tile<T, TS, false>(o, b, x2o, bc, c0, c1, w, x0, y0, sb, s1, s2);
    }
}

// one TS x TS tile of k2_up (IN: f2_inner holds for the tile)
template <typename T, int TS, bool DOTS, bool IN>
__device__ __forceinline__ void k2_up_tile(const GOp<T>& o, const T* __restrict__ x2, const T* __restrict__ xc,
                                           int c0, int c1, const T* __restrict__ b, T* __restrict__ out, T w,
                                           const T* __restrict__ q, int x0, int y0, T* s3, T* sb, T* s4, double& d0,
                                           double& d1, double& d2) {
    constexpr int W2 = TS + 4, W1 = TS + 2;
    const int cn[3] = {c0, c1, 1};
    // x3 = x2 + P x_c on the tile + 2
    for (int p = threadIdx.x; p < W2 * W2; p += blockDim.x) {
        int i, j;
        T x3 = 0;
        if (IN) {
            i = x0 - 2 + p % W2;
            j = y0 - 2 + p / W2;
            x3 = x2[(uint32_t)i + o.s1 * (uint32_t)j] + prolong_in<T>(xc, c0, i, j);
        } else if (f2_map(x0 - 2 + p % W2, o.n0, o.per[0], i) && f2_map(y0 - 2 + p / W2, o.n1, o.per[1], j)) {
            x3 = x2[(uint32_t)i + o.s1 * (uint32_t)j] + prolong_at<T, 2>(o, xc, cn, i, j, 0);
        }
        s3[p] = x3;
    }
    // b on the tile + 1
    for (int p = threadIdx.x; p < W1 * W1; p += blockDim.x) {
        int i, j;
        bool ok = true;
        if (IN) {
            i = x0 - 1 + p % W1;
            j = y0 - 1 + p / W1;
        } else {
            ok = f2_map(x0 - 1 + p % W1, o.n0, o.per[0], i) && f2_map(y0 - 1 + p / W1, o.n1, o.per[1], j);
        }
        sb[p] = ok ? b[(uint32_t)i + o.s1 * (uint32_t)j] : (T)0;
    }
    __syncthreads();
    // x4 on the tile + 1
    for (int p = threadIdx.x; p < W1 * W1; p += blockDim.x) {
        const int li = p % W1, lj = p / W1, c = (lj + 1) * W2 + li + 1;
        T x4 = 0;
        if (IN) {
            x4 = s3[c] + o.wdi * (sb[p] - f2_apply_in<T>(o, s3, c, W2));
        } else {
            int i, j;
            if (f2_map(x0 - 1 + li, o.n0, o.per[0], i) && f2_map(y0 - 1 + lj, o.n1, o.per[1], j)) {
                T D;
                const T ax = f2_apply<T>(o, i, j, s3, c, W2, D);
                x4 = s3[c] + f2_corr<T>(o, i, j, sb[p] - ax, D, w);
            }
        }
        s4[p] = x4;
    }
    __syncthreads();
    // x5 on the tile -> out (+ the dot products)
    for (int p = threadIdx.x; p < TS * TS; p += blockDim.x) {
        const int li = p % TS, lj = p / TS, i = x0 + li, j = y0 + lj;
        if (!IN && (i >= o.n0 || j >= o.n1)) continue;
        const int c = (lj + 1) * W1 + li + 1;
        const T bv = sb[c];
        T x5;
        if (IN) {
            x5 = s4[c] + o.wdi * (bv - f2_apply_in<T>(o, s4, c, W1));
        } else {
            T D;
            const T ax = f2_apply<T>(o, i, j, s4, c, W1, D);
            x5 = s4[c] + f2_corr<T>(o, i, j, bv - ax, D, w);
        }
        const uint32_t t = (uint32_t)i + o.s1 * (uint32_t)j;
        out[t] = x5;
        if (DOTS) {
            const double zv = (double)x5;
            d0 += (double)bv * zv;
            d1 += zv * (double)q[t];
            d2 += zv;
        }
    }
    __syncthreads();
}

template <typename T, int TS, int NT, bool DOTS>
__global__ __launch_bounds__(NT) void k2_up(GOp<T> o, const T* __restrict__ x2, const T* __restrict__ xc, int c0,
                                            int c1, const T* __restrict__ b, T* __restrict__ out, T w,
                                            const T* __restrict__ q, double* __restrict__ part, F2Tile ft, Bat bt) {
    constexpr int W2 = TS + 4, W1 = TS + 2;
    __shared__ T s3[W2 * W2], sb[W1 * W1], s4[W1 * W1];
    BAT_SLOT(bt);
    x2 += vo;
    b += vo;
    out += vo;
    xc += (long long)slot * ((long long)c0 * c1);
    if (DOTS) {
        q += vo;
        part += (long long)slot * bt.ps;
    }
    double d0 = 0, d1 = 0, d2 = 0;
    for (int tile = (int)xcd_block(blockIdx.x, gridDim.x); tile < ft.tiles; tile += gridDim.x) {
        const int x0 = (tile % ft.tx) * TS, y0 = (tile / ft.tx) * TS;
        if (f2_inner<T, TS>(o, x0, y0))
            k2_up_tile<T, TS, DOTS, true>(o, x2, xc, c0, c1, b, out, w, q, x0, y0, s3, sb, s4, d0, d1, d2);
        else
            k2_up_tile<T, TS, DOTS, false>(o, x2, xc, c0, c1, b, out, w, q, x0, y0, s3, sb, s4, d0, d1, d2);
    }
    if (DOTS) {
        d0 = gmg_block_sum(d0);
        d1 = gmg_block_sum(d1);
        d2 = gmg_block_sum(d2);
        if (threadIdx.x == 0) {
            part[blockIdx.x] = d0;
            part[GMG_MAXB + blockIdx.x] = d1;
            part[2 * GMG_MAXB + blockIdx.x] = d2;
        }
    }
}

// ---- coarse tail: every level of at most GMG_TAIL_N cells in one workgroup per slot.
// A V-cycle's coarse levels are tiny and launch-bound (a handful of cells per
// CU); here one 1024-thread workgroup runs their smoothing sweeps, residual
// restrictions, the dense coarsest solve, prolongations and post-smoothing in
// sequence with barriers instead of launches, on copies of the levels' vectors
// in LDS (the first tail level's b comes in from HBM, its x goes back). Per cell
// the arithmetic is that of the per-level kernels' general path (apply_op), and
// the buffers alternate as in ph_vcycle, so the last sweep of every level lands
// in its x.
constexpr int GMG_TAIL_N = 1024;      // cells of the first tail level at most
constexpr int GMG_TAIL_CELLS = 1400;  // cells of all tail levels together at most (LDS: 3 vectors each)
constexpr int GMG_TAIL_MAXL = 12;
constexpr int GMG_TAIL_NT = 1024;

template <typename T>
struct TailLevel {
    GOp<T> o;
    int N;
    int ofs;  // LDS offset of x (b follows at ofs + N, t at ofs + 2 N)
};

template <typename T>
struct TailArgs {
    TailLevel<T> lv[GMG_TAIL_MAXL];
    int nl;  // levels in the tail; the last is the coarsest (dense inverse)
    const T* b0;  // the first tail level's b and x in HBM (slot stride lv[0].N)
    T* x0;
    const double* cinv;
    int cN;
    T w;
    int pre, post;
};

// one weighted Jacobi sweep over a level (zero: from x = 0)
template <typename T, bool VAR, int DIM>
__device__ void tail_sweep(const GOp<T>& o, int N, const T* xi, const T* b, T* xo, T w, bool zero) {
    for (int t = threadIdx.x; t < N; t += blockDim.x) {
        int i, j, k;
        cell_of((uint32_t)t, o.f0, o.f1, i, j, k);
        T D;
        // constant mode off the boundary: the interior scaling o.wdi, as the per-level kernels
        const bool inner = !VAR && interior3<T, DIM>(o, i, j, k);
        if (zero) {
            auto zv = [&](uint32_t) { return (T)0; };
            apply_op<T, VAR, DIM>(o, i, j, k, (uint32_t)t, (T)0, zv, D);
            xo[t] = inner ? o.wdi * b[t] : w * b[t] / D;
        } else {
            auto xv = [&](uint32_t l) { return xi[l]; };
            const T xc = xi[t];
            const T ax = apply_op<T, VAR, DIM>(o, i, j, k, (uint32_t)t, xc, xv, D);
            const T r = b[t] - ax;
            xo[t] = xc + (inner ? o.wdi * r : w * r / D);
        }
    }
    __syncthreads();
}

// coarse b = average over the 2^d children of the fine residual b - A x
template <typename T, bool VAR, int DIM>
__device__ void tail_restrict(const GOp<T>& o, const T* x, const T* b, T* bc, const GOp<T>& oc) {
    const int NC = oc.n0 * oc.n1 * oc.n2;
    auto xv = [&](uint32_t l) { return x[l]; };
    for (int t = threadIdx.x; t < NC; t += blockDim.x) {
        int I, J, K;
        cell_of((uint32_t)t, oc.f0, oc.f1, I, J, K);
        T s = 0;
        const int nz = DIM == 3 ? 2 : 1;
        for (int dk = 0; dk < nz; ++dk)
            for (int dj = 0; dj < 2; ++dj)
                for (int di = 0; di < 2; ++di) {
                    const int i = 2 * I + di, j = 2 * J + dj, k = DIM == 3 ? 2 * K + dk : 0;
                    const uint32_t f = (uint32_t)i + o.s1 * j + o.s2 * k;
                    T D;
                    s += b[f] - apply_op<T, VAR, DIM>(o, i, j, k, f, x[f], xv, D);
                }
        bc[t] = s * (T)(DIM == 3 ? 0.125 : 0.25);
    }
    __syncthreads();
}

// x_f += P x_c (the general path of k_prolong)
template <typename T, int DIM>
__device__ void tail_prolong(const GOp<T>& o, T* xf, const T* xc, const GOp<T>& oc) {
    const int N = o.n0 * o.n1 * o.n2;
    const int cn[3] = {oc.n0, oc.n1, oc.n2};
    for (int t = threadIdx.x; t < N; t += blockDim.x) {
        int i, j, k;
        cell_of((uint32_t)t, o.f0, o.f1, i, j, k);
        const int f[3] = {i, j, k};
        int par[3], oth[3];
        for (int d = 0; d < 3; ++d) {
            par[d] = f[d] >> 1;
            oth[d] = (f[d] & 1) ? par[d] + 1 : par[d] - 1;
        }
        T s = 0;
        for (int m = 0; m < (1 << DIM); ++m) {
            T wgt = 1, fac = 1;
            int q[3] = {0, 0, 0};
            for (int d = 0; d < DIM; ++d) {
                const bool use_oth = (m >> d) & 1;
                wgt *= use_oth ? (T)0.25 : (T)0.75;
                int c = use_oth ? oth[d] : par[d];
                if (c < 0 || c >= cn[d]) {
                    if (o.per[d]) c = c < 0 ? c + cn[d] : c - cn[d];
                    else {
                        fac *= (T)1 - o.bfac[d][c < 0 ? 0 : 1];
                        c = par[d];
                    }
                }
                q[d] = c;
            }
            s += wgt * fac * xc[(uint32_t)q[0] + (uint32_t)cn[0] * ((uint32_t)q[1] + (uint32_t)cn[1] * q[2])];
        }
        xf[t] += s;
    }
    __syncthreads();
}

template <typename T, bool VAR, int DIM>
__global__ __launch_bounds__(GMG_TAIL_NT) void k_vtail(TailArgs<T> ta, Bat bt) {
    __shared__ T lds[3 * GMG_TAIL_CELLS];
    BAT_SLOT(bt);
    const int nsw = ta.pre + ta.post;  // sweeps per level; sweep m writes x when (nsw - 1 - m) is even
    auto X = [&](const TailLevel<T>& L) { return lds + L.ofs; };
    auto Bv = [&](const TailLevel<T>& L) { return lds + L.ofs + L.N; };
    auto buf = [&](const TailLevel<T>& L, int m) -> T* { return lds + L.ofs + (((nsw - 1 - m) & 1) ? 2 * L.N : 0); };
    {
        const TailLevel<T>& L = ta.lv[0];
        const T* b0 = ta.b0 + (size_t)slot * L.N;
        for (int t = threadIdx.x; t < L.N; t += blockDim.x) Bv(L)[t] = b0[t];
        __syncthreads();
    }
    for (int l = 0; l + 1 < ta.nl; ++l) {  // down: pre-smoothing from zero, restriction
        const TailLevel<T>& L = ta.lv[l];
        for (int m = 0; m < ta.pre; ++m)
            tail_sweep<T, VAR, DIM>(L.o, L.N, m ? buf(L, m - 1) : nullptr, Bv(L), buf(L, m), ta.w, m == 0);
        const TailLevel<T>& C = ta.lv[l + 1];
        tail_restrict<T, VAR, DIM>(L.o, buf(L, ta.pre - 1), Bv(L), Bv(C), C.o);
    }
    {  // coarsest: x = Cinv b in FP64
        const TailLevel<T>& L = ta.lv[ta.nl - 1];
        const T* bl = Bv(L);
        dense_apply<T>(ta.cinv, ta.cN, [&](int c) { return (double)bl[c]; }, X(L));
        __syncthreads();
    }
    for (int l = ta.nl - 2; l >= 0; --l) {  // up: prolongation, post-smoothing
        const TailLevel<T>& L = ta.lv[l];
        const TailLevel<T>& C = ta.lv[l + 1];
        tail_prolong<T, DIM>(L.o, buf(L, ta.pre - 1), X(C), C.o);
        for (int m = ta.pre; m < nsw; ++m)
            tail_sweep<T, VAR, DIM>(L.o, L.N, buf(L, m - 1), Bv(L), buf(L, m), ta.w, false);
    }
    {
        const TailLevel<T>& L = ta.lv[0];
        T* x0 = ta.x0 + (size_t)slot * L.N;
        for (int t = threadIdx.x; t < L.N; t += blockDim.x) x0[t] = X(L)[t];
    }
}

__global__ __launch_bounds__(1024) void k_finish(const double* __restrict__ part, int nb, int mode,
                                                 double* __restrict__ s, double invN,
                         int singular, Bat bt, int* __restrict__ mk_it, int* __restrict__ mk_rep, int maxit) {
    BAT_SLOT(bt);
    if (mode == FIN_RR_DECIDE && !mk_it[slot]) {  // launched for every slot: a stopped one is not updated
        if (threadIdx.x == 0) mk_rep[slot] = 0;
        return;
    }
    part += (long long)slot * bt.ps;
    s += (long long)slot * bt.ss;
    // up to 16384 partials per value: four accumulators per value per thread, the
    // loads of every value issued together (one latency round per 4 x blockDim
    // partials); a fixed order, so the sums are reproducible
    const int nv = (mode == FIN_DOT3 || mode == FIN_DOT3_FIRST) ? 3 : 1;
    double v[3] = {0, 0, 0};
    {
        double a[3][4] = {};
        const int step = blockDim.x;
        int t = threadIdx.x;
        for (; t + 3 * step < nb; t += 4 * step)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                if (q < nv) {
                    const double* pq = part + q * GMG_MAXB;
                    a[q][0] += pq[t];
                    a[q][1] += pq[t + step];
                    a[q][2] += pq[t + 2 * step];
                    a[q][3] += pq[t + 3 * step];
                }
        for (; t < nb; t += step)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                if (q < nv) a[q][0] += part[q * GMG_MAXB + t];
#pragma unroll
        for (int q = 0; q < 3; ++q) v[q] = (a[q][0] + a[q][1]) + (a[q][2] + a[q][3]);
    }
    for (int q = 0; q < nv; ++q) v[q] = gmg_block_sum(v[q]);
    if (threadIdx.x != 0) return;
    switch (mode) {
        case FIN_PQ:
            s[S_PQ] = v[0];
            s[S_ALPHA] = v[0] != 0.0 ? s[S_RZ] / v[0] : 0.0;
            break;
        case FIN_RR: s[S_RR] = v[0]; break;
        case FIN_BB: s[S_BB] = v[0]; break;
        case FIN_RR0: s[S_RR0] = v[0]; break;
        case FIN_REP:
            s[S_RR] = v[0];
            s[S_RREF] = v[0];
            break;
        case FIN_RR_DECIDE: {
            // count the iteration; fold e into x and recompute the true residual when the
            // recurrence residual fell by delta2 since the last update, looks converged, or
            // the cap is reached (!(a >= b) also catches NaN)
            const double rr = v[0];
            s[S_RR] = rr;
            const double its = s[S_ITS] + 1.0;
            s[S_ITS] = its;
            const int rep = !(rr >= s[S_DLT2] * s[S_RREF]) || rr <= s[S_TGT] || its >= (double)maxit;
            mk_rep[slot] = rep;
            if (rep) {
                s[S_NREP] += 1.0;
                s[S_RREC] = rr;
            }
        } break;
        case FIN_REP_DONE: {
            // after the update's true residual: stop a slot that meets tol (or hit the cap /
            // broke down); else adapt the update threshold to the drift of the recurrence
            const double rr = v[0];
            s[S_RR] = rr;
            s[S_RREF] = rr;
            if (rr <= s[S_TGT] || s[S_ITS] >= (double)maxit || !isfinite(rr)) {
                mk_it[slot] = 0;
                s[S_DONE] = 1.0;
                break;
            }
            const double d2 = s[S_DLT2];
            if (d2 > 0.0) {
                const double drift = rr / s[S_RREC];
                if (drift < 4.0) s[S_DLT2] = fmax(d2 * 1e-2, 1e-12);
                else if (drift > 100.0) s[S_DLT2] = fmin(d2 * 1e2, 1e-2);
            }
        } break;
        case FIN_MEAN: s[S_MEAN] = v[0]; break;
        case FIN_DOT3_FIRST:
            s[S_RZ] = v[0];
            s[S_ZMEAN] = singular ? v[2] * invN : 0.0;
            s[S_BETA] = 0.0;
            break;
        case FIN_DOT3: {
            double zq = v[1];
            s[S_ZQ] = zq;
            s[S_BETA] = s[S_RZ] != 0.0 ? -s[S_ALPHA] * zq / s[S_RZ] : 0.0;
            s[S_RZ] = v[0];
            s[S_ZMEAN] = singular ? v[2] * invN : 0.0;
        } break;
    }
}

// p_new = (z - zmean) + beta p_old ; q = A p_new ; partial (p_new, q)
template <typename T, bool VAR, int DIM>
__global__ __launch_bounds__(256) void k_spmv_p(GOp<T> o, const T* __restrict__ z, const T* __restrict__ pold,
                                                T* __restrict__ pnew, T* __restrict__ q,
                                                const double* __restrict__ s, double* __restrict__ part, Bat bt) {
    BAT_SLOT(bt);
    const uint32_t N = o.s2 * (uint32_t)o.n2;
    z += vo;
    pold += vo;
    pnew += vo;
    q += vo;
    s += (long long)slot * bt.ss;
    part += (long long)slot * bt.ps;
    const T beta = (T)s[S_BETA];
    const T zm = (T)s[S_ZMEAN];
    double acc = 0;
    uint32_t tb, te;
    block_range(N, tb, te);
    for (uint32_t t = tb + threadIdx.x; t < te; t += blockDim.x) {
        int i, j, k;
        cell_of(t, o.f0, o.f1, i, j, k);
        auto pv = [&](uint32_t l) { return (z[l] - zm) + beta * pold[l]; };
        T pc = pv(t);
        T D;
        T qc = apply_op<T, VAR, DIM>(o, i, j, k, t, pc, pv, D);
        pnew[t] = pc;
        q[t] = qc;
        acc += (double)pc * (double)qc;
    }
    acc = gmg_block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// x += alpha p ; r -= alpha q ; partial (r, r)
template <typename T>
__global__ __launch_bounds__(256) void k_axpy2(long long N, T* __restrict__ x, const T* __restrict__ p,
                                               T* __restrict__ r, const T* __restrict__ q,
                                               const double* __restrict__ s, double* __restrict__ part, Bat bt) {
    BAT_SLOT(bt);
    x += vo;
    p += vo;
    r += vo;
    q += vo;
    s += (long long)slot * bt.ss;
    part += (long long)slot * bt.ps;
    const T alpha = (T)s[S_ALPHA];
    double acc = 0;
    uint32_t tb, te;
    block_range((uint32_t)N, tb, te);
    for (uint32_t t = tb + threadIdx.x; t < te; t += blockDim.x) {
        x[t] += alpha * p[t];
        T rv = r[t] - alpha * q[t];
        r[t] = rv;
        acc += (double)rv * (double)rv;
    }
    acc = gmg_block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// partials of (r, z), (z, q), sum(z)
template <typename T>
__global__ __launch_bounds__(256) void k_dot3(long long N, const T* __restrict__ r, const T* __restrict__ z,
                                              const T* __restrict__ q, double* __restrict__ part, Bat bt) {
    BAT_SLOT(bt);
    r += vo;
    z += vo;
    q += vo;
    part += (long long)slot * bt.ps;
    double a = 0, b = 0, c = 0;
    uint32_t tb, te;
    block_range((uint32_t)N, tb, te);
    for (uint32_t t = tb + threadIdx.x; t < te; t += blockDim.x) {
        double zv = (double)z[t];
        a += (double)r[t] * zv;
        b += zv * (double)q[t];
        c += zv;
    }
    a = gmg_block_sum(a);
    b = gmg_block_sum(b);
    c = gmg_block_sum(c);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[GMG_MAXB + blockIdx.x] = b;
        part[2 * GMG_MAXB + blockIdx.x] = c;
    }
}

// x (FP64) += e (FP32 or FP64)
template <typename TE>
__global__ void k_add_to64(long long N, double* __restrict__ x, const TE* __restrict__ e, Bat bt) {
    BAT_SLOT(bt);
    x += vo;
    e += vo;
    const uint32_t t = cell_id();
    if (t < (uint32_t)N) x[t] += (double)e[t];
}

template <typename T>
__global__ void k_fill(long long N, T* __restrict__ x, T v, Bat bt) {
    BAT_SLOT(bt);
    x += vo;
    const uint32_t t = cell_id();
    if (t < (uint32_t)N) x[t] = v;
}

// reliable-update control (one thread per slot; see ph_rel_start / ph_rel_iter)
__global__ void k_seti(int* __restrict__ p, int n, int v) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// slots in m_new, after ||b||^2 -> S_BB and the first true residual -> S_RR0
__global__ void k_rel_init(int B, const int* __restrict__ m_new, int* __restrict__ m_it, int* __restrict__ m_start,
                           double* __restrict__ scal, double tol2, double delta2, int x_zero) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    if (!m_new[k]) {
        m_start[k] = 0;
        return;
    }
    double* s = scal + (size_t)k * GMG_SSTRIDE;
    const double bb = s[S_BB];
    if (x_zero) s[S_RR0] = bb;
    const double rr = s[S_RR0];
    s[S_DLT2] = delta2;
    s[S_RR] = rr;
    s[S_RREF] = rr;
    s[S_TGT] = tol2 * bb;
    s[S_ITS] = 0.0;
    s[S_NREP] = 1.0;
    const int go = bb > 0.0 && rr > tol2 * bb && isfinite(rr);
    s[S_DONE] = go ? 0.0 : 1.0;
    m_it[k] = go;
    m_start[k] = go;
}

// x (FP64) += e ; e = 0
template <typename TE>
__global__ void k_add_zero(long long N, double* __restrict__ x, TE* __restrict__ e, Bat bt) {
    BAT_SLOT(bt);
    x += vo;
    e += vo;
    // a capped grid striding over the cells: the launch sits in every iteration and is
    // masked off in most, where fewer blocks cost less to dispatch
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; t < (uint32_t)N; t += stride) {
        x[t] += (double)e[t];
        e[t] = (TE)0;
    }
}

// subtract a device scalar mean (singular problems): x -= s[slot] * scale
template <typename T>
__global__ void k_sub_mean(long long N, T* __restrict__ x, const double* __restrict__ s, int which, double scale,
                           Bat bt) {
    BAT_SLOT(bt);
    x += vo;
    s += (long long)slot * bt.ss;
    const uint32_t t = cell_id();
    if (t < (uint32_t)N) x[t] -= (T)(s[which] * scale);
}

template <typename T>
__global__ void k_sum(long long N, const T* __restrict__ x, double* __restrict__ part, Bat bt) {
    BAT_SLOT(bt);
    x += vo;
    part += (long long)slot * bt.ps;
    double a = 0;
    uint32_t tb, te;
    block_range((uint32_t)N, tb, te);
    for (uint32_t t = tb + threadIdx.x; t < te; t += blockDim.x)
        a += (double)x[t];
    a = gmg_block_sum(a);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// ------------------------------------------------------------------ host side
static inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }
static inline int rblk(long long n) { return (int)std::min<long long>(GMG_RBLK, (n + 255) / 256); }

// z-marching kernels for the big 3-D constant-coefficient levels
static bool use_z(const Hierarchy& H, int l) {
    const GLevelHost& L = H.host[l];
    return !H.var && L.n[0] >= 64 && L.N >= (H.dim == 3 ? (1 << 18) : (1 << 16));
}
// blocks of a z-marching launch: one per tile item (each block's tiles stay in its XCD's
// L2; striding fewer blocks over the items measured slower), at most GMG_MAXB
static inline int zgrid(const ZTile& z) { return std::min(z.items, GMG_MAXB); }
static ZTile ztile(const GLevelHost& L) {
    ZTile z;
    z.zc = L.n[2] == 1 ? 1 : L.n[2] >= 256 ? 8 : 4;
    z.tx = (L.n[0] + 63) / 64;
    z.ty = (L.n[1] + 3) / 4;
    z.nz = (L.n[2] + z.zc - 1) / z.zc;
    z.items = z.tx * z.ty * z.nz;
    return z;
}
// z-pipelined LDS-ring passes (gmg_zp.hpp) for big 3-D constant-coefficient levels
template <typename T>
static ZpGrid zp_grid(const GLevelHost& L) {
    using C = ZpCfg<T>;
    ZpGrid g;
    g.tx = (L.n[0] + C::TX - 1) / C::TX;
    g.ty = (L.n[1] + C::TY - 1) / C::TY;
    const long long tiles = (long long)g.tx * g.ty;
    // chunks of 16 planes, fewer when that leaves under ~8 workgroups per CU
    int zc = 16;
    while (zc > 2 && tiles * ((L.n[2] + zc - 1) / zc) < 2048) zc /= 2;
    while (zc < L.n[2] && tiles * ((L.n[2] + zc - 1) / zc) > GMG_MAXB) zc *= 2;
    g.zc = zc;
    g.nch = (L.n[2] + zc - 1) / zc;
    g.blocks = (int)std::min<long long>(tiles * g.nch, 1LL << 30);
    return g;
}
static bool use_zp(const Hierarchy& H, int l) {
    const GLevelHost& L = H.host[l];
    return H.dim == 3 && !H.var && L.n[0] >= 64 && L.n[1] >= 16 && L.n[2] >= 8 && L.N >= (1 << 18) &&
           zp_grid<double>(L).blocks <= GMG_MAXB && zp_grid<float>(L).blocks <= GMG_MAXB;
}
// the fused restriction also needs even extents (children pair up inside a tile and a chunk)
static bool use_zp_restrict(const Hierarchy& H, int l) {
    const GLevelHost& L = H.host[l];
    return use_zp(H, l) && L.n[0] % 2 == 0 && L.n[1] % 2 == 0 && L.n[2] % 2 == 0;
}

// blocks (= partial sums per value) of every finest-level reduction launch in precision T
template <typename T>
static int red_count(const Hierarchy& H) {
    if (use_zp(H, 0)) return zp_grid<T>(H.host[0]).blocks;
    if (use_z(H, 0)) return zgrid(ztile(H.host[0]));
    return rblk(H.host[0].N);
}

template <typename T>
static GOp<T> make_op(const Hierarchy& H, int l) {
    const GLevelHost& L = H.host[l];
    const GLevelDev& D = H.dev[l];
    GOp<T> o{};
    o.n0 = L.n[0];
    o.n1 = L.n[1];
    o.n2 = L.n[2];
    o.s1 = (uint32_t)L.n[0];
    o.s2 = (uint32_t)L.n[0] * (uint32_t)L.n[1];
    o.f0 = FastDiv((uint32_t)L.n[0]);
    o.f1 = FastDiv((uint32_t)L.n[1]);
    o.dim = H.dim;
    for (int d = 0; d < 3; ++d) {
        o.per[d] = H.per[d];
        o.a[d] = (T)L.a[d];
        for (int s = 0; s < 2; ++s) o.bfac[d][s] = (T)(1.0 - H.alpha[d][s]);
        if (std::is_same<T, double>::value) o.af[d] = (const T*)D.af64[d];
        else o.af[d] = (const T*)D.af32[d];
    }
    o.sigma = (T)L.sigma;
    o.sg = std::is_same<T, double>::value ? (const T*)D.sg64 : (const T*)D.sg32;
    // interior diagonal, accumulated in precision T in the order the kernels use
    T di = o.sigma;
    for (int d = 0; d < H.dim; ++d) {
        di += o.a[d];
        di += o.a[d];
    }
    o.dint = di;
    o.wdi = (T)H.omega / di;
    return o;
}

static long long face_count(const int n[3], int d) {
    return (long long)(n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2));
}

// dense operator of a level (host, FP64), for the coarsest solve
static void dense_matrix(const Hierarchy& H, const GLevelHost& L, std::vector<double>& A) {
    int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    long long N = L.N;
    A.assign(N * N, 0.0);
    for (long long c = 0; c < N; ++c) {
        int i = (int)(c % n0), j = (int)((c / n0) % n1), k = (int)(c / ((long long)n0 * n1));
        int cc[3] = {i, j, k};
        double D = L.var ? L.sg[c] : L.sigma;
        for (int d = 0; d < H.dim; ++d) {
            for (int s = 0; s < 2; ++s) {
                int fc[3] = {i, j, k};
                if (s == 1) fc[d] += 1;
                double a = L.var ? L.af[d][face_index(d, fc[0], fc[1], fc[2], n0, n1)] : L.a[d];
                int q[3] = {i, j, k};
                q[d] += s == 0 ? -1 : 1;
                bool inside = q[d] >= 0 && q[d] < L.n[d];
                if (!inside && H.per[d]) {
                    q[d] = q[d] < 0 ? L.n[d] - 1 : 0;
                    inside = true;
                }
                if (inside) {
                    D += a;
                    long long nb = q[0] + (long long)n0 * (q[1] + (long long)n1 * q[2]);
                    A[c * N + nb] -= a;
                } else {
                    D += a * (1.0 - H.alpha[d][s]);
                }
            }
        }
        (void)cc;
        A[c * N + c] += D;
    }
}

// Gauss-Jordan inverse with partial pivoting (small dense matrices)
static bool invert(std::vector<double>& A, int n, std::vector<double>& inv) {
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)p * n + c])) p = r;
        if (A[(size_t)p * n + c] == 0.0) return false;
        if (p != c)
            for (int k = 0; k < n; ++k) {
                std::swap(A[(size_t)p * n + k], A[(size_t)c * n + k]);
                std::swap(inv[(size_t)p * n + k], inv[(size_t)c * n + k]);
            }
        double d = 1.0 / A[(size_t)c * n + c];
        for (int k = 0; k < n; ++k) {
            A[(size_t)c * n + k] *= d;
            inv[(size_t)c * n + k] *= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = A[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
                inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
            }
        }
    }
    return true;
}

// the next coarser level's shape and constant-mode coefficients; its variable-mode
// arrays too when `arrays` (host master copy)
static void coarsen(const Hierarchy& H, const GLevelHost& F, GLevelHost& C, bool arrays = true) {
    for (int d = 0; d < 3; ++d) C.n[d] = d < H.dim ? F.n[d] / 2 : 1;
    C.N = (long long)C.n[0] * C.n[1] * C.n[2];
    C.var = F.var;
    for (int d = 0; d < 3; ++d) C.a[d] = F.a[d] * 0.25;
    C.sigma = F.sigma;
    if (!F.var || !arrays) return;
    const int tw = H.dim == 3 ? 4 : 2;  // fine faces per coarse face
    for (int d = 0; d < H.dim; ++d) {
        C.af[d].assign(face_count(C.n, d), 0.0);
        int e0 = C.n[0] + (d == 0), e1 = C.n[1] + (d == 1), e2 = C.n[2] + (d == 2);
        for (int K = 0; K < e2; ++K)
            for (int J = 0; J < e1; ++J)
                for (int I = 0; I < e0; ++I) {
                    double s = 0;
                    int nz = (H.dim == 3 && d != 2) ? 2 : 1;
                    int ny = d != 1 ? 2 : 1;
                    int nx = d != 0 ? 2 : 1;
                    for (int dk = 0; dk < nz; ++dk)
                        for (int dj = 0; dj < ny; ++dj)
                            for (int di = 0; di < nx; ++di) {
                                int fi = 2 * I + di, fj = 2 * J + dj, fk = H.dim == 3 ? 2 * K + dk : 0;
                                s += F.af[d][face_index(d, fi, fj, fk, F.n[0], F.n[1])];
                            }
                    C.af[d][face_index(d, I, J, K, C.n[0], C.n[1])] = s / tw * 0.25;
                }
    }
    C.sg.assign(C.N, 0.0);
    for (long long c = 0; c < C.N; ++c) {
        int I = (int)(c % C.n[0]), J = (int)((c / C.n[0]) % C.n[1]), K = (int)(c / ((long long)C.n[0] * C.n[1]));
        double s = 0;
        int nz = H.dim == 3 ? 2 : 1;
        for (int dk = 0; dk < nz; ++dk)
            for (int dj = 0; dj < 2; ++dj)
                for (int di = 0; di < 2; ++di) {
                    long long f = (2 * I + di) + (long long)F.n[0] * ((2 * J + dj) + (long long)F.n[1] * (H.dim == 3 ? 2 * K + dk : 0));
                    s += F.sg[f];
                }
        C.sg[c] = s / (1 << H.dim);
    }
}

template <typename T>
static int alloc(void** p, long long n, size_t& bytes) {
    OPF_HIP(hipMalloc(p, sizeof(T) * std::max<long long>(n, 1)));
    bytes += sizeof(T) * std::max<long long>(n, 1);
    return OPF_OK;
}

// parameters and singularity of H and the host description of its levels (the
// coarse levels' variable-mode arrays only when `arrays`); shift: some cell of
// the fine level has a nonzero sigma
static int hier_levels(Hierarchy& H, const GLevelHost& fine, const int per[3], const double alpha[3][2],
                       int precision, int slots, bool shift, bool arrays) {
    H.precision = precision;
    H.B = std::max(1, slots);
    H.var = fine.var;
    for (int d = 0; d < 3; ++d) {
        H.per[d] = per[d];
        for (int s = 0; s < 2; ++s) H.alpha[d][s] = alpha[d][s];
    }
    H.dim = fine.n[2] > 1 ? 3 : 2;
    if (fine.n[1] == 1) return (set_error("gmg: 1-D grids are not supported"), OPF_ERR_UNSUPPORTED);
    OPF_CHECK(fine.N < (1LL << 31), OPF_ERR_UNSUPPORTED, "gmg: %lld cells per system (kernels index cells with 32 bits)",
              fine.N);
    H.omega = H.dim == 3 ? 6.0 / 7.0 : 0.8;
    // singular: no Dirichlet-like face and no diagonal shift
    bool anchored = false;
    for (int d = 0; d < H.dim; ++d)
        for (int s = 0; s < 2; ++s) anchored |= !per[d] && alpha[d][s] != 1.0;
    H.singular = !shift && !anchored;
    H.host.clear();
    H.host.push_back(fine);
    for (;;) {
        const GLevelHost& F = H.host.back();
        bool can = F.N > 64;
        for (int d = 0; d < H.dim; ++d) can &= (F.n[d] % 2 == 0) && F.n[d] >= 4;
        if (!can) break;
        GLevelHost C;
        coarsen(H, F, C, arrays);
        H.host.push_back(C);
    }
    const GLevelHost& Lc = H.host.back();
    OPF_CHECK(Lc.N <= 1024, OPF_ERR_UNSUPPORTED,
              "gmg: coarsest grid has %lld cells (> 1024); use sizes divisible by a power of two", Lc.N);
    return OPF_OK;
}

// dense inverse of the coarsest level (its arrays on the host in variable mode)
// into H.cinv (allocated on first use)
static int hier_coarsest(Hierarchy& H) {
    const GLevelHost& Lc = H.host.back();
    std::vector<double> A, inv;
    dense_matrix(H, Lc, A);
    if (H.singular) {
        double dm = 0;
        for (long long c = 0; c < Lc.N; ++c) dm += A[c * Lc.N + c];
        dm /= (double)Lc.N * Lc.N;
        for (auto& v : A) v += dm;
    }
    OPF_CHECK(invert(A, (int)Lc.N, inv), OPF_ERR_ARG, "gmg: singular coarsest operator");
    if (!H.cinv) {
        H.cN = (int)Lc.N;
        OPF_TRY(alloc<double>((void**)&H.cinv, (long long)H.cN * H.cN, H.bytes));
    }
    OPF_HIP(hipMemcpy(H.cinv, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice));
    return OPF_OK;
}

// work vectors of every level (x, b below the finest, t), the Krylov vectors and
// the scalars of the B slots
static int hier_vectors(Hierarchy& H) {
    const bool mixed = H.precision == OPF_PREC_MIXED;
    const size_t ts = mixed ? sizeof(float) : sizeof(double);
    size_t& bytes = H.bytes;
    for (size_t l = 0; l < H.host.size(); ++l) {
        GLevelDev& D = H.dev[l];
        const size_t BN = (size_t)H.B * H.host[l].N;
        if (l > 0) {
            OPF_HIP(hipMalloc(&D.x, ts * BN));
            OPF_HIP(hipMalloc(&D.b, ts * BN));
            bytes += 2 * ts * BN;
        }
        OPF_HIP(hipMalloc(&D.t, ts * BN));
        bytes += ts * BN;
    }
    const size_t BN = (size_t)H.B * H.host[0].N;
    OPF_HIP(hipMalloc(&H.r, ts * BN));
    OPF_HIP(hipMalloc(&H.z, ts * BN));
    OPF_HIP(hipMalloc(&H.p[0], ts * BN));
    OPF_HIP(hipMalloc(&H.p[1], ts * BN));
    OPF_HIP(hipMalloc(&H.q, ts * BN));
    bytes += 5 * ts * BN;
    OPF_HIP(hipMalloc(&H.e, ts * BN));  // inner correction (mixed) / serving correction
    bytes += ts * BN;
    OPF_HIP(hipMalloc((void**)&H.r64, sizeof(double) * BN));
    bytes += sizeof(double) * BN;
    OPF_HIP(hipMalloc((void**)&H.part, sizeof(double) * GMG_PSTRIDE * H.B));
    OPF_HIP(hipMalloc((void**)&H.scal, sizeof(double) * GMG_SSTRIDE * H.B));
    OPF_HIP(hipMemset(H.scal, 0, sizeof(double) * GMG_SSTRIDE * H.B));
    OPF_HIP(hipHostMalloc((void**)&H.hscal, sizeof(double) * GMG_SSTRIDE * H.B, hipHostMallocDefault));
    OPF_HIP(hipMalloc((void**)&H.mk, sizeof(int) * MK_COUNT * H.B));
    OPF_HIP(hipMemset(H.mk, 0, sizeof(int) * MK_COUNT * H.B));
    return OPF_OK;
}

int gmg_build(opf_ctx* ctx, Hierarchy& H, const GLevelHost& fine, const int per[3], const double alpha[3][2],
              int precision, int slots) {
    bool shift = fine.var ? false : fine.sigma != 0.0;
    if (fine.var)
        for (double v : fine.sg) shift |= v != 0.0;
    OPF_TRY(hier_levels(H, fine, per, alpha, precision, slots, shift, true));
    H.bytes = 0;
    OPF_TRY(hier_coarsest(H));
    // device levels
    H.dev.assign(H.host.size(), GLevelDev{});
    const bool mixed = precision == OPF_PREC_MIXED;
    size_t& bytes = H.bytes;
    for (size_t l = 0; l < H.host.size(); ++l) {
        GLevelHost& L = H.host[l];
        GLevelDev& D = H.dev[l];
        for (int d = 0; d < 3; ++d) D.n[d] = L.n[d];
        D.N = L.N;
        if (L.var) {
            for (int d = 0; d < H.dim; ++d) {
                long long nf = face_count(L.n, d);
                OPF_TRY(alloc<double>((void**)&D.af64[d], nf, bytes));
                OPF_HIP(hipMemcpy(D.af64[d], L.af[d].data(), nf * sizeof(double), hipMemcpyHostToDevice));
                if (mixed) {
                    std::vector<float> f(L.af[d].begin(), L.af[d].end());
                    OPF_TRY(alloc<float>((void**)&D.af32[d], nf, bytes));
                    OPF_HIP(hipMemcpy(D.af32[d], f.data(), nf * sizeof(float), hipMemcpyHostToDevice));
                }
            }
            OPF_TRY(alloc<double>((void**)&D.sg64, L.N, bytes));
            OPF_HIP(hipMemcpy(D.sg64, L.sg.data(), L.N * sizeof(double), hipMemcpyHostToDevice));
            if (mixed) {
                std::vector<float> f(L.sg.begin(), L.sg.end());
                OPF_TRY(alloc<float>((void**)&D.sg32, L.N, bytes));
                OPF_HIP(hipMemcpy(D.sg32, f.data(), L.N * sizeof(float), hipMemcpyHostToDevice));
            }
            if (l > 0) {  // the fine arrays are kept only where needed
                L.af[0].clear(); L.af[1].clear(); L.af[2].clear(); L.sg.clear();
            }
        }
    }
    OPF_TRY(hier_vectors(H));
    (void)ctx;
    return OPF_OK;
}

int gmg_build_dev(opf_ctx* ctx, Hierarchy& H, const DevLevel& fine, const int per[3], const double alpha[3][2],
                  int precision, int slots) {
    hipStream_t st = ctx->stream;
    GLevelHost F{};
    for (int d = 0; d < 3; ++d) {
        F.n[d] = fine.n[d];
        F.a[d] = fine.a[d];
    }
    F.N = fine.N;
    F.var = fine.var;
    F.sigma = fine.sigma;
    // an existing hierarchy of the same layout keeps its buffers (and graph)
    bool same = !H.dev.empty() && H.var == fine.var && H.precision == precision && H.B == std::max(1, slots);
    for (int d = 0; d < 3 && same; ++d)
        same &= H.host[0].n[d] == fine.n[d] && H.per[d] == per[d] && H.alpha[d][0] == alpha[d][0] &&
                H.alpha[d][1] == alpha[d][1];
    if (same && fine.var) {
        same = H.dev[0].borrowed && H.dev[0].sg64 == fine.sg;
        for (int d = 0; d < 3; ++d) same &= H.dev[0].af64[d] == fine.af[d];
        bool anchored = false;
        for (int d = 0; d < H.dim; ++d)
            for (int s = 0; s < 2; ++s) anchored |= !per[d] && alpha[d][s] != 1.0;
        same &= H.singular == (!fine.shift && !anchored);
    }
    if (same && !fine.var) {
        same = H.host[0].sigma == fine.sigma;
        for (int d = 0; d < 3; ++d) same &= H.host[0].a[d] == fine.a[d];
        if (same) return OPF_OK;  // constant mode: the same operator
    }
    if (!same) {
        const int pre = H.pre, post = H.post;
        if (!H.dev.empty()) gmg_free(H);
        H = Hierarchy{};
        H.pre = pre;
        H.post = post;
    }
    if (!fine.var) return gmg_build(ctx, H, F, per, alpha, precision, slots);  // two numbers per level
    const bool mixed = precision == OPF_PREC_MIXED;
    if (!same) {
        OPF_TRY(hier_levels(H, F, per, alpha, precision, slots, fine.shift, false));
        H.bytes = 0;
        H.dev.assign(H.host.size(), GLevelDev{});
        for (size_t l = 0; l < H.host.size(); ++l) {
            const GLevelHost& L = H.host[l];
            GLevelDev& D = H.dev[l];
            for (int d = 0; d < 3; ++d) D.n[d] = L.n[d];
            D.N = L.N;
            if (l == 0) {
                D.borrowed = true;
                for (int d = 0; d < 3; ++d) D.af64[d] = fine.af[d];
                D.sg64 = fine.sg;
            } else {
                for (int d = 0; d < H.dim; ++d) OPF_TRY(alloc<double>((void**)&D.af64[d], face_count(L.n, d), H.bytes));
                OPF_TRY(alloc<double>((void**)&D.sg64, L.N, H.bytes));
            }
            if (mixed) {
                for (int d = 0; d < H.dim; ++d) OPF_TRY(alloc<float>((void**)&D.af32[d], face_count(L.n, d), H.bytes));
                OPF_TRY(alloc<float>((void**)&D.sg32, L.N, H.bytes));
            }
        }
        OPF_TRY(hier_vectors(H));
    }
    // coefficients of the coarse levels, FP32 copies, coarsest inverse
    for (size_t l = 1; l < H.host.size(); ++l) {
        const GLevelDev &P = H.dev[l - 1], &D = H.dev[l];
        const double* faf[3] = {P.af64[0], P.af64[1], P.af64[2]};
        double* const caf[3] = {D.af64[0], D.af64[1], D.af64[2]};
        coarsen_dev(st, H.dim, P.n, faf, P.sg64, D.n, caf, D.sg64);
    }
    if (mixed)
        for (size_t l = 0; l < H.host.size(); ++l) {
            const GLevelDev& D = H.dev[l];
            for (int d = 0; d < H.dim; ++d) to_f32_dev(st, face_count(D.n, d), D.af64[d], D.af32[d]);
            to_f32_dev(st, D.N, D.sg64, D.sg32);
        }
    OPF_HIP(hipGetLastError());
    GLevelHost& Lc = H.host.back();
    const GLevelDev& Dc = H.dev.back();
    for (int d = 0; d < H.dim; ++d) {
        Lc.af[d].resize(face_count(Lc.n, d));
        OPF_HIP(hipMemcpyAsync(Lc.af[d].data(), Dc.af64[d], sizeof(double) * Lc.af[d].size(), hipMemcpyDeviceToHost, st));
    }
    Lc.sg.resize(Lc.N);
    OPF_HIP(hipMemcpyAsync(Lc.sg.data(), Dc.sg64, sizeof(double) * Lc.N, hipMemcpyDeviceToHost, st));
    OPF_HIP(hipStreamSynchronize(st));
    int rc = hier_coarsest(H);
    for (int d = 0; d < 3; ++d) std::vector<double>().swap(Lc.af[d]);
    std::vector<double>().swap(Lc.sg);
    return rc;
}

void gmg_free(Hierarchy& H) {
    for (auto& D : H.dev) {
        for (int d = 0; d < 3; ++d) {
            if (D.af64[d] && !D.borrowed) hipFree(D.af64[d]);
            if (D.af32[d]) hipFree(D.af32[d]);
        }
        if (D.sg64 && !D.borrowed) hipFree(D.sg64);
        if (D.sg32) hipFree(D.sg32);
        if (D.x) hipFree(D.x);
        if (D.b) hipFree(D.b);
        if (D.t) hipFree(D.t);
    }
    H.dev.clear();
    void* ptrs[] = {H.cinv, H.r, H.z, H.p[0], H.p[1], H.q, H.e, H.r64, H.part, H.scal};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (H.hscal) hipHostFree(H.hscal);
    if (H.mk) hipFree(H.mk);
    if (H.iter_exec) hipGraphExecDestroy(H.iter_exec);
    if (H.iter_graph) hipGraphDestroy(H.iter_graph);
    H.iter_exec = nullptr;
    H.iter_graph = nullptr;
    H.mk = nullptr;
    H.cinv = nullptr;
    H.r = H.z = H.p[0] = H.p[1] = H.q = H.e = nullptr;
    H.r64 = H.part = H.scal = H.hscal = nullptr;
}

// ---- launch wrappers dispatching on (VAR, DIM)
#define GMG_DISPATCH(VAR, DIM, ...)                                                 \
    do {                                                                            \
        if (VAR) {                                                                  \
            if (DIM == 3) { constexpr bool V_ = true; constexpr int D_ = 3; __VA_ARGS__; }  \
            else { constexpr bool V_ = true; constexpr int D_ = 2; __VA_ARGS__; }           \
        } else {                                                                    \
            if (DIM == 3) { constexpr bool V_ = false; constexpr int D_ = 3; __VA_ARGS__; } \
            else { constexpr bool V_ = false; constexpr int D_ = 2; __VA_ARGS__; }          \
        }                                                                           \
    } while (0)

static inline Bat bat(const Hierarchy& H, const int* mask, long long vs) {
    return Bat{mask, vs, GMG_SSTRIDE, GMG_PSTRIDE};
}

// one z-pipelined pass over level l (every slot of the batch)
template <typename T, int MODE, typename TO = T>
static void zp_launch(const Hierarchy& H, int l, hipStream_t st, const Bat& bt, const T* in0, const T* in1,
                      const T* cb, const T* cq, T* out0, TO* out1, double* part, long long cvs = 0) {
    const ZpGrid g = zp_grid<T>(H.host[l]);
    const GOp<T> o = make_op<T>(H, l);
    hipLaunchKernelGGL((zp_pass<T, MODE, TO>), dim3(g.blocks, (unsigned)H.B), dim3(ZpCfg<T>::NT), 0, st, o, in0, in1,
                       cb, cq, out0, out1, (const double*)H.scal, part, (T)H.omega, g, cvs, bt);
}

// the up half of a V(2,2) level as one zp2_pass: FP64, levels of >= 8-plane chunks
template <typename T>
static bool use_zp2_up(const Hierarchy& H, int l) {
    return std::is_same<T, double>::value && use_zp_restrict(H, l) && H.pre == 2 && H.post == 2 &&
           zp_grid<T>(H.host[l]).zc >= 8;
}
template <typename T, int M2>
static void zp2_launch(const Hierarchy& H, int l, hipStream_t st, const Bat& bt, const T* xin, const T* xc, const T* cb,
                       const T* cq, T* out0, double* part, long long cvs) {
    const ZpGrid g = zp_grid<T>(H.host[l]);
    const GOp<T> o = make_op<T>(H, l);
    hipLaunchKernelGGL((zp2_pass<T, M2>), dim3(g.blocks, (unsigned)H.B), dim3(ZpCfg<T>::NT), 0, st, o, xin, xc, cb, cq,
                       out0, part, (T)H.omega, g, cvs, bt);
}

// first level (>= 1) handled by the coarse tail, or -1: levels of at most GMG_TAIL_N cells,
// at least two of them (the coarsest included)
static int tail_level(const Hierarchy& H) {
    const int nl = (int)H.dev.size();
    if (H.pre < 1 || H.cN > GMG_TAIL_CELLS) return -1;
    for (int l = 1; l < nl; ++l) {
        if (H.host[l].N > GMG_TAIL_N) continue;
        long long cells = 0;
        for (int q = l; q < nl; ++q) cells += H.host[q].N;
        return (nl - l >= 2 && nl - l <= GMG_TAIL_MAXL && cells <= GMG_TAIL_CELLS) ? l : -1;
    }
    return -1;
}

template <typename T>
static void ph_vtail(const Hierarchy& H, hipStream_t st, const int* mask, int l0) {
    TailArgs<T> ta{};
    ta.nl = (int)H.dev.size() - l0;
    int ofs = 0;
    for (int q = 0; q < ta.nl; ++q) {
        const GLevelDev& D = H.dev[l0 + q];
        ta.lv[q].o = make_op<T>(H, l0 + q);
        ta.lv[q].N = (int)D.N;
        ta.lv[q].ofs = ofs;
        ofs += 3 * (int)D.N;
    }
    ta.b0 = (const T*)H.dev[l0].b;
    ta.x0 = (T*)H.dev[l0].x;
    ta.cinv = H.cinv;
    ta.cN = H.cN;
    ta.w = (T)H.omega;
    ta.pre = H.pre;
    ta.post = H.post;
    const Bat bt = bat(H, mask, H.dev[l0].N);
    GMG_DISPATCH(H.var, H.dim, hipLaunchKernelGGL((k_vtail<T, V_, D_>), dim3(1, (unsigned)H.B), dim3(GMG_TAIL_NT), 0,
                                                  st, ta, bt));
}

// 2-D constant-coefficient V(2,2) levels run as k2_down / k2_up: one block per tile
// (at most GMG_MAXB, then striding); on the finest level k2_up writes the partials
// of the three dot products, one per block
static bool fused2d(const Hierarchy& H) { return H.dim == 2 && !H.var && H.pre == 2 && H.post == 2; }
// tile side of a level's k2 kernels (threads: 256 for 32 and 16, 64 for 8); measured on
// the 2048^2 hierarchy: 16 x 16 tiles took the 1024^2 level's pair 26.2 -> 25.1 us and
// 8 x 8 the 128^2 and 64^2 levels' 8-10 -> 5.5 us each, while a 32^2 level stays in the
// coarse tail (its 8 x 8 pair costs 11 us, the tail 5 us more than without it)
static int f2_ts(const GLevelHost& L) { return L.N >= (1LL << 21) ? 32 : L.N >= (1LL << 16) ? 16 : 8; }
static F2Tile f2tile(const GLevelHost& L) {
    const int ts = f2_ts(L);
    F2Tile ft;
    ft.tx = (L.n[0] + ts - 1) / ts;
    ft.ty = (L.n[1] + ts - 1) / ts;
    ft.tiles = ft.tx * ft.ty;
    return ft;
}
// partial sums per value that the finest level's V-cycle leaves for the DOT3 finish
template <typename T>
static int dot3_parts(const Hierarchy& H) {
    if (fused2d(H) && H.dev.size() > 1) return std::min(f2tile(H.host[0]).tiles, GMG_MAXB);
    return red_count<T>(H);
}

template <typename T>
void ph_vcycle(const Hierarchy& H, hipStream_t st, const int* mask, int l, const T* rhs, T* out,
               const T* dot_q = nullptr) {
    if (l >= 1 && l == tail_level(H)) {  // rhs / out are this level's b / x
        ph_vtail<T>(H, st, mask, l);
        return;
    }
    const GLevelDev& D = H.dev[l];
    const GLevelHost& L = H.host[l];
    const long long N = L.N;
    const unsigned B = (unsigned)H.B;
    if (l == (int)H.dev.size() - 1) {
        hipLaunchKernelGGL(k_dense<T>, dim3(1, B), dim3(256), 0, st, H.cinv, H.cN, rhs, out, bat(H, mask, N));
        if (dot_q)  // a one-level hierarchy: the dot products follow the dense solve
            hipLaunchKernelGGL(k_dot3<T>, dim3(red_count<T>(H), B), dim3(256), 0, st, N, rhs, (const T*)out, dot_q,
                               H.part, bat(H, mask, N));
        return;
    }
    GOp<T> o = make_op<T>(H, l);
    const T w = (T)H.omega;
    if (fused2d(H)) {
        // the level in two launches (k2_down / k2_up): x2 kept in the level temp
        const GLevelDev& C = H.dev[l + 1];
        const F2Tile ft = f2tile(L);
        const unsigned G = (unsigned)std::min(ft.tiles, GMG_MAXB);
        const Bat bt = bat(H, mask, N);
        const int ts = f2_ts(L);
        auto by_ts = [&](auto launch) {
            if (ts == 32) launch(std::integral_constant<int, 32>{}, std::integral_constant<int, 256>{});
            else if (ts == 16) launch(std::integral_constant<int, 16>{}, std::integral_constant<int, 256>{});
            else launch(std::integral_constant<int, 8>{}, std::integral_constant<int, 64>{});
        };
        by_ts([&](auto TS_, auto NT_) {
            constexpr int TS = decltype(TS_)::value, NT = decltype(NT_)::value;
            hipLaunchKernelGGL((k2_down<T, TS, NT>), dim3(G, B), dim3(NT), 0, st, o, rhs, (T*)D.t, (T*)C.b, C.n[0],
                               C.n[1], w, ft, bt);
        });
        ph_vcycle<T>(H, st, mask, l + 1, (const T*)C.b, (T*)C.x);
        by_ts([&](auto TS_, auto NT_) {
            constexpr int TS = decltype(TS_)::value, NT = decltype(NT_)::value;
            if (dot_q)  // one partial per block (dot3_parts)
                hipLaunchKernelGGL((k2_up<T, TS, NT, true>), dim3(G, B), dim3(NT), 0, st, o, (const T*)D.t,
                                   (const T*)C.x, C.n[0], C.n[1], rhs, out, w, dot_q, H.part, ft, bt);
            else
                hipLaunchKernelGGL((k2_up<T, TS, NT, false>), dim3(G, B), dim3(NT), 0, st, o, (const T*)D.t,
                                   (const T*)C.x, C.n[0], C.n[1], rhs, out, w, (const T*)nullptr, (double*)nullptr,
                                   ft, bt);
        });
        return;
    }
    // Jacobi sweeps ping-pong between `out` and the level temp; start on the
    // buffer that makes the last sweep land in `out` (with the fused up half: the
    // pre-smoothed x in the temp, read by the one pass that writes `out`)
    const bool up2 = use_zp2_up<T>(H, l);
    T* cur = out;
    T* tmp = (T*)D.t;
    if (!up2 && (((H.pre - 1) + H.post) & 1)) std::swap(cur, tmp);
    const Bat bt = bat(H, mask, N);
    const bool zm = use_z(H, l);
    const bool zp = use_zp(H, l);
    const ZTile zt = ztile(L);
    int s0 = 1;
    if (!H.var && H.pre >= 2) {
        // sweeps 1 and 2 from zero in one pass, landing where the second sweep of
        // the two-launch sequence would (the buffer after `cur`)
        if (zp)
            zp_launch<T, ZP_JAC2Z>(H, l, st, bt, rhs, nullptr, rhs, nullptr, tmp, (T*)nullptr, nullptr);
        else if (zm && H.dim == 3)
            hipLaunchKernelGGL((zk_jacobi2z<T, 3>), dim3(zgrid(zt), B), dim3(256), 0, st, o, rhs, tmp, w, zt, bt);
        else if (zm)
            hipLaunchKernelGGL((zk_jacobi2z<T, 2>), dim3(zgrid(zt), B), dim3(256), 0, st, o, rhs, tmp, w, zt, bt);
        else if (H.dim == 3)
            hipLaunchKernelGGL((k_jacobi2z<T, 3>), dim3(nblk(N), B), dim3(256), 0, st, o, rhs, tmp, w, bt);
        else
            hipLaunchKernelGGL((k_jacobi2z<T, 2>), dim3(nblk(N), B), dim3(256), 0, st, o, rhs, tmp, w, bt);
        std::swap(cur, tmp);
        s0 = 2;
    } else {
        GMG_DISPATCH(H.var, H.dim, hipLaunchKernelGGL((k_jacobi<T, V_, D_, true>), dim3(nblk(N), B), dim3(256), 0, st,
                                                      o, (const T*)nullptr, rhs, cur, w, bt));
    }
    for (int s = s0; s < H.pre; ++s) {
        if (zp)
            zp_launch<T, ZP_JAC>(H, l, st, bt, (const T*)cur, nullptr, rhs, nullptr, tmp, (T*)nullptr, nullptr);
        else if (zm)
            GMG_DISPATCH(false, H.dim, hipLaunchKernelGGL((zk_jacobi<T, D_>), dim3(zgrid(zt), B), dim3(256), 0, st, o,
                                                          (const T*)cur, rhs, tmp, w, zt, bt));
        else
            GMG_DISPATCH(H.var, H.dim, hipLaunchKernelGGL((k_jacobi<T, V_, D_, false>), dim3(nblk(N), B), dim3(256),
                                                          0, st, o, (const T*)cur, rhs, tmp, w, bt));
        std::swap(cur, tmp);
    }
    const GLevelDev& C = H.dev[l + 1];
    const FastDiv fc0((uint32_t)C.n[0]), fc1((uint32_t)C.n[1]);
    if (use_zp_restrict(H, l))
        zp_launch<T, ZP_RESTRICT>(H, l, st, bt, (const T*)cur, nullptr, rhs, nullptr, (T*)C.b, (T*)nullptr, nullptr,
                                  C.N);
    else if (N <= (1LL << 20))  // small levels: one thread per child
        GMG_DISPATCH(H.var, H.dim, hipLaunchKernelGGL((k_restrict_ch<T, V_, D_>), dim3(nblk(N), B), dim3(256), 0, st,
                                                      o, (const T*)cur, rhs, (T*)C.b, C.n[0], C.n[1], C.n[2], fc0,
                                                      fc1, bt));
    else
        GMG_DISPATCH(H.var, H.dim, hipLaunchKernelGGL((k_restrict<T, V_, D_>), dim3(nblk(C.N), B), dim3(256), 0, st,
                                                      o, (const T*)cur, rhs, (T*)C.b, C.n[0], C.n[1], C.n[2], fc0,
                                                      fc1, bt));
    ph_vcycle<T>(H, st, mask, l + 1, (const T*)C.b, (T*)C.x);
    if constexpr (std::is_same<T, double>::value) {
        if (up2) {
            if (dot_q)  // one partial per block (red_count: the zp_pass grid)
                zp2_launch<T, Z2_UP_DOT3>(H, l, st, bt, (const T*)cur, (const T*)C.x, rhs, dot_q, out, H.part, C.N);
            else
                zp2_launch<T, Z2_UP>(H, l, st, bt, (const T*)cur, (const T*)C.x, rhs, nullptr, out, nullptr, C.N);
            return;
        }
    }
    int p0 = 0;
    if (zp && (H.post >= 2 || (H.post == 1 && !dot_q))) {
        // the prolongation x += P x_c happens inside the first post-smoothing sweep
        zp_launch<T, ZP_PJAC>(H, l, st, bt, (const T*)cur, (const T*)C.x, rhs, nullptr, tmp, (T*)nullptr, nullptr,
                              C.N);
        std::swap(cur, tmp);
        p0 = 1;
    } else if (H.dim == 2 && L.N <= (1 << 18) && (H.post >= 2 || (H.post == 1 && !dot_q))) {
        // 2-D levels up to 512^2, where a launch costs more than the work: one launch
        // fewer per level. Larger levels (and 3-D, with 7 prolongations of 8 coarse
        // values per cell) keep the separate prolongation, which streams faster.
        if (H.var)
            hipLaunchKernelGGL((k_pjacobi<T, true, 2>), dim3(nblk(N), B), dim3(256), 0, st, o, (const T*)cur,
                               (const T*)C.x, C.n[0], C.n[1], C.n[2], rhs, tmp, w, bt);
        else
            hipLaunchKernelGGL((k_pjacobi<T, false, 2>), dim3(nblk(N), B), dim3(256), 0, st, o, (const T*)cur,
                               (const T*)C.x, C.n[0], C.n[1], C.n[2], rhs, tmp, w, bt);
        std::swap(cur, tmp);
        p0 = 1;
    } else if (H.dim == 3) {
        hipLaunchKernelGGL((k_prolong<T, 3>), dim3(nblk(N), B), dim3(256), 0, st, cur, (const T*)C.x, C.n[0], C.n[1],
                           C.n[2], o, bt);
    } else {
        hipLaunchKernelGGL((k_prolong<T, 2>), dim3(nblk(N), B), dim3(256), 0, st, cur, (const T*)C.x, C.n[0], C.n[1],
                           C.n[2], o, bt);
    }
    for (int s = p0; s < H.post; ++s) {
        if (dot_q && s == H.post - 1) {
            // last sweep also forms the partials of (r, z), (z, q), sum(z) (finest level only)
            const int G = red_count<T>(H);
            if (zp)
                zp_launch<T, ZP_JAC_DOT3>(H, l, st, bt, (const T*)cur, nullptr, rhs, dot_q, tmp, (T*)nullptr, H.part);
            else if (zm)
                GMG_DISPATCH(false, H.dim, hipLaunchKernelGGL((zk_jacobi_dot3<T, D_>), dim3(G, B), dim3(256), 0, st, o,
                                                              (const T*)cur, rhs, tmp, w, dot_q, H.part, zt, bt));
            else
                GMG_DISPATCH(H.var, H.dim, hipLaunchKernelGGL((k_jacobi_dot3<T, V_, D_>), dim3(G, B), dim3(256), 0,
                                                              st, o, (const T*)cur, rhs, tmp, w, dot_q, H.part, bt));
        } else if (zp) {
            zp_launch<T, ZP_JAC>(H, l, st, bt, (const T*)cur, nullptr, rhs, nullptr, tmp, (T*)nullptr, nullptr);
        } else if (zm) {
            GMG_DISPATCH(false, H.dim, hipLaunchKernelGGL((zk_jacobi<T, D_>), dim3(zgrid(zt), B), dim3(256), 0, st, o,
                                                          (const T*)cur, rhs, tmp, w, zt, bt));
        } else {
            GMG_DISPATCH(H.var, H.dim, hipLaunchKernelGGL((k_jacobi<T, V_, D_, false>), dim3(nblk(N), B), dim3(256),
                                                          0, st, o, (const T*)cur, rhs, tmp, w, bt));
        }
        std::swap(cur, tmp);
    }
}

// finish a reduction whose partials came from a finest-level launch in precision T
template <typename T>
static void ph_finish(const Hierarchy& H, hipStream_t st, const int* mask, int mode, int singular,
                      int* m_it = nullptr, int* m_rep = nullptr, int maxit = 0) {
    const long long N = H.host[0].N;
    const int nb = mode == FIN_DOT3 || mode == FIN_DOT3_FIRST ? dot3_parts<T>(H) : red_count<T>(H);
    hipLaunchKernelGGL(k_finish, dim3(1, (unsigned)H.B), dim3(1024), 0, st, H.part, nb, mode, H.scal, 1.0 / N,
                       singular, bat(H, mask, N), m_it, m_rep, maxit);
}

// start of an inner FCG solve on r: p_old = 0, z = M r, rz = (r, z)
template <typename T>
void ph_fcg_start(const Hierarchy& H, hipStream_t st, const int* mask) {
    const long long N = H.host[0].N;
    const unsigned B = (unsigned)H.B;
    const Bat bt = bat(H, mask, N);
    hipLaunchKernelGGL(k_fill<T>, dim3(nblk(N), B), dim3(256), 0, st, N, (T*)H.p[0], (T)0, bt);
    if (H.post > 0) {
        // (z, q) is not used by the first finish: any vector other than z serves as q
        ph_vcycle<T>(H, st, mask, 0, (const T*)H.r, (T*)H.z, (const T*)H.r);
    } else {
        ph_vcycle<T>(H, st, mask, 0, (const T*)H.r, (T*)H.z);
        hipLaunchKernelGGL(k_dot3<T>, dim3(red_count<T>(H), B), dim3(256), 0, st, N, (const T*)H.r, (const T*)H.z,
                           (const T*)H.z, H.part, bt);
    }
    ph_finish<T>(H, st, mask, FIN_DOT3_FIRST, H.singular);
}

// one FCG iteration on (e, r) with p ping-pong index pc
template <typename T>
void ph_fcg_iter(const Hierarchy& H, hipStream_t st, const int* mask, T* e, int pc) {
    const long long N = H.host[0].N;
    const unsigned B = (unsigned)H.B;
    const int G = red_count<T>(H);
    const Bat bt = bat(H, mask, N);
    GOp<T> o = make_op<T>(H, 0);
    T* z = (T*)H.z;
    T* q = (T*)H.q;
    T* r = (T*)H.r;
    T* pold = (T*)H.p[pc];
    T* pnew = (T*)H.p[1 - pc];
    if (use_zp(H, 0))
        zp_launch<T, ZP_SPMV_P>(H, 0, st, bt, (const T*)z, (const T*)pold, nullptr, nullptr, pnew, q, H.part);
    else if (use_z(H, 0))
        GMG_DISPATCH(false, H.dim, hipLaunchKernelGGL((zk_spmv_p<T, D_>), dim3(G, B), dim3(256), 0, st, o, (const T*)z,
                                                      (const T*)pold, pnew, q, H.scal, H.part, ztile(H.host[0]), bt));
    else
        GMG_DISPATCH(H.var, H.dim, hipLaunchKernelGGL((k_spmv_p<T, V_, D_>), dim3(G, B), dim3(256), 0, st, o,
                                                      (const T*)z, (const T*)pold, pnew, q, H.scal, H.part, bt));
    ph_finish<T>(H, st, mask, FIN_PQ, H.singular);
    hipLaunchKernelGGL(k_axpy2<T>, dim3(G, B), dim3(256), 0, st, N, e, (const T*)pnew, r, (const T*)q, H.scal, H.part,
                       bt);
    ph_finish<T>(H, st, mask, FIN_RR, H.singular);
    if (H.post > 0) {
        ph_vcycle<T>(H, st, mask, 0, (const T*)r, z, (const T*)q);  // last sweep forms the dot products
    } else {
        ph_vcycle<T>(H, st, mask, 0, (const T*)r, z);
        hipLaunchKernelGGL(k_dot3<T>, dim3(G, B), dim3(256), 0, st, N, (const T*)r, (const T*)z, (const T*)q, H.part,
                           bt);
    }
    ph_finish<T>(H, st, mask, FIN_DOT3, H.singular);
}

// r_T = b - A x in FP64 (x may be null), partial sums of r^2 -> finish `mode`
template <typename TO>
void ph_resid64(const Hierarchy& H, hipStream_t st, const int* mask, const double* x, const double* b, TO* r,
                int mode, int* m_it = nullptr, int maxit = 0) {
    GOp<double> o = make_op<double>(H, 0);
    const long long N = H.host[0].N;
    if (x && use_zp(H, 0))
        zp_launch<double, ZP_RESID, TO>(H, 0, st, bat(H, mask, N), x, nullptr, b, nullptr, (double*)nullptr, r,
                                        H.part);
    else if (use_z(H, 0))
        GMG_DISPATCH(false, H.dim, hipLaunchKernelGGL((zk_resid<double, TO, D_>), dim3(red_count<double>(H), (unsigned)H.B),
                                                      dim3(256), 0, st, o, x, b, r, H.part, ztile(H.host[0]),
                                                      bat(H, mask, N)));
    else
        GMG_DISPATCH(H.var, H.dim, hipLaunchKernelGGL((k_resid<double, TO, V_, D_>), dim3(red_count<double>(H), (unsigned)H.B),
                                                      dim3(256), 0, st, o, x, b, r, H.part, bat(H, mask, N)));
    ph_finish<double>(H, st, mask, mode, H.singular, m_it, nullptr, maxit);
}

template <typename T>
void ph_project(const Hierarchy& H, hipStream_t st, const int* mask, T* r) {
    const long long N = H.host[0].N;
    const Bat bt = bat(H, mask, N);
    hipLaunchKernelGGL(k_sum<T>, dim3(red_count<T>(H), (unsigned)H.B), dim3(256), 0, st, N, (const T*)r, H.part, bt);
    ph_finish<T>(H, st, mask, FIN_MEAN, 1);
    hipLaunchKernelGGL(k_sub_mean<T>, dim3(nblk(N), (unsigned)H.B), dim3(256), 0, st, N, r, H.scal, (int)S_MEAN,
                       1.0 / N, bt);
}

template <typename T>
static void ph_zero(const Hierarchy& H, hipStream_t st, const int* mask, T* v) {
    const long long N = H.host[0].N;
    hipLaunchKernelGGL(k_fill<T>, dim3(nblk(N), (unsigned)H.B), dim3(256), 0, st, N, v, (T)0, bat(H, mask, N));
}

static inline unsigned slot_blocks(const Hierarchy& H) { return (unsigned)((H.B + 63) / 64); }

template <typename T>
void ph_rel_start(const Hierarchy& H, hipStream_t st, double* x64, const double* b64, double tol2, double delta2,
                  bool x_zero) {
    const int B = H.B;
    const int* m_new = H.mk + MK_NEW * B;
    const int* m_start = H.mk + MK_START * B;
    if (x_zero) {
        ph_resid64<T>(H, st, m_new, nullptr, b64, (T*)H.r, FIN_BB);  // r = b, ||b||^2
    } else {
        ph_resid64<T>(H, st, m_new, nullptr, b64, (T*)nullptr, FIN_BB);
        ph_resid64<T>(H, st, m_new, x64, b64, (T*)H.r, FIN_RR0);
    }
    if (H.singular) ph_project<T>(H, st, m_new, (T*)H.r);
    hipLaunchKernelGGL(k_rel_init, dim3(slot_blocks(H)), dim3(64), 0, st, B, m_new, H.mk + MK_IT * B,
                       H.mk + MK_START * B, H.scal, tol2, delta2, (int)x_zero);
    ph_zero<T>(H, st, m_start, (T*)H.e);
    ph_fcg_start<T>(H, st, m_start);
}

template <typename T>
void ph_rel_iter(const Hierarchy& H, hipStream_t st, double* x64, const double* b64, int pc, int maxit) {
    const long long N = H.host[0].N;
    const int B = H.B;
    int* m_it = H.mk + MK_IT * B;
    int* m_rep = H.mk + MK_REP * B;
    const int G = red_count<T>(H);
    const Bat bt = bat(H, m_it, N);
    GOp<T> o = make_op<T>(H, 0);
    T* e = (T*)H.e;
    T* z = (T*)H.z;
    T* q = (T*)H.q;
    T* r = (T*)H.r;
    T* pold = (T*)H.p[pc];
    T* pnew = (T*)H.p[1 - pc];
    if (use_zp(H, 0))
        zp_launch<T, ZP_SPMV_P>(H, 0, st, bt, (const T*)z, (const T*)pold, nullptr, nullptr, pnew, q, H.part);
    else if (use_z(H, 0))
        GMG_DISPATCH(false, H.dim, hipLaunchKernelGGL((zk_spmv_p<T, D_>), dim3(G, B), dim3(256), 0, st, o, (const T*)z,
                                                      (const T*)pold, pnew, q, H.scal, H.part, ztile(H.host[0]), bt));
    else
        GMG_DISPATCH(H.var, H.dim, hipLaunchKernelGGL((k_spmv_p<T, V_, D_>), dim3(G, B), dim3(256), 0, st, o,
                                                      (const T*)z, (const T*)pold, pnew, q, H.scal, H.part, bt));
    ph_finish<T>(H, st, m_it, FIN_PQ, H.singular);
    // FP64: x advances in place (no correction vector, no fold-in launch); FP32: the
    // correction e accumulates in FP32 and is folded into x at each reliable update
    constexpr bool direct = std::is_same<T, double>::value;
    hipLaunchKernelGGL(k_axpy2<T>, dim3(G, B), dim3(256), 0, st, N, direct ? reinterpret_cast<T*>(x64) : e,
                       (const T*)pnew, r, (const T*)q, H.scal, H.part, bt);
    // (r, r), the iteration counted and the reliable-update decision (sets m_rep)
    ph_finish<T>(H, st, nullptr, FIN_RR_DECIDE, H.singular, m_it, m_rep, maxit);
    // reliable update: x += e, e = 0 (FP32), r = b - A x in FP64 (the direction p is
    // kept); its finish stops the slots that meet the tolerance (clears m_it)
    if (!direct)
        hipLaunchKernelGGL(k_add_zero<T>, dim3((unsigned)rblk(N), (unsigned)B), dim3(256), 0, st, N, x64, e,
                           bat(H, m_rep, N));
    ph_resid64<T>(H, st, m_rep, x64, b64, r, FIN_REP_DONE, m_it, maxit);
    if (H.singular) ph_project<T>(H, st, m_rep, r);
    // next preconditioned residual for the slots still iterating
    if (H.post > 0) {
        ph_vcycle<T>(H, st, m_it, 0, (const T*)r, z, (const T*)q);
    } else {
        ph_vcycle<T>(H, st, m_it, 0, (const T*)r, z);
        hipLaunchKernelGGL(k_dot3<T>, dim3(G, B), dim3(256), 0, st, N, (const T*)r, (const T*)z, (const T*)q, H.part,
                           bt);
    }
    ph_finish<T>(H, st, m_it, FIN_DOT3, H.singular);
}

int read_scal(const Hierarchy& H, hipStream_t st) {
    OPF_HIP(hipMemcpyAsync(H.hscal, H.scal, sizeof(double) * GMG_SSTRIDE * H.B, hipMemcpyDeviceToHost, st));
    OPF_HIP(hipStreamSynchronize(st));
    return OPF_OK;
}

// Solves of every slot of H (vectors slot-major) in Krylov precision T: flexible CG
// with reliable updates (FP32: FP64 update whenever the recurrence residual fell by
// the slot's threshold; FP64: at the convergence test only), stopped on the device
// per slot. Two iterations are replayed as a graph per host poll; a stopped slot
// skips the work inside them.
template <typename T>
static int rel_solve(opf_ctx* ctx, Hierarchy& H, const double* b, double* x, double tol, int maxit, int verbose) {
    hipStream_t s = ctx->stream;
    auto all_done = [&]() {
        for (int q = 0; q < H.B; ++q)
            if (H.hscal[(size_t)q * GMG_SSTRIDE + S_DONE] == 0.0) return false;
        return true;
    };
    hipLaunchKernelGGL(k_seti, dim3(1), dim3(64), 0, s, H.mk + MK_NEW * H.B, H.B, 1);
    ph_rel_start<T>(H, s, x, b, tol * tol, rel_delta2<T>(), false);
    OPF_TRY(read_scal(H, s));
    if (all_done()) return OPF_OK;
    if (!H.iter_exec || H.iter_x != x || H.iter_b != b || H.iter_maxit != maxit) {
        if (H.iter_exec) OPF_HIP(hipGraphExecDestroy(H.iter_exec));
        if (H.iter_graph) OPF_HIP(hipGraphDestroy(H.iter_graph));
        H.iter_exec = nullptr;
        H.iter_graph = nullptr;
        OPF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        ph_rel_iter<T>(H, s, x, b, 0, maxit);
        ph_rel_iter<T>(H, s, x, b, 1, maxit);
        OPF_HIP(hipStreamEndCapture(s, &H.iter_graph));
        OPF_HIP(hipGraphInstantiate(&H.iter_exec, H.iter_graph, nullptr, nullptr, 0));
        H.iter_x = x;
        H.iter_b = b;
        H.iter_maxit = maxit;
    }
    for (int it = 0; it < maxit; it += 2) {
        OPF_HIP(hipGraphLaunch(H.iter_exec, s));
        OPF_TRY(read_scal(H, s));
        if (verbose)
            fprintf(stderr, "[gmg] it %d updates %d relres %.3e\n", (int)H.hscal[S_ITS], (int)H.hscal[S_NREP],
                    std::sqrt(H.hscal[S_RR] / H.hscal[S_BB]));
        if (all_done()) break;
    }
    return OPF_OK;
}

int gmg_solve_slots(opf_ctx* ctx, Hierarchy& H, const double* b, double* x, double tol, int maxit,
                    opf_solve_state* st, int verbose) {
    hipStream_t s = ctx->stream;
    if (H.precision == OPF_PREC_MIXED) OPF_TRY(rel_solve<float>(ctx, H, b, x, tol, maxit, verbose));
    else OPF_TRY(rel_solve<double>(ctx, H, b, x, tol, maxit, verbose));
    std::vector<int> its(H.B), reps(H.B);
    for (int q = 0; q < H.B; ++q) {
        its[q] = (int)H.hscal[(size_t)q * GMG_SSTRIDE + S_ITS];
        reps[q] = (int)H.hscal[(size_t)q * GMG_SSTRIDE + S_NREP];
    }
    // final true residual in FP64
    ph_resid64<double>(H, s, nullptr, x, b, H.r64, FIN_RR);
    OPF_TRY(read_scal(H, s));
    OPF_HIP(hipGetLastError());
    for (int q = 0; q < H.B; ++q) {
        const double rr = H.hscal[(size_t)q * GMG_SSTRIDE + S_RR], bb = H.hscal[(size_t)q * GMG_SSTRIDE + S_BB];
        opf_solve_state& o = st[q];
        o.niter = its[q];
        o.nrefine = reps[q];
        o.abserr = std::sqrt(rr);
        o.relerr = bb > 0 ? std::sqrt(rr / bb) : 0.0;
        o.converged = o.relerr <= tol * (1 + 1e-12) || bb == 0;
        if (verbose) fprintf(stderr, "[gmg] slot %d iters %d refine %d relres %.3e\n", q, o.niter, o.nrefine, o.relerr);
    }
    return OPF_OK;
}

int gmg_solve(opf_ctx* ctx, Hierarchy& H, const double* b, double* x, double tol, int maxit, opf_solve_state* st,
              int verbose) {
    OPF_CHECK(H.B == 1, OPF_ERR_ARG, "gmg_solve drives single-slot hierarchies");
    return gmg_solve_slots(ctx, H, b, x, tol, maxit, st, verbose);
}

int gmg_vcycle64(opf_ctx* ctx, Hierarchy& H, const double* r, double* z) {
    OPF_CHECK(H.precision == OPF_PREC_FP64, OPF_ERR_ARG, "gmg_vcycle64 needs an FP64 hierarchy");
    ph_vcycle<double>(H, ctx->stream, nullptr, 0, r, z);
    OPF_HIP(hipGetLastError());
    return OPF_OK;
}

int gmg_apply64(opf_ctx* ctx, Hierarchy& H, const double* x, double* y) {
    GOp<double> o = make_op<double>(H, 0);
    const long long N = H.host[0].N;
    if (use_zp(H, 0) && H.B == 1) {
        zp_launch<double, ZP_APPLY>(H, 0, ctx->stream, bat(H, nullptr, N), x, nullptr, nullptr, nullptr, y,
                                    (double*)nullptr, nullptr);
        OPF_HIP(hipGetLastError());
        return OPF_OK;
    }
    GMG_DISPATCH(H.var, H.dim, hipLaunchKernelGGL((k_apply<double, V_, D_>), dim3(nblk(N)), dim3(256), 0,
                                                  ctx->stream, o, x, y));
    OPF_HIP(hipGetLastError());
    return OPF_OK;
}

// explicit instantiations used by the serving loop
template void ph_rel_start<float>(const Hierarchy&, hipStream_t, double*, const double*, double, double, bool);
template void ph_rel_start<double>(const Hierarchy&, hipStream_t, double*, const double*, double, double, bool);
template void ph_rel_iter<float>(const Hierarchy&, hipStream_t, double*, const double*, int, int);
template void ph_rel_iter<double>(const Hierarchy&, hipStream_t, double*, const double*, int, int);

}  // namespace opf

// ------------------------------------------------------------------ C ABI
using namespace opf;

struct opf_gmg {
    opf_ctx* ctx;
    Hierarchy H;
    opf_poisson_desc d;
};

static int poisson_level(const opf_poisson_desc* d, GLevelHost& L, int per[3], double alpha[3][2]) {
    OPF_CHECK(d && (d->dim == 2 || d->dim == 3), OPF_ERR_ARG, "gmg: dim must be 2 or 3");
    L = GLevelHost{};
    for (int q = 0; q < 3; ++q) L.n[q] = q < d->dim ? d->n[q] : 1;
    L.N = (long long)L.n[0] * L.n[1] * L.n[2];
    L.var = false;
    L.sigma = d->sigma;
    for (int q = 0; q < 3; ++q) {
        per[q] = 0;
        alpha[q][0] = alpha[q][1] = 1.0;
        L.a[q] = 0.0;
        if (q >= d->dim) continue;
        OPF_CHECK(d->n[q] >= 2 && d->h[q] > 0, OPF_ERR_ARG, "gmg: bad n/h in dim %d", q);
        L.a[q] = 1.0 / (d->h[q] * d->h[q]);
        for (int s = 0; s < 2; ++s) {
            int t = d->bc[q][s];
            OPF_CHECK(t == OPF_BC_DIRC || t == OPF_BC_NEUM || t == OPF_BC_PERIODIC, OPF_ERR_UNSUPPORTED,
                      "gmg: bc %d unsupported", t);
            alpha[q][s] = t == OPF_BC_DIRC ? -1.0 : 1.0;
        }
        per[q] = d->bc[q][0] == OPF_BC_PERIODIC;
        OPF_CHECK(per[q] == (d->bc[q][1] == OPF_BC_PERIODIC), OPF_ERR_ARG, "gmg: periodic on one side only");
    }
    return OPF_OK;
}

extern "C" size_t opf_gmg_footprint(const opf_poisson_desc* d, int precision) {
    if (!d) return 0;
    long long N = 1;
    for (int q = 0; q < d->dim; ++q) N *= d->n[q];
    size_t ts = precision == OPF_PREC_MIXED ? 4 : 8;
    double lv = d->dim == 3 ? 8.0 / 7.0 : 4.0 / 3.0;
    // 3 vectors per level (x, b, t) + 5/6 Krylov vectors + FP64 residual
    return (size_t)(N * ts * (3 * lv + 6) + N * 8) + (1 << 20);
}

extern "C" int opf_gmg_create(opf_ctx* ctx, const opf_poisson_desc* d, int precision, opf_gmg** out) {
    OPF_CHECK(ctx && d && out, OPF_ERR_ARG, "opf_gmg_create: null arg");
    OPF_CHECK(precision == OPF_PREC_FP64 || precision == OPF_PREC_MIXED, OPF_ERR_ARG, "bad precision");
    GLevelHost L;
    int per[3];
    double alpha[3][2];
    OPF_TRY(poisson_level(d, L, per, alpha));
    auto* g = new opf_gmg();
    g->ctx = ctx;
    g->d = *d;
    int rc = gmg_build(ctx, g->H, L, per, alpha, precision);
    if (rc != OPF_OK) {
        gmg_free(g->H);
        delete g;
        return rc;
    }
    *out = g;
    return OPF_OK;
}

extern "C" int opf_gmg_solve(opf_gmg* g, const double* b, double* x, const opf_solve_params* p,
                             opf_solve_state* st) {
    OPF_CHECK(g && b && x && st, OPF_ERR_ARG, "opf_gmg_solve: null arg");
    double tol = p && p->tol > 0 ? p->tol : 1e-6;
    int maxit = p && p->max_iter > 0 ? p->max_iter : 200;
    if (p && p->pre_smooth > 0) g->H.pre = p->pre_smooth;
    if (p && p->post_smooth > 0) g->H.post = p->post_smooth;
    return gmg_solve(g->ctx, g->H, b, x, tol, maxit, st, p ? p->verbose : 0);
}

extern "C" int opf_gmg_apply(opf_gmg* g, const double* x, double* y) {
    OPF_CHECK(g && x && y, OPF_ERR_ARG, "opf_gmg_apply: null arg");
    return gmg_apply64(g->ctx, g->H, x, y);
}

extern "C" int opf_gmg_destroy(opf_gmg* g) {
    if (!g) return OPF_OK;
    hipStreamSynchronize(g->ctx->stream);
    gmg_free(g->H);
    delete g;
    return OPF_OK;
}
