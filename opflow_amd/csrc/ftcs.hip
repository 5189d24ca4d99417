// FTCS heat update fast path:  u = u + coef * (d2x(u) + d2y(u) [+ d2z(u)])
// (examples/FTCS2D/FTCS-OMP.cpp:26, D2SecondOrderCentered.hpp:160-171,
//  FieldAssigner.hpp:26-86). One launch per step reads u^n (interior + ghosts)
// and writes u^{n+1} together with the ghost values updatePadding would give
// it, so the field is complete after every launch. On a uniform mesh the
// reference's dx_l = dx_r = dx_c = h exactly, so D2 = ((r-c)/h - (c-l)/h)/h;
// when h is a power of two the divisions are replaced by exact multiplications
// by 1/h (same bits).
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "fastdiv.hpp"
#include "opf_internal.hpp"

using namespace opf;

namespace {

// storage index of cell q (interior origin 0, one ghost layer); z is padded only in 3-D
template <int DIM>
__host__ __device__ __forceinline__ long long ftcs_index(int e0, int e1, const int q[3]) {
    return (q[0] + 1) + (long long)e0 * ((DIM >= 2 ? q[1] + 1 : 0) + (long long)e1 * (DIM == 3 ? q[2] + 1 : 0));
}

template <typename T>
struct FtcsArgs {
    int n[3];       // interior cells
    int e0, e1;     // storage extents (n+2)
    T coef;
    T coefh;        // coef / h^2 (FTCS_FACTORED: exact, h a power of two)
    T hh[3];        // 1/h (FTCS_MUL, FTCS_FACTORED) or h (FTCS_DIV)
    int per[3];
    int rtype[3][2];  // ghost rule per dim/side: 1 dirichlet, 2 neumann
    T rc1[3][2], rc2[3][2], rc3[3][2];
};

// Arithmetic modes of the update (all give the reference's bits):
//  FTCS_DIV      D2 = ((r-c)/h - (c-l)/h)/h per dimension, as D2SecondOrderCentered computes it;
//  FTCS_MUL      h a power of two: the divisions become multiplications by 1/h (exact);
//  FTCS_FACTORED h a power of two and equal in every dimension. Scaling by a power of two
//                commutes with rounding (no overflow: |u| h^-2 below the largest finite
//                value), so ((r-c)*k - (c-l)*k)*k = ((r-c) - (c-l)) * k^2 exactly, the sum
//                over dimensions keeps the common factor k^2, and coef * (k^2 S) =
//                (coef k^2) * S with coef k^2 exact: u + coefh * (Dx + Dy [+ Dz]) with
//                D = (r-c) - (c-l) -- 10 operations per 3-D cell instead of 19.
enum { FTCS_DIV = 0, FTCS_MUL = 1, FTCS_FACTORED = 2 };

template <typename T, int MODE>
__device__ __forceinline__ T d2(T l, T c, T r, T hh) {
    if constexpr (MODE != FTCS_DIV) return ((r - c) * hh - (c - l) * hh) * hh;
    else return ((r - c) / hh - (c - l) / hh) / hh;
}

// u^{n+1} at one cell from its value and its x / y (/ z) neighbours
template <typename T, int MODE>
__device__ __forceinline__ T ftcs2(T c, T xl, T xr, T yl, T yr, const FtcsArgs<T>& a) {
    if constexpr (MODE == FTCS_FACTORED) return c + a.coefh * (((xr - c) - (c - xl)) + ((yr - c) - (c - yl)));
    else return c + a.coef * (d2<T, MODE>(xl, c, xr, a.hh[0]) + d2<T, MODE>(yl, c, yr, a.hh[1]));
}
template <typename T, int MODE>
__device__ __forceinline__ T ftcs3(T c, T xl, T xr, T yl, T yr, T zl, T zr, const FtcsArgs<T>& a) {
    if constexpr (MODE == FTCS_FACTORED) {
        T s = ((xr - c) - (c - xl)) + ((yr - c) - (c - yl));
        s = s + ((zr - c) - (c - zl));
        return c + a.coefh * s;
    } else {
        T s = d2<T, MODE>(xl, c, xr, a.hh[0]) + d2<T, MODE>(yl, c, yr, a.hh[1]);
        s = s + d2<T, MODE>(zl, c, zr, a.hh[2]);
        return c + a.coef * s;
    }
}

template <typename T>
__device__ __forceinline__ T rule(const FtcsArgs<T>& a, int d, int s, T v) {
    int t = a.rtype[d][s];
    if (t == 1) return (a.rc1[d][s] * v - a.rc2[d][s]) / a.rc3[d][s];
    if (t == 2) return v + a.rc1[d][s];
    return v;
}

// Write the ghost images whose source is interior cell p (value v), in the
// closed form of updatePaddingImpl_final for one ghost layer.
template <typename T, int DIM>
__device__ __forceinline__ void write_images(T* __restrict__ out, const FtcsArgs<T>& a, const int p[3], T v) {
    int lo[3], hi[3];  // is p the source of the lower / upper ghost of dim d
    bool any = false;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        if (a.per[d]) {
            lo[d] = p[d] == a.n[d] - 1;
            hi[d] = p[d] == 0;
        } else {
            lo[d] = p[d] == 0;
            hi[d] = p[d] == a.n[d] - 1;
        }
        any |= lo[d] | hi[d];
    }
    if (!any) return;
    constexpr int NC = DIM == 1 ? 3 : DIM == 2 ? 9 : 27;
    for (int c = 1; c < NC; ++c) {
        int ch[3] = {c % 3, (c / 3) % 3, c / 9};
        int q[3] = {p[0], p[1], p[2]};
        T w = v;
        bool ok = true;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            if (ch[d] == 0) continue;
            int s = ch[d] - 1;
            if (!(s == 0 ? lo[d] : hi[d])) { ok = false; break; }
            q[d] = s == 0 ? -1 : a.n[d];
            if (!a.per[d]) w = rule(a, d, s, w);
        }
        if (!ok) continue;
        out[ftcs_index<DIM>(a.e0, a.e1, q)] = w;
    }
}

// 2-D: one cell per thread, block 64 x 4.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void ftcs2d_kernel(const T* __restrict__ in, T* __restrict__ out, FtcsArgs<T> a) {
    int i = blockIdx.x * 64 + threadIdx.x;
    int j = blockIdx.y * 4 + threadIdx.y;
    if (i >= a.n[0] || j >= a.n[1]) return;
    long long l = (i + 1) + (long long)a.e0 * (j + 1);
    T c = in[l];
    T v = ftcs2<T, MODE>(c, in[l - 1], in[l + 1], in[l - a.e0], in[l + a.e0], a);
    out[l] = v;
    int p[3] = {i, j, 0};
    write_images<T, 2>(out, a, p, v);
}


// 2-D temporal blocking: one workgroup advances a TS x TS output tile by `nt`
// (<= TB) steps inside LDS, from a region padded by TB cells per side (clamped
// to the ghost layer on non-periodic sides, wrapped on periodic ones). Every
// step recomputes all region cells whose stencil lies inside the region; the
// values at distance < s from an open region edge go stale, which never reaches
// the tile. Ghost cells inside the region are refilled from the fresh interior
// with the BC rules, x first then y (the corner composition of updatePadding).
// Same arithmetic per cell and step as ftcs2d_kernel, so the bits are equal.
template <int TS, int TB>
struct Ftcs2dTile {
    static constexpr int W = TS + 2 * TB;
};

template <typename T, int MODE, int TS, int TB>
__global__ __launch_bounds__(256) void ftcs2d_tb_kernel(const T* __restrict__ in, T* __restrict__ out, FtcsArgs<T> a,
                                                        int nt) {
    constexpr int W = Ftcs2dTile<TS, TB>::W;
    __shared__ T buf[2][W][W];
    const int tilesx = (a.n[0] + TS - 1) / TS;
    const int tb = (int)xcd_block(blockIdx.x, gridDim.x);  // x-neighbouring tiles share an XCD's L2
    const int x0 = (tb % tilesx) * TS, y0 = (tb / tilesx) * TS;
    // region origin in global cell coordinates; non-periodic sides clamp to the ghost layer
    int lo[2], hi[2];
    const int org[2] = {x0, y0};
    for (int d = 0; d < 2; ++d) {
        lo[d] = org[d] - TB;
        hi[d] = org[d] + TS + TB;
        if (!a.per[d]) {
            lo[d] = max(lo[d], -1);
            hi[d] = min(hi[d], a.n[d] + 1);
        }
    }
    const int wx = hi[0] - lo[0], wy = hi[1] - lo[1];
    auto gidx = [&](int gx, int gy) -> long long {
        if (a.per[0]) gx = ((gx % a.n[0]) + a.n[0]) % a.n[0];
        if (a.per[1]) gy = ((gy % a.n[1]) + a.n[1]) % a.n[1];
        return (gx + 1) + (long long)a.e0 * (gy + 1);
    };
    // positions walk the full W x W square (a compile-time divisor); clamped regions skip the rest
    for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
        const int lx = t % W, ly = t / W;
        if (lx >= wx || ly >= wy) continue;
        buf[0][ly][lx] = in[gidx(lo[0] + lx, lo[1] + ly)];
    }
    __syncthreads();
    int cur = 0;
    for (int s = 0; s < nt; ++s) {
        const int nxt = 1 - cur;
        // interior cells with all four neighbours inside the region
        for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
            const int lx = t % W, ly = t / W;
            if (lx >= wx || ly >= wy) continue;
            int gx = lo[0] + lx, gy = lo[1] + ly;
            bool inx = a.per[0] || (gx >= 0 && gx < a.n[0]);
            bool iny = a.per[1] || (gy >= 0 && gy < a.n[1]);
            if (!inx || !iny || lx == 0 || ly == 0 || lx == wx - 1 || ly == wy - 1) continue;
            T c = buf[cur][ly][lx];
            T v = ftcs2<T, MODE>(c, buf[cur][ly][lx - 1], buf[cur][ly][lx + 1], buf[cur][ly - 1][lx],
                                 buf[cur][ly + 1][lx], a);
            buf[nxt][ly][lx] = v;
        }
        __syncthreads();
        // ghosts of non-periodic x sides (rows inside the y interior), then y sides (all
        // columns); only tiles whose region reaches a non-periodic side have any (a
        // condition uniform over the workgroup, so the barriers stay uniform)
        if (!a.per[0] && (lo[0] == -1 || hi[0] == a.n[0] + 1)) {
            for (int ly = threadIdx.x; ly < wy; ly += blockDim.x) {
                int gy = lo[1] + ly;
                if (!a.per[1] && (gy < 0 || gy >= a.n[1])) continue;
                if (lo[0] == -1) buf[nxt][ly][0] = rule(a, 0, 0, buf[nxt][ly][1]);
                if (hi[0] == a.n[0] + 1) buf[nxt][ly][wx - 1] = rule(a, 0, 1, buf[nxt][ly][wx - 2]);
            }
            __syncthreads();
        }
        if (!a.per[1] && (lo[1] == -1 || hi[1] == a.n[1] + 1)) {
            for (int lx = threadIdx.x; lx < wx; lx += blockDim.x) {
                if (lo[1] == -1) buf[nxt][0][lx] = rule(a, 1, 0, buf[nxt][1][lx]);
                if (hi[1] == a.n[1] + 1) buf[nxt][wy - 1][lx] = rule(a, 1, 1, buf[nxt][wy - 2][lx]);
            }
            __syncthreads();
        }
        cur = nxt;
    }
    // write the tile and the ghost images it is the source of
    for (int t = threadIdx.x; t < TS * TS; t += blockDim.x) {
        int i = x0 + t % TS, j = y0 + t / TS;
        if (i >= a.n[0] || j >= a.n[1]) continue;
        T v = buf[cur][j - lo[1]][i - lo[0]];
        int p[3] = {i, j, 0};
        out[ftcs_index<2>(a.e0, a.e1, p)] = v;
        write_images<T, 2>(out, a, p, v);
    }
}

// 3-D single step, LDS-tiled: a 64 x NTY workgroup owns a 64 x (NTY * RY) x-y
// tile (each thread RY rows) and marches ZC planes in z. The current plane of
// the tile and its one-cell x/y halo sit in LDS (double-buffered, one barrier
// per plane); each thread keeps its cells' z-1 / z / z+1 values in registers
// and loads plane k+2 while plane k is computed. The input's ghost layer
// supplies the halo of boundary tiles, so loads never wrap; the kernel writes
// interior cells only and the host loop refills the output's ghosts before
// they are read. Per cell: d2x + d2y, then + d2z, then u + coef * s -- the
// arithmetic every FTCS kernel here uses, so all give the same bits.
// Launch shape swept at 512^3 on MI355X (64 x 8 threads, 2 rows each, 8 planes).
constexpr int FT3_TX = 64, FT3_NTY = 8, FT3_RY = 2, FT3_TY = FT3_NTY * FT3_RY, FT3_ZC = 8;

template <typename T, int MODE, int ZC, int NTY, int RY>
__global__ __launch_bounds__(64 * NTY) void ftcs3d_lds_kernel(const T* __restrict__ in, T* __restrict__ out,
                                                              FtcsArgs<T> a) {
    constexpr int TY = NTY * RY;
    static_assert(128 + 2 * TY <= 64 * NTY, "one halo element per thread at most");
    __shared__ T lds[2][TY + 2][FT3_TX + 2];
    const int tx = threadIdx.x, ty = threadIdx.y, t = ty * 64 + tx;
    // tile of this block: the flat id through xcd_block, so that each XCD's L2 holds a
    // contiguous slab whose x / y halo lines its own tiles share
    const uint32_t L = xcd_block(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                                 gridDim.x * gridDim.y * gridDim.z);
    const int x0 = (int)(L % gridDim.x) * FT3_TX, y0 = (int)((L / gridDim.x) % gridDim.y) * TY,
              k0 = (int)(L / (gridDim.x * gridDim.y)) * ZC;
    const int kend = min(k0 + ZC, a.n[2]);
    const long long sy = a.e0, sz = (long long)a.e0 * a.e1;
    const int i = x0 + tx;
    // storage index of (x, y, k) (ghost layer at -1 and n)
    auto at = [&](int x, int y, int k) { return (long long)(x + 1) + sy * (y + 1) + sz * (k + 1); };
    // valid storage positions: interior and the ghost layer
    auto ok = [&](int x, int y) { return x >= -1 && x <= a.n[0] && y >= -1 && y <= a.n[1]; };
    // this thread's halo element: rows y0 - 1 and y0 + TY (threads 0..127), then the
    // columns x0 - 1 and x0 + TX over the tile's rows; corners are not read
    int hx = 0, hy = 0, hlx = 0, hly = 0;
    if (t < 128) {
        hx = x0 + (t & 63);
        hy = t < 64 ? y0 - 1 : y0 + TY;
        hlx = (t & 63) + 1;
        hly = t < 64 ? 0 : TY + 1;
    } else if (t < 128 + 2 * TY) {
        const int q = t - 128, side = q / TY, r = q % TY;
        hx = side == 0 ? x0 - 1 : x0 + FT3_TX;
        hy = y0 + r;
        hlx = side == 0 ? 0 : FT3_TX + 1;
        hly = r + 1;
    }
    const bool hl = t < 128 + 2 * TY && ok(hx, hy);
    bool own[RY];
    T zm[RY], zc[RY], zp[RY], zq[RY];
    T hp = T(0), hq = T(0);
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int j = y0 + ty * RY + r;
        own[r] = ok(i, j);
        zm[r] = own[r] ? in[at(i, j, k0 - 1)] : T(0);
        zc[r] = own[r] ? in[at(i, j, k0)] : T(0);
        zp[r] = own[r] ? in[at(i, j, k0 + 1)] : T(0);  // k0 + 1 <= n[2]: the ghost plane at most
        lds[0][ty * RY + r + 1][tx + 1] = zc[r];
    }
    if (hl) {
        lds[0][hly][hlx] = in[at(hx, hy, k0)];
        hp = in[at(hx, hy, k0 + 1)];
    }
    __syncthreads();
    int cur = 0;
    for (int k = k0; k < kend; ++k) {
        // plane k + 2 is loaded one step ahead of its use (only while a next step needs it)
        if (k + 1 < kend) {
#pragma unroll
            for (int r = 0; r < RY; ++r) zq[r] = own[r] ? in[at(i, y0 + ty * RY + r, k + 2)] : T(0);
            if (hl) hq = in[at(hx, hy, k + 2)];
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int ly = ty * RY + r + 1, j = y0 + ly - 1;
            if (i >= a.n[0] || j >= a.n[1]) continue;
            const T c = zc[r];
            out[at(i, j, k)] = ftcs3<T, MODE>(c, lds[cur][ly][tx], lds[cur][ly][tx + 2], lds[cur][ly - 1][tx + 1],
                                              lds[cur][ly + 1][tx + 1], zm[r], zp[r], a);
        }
        const int nxt = cur ^ 1;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            lds[nxt][ty * RY + r + 1][tx + 1] = zp[r];
            zm[r] = zc[r];
            zc[r] = zp[r];
            zp[r] = zq[r];
        }
        if (hl) lds[nxt][hly][hlx] = hp;
        hp = hq;
        __syncthreads();
        cur = nxt;
    }
}

// 3-D temporal blocking, K steps per launch. A workgroup owns a TX x TY column
// of the x-y plane and the z chunk [z0, z0 + zc) and streams the planes of u^n
// with a K-cell halo in x and y. Every thread owns fixed positions of the
// (TX+2K) x (TY+2K) region and keeps, for each, the last three planes of
// levels 0 .. K-1 (u^n .. u^{n+K-1}) in registers, so the centre and the z
// neighbours never go through LDS; LDS holds only the current plane of each
// level (double-buffered), read for the x and y neighbours. Step zz stores
// level-0 plane zz and computes level L at plane zz - L (L = 1 .. K) on the
// region shrunk by L cells; level K is exactly the tile and goes straight to
// global memory. One barrier per step separates a step's LDS writes from the
// next step's reads (levels read the plane the level below wrote one step
// earlier). Level-0 planes are loaded three planes ahead; the unroll by 3 keeps
// the history slots static; every load is unconditional so the compiler keeps
// counted load waits.
// Periodic dimensions wrap the halo to interior cells (the input's ghosts are
// never read). Dirichlet / Neumann dimensions (ALLPER = false) clamp the halo to
// the ghost layer: level 0 reads the input's ghosts (refilled by the host before
// the launch), and the ghosts of the intermediate levels 1 .. K-1 are rebuilt
// from the level's fresh boundary cells with the BC rules, as updatePadding
// does after every step: x / y face ghosts in LDS at the start of the next step
// (one extra barrier, only in workgroups whose region reaches such a side), z
// ghost planes in the register history when the level computes plane 0 or
// reaches plane n. The 7-point stencil reads face ghosts only, so edge and
// corner ghosts are never needed; region cells beyond the ghost layer hold
// values that never reach a valid cell. The output's ghosts are left stale and
// refilled by the host loop before anything reads them. Per cell and step the
// arithmetic is that of ftcs3d_lds_kernel, so the result has the same bits as K
// single steps.
template <typename T, int MODE, int TX, int TY, int NT, int K, bool XCD, bool ALLPER>
__global__ __launch_bounds__(NT) void ftcs3d_tb_kernel(const T* __restrict__ in, T* __restrict__ out, FtcsArgs<T> a,
                                                       int zc) {
    static_assert(K >= 2, "one step per launch is ftcs3d_lds_kernel");
    constexpr int RX = TX + 2 * K, RY = TY + 2 * K, RP = RX * RY;
    constexpr int NL = (RP + NT - 1) / NT;
    constexpr int PF = 3;
    __shared__ T lds[K][2][NL * NT];  // current plane of levels 0 .. K-1
    const int tilesx = (a.n[0] + TX - 1) / TX, tilesy = (a.n[1] + TY - 1) / TY;
    int b = XCD ? (int)xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;  // neighbouring tiles share an XCD's L2
    const int x0 = (b % tilesx) * TX;
    b /= tilesx;
    const int y0 = (b % tilesy) * TY;
    const int z0 = (b / tilesy) * zc;
    const int z1 = min(z0 + zc, a.n[2]);
    const long long sz = (long long)a.e0 * a.e1;
    const int tid = threadIdx.x;
    // this thread's region positions: input offset, region coordinates, output offset;
    // gh: 0 interior, 1 / 2 the low / high x face ghost, 3 / 4 the low / high y face ghost
    int off[NL], lxm[NL], lym[NL], gh[NL];
    bool valid[NL], inr[NL];
    long long oo[NL];
#pragma unroll
    for (int m = 0; m < NL; ++m) {
        const int e = tid + m * NT;
        valid[m] = e < RP;
        const int ee = valid[m] ? e : 0;
        const int lx = ee % RX, ly = ee / RX;
        lxm[m] = lx;
        lym[m] = ly;
        int gx = x0 - K + lx, gy = y0 - K + ly;
        gh[m] = 0;
        if (ALLPER || a.per[0]) gx = (gx % a.n[0] + a.n[0]) % a.n[0];
        if (ALLPER || a.per[1]) gy = (gy % a.n[1] + a.n[1]) % a.n[1];
        if (!ALLPER) {
            const bool inx = gx >= 0 && gx < a.n[0], iny = gy >= 0 && gy < a.n[1];
            if (iny && (gx == -1 || gx == a.n[0])) gh[m] = gx < 0 ? 1 : 2;
            if (inx && (gy == -1 || gy == a.n[1])) gh[m] = gy < 0 ? 3 : 4;
            gx = min(max(gx, -1), a.n[0]);  // clamped: inside the storage
            gy = min(max(gy, -1), a.n[1]);
        }
        off[m] = (gx + 1) + a.e0 * (gy + 1);
        const int i = x0 + lx - K, j = y0 + ly - K;
        inr[m] = i < a.n[0] && j < a.n[1];
        oo[m] = (long long)(i + 1) + (long long)a.e0 * (j + 1);
    }
    // position m takes part in level L (region shrunk by L cells)
    auto at_level = [&](int m, int L) {
        return valid[m] && lxm[m] >= L && lxm[m] <= RX - 1 - L && lym[m] >= L && lym[m] <= RY - 1 - L;
    };
    // the region reaches a non-periodic x / y side: face ghosts to rebuild (uniform over the block)
    const bool edge = !ALLPER && ((!a.per[0] && (x0 - K < 0 || x0 + TX + K > a.n[0])) ||
                                  (!a.per[1] && (y0 - K < 0 || y0 + TY + K > a.n[1])));
    const bool zper = ALLPER || a.per[2];
    const int zb = z0 - K, ze = z1 + K;  // level-0 planes zb .. ze - 1
    T pre[PF][NL], hist[K][NL][3];       // history slot of plane p: (p - zb) % 3
    auto load = [&](T(&r)[NL], int z) {
        if (zper) z = (z % a.n[2] + a.n[2]) % a.n[2];
        else z = min(max(z, -1), a.n[2]);
        const T* p = in + sz * (z + 1);
#pragma unroll
        for (int m = 0; m < NL; ++m) r[m] = p[off[m]];
    };
#pragma unroll
    for (int u = 0; u < PF; ++u) load(pre[u], min(zb + u, ze - 1));
    for (int zz0 = zb; zz0 < ze; zz0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int zz = zz0 + u;  // (zz - zb) % 3 == u
            if (zz >= ze) break;     // uniform over the block
            __syncthreads();
            T* w0 = lds[0][zz & 1];
#pragma unroll
            for (int m = 0; m < NL; ++m) {
                hist[0][m][u] = pre[u][m];
                w0[tid + m * NT] = pre[u][m];
            }
            load(pre[u], min(zz + PF, ze - 1));
            if (!ALLPER && edge) {
                // face ghosts of the planes levels 1 .. K-1 wrote in the last step, from their
                // boundary cells (read after the barrier above), before the next level reads them
#pragma unroll
                for (int L = 1; L < K; ++L) {
                    if (zz - 1 < z0 - K + 2 * L) continue;  // level L had no plane yet (uniform)
                    T* P = lds[L][(zz - 1 - L) & 1];
#pragma unroll
                    for (int m = 0; m < NL; ++m) {
                        if (!gh[m] || !at_level(m, L)) continue;
                        const int e = tid + m * NT;
                        const int g = gh[m];
                        const int nb = g == 1 ? e + 1 : g == 2 ? e - 1 : g == 3 ? e + RX : e - RX;
                        P[e] = rule(a, g <= 2 ? 0 : 1, (g - 1) & 1, P[nb]);
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int L = 1; L <= K; ++L) {
                if (zz < z0 - K + 2 * L) break;  // plane zz - L of level L not computable yet (uniform)
                // level L-1 planes zz-L+1, zz-L, zz-L-1 sit in slots sp, sc, sm
                const int sp = ((u - L + 1) % 3 + 3) % 3, sc = ((u - L) % 3 + 3) % 3, sm = ((u - L - 1) % 3 + 3) % 3;
                const T* P = lds[L - 1][(zz - L) & 1];
                const int pz = zz - L;  // the plane level L computes
#pragma unroll
                for (int m = 0; m < NL; ++m) {
                    if (L < K ? !at_level(m, L) : !(at_level(m, K) && inr[m])) continue;
                    const int e = tid + m * NT;
                    const T c = hist[L - 1][m][sc];
                    T v = ftcs3<T, MODE>(c, P[e - 1], P[e + 1], P[e - RX], P[e + RX], hist[L - 1][m][sm],
                                         hist[L - 1][m][sp], a);
                    if (L < K) {
                        if (!zper && pz == a.n[2]) v = rule(a, 2, 1, hist[L][m][sm]);  // ghost plane n
                        hist[L][m][sc] = v;
                        if (!zper && pz == 0) hist[L][m][sm] = rule(a, 2, 0, v);  // ghost plane -1
                        lds[L][(zz - L) & 1][e] = v;
                    } else {
                        out[sz * (zz - K + 1) + oo[m]] = v;
                    }
                }
            }
        }
    }
}

// launch shapes: 2-D tiles of FT_TS2^2 cells advanced FT_TB2 steps per launch
// (swept at 1024^2..4096^2: 32x32 with 4 steps beat 2, 3, 5, 6, 8 and 10 steps
// and 16..64-cell tiles; 4096^2 FP64 288 -> 390 G cell-updates/s against 8 steps);
// grids with fewer than 512 such tiles use FT_TS2S^2 tiles and FT_TB2S steps, so
// that every CU gets several workgroups (256^2 FP64 x 1000 steps: 16x16/8 2.06 ms,
// 8x8/5 1.73 ms; 8x8 with 3..7 steps and 4x4 / 12x12 tiles were slower)
constexpr int FT_TS2 = 32, FT_TB2 = 4, FT_TS2S = 8, FT_TB2S = 5;

// steps per launch of the 2-D blocked kernel for an n[0] x n[1] grid
inline int tb2_steps(const int n[3]) {
    const int tiles = ((n[0] + FT_TS2 - 1) / FT_TS2) * ((n[1] + FT_TS2 - 1) / FT_TS2);
    return tiles < 512 ? FT_TB2S : FT_TB2;
}

// 3-D blocked launch shape per precision (swept at 512^3 on MI355X): TX x TY
// columns, NT threads, K steps per launch (a remainder of two steps takes one K = 2
// launch of the same tile, one step the single-step kernel); the z chunk is chosen per
// launch (ftcs_t)
template <typename T>
struct Ftcs3dCfg;
template <>
struct Ftcs3dCfg<double> {
    static constexpr int TX = 64, TY = 32, NT = 1024, K = 3;
    static constexpr bool XCD = true;
};
template <>
struct Ftcs3dCfg<float> {
    static constexpr int TX = 128, TY = 32, NT = 1024, K = 3;
    static constexpr bool XCD = true;
};

// Steps the next launch advances: the launch plan shared by the sweep and
// opf_ftcs_plan. tblock = 1 forces single steps.
int launch_steps(int dim, int dtype, const int n[3], const int per[3], long long plane, int remaining, int tblock) {
    if (tblock == 1) return 1;
    if (dim == 2) {
        const int tb = tb2_steps(n);
        if (remaining >= 2 && n[0] >= tb && n[1] >= tb) return std::min(tb, remaining);
        return 1;
    }
    const int k3 = dtype == OPF_F64 ? Ftcs3dCfg<double>::K : Ftcs3dCfg<float>::K;
    (void)per;
    if (dim == 3 && plane < (1LL << 31) && remaining >= 2) return std::min(k3, remaining);
    return 1;
}

bool is_pow2(double h) {
    if (!(h > 0)) return false;
    int e;
    double m = std::frexp(h, &e);
    return m == 0.5;
}

// ext_spare: null = the field's own second buffer (allocated on first use);
// otherwise a caller-provided device buffer of at least f->nbytes used as the
// second time level, with the result copied back into f->data at the end.
template <typename T>
int ftcs_t(opf_field* f, double coef, int nsteps, int tblock, void* ext_spare = nullptr) {
    const opf_field_desc& d = f->d;
    const opf_mesh* m = f->mesh;
    FtcsArgs<T> a{};
    bool pow2 = true;
    for (int q = 0; q < 3; ++q) {
        a.n[q] = q < d.dim ? d.local_end[q] - d.local_start[q] : 1;
        a.per[q] = 0;
    }
    a.e0 = d.extent[0];
    a.e1 = d.extent[1];
    a.coef = (T)coef;
    GhostInfo gi;
    OPF_TRY(build_ghost_info(f, &gi));
    for (int q = 0; q < d.dim; ++q) {
        OPF_CHECK(d.loc[q] == OPF_LOC_CENTER, OPF_ERR_UNSUPPORTED, "ftcs: centre location required");
        OPF_CHECK(d.padding == 1 && d.ext[q][0] == 1 && d.ext[q][1] == 1, OPF_ERR_UNSUPPORTED,
                  "ftcs: padding = ext = 1 required");
        OPF_CHECK(m->uniform[q], OPF_ERR_UNSUPPORTED, "ftcs: uniform mesh required");
        OPF_CHECK(d.assign_start[q] == d.local_start[q] && d.assign_end[q] == d.local_end[q] && d.local_start[q] == 0,
                  OPF_ERR_UNSUPPORTED, "ftcs: assignable range must be the whole interior");
        a.per[q] = gi.per[q];
        for (int s = 0; s < 2; ++s) {
            int t = d.bc_type[q][s];
            OPF_CHECK(t == OPF_BC_DIRC || t == OPF_BC_NEUM || t == OPF_BC_PERIODIC, OPF_ERR_UNSUPPORTED,
                      "ftcs: bc %d unsupported", t);
            const GhostRule& r = gi.rule[q][s][0];
            a.rtype[q][s] = r.type;
            a.rc1[q][s] = (T)r.c1;
            a.rc2[q][s] = (T)r.c2;
            a.rc3[q][s] = (T)r.c3;
        }
        double h = m->h[q];
        // D2 on a uniform mesh: dx_l = (h+h)*0.5 = h, dx_r = h, dx_c = h exactly
        pow2 &= is_pow2(h);
        a.hh[q] = (T)h;
    }
    bool same = true;
    for (int q = 1; q < d.dim; ++q) same &= m->h[q] == m->h[0];
    const int mode = !pow2 ? FTCS_DIV : same ? FTCS_FACTORED : FTCS_MUL;
    if (pow2) {
        for (int q = 0; q < d.dim; ++q) a.hh[q] = (T)(1.0 / m->h[q]);
        const T k2 = a.hh[0] * a.hh[0];  // exact: a power of two
        a.coefh = a.coef * k2;            // exact: a power-of-two scaling of the rounded coef
    }
    hipStream_t st = f->ctx->stream;
    void* cur = f->data;  // time level the next launch reads
    void* nxt = nullptr;  // time level it writes
    if (ext_spare) {
        // every cell the kernels do not write keeps the field's value
        OPF_HIP(hipMemcpyAsync(ext_spare, f->data, f->nbytes, hipMemcpyDeviceToDevice, st));
        nxt = ext_spare;
    } else {
        if (!f->spare) {
            OPF_HIP(hipMalloc(&f->spare, f->nbytes));
            f->spare_valid = false;
        }
        if (!f->spare_valid) {
            OPF_HIP(hipMemcpyAsync(f->spare, f->data, f->nbytes, hipMemcpyDeviceToDevice, st));
            f->spare_valid = true;
        }
        nxt = f->spare;
    }
    bool ghosts_stale = false;  // the 3-D blocked kernel writes interior cells only
    for (int it = 0; it < nsteps;) {
        const T* in = (const T*)cur;
        T* out = (T*)nxt;
        const int k = launch_steps(d.dim, d.dtype, a.n, a.per, (long long)a.e0 * a.e1, nsteps - it, tblock);
        if (ghosts_stale && !(d.dim == 3 && k > 1 && a.per[0] && a.per[1] && a.per[2])) {
            OPF_TRY(launch_ghost_fill(f, cur));
            ghosts_stale = false;
        }
        // launches the kernel instantiated for the arithmetic mode
        auto by_mode = [&](auto launch) {
            if (mode == FTCS_FACTORED) launch(std::integral_constant<int, FTCS_FACTORED>{});
            else if (mode == FTCS_MUL) launch(std::integral_constant<int, FTCS_MUL>{});
            else launch(std::integral_constant<int, FTCS_DIV>{});
        };
        if (d.dim == 2 && k > 1) {
            // nt = k <= the kernel's TB steps (a remainder runs the same tile shape)
            int tiles = ((a.n[0] + FT_TS2 - 1) / FT_TS2) * ((a.n[1] + FT_TS2 - 1) / FT_TS2);
            if (tb2_steps(a.n) == FT_TB2S) {
                // small grids: small tiles so that every CU gets several workgroups
                int ts = ((a.n[0] + FT_TS2S - 1) / FT_TS2S) * ((a.n[1] + FT_TS2S - 1) / FT_TS2S);
                by_mode([&](auto M) {
                    hipLaunchKernelGGL((ftcs2d_tb_kernel<T, decltype(M)::value, FT_TS2S, FT_TB2S>), dim3(ts), dim3(256),
                                       0, st, in, out, a, k);
                });
            } else {
                by_mode([&](auto M) {
                    hipLaunchKernelGGL((ftcs2d_tb_kernel<T, decltype(M)::value, FT_TS2, FT_TB2>), dim3(tiles), dim3(256),
                                       0, st, in, out, a, k);
                });
            }
        } else if (d.dim == 2) {
            dim3 grid((a.n[0] + 63) / 64, (a.n[1] + 3) / 4);
            by_mode([&](auto M) {
                hipLaunchKernelGGL((ftcs2d_kernel<T, decltype(M)::value>), grid, dim3(64, 4), 0, st, in, out, a);
            });
        } else if (d.dim == 3 && k > 1) {
            using C3 = Ftcs3dCfg<T>;
            const bool allper = a.per[0] && a.per[1] && a.per[2];
            // z chunk: periodic fields take the longest (each chunk recomputes 2K halo
            // planes) that still gives every CU a workgroup (one 147 KB workgroup fills a
            // CU). 512^3 FP64: 256 planes, 256 workgroups, one round: 750 -> 800 G
            // cell-updates/s against 64-plane chunks (four rounds); FP32 (128-wide tiles)
            // 128 planes: 1515 -> 1610. With Dirichlet / Neumann sides the workgroups that
            // rebuild ghosts take longer, and 64-plane chunks (more workgroups than CUs,
            // balanced by the dispatcher) stay faster (512^3 FP64: 376 vs 330 G/s)
            const int tiles = ((a.n[0] + C3::TX - 1) / C3::TX) * ((a.n[1] + C3::TY - 1) / C3::TY);
            int zc = allper ? 256 : 64;
            while (zc > 16 && (long long)tiles * ((a.n[2] + zc - 1) / zc) < f->ctx->cus) zc /= 2;
            const int nb = tiles * ((a.n[2] + zc - 1) / zc);
            static_assert(C3::K == 3, "remainder launches assume K = 3");
            by_mode([&](auto M) {
                constexpr int MD = decltype(M)::value;
                if (k == C3::K && allper)
                    hipLaunchKernelGGL((ftcs3d_tb_kernel<T, MD, C3::TX, C3::TY, C3::NT, C3::K, C3::XCD, true>),
                                       dim3(nb), dim3(C3::NT), 0, st, in, out, a, zc);
                else if (k == C3::K)
                    hipLaunchKernelGGL((ftcs3d_tb_kernel<T, MD, C3::TX, C3::TY, C3::NT, C3::K, C3::XCD, false>),
                                       dim3(nb), dim3(C3::NT), 0, st, in, out, a, zc);
                else if (allper)
                    hipLaunchKernelGGL((ftcs3d_tb_kernel<T, MD, C3::TX, C3::TY, C3::NT, 2, C3::XCD, true>),
                                       dim3(nb), dim3(C3::NT), 0, st, in, out, a, zc);
                else
                    hipLaunchKernelGGL((ftcs3d_tb_kernel<T, MD, C3::TX, C3::TY, C3::NT, 2, C3::XCD, false>),
                                       dim3(nb), dim3(C3::NT), 0, st, in, out, a, zc);
            });
            ghosts_stale = true;
        } else if (d.dim == 3) {
            dim3 grid((a.n[0] + FT3_TX - 1) / FT3_TX, (a.n[1] + FT3_TY - 1) / FT3_TY, (a.n[2] + FT3_ZC - 1) / FT3_ZC);
            by_mode([&](auto M) {
                hipLaunchKernelGGL((ftcs3d_lds_kernel<T, decltype(M)::value, FT3_ZC, FT3_NTY, FT3_RY>), grid,
                                   dim3(64, FT3_NTY), 0, st, in, out, a);
            });
            ghosts_stale = true;
        } else {
            set_error("ftcs: dim %d not supported by the fast path", d.dim);
            return OPF_ERR_UNSUPPORTED;
        }
        it += k;
        std::swap(cur, nxt);
    }
    if (ghosts_stale) OPF_TRY(launch_ghost_fill(f, cur));
    OPF_HIP(hipGetLastError());
    if (ext_spare) {
        if (cur != f->data) OPF_HIP(hipMemcpyAsync(f->data, cur, f->nbytes, hipMemcpyDeviceToDevice, st));
    } else {
        f->data = cur;
        f->spare = nxt;
    }
    return OPF_OK;
}

}  // namespace

extern "C" int opf_ftcs(opf_field* u, double coef, int nsteps) {
    OPF_CHECK(u && nsteps >= 0, OPF_ERR_ARG, "opf_ftcs: bad args");
    if (nsteps == 0) return OPF_OK;
    return u->d.dtype == OPF_F64 ? ftcs_t<double>(u, coef, nsteps, 0) : ftcs_t<float>(u, coef, nsteps, 0);
}

extern "C" int opf_ftcs_noalloc(opf_field* u, double coef, int nsteps) {
    OPF_CHECK(u && nsteps >= 0, OPF_ERR_ARG, "opf_ftcs_noalloc: bad args");
    OPF_CHECK(u->nbytes <= (size_t)OPF_FTCS_NOALLOC_MAX_BYTES, OPF_ERR_NOMEM,
              "opf_ftcs_noalloc: field storage of %zu bytes exceeds the %zu-byte device scratch", u->nbytes,
              (size_t)OPF_FTCS_NOALLOC_MAX_BYTES);
    OPF_CHECK(u->ctx->noalloc_scratch, OPF_ERR_HIP, "opf_ftcs_noalloc: no scratch address for device %d",
              u->ctx->device);
    if (nsteps == 0) return OPF_OK;
    void* s = u->ctx->noalloc_scratch;
    return u->d.dtype == OPF_F64 ? ftcs_t<double>(u, coef, nsteps, 0, s) : ftcs_t<float>(u, coef, nsteps, 0, s);
}

extern "C" int opf_ftcs_plan(const opf_field* u, int nsteps, int tblock, int* launches, int* max_steps) {
    OPF_CHECK(u && nsteps >= 0 && tblock >= 0 && launches && max_steps, OPF_ERR_ARG, "opf_ftcs_plan: bad args");
    const opf_field_desc& d = u->d;
    int n[3] = {1, 1, 1}, per[3] = {0, 0, 0};
    for (int q = 0; q < d.dim; ++q) {
        n[q] = d.local_end[q] - d.local_start[q];
        per[q] = d.bc_type[q][0] == OPF_BC_PERIODIC;
    }
    *launches = 0;
    *max_steps = 0;
    for (int it = 0; it < nsteps;) {
        int k = launch_steps(d.dim, d.dtype, n, per, (long long)d.extent[0] * d.extent[1], nsteps - it, tblock);
        ++*launches;
        *max_steps = std::max(*max_steps, k);
        it += k;
    }
    return OPF_OK;
}

extern "C" int opf_ftcs_blocked(opf_field* u, double coef, int nsteps, int tblock) {
    OPF_CHECK(u && nsteps >= 0 && tblock >= 0, OPF_ERR_ARG, "opf_ftcs_blocked: bad args");
    if (nsteps == 0) return OPF_OK;
    return u->d.dtype == OPF_F64 ? ftcs_t<double>(u, coef, nsteps, tblock) : ftcs_t<float>(u, coef, nsteps, tblock);
}
