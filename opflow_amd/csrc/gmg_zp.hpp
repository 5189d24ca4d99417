// Z-pipelined stencil passes over 3-D constant-coefficient multigrid levels
// (included by gmg.hip after its helpers; namespace opf).
//
// A workgroup owns a TX x TY column of the x-y plane and a chunk of planes
// [z0, z1). It streams the planes of the stencilled vector, each with a
// one-cell halo in x and y, through a 4-plane LDS ring: the loads of plane
// z + PF are in flight (registers) while plane z - 1 is computed from planes
// z - 2 .. z of the ring. Four slots let one barrier per plane separate the
// store of plane z from the last reads of plane z - 4.
//
// Boundary conditions are folded into the halo: a periodic side wraps and a
// Dirichlet / Neumann side stores the ghost value alpha * (boundary cell)
// (alpha = 1 - bfac: -1 Dirichlet, +1 Neumann). Then
//   (A v)_c = dint v_c - sum_faces a_d v_nb
// holds at every cell with one straight-line expression; only the
// smoother's diagonal differs at boundary cells:
//   D_c = dint - sum over the cell's boundary faces of a_d alpha.
// Every load is unconditional (padding lanes and lanes past a ragged edge read
// a valid cell), so the compiler keeps counted waits on the prefetch.
//
// Modes (one kernel template, fused epilogues):
//   ZP_JAC      xo = x + (w / D)(b - A x)
//   ZP_JAC_DOT3 the same + partials of (b, xo), (xo, q), sum(xo)
//   ZP_JAC2Z    two sweeps from zero: x1 = (w / D) b in the ring, xo = x1 + (w / D)(b - A x1)
//   ZP_SPMV_P   p = (z - zmean) + beta p_old in the ring; p_new = p, q = A p, partial (p, q)
//   ZP_RESID    r = b - A x (optional store, precision TO), partial r^2
//   ZP_RESTRICT coarse b = average over the 2 x 2 x 2 children of b - A x (even extents)
//   ZP_APPLY    y = A x
//   ZP_PJAC     prolongation fused into the first post-smoothing sweep: the ring holds
//               x_f + P x_c (trilinear, coarse ghosts as the separate prolongation), and
//               xo = that + (w / D)(b - A that); in1 is the coarse correction x_c

template <typename T>
struct ZpCfg;
template <>
struct ZpCfg<double> {
    static constexpr int TX = 32, TY = 16, NT = 256, PF = 2;
};
template <>
struct ZpCfg<float> {
    static constexpr int TX = 64, TY = 8, NT = 256, PF = 4;
};

struct ZpGrid {
    int tx, ty;  // tiles in x and y
    int zc;      // planes per chunk (even)
    int nch;     // chunks in z
    int blocks;  // tx * ty * nch
};

enum { ZP_JAC = 0, ZP_JAC_DOT3, ZP_JAC2Z, ZP_SPMV_P, ZP_RESID, ZP_RESTRICT, ZP_APPLY, ZP_PJAC };

template <typename T>
__device__ __forceinline__ T zp_alpha(const GOp<T>& o, int d, int s) {
    return (T)1 - o.bfac[d][s];
}

// in-grid cell whose value stands at coordinate g in [-1, n] of dimension d
// (times f: the ghost factor); coordinates past n (ragged tiles) clamp to a
// valid cell whose value is never used
template <typename T>
__device__ __forceinline__ int zp_mirror(const GOp<T>& o, int d, int g, int n, T& f) {
    if (g < 0) {
        if (o.per[d]) return g + n;
        f *= zp_alpha(o, d, 0);
        return 0;
    }
    if (g >= n) {
        if (g == n) {
            if (o.per[d]) return 0;
            f *= zp_alpha(o, d, 1);
        }
        return n - 1;
    }
    return g;
}

// boundary-face part of the diagonal of a cell at coordinate c (0 <= c < n)
template <typename T>
__device__ __forceinline__ T zp_dadj(const GOp<T>& o, int d, int c, int n) {
    if (o.per[d]) return (T)0;
    T r = 0;
    if (c == 0) r += o.a[d] * zp_alpha(o, d, 0);
    if (c == n - 1) r += o.a[d] * zp_alpha(o, d, 1);
    return r;
}

template <typename T, int MODE, typename TO>
__global__ __launch_bounds__(256) void zp_pass(GOp<T> o, const T* __restrict__ in0, const T* __restrict__ in1,
                                               const T* __restrict__ cb, const T* __restrict__ cq,
                                               T* __restrict__ out0, TO* __restrict__ out1,
                                               const double* __restrict__ scal, double* __restrict__ part, T w,
                                               ZpGrid g, long long cvs, Bat bt) {
    using C = ZpCfg<T>;
    constexpr int TX = C::TX, TY = C::TY, NT = C::NT, PF = C::PF;
    constexpr int RX = TX + 2, RY = TY + 2, RP = RX * RY;
    constexpr int NL = (RP + NT - 1) / NT;  // ring elements per thread
    constexpr int CPT = TX * TY / NT;       // output cells per thread and plane
    constexpr int RS = NT / TX;             // rows between a thread's output cells
    static_assert((TX * TY) % NT == 0 && NT % TX == 0 && TX % 2 == 0 && TY % 2 == 0, "tile shape");
    constexpr bool RING1 = MODE == ZP_SPMV_P;
    constexpr bool PJ = MODE == ZP_PJAC;
    constexpr bool SMOOTH = MODE == ZP_JAC || MODE == ZP_JAC_DOT3 || MODE == ZP_JAC2Z || PJ;
    static_assert(!PJ || PF % 2 == 0, "the coarse window shifts on odd planes: PF must be even");
    constexpr bool HASB = MODE != ZP_SPMV_P && MODE != ZP_APPLY;
    constexpr bool HASQ = MODE == ZP_JAC_DOT3;
    constexpr bool RED1 = MODE == ZP_SPMV_P || MODE == ZP_RESID;
    __shared__ T ring[4][NL * NT];
    __shared__ T rbuf[MODE == ZP_RESTRICT ? 2 * TX * TY : 1];

    BAT_SLOT(bt);
    in0 += vo;
    if (RING1) in1 += vo;
    if (PJ) in1 += (long long)slot * cvs;
    if (HASB) cb += vo;
    if (HASQ) cq += vo;
    if (MODE == ZP_RESTRICT) out0 += (long long)slot * cvs;
    else if (out0) out0 += vo;
    if (out1) out1 += vo;
    if (part) part += (long long)slot * bt.ps;
    T beta = 0, zmean = 0;
    if (RING1) {
        const double* sc = scal + (long long)slot * bt.ss;
        beta = (T)sc[S_BETA];
        zmean = (T)sc[S_ZMEAN];
    }

    const int tid = threadIdx.x;
    int bi = (int)xcd_block(blockIdx.x, gridDim.x);  // neighbouring tiles share an XCD's L2
    const int x0 = (bi % g.tx) * TX;
    bi /= g.tx;
    const int y0 = (bi % g.ty) * TY;
    const int z0 = (bi / g.ty) * g.zc;
    const int z1 = min(z0 + g.zc, o.n2);

    // ring elements of this thread: in-plane offset of the cell it reads and its ghost factor
    uint32_t off[NL];
    T fac[NL];
    T dxe[NL], wde[NL];  // ZP_JAC2Z: x/y diagonal part and w / D of the cell read (interior z)
#pragma unroll
    for (int m = 0; m < NL; ++m) {
        int e = tid + m * NT;
        if (e >= RP) e = 0;
        const int lx = e % RX, ly = e / RX;
        T f = 1;
        const int mx = zp_mirror(o, 0, x0 - 1 + lx, o.n0, f);
        const int my = zp_mirror(o, 1, y0 - 1 + ly, o.n1, f);
        off[m] = (uint32_t)mx + o.s1 * (uint32_t)my;
        fac[m] = f;
        if (MODE == ZP_JAC2Z) {
            dxe[m] = zp_dadj(o, 0, mx, o.n0) + zp_dadj(o, 1, my, o.n1);
            wde[m] = dxe[m] != (T)0 ? w / (o.dint - dxe[m]) : o.wdi;
        }
    }
    // ZP_PJAC: coarse cells of each ring element's interpolation in x and y (parent, other
    // neighbour; a non-periodic side's ghost is alpha * parent) and the weights of the others
    const int cn0 = o.n0 >> 1, cn1 = o.n1 >> 1, cn2 = o.n2 >> 1;
    uint32_t co[PJ ? NL : 1][4];
    T gxe[PJ ? NL : 1], gye[PJ ? NL : 1];
    if (PJ) {
#pragma unroll
        for (int m = 0; m < NL; ++m) {
            int e = tid + m * NT;
            if (e >= RP) e = 0;
            T f = 1;
            const int mx = zp_mirror(o, 0, x0 - 1 + e % RX, o.n0, f);
            const int my = zp_mirror(o, 1, y0 - 1 + e / RX, o.n1, f);
            int q[2][2];  // [dim][parent, other]
            T gq[2];
            const int mm[2] = {mx, my}, cn[2] = {cn0, cn1};
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const int par = mm[d] >> 1;
                int oth = (mm[d] & 1) ? par + 1 : par - 1;
                T gf = (T)0.25;
                if (oth < 0 || oth >= cn[d]) {
                    if (o.per[d]) oth = oth < 0 ? oth + cn[d] : oth - cn[d];
                    else {
                        gf *= zp_alpha(o, d, oth < 0 ? 0 : 1);
                        oth = par;
                    }
                }
                q[d][0] = par;
                q[d][1] = oth;
                gq[d] = gf;
            }
            co[m][0] = (uint32_t)q[0][0] + (uint32_t)cn0 * (uint32_t)q[1][0];
            co[m][1] = (uint32_t)q[0][1] + (uint32_t)cn0 * (uint32_t)q[1][0];
            co[m][2] = (uint32_t)q[0][0] + (uint32_t)cn0 * (uint32_t)q[1][1];
            co[m][3] = (uint32_t)q[0][1] + (uint32_t)cn0 * (uint32_t)q[1][1];
            gxe[m] = gq[0];
            gye[m] = gq[1];
        }
    }
    // output cells of this thread: (lx, ly0 + m * RS)
    const int lx = tid % TX, ly0 = tid / TX;
    uint32_t coff[CPT];
    bool cval[CPT];
    T dxc[CPT], wdc[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
        const int gx = x0 + lx, gy = y0 + ly0 + m * RS;
        cval[m] = gx < o.n0 && gy < o.n1;
        const int cx = min(gx, o.n0 - 1), cy = min(gy, o.n1 - 1);
        coff[m] = (uint32_t)cx + o.s1 * (uint32_t)cy;
        if (SMOOTH) {
            dxc[m] = zp_dadj(o, 0, cx, o.n0) + zp_dadj(o, 1, cy, o.n1);
            wdc[m] = dxc[m] != (T)0 ? w / (o.dint - dxc[m]) : o.wdi;
        }
    }

    const int zb = z0 - 1, ze = z1 + 1;  // ring planes zb .. ze - 1
    T pre[PF][NL];
    T pre1[RING1 ? PF : 1][NL];
    T fzp[PF];
    int mzp[PF];
    T cbp[HASB ? PF : 1][CPT];
    T cqp[HASQ ? PF : 1][CPT];
    auto ring_load = [&](T(&r)[NL], T(&r1)[NL], T& fz, int& mzo, int z) {
        T f = 1;
        const int mz = zp_mirror(o, 2, z, o.n2, f);
        fz = f;
        mzo = mz;
        const T* p = in0 + (size_t)o.s2 * (uint32_t)mz;
#pragma unroll
        for (int m = 0; m < NL; ++m) r[m] = p[off[m]];
        if (RING1) {
            const T* p1 = in1 + (size_t)o.s2 * (uint32_t)mz;
#pragma unroll
            for (int m = 0; m < NL; ++m) r1[m] = p1[off[m]];
        }
    };
    auto ctr_load = [&](T(&rb)[CPT], T(&rq)[CPT], int z) {
        z = min(max(z, z0), z1 - 1);
        const size_t po = (size_t)o.s2 * (uint32_t)z;
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            if (HASB) rb[m] = cb[po + coff[m]];
            if (HASQ) rq[m] = cq[po + coff[m]];
        }
    };
    // ZP_PJAC: x-y interpolated coarse planes Q(c) at the ring elements, for the pair
    // (cA, cA + 1), cA = floor((zz - 1) / 2), that ring plane zz interpolates between
    // in z; virtual coarse planes -1 and cn2 follow the coarse ghost rule. The raw
    // values of the next plane are loaded two fine planes ahead.
    T qa[PJ ? NL : 1], qb[PJ ? NL : 1], raw[PJ ? 4 : 1][PJ ? NL : 1];
    T rawf = 1;
    auto coarse_load = [&](int c) {
        c = min(max(c, -1), cn2);
        T f = 1;
        const int mc = zp_mirror(o, 2, c, cn2, f);
        rawf = f;
        const T* p = in1 + (size_t)cn0 * (size_t)cn1 * (uint32_t)mc;
#pragma unroll
        for (int m = 0; m < NL; ++m)
#pragma unroll
            for (int k = 0; k < 4; ++k) raw[k][m] = p[co[m][k]];
    };
    auto coarse_q = [&](T(&qo)[PJ ? NL : 1]) {
#pragma unroll
        for (int m = 0; m < NL; ++m) {
            const T ry0 = (T)0.75 * raw[0][m] + gxe[m] * raw[1][m];
            const T ry1 = (T)0.75 * raw[2][m] + gxe[m] * raw[3][m];
            qo[m] = rawf * ((T)0.75 * ry0 + gye[m] * ry1);
        }
    };
    if (PJ) {
        coarse_load(z0 / 2 - 1);
        coarse_q(qa);
        coarse_load(z0 / 2);
        coarse_q(qb);
        coarse_load(z0 / 2 + 1);
    }
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        ring_load(pre[u], pre1[RING1 ? u : 0], fzp[u], mzp[u], min(zb + u, ze - 1));
        ctr_load(cbp[HASB ? u : 0], cqp[HASQ ? u : 0], zb + u - 1);
    }
    double s0 = 0, s1 = 0, s2 = 0;
    for (int zz0 = zb; zz0 < ze; zz0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int zz = zz0 + u;
            if (zz >= ze) break;  // uniform over the block
            T* rs = ring[(zz - zb) & 3];
            {
                const T fz = fzp[u];
                if (MODE == ZP_JAC2Z) {
                    const int mz = mzp[u];
                    const T dz = zp_dadj(o, 2, mz, o.n2);
                    if (dz != (T)0) {  // a boundary plane of a non-periodic z (uniform over the block)
#pragma unroll
                        for (int m = 0; m < NL; ++m)
                            rs[tid + m * NT] = (fac[m] * fz) * ((w / (o.dint - (dxe[m] + dz))) * pre[u][m]);
                    } else {
#pragma unroll
                        for (int m = 0; m < NL; ++m) rs[tid + m * NT] = (fac[m] * fz) * (wde[m] * pre[u][m]);
                    }
                } else if (RING1) {
#pragma unroll
                    for (int m = 0; m < NL; ++m)
                        rs[tid + m * NT] = (fac[m] * fz) * ((pre[u][m] - zmean) + beta * pre1[u][m]);
                } else if (PJ) {
                    if (u % 2 == 0 && zz > zb) {  // odd plane: the pair moves up by one coarse plane
#pragma unroll
                        for (int m = 0; m < NL; ++m) qa[m] = qb[m];
                        coarse_q(qb);
                        coarse_load((zz + 1) / 2 + 1);
                    }
                    // weights by the parity of the cell interpolated at (the mirror plane)
                    const bool odd = mzp[u] & 1;
                    const T wa = odd ? (T)0.75 : (T)0.25, wb = odd ? (T)0.25 : (T)0.75;
#pragma unroll
                    for (int m = 0; m < NL; ++m)
                        rs[tid + m * NT] = (fac[m] * fz) * (pre[u][m] + (wa * qa[m] + wb * qb[m]));
                } else {
#pragma unroll
                    for (int m = 0; m < NL; ++m) rs[tid + m * NT] = (fac[m] * fz) * pre[u][m];
                }
            }
            ring_load(pre[u], pre1[RING1 ? u : 0], fzp[u], mzp[u], min(zz + PF, ze - 1));
            __syncthreads();
            const int zl = zz - 1;  // output plane
            if (zl < z0) {          // pipeline still filling (uniform)
                ctr_load(cbp[HASB ? u : 0], cqp[HASQ ? u : 0], zl + PF);
                continue;
            }
            const T* pm = ring[(zl - 1 - zb) & 3];
            const T* pc = ring[(zl - zb) & 3];
            const T* pp = ring[(zl + 1 - zb) & 3];
            const size_t po = (size_t)o.s2 * (uint32_t)zl;
            T wdp[CPT];  // w / D of this plane's output cells
            if (SMOOTH) {
                const T dz = zp_dadj(o, 2, zl, o.n2);
                if (dz != (T)0) {  // boundary plane of a non-periodic z (uniform)
#pragma unroll
                    for (int m = 0; m < CPT; ++m) wdp[m] = w / (o.dint - (dxc[m] + dz));
                } else {
#pragma unroll
                    for (int m = 0; m < CPT; ++m) wdp[m] = wdc[m];
                }
            }
#pragma unroll
            for (int m = 0; m < CPT; ++m) {
                const int ly = ly0 + m * RS;
                const int e = (lx + 1) + RX * (ly + 1);
                const T c = pc[e];
                T acc = o.a[0] * pc[e - 1];
                acc += o.a[0] * pc[e + 1];
                acc += o.a[1] * pc[e - RX];
                acc += o.a[1] * pc[e + RX];
                acc += o.a[2] * pm[e];
                acc += o.a[2] * pp[e];
                const T av = o.dint * c - acc;
                if (SMOOTH) {
                    const T bv = cbp[HASB ? u : 0][m];
                    const T xo = c + wdp[m] * (bv - av);
                    if (cval[m]) {
                        out0[po + coff[m]] = xo;
                        if (MODE == ZP_JAC_DOT3) {
                            const double zd = (double)xo;
                            s0 += (double)bv * zd;
                            s1 += zd * (double)cqp[HASQ ? u : 0][m];
                            s2 += zd;
                        }
                    }
                } else if (MODE == ZP_SPMV_P) {
                    if (cval[m]) {
                        out0[po + coff[m]] = c;
                        out1[po + coff[m]] = (TO)av;
                        s0 += (double)c * (double)av;
                    }
                } else if (MODE == ZP_RESID) {
                    const T v = cbp[HASB ? u : 0][m] - av;
                    if (cval[m]) {
                        if (out1) out1[po + coff[m]] = (TO)v;
                        s0 += (double)v * (double)v;
                    }
                } else if (MODE == ZP_APPLY) {
                    if (cval[m]) out0[po + coff[m]] = av;
                } else if (MODE == ZP_RESTRICT) {
                    const T v = cbp[HASB ? u : 0][m] - av;
                    T* rb = rbuf + ((zl >> 1) & 1) * (TX * TY) + lx + TX * ly;
                    if (zl & 1) *rb += v;
                    else *rb = v;
                }
            }
            if (MODE == ZP_RESTRICT && (zl & 1)) {
                // both fine planes of coarse plane zl / 2 are in rbuf: sum the 2 x 2 x 2 children
                __syncthreads();
                constexpr int CX = TX / 2, CY = TY / 2;
                const int c0 = o.n0 >> 1, c1 = o.n1 >> 1;
                const T* rb = rbuf + ((zl >> 1) & 1) * (TX * TY);
                for (int t = tid; t < CX * CY; t += NT) {
                    const int I = t % CX, J = t / CX;
                    const int gI = (x0 >> 1) + I, gJ = (y0 >> 1) + J;
                    const T* r0 = rb + 2 * I + TX * (2 * J);
                    const T s = (r0[0] + r0[1]) + (r0[TX] + r0[TX + 1]);
                    if (gI < c0 && gJ < c1)
                        out0[(uint32_t)gI + (uint32_t)c0 * ((uint32_t)gJ + (uint32_t)c1 * (uint32_t)(zl >> 1))] =
                            s * (T)0.125;
                }
            }
            ctr_load(cbp[HASB ? u : 0], cqp[HASQ ? u : 0], zl + PF);
        }
    }
    if (MODE == ZP_JAC_DOT3) {
        s0 = gmg_block_sum(s0);
        s1 = gmg_block_sum(s1);
        s2 = gmg_block_sum(s2);
        if (tid == 0) {
            part[blockIdx.x] = s0;
            part[GMG_MAXB + blockIdx.x] = s1;
            part[2 * GMG_MAXB + blockIdx.x] = s2;
        }
    } else if (RED1) {
        if (part) {
            s0 = gmg_block_sum(s0);
            if (tid == 0) part[blockIdx.x] = s0;
        }
    }
}

// ---------------------------------------------------------------------------------------
// The up half of a V(2,2) level in one pass (same tiling and chunking as zp_pass, so the
// same partial count): x3 = x2 + P x_c, x4 = x3 + (w / D)(b - A x3), x5 = x4 + (w / D)(b -
// A x4) stored; Z2_UP_DOT3 also the partials of (b, x5), (x5, q), sum(x5).
// Two LDS rings of 4 planes: level 0 (x3) on the tile plus a halo of 2 in x and y, level 1
// (x4) on the tile plus 1. Step zz stores level-0 plane zz, computes level 1 at plane
// zz - 1 and the output at plane zz - 3 (its level-1 planes were written in earlier
// steps), so one barrier per step separates every store from its readers. Ghost values
// follow zp_pass: a non-periodic side's ghost of either level is alpha times the level's
// boundary cell, so a level-1 ghost position computes the sweep at its mirror cell
// (inside the tile, its level-0 neighbours in the halo-2 region) times alpha, and the
// ghost planes -1 / n of level 1 are alpha times its planes 0 / n - 1. Every cell gets
// the arithmetic of ZP_PJAC followed by ZP_JAC(_DOT3), so the results are the same bits
// with one pass fewer over the level's vectors. Used in FP64 on levels of >= 8-plane
// chunks (256^3: 237 -> 191 us per V-cycle for the pair); in FP32, and with the down half
// (ZP_JAC2Z + ZP_RESTRICT) fused the same way, the two-level kernel measured slower than
// the separate passes (its registers allow 2 waves per SIMD).
enum { Z2_UP = 0, Z2_UP_DOT3 };

// cell whose value stands at coordinate g (region coordinates run past the ghost layer
// by one in halo-2 regions): periodic dims wrap any g, a non-periodic ghost is alpha times
// the boundary cell, positions further out (never used) clamp to a boundary cell
template <typename T>
__device__ __forceinline__ int zp2_mirror(const GOp<T>& o, int d, int g, int n, T& f) {
    if (o.per[d]) return ((g % n) + n) % n;
    if (g < 0) {
        if (g == -1) f *= zp_alpha(o, d, 0);
        return 0;
    }
    if (g >= n) {
        if (g == n) f *= zp_alpha(o, d, 1);
        return n - 1;
    }
    return g;
}

// level-1 position at coordinate g: the coordinate whose sweep it holds (itself, or the
// mirror cell of a non-periodic ghost, factor alpha; further out: a boundary cell, unused)
template <typename T>
__device__ __forceinline__ int zp2_src(const GOp<T>& o, int d, int g, int n, T& f) {
    if (o.per[d]) return g;
    if (g < 0) {
        f *= zp_alpha(o, d, 0);
        return 0;
    }
    if (g >= n) {
        if (g == n) f *= zp_alpha(o, d, 1);
        return n - 1;
    }
    return g;
}

template <typename T, int M2>
__global__ __launch_bounds__(256) void zp2_pass(GOp<T> o, const T* __restrict__ xin, const T* __restrict__ xc,
                                                const T* __restrict__ cb, const T* __restrict__ cq,
                                                T* __restrict__ out0, double* __restrict__ part, T w, ZpGrid g,
                                                long long cvs, Bat bt) {
    using C = ZpCfg<T>;
    // prefetch distance 2 in both precisions: the second level and the b / q prefetch
    // take the registers a deeper level-0 prefetch would
    constexpr int TX = C::TX, TY = C::TY, NT = C::NT, PF = 2;
    constexpr int RX0 = TX + 4, RY0 = TY + 4, RP0 = RX0 * RY0, NL0 = (RP0 + NT - 1) / NT;
    constexpr int RX1 = TX + 2, RY1 = TY + 2, RP1 = RX1 * RY1, NL1 = (RP1 + NT - 1) / NT;
    constexpr int CPT = TX * TY / NT;  // output cells per thread and plane
    constexpr int RS = NT / TX;        // rows between a thread's output cells
    constexpr bool DOTS = M2 == Z2_UP_DOT3;
    static_assert((TX * TY) % NT == 0 && NT % TX == 0 && TX % 2 == 0 && TY % 2 == 0, "tile shape");
    static_assert(PF % 2 == 0, "the coarse window shifts every second plane: PF must be even");
    __shared__ T ring0[4][RP0];
    __shared__ T ring1[4][RP1];

    BAT_SLOT(bt);
    xin += vo;
    xc += (long long)slot * cvs;
    cb += vo;
    if (DOTS) {
        cq += vo;
        part += (long long)slot * bt.ps;
    }
    out0 += vo;

    const int tid = threadIdx.x;
    int bi = (int)xcd_block(blockIdx.x, gridDim.x);  // neighbouring tiles share an XCD's L2
    const int x0 = (bi % g.tx) * TX;
    bi /= g.tx;
    const int y0 = (bi % g.ty) * TY;
    const int z0 = (bi / g.ty) * g.zc;
    const int z1 = min(z0 + g.zc, o.n2);
    const bool zper = o.per[2] != 0;

    // level-0 elements: in-plane offset of the cell they hold, ghost factor, the cell
    // (x | y << 16: its coarse cells are recomputed per use)
    uint32_t off0[NL0];
    T fac0[NL0];
    uint32_t mxy[NL0];
    const int cn0 = o.n0 >> 1, cn1 = o.n1 >> 1, cn2 = o.n2 >> 1;
#pragma unroll
    for (int m = 0; m < NL0; ++m) {
        int e = tid + m * NT;
        if (e >= RP0) e = 0;
        T f = 1;
        const int mx = zp2_mirror(o, 0, x0 - 2 + e % RX0, o.n0, f);
        const int my = zp2_mirror(o, 1, y0 - 2 + e / RX0, o.n1, f);
        off0[m] = (uint32_t)mx + o.s1 * (uint32_t)my;
        fac0[m] = f;
        mxy[m] = (uint32_t)mx | ((uint32_t)my << 16);
    }
    // coarse cell (parent or other) of the x-y interpolation at fine coordinate mm
    // in dim d, and the weight of the other (a non-periodic coarse ghost is alpha * parent)
    auto coarse_of = [&](int d, int mm, int cn, bool other, T& gf) {
        const int par = mm >> 1;
        int oth = (mm & 1) ? par + 1 : par - 1;
        gf = (T)0.25;
        if (oth < 0 || oth >= cn) {
            if (o.per[d]) oth = oth < 0 ? oth + cn : oth - cn;
            else {
                gf *= zp_alpha(o, d, oth < 0 ? 0 : 1);
                oth = par;
            }
        }
        return other ? oth : par;
    };
    // level-1 elements: the level-0 ring index of the cell whose sweep they hold, its
    // in-plane offset (b and the diagonal), the ghost factor
    int src1[NL1];
    uint32_t off1[NL1];
    T fac1[NL1], dx1[NL1];
#pragma unroll
    for (int m = 0; m < NL1; ++m) {
        int e = tid + m * NT;
        if (e >= RP1) e = 0;
        T f = 1;
        const int sx = zp2_src(o, 0, x0 - 1 + e % RX1, o.n0, f);
        const int sy = zp2_src(o, 1, y0 - 1 + e / RX1, o.n1, f);
        src1[m] = (sx - (x0 - 2)) + RX0 * (sy - (y0 - 2));
        T fd = 1;
        const int cx = zp2_mirror(o, 0, sx, o.n0, fd), cy = zp2_mirror(o, 1, sy, o.n1, fd);
        off1[m] = (uint32_t)cx + o.s1 * (uint32_t)cy;
        fac1[m] = f;
        dx1[m] = zp_dadj(o, 0, cx, o.n0) + zp_dadj(o, 1, cy, o.n1);
    }
    // output cells of this thread: (lx, ly0 + m * RS)
    const int lx = tid % TX, ly0 = tid / TX;
    uint32_t coff[CPT];
    bool cval[CPT];
    T dxc[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
        const int gx = x0 + lx, gy = y0 + ly0 + m * RS;
        cval[m] = gx < o.n0 && gy < o.n1;
        const int cx = min(gx, o.n0 - 1), cy = min(gy, o.n1 - 1);
        coff[m] = (uint32_t)cx + o.s1 * (uint32_t)cy;
        dxc[m] = zp_dadj(o, 0, cx, o.n0) + zp_dadj(o, 1, cy, o.n1);
    }
    // w / D of a cell whose boundary diagonal part is dd
    auto wdiag = [&](T dd) { return dd != (T)0 ? w / (o.dint - dd) : o.wdi; };
    // plane of the vectors a level-1 / output plane p reads (periodic z wraps; a
    // non-periodic plane outside [0, n) is never computed from them: clamped)
    auto plane_of = [&](int p) { return zper ? ((p % o.n2) + o.n2) % o.n2 : min(max(p, 0), o.n2 - 1); };

    const int zb = z0 - 2, zl0 = z1 + 1;  // level-0 planes zb .. zl0; steps zz = zb .. z1 + 2
    T pre[PF][NL0];
    T fzp[PF];
    int mzp[PF];
    auto l0_load = [&](T(&r)[NL0], T& fz, int& mzo, int z) {
        T f = 1;
        const int mz = zp2_mirror(o, 2, z, o.n2, f);
        fz = f;
        mzo = mz;
        const T* p = xin + (size_t)o.s2 * (uint32_t)mz;
#pragma unroll
        for (int m = 0; m < NL0; ++m) r[m] = p[off0[m]];
    };
    // x-y interpolated coarse planes Q(c) at the level-0 elements for the pair
    // (cA, cA + 1), cA = floor((zz - 1) / 2), that plane zz interpolates between in z;
    // coarse planes outside [0, cn2) follow the coarse ghost rule; the raw values of the
    // next coarse plane are loaded two fine planes ahead
    T qa[NL0], qb[NL0], raw[4][NL0];
    T rawf = 1;
    auto coarse_load = [&](int c) {
        if (!o.per[2]) c = min(max(c, -1), cn2);  // a periodic z wraps any plane
        T f = 1;
        const int mc = zp2_mirror(o, 2, c, cn2, f);
        rawf = f;
        const T* p = xc + (size_t)cn0 * (size_t)cn1 * (uint32_t)mc;
#pragma unroll
        for (int m = 0; m < NL0; ++m) {
            const int mx = (int)(mxy[m] & 0xffffu), my = (int)(mxy[m] >> 16);
            T gx, gy;
            const uint32_t px = (uint32_t)coarse_of(0, mx, cn0, false, gx), ox = (uint32_t)coarse_of(0, mx, cn0, true, gx);
            const uint32_t py = (uint32_t)coarse_of(1, my, cn1, false, gy), oy = (uint32_t)coarse_of(1, my, cn1, true, gy);
            raw[0][m] = p[px + (uint32_t)cn0 * py];
            raw[1][m] = p[ox + (uint32_t)cn0 * py];
            raw[2][m] = p[px + (uint32_t)cn0 * oy];
            raw[3][m] = p[ox + (uint32_t)cn0 * oy];
        }
    };
    auto coarse_q = [&](T(&qo)[NL0]) {
#pragma unroll
        for (int m = 0; m < NL0; ++m) {
            T gx, gy;
            coarse_of(0, (int)(mxy[m] & 0xffffu), cn0, true, gx);
            coarse_of(1, (int)(mxy[m] >> 16), cn1, true, gy);
            const T ry0 = (T)0.75 * raw[0][m] + gx * raw[1][m];
            const T ry1 = (T)0.75 * raw[2][m] + gx * raw[3][m];
            qo[m] = rawf * ((T)0.75 * ry0 + gy * ry1);
        }
    };
    // plane zb = z0 - 2 (even) interpolates between coarse planes z0 / 2 - 2 and z0 / 2 - 1
    coarse_load(z0 / 2 - 2);
    coarse_q(qa);
    coarse_load(z0 / 2 - 1);
    coarse_q(qb);
    coarse_load(z0 / 2);
    // b of step zz's level-1 plane (zz - 1) and output plane (zz - 3), q of the output
    // plane: loaded PF steps ahead like the level-0 planes
    T bl1p[PF][NL1], bop[PF][CPT], qop[DOTS ? PF : 1][CPT];
    auto b_load = [&](T(&bl)[NL1], T(&bo)[CPT], T(&qo)[CPT], int zz) {
        const size_t po1 = (size_t)o.s2 * (uint32_t)plane_of(zz - 1);
#pragma unroll
        for (int m = 0; m < NL1; ++m) bl[m] = cb[po1 + off1[m]];
        const size_t po2 = (size_t)o.s2 * (uint32_t)plane_of(zz - 3);
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            bo[m] = cb[po2 + coff[m]];
            if (DOTS) qo[m] = cq[po2 + coff[m]];
        }
    };
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        l0_load(pre[u], fzp[u], mzp[u], min(zb + u, zl0));
        b_load(bl1p[u], bop[u], qop[DOTS ? u : 0], zb + u);
    }
    double s0 = 0, s1 = 0, s2 = 0;
    for (int zz0 = zb; zz0 <= z1 + 2; zz0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int zz = zz0 + u;
            if (zz > z1 + 2) break;  // uniform over the block
            const int p1 = zz - 1, p2 = zz - 3;
            const T(&bl1)[NL1] = bl1p[u];
            const T(&bo)[CPT] = bop[u];
            const T(&qo)[CPT] = qop[DOTS ? u : 0];
            // level-0 plane zz
            if (zz <= zl0) {
                T* rs = ring0[zz & 3];
                const T fz = fzp[u];
                if (u % 2 == 1) {  // odd plane (zb is even): the pair moves up by one coarse plane
#pragma unroll
                    for (int m = 0; m < NL0; ++m) qa[m] = qb[m];
                    coarse_q(qb);
                    coarse_load((zz + 1) / 2 + 1);
                }
                // weights by the parity of the cell interpolated at (the mirror plane)
                const bool odd = mzp[u] & 1;
                const T wa = odd ? (T)0.75 : (T)0.25, wb = odd ? (T)0.25 : (T)0.75;
#pragma unroll
                for (int m = 0; m < NL0; ++m) {
                    const int e = tid + m * NT;
                    if (e < RP0) rs[e] = (fac0[m] * fz) * (pre[u][m] + (wa * qa[m] + wb * qb[m]));
                }
            }
            l0_load(pre[u], fzp[u], mzp[u], min(zz + PF, zl0));
            __syncthreads();
            // level 1 at plane p1 (level-0 planes p1 - 1 .. p1 + 1)
            if (p1 >= z0 - 1 && p1 <= z1 && (zper || p1 >= 0)) {
                T* r1 = ring1[p1 & 3];
                if (!zper && p1 == o.n2) {  // ghost plane n: alpha * plane n - 1
                    const T az = zp_alpha(o, 2, 1);
                    const T* r1m = ring1[(p1 - 1) & 3];
#pragma unroll
                    for (int m = 0; m < NL1; ++m) {
                        const int e = tid + m * NT;
                        if (e < RP1) r1[e] = az * r1m[e];
                    }
                } else {
                    const T* pm = ring0[(p1 - 1) & 3];
                    const T* pc = ring0[p1 & 3];
                    const T* pp = ring0[(p1 + 1) & 3];
                    const T dz = zp_dadj(o, 2, plane_of(p1), o.n2);
                    const bool g0 = !zper && p1 == 0;  // also the ghost plane -1
                    T* r1g = ring1[(p1 - 1) & 3];
                    const T az0 = zp_alpha(o, 2, 0);
#pragma unroll
                    for (int m = 0; m < NL1; ++m) {
                        const int e = tid + m * NT;
                        const int s = src1[m];
                        const T c = pc[s];
                        T acc = o.a[0] * pc[s - 1];
                        acc += o.a[0] * pc[s + 1];
                        acc += o.a[1] * pc[s - RX0];
                        acc += o.a[1] * pc[s + RX0];
                        acc += o.a[2] * pm[s];
                        acc += o.a[2] * pp[s];
                        const T av = o.dint * c - acc;
                        const T v = fac1[m] * (c + wdiag(dx1[m] + dz) * (bl1[m] - av));
                        if (e < RP1) {
                            r1[e] = v;
                            if (g0) r1g[e] = az0 * v;
                        }
                    }
                }
            }
            // output plane p2 (level-1 planes p2 - 1 .. p2 + 1, written in earlier steps)
            if (p2 >= z0 && p2 < z1) {
                const T* pm = ring1[(p2 - 1) & 3];
                const T* pc = ring1[p2 & 3];
                const T* pp = ring1[(p2 + 1) & 3];
                const size_t po = (size_t)o.s2 * (uint32_t)p2;
                const T dz = zp_dadj(o, 2, p2, o.n2);
#pragma unroll
                for (int m = 0; m < CPT; ++m) {
                    const int ly = ly0 + m * RS;
                    const int e = (lx + 1) + RX1 * (ly + 1);
                    const T c = pc[e];
                    T acc = o.a[0] * pc[e - 1];
                    acc += o.a[0] * pc[e + 1];
                    acc += o.a[1] * pc[e - RX1];
                    acc += o.a[1] * pc[e + RX1];
                    acc += o.a[2] * pm[e];
                    acc += o.a[2] * pp[e];
                    const T av = o.dint * c - acc;
                    const T xo = c + wdiag(dxc[m] + dz) * (bo[m] - av);
                    if (cval[m]) {
                        out0[po + coff[m]] = xo;
                        if (DOTS) {
                            const double zd = (double)xo;
                            s0 += (double)bo[m] * zd;
                            s1 += zd * (double)qo[DOTS ? m : 0];
                            s2 += zd;
                        }
                    }
                }
            }
            b_load(bl1p[u], bop[u], qop[DOTS ? u : 0], zz + PF);
        }
    }
    if (DOTS) {
        s0 = gmg_block_sum(s0);
        s1 = gmg_block_sum(s1);
        s2 = gmg_block_sum(s2);
        if (tid == 0) {
            part[blockIdx.x] = s0;
            part[GMG_MAXB + blockIdx.x] = s1;
            part[2 * GMG_MAXB + blockIdx.x] = s2;
        }
    }
}
