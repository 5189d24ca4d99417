/*
 * oracle_solve.c -- CPU restatement of OpFlow's implicit path (test
 * infrastructure only; see oracle.h):
 *
 *   orc_csr_poisson   StencilField pad evaluation + CSRMatrixGenerator::generate_s
 *                     for `c == d2x(e) + d2y(e) [+ d2z(e)]` on a centre field
 *                     (StencilField.hpp:123-350, StencilPad.hpp:136-298,
 *                     CSRMatrixGenerator.hpp:140-196). Pinned by the golden CSR
 *                     arrays of CSRMatrixGeneratorTest.cpp:57-161.
 *   orc_poisson_apply the SPD operator -sum_d d2_d(x) + sigma x with OpFlow's
 *                     centre ghost rules (D2SecondOrderCentered.hpp:160-171).
 *   orc_pcg_mg        CPU baseline solver: flexible PCG preconditioned by a
 *                     cell-centred geometric V(1,1)-cycle with weighted Jacobi
 *                     (the PFMG defaults, hypre struct_ls/pfmg.c:25-39), OpenMP.
 *                     Pinned by the HYPRE/AMGCL solution values of the 4x4
 *                     Dirichlet problem and by oracle/_ref (AMGCL).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---------------------------------------------------------------- pads */
typedef struct {
    long idx[32];
    double v[32];
    int n;
    double bias;
} pad_t;

static void pad_unit(pad_t* p, long idx) {
    p->n = 1;
    p->idx[0] = idx;
    p->v[0] = 1.0;
    p->bias = 0.0;
}
/* StencilPad::operator+= / -= (StencilPad.hpp:190-214): merge by key */
static void pad_acc(pad_t* a, const pad_t* b, double sgn) {
    for (int q = 0; q < b->n; ++q) {
        int f = -1;
        for (int t = 0; t < a->n; ++t)
            if (a->idx[t] == b->idx[q]) { f = t; break; }
        if (f >= 0) a->v[f] = sgn > 0 ? a->v[f] + b->v[q] : a->v[f] - b->v[q];
        else { a->idx[a->n] = b->idx[q]; a->v[a->n] = sgn > 0 ? b->v[q] : -b->v[q]; a->n++; }
    }
    a->bias = sgn > 0 ? a->bias + b->bias : a->bias - b->bias;
}
static void pad_scale(pad_t* a, double r, int div) {
    for (int q = 0; q < a->n; ++q) a->v[q] = div ? a->v[q] / r : a->v[q] * r;
    a->bias = div ? a->bias / r : a->bias * r;
}

#define MX(d, i) (m->x[d][(i) - m->org[d]])
#define MDX(d, i) (m->dx[d][(i) - m->org[d]])

/* StencilField::evalAtImpl_final for a centre field, no pin (StencilField.hpp:123-350) */
static void stencil_eval(const opf_field_desc* f, const orc_mesh* m, const int p[3], pad_t* out) {
    int q[3] = {p[0], p[1], p[2]};
    for (int d = 0; d < f->dim; ++d) {
        if (q[d] >= f->access_start[d] && q[d] < f->access_end[d]) continue;
        int upper = q[d] >= f->access_end[d];
        int type = f->bc_type[d][upper];
        if (type == OPF_BC_PERIODIC) {
            int L = f->access_end[d] - f->access_start[d];
            q[d] += upper ? -L : L;
            continue;
        }
        int mir[3] = {q[0], q[1], q[2]};
        mir[d] = upper ? 2 * f->access_end[d] - 1 - q[d] : 2 * f->access_start[d] - 1 - q[d];
        pad_t pm;
        stencil_eval(f, m, mir, &pm);
        double bc = f->bc_value[d][upper];
        if (type == OPF_BC_DIRC) {
            /* Interpolator1D::intp(x1, bc, x2, pad, x) = ((x1-x)*pad - (x2-x)*bc) / ((x1-x) - (x2-x)) */
            double x1 = MX(d, upper ? f->access_end[d] : f->access_start[d]);
            double x2 = MX(d, mir[d]) + MDX(d, mir[d]) / 2.;
            double x = MX(d, q[d]) + MDX(d, q[d]) / 2.;
            double A = x1 - x, B = x2 - x;
            pad_scale(&pm, A, 0);
            pm.bias -= B * bc;
            pad_scale(&pm, A - B, 1);
        } else if (type == OPF_BC_NEUM) {
            pm.bias += bc * (MX(d, q[d]) + MDX(d, q[d]) / 2. - MX(d, mir[d]) - MDX(d, mir[d]) / 2.);
        }
        *out = pm;
        return;
    }
    long lin = (long)(q[0] - f->access_start[0]) +
               (long)(f->access_end[0] - f->access_start[0]) *
                   ((long)(q[1] - f->access_start[1]) + (long)(f->access_end[1] - f->access_start[1]) * (q[2] - f->access_start[2]));
    pad_unit(out, lin);
}

static void d2_pad(const opf_field_desc* f, const orc_mesh* m, const int p[3], int d, pad_t* out) {
    int l[3] = {p[0], p[1], p[2]}, r[3] = {p[0], p[1], p[2]};
    l[d] -= 1;
    r[d] += 1;
    pad_t pl, pc, pr, A, B;
    stencil_eval(f, m, l, &pl);
    stencil_eval(f, m, p, &pc);
    stencil_eval(f, m, r, &pr);
    int i = p[d];
    double dxl = (MDX(d, i - 1) + MDX(d, i)) * 0.5, dxr = (MDX(d, i) + MDX(d, i + 1)) * 0.5;
    double dxc = (dxl + dxr) * 0.5;
    A = pr; pad_acc(&A, &pc, -1); pad_scale(&A, dxr, 1);
    B = pc; pad_acc(&B, &pl, -1); pad_scale(&B, dxl, 1);
    pad_acc(&A, &B, -1);
    pad_scale(&A, dxc, 1);
    *out = A;
}

static int cmp_pair(const void* a, const void* b) {
    long x = *(const long*)a, y = *(const long*)b;
    return x < y ? -1 : x > y;
}

/* CSRMatrixGenerator::generate_s for uniEqn = c - (d2x(e)+d2y(e)[+d2z(e)]).
 * Returns nnz; ptr has ncell+1 entries. pin_last: last row -> identity, rhs 0. */
long orc_csr_poisson(const opf_field_desc* f, const orc_mesh* m, double c, int pin_last, long* ptr, long* col,
                     double* val, double* rhs) {
    long row = 0, nnz = 0;
    long ncell = 1;
    for (int d = 0; d < f->dim; ++d) ncell *= f->access_end[d] - f->access_start[d];
    ptr[0] = 0;
    int p[3];
    for (p[2] = f->access_start[2]; p[2] < f->access_end[2]; ++p[2])
        for (p[1] = f->access_start[1]; p[1] < f->access_end[1]; ++p[1])
            for (p[0] = f->access_start[0]; p[0] < f->access_end[0]; ++p[0], ++row) {
                if (pin_last && row == ncell - 1) {
                    col[nnz] = row;
                    val[nnz] = 1.0;
                    nnz++;
                    rhs[row] = 0;
                    ptr[row + 1] = nnz;
                    continue;
                }
                pad_t sum, t;
                d2_pad(f, m, p, 0, &sum);
                for (int d = 1; d < f->dim; ++d) {
                    d2_pad(f, m, p, d, &t);
                    pad_acc(&sum, &t, +1);
                }
                /* scalar - pad */
                pad_t e;
                e.n = 0;
                e.bias = 0;
                pad_acc(&e, &sum, -1);
                e.bias = c - sum.bias;
                long pairs[64];
                double vals[32];
                for (int q = 0; q < e.n; ++q) {
                    pairs[2 * q] = e.idx[q];
                    pairs[2 * q + 1] = q;
                }
                qsort(pairs, e.n, 2 * sizeof(long), cmp_pair);
                for (int q = 0; q < e.n; ++q) vals[q] = e.v[pairs[2 * q + 1]];
                for (int q = 0; q < e.n; ++q) {
                    col[nnz] = pairs[2 * q];
                    val[nnz] = vals[q];
                    nnz++;
                }
                rhs[row] = -e.bias;
                ptr[row + 1] = nnz;
            }
    return nnz;
}

/* ---------------------------------------------------------------- operator */
typedef struct {
    int dim, n[3];
    double h[3];
    int bc[3][2];
    double sigma;
} pois_t;

static inline long cidx(const pois_t* P, int i, int j, int k) { return i + (long)P->n[0] * (j + (long)P->n[1] * k); }

/* value of x at neighbour of (i,j,k) in dim d, side s, with centre ghost rules
 * (homogeneous: Dirichlet ghost = -mirror, Neumann ghost = mirror, periodic wrap) */
static inline double nbval(const pois_t* P, const double* x, int i, int j, int k, int d, int s) {
    int q[3] = {i, j, k};
    q[d] += s ? 1 : -1;
    if (q[d] >= 0 && q[d] < P->n[d]) return x[cidx(P, q[0], q[1], q[2])];
    int t = P->bc[d][s];
    if (t == OPF_BC_PERIODIC) {
        q[d] = q[d] < 0 ? P->n[d] - 1 : 0;
        return x[cidx(P, q[0], q[1], q[2])];
    }
    double xm = x[cidx(P, i, j, k)];
    return t == OPF_BC_DIRC ? -xm : xm;
}

/* y = sigma x - sum_d ((r - c)/h - (c - l)/h)/h  (SPD form of -d2 with folded ghosts) */
void orc_poisson_apply(int dim, const int n[3], const double h[3], const int bc[3][2], double sigma, const double* x,
                       double* y) {
    pois_t P = {dim, {n[0], dim > 1 ? n[1] : 1, dim > 2 ? n[2] : 1}, {h[0], h[1], h[2]}, {{0}}, sigma};
    memcpy(P.bc, bc, sizeof(P.bc));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < P.n[2]; ++k)
        for (int j = 0; j < P.n[1]; ++j)
            for (int i = 0; i < P.n[0]; ++i) {
                double c = x[cidx(&P, i, j, k)];
                double s = 0;
                for (int d = 0; d < dim; ++d) {
                    double l = nbval(&P, x, i, j, k, d, 0), r = nbval(&P, x, i, j, k, d, 1);
                    s = s + ((r - c) / h[d] - (c - l) / h[d]) / h[d];
                }
                y[cidx(&P, i, j, k)] = sigma * c - s;
            }
}

/* ---------------------------------------------------------------- CPU GMG-FCG */
typedef struct {
    pois_t P;
    double *x, *b, *t;
} lvl_t;

static double diag_of(const pois_t* P, int i, int j, int k) {
    double D = P->sigma;
    int c[3] = {i, j, k};
    for (int d = 0; d < P->dim; ++d) {
        double a = 1.0 / (P->h[d] * P->h[d]);
        for (int s = 0; s < 2; ++s) {
            int inside = s ? c[d] < P->n[d] - 1 : c[d] > 0;
            int t = P->bc[d][s];
            if (inside || t == OPF_BC_PERIODIC) D += a;
            else if (t == OPF_BC_DIRC) D += 2 * a;
        }
    }
    return D;
}

static void apply_l(const pois_t* P, const double* x, double* y) {
    orc_poisson_apply(P->dim, P->n, P->h, P->bc, P->sigma, x, y);
}

static void jacobi(const pois_t* P, const double* x, const double* b, double* xo, double w, int zero) {
    long N = (long)P->n[0] * P->n[1] * P->n[2];
    double* ax = NULL;
    if (!zero) {
        ax = (double*)malloc(sizeof(double) * N);
        apply_l(P, x, ax);
    }
#pragma omp parallel for schedule(static)
    for (long c = 0; c < N; ++c) {
        int i = (int)(c % P->n[0]), j = (int)((c / P->n[0]) % P->n[1]), k = (int)(c / ((long)P->n[0] * P->n[1]));
        double D = diag_of(P, i, j, k);
        xo[c] = zero ? w * b[c] / D : x[c] + w * (b[c] - ax[c]) / D;
    }
    free(ax);
}

static void vcyc(lvl_t* L, int nl, int l, const double* rhs, double* out, double w) {
    pois_t* P = &L[l].P;
    long N = (long)P->n[0] * P->n[1] * P->n[2];
    if (l == nl - 1) { /* coarsest: many Jacobi sweeps */
        jacobi(P, NULL, rhs, out, w, 1);
        for (int s = 0; s < 200; ++s) {
            jacobi(P, out, rhs, L[l].t, w, 0);
            memcpy(out, L[l].t, sizeof(double) * N);
        }
        return;
    }
    jacobi(P, NULL, rhs, out, w, 1);
    double* r = (double*)malloc(sizeof(double) * N);
    apply_l(P, out, r);
    for (long c = 0; c < N; ++c) r[c] = rhs[c] - r[c];
    pois_t* C = &L[l + 1].P;
    long NC = (long)C->n[0] * C->n[1] * C->n[2];
    double fac = P->dim == 3 ? 0.125 : 0.25;
    for (long c = 0; c < NC; ++c) {
        int I = (int)(c % C->n[0]), J = (int)((c / C->n[0]) % C->n[1]), K = (int)(c / ((long)C->n[0] * C->n[1]));
        double s = 0;
        for (int dk = 0; dk < (P->dim == 3 ? 2 : 1); ++dk)
            for (int dj = 0; dj < 2; ++dj)
                for (int di = 0; di < 2; ++di) s += r[cidx(P, 2 * I + di, 2 * J + dj, P->dim == 3 ? 2 * K + dk : 0)];
        L[l + 1].b[c] = s * fac;
    }
    free(r);
    vcyc(L, nl, l + 1, L[l + 1].b, L[l + 1].x, w);
    /* linear prolongation with coarse ghosts */
    for (long c = 0; c < N; ++c) {
        int f[3] = {(int)(c % P->n[0]), (int)((c / P->n[0]) % P->n[1]), (int)(c / ((long)P->n[0] * P->n[1]))};
        double s = 0;
        for (int mbits = 0; mbits < (1 << P->dim); ++mbits) {
            double wt = 1, fac2 = 1;
            int q[3] = {0, 0, 0};
            for (int d = 0; d < P->dim; ++d) {
                int par = f[d] >> 1, oth = (f[d] & 1) ? par + 1 : par - 1;
                int use = (mbits >> d) & 1;
                wt *= use ? 0.25 : 0.75;
                int cc = use ? oth : par;
                if (cc < 0 || cc >= C->n[d]) {
                    int t = C->bc[d][cc < 0 ? 0 : 1];
                    if (t == OPF_BC_PERIODIC) cc = cc < 0 ? C->n[d] - 1 : 0;
                    else { fac2 *= t == OPF_BC_DIRC ? -1.0 : 1.0; cc = par; }
                }
                q[d] = cc;
            }
            s += wt * fac2 * L[l + 1].x[cidx(C, q[0], q[1], q[2])];
        }
        out[c] += s;
    }
    jacobi(P, out, rhs, L[l].t, w, 0);
    memcpy(out, L[l].t, sizeof(double) * N);
}

static double dot(long N, const double* a, const double* b) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long c = 0; c < N; ++c) s += a[c] * b[c];
    return s;
}

/* Flexible PCG with a geometric V(1,1) preconditioner. Stops when ||b-Ax||/||b|| <= tol.
 * Returns iterations; *relres = final true relative residual. */
int orc_pcg_mg(int dim, const int n[3], const double h[3], const int bc[3][2], double sigma, const double* b,
               double* x, double tol, int maxit, double* relres) {
    lvl_t L[24];
    int nl = 0;
    pois_t P = {dim, {n[0], dim > 1 ? n[1] : 1, dim > 2 ? n[2] : 1}, {h[0], h[1], dim > 2 ? h[2] : 1}, {{0}}, sigma};
    memcpy(P.bc, bc, sizeof(P.bc));
    for (;;) {
        long N = (long)P.n[0] * P.n[1] * P.n[2];
        L[nl].P = P;
        L[nl].x = (double*)calloc(N, sizeof(double));
        L[nl].b = (double*)calloc(N, sizeof(double));
        L[nl].t = (double*)calloc(N, sizeof(double));
        nl++;
        int can = N > 64 && nl < 24;
        for (int d = 0; d < dim; ++d) can &= P.n[d] % 2 == 0 && P.n[d] >= 4;
        if (!can) break;
        for (int d = 0; d < dim; ++d) {
            P.n[d] /= 2;
            P.h[d] *= 2;
        }
    }
    double w = dim == 3 ? 6.0 / 7.0 : 0.8;
    long N = (long)n[0] * (dim > 1 ? n[1] : 1) * (dim > 2 ? n[2] : 1);
    double *r = malloc(N * 8), *z = malloc(N * 8), *p = calloc(N, 8), *q = malloc(N * 8);
    apply_l(&L[0].P, x, r);
    for (long c = 0; c < N; ++c) r[c] = b[c] - r[c];
    double bb = dot(N, b, b), rr = dot(N, r, r);
    int it = 0;
    if (rr > tol * tol * bb) {
        vcyc(L, nl, 0, r, z, w);
        double rz = dot(N, r, z), beta = 0;
        for (it = 1; it <= maxit; ++it) {
            for (long c = 0; c < N; ++c) p[c] = z[c] + beta * p[c];
            apply_l(&L[0].P, p, q);
            double alpha = rz / dot(N, p, q);
            for (long c = 0; c < N; ++c) {
                x[c] += alpha * p[c];
                r[c] -= alpha * q[c];
            }
            rr = dot(N, r, r);
            if (rr <= tol * tol * bb) break;
            vcyc(L, nl, 0, r, z, w);
            double rzn = dot(N, r, z), zq = dot(N, z, q);
            beta = -alpha * zq / rz;
            rz = rzn;
        }
    }
    apply_l(&L[0].P, x, r);
    for (long c = 0; c < N; ++c) r[c] = b[c] - r[c];
    *relres = bb > 0 ? sqrt(dot(N, r, r) / bb) : 0.0;
    free(r); free(z); free(p); free(q);
    for (int l = 0; l < nl; ++l) { free(L[l].x); free(L[l].b); free(L[l].t); }
    return it;
}
