/*
 * oracle.c -- CPU restatement of OpFlow's hot path, used ONLY as test
 * infrastructure: the checker in tests/, in __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py. Nothing in the product (opflow_amd/,
 * include/) links, loads or calls this code.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the OpFlow source tree). Pins (see the header comment of oracle.h):
 *   orc_desc_finalize     CartesianFieldTest ranges, closed forms       pinned
 *   orc_update_padding    DircBCTest linear-extrapolation, CenterPeriodicValueCheck,
 *                         brute-force pass-by-pass replay (tests)       pinned
 *   orc_ftcs              discrete eigenmode closed forms (2-D Dirichlet,
 *                         3-D periodic Fourier modes)                   pinned
 *   orc_reduce            RangeReduceTest known sums                    pinned
 *   orc_csr_assemble      CSRMatrixGeneratorTest golden arrays          pinned
 *   orc_pcg_mg            manufactured solutions + AMGCL (oracle/_ref)  pinned
 *
 * Built with -ffp-contract=off so every a*b+c rounds twice, as the
 * reference's expressions do without FMA contraction.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* ---- mesh: MeshBuilder::set1DMesh(min,max) + setExtMesh(Symm), CartesianMesh.hpp:235-248,279-303 */
void orc_mesh_uniform(int nnodes, double lo, double hi, int pad, double* x, double* dx) {
    for (int i = 0; i < nnodes; ++i) x[i + pad] = (hi - lo) / (nnodes - 1) * (i - 0) + lo;
    for (int j = 0; j < nnodes - 1; ++j) dx[j + pad] = (hi - lo) / (nnodes - 1);
    for (int i = -pad; i < 0; ++i) dx[i + pad] = dx[2 * 0 - 1 - i + pad];
    for (int i = nnodes - 1; i < nnodes + pad - 1; ++i) dx[i + pad] = dx[2 * nnodes - 3 - i + pad];
    for (int i = -1; i >= -pad; --i) x[i + pad] = x[i + 1 + pad] - dx[i + pad];
    for (int i = nnodes; i < nnodes + pad; ++i) x[i + pad] = x[i - 1 + pad] + dx[i - 1 + pad];
}

/* ---- ranges: ExprBuilder::calculateRanges/validateRanges/build, CartesianField.hpp:929-1029 */
int orc_desc_finalize(opf_field_desc* f, const int mesh_start[3], const int mesh_nnodes[3]) {
    for (int d = 0; d < f->dim; ++d)
        for (int s = 0; s < 2; ++s) f->padding = imax(f->padding, f->ext[d][s]);
    for (int d = 0; d < 3; ++d) {
        if (d >= f->dim) {
            f->mesh_start[d] = f->local_start[d] = f->assign_start[d] = f->access_start[d] = f->logical_start[d] = 0;
            f->mesh_end[d] = f->local_end[d] = f->assign_end[d] = f->access_end[d] = f->logical_end[d] = 1;
            f->offset[d] = 0;
            f->extent[d] = 1;
            continue;
        }
        /* logical = assignable = local = accessible = mesh range */
        int ms = mesh_start[d], me = mesh_start[d] + mesh_nnodes[d];
        int as = ms, ae = me, cs = ms, ce = me, gs, ge, ls, le;
        int loc = f->loc[d];
        switch (f->bc_type[d][0]) {
            case OPF_BC_DIRC:
                if (loc == OPF_LOC_CORNER) as++;
                break;
            case OPF_BC_NEUM: case OPF_BC_UNDEFINED: case OPF_BC_SYMM: case OPF_BC_ASYMM: case OPF_BC_PERIODIC:
                break;
            default: return -3;
        }
        switch (f->bc_type[d][1]) {
            case OPF_BC_DIRC:
                if (loc == OPF_LOC_CORNER) ae--;
                else if (loc == OPF_LOC_CENTER) { ce--; ae--; }
                break;
            case OPF_BC_NEUM: case OPF_BC_UNDEFINED: case OPF_BC_SYMM: case OPF_BC_ASYMM:
                if (loc == OPF_LOC_CENTER) { ce--; ae--; }
                break;
            case OPF_BC_PERIODIC:
                ce--; ae--;
                break;
            default: return -3;
        }
        gs = cs - f->ext[d][0];
        ge = ce + f->ext[d][1];
        ls = cs; le = ce;                     /* localRange = accessibleRange */
        cs = imax(cs, gs); ce = imin(ce, ge); /* accessible <= logical */
        ls = imax(ls, gs); le = imin(le, ge); /* local <= logical */
        as = imax(as, cs); ae = imin(ae, ce); /* assignable <= accessible */
        f->mesh_start[d] = ms; f->mesh_end[d] = me;
        f->assign_start[d] = as; f->assign_end[d] = ae;
        f->access_start[d] = cs; f->access_end[d] = ce;
        f->logical_start[d] = gs; f->logical_end[d] = ge;
        f->local_start[d] = ls; f->local_end[d] = le;
        f->offset[d] = ls - f->padding;
        f->extent[d] = le - ls + 2 * f->padding;
    }
    return 0;
}

static inline size_t lin(const opf_field_desc* f, int i, int j, int k) {
    return (size_t)(i - f->offset[0]) +
           (size_t)f->extent[0] * ((size_t)(j - f->offset[1]) + (size_t)f->extent[1] * (size_t)(k - f->offset[2]));
}
static inline size_t lin3(const opf_field_desc* f, const int p[3]) { return lin(f, p[0], p[1], p[2]); }

#define MX(d, i) (m->x[d][(i) - m->org[d]])
#define MDX(d, i) (m->dx[d][(i) - m->org[d]])

/* Interpolator1D::intp(x1,y1,x2,y2,x) -> intp(x1-x,y1,x2-x,y2) = (x1 y2 - x2 y1)/(x1-x2), Interpolator.hpp:21-40 */
static inline double intp5(double x1, double y1, double x2, double y2, double x) {
    double a = x1 - x, b = x2 - x;
    return (a * y2 - b * y1) / (a - b);
}

typedef struct { int s[3], e[3]; } box;

/* One rangeFor over box r applying the rule of (dim i, side) -- CartesianField.hpp:385-481 (lower), 496-593 (upper) */
static void pad_zone(const opf_field_desc* f, const orc_mesh* m, double* u, box r, int i, int upper) {
    int ls = f->local_start[i], le = f->local_end[i];
    int type = f->bc_type[i][upper];
    double bc_v = f->bc_value[i][upper];
    int corner = f->loc[i] == OPF_LOC_CORNER;
    int p[3];
    for (p[2] = r.s[2]; p[2] < r.e[2]; ++p[2])
        for (p[1] = r.s[1]; p[1] < r.e[1]; ++p[1])
            for (p[0] = r.s[0]; p[0] < r.e[0]; ++p[0]) {
                int q[3] = {p[0], p[1], p[2]};
                int c = p[i];
                int mi;
                if (!upper) mi = corner ? 2 * ls - c : 2 * ls - 1 - c;
                else mi = corner ? 2 * le - 2 - c : 2 * le - 1 - c;
                q[i] = mi;
                double um = u[lin3(f, q)], v;
                switch (type) {
                    case OPF_BC_DIRC:
                        if (corner)
                            v = intp5(MX(i, upper ? le - 1 : ls), bc_v, MX(i, mi), um, MX(i, c));
                        else
                            v = intp5(MX(i, upper ? le : ls), bc_v, MX(i, mi) + MDX(i, mi) / 2., um,
                                      MX(i, c) + MDX(i, c) / 2.);
                        break;
                    case OPF_BC_NEUM:
                        if (corner) v = um + bc_v * (MX(i, c) - MX(i, mi));
                        else v = um + bc_v * (MX(i, c) + MDX(i, c) / 2. - MX(i, mi) - MDX(i, mi) / 2.);
                        break;
                    case OPF_BC_SYMM: v = um; break;
                    case OPF_BC_ASYMM: v = -um; break;
                    default: v = u[lin3(f, p)];
                }
                u[lin3(f, p)] = v;
            }
}

/* CartesianField::updatePaddingImpl_final, single-process branch (CartesianField.hpp:349-629) */
void orc_update_padding(const opf_field_desc* f, const orc_mesh* m, double* u) {
    int dim = f->dim;
    /* step 0: Dirichlet boundary nodes of corner-located dims (:350-363) */
    for (int i = 0; i < dim; ++i) {
        for (int s = 0; s < 2; ++s) {
            int touches = s == 0 ? f->local_start[i] == f->access_start[i] : f->local_end[i] == f->access_end[i];
            if (!touches || f->bc_type[i][s] != OPF_BC_DIRC || f->loc[i] != OPF_LOC_CORNER) continue;
            box r;
            for (int d = 0; d < 3; ++d) { r.s[d] = f->local_start[d]; r.e[d] = f->local_end[d]; }
            r.s[i] = s == 0 ? f->local_start[i] : f->local_end[i] - 1;
            r.e[i] = r.s[i] + 1;
            int p[3];
            for (p[2] = r.s[2]; p[2] < r.e[2]; ++p[2])
                for (p[1] = r.s[1]; p[1] < r.e[1]; ++p[1])
                    for (p[0] = r.s[0]; p[0] < r.e[0]; ++p[0]) u[lin3(f, p)] = f->bc_value[i][s];
        }
    }
    /* step 1: per dim, lower then upper, padding zones extended over earlier dims (:365-607) */
    int start[3], end[3];
    for (int i = 0; i < dim; ++i) {
        if (f->local_start[i] == f->access_start[i] && f->bc_type[i][0] != OPF_BC_UNDEFINED &&
            f->bc_type[i][0] != OPF_BC_PERIODIC) {
            start[i] = f->logical_start[i];
            box r;
            for (int d = 0; d < 3; ++d) { r.s[d] = f->local_start[d]; r.e[d] = f->local_end[d]; }
            for (int j = 0; j < i; ++j) { r.s[j] = start[j]; r.e[j] = end[j]; }
            r.s[i] = start[i];
            r.e[i] = f->local_start[i];
            pad_zone(f, m, u, r, i, 0);
        } else {
            start[i] = f->local_start[i];
        }
        if (f->local_end[i] == f->access_end[i] && f->bc_type[i][1] != OPF_BC_UNDEFINED &&
            f->bc_type[i][1] != OPF_BC_PERIODIC) {
            end[i] = f->logical_end[i];
            box r;
            for (int d = 0; d < 3; ++d) { r.s[d] = f->local_start[d]; r.e[d] = f->local_end[d]; }
            for (int j = 0; j < i; ++j) { r.s[j] = start[j]; r.e[j] = end[j]; }
            r.s[i] = f->local_end[i];
            r.e[i] = f->logical_end[i];
            pad_zone(f, m, u, r, i, 1);
        } else {
            end[i] = f->local_end[i];
        }
    }
    /* step 2: periodic wrap over the whole logical range (:609-629) */
    for (int i = 0; i < dim; ++i) {
        if (f->bc_type[i][0] != OPF_BC_PERIODIC) continue;
        int L = f->access_end[i] - f->access_start[i];
        for (int side = 0; side < 2; ++side) {
            box r;
            for (int d = 0; d < 3; ++d) { r.s[d] = f->logical_start[d]; r.e[d] = f->logical_end[d]; }
            if (side == 0) { r.s[i] = f->logical_start[i]; r.e[i] = f->access_start[i]; }
            else { r.s[i] = f->access_end[i]; r.e[i] = f->logical_end[i]; }
            int p[3];
            for (p[2] = r.s[2]; p[2] < r.e[2]; ++p[2])
                for (p[1] = r.s[1]; p[1] < r.e[1]; ++p[1])
                    for (p[0] = r.s[0]; p[0] < r.e[0]; ++p[0]) {
                        int q[3] = {p[0], p[1], p[2]};
                        q[i] += side == 0 ? L : -L;
                        u[lin3(f, p)] = u[lin3(f, q)];
                    }
        }
    }
}

/* D2SecondOrderCentered<d>::eval (D2SecondOrderCentered.hpp:160-171) on a field */
static inline double d2c(const opf_field_desc* f, const orc_mesh* m, const double* u, int d, int i, int j, int k) {
    int p[3] = {i, j, k}, q[3] = {i, j, k}, r[3] = {i, j, k};
    q[d] -= 1;
    r[d] += 1;
    double l = u[lin3(f, q)], c = u[lin3(f, p)], rr = u[lin3(f, r)];
    int ii = p[d];
    double dxl = f->loc[d] == OPF_LOC_CORNER ? MDX(d, ii - 1) : (MDX(d, ii - 1) + MDX(d, ii)) * 0.5;
    double dxr = f->loc[d] == OPF_LOC_CORNER ? MDX(d, ii) : (MDX(d, ii) + MDX(d, ii + 1)) * 0.5;
    double dxc = (dxl + dxr) * 0.5;
    return ((rr - c) / dxr - (c - l) / dxl) / dxc;
}

/* One FTCS assignment  u = u + coef*(d2x(u)+d2y(u)[+d2z(u)])  with OpFlow's
 * FieldAssigner sequence: temp = u; temp[i] = rhs(i) over
 * assignable∩local; temp.updatePadding(); u[i] = temp[i]; u.updatePadding();
 * then assignImpl_final's extra u.updatePadding()
 * (FieldAssigner.hpp:26-36,39-86; CartesianField.hpp:225-233). */
void orc_ftcs(const opf_field_desc* f, const orc_mesh* m, double* u, double coef, int nsteps, int nthreads) {
    size_t n = (size_t)f->extent[0] * f->extent[1] * f->extent[2];
    double* temp = (double*)malloc(n * sizeof(double));
    int s[3], e[3];
    for (int d = 0; d < 3; ++d) {
        s[d] = imax(f->assign_start[d], f->local_start[d]);
        e[d] = imin(f->assign_end[d], f->local_end[d]);
    }
#ifdef _OPENMP
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();  /* this call only */
#else
    const int nt = 1;
    (void)nthreads;
#endif
    for (int it = 0; it < nsteps; ++it) {
        memcpy(temp, u, n * sizeof(double));
        int dim = f->dim;
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int k = s[2]; k < e[2]; ++k)
            for (int j = s[1]; j < e[1]; ++j)
                for (int i = s[0]; i < e[0]; ++i) {
                    double acc = d2c(f, m, u, 0, i, j, k);
                    if (dim > 1) acc = acc + d2c(f, m, u, 1, i, j, k);
                    if (dim > 2) acc = acc + d2c(f, m, u, 2, i, j, k);
                    temp[lin(f, i, j, k)] = u[lin(f, i, j, k)] + coef * acc;
                }
        orc_update_padding(f, m, temp);
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int k = s[2]; k < e[2]; ++k)
            for (int j = s[1]; j < e[1]; ++j)
                for (int i = s[0]; i < e[0]; ++i) u[lin(f, i, j, k)] = temp[lin(f, i, j, k)];
        orc_update_padding(f, m, u);
        orc_update_padding(f, m, u);
    }
    free(temp);
}

/* rangeReduce with plus / max / min (RangeFor.hpp:86-121), serial order */
double orc_reduce(const opf_field_desc* f, const double* u, int red, const int rs[3], const int re[3]) {
    double acc = red == OPF_RED_MAX ? -INFINITY : red == OPF_RED_MIN ? INFINITY : 0.0;
    for (int k = rs[2]; k < re[2]; ++k)
        for (int j = rs[1]; j < re[1]; ++j)
            for (int i = rs[0]; i < re[0]; ++i) {
                double v = u[lin(f, i, j, k)];
                switch (red) {
                    case OPF_RED_SUM: acc += v; break;
                    case OPF_RED_SUMSQ: acc += v * v; break;
                    case OPF_RED_MAX: acc = v > acc ? v : acc; break;
                    case OPF_RED_MIN: acc = v < acc ? v : acc; break;
                    case OPF_RED_MAXABS: acc = fabs(v) > acc ? fabs(v) : acc; break;
                }
            }
    return acc;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Thread count of later parallel regions (the CPU-baseline timings at one
 * thread and at all threads); n <= 0 restores the count the process started with. */
void orc_set_threads(int n) {
#ifdef _OPENMP
    static int initial = 0;
    if (!initial) initial = omp_get_max_threads();
    omp_set_num_threads(n > 0 ? n : initial);
#else
    (void)n;
#endif
}
