// Matrix-free geometric multigrid + flexible CG for cell-centred 5/7-point
// operators in divergence form (the SPD form of the StencilPad systems OpFlow
// assembles for Laplacian-type equations, HYPREEqnSolveHandler.hpp:125-166):
//
//   (A x)_i = sigma_i x_i + sum_faces a_f (x_i - x_nb)
//
// where a boundary face's neighbour is the ghost value alpha * x_i (alpha = -1
// for a centre Dirichlet face on a uniform mesh, +1 for Neumann) and a
// periodic face wraps. Coefficients are either constant per dimension or
// per-face arrays (variable coefficients).
#pragma once

#include <cmath>
#include <vector>

#include "fastdiv.hpp"
#include "opf_internal.hpp"

namespace opf {

constexpr int GMG_MAXB = 16384;  // max blocks of a reduction kernel (partials per value)
constexpr int GMG_RBLK = 2048;   // blocks of a one-dimensional (streaming) reduction launch
constexpr int GMG_SSTRIDE = 24;            // scalars per slot
constexpr int GMG_PSTRIDE = 4 * GMG_MAXB;  // reduction partials per slot

// Operator of one level as seen by the kernels (T = float or double).
template <typename T>
struct GOp {
    int n0, n1, n2;
    uint32_t s1, s2;  // linear strides of dims 1 and 2 (n0, n0 * n1); cells per slot < 2^31
    FastDiv f0, f1;   // division by n0 and n1 (cell index -> coordinates)
    int dim;
    int per[3];
    T a[3];          // constant face coefficient per dim (const mode)
    T sigma;         // constant diagonal shift (const mode)
    T bfac[3][2];    // boundary face factor (1 - alpha) per dim/side (non-periodic)
    const T* af[3];  // per-face coefficients, face array dims n + e_d (variable mode)
    const T* sg;     // per-cell sigma (variable mode)
    T dint;          // diagonal of a cell with no boundary face (const mode)
    T wdi;           // omega / dint: the smoother's interior scaling (const mode)
};

// Host description of one level (double master copy of the coefficients).
struct GLevelHost {
    int n[3];
    long long N;
    bool var;
    double a[3], sigma;
    std::vector<double> af[3], sg;  // variable mode
};

// A level built on the device (rows_to_level_dev): constant mode (a, sigma) or
// variable mode (FP64 device arrays af over the faces of each dim, sg per cell).
struct DevLevel {
    int n[3] = {1, 1, 1};
    long long N = 0;
    bool var = false;
    bool shift = false;  // some cell has a nonzero sigma
    double a[3] = {0, 0, 0}, sigma = 0;
    double* af[3] = {nullptr, nullptr, nullptr};
    double* sg = nullptr;
};

struct GLevelDev {
    int n[3];
    long long N;
    // coefficients in both precisions (variable mode arrays); borrowed: af64 / sg64
    // belong to the caller's DevLevel and are not freed with the hierarchy
    bool borrowed = false;
    double* af64[3] = {nullptr, nullptr, nullptr};
    float* af32[3] = {nullptr, nullptr, nullptr};
    double* sg64 = nullptr;
    float* sg32 = nullptr;
    // work vectors of the V-cycle precision: x (solution), b (rhs), t (temp)
    void* x = nullptr;
    void* b = nullptr;
    void* t = nullptr;
};

struct Hierarchy {
    int dim = 2;
    int per[3] = {0, 0, 0};
    double alpha[3][2] = {{-1, -1}, {-1, -1}, {-1, -1}};
    bool var = false;
    bool singular = false;
    int B = 1;  // request slots sharing this hierarchy (vectors are slot-major)
    int precision = OPF_PREC_FP64;  // V-cycle / inner-Krylov precision
    double omega = 0.8;
    int pre = 2, post = 2;
    std::vector<GLevelHost> host;
    std::vector<GLevelDev> dev;
    // coarsest dense inverse (FP64)
    double* cinv = nullptr;
    int cN = 0;
    // Krylov vectors at level 0 (precision of the inner solve) and FP64 outer vectors
    void* r = nullptr;
    void* z = nullptr;
    void* p[2] = {nullptr, nullptr};
    void* q = nullptr;
    void* e = nullptr;      // inner solution (mixed)
    double* r64 = nullptr;  // outer residual (mixed) / work
    double* part = nullptr; // reduction partials [4][GMG_MAXB] per slot
    double* scal = nullptr; // device scalars
    double* hscal = nullptr;  // pinned host mirror
    int* mk = nullptr;        // device slot masks [MK_COUNT][B]
    // single solves: two device-controlled iterations captured as a graph for the
    // (x, b, maxit) they were captured with
    hipGraph_t iter_graph = nullptr;
    hipGraphExec_t iter_exec = nullptr;
    const void* iter_x = nullptr;
    const void* iter_b = nullptr;
    int iter_maxit = 0;
    size_t bytes = 0;
};

int gmg_build(opf_ctx* ctx, Hierarchy& H, const GLevelHost& fine, const int per[3], const double alpha[3][2],
              int precision, int slots = 1);
// Hierarchy over a level built on the device: the coarse levels are coarsened by
// kernels and a variable-mode fine level's arrays are borrowed (they must outlive
// H). When H already holds a variable-mode hierarchy of the same shape, precision,
// slots, boundaries and singularity over the same arrays, its buffers (and a
// captured iteration graph) are kept and only the coefficients are recomputed.
int gmg_build_dev(opf_ctx* ctx, Hierarchy& H, const DevLevel& fine, const int per[3], const double alpha[3][2],
                  int precision, int slots = 1);

// ---- device level construction (level.hip)
// face arrays of dim d have extents n + e_d
__host__ __device__ __forceinline__ long long face_index(int d, int i, int j, int k, int n0, int n1) {
    int f0 = n0 + (d == 0), f1 = n1 + (d == 1);
    return i + (long long)f0 * (j + (long long)f1 * k);
}
long long level_face_count(const int n[3], int d);
// The rows of one diagonal block (PAD_K device arrays of N rows) as a level:
// face coefficients from the neighbour slots, the rest of the oriented diagonal as
// sigma. *sign orients the diagonal positive. A face whose two rows disagree is
// given their mean (the level is then the operator's symmetric part) and
// *symmetric is false. L.af / L.sg are allocated unless already set (reused); a
// uniform operator comes back in constant mode with its arrays still allocated.
int rows_to_level_dev(hipStream_t st, const double* coef, long long N, const int nn[3], int dim, const int per[3],
                      double* sign, DevLevel& L, bool* symmetric);
int dev_levels_equal(hipStream_t st, const DevLevel& a, const DevLevel& b, int dim, bool* eq);
void dev_level_free(DevLevel& L);
void coarsen_dev(hipStream_t st, int dim, const int fn[3], const double* const faf[3], const double* fsg,
                 const int cn[3], double* const caf[3], double* csg);
void to_f32_dev(hipStream_t st, long long n, const double* src, float* dst);
int read_scal(const Hierarchy& H, hipStream_t st);
// Krylov solve with reliable updates, batched over the slots of H (vectors slot-major).
// The Krylov vectors and the V-cycle are in precision T; x64 / b64 are the FP64
// solution and right-hand side. Start: for the slots in mk[MK_NEW], r = b - A x
// (FP64, stored in T), ||b||^2, and the first preconditioned direction; slots
// already within tolerance stop at once (S_DONE); x_zero: x64 is known to be zero
// (r = b, one pass fewer). Iteration: one flexible-CG step for the slots in
// mk[MK_IT]; the device folds the correction e into x64 and recomputes the true
// residual (keeping the search direction) whenever the recurrence residual has
// fallen by the slot's threshold since the last such update, or looks converged,
// or the iteration cap is reached; a slot whose true residual meets tol stops
// (mk[MK_IT] cleared, S_DONE set) without a host round trip. The threshold starts
// at delta2 and follows the drift the updates measure between the recurrence and
// the true residual: 1e-2 lower while they agree within a factor 2 (in norm),
// 1e2 higher when they differ by more than 10.
template <typename T>
void ph_rel_start(const Hierarchy& H, hipStream_t st, double* x64, const double* b64, double tol2, double delta2,
                  bool x_zero);
template <typename T>
void ph_rel_iter(const Hierarchy& H, hipStream_t st, double* x64, const double* b64, int pc, int maxit);
// initial reliable-update threshold (squared) for Krylov precision T (FP64: only
// the convergence test triggers an update)
template <typename T>
inline double rel_delta2() {
    // FP32: update when the residual norm fell ~32x (measured: 1e-4 took 9 iterations
    // at 2048^2 where 1e-3 takes 8, other grids unchanged; 1e-2 cost iterations at 1024^2)
    return sizeof(T) == 4 ? 1e-3 : 0.0;
}
// per-slot device scalars (GMG_SSTRIDE doubles per slot)
enum {
    S_RZ = 0, S_PQ, S_ALPHA, S_RR, S_BETA, S_ZQ, S_ZMEAN, S_BB, S_RR0,
    S_RREF,  // true residual^2 at the last reliable update
    S_TGT,   // tol^2 * ||b||^2
    S_ITS,   // Krylov iterations done
    S_NREP,  // true-residual evaluations (start + reliable updates)
    S_DONE,  // 1 once the slot has stopped (converged, iteration cap or breakdown)
    S_MEAN,  // scratch mean (singular projections)
    S_RREC,  // recurrence residual^2 when the last update was decided
    S_DLT2,  // current update threshold: update when rr < S_DLT2 * S_RREF
    S_N
};
static_assert(S_N <= GMG_SSTRIDE, "scalars per slot");
// FIN_RR_DECIDE: FIN_RR plus the reliable-update decision of the iteration (it counts
// the iteration and sets the slot's MK_REP flag); FIN_REP_DONE: FIN_REP plus the stop
// decision after the update's true residual (clears MK_IT of a finished slot)
enum { FIN_PQ = 0, FIN_RR, FIN_DOT3, FIN_DOT3_FIRST, FIN_BB, FIN_RR0, FIN_REP, FIN_MEAN, FIN_RR_DECIDE, FIN_REP_DONE };
// device masks per slot, rows of B ints (Hierarchy::mk)
enum { MK_NEW = 0, MK_IT, MK_REP, MK_START, MK_COUNT };
void gmg_free(Hierarchy& H);
// solve A x = b on the stream: b, x FP64 device arrays (N each). Returns iterations etc.
int gmg_solve(opf_ctx* ctx, Hierarchy& H, const double* b, double* x, double tol, int maxit, opf_solve_state* st,
              int verbose);
// every slot of H at once (b, x slot-major, B * N each); st[B]
int gmg_solve_slots(opf_ctx* ctx, Hierarchy& H, const double* b, double* x, double tol, int maxit,
                    opf_solve_state* st, int verbose);
// z = one V-cycle on r from a zero guess, every slot (FP64 hierarchies): the
// multigrid as a preconditioner of an outer Krylov method
int gmg_vcycle64(opf_ctx* ctx, Hierarchy& H, const double* r, double* z);
int gmg_apply64(opf_ctx* ctx, Hierarchy& H, const double* x, double* y);

}  // namespace opf
