/*
 * opflow_hip.h -- C ABI of the MI355X (gfx950) execution engine for OpFlow's
 * data-parallel hot path: expression-template evaluation over CartesianField
 * (`u = expr` through rangeFor + updatePadding) and the implicit solve that
 * StencilPad assembles for Laplacian-type operators.
 *
 * Conventions
 *   - every entry point returns int: OPF_OK (0) or a negative OPF_ERR_* code;
 *     opf_last_error() gives a message for the calling thread;
 *   - no exceptions cross this boundary, no torch or HIP types in signatures;
 *   - storage layout of a field is OpFlow's PlainTensor layout: an x-fastest
 *     block covering localRange +- padding, index origin `offset`
 *     (src/DataStructures/Arrays/Tensor/PlainTensor.hpp:203-212,
 *      src/Core/Field/MeshBased/Structured/CartesianField.hpp:929-938);
 *   - all device work of a context is ordered on that context's stream.
 *
 * Each declaration cites the reference interface it replaces. Paths are
 * relative to the OpFlow source tree.
 */
#ifndef OPFLOW_HIP_H
#define OPFLOW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define OPF_OK 0
#define OPF_ERR_ARG (-1)         /* invalid argument / inconsistent descriptor   */
#define OPF_ERR_HIP (-2)         /* HIP runtime error                            */
#define OPF_ERR_UNSUPPORTED (-3) /* valid OpFlow input this engine does not run  */
#define OPF_ERR_RANGE (-4)       /* a read would leave the field's storage       */
#define OPF_ERR_NOMEM (-5)       /* device allocation failed / over budget       */
#define OPF_ERR_NOTREADY (-6)    /* serving: nothing to poll yet                 */

/* ---- enums mirroring OpFlow's (same numeric order) ---------------------- */
enum { OPF_F64 = 0, OPF_F32 = 1 };
/* LocOnMesh, src/Core/Constants.hpp:50 */
enum { OPF_LOC_CORNER = 0, OPF_LOC_CENTER = 1 };
/* BCType, src/Core/BC/BCBase.hpp:24 */
enum {
    OPF_BC_UNDEFINED = 0,
    OPF_BC_DIRC = 1,
    OPF_BC_NEUM = 2,
    OPF_BC_PERIODIC = 3,
    OPF_BC_INTERNAL = 4,
    OPF_BC_SYMM = 5,
    OPF_BC_ASYMM = 6
};
/* BasicArithOp subset, src/Core/Constants.hpp:53 (Eq, Add, Minus, Mul, Div) */
enum { OPF_OP_EQ = 0, OPF_OP_ADD = 1, OPF_OP_MINUS = 2, OPF_OP_MUL = 3, OPF_OP_DIV = 4 };
/* reductions for rangeReduce, src/Core/Loops/RangeFor.hpp:86-121 */
enum { OPF_RED_SUM = 0, OPF_RED_MAX = 1, OPF_RED_MIN = 2, OPF_RED_SUMSQ = 3, OPF_RED_MAXABS = 4 };

/* ---- opaque handles ------------------------------------------------------ */
typedef struct opf_ctx opf_ctx;
typedef struct opf_mesh opf_mesh;
typedef struct opf_field opf_field;
typedef struct opf_eqn opf_eqn;
typedef struct opf_gmg opf_gmg;
typedef struct opf_srv opf_srv;

/* ---- field descriptor ---------------------------------------------------
 * Inputs: dim, dtype, loc, bc_type/bc_value, ext, padding. The range fields
 * are computed by opf_field_desc_finalize exactly as ExprBuilder::build does
 * (CartesianField.hpp:929-1029, single-process branch).                     */
typedef struct {
    int dim;
    int dtype;
    int loc[3];
    int bc_type[3][2];
    double bc_value[3][2];
    int ext[3][2];
    int padding;
    /* derived */
    int mesh_start[3], mesh_end[3];
    int local_start[3], local_end[3];
    int assign_start[3], assign_end[3];
    int access_start[3], access_end[3];
    int logical_start[3], logical_end[3];
    int offset[3]; /* index of storage element 0 (localRange.start - padding) */
    int extent[3]; /* storage extents (localRange extents + 2 padding)       */
} opf_field_desc;

/* ---- expression tree node (post-order list; children precede parents) ---
 * Lowered inside the library to a straight-line stack program. Kinds follow
 * OpFlow's operator set:                                                    */
enum {
    OPF_N_FIELD = 0,    /* a = slot in the fields[] array                         */
    OPF_N_CONST = 1,    /* value (ScalarExpr, Expr/ScalarExpr.hpp:24-61)          */
    OPF_N_ADD = 2,      /* a + b  (BinOpDefMacros.hpp.in, Arithmetic/AMDS.hpp)    */
    OPF_N_SUB = 3,
    OPF_N_MUL = 4,
    OPF_N_DIV = 5,
    OPF_N_NEG = 6,      /* -a                                                     */
    OPF_N_D2C = 7,      /* d2<D2SecondOrderCentered<dim>>(a)  D2SecondOrderCentered.hpp:160-171 */
    OPF_N_D1C = 8,      /* d1<D1FirstOrderCentered<dim>>(a)   D1FirstOrderCentered.hpp:30-37   */
    OPF_N_D1UP = 9,     /* D1FirstOrderBiasedUpwind           D1FirstOrderBiasedUpwind.hpp:53-58 */
    OPF_N_D1DOWN = 10,  /* D1FirstOrderBiasedDownwind         D1FirstOrderBiasedDownwind.hpp:52-57 */
    OPF_N_WENO_UP = 11, /* D1WENO53Upwind                     D1WENO53Upwind.hpp:74-151 */
    OPF_N_WENO_DOWN = 12, /* D1WENO53Downwind                 D1WENO53Downwind.hpp:75-154 */
    OPF_N_INTP_C2N = 13,  /* D1Linear<dim, Cen2Cor>           Interpolator/D1Linear.hpp:34-42 */
    OPF_N_INTP_N2C = 14,  /* D1Linear<dim, Cor2Cen>           Interpolator/D1Linear.hpp:43-45 */
    OPF_N_UNKNOWN = 15,   /* the StencilField of the solve target (StencilField.hpp:123-350);
                             a = the target's index in an equation set (0 otherwise) */
    /* D1FluxLimiterGen<Kernel>::Op<dim, dir>(a, b) (Interpolator/D1FluxLimiter.hpp:43-202): b
     * interpolated to the faces in `dim`, upwinded by the sign of a (a > 0: upwind form, else
     * downwind form); Cen2Cor when b is centre-located in `dim`, Cor2Cen when corner-located;
     * value = the limiter kernel OPF_LIM_* (FluxLimiterKernels.hpp) + 32 x the direction the
     * caller names (0: from b's location, 1: Cen2Cor, 2: Cor2Cen; a mismatch is OPF_ERR_ARG,
     * as the reference's prepare() asserts)                                          */
    OPF_N_FLUXLIM = 16,
    OPF_N_MIN = 17,       /* std::min(a, b) = b < a ? b : a                          */
    OPF_N_MAX = 18,       /* std::max(a, b) = a < b ? b : a                          */
    OPF_N_ABS = 19,       /* std::fabs(a)                                            */
    OPF_N_GT = 20,        /* a > b as 1 / 0 (the condition of conditional)           */
    OPF_N_LT = 21,        /* a < b                                                   */
    OPF_N_GE = 22,        /* a >= b                                                  */
    OPF_N_LE = 23,        /* a <= b                                                  */
    /* conditional(c, t, f) (Operator/Conditional.hpp:22-60): a = c, b = t, dim = the node index
     * of f; c != 0 selects t                                                          */
    OPF_N_COND = 24,
    /* f(a) for the <cmath> function dim = OPF_FN_* (Arithmetic/AMDS.hpp:52-90, DEFINE_UNIFUNC);
     * OPF_FN_SMOOTH_DELTA / OPF_FN_SMOOTH_HEAVISIDE take value = eps (Math::smoothDelta /
     * smoothHeviside, Math/Function/Numeric.hpp:86-96, the functors of UniOpAdaptor)       */
    OPF_N_FUNC1 = 25,
    /* f(a, b) for the <cmath> function dim = OPF_FN2_* (Arithmetic/AMDS.hpp:39-50, DEFINE_BINFUNC) */
    OPF_N_FUNC2 = 26,
    /* conv(a, kernel) (Convolution/Convolution.hpp:20-95): the kernel's K = n0 n1 n2 weights
     * are K consecutive OPF_N_CONST nodes from index b, x-fastest; dim = n0 | n1 << 8 |
     * n2 << 16 (odd extents, 1 in unused dims); the sum runs x-fastest over the window,
     * weight * value, the first term alone and each next one added (rangeReduce_s)   */
    OPF_N_CONV = 27
};
/* functions of OPF_N_FUNC1 */
enum {
    OPF_FN_EXP = 0, OPF_FN_EXP2 = 1, OPF_FN_EXPM1 = 2, OPF_FN_LOG = 3, OPF_FN_LOG10 = 4, OPF_FN_LOG2 = 5,
    OPF_FN_LOG1P = 6, OPF_FN_SQRT = 7, OPF_FN_CBRT = 8, OPF_FN_SIN = 9, OPF_FN_COS = 10, OPF_FN_TAN = 11,
    OPF_FN_ASIN = 12, OPF_FN_ACOS = 13, OPF_FN_ATAN = 14, OPF_FN_SINH = 15, OPF_FN_COSH = 16, OPF_FN_TANH = 17,
    OPF_FN_ASINH = 18, OPF_FN_ACOSH = 19, OPF_FN_ATANH = 20, OPF_FN_ERF = 21, OPF_FN_ERFC = 22,
    OPF_FN_TGAMMA = 23, OPF_FN_LGAMMA = 24, OPF_FN_CEIL = 25, OPF_FN_FLOOR = 26, OPF_FN_TRUNC = 27,
    OPF_FN_ROUND = 28, OPF_FN_NEARBYINT = 29, OPF_FN_RINT = 30, OPF_FN_LOGB = 31,
    OPF_FN_SMOOTH_DELTA = 32, OPF_FN_SMOOTH_HEAVISIDE = 33
};
/* functions of OPF_N_FUNC2 */
enum {
    OPF_FN2_POW = 0, OPF_FN2_FMOD = 1, OPF_FN2_REMAINDER = 2, OPF_FN2_FDIM = 3, OPF_FN2_HYPOT = 4,
    OPF_FN2_ATAN2 = 5, OPF_FN2_COPYSIGN = 6, OPF_FN2_NEXTAFTER = 7
};
/* limiter kernels of OPF_N_FLUXLIM (Interpolator/FluxLimiterKernels.hpp); the kappa schemes use
 * the linear form of D1FluxLimiter.hpp, the others the ratio form                    */
enum {
    OPF_LIM_CDS = 0, OPF_LIM_QUICK = 1, OPF_LIM_CUI = 2, OPF_LIM_FROMM = 3, OPF_LIM_LUI = 4,
    OPF_LIM_MINMOD = 5, OPF_LIM_SUPERBEE = 6, OPF_LIM_MUSCL = 7, OPF_LIM_HARMONIC = 8, OPF_LIM_VANALBADA = 9,
    OPF_LIM_UMIST = 10, OPF_LIM_SPL13 = 11, OPF_LIM_SPLMAX12 = 12, OPF_LIM_SPLMAX13 = 13,
    OPF_LIM_KOREN = 14, OPF_LIM_SMART = 15, OPF_LIM_BSOU = 16
};
typedef struct {
    int kind;
    int a, b; /* child node indices (FIELD: a = field slot)  */
    int dim;  /* operator direction for derivative / interpolation kinds */
    double value;
} opf_expr_node;

/* ---- context ------------------------------------------------------------ */
/* Replaces: the global ParallelPlan / TBB arena (src/Core/Environment.hpp:46-62,
 * src/Core/Loops/RangeFor.hpp:68-84). One context = one device + one stream. */
int opf_ctx_create(int device, opf_ctx** out);
int opf_ctx_destroy(opf_ctx* ctx);
int opf_ctx_sync(opf_ctx* ctx);
void* opf_ctx_stream(opf_ctx* ctx); /* hipStream_t as void* */
/* Enqueue the context's work on the caller's HIP stream (hipStream_t as void*;
 * null: back to the context's own stream), ordered after the work already
 * enqueued. No counterpart in the reference, which has no stream concept; it
 * lets a host that owns a stream (a framework's current stream, a graph
 * capture) place the engine's work on it. The stream must stay valid until it
 * is replaced or the context is destroyed.                                  */
int opf_ctx_set_stream(opf_ctx* ctx, void* stream);
/* Engine of expression sweeps (opf_sweep and equation right-hand sides). No
 * counterpart in the reference, whose expressions are C++ templates compiled
 * ahead of time: 0 = the stack-program interpreter, 1 = (default) a kernel
 * generated from the expression and compiled at run time (hipRTC, cached in the
 * process and under $OPF_JIT_CACHE or ~/.cache/opflow_amd/jit) for sweeps of at
 * least 2^18 cells, 2 = that kernel for every sweep. The compiled kernel has
 * two forms, chosen by grid size: fields staged in LDS while a workgroup marches
 * z (grids of at least 512 64 x 16-cell tiles x 8-plane chunks) or loads straight
 * from HBM; 3 / 4 = the compiled kernel for every sweep in the staged / direct
 * form (tuning and tests). Every mode gives the same bits. */
int opf_ctx_set_jit(opf_ctx* ctx, int mode);
/* Counters of the engine's caches (no counterpart in the reference):
 * stats[0] hipRTC compilations in this process, [1] compiled sweep kernels
 * loaded in this process, [2] of them loaded now (at most 512 per device,
 * least recently used unloaded first), [3] sweep/equation programs this context
 * uploaded, [4] of them resident now (sweep programs: at most 1024, least
 * recently used freed first). Scalar constants of an expression are launch
 * arguments, not part of a program or kernel, so a sweep repeated with a new
 * constant (an adaptive dt) reuses both. A HIP graph that captured a sweep must
 * not be replayed after its program or kernel was evicted.                  */
int opf_cache_stats(opf_ctx* ctx, long long* stats);
const char* opf_last_error(void);
const char* opf_version(void);

/* ---- mesh -------------------------------------------------------------- */
/* Host-only: node coordinates and spacings of a uniform mesh dimension with
 * `pad` ghost nodes per side, restating MeshBuilder::set1DMesh(min,max) and
 * setExtMesh(Symm) (src/Core/Mesh/Structured/CartesianMesh.hpp:209-303).
 * x has nnodes+2*pad entries, dx nnodes+2*pad-1.                           */
int opf_mesh_uniform_arrays(int nnodes, double lo, double hi, int pad, double* x, double* dx);
/* Replaces CartesianMesh (CartesianMesh.hpp:35-120): per-dim node arrays
 * over the extended range [start-pad, start+nnodes+pad).                   */
int opf_mesh_create(opf_ctx* ctx, int dim, const int start[3], const int nnodes[3], int pad,
                    const double* const x[3], const double* const dx[3], opf_mesh** out);
int opf_mesh_destroy(opf_mesh* m);

/* ---- fields --------------------------------------------------------------- */
/* Host-only: ExprBuilder::calculateRanges + validateRanges + storage shape
 * (CartesianField.hpp:929-1029). Fills the derived members of *d.           */
int opf_field_desc_finalize(opf_field_desc* d, const int mesh_start[3], const int mesh_nnodes[3]);
/* Replaces ExprBuilder<CartesianField>::build (CartesianField.hpp:929-938):
 * allocates device storage, zero-filled, then updatePadding.                 */
int opf_field_create(opf_mesh* mesh, const opf_field_desc* desc, opf_field** out);
int opf_field_destroy(opf_field* f);
const opf_field_desc* opf_field_get_desc(const opf_field* f);
size_t opf_field_nbytes(const opf_field* f);
void* opf_field_device_ptr(opf_field* f);
/* whole storage block host<->device (PlainTensor layout) */
int opf_field_upload(opf_field* f, const void* host, size_t nbytes);
int opf_field_download(const opf_field* f, void* host, size_t nbytes);
/* Replaces CartesianField::updatePaddingImpl_final (CartesianField.hpp:349-629). */
int opf_field_update_padding(opf_field* f);
/* Replaces assignImpl_final(const D& c) (CartesianField.hpp:237-280): f (op)= c
 * over assignableRange ∩ localRange, then updatePadding.                    */
int opf_field_assign_const(opf_field* f, int op, double c);
/* Replaces the copy constructor / assignImpl_final(const CartesianField&)
 * when both fields share a descriptor: device copy of the whole storage.    */
int opf_field_copy(opf_field* dst, const opf_field* src);
/* exchange the storage of two fields with identical descriptors (O(1)) */
int opf_field_swap(opf_field* a, opf_field* b);

/* ---- explicit evaluation ------------------------------------------------ */
/* Replaces Expr::operator= -> CartesianField::assignImpl_final<Op>(T&&) ->
 * internal::FieldAssigner::assign<Op> -> rangeFor(evalAt) -> updatePadding
 * (Expr/Expr.hpp:53-117, CartesianField.hpp:195-234, Loops/FieldAssigner.hpp:26-86).
 * `out (op)= tree(nodes, root)` over out's assignableRange ∩ localRange, then
 * out's updatePadding. `out` may appear among fields[] (handled as OpFlow's
 * copy-to-temporary case, without the copies).                             */
int opf_sweep(const opf_expr_node* nodes, int nnodes, int root, opf_field* const* fields, int nfields,
              opf_field* out, int op);
/* Fast path for the FTCS heat update of examples/FTCS2D/FTCS-OMP.cpp:26:
 *   repeat nsteps: u = u + coef * (d2x<D2SecondOrderCentered>(u) + d2y(u) [+ d2z(u)])
 * Requires a uniform mesh, centre location, padding = ext = 1 and Dirc/Neum/
 * Periodic BCs; returns OPF_ERR_UNSUPPORTED otherwise (use opf_sweep).      */
int opf_ftcs(opf_field* u, double coef, int nsteps);
/* Same update with the temporal-blocking choice made by the caller:
 * tblock = 1 runs one launch per step; tblock = 0 lets the library fuse steps
 * inside one launch where it can (2-D: 4 steps per launch, 5 on small grids, any BC;
 * 3-D: 3 steps per launch, the ghosts of non-periodic sides rebuilt inside the launch),
 * which is what opf_ftcs does. Results are bit-identical either way.       */
int opf_ftcs_blocked(opf_field* u, double coef, int nsteps, int tblock);
/* Launch plan of opf_ftcs_blocked(u, ., nsteps, tblock) without running it:
 * number of kernel launches and the most steps one launch advances. No
 * counterpart in the reference (its FTCS loop is one rangeFor pass per step);
 * it lets a caller relate launch timings to steps.                          */
int opf_ftcs_plan(const opf_field* u, int nsteps, int tblock, int* launches, int* max_steps);
/* The update of opf_ftcs with no device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, hosts that forbid hipMalloc). opf_ftcs keeps a
 * second time level per field, allocated on the field's first update; this
 * entry point uses a fixed device scratch of OPF_FTCS_NOALLOC_MAX_BYTES per
 * device instead (held from library load to process end), copies the field
 * into it, ping-pongs between the two and leaves the result in the field's own
 * storage. It calls no hipMalloc, hipFree or other allocating routine. A field
 * whose storage exceeds the scratch is rejected with OPF_ERR_NOMEM before any
 * work. Results are bit-identical to opf_ftcs. The scratch is shared by every
 * call on a device: calls that overlap on one device (other threads, other
 * contexts' streams) race on it; calls ordered on one stream do not.       */
#define OPF_FTCS_NOALLOC_MAX_BYTES (64u << 20)
int opf_ftcs_noalloc(opf_field* u, double coef, int nsteps);

/* ---- reductions ----------------------------------------------------------- */
/* Replaces rangeReduce (Loops/RangeFor.hpp:86-121) for the common reducers;
 * deterministic (fixed two-pass tree). range = [rs, re) in field indices.   */
int opf_reduce(const opf_field* f, int red, const int rs[3], const int re[3], double* out);

/* ---- raw device buffers (for the direct solver API) ---------------------- */
int opf_buf_create(opf_ctx* ctx, size_t nbytes, void** dptr);
int opf_buf_destroy(opf_ctx* ctx, void* dptr);
int opf_buf_upload(opf_ctx* ctx, void* dptr, const void* host, size_t nbytes);
int opf_buf_download(opf_ctx* ctx, const void* dptr, void* host, size_t nbytes);
/* zero nbytes of a device buffer, asynchronously on the context's stream */
int opf_buf_zero(opf_ctx* ctx, void* dptr, size_t nbytes);

/* ---- synthetic inputs (host-only) ----------------------------------------- */
/* SplitMix64 stream, u = (x >> 11) * 2^-53; lo + (hi-lo) u. */
int opf_gen_uniform(uint64_t seed, size_t n, double lo, double hi, double* out);
/* Fill the storage block `out` of field desc d on mesh m with
 * sum_k A_k exp(-|x - c_k|^2 / (2 w_k^2)) at cell centres (interior only),
 * c ~ U[0.2,0.8]^dim, w ~ U[0.03,0.1], A ~ U[0.5,1], plus noise * N(0,1)
 * (Box-Muller) when noise != 0. Ghosts are left zero.                      */
int opf_gen_bumps(const opf_mesh* m, const opf_field_desc* d, uint64_t seed, int nbumps, double noise,
                  double* out);

/* ---- implicit solve ----------------------------------------------------- */
/* Solver kinds, StructSolverType order (src/Core/Solvers/Struct/StructSolver.hpp:23-34) */
enum {
    OPF_SOLVER_NONE = 0,
    OPF_SOLVER_JACOBI = 1,
    OPF_SOLVER_SMG = 2,
    OPF_SOLVER_PFMG = 3,
    OPF_SOLVER_CYCRED = 4,
    OPF_SOLVER_PCG = 5,
    OPF_SOLVER_GMRES = 6,
    OPF_SOLVER_FGMRES = 7,
    OPF_SOLVER_LGMRES = 8,
    OPF_SOLVER_BICGSTAB = 9
};
/* precision policy of the multigrid-preconditioned Krylov solve */
enum { OPF_PREC_FP64 = 0, OPF_PREC_MIXED = 1 };
/* Mirrors StructSolverParamsBase (StructSolver.hpp:39-51) plus the engine's
 * own knobs. tol is a relative residual ||b-Ax||_2/||b||_2.               */
typedef struct {
    int solver;       /* OPF_SOLVER_*: Krylov kind requested by the caller        */
    int precond;      /* OPF_SOLVER_*: preconditioner requested (PFMG/SMG -> GMG)  */
    double tol;       /* <= 0: HYPRE default 1e-6                                  */
    int max_iter;     /* <= 0: 200                                                 */
    int static_mat;   /* StructSolverParamsBase::staticMat                         */
    int pin_value;    /* StructSolverParamsBase::pinValue                          */
    int precision;    /* OPF_PREC_*                                                */
    int pre_smooth;   /* <= 0: 2                                                   */
    int post_smooth;  /* <= 0: 2                                                   */
    int verbose;
} opf_solve_params;
/* Mirrors EqnSolveState (src/Core/Equation/EqnSolveHandler.hpp:17-25) */
typedef struct {
    int niter;        /* total Krylov iterations (inner iterations when mixed) */
    double relerr;    /* final ||b-Ax||_2 / ||b||_2 evaluated in FP64          */
    double abserr;    /* final ||b-Ax||_2                                      */
    int nrefine;      /* FP64 true-residual evaluations (mixed: the start and
                         each reliable update x += e, r = b - A x)           */
    int converged;
} opf_solve_state;

/* ---- field streams (OpFlow's raw binary format) -------------------------- */
/* Replaces Utils::RawBinaryOStream::operator<< (Utils/Writers/RawBinaryStream.hpp:84-155):
 * writes `<dir>/<name>_<count>.cart` (name length + name, dim, ranks = 1, time,
 * mesh range, node coordinates, accessible range, local range, then the local
 * range's values x-fastest in the field's element type), streamed from device
 * memory through pinned double buffers. dir null or "" = ".".               */
int opf_field_write_cart(opf_field* f, const char* dir, const char* name, int count, double time);
/* Replaces Utils::RawBinaryIStream::operator>> (RawBinaryStream.hpp:157-224):
 * reads `<dir>/<name>_<count>.cart` into f's local range after checking dim,
 * rank count, mesh range and coordinates, accessible and local ranges (a name
 * mismatch only warns, as the reference's OP_EXPECT_MSG does), then
 * updatePadding. time (optional) receives the stored time stamp.           */
int opf_field_read_cart(opf_field* f, const char* dir, const char* name, int count, double* time);

/* Replaces makeEqnSolveHandler(getter, target, solver) -> HYPREEqnSolveHandler
 * ::init/initStencil (HYPREEqnSolveHandler.hpp:44-117) and the handler's
 * StencilPad evaluation (StencilField.hpp:123-350): the equation
 * `tree(lhs_root) == tree(rhs_root)`, in which OPF_N_UNKNOWN stands for the
 * target's StencilField, is lowered once; its per-cell coefficients are
 * evaluated on the device.                                                */
int opf_eqn_create(const opf_expr_node* nodes, int nnodes, int lhs_root, int rhs_root,
                   opf_field* const* fields, int nfields, opf_field* target, const opf_solve_params* params,
                   opf_eqn** out);
/* Replaces HYPREEqnSolveHandler::solve (HYPREEqnSolveHandler.hpp:190-209):
 * regenerate b (and A unless static_mat), solve from the target's current
 * values, write the solution back, updatePadding. Symmetric operators run the
 * multigrid-preconditioned flexible CG (FP64 or mixed); non-symmetric ones
 * (advection terms) run BiCGStab in FP64 preconditioned by one V-cycle of the
 * operator's symmetric part, with pin_value's row kept literally.          */
int opf_eqn_solve(opf_eqn* e, opf_solve_state* state);
/* The equation's linear system in the CSR form CSRMatrixGenerator::generate /
 * generate_rhs produce (CSRMatrixGenerator.hpp:55-220, the AMGCL path of
 * AMGCLEqnSolveHandler.hpp:92-99): one row per unknown, x-fastest over the
 * target's assignable range; row = StencilPad of lhs - rhs with columns
 * ascending; rhs = -bias; with params.pin_value the LAST row becomes
 * x_last = 0 (rhs 0). The pads are evaluated on the device from the current
 * field values; the rows are assembled on the host. Call with null arrays to
 * get nrows / nnz, then with ptr[nrows+1], col[nnz], val[nnz], rhs[nrows].  */
int opf_eqn_csr(opf_eqn* e, long long* nrows, long long* nnz, long long* ptr, long long* col, double* val,
                double* rhs);
int opf_eqn_destroy(opf_eqn* e);

/* One equation of a set: `tree(lhs_root) == tree(rhs_root)` over post-order
 * nodes; OPF_N_FIELD slots index fields[]; OPF_N_UNKNOWN's `a` is the index of
 * the target (in the set's targets[]) whose StencilField it stands for, so an
 * equation may read several unknowns (EqnHolder's getters take every target's
 * StencilField, Equation/EquationHolder.hpp:31-99).                        */
typedef struct {
    const opf_expr_node* nodes;
    int nnodes;
    int lhs_root, rhs_root;
    opf_field* const* fields;
    int nfields;
} opf_eqn_desc;
/* Replaces SolveEqns<Solver>(getters..., targets..., mapper, params)
 * (Equation/UnifiedSolve.hpp:69-100): equation i is solved for targets[i]
 * (FP64, centre-located, one shared layout, at most 16). The system is the one
 * CSRMatrixGenerator builds (opf_eqns_csr), solved from x = 0 to params[0].tol
 * (default 1e-8) on the whole system's relative residual, then every target is
 * written back and its padding updated. params[i].pin_value pins block i's
 * last unknown at 0. Blocks that read only their own unknown are solved by
 * multigrid, blocks with equal operators batched as slots of one hierarchy;
 * coupled sets, and sets with a non-symmetric diagonal block, by BiCGStab
 * preconditioned with one V-cycle per diagonal block (its symmetric part).
 * state: Krylov iterations (the most of any block when block-diagonal) and the
 * whole system's residual.                                                  */
int opf_eqns_solve(int neq, const opf_eqn_desc* eqs, opf_field* const* targets, const opf_solve_params* params,
                   opf_solve_state* state);
/* Replaces CSRMatrixGenerator::generate(StencilHolder, ColoredMDRangeMapper,
 * pin) (Equation/CSRMatrixGenerator.hpp:26-41, 55-193) for a set: block i
 * holds targets[i]'s cells x-fastest (ColoredMDRangeMapper.hpp:20-41); a row
 * holds the (unknown, cell) entries the equation's pad touches, columns
 * ascending (block offset + cell rank), rhs = -bias; pin[i] makes block i's
 * LAST row x = 0. pin may be null. Null arrays return nrows / nnz only.      */
int opf_eqns_csr(int neq, const opf_eqn_desc* eqs, opf_field* const* targets, const int* pin, long long* nrows,
                 long long* nnz, long long* ptr, long long* col, double* val, double* rhs);

/* Direct multigrid API on device vectors (used by the serving loop and the
 * benchmarks): constant-coefficient operator
 *   A x = sigma x - sum_d (x_{i+e_d} - 2 x_i + x_{i-e_d}) / h_d^2
 * on n[0] x n[1] (x n[2]) cells, boundary kind per face (DIRC: ghost = -mirror,
 * NEUM: ghost = mirror, PERIODIC: wrap), i.e. the SPD form of OpFlow's
 * centre-located d2x+d2y(+d2z) with homogeneous BCs after StencilPad folding. */
typedef struct {
    int dim;
    int n[3];
    double h[3];
    int bc[3][2];
    double sigma;
} opf_poisson_desc;
int opf_gmg_create(opf_ctx* ctx, const opf_poisson_desc* d, int precision, opf_gmg** out);
/* b, x: device arrays of prod(n) doubles, x-fastest, no ghosts. */
int opf_gmg_solve(opf_gmg* g, const double* b, double* x, const opf_solve_params* p, opf_solve_state* st);
/* y = A x on device arrays (FP64), for residual checks */
int opf_gmg_apply(opf_gmg* g, const double* x, double* y);
int opf_gmg_destroy(opf_gmg* g);
size_t opf_gmg_footprint(const opf_poisson_desc* d, int precision);

/* ---- serving loop ---------------------------------------------------------- */
typedef struct {
    size_t mem_budget;   /* bytes of HBM the server may hold for requests          */
    int max_inflight;    /* concurrent requests (slots)                            */
    int precision;       /* OPF_PREC_*                                             */
    int batched;         /* 1: cross-request batched smoothing launches             */
    double tol;          /* relative residual target (1e-8)                        */
    int max_iter;
} opf_srv_config;
typedef struct {
    opf_poisson_desc desc;
    uint64_t seed;       /* RHS = nbumps seeded Gaussian bumps                      */
    int nbumps;
    double noise;
    const double* rhs;   /* optional host RHS (prod(n) doubles); NULL -> generated on the device from seed/nbumps/noise */
    double* out_x;       /* optional host buffer: the solution is copied here when done */
    double t_submit;     /* arrival time on the server clock (< 0: now)             */
} opf_srv_request;
typedef struct {
    uint64_t id;
    int status;          /* OPF_OK or error code                                    */
    opf_solve_state state;
    double t_submit, t_admit, t_done; /* seconds on the server clock             */
    double checksum;     /* sum of the solution                                     */
} opf_srv_result;
/* bumps (+ noise) on the cell centres (h*i + h/2) of the unit-cube grid of d, x-fastest */
int opf_gen_bumps_grid(const opf_poisson_desc* d, uint64_t seed, int nbumps, double noise, double* out);
/* The same right-hand side generated on the device into dev_out (prod(n)
 * doubles) on the context's stream; at most 32 bumps. Values equal the host
 * generator's up to the last bit of exp/log/cos. Synthetic-input helper with
 * no counterpart in the reference.                                          */
int opf_gen_bumps_grid_dev(opf_ctx* ctx, const opf_poisson_desc* d, uint64_t seed, int nbumps, double noise,
                           double* dev_out);
int opf_srv_create(opf_ctx* ctx, const opf_srv_config* cfg, opf_srv** out);
int opf_srv_submit(opf_srv* s, const opf_srv_request* req, uint64_t* id);
/* advance the device work by one scheduling round; returns number of finished */
int opf_srv_step(opf_srv* s, int* nfinished);
int opf_srv_poll(opf_srv* s, opf_srv_result* out, int max, int* n);
int opf_srv_pending(opf_srv* s, int* queued, int* inflight);
double opf_srv_clock(opf_srv* s);
int opf_srv_destroy(opf_srv* s);

#ifdef __cplusplus
}
#endif
#endif /* OPFLOW_HIP_H */
