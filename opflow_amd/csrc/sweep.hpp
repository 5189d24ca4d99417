// Stack program format shared by the lowering (host) and the sweep kernels.
#pragma once
#include <array>
#include <vector>

#include "opf_internal.hpp"

namespace opf {

constexpr int SWEEP_MAX_FIELDS = 8;
constexpr int SWEEP_MAX_DEPTH = 24;
// scalar constants of an expression (ScalarExpr nodes) passed at launch, not
// baked into the program: a new dt re-uses the program and the compiled kernel
constexpr int SWEEP_MAX_CONSTS = 32;
// programs longer than this run on the interpreter (compile time and code size
// of the generated kernel grow with the program times the cells per thread)
constexpr int SWEEP_JIT_MAX_INSNS = 16384;

enum InsnOp {
    I_LD = 0,     // push field[a] at linear offset b from the evaluation cell (c: the offset per
                  // dimension, (dx + 64) | (dy + 64) << 8 | (dz + 64) << 16)
    I_LDU = 1,    // push unknown a (pad mode) at packed offset c
    I_CONST = 2,  // push imm; a = 1: push the launch-time constant slot b (imm keeps its value)
    I_DX = 3,     // push mesh dx of dim a at (cell index + b)
    I_X = 4,      // push mesh x of dim a at (cell index + b)
    I_ADD = 5,
    I_SUB = 6,
    I_MUL = 7,
    I_DIV = 8,
    I_NEG = 9,
    I_WENO_UP = 10,
    I_WENO_DOWN = 11,
    I_MIN = 12,   // b < a ? b : a  (std::min(a, b))
    I_MAX = 13,   // a < b ? b : a  (std::max(a, b))
    I_ABS = 14,
    I_GT = 15,    // a > b ? 1 : 0
    I_LT = 16,
    I_GE = 17,
    I_LE = 18,
    I_SEL = 19,   // pops f, t, c: c != 0 ? t : f
    I_FN1 = 20,   // f(a) for OPF_FN_* a (opf_fn1)
    I_FN2 = 21    // f(a, b) for OPF_FN2_* a (opf_fn2)
};

struct Insn {
    int op, a, b, c;
    double imm;
};

struct LoweredProgram {
    std::vector<Insn> code;
    std::vector<double> consts;  // launch-time constants by slot (I_CONST with a = 1)
    int maxdepth = 0;
    int omin[SWEEP_MAX_FIELDS][3], omax[SWEEP_MAX_FIELDS][3];
    int tmin[3], tmax[3];
    int loc[3];
};

#define OPF_SHARED_SRC(...) __VA_ARGS__
#include "sweep_args.inc"
#undef OPF_SHARED_SRC
static_assert(sizeof(((SweepArgs*)nullptr)->f) / sizeof(FieldArg) == SWEEP_MAX_FIELDS, "field slots");
static_assert(sizeof(((SweepArgs*)nullptr)->cst) / sizeof(double) == SWEEP_MAX_CONSTS, "constant slots");
// the program bytes a sweep uploads: launch-time constants zeroed, so the
// device copy (and its cache key) depends on the expression's structure only
std::vector<Insn> structural_code(const LoweredProgram& P);

// unknown_slot: the field slot OPF_N_UNKNOWN reads; pad mode (unknown_as_field = false)
// emits I_LDU for it, an ordinary sweep (true) reads that slot like any field
int lower_tree(const opf_expr_node* nodes, int nnodes, int root, opf_field* const* fields, int nfields,
               const opf_mesh* mesh, int unknown_slot, LoweredProgram* out, bool unknown_as_field = false);
// several unknowns (equation sets): OPF_N_UNKNOWN's `a` indexes unknown_slots, the field
// slot whose layout that unknown has; pad mode keeps the index in the I_LDU instruction
int lower_tree_set(const opf_expr_node* nodes, int nnodes, int root, opf_field* const* fields, int nfields,
                   const opf_mesh* mesh, const int* unknown_slots, int nunknown, LoweredProgram* out,
                   bool unknown_as_field = false);
// the sweep of program P over the box [a.s, a.s + a.n) on the context's stream, by the
// engine the context's jit mode selects (interpreter or the compiled kernel of P)
int launch_sweep(opf_ctx* ctx, int dtype, const SweepArgs& a, const LoweredProgram& P);
// the compiled kernel of P (hipRTC, cached) and its launch; form 0 = by grid size,
// 1 = direct (loads from HBM), 2 = staged in LDS where the program allows
int jit_launch_sweep(hipStream_t st, const LoweredProgram& P, int dtype, const SweepArgs& a, int form = 0);
// the StencilPads of a pad-mode program, compiled (eqn.hip's PadArgs); OPF_ERR_UNSUPPORTED
// when the program is not a linear form the compiled kernel covers (the interpreter runs it)
struct PadArgs;
int jit_launch_pad(hipStream_t st, const LoweredProgram& P, const PadArgs& a);
// process-wide: hipRTC compilations, compiled kernels loaded, kernels loaded now
void jit_stats(long long* compiles, long long* loaded, long long* resident);
int check_bounds(const LoweredProgram& P, opf_field* const* fields, int nfields, const opf_mesh* mesh, const int s[3],
                 const int n[3]);
int mesh_arrays(const opf_mesh* m, int dtype, const void* x[3], const void* dx[3]);
// dims of m whose spacing is uniform and a power of two with a normal reciprocal in
// precision dtype: a.umask, a.hdx, a.rdx (compiled sweeps divide by h as a
// multiplication by 1 / h, which gives the same bits)
void uniform_spacing(const opf_mesh* m, int dtype, SweepArgs& a);

}  // namespace opf
