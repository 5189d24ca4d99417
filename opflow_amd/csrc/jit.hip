// Runtime specialisation of lowered sweep programs (hipRTC).
//
// The stack program of a sweep (lower_tree) becomes straight-line HIP source:
// one named value per instruction, every field and mesh array addressed once
// per cell, the operations and their order exactly those of the interpreter
// (sweep_kernel), so results have the same bits. It plays the role of the
// reference's compile-time expression templates (Expression.hpp:51-65), at
// run time: the source is compiled for gfx950 with hipRTC (-O3, no FMA
// contraction, no packed FP32), loaded as a module and cached in the process
// and on disk (code objects keyed by a hash of source, options, compiler
// version and device), so
// each expression is compiled once per machine. The context's jit mode picks
// the engine per sweep: the interpreter, the compiled kernel for sweeps of at
// least 2^18 cells (default), or the compiled kernel always.
#include <dirent.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <utility>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "sweep.hpp"

namespace opf {
#define OPF_SHARED_SRC(...) __VA_ARGS__
#include "pad_args.inc"
#undef OPF_SHARED_SRC
}  // namespace opf

namespace opf {

namespace {

#define OPF_SHARED_SRC(...) #__VA_ARGS__
const char* const kArgsSrc =
#include "sweep_args.inc"
    ;
const char* const kWenoSrc =
#include "weno_device.inc"
    ;
const char* const kMathSrc =
#include "math_device.inc"
    ;
const char* const kPadArgsSrc =
#include "pad_args.inc"
    ;
#undef OPF_SHARED_SRC

const char* const kOpts[] = {"--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-Xclang", "-target-feature",
                             "-Xclang", "-packed-fp32-ops"};
constexpr int kNOpts = sizeof(kOpts) / sizeof(kOpts[0]);

uint64_t fnv1a64(const void* p, size_t n, uint64_t h = 1469598103934665603ULL) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t q = 0; q < n; ++q) h = (h ^ b[q]) * 1099511628211ULL;
    return h;
}

std::string hexkey(uint64_t h) {
    char buf[32];
    snprintf(buf, sizeof buf, "%016llx", (unsigned long long)h);
    return buf;
}

// Values a generator knows to be s * h_d, a power-of-two multiple s of the uniform
// power-of-two spacing h_d of dim d (a dx of that dim, sums of two such, scalings by
// a power-of-two literal), and literal constants. A division by s * h_d is emitted
// as the multiplication by 1 / (s * h_d): both round the same real number, so the
// bits are the reference's.
struct Known {
    std::vector<signed char> d;  // dim of a known multiple, -1 none
    std::vector<double> s;       // its multiple
    std::vector<char> lit;       // a literal constant
    std::vector<double> c;       // its value
    explicit Known(size_t n) : d(n, -1), s(n, 0.0), lit(n, 0), c(n, 0.0) {}
    static bool pow2(double v, int lo, int hi) {
        int e;
        return v > 0 && std::frexp(v, &e) == 0.5 && e - 1 >= lo && e - 1 <= hi;
    }
    void dx(int id, int dim) { d[id] = (signed char)dim, s[id] = 1.0; }
    void literal(int id, double v) { lit[id] = 1, c[id] = v; }
    // result id of x op y
    void arith(int op, int id, int x, int y) {
        if (op == I_ADD && d[x] >= 0 && d[x] == d[y] && pow2(s[x] + s[y], -2, 2)) d[id] = d[x], s[id] = s[x] + s[y];
        if (op == I_MUL) {
            if (d[x] >= 0 && lit[y] && pow2(s[x] * c[y], -2, 2)) d[id] = d[x], s[id] = s[x] * c[y];
            if (d[y] >= 0 && lit[x] && pow2(s[y] * c[x], -2, 2)) d[id] = d[y], s[id] = s[y] * c[x];
        }
    }
    // the multiplication replacing x / y, or an empty string
    std::string div(int x, int y) const {
        if (d[y] < 0) return "";
        char b[96];
        snprintf(b, sizeof b, "v%d * ((T)a.rdx[%d] * (T)%a)", x, d[y], 1.0 / s[y]);
        return b;
    }
};

// the straight-line kernel of program P in precision T, writing out (op) value
// the block's tile (bx, by, bz) from its hardware id: workgroups are dispatched round-robin
// over the 8 XCDs (each with its own L2), so the flat id is remapped to give every XCD one
// contiguous range of tiles (xcd_block, fastdiv.hpp): x and y neighbours, whose halo rows
// are the same cache lines, then share one L2 instead of each fetching them from HBM
const char* const kXcdBlockSrc =
    "  const unsigned gx_ = gridDim.x, gy_ = gridDim.y, G_ = gx_ * gy_ * gridDim.z;\n"
    "  const unsigned lin_ = blockIdx.x + gx_ * (blockIdx.y + gy_ * blockIdx.z);\n"
    "  const unsigned q_ = G_ >> 3, r_ = G_ & 7u, xb_ = lin_ & 7u;\n"
    "  const unsigned L_ = xb_ * q_ + (xb_ < r_ ? xb_ : r_) + (lin_ >> 3);\n"
    "  const int bx = (int)(L_ % gx_), by = (int)((L_ / gx_) % gy_), bz = (int)(L_ / (gx_ * gy_));\n";

int gen_source(const LoweredProgram& P, int dtype, int op, int R, const SweepArgs& a, std::string& s) {
    s.clear();
    s += dtype == OPF_F64 ? "typedef double T;\n" : "typedef float T;\n";
    s += kArgsSrc;
    s += "\n";
    s += kWenoSrc;
    s += kMathSrc;
    {
        char hdr[2048];
        snprintf(hdr, sizeof hdr,
                 "\nextern \"C\" __global__ __launch_bounds__(256) void opf_sweep_jit(SweepArgs a) {\n"
                 "%s"
                 "  const int i = a.s[0] + bx * 64 + (int)threadIdx.x;\n"
                 "  const int j = a.s[1] + by * 4 + (int)threadIdx.y;\n"
                 "  if (i >= a.s[0] + a.n[0] || j >= a.s[1] + a.n[1]) return;\n"
                 "  const int k = a.s[2] + bz * %d;\n",
                 kXcdBlockSrc, R);
        s += hdr;
    }
    bool fu[SWEEP_MAX_FIELDS] = {}, dxu[3] = {}, xu[3] = {};
    for (const Insn& in : P.code) {
        if (in.op == I_LD) {
            OPF_CHECK(in.a >= 0 && in.a < SWEEP_MAX_FIELDS, OPF_ERR_ARG, "jit: field slot %d", in.a);
            fu[in.a] = true;
        }
        if ((in.op == I_DX || in.op == I_X) && (in.a < 0 || in.a > 2)) return OPF_ERR_ARG;
        if (in.op == I_DX) dxu[in.a] = true;
        if (in.op == I_X) xu[in.a] = true;
        OPF_CHECK(in.op != I_LDU, OPF_ERR_ARG, "jit: pad-mode program");
    }
    char line[512];
    for (int f = 0; f < SWEEP_MAX_FIELDS; ++f) {
        if (!fu[f]) continue;
        snprintf(line, sizeof line,
                 "  const T* __restrict__ F%d = (const T*)a.f[%d].ptr + ((i - a.f[%d].off0) + (long long)a.f[%d].e0 * "
                 "((j - a.f[%d].off1) + (long long)a.f[%d].e1 * (k - a.f[%d].off2)));\n",
                 f, f, f, f, f, f, f);
        s += line;
    }
    const char* cell[3] = {"i", "j", "k"};
    for (int d = 0; d < 3; ++d) {
        if (dxu[d]) {
            snprintf(line, sizeof line, "  const T* __restrict__ DX%d = (const T*)a.dx[%d] + (%s - a.morg[%d]);\n", d, d,
                     cell[d], d);
            s += line;
        }
        if (xu[d]) {
            snprintf(line, sizeof line, "  const T* __restrict__ X%d = (const T*)a.x[%d] + (%s - a.morg[%d]);\n", d, d,
                     cell[d], d);
            s += line;
        }
    }
    std::vector<int> st;  // stack of value ids
    std::vector<int> res;  // result value of each of the R cells
    const int NV = (int)P.code.size();
    Known kn((size_t)R * NV);
    auto pop = [&](int& v) -> bool {
        if (st.empty()) return false;
        v = st.back();
        st.pop_back();
        return true;
    };
    // cell r of the thread's R cells (k + r) uses value ids r * NV + (instruction index)
    for (int r = 0; r < R; ++r) {
        int nv = r * NV;
        st.clear();
        for (const Insn& in : P.code) {
            const int id = nv;
            switch (in.op) {
                case I_LD:
                    if (r == 0) snprintf(line, sizeof line, "  const T v%d = F%d[%d];\n", id, in.a, in.b);
                    else  // a constant offset, so loads that cells share are one load
                        snprintf(line, sizeof line, "  const T v%d = F%d[%lldLL];\n", id, in.a,
                                 in.b + (long long)r * a.f[in.a].e0 * a.f[in.a].e1);
                    break;
                case I_CONST: {
                    if (in.a) {  // launch-time constant: the source does not depend on its value
                        snprintf(line, sizeof line, "  const T v%d = (T)a.cst[%d];\n", id, in.b);
                        break;
                    }
                    long long bits;
                    std::memcpy(&bits, &in.imm, 8);
                    snprintf(line, sizeof line, "  const T v%d = (T)__longlong_as_double(%lldLL);\n", id, bits);
                    kn.literal(id, in.imm);
                } break;
                case I_DX:
                    if (a.umask >> in.a & 1) {
                        snprintf(line, sizeof line, "  const T v%d = (T)a.hdx[%d];\n", id, in.a);
                        kn.dx(id, in.a);
                    } else {
                        snprintf(line, sizeof line, "  const T v%d = DX%d[%d];\n", id, in.a, in.b + (in.a == 2 ? r : 0));
                    }
                    break;
                case I_X:
                    snprintf(line, sizeof line, "  const T v%d = X%d[%d];\n", id, in.a, in.b + (in.a == 2 ? r : 0));
                    break;
                case I_ADD:
                case I_SUB:
                case I_MUL:
                case I_DIV: {
                    int x, y;
                    if (!pop(y) || !pop(x)) return OPF_ERR_ARG;
                    const char o = in.op == I_ADD ? '+' : in.op == I_SUB ? '-' : in.op == I_MUL ? '*' : '/';
                    const std::string mul = in.op == I_DIV ? kn.div(x, y) : std::string();
                    if (!mul.empty())
                        snprintf(line, sizeof line, "  const T v%d = %s;\n", id, mul.c_str());
                    else
                        snprintf(line, sizeof line, "  const T v%d = v%d %c v%d;\n", id, x, o, y);
                    kn.arith(in.op, id, x, y);
                } break;
                case I_NEG: {
                    int x;
                    if (!pop(x)) return OPF_ERR_ARG;
                    snprintf(line, sizeof line, "  const T v%d = -v%d;\n", id, x);
                } break;
                case I_MIN:
                case I_MAX:
                case I_GT:
                case I_LT:
                case I_GE:
                case I_LE: {
                    int x, y;
                    if (!pop(y) || !pop(x)) return OPF_ERR_ARG;
                    if (in.op == I_MIN)  // std::min(a, b) = b < a ? b : a
                        snprintf(line, sizeof line, "  const T v%d = v%d < v%d ? v%d : v%d;\n", id, y, x, y, x);
                    else if (in.op == I_MAX)  // std::max(a, b) = a < b ? b : a
                        snprintf(line, sizeof line, "  const T v%d = v%d < v%d ? v%d : v%d;\n", id, x, y, y, x);
                    else
                        snprintf(line, sizeof line, "  const T v%d = v%d %s v%d ? (T)1 : (T)0;\n", id, x,
                                 in.op == I_GT ? ">" : in.op == I_LT ? "<" : in.op == I_GE ? ">=" : "<=", y);
                } break;
                case I_ABS: {
                    int x;
                    if (!pop(x)) return OPF_ERR_ARG;
                    snprintf(line, sizeof line, "  const T v%d = (T)fabs((double)v%d);\n", id, x);
                } break;
                case I_FN1: {
                    int x;
                    if (!pop(x)) return OPF_ERR_ARG;
                    snprintf(line, sizeof line, "  const T v%d = opf_fn1<T>(%d, v%d);\n", id, in.a, x);
                } break;
                case I_FN2: {
                    int x, y;
                    if (!pop(y) || !pop(x)) return OPF_ERR_ARG;
                    snprintf(line, sizeof line, "  const T v%d = opf_fn2<T>(%d, v%d, v%d);\n", id, in.a, x, y);
                } break;
                case I_SEL: {
                    int c, t_, f_;
                    if (!pop(f_) || !pop(t_) || !pop(c)) return OPF_ERR_ARG;
                    snprintf(line, sizeof line, "  const T v%d = v%d != (T)0 ? v%d : v%d;\n", id, c, t_, f_);
                } break;
                case I_WENO_UP:
                case I_WENO_DOWN: {
                    int h, v[6];
                    if (!pop(h)) return OPF_ERR_ARG;
                    for (int q = 5; q >= 0; --q)
                        if (!pop(v[q])) return OPF_ERR_ARG;
                    snprintf(line, sizeof line, "  const T v%d = weno_kernel<T>(v%d, v%d, v%d, v%d, v%d, v%d, v%d, %s);\n",
                             id, v[0], v[1], v[2], v[3], v[4], v[5], h, in.op == I_WENO_UP ? "true" : "false");
                } break;
                default: return OPF_ERR_ARG;
            }
            s += line;
            st.push_back(id);
            ++nv;
        }
        if (st.size() != 1) return OPF_ERR_ARG;
        res.push_back(st.back());
    }
    // every load of the R cells precedes the first store
    s += "  const long long lo = (i - a.oo[0]) + (long long)a.oe0 * ((j - a.oo[1]) + (long long)a.oe1 * (k - a.oo[2]));\n";
    if (R > 1) s += "  const long long OP = (long long)a.oe0 * a.oe1;\n";
    for (int r = 0; r < R; ++r) {
        char at[48];
        snprintf(at, sizeof at, r ? "lo + %d * OP" : "lo", r);
        const int v = res[r];
        switch (op) {
            case OPF_OP_EQ: snprintf(line, sizeof line, "  ((T*)a.out)[%s] = v%d;\n", at, v); break;
            case OPF_OP_ADD: snprintf(line, sizeof line, "  ((T*)a.out)[%s] = ((const T*)a.old)[%s] + v%d;\n", at, at, v); break;
            case OPF_OP_MINUS: snprintf(line, sizeof line, "  ((T*)a.out)[%s] = ((const T*)a.old)[%s] - v%d;\n", at, at, v); break;
            case OPF_OP_MUL: snprintf(line, sizeof line, "  ((T*)a.out)[%s] = ((const T*)a.old)[%s] * v%d;\n", at, at, v); break;
            case OPF_OP_DIV:
                snprintf(line, sizeof line, "  ((T*)a.out)[%s] = ((const T*)a.old)[%s] / v%d;\n", at, at, v);
                break;
            default: return OPF_ERR_ARG;
        }
        s += line;
    }
    s += "}\n";
    return OPF_OK;
}

// ---------------------------------------------------------------------------
// Tiled plane-marching form. A 64 x NTY workgroup owns a 64 x TY tile of the
// x-y plane (TY = NTY * RY: each thread RY rows of one column) and marches a
// chunk of a.zc planes in z (a launch argument). Every field the program reads is staged in LDS as a
// ring of planes covering the tile plus the program's x/y read offsets; the
// ring holds the planes at the z offsets the program reads plus one, so one
// barrier per plane separates a plane's store from its reads (the slot written
// at plane k + 1 was last read at plane k - 1). The next plane is loaded into
// registers while the current one is computed. Each load of the program reads
// its field's ring at a constant offset from the thread's base. Same operations
// per cell as the direct form (and as the interpreter), so the same bits.
struct TileField {
    bool used = false;
    int oxn = 0, oxp = 0, oyn = 0, oyp = 0, ozn = 0, ozp = 0;  // read offsets (min, max) per dimension
    int W = 0, H = 0, NZ = 0, NQ = 0;                          // tile width, height, ring planes, loads per thread
};
constexpr int kTileNTY = 4;

bool tile_layout(const LoweredProgram& P, bool plane, int RY, TileField (&tf)[SWEEP_MAX_FIELDS], size_t* lds_bytes,
                 int tsize) {
    for (auto& f : tf) f = TileField{};
    for (const Insn& in : P.code) {
        if (in.op != I_LD) continue;
        TileField& f = tf[in.a];
        const int dx = (in.c & 255) - 64, dy = ((in.c >> 8) & 255) - 64, dz = ((in.c >> 16) & 255) - 64;
        if (!f.used) {
            f.used = true;
            f.oxn = f.oxp = dx;
            f.oyn = f.oyp = dy;
            f.ozn = f.ozp = dz;
        }
        f.oxn = std::min(f.oxn, dx), f.oxp = std::max(f.oxp, dx);
        f.oyn = std::min(f.oyn, dy), f.oyp = std::max(f.oyp, dy);
        f.ozn = std::min(f.ozn, dz), f.ozp = std::max(f.ozp, dz);
    }
    size_t bytes = 0;
    for (auto& f : tf) {
        if (!f.used) continue;
        if (plane && (f.ozn != 0 || f.ozp != 0)) return false;
        f.W = 64 + f.oxp - f.oxn;
        f.H = kTileNTY * RY + f.oyp - f.oyn;
        f.NZ = plane ? 1 : f.ozp - f.ozn + 2;
        f.NQ = (f.W * f.H + 255) / 256;
        bytes += (size_t)f.NZ * f.W * f.H * tsize;
    }
    *lds_bytes = bytes;
    return true;
}

int gen_source_tile(const LoweredProgram& P, int dtype, int op, int RY, bool plane, int umask,
                    const TileField (&tf)[SWEEP_MAX_FIELDS], std::string& s) {
    s.clear();
    s += dtype == OPF_F64 ? "typedef double T;\n" : "typedef float T;\n";
    s += kArgsSrc;
    s += "\n";
    s += kWenoSrc;
    s += kMathSrc;
    const int TY = kTileNTY * RY;
    char line[2048];
    snprintf(line, sizeof line,
             "\nextern \"C\" __global__ __launch_bounds__(256) void opf_sweep_jit(SweepArgs a) {\n"
             "%s"
             "  const int tx = (int)threadIdx.x, ty = (int)threadIdx.y, t = ty * 64 + tx;\n"
             "  const int x0 = a.s[0] + bx * 64, y0 = a.s[1] + by * %d;\n"
             "  const int kb = a.s[2] + bz * a.zc;\n"
             "  const int ke = min(kb + a.zc, a.s[2] + a.n[2]);\n"
             "  const int xe = a.s[0] + a.n[0], ye = a.s[1] + a.n[1];\n"
             "  const int i = x0 + tx, jb = y0 + ty * %d;\n",
             kXcdBlockSrc, TY, RY);
    s += line;
    bool dxu[3] = {}, xu[3] = {};
    for (const Insn& in : P.code) {
        if (in.op == I_DX) dxu[in.a] = true;
        if (in.op == I_X) xu[in.a] = true;
        OPF_CHECK(in.op != I_LDU, OPF_ERR_ARG, "jit: pad-mode program");
    }
    // per field: the ring, the tile origin in storage coordinates, this thread's load slots
    for (int f = 0; f < SWEEP_MAX_FIELDS; ++f) {
        const TileField& F = tf[f];
        if (!F.used) continue;
        const int HW = F.W * F.H;
        snprintf(line, sizeof line,
                 "  __shared__ T L%d[%d];\n"
                 "  const T* __restrict__ G%d = (const T*)a.f[%d].ptr;\n"
                 "  const long long PS%d = (long long)a.f[%d].e0 * a.f[%d].e1;\n"
                 "  const int gx%d = x0 + (%d) - a.f[%d].off0, gy%d = y0 + (%d) - a.f[%d].off1;\n"
                 "  const int xl%d = xe - 1 + (%d) - a.f[%d].off0, yl%d = ye - 1 + (%d) - a.f[%d].off1;\n",
                 f, F.NZ * HW, f, f, f, f, f, f, F.oxn, f, f, F.oyn, f, f, F.oxp, f, f, F.oyp, f);
        s += line;
        for (int q = 0; q < F.NQ; ++q) {
            // position e = t + 256 q of the W x H tile: valid inside the tile and the sweep's read box
            snprintf(line, sizeof line,
                     "  const int e%d_%d = t + %d, lx%d_%d = e%d_%d %% %d, ly%d_%d = e%d_%d / %d;\n"
                     "  const bool ok%d_%d = e%d_%d < %d && gx%d + lx%d_%d <= xl%d && gy%d + ly%d_%d <= yl%d;\n"
                     "  const long long o%d_%d = (long long)(gx%d + lx%d_%d) + (long long)a.f[%d].e0 * (gy%d + ly%d_%d);\n",
                     f, q, 256 * q, f, q, f, q, F.W, f, q, f, q, F.W,
                     f, q, f, q, HW, f, f, q, f, f, f, q, f,
                     f, q, f, f, q, f, f, f, q);
            s += line;
        }
        snprintf(line, sizeof line, "  const int tb%d = (ty * %d - (%d)) * %d + (tx - (%d));\n", f, RY, F.oyn, F.W,
                 F.oxn);
        s += line;
    }
    // mesh tables (clamped to the last live cell, so dead lanes read inside the tables)
    if (dxu[0] || xu[0]) s += "  const int ic = min(i, xe - 1);\n";
    for (int d = 0; d < 2; ++d) {
        if (dxu[d] && d == 0) s += "  const T* __restrict__ DX0 = (const T*)a.dx[0] + (ic - a.morg[0]);\n";
        if (xu[d] && d == 0) s += "  const T* __restrict__ X0 = (const T*)a.x[0] + (ic - a.morg[0]);\n";
    }
    if (dxu[1] || xu[1])
        for (int r = 0; r < RY; ++r) {
            if (dxu[1]) {
                snprintf(line, sizeof line,
                         "  const T* __restrict__ DX1_%d = (const T*)a.dx[1] + (min(jb + %d, ye - 1) - a.morg[1]);\n",
                         r, r);
                s += line;
            }
            if (xu[1]) {
                snprintf(line, sizeof line,
                         "  const T* __restrict__ X1_%d = (const T*)a.x[1] + (min(jb + %d, ye - 1) - a.morg[1]);\n", r,
                         r);
                s += line;
            }
        }
    // prologue: planes kb + ozn .. kb + ozp - 1 into their slots, plane kb + ozp into registers
    for (int f = 0; f < SWEEP_MAX_FIELDS; ++f) {
        const TileField& F = tf[f];
        if (!F.used) continue;
        const int HW = F.W * F.H;
        if (!plane) {
            snprintf(line, sizeof line, "  for (int p = kb + (%d); p < kb + (%d); ++p) {\n", F.ozn, F.ozp);
            s += line;
            for (int q = 0; q < F.NQ; ++q) {
                snprintf(line, sizeof line,
                         "    if (ok%d_%d) L%d[((p - kb - (%d)) %% %d) * %d + e%d_%d] = G%d[o%d_%d + PS%d * (p - a.f[%d].off2)];\n",
                         f, q, f, F.ozn, F.NZ, HW, f, q, f, f, q, f, f);
                s += line;
            }
            s += "  }\n";
        }
        for (int q = 0; q < F.NQ; ++q) {
            snprintf(line, sizeof line,
                     "  T pf%d_%d = ok%d_%d ? G%d[o%d_%d + PS%d * (kb + (%d) - a.f[%d].off2)] : (T)0;\n", f, q, f, q, f,
                     f, q, f, F.ozp, f);
            s += line;
        }
    }
    s += "  int ring = 0;  // (k - kb) modulo each ring's length, advanced per plane\n";
    s += "  for (int k = kb; k < ke; ++k) {\n";
    for (int f = 0; f < SWEEP_MAX_FIELDS; ++f) {
        const TileField& F = tf[f];
        if (!F.used) continue;
        const int HW = F.W * F.H;
        // store the prefetched plane k + ozp into its slot
        snprintf(line, sizeof line, "    { int sl = ring %% %d + (%d); sl -= sl >= %d ? %d : 0;\n", F.NZ, F.ozp - F.ozn,
                 F.NZ, F.NZ);
        s += line;
        for (int q = 0; q < F.NQ; ++q) {
            snprintf(line, sizeof line, "      if (ok%d_%d) L%d[sl * %d + e%d_%d] = pf%d_%d;\n", f, q, f, HW, f, q, f, q);
            s += line;
        }
        s += "    }\n";
    }
    s += "    __syncthreads();\n";
    for (int f = 0; f < SWEEP_MAX_FIELDS; ++f) {
        const TileField& F = tf[f];
        if (!F.used || plane) continue;
        s += "    if (k + 1 < ke) {\n";
        for (int q = 0; q < F.NQ; ++q) {
            snprintf(line, sizeof line, "      if (ok%d_%d) pf%d_%d = G%d[o%d_%d + PS%d * (k + 1 + (%d) - a.f[%d].off2)];\n",
                     f, q, f, q, f, f, q, f, F.ozp, f);
            s += line;
        }
        s += "    }\n";
    }
    // ring bases of the planes the program reads
    for (int f = 0; f < SWEEP_MAX_FIELDS; ++f) {
        const TileField& F = tf[f];
        if (!F.used) continue;
        const int HW = F.W * F.H;
        for (int dz = F.ozn; dz <= F.ozp; ++dz) {
            snprintf(line, sizeof line,
                     "    const T* __restrict__ B%d_%d; { int sl = ring %% %d + (%d); sl -= sl >= %d ? %d : 0; "
                     "B%d_%d = L%d + sl * %d + tb%d; }\n",
                     f, dz + 64, F.NZ, dz - F.ozn, F.NZ, F.NZ, f, dz + 64, f, HW, f);
            s += line;
        }
    }
    if (dxu[2]) s += "    const T* __restrict__ DX2 = (const T*)a.dx[2] + (k - a.morg[2]);\n";
    if (xu[2]) s += "    const T* __restrict__ X2 = (const T*)a.x[2] + (k - a.morg[2]);\n";
    std::vector<int> st;
    std::vector<int> res;
    const int NV = (int)P.code.size();
    Known kn((size_t)RY * NV);
    auto pop = [&](int& v) -> bool {
        if (st.empty()) return false;
        v = st.back();
        st.pop_back();
        return true;
    };
    for (int r = 0; r < RY; ++r) {
        int nv = r * NV;
        st.clear();
        for (const Insn& in : P.code) {
            const int id = nv;
            switch (in.op) {
                case I_LD: {
                    const TileField& F = tf[in.a];
                    const int dx = (in.c & 255) - 64, dy = ((in.c >> 8) & 255) - 64, dz = ((in.c >> 16) & 255) - 64;
                    snprintf(line, sizeof line, "    const T v%d = B%d_%d[%d];\n", id, in.a, dz + 64,
                             (r + dy) * F.W + dx);
                } break;
                case I_CONST: {
                    if (in.a) {
                        snprintf(line, sizeof line, "    const T v%d = (T)a.cst[%d];\n", id, in.b);
                        break;
                    }
                    long long bits;
                    std::memcpy(&bits, &in.imm, 8);
                    snprintf(line, sizeof line, "    const T v%d = (T)__longlong_as_double(%lldLL);\n", id, bits);
                    kn.literal(id, in.imm);
                } break;
                case I_DX:
                case I_X: {
                    if (in.op == I_DX && (umask >> in.a & 1)) {
                        snprintf(line, sizeof line, "    const T v%d = (T)a.hdx[%d];\n", id, in.a);
                        kn.dx(id, in.a);
                        break;
                    }
                    const char* tb = in.op == I_DX ? "DX" : "X";
                    if (in.a == 1) snprintf(line, sizeof line, "    const T v%d = %s1_%d[%d];\n", id, tb, r, in.b);
                    else snprintf(line, sizeof line, "    const T v%d = %s%d[%d];\n", id, tb, in.a, in.b);
                } break;
                case I_ADD:
                case I_SUB:
                case I_MUL:
                case I_DIV: {
                    int x, y;
                    if (!pop(y) || !pop(x)) return OPF_ERR_ARG;
                    const char o = in.op == I_ADD ? '+' : in.op == I_SUB ? '-' : in.op == I_MUL ? '*' : '/';
                    const std::string mul = in.op == I_DIV ? kn.div(x, y) : std::string();
                    if (!mul.empty())
                        snprintf(line, sizeof line, "    const T v%d = %s;\n", id, mul.c_str());
                    else
                        snprintf(line, sizeof line, "    const T v%d = v%d %c v%d;\n", id, x, o, y);
                    kn.arith(in.op, id, x, y);
                } break;
                case I_NEG: {
                    int x;
                    if (!pop(x)) return OPF_ERR_ARG;
                    snprintf(line, sizeof line, "    const T v%d = -v%d;\n", id, x);
                } break;
                case I_MIN:
                case I_MAX:
                case I_GT:
                case I_LT:
                case I_GE:
                case I_LE: {
                    int x, y;
                    if (!pop(y) || !pop(x)) return OPF_ERR_ARG;
                    if (in.op == I_MIN)  // std::min(a, b) = b < a ? b : a
                        snprintf(line, sizeof line, "      const T v%d = v%d < v%d ? v%d : v%d;\n", id, y, x, y, x);
                    else if (in.op == I_MAX)  // std::max(a, b) = a < b ? b : a
                        snprintf(line, sizeof line, "      const T v%d = v%d < v%d ? v%d : v%d;\n", id, x, y, y, x);
                    else
                        snprintf(line, sizeof line, "      const T v%d = v%d %s v%d ? (T)1 : (T)0;\n", id, x,
                                 in.op == I_GT ? ">" : in.op == I_LT ? "<" : in.op == I_GE ? ">=" : "<=", y);
                } break;
                case I_ABS: {
                    int x;
                    if (!pop(x)) return OPF_ERR_ARG;
                    snprintf(line, sizeof line, "      const T v%d = (T)fabs((double)v%d);\n", id, x);
                } break;
                case I_FN1: {
                    int x;
                    if (!pop(x)) return OPF_ERR_ARG;
                    snprintf(line, sizeof line, "      const T v%d = opf_fn1<T>(%d, v%d);\n", id, in.a, x);
                } break;
                case I_FN2: {
                    int x, y;
                    if (!pop(y) || !pop(x)) return OPF_ERR_ARG;
                    snprintf(line, sizeof line, "      const T v%d = opf_fn2<T>(%d, v%d, v%d);\n", id, in.a, x, y);
                } break;
                case I_SEL: {
                    int c, t_, f_;
                    if (!pop(f_) || !pop(t_) || !pop(c)) return OPF_ERR_ARG;
                    snprintf(line, sizeof line, "      const T v%d = v%d != (T)0 ? v%d : v%d;\n", id, c, t_, f_);
                } break;
                case I_WENO_UP:
                case I_WENO_DOWN: {
                    int h, v[6];
                    if (!pop(h)) return OPF_ERR_ARG;
                    for (int q = 5; q >= 0; --q)
                        if (!pop(v[q])) return OPF_ERR_ARG;
                    snprintf(line, sizeof line,
                             "    const T v%d = weno_kernel<T>(v%d, v%d, v%d, v%d, v%d, v%d, v%d, %s);\n", id, v[0],
                             v[1], v[2], v[3], v[4], v[5], h, in.op == I_WENO_UP ? "true" : "false");
                } break;
                default: return OPF_ERR_ARG;
            }
            s += line;
            st.push_back(id);
            ++nv;
        }
        if (st.size() != 1) return OPF_ERR_ARG;
        res.push_back(st.back());
    }
    s += "    if (i < xe) {\n";
    s += "      const long long lo = (i - a.oo[0]) + (long long)a.oe0 * ((jb - a.oo[1]) + (long long)a.oe1 * (k - a.oo[2]));\n";
    for (int r = 0; r < RY; ++r) {
        char at[48];
        snprintf(at, sizeof at, "lo + %d * a.oe0", r);
        const int v = res[r];
        char val[160];
        switch (op) {
            case OPF_OP_EQ: snprintf(val, sizeof val, "v%d", v); break;
            case OPF_OP_ADD: snprintf(val, sizeof val, "((const T*)a.old)[%s] + v%d", at, v); break;
            case OPF_OP_MINUS: snprintf(val, sizeof val, "((const T*)a.old)[%s] - v%d", at, v); break;
            case OPF_OP_MUL: snprintf(val, sizeof val, "((const T*)a.old)[%s] * v%d", at, v); break;
            case OPF_OP_DIV: snprintf(val, sizeof val, "((const T*)a.old)[%s] / v%d", at, v); break;
            default: return OPF_ERR_ARG;
        }
        snprintf(line, sizeof line, "      if (jb + %d < ye) ((T*)a.out)[%s] = %s;\n", r, at, val);
        s += line;
    }
    s += "    }\n";
    s += "    ring = ring + 1;\n";
    s += "  }\n}\n";
    return OPF_OK;
}

std::string cache_dir() {
    const char* e = getenv("OPF_JIT_CACHE");
    if (e && *e) return e;
    const char* home = getenv("HOME");
    return home && *home ? std::string(home) + "/.cache/opflow_amd/jit" : std::string();
}

bool mkdirs(const std::string& path) {
    for (size_t p = 1; p <= path.size(); ++p) {
        if (p == path.size() || path[p] == '/') {
            const std::string sub = path.substr(0, p);
            if (mkdir(sub.c_str(), 0755) != 0 && errno != EEXIST) return false;
        }
    }
    return true;
}

bool read_file(const std::string& path, std::vector<char>& out) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    bool ok = n > 0;
    if (ok) {
        out.resize((size_t)n);
        ok = fread(out.data(), 1, (size_t)n, f) == (size_t)n;
    }
    fclose(f);
    return ok;
}

// at most this many code objects stay in the disk cache (oldest removed first)
constexpr size_t kMaxDiskEntries = 1024;

void prune_dir(const std::string& dir) {
    DIR* dp = opendir(dir.c_str());
    if (!dp) return;
    std::vector<std::pair<double, std::string>> files;
    while (dirent* de = readdir(dp)) {
        const std::string n = de->d_name;
        if (n.size() < 10 || n.compare(0, 6, "sweep_") != 0 || n.compare(n.size() - 3, 3, ".co") != 0) continue;
        struct stat sb;
        if (stat((dir + "/" + n).c_str(), &sb) == 0)
            files.emplace_back((double)sb.st_mtim.tv_sec + 1e-9 * sb.st_mtim.tv_nsec, n);
    }
    closedir(dp);
    if (files.size() <= kMaxDiskEntries) return;
    std::sort(files.begin(), files.end());
    for (size_t q = 0; q + kMaxDiskEntries < files.size(); ++q) unlink((dir + "/" + files[q].second).c_str());
}

void write_file(const std::string& dir, const std::string& name, const std::vector<char>& code) {
    if (dir.empty() || !mkdirs(dir)) return;  // the disk cache is an optimisation only
    const std::string tmp = dir + "/" + name + ".tmp" + std::to_string((long)getpid());
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return;
    const bool ok = fwrite(code.data(), 1, code.size(), f) == code.size();
    fclose(f);
    if (ok) rename(tmp.c_str(), (dir + "/" + name).c_str());
    else unlink(tmp.c_str());
    if (ok) prune_dir(dir);
}

int compile(const std::string& src, std::vector<char>& code) {
    hiprtcProgram prog;
    OPF_CHECK(hiprtcCreateProgram(&prog, src.c_str(), "opf_sweep_jit.hip", 0, nullptr, nullptr) == HIPRTC_SUCCESS,
              OPF_ERR_HIP, "jit: hiprtcCreateProgram failed");
    const hiprtcResult r = hiprtcCompileProgram(prog, kNOpts, kOpts);
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        set_error("jit: compile failed (%s): %.400s", hiprtcGetErrorString(r), log.c_str());
        return OPF_ERR_HIP;
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    code.resize(n);
    const bool ok = n > 0 && hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
    hiprtcDestroyProgram(&prog);
    OPF_CHECK(ok, OPF_ERR_HIP, "jit: no code object");
    return OPF_OK;
}

// Loaded kernels, key: source, options, compiler version, device. At most
// kMaxModules stay loaded; past that the least recently used one is unloaded
// after the device has drained (a graph that captured it must not be replayed).
constexpr size_t kMaxModules = 512;
struct Loaded {
    hipFunction_t fn;
    hipModule_t mod;
    int dev;
    uint64_t used;
};
std::mutex g_mu;
std::unordered_map<uint64_t, Loaded> g_fns;
uint64_t g_tick = 0;
long long g_compiles = 0, g_loads = 0;

int evict_lru_module(int dev) {
    auto victim = g_fns.end();
    for (auto it = g_fns.begin(); it != g_fns.end(); ++it)
        if (it->second.dev == dev && (victim == g_fns.end() || it->second.used < victim->second.used)) victim = it;
    if (victim == g_fns.end()) return OPF_OK;
    OPF_HIP(hipDeviceSynchronize());
    OPF_HIP(hipModuleUnload(victim->second.mod));
    g_fns.erase(victim);
    return OPF_OK;
}

}  // namespace

void jit_stats(long long* compiles, long long* loaded, long long* resident) {
    std::lock_guard<std::mutex> lock(g_mu);
    *compiles = g_compiles;
    *loaded = g_loads;
    *resident = (long long)g_fns.size();
}

namespace {
int function_for_source(const std::string& src, hipFunction_t* fn, uint64_t* key_out = nullptr,
                        const char* kname = "opf_sweep_jit");
}  // namespace

namespace {
// the loaded kernel of a generated source: in-process cache, disk cache, then hipRTC
int function_for_source(const std::string& src, hipFunction_t* fn, uint64_t* key_out, const char* kname) {
    int dev = 0;
    OPF_HIP(hipGetDevice(&dev));
    uint64_t h = fnv1a64(src.data(), src.size());
    for (int q = 0; q < kNOpts; ++q) h = fnv1a64(kOpts[q], strlen(kOpts[q]), h);
    int ver[2] = {0, 0};  // a compiler update invalidates the disk cache
    if (hiprtcVersion(&ver[0], &ver[1]) == HIPRTC_SUCCESS) h = fnv1a64(ver, sizeof ver, h);
    const uint64_t key = fnv1a64(&dev, sizeof dev, h);
    if (key_out) *key_out = key;
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_fns.find(key);
    if (it != g_fns.end()) {
        it->second.used = ++g_tick;
        *fn = it->second.fn;
        return OPF_OK;
    }
    const std::string dir = cache_dir(), name = "sweep_" + hexkey(h) + ".co";
    std::vector<char> code;
    if (dir.empty() || !read_file(dir + "/" + name, code)) {
        OPF_TRY(compile(src, code));
        ++g_compiles;
        write_file(dir, name, code);
    }
    size_t on_dev = 0;
    for (auto& e : g_fns) on_dev += e.second.dev == dev;
    if (on_dev >= kMaxModules) OPF_TRY(evict_lru_module(dev));
    hipModule_t mod;
    OPF_HIP(hipModuleLoadData(&mod, code.data()));
    hipFunction_t f;
    if (hipModuleGetFunction(&f, mod, kname) != hipSuccess) {
        hipModuleUnload(mod);
        set_error("jit: no %s in the code object", kname);
        return OPF_ERR_HIP;
    }
    g_fns.emplace(key, Loaded{f, mod, dev, ++g_tick});
    ++g_loads;
    *fn = f;
    return OPF_OK;
}
}  // namespace

namespace {
// What a launch needs once its source has been generated and loaded: the module
// (by its key in g_fns) and the form's rows per thread. Keyed by everything the
// generated source depends on, so a repeated sweep skips generating and hashing it.
struct Plan {
    uint64_t mod;
    bool tile;
    int rows;
};
std::unordered_map<uint64_t, Plan> g_plans;

uint64_t plan_key(const LoweredProgram& P, int dtype, const SweepArgs& a, int dev, int form) {
    uint64_t h = 1469598103934665603ULL;
    for (const Insn& in : P.code) {
        const int w[4] = {in.op, in.a, in.b, in.c};
        h = fnv1a64(w, sizeof w, h);
        if (!(in.op == I_CONST && in.a)) h = fnv1a64(&in.imm, sizeof in.imm, h);
    }
    const int q[6] = {dtype, a.op, a.umask, a.n[2] == 1, a.n[2] % 4, dev * 4 + form};
    h = fnv1a64(q, sizeof q, h);
    for (const FieldArg& f : a.f) {  // the direct form bakes plane strides into its loads
        const long long ps = (long long)f.e0 * f.e1;
        h = fnv1a64(&ps, sizeof ps, h);
    }
    return h;
}

}  // namespace

int jit_launch_sweep(hipStream_t st, const LoweredProgram& P, int dtype, const SweepArgs& a, int form) {
    hipFunction_t fn;
    const bool plane = a.n[2] == 1;
    int dev = 0;
    OPF_HIP(hipGetDevice(&dev));
    // the tiled form needs enough workgroups to fill the device: tiles of 64 x 16 cells
    // and chunks of 32 .. 8 planes, at least 512 workgroups (else the direct form)
    const long long tiles = (long long)((a.n[0] + 63) / 64) * ((a.n[1] + 15) / 16);
    if (form == 0 && tiles * (plane ? 1 : (a.n[2] + 7) / 8) < 512) form = 1;
    const uint64_t pk = plan_key(P, dtype, a, dev, form);
    auto launch = [&](const Plan& pl) -> int {
        SweepArgs args = a;
        void* params[] = {&args};
        if (pl.tile) {
            const int TY = kTileNTY * pl.rows;
            const long long t2 = (long long)((a.n[0] + 63) / 64) * ((a.n[1] + TY - 1) / TY);
            int zc = 32;
            while (zc > 8 && t2 * ((a.n[2] + zc - 1) / zc) < 2048) zc /= 2;
            args.zc = plane ? 1 : zc;
            OPF_HIP(hipModuleLaunchKernel(fn, (unsigned)((a.n[0] + 63) / 64), (unsigned)((a.n[1] + TY - 1) / TY),
                                          (unsigned)(plane ? 1 : (a.n[2] + zc - 1) / zc), 64, kTileNTY, 1, 0, st,
                                          params, nullptr));
        } else {
            OPF_HIP(hipModuleLaunchKernel(fn, (unsigned)((a.n[0] + 63) / 64), (unsigned)((a.n[1] + 3) / 4),
                                          (unsigned)(a.n[2] / pl.rows), 64, 4, 1, 0, st, params, nullptr));
        }
        return OPF_OK;
    };
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = g_plans.find(pk);
        if (it != g_plans.end()) {
            auto m = g_fns.find(it->second.mod);
            if (m != g_fns.end()) {
                m->second.used = ++g_tick;
                fn = m->second.fn;
                const Plan pl = it->second;
                return launch(pl);
            }
            g_plans.erase(it);  // its module was unloaded
        }
        if (g_plans.size() > 4 * kMaxModules) g_plans.clear();
    }
    auto remember = [&](const Plan& pl) {
        std::lock_guard<std::mutex> lock(g_mu);
        g_plans[pk] = pl;
    };
    // tiled form: fields staged in LDS rings, 4 / 2 / 1 rows per thread by LDS need
    const int tsize = dtype == OPF_F64 ? 8 : 4;
    for (int RY = 4; RY >= 1 && form != 1; RY /= 2) {
        TileField tf[SWEEP_MAX_FIELDS];
        size_t lds = 0;
        if (!tile_layout(P, plane, RY, tf, &lds, tsize)) break;
        int nq = 0;
        for (auto& f : tf) nq += f.used ? f.NQ : 0;
        if (lds > (size_t)64 * 1024 || nq > 24 || P.code.size() * RY > (size_t)SWEEP_JIT_MAX_INSNS) continue;
        std::string src;
        OPF_CHECK(gen_source_tile(P, dtype, a.op, RY, plane, a.umask, tf, src) == OPF_OK, OPF_ERR_ARG,
                  "jit: malformed sweep program");
        uint64_t mk = 0;
        OPF_TRY(function_for_source(src, &fn, &mk));
        const Plan pl{mk, true, RY};
        remember(pl);
        return launch(pl);
    }
    // direct form: loads from global memory
    // each thread computes R cells along z (R = 4, 2 or 1, dividing the z extent):
    // 512^3 FTCS expression 126 -> 152 G cell-updates/s FP64, 209 -> 242 FP32 at R = 4; R = 8 is slower
    int R = 1;
    for (int c = 4; c > 1; c /= 2)
        if (a.n[2] % c == 0 && P.code.size() * c <= (size_t)SWEEP_JIT_MAX_INSNS) { R = c; break; }
    std::string src;
    OPF_CHECK(gen_source(P, dtype, a.op, R, a, src) == OPF_OK, OPF_ERR_ARG, "jit: malformed sweep program");
    uint64_t mk = 0;
    OPF_TRY(function_for_source(src, &fn, &mk));
    const Plan pl{mk, false, R};
    remember(pl);
    return launch(pl);
}


// ---------------------------------------------------------------------------
// StencilPads of an equation, compiled: the pad-mode program (the interpreter in
// eqn.hip carries every value as PAD_K coefficients + bias + a linear flag on an LDS
// stack) becomes straight-line code in which each value is its bias and only the
// coefficients of the slots it can reach, all in registers. Which slots a value
// holds and whether it is linear in the unknowns are known from the program, so the
// generated code does per row only the arithmetic the interpreter does on those
// entries, in the same order (same bits). Forms that are not linear in the unknowns
// (a product of two linear values, a division by one, WENO or a nonlinear function
// of one, a read off the 5/7-point stencil) are not compiled: the interpreter runs
// them and reports the error.
namespace {
struct PadVal {
    bool lin = false;
    unsigned mask = 0;  // slots with a coefficient
};

int gen_pad_source(const LoweredProgram& P, std::string& s) {
    s.clear();
    s += "typedef double T;\n";
    s += kArgsSrc;
    s += "\n";
    s += kPadArgsSrc;
    s += "\n";
    s += kMathSrc;
    s += "\nextern \"C\" __global__ __launch_bounds__(256) void opf_pad_jit(PadArgs a) {\n"
         "  const long long N = (long long)a.n[0] * a.n[1] * a.n[2];\n"
         "  const long long row = (long long)blockIdx.x * 256 + threadIdx.x;\n"
         "  if (row >= N) return;\n"
         "  const int ci = a.s[0] + (int)(row % a.n[0]);\n"
         "  const int cj = a.s[1] + (int)((row / a.n[0]) % a.n[1]);\n"
         "  const int ck = a.s[2] + (int)(row / ((long long)a.n[0] * a.n[1]));\n";
    bool fu[SWEEP_MAX_FIELDS] = {};
    for (const Insn& in : P.code)
        if (in.op == I_LD) {
            if (in.a < 0 || in.a >= SWEEP_MAX_FIELDS) return OPF_ERR_ARG;
            fu[in.a] = true;
        }
    char line[512];
    for (int f = 0; f < SWEEP_MAX_FIELDS; ++f) {
        if (!fu[f]) continue;
        snprintf(line, sizeof line,
                 "  const double* __restrict__ F%d = (const double*)a.f[%d].ptr + ((ci - a.f[%d].off0) + "
                 "(long long)a.f[%d].e0 * ((cj - a.f[%d].off1) + (long long)a.f[%d].e1 * (ck - a.f[%d].off2)));\n",
                 f, f, f, f, f, f, f);
        s += line;
    }
    std::vector<PadVal> val(P.code.size());
    std::vector<int> st;
    auto pop = [&](int& v) -> bool {
        if (st.empty()) return false;
        v = st.back();
        st.pop_back();
        return true;
    };
    auto k = [](int id, int q) { return "v" + std::to_string(id) + "k" + std::to_string(q); };
    auto c = [](int id) { return "v" + std::to_string(id) + "c"; };
    auto emit = [&](const std::string& t) { s += "  " + t + ";\n"; };
    for (size_t pc = 0; pc < P.code.size(); ++pc) {
        const Insn& in = P.code[pc];
        const int id = (int)pc;
        PadVal& V = val[pc];
        switch (in.op) {
            case I_LD:
                snprintf(line, sizeof line, "double %s = F%d[%d]", c(id).c_str(), in.a, in.b);
                emit(line);
                break;
            case I_CONST: {
                if (in.a) {
                    snprintf(line, sizeof line, "double %s = a.cst[%d]", c(id).c_str(), in.b);
                } else {
                    long long bits;
                    std::memcpy(&bits, &in.imm, 8);
                    snprintf(line, sizeof line, "double %s = __longlong_as_double(%lldLL)", c(id).c_str(), bits);
                }
                emit(line);
            } break;
            case I_DX:
            case I_X: {
                const char* tb = in.op == I_DX ? "dx" : "x";
                const char* cc = in.a == 0 ? "ci" : in.a == 1 ? "cj" : "ck";
                if (in.a < 0 || in.a > 2) return OPF_ERR_ARG;
                snprintf(line, sizeof line, "double %s = a.%s[%d][%s + %d - a.morg[%d]]", c(id).c_str(), tb, in.a, cc,
                         in.b, in.a);
                emit(line);
            } break;
            case I_LDU: {
                const int u = in.a, ox = (in.c & 31) - 8, oy = ((in.c >> 5) & 31) - 8, oz = ((in.c >> 10) & 31) - 8;
                int sl = -1;
                if (ox == 0 && oy == 0 && oz == 0) sl = 0;
                else if (oy == 0 && oz == 0) sl = ox == -1 ? 1 : ox == 1 ? 2 : -1;
                else if (ox == 0 && oz == 0) sl = oy == -1 ? 3 : oy == 1 ? 4 : -1;
                else if (ox == 0 && oy == 0) sl = oz == -1 ? 5 : oz == 1 ? 6 : -1;
                if (sl < 0 || u < 0 || u >= EQS_MAX) return OPF_ERR_UNSUPPORTED;
                V.lin = true;
                V.mask = 1u << sl;
                snprintf(line, sizeof line, "double %s = (a.live < 0 || a.live == %d) ? 1.0 : 0.0", k(id, sl).c_str(),
                         u);
                emit(line);
                snprintf(line, sizeof line, "double %s = 0.0", c(id).c_str());
                emit(line);
                const int off[3] = {ox, oy, oz};
                const char* cc[3] = {"ci", "cj", "ck"};
                for (int d = 0; d < 3; ++d) {  // the boundary constant of the last dim outside, as the interpreter
                    snprintf(line, sizeof line,
                             "{ const int p = %s + %d; if (!(p >= a.tl[%d] && p < a.te[%d]) && !a.bc[%d].per[%d]) "
                             "%s = a.bc[%d].beta[%d][p < a.tl[%d] ? 0 : 1]; }",
                             cc[d], off[d], d, d, u, d, c(id).c_str(), u, d, d);
                    emit(line);
                }
            } break;
            case I_ADD:
            case I_SUB: {
                int x, y;
                if (!pop(y) || !pop(x)) return OPF_ERR_ARG;
                const char o = in.op == I_ADD ? '+' : '-';
                V.lin = val[x].lin || val[y].lin;
                V.mask = val[x].mask | val[y].mask;
                for (int q = 0; q < PAD_K; ++q) {
                    if (!(V.mask >> q & 1)) continue;
                    const std::string l = (val[x].mask >> q & 1) ? k(x, q) : "0.0";
                    const std::string r = (val[y].mask >> q & 1) ? k(y, q) : "0.0";
                    emit("double " + k(id, q) + " = " + l + " " + o + " " + r);
                }
                emit("double " + c(id) + " = " + c(x) + " " + o + " " + c(y));
            } break;
            case I_MUL: {
                int x, y;
                if (!pop(y) || !pop(x)) return OPF_ERR_ARG;
                if (val[x].lin && val[y].lin) return OPF_ERR_UNSUPPORTED;
                if (val[y].lin) {  // const * pad: every entry of y times x's constant
                    V = val[y];
                    for (int q = 0; q < PAD_K; ++q)
                        if (V.mask >> q & 1) emit("double " + k(id, q) + " = " + c(x) + " * " + k(y, q));
                    emit("double " + c(id) + " = " + c(x) + " * " + c(y));
                } else {  // pad * const (or const * const)
                    V = val[x];
                    for (int q = 0; q < PAD_K; ++q)
                        if (V.mask >> q & 1) emit("double " + k(id, q) + " = " + k(x, q) + " * " + c(y));
                    emit("double " + c(id) + " = " + c(x) + " * " + c(y));
                }
            } break;
            case I_DIV: {
                int x, y;
                if (!pop(y) || !pop(x)) return OPF_ERR_ARG;
                if (val[y].lin) return OPF_ERR_UNSUPPORTED;
                V = val[x];
                for (int q = 0; q < PAD_K; ++q)
                    if (V.mask >> q & 1) emit("double " + k(id, q) + " = " + k(x, q) + " / " + c(y));
                emit("double " + c(id) + " = " + c(x) + " / " + c(y));
            } break;
            case I_NEG: {
                int x;
                if (!pop(x)) return OPF_ERR_ARG;
                V = val[x];
                for (int q = 0; q < PAD_K; ++q)
                    if (V.mask >> q & 1) emit("double " + k(id, q) + " = -" + k(x, q));
                emit("double " + c(id) + " = -" + c(x));
            } break;
            case I_ABS:
            case I_FN1: {
                int x;
                if (!pop(x)) return OPF_ERR_ARG;
                if (val[x].lin) return OPF_ERR_UNSUPPORTED;
                if (in.op == I_ABS) emit("double " + c(id) + " = fabs(" + c(x) + ")");
                else emit("double " + c(id) + " = opf_fn1<double>(" + std::to_string(in.a) + ", " + c(x) + ")");
            } break;
            case I_MIN:
            case I_MAX:
            case I_GT:
            case I_LT:
            case I_GE:
            case I_LE:
            case I_FN2: {
                int x, y;
                if (!pop(y) || !pop(x)) return OPF_ERR_ARG;
                if (val[x].lin || val[y].lin) return OPF_ERR_UNSUPPORTED;
                const std::string X = c(x), Y = c(y);
                std::string e;
                switch (in.op) {
                    case I_MIN: e = Y + " < " + X + " ? " + Y + " : " + X; break;
                    case I_MAX: e = X + " < " + Y + " ? " + Y + " : " + X; break;
                    case I_GT: e = X + " > " + Y + " ? 1.0 : 0.0"; break;
                    case I_LT: e = X + " < " + Y + " ? 1.0 : 0.0"; break;
                    case I_GE: e = X + " >= " + Y + " ? 1.0 : 0.0"; break;
                    case I_LE: e = X + " <= " + Y + " ? 1.0 : 0.0"; break;
                    default: e = "opf_fn2<double>(" + std::to_string(in.a) + ", " + X + ", " + Y + ")";
                }
                emit("double " + c(id) + " = " + e);
            } break;
            case I_SEL: {
                int cnd, t_, f_;
                if (!pop(f_) || !pop(t_) || !pop(cnd)) return OPF_ERR_ARG;
                if (val[cnd].lin || val[t_].lin || val[f_].lin) return OPF_ERR_UNSUPPORTED;
                emit("double " + c(id) + " = " + c(cnd) + " != 0.0 ? " + c(t_) + " : " + c(f_));
            } break;
            default:  // WENO on pads and anything else: the interpreter
                return OPF_ERR_UNSUPPORTED;
        }
        st.push_back(id);
    }
    if (st.size() != 1) return OPF_ERR_ARG;
    const int r = st.back();
    for (int q = 0; q < PAD_K; ++q) {
        snprintf(line, sizeof line, "  a.coef[(long long)%d * N + row] = %s;\n", q,
                 (val[r].mask >> q & 1) ? k(r, q).c_str() : "0.0");
        s += line;
    }
    s += "  a.bias[row] = " + c(r) + ";\n}\n";
    return OPF_OK;
}

std::unordered_map<uint64_t, uint64_t> g_pad_plans;  // program structure -> module key (0: not compiled)
}  // namespace

int jit_launch_pad(hipStream_t st, const LoweredProgram& P, const PadArgs& a) {
    int dev = 0;
    OPF_HIP(hipGetDevice(&dev));
    uint64_t pk = 1469598103934665603ULL;
    for (const Insn& in : P.code) {
        const int w[4] = {in.op, in.a, in.b, in.c};
        pk = fnv1a64(w, sizeof w, pk);
        if (!(in.op == I_CONST && in.a)) pk = fnv1a64(&in.imm, sizeof in.imm, pk);
    }
    pk = fnv1a64(&dev, sizeof dev, pk);
    hipFunction_t fn = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = g_pad_plans.find(pk);
        if (it != g_pad_plans.end()) {
            if (it->second == 0) return OPF_ERR_UNSUPPORTED;  // known not to compile
            auto m = g_fns.find(it->second);
            if (m != g_fns.end()) {
                m->second.used = ++g_tick;
                fn = m->second.fn;
            }
        }
        if (g_pad_plans.size() > 4 * kMaxModules) g_pad_plans.clear();
    }
    if (!fn) {
        std::string src;
        const int rc = gen_pad_source(P, src);
        if (rc != OPF_OK) {
            std::lock_guard<std::mutex> lock(g_mu);
            g_pad_plans[pk] = 0;
            return OPF_ERR_UNSUPPORTED;
        }
        uint64_t mk = 0;
        OPF_TRY(function_for_source(src, &fn, &mk, "opf_pad_jit"));
        std::lock_guard<std::mutex> lock(g_mu);
        g_pad_plans[pk] = mk;
    }
    const long long N = (long long)a.n[0] * a.n[1] * a.n[2];
    PadArgs args = a;
    void* params[] = {&args};
    OPF_HIP(hipModuleLaunchKernel(fn, (unsigned)((N + 255) / 256), 1, 1, 256, 1, 1, 0, st, params, nullptr));
    return OPF_OK;
}

}  // namespace opf
