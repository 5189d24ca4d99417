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
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

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

namespace {

#define OPF_SHARED_SRC(...) #__VA_ARGS__
const char* const kArgsSrc =
#include "sweep_args.inc"
    ;
const char* const kWenoSrc =
#include "weno_device.inc"
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

// the straight-line kernel of program P in precision T, writing out (op) value
int gen_source(const LoweredProgram& P, int dtype, int op, int R, const SweepArgs& a, std::string& s) {
    s.clear();
    s += dtype == OPF_F64 ? "typedef double T;\n" : "typedef float T;\n";
    s += kArgsSrc;
    s += "\n";
    s += kWenoSrc;
    {
        char hdr[512];
        snprintf(hdr, sizeof hdr,
                 "\nextern \"C\" __global__ __launch_bounds__(256) void opf_sweep_jit(SweepArgs a) {\n"
                 "  const int i = a.s[0] + (int)blockIdx.x * 64 + (int)threadIdx.x;\n"
                 "  const int j = a.s[1] + (int)blockIdx.y * 4 + (int)threadIdx.y;\n"
                 "  if (i >= a.s[0] + a.n[0] || j >= a.s[1] + a.n[1]) return;\n"
                 "  const int k = a.s[2] + (int)blockIdx.z * %d;\n",
                 R);
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
                    long long bits;
                    std::memcpy(&bits, &in.imm, 8);
                    snprintf(line, sizeof line, "  const T v%d = (T)__longlong_as_double(%lldLL);\n", id, bits);
                } break;
                case I_DX:
                    snprintf(line, sizeof line, "  const T v%d = DX%d[%d];\n", id, in.a, in.b + (in.a == 2 ? r : 0));
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
                    snprintf(line, sizeof line, "  const T v%d = v%d %c v%d;\n", id, x, o, y);
                } break;
                case I_NEG: {
                    int x;
                    if (!pop(x)) return OPF_ERR_ARG;
                    snprintf(line, sizeof line, "  const T v%d = -v%d;\n", id, x);
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

void write_file(const std::string& dir, const std::string& name, const std::vector<char>& code) {
    if (dir.empty() || !mkdirs(dir)) return;  // the disk cache is an optimisation only
    const std::string tmp = dir + "/" + name + ".tmp" + std::to_string((long)getpid());
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return;
    const bool ok = fwrite(code.data(), 1, code.size(), f) == code.size();
    fclose(f);
    if (ok) rename(tmp.c_str(), (dir + "/" + name).c_str());
    else unlink(tmp.c_str());
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

std::mutex g_mu;
std::unordered_map<uint64_t, hipFunction_t> g_fns;  // key: source, options, device

}  // namespace

int jit_sweep_function(const LoweredProgram& P, int dtype, int op, int R, const SweepArgs& a, hipFunction_t* fn) {
    std::string src;
    OPF_CHECK(gen_source(P, dtype, op, R, a, src) == OPF_OK, OPF_ERR_ARG, "jit: malformed sweep program");
    int dev = 0;
    OPF_HIP(hipGetDevice(&dev));
    uint64_t h = fnv1a64(src.data(), src.size());
    for (int q = 0; q < kNOpts; ++q) h = fnv1a64(kOpts[q], strlen(kOpts[q]), h);
    int ver[2] = {0, 0};  // a compiler update invalidates the disk cache
    if (hiprtcVersion(&ver[0], &ver[1]) == HIPRTC_SUCCESS) h = fnv1a64(ver, sizeof ver, h);
    const uint64_t key = fnv1a64(&dev, sizeof dev, h);
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_fns.find(key);
    if (it != g_fns.end()) {
        *fn = it->second;
        return OPF_OK;
    }
    const std::string dir = cache_dir(), name = "sweep_" + hexkey(h) + ".co";
    std::vector<char> code;
    if (dir.empty() || !read_file(dir + "/" + name, code)) {
        OPF_TRY(compile(src, code));
        write_file(dir, name, code);
    }
    hipModule_t mod;
    OPF_HIP(hipModuleLoadData(&mod, code.data()));
    hipFunction_t f;
    OPF_HIP(hipModuleGetFunction(&f, mod, "opf_sweep_jit"));
    g_fns.emplace(key, f);
    *fn = f;
    return OPF_OK;
}

int jit_launch_sweep(hipStream_t st, const LoweredProgram& P, int dtype, const SweepArgs& a) {
    hipFunction_t fn;
    // each thread computes R cells along z (R = 4, 2 or 1, dividing the z extent):
    // 512^3 FTCS expression 126 -> 152 G cell-updates/s FP64, 209 -> 242 FP32 at R = 4; R = 8 is slower
    int R = 1;
    for (int c = 4; c > 1; c /= 2)
        if (a.n[2] % c == 0) { R = c; break; }
    OPF_TRY(jit_sweep_function(P, dtype, a.op, R, a, &fn));
    SweepArgs args = a;
    void* params[] = {&args};
    OPF_HIP(hipModuleLaunchKernel(fn, (unsigned)((a.n[0] + 63) / 64), (unsigned)((a.n[1] + 3) / 4), (unsigned)(a.n[2] / R),
                                  64, 4, 1, 0, st, params, nullptr));
    return OPF_OK;
}

}  // namespace opf
