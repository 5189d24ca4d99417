// Internal declarations shared by the translation units of libopflow_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/opflow_hip.h"
#include "ghost_index.hpp"

namespace opf {

void set_error(const char* fmt, ...);

#define OPF_HIP(call)                                                                               \
    do {                                                                                            \
        hipError_t _e = (call);                                                                     \
        if (_e != hipSuccess) {                                                                     \
            opf::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call, hipGetErrorString(_e));     \
            return OPF_ERR_HIP;                                                                     \
        }                                                                                           \
    } while (0)

#define OPF_CHECK(cond, code, ...)                                                                  \
    do {                                                                                            \
        if (!(cond)) {                                                                              \
            opf::set_error(__VA_ARGS__);                                                            \
            return (code);                                                                          \
        }                                                                                           \
    } while (0)

#define OPF_TRY(expr)                                                                               \
    do {                                                                                            \
        int _rc = (expr);                                                                           \
        if (_rc != OPF_OK) return _rc;                                                              \
    } while (0)

inline size_t dtype_size(int dt) { return dt == OPF_F32 ? 4 : 8; }

}  // namespace opf

struct opf_ctx {
    int device = 0;
    int cus = 256;  // compute units of the device (launch shapes that want one workgroup per CU)
    hipStream_t stream = nullptr;      // where the context's work goes
    hipStream_t own_stream = nullptr;  // the stream the context created (destroyed with it)
    // scratch for deterministic reductions (partials + result)
    double* red_buf = nullptr;
    size_t red_cap = 0;
    double* red_host = nullptr;  // pinned
    // cache of lowered device programs keyed by content (the bytes a sweep uploads
    // carry no constant values). Sweep programs are evictable: at most kMaxPrograms
    // of them stay resident, least recently used freed first. Programs that an
    // equation handler keeps using are not evictable and live as long as the context.
    struct ProgEntry {
        uint64_t hash;
        void* dev;
        size_t bytes;
        uint64_t used;
        bool evictable;
        std::vector<unsigned char> host;
    };
    static constexpr size_t kMaxPrograms = 1024;
    std::vector<ProgEntry> progs;
    uint64_t prog_tick = 0;
    long long prog_uploads = 0;
    // this device's copy of the allocation-free FTCS scratch (core.hip), looked up at creation
    void* noalloc_scratch = nullptr;
    // expression sweeps: 0 = interpreter, 1 = compiled kernel for >= 2^18 cells, 2 = compiled kernel always
    int jit_mode = 1;
};

struct opf_mesh {
    opf_ctx* ctx = nullptr;
    int dim = 0;
    int start[3] = {0, 0, 0};
    int nnodes[3] = {1, 1, 1};
    int pad = 0;
    std::vector<double> hx[3], hdx[3];  // host copies, index k <-> node start-pad+k
    double* dx_dev[3] = {nullptr, nullptr, nullptr};
    double* x_dev[3] = {nullptr, nullptr, nullptr};
    float* dx_f32[3] = {nullptr, nullptr, nullptr};  // lazily made for FP32 sweeps
    float* x_f32[3] = {nullptr, nullptr, nullptr};
    bool uniform[3] = {false, false, false};
    double h[3] = {0, 0, 0};
    double x(int d, int i) const { return hx[d][i - (start[d] - pad)]; }
    double dx(int d, int i) const { return hdx[d][i - (start[d] - pad)]; }
    bool has_node(int d, int i) const {
        return i - (start[d] - pad) >= 0 && i - (start[d] - pad) < (int)hx[d].size();
    }
    bool has_dx(int d, int i) const {
        return i - (start[d] - pad) >= 0 && i - (start[d] - pad) < (int)hdx[d].size();
    }
};

struct opf_field {
    opf_ctx* ctx = nullptr;
    opf_mesh* mesh = nullptr;
    opf_field_desc d{};
    void* data = nullptr;   // device storage
    void* spare = nullptr;  // second buffer for in-place updates (lazy)
    bool spare_valid = false;
    size_t nbytes = 0;
    size_t count = 0;
};

namespace opf {

int build_ghost_info(const opf_field* f, GhostInfo* gi);
int launch_ghost_fill(const opf_field* f, void* data);

// seeded Gaussian bumps on a unit-cube cell-centred grid (host)
void gen_bumps_grid(int dim, const int n[3], uint64_t seed, int nbumps, double noise, double* out);
// the same field generated on the device into `out` (device array, N doubles), on stream st
int gen_bumps_grid_dev(hipStream_t st, int dim, const int n[3], uint64_t seed, int nbumps, double noise, double* out);

// device program cache
// evictable: the caller uses *dev only for launches enqueued right away (a sweep),
// so the cache may free it later, after a drain of the context's stream
int upload_program(opf_ctx* ctx, const void* bytes, size_t n, void** dev, bool evictable = false);

}  // namespace opf
