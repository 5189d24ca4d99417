// Context, mesh, field storage, ghost filling (updatePadding), constant
// assignment, deterministic reductions and synthetic-input generators.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "opf_internal.hpp"
#include "ghost_index.hpp"

namespace opf {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

}  // namespace opf

using namespace opf;

extern "C" const char* opf_last_error(void) { return g_err; }
extern "C" const char* opf_version(void) { return "opflow-hip 0.1 (gfx950)"; }

// ---------------------------------------------------------------- context
namespace {
// Device scratch of the allocation-free entry point opf_ftcs_noalloc: the
// second time-level buffer of the FTCS ping-pong. Every device the library runs
// on holds its own copy until the process ends. Host code never passes
// `scratch` itself to a kernel or a copy (on the host the name is only a
// stand-in); it uses the per-device address looked up once below.
__device__ __align__(256) unsigned char scratch[OPF_FTCS_NOALLOC_MAX_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Once per device, with that device current: the address of its copy.
int scratch_setup(int device, void** out) {
    OPF_CHECK(device >= 0 && device < kMaxDevices, OPF_ERR_ARG, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        OPF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        OPF_CHECK(p != nullptr, OPF_ERR_HIP, "hipGetSymbolAddress(scratch) gave a null address");
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return OPF_OK;
}
}  // namespace

extern "C" int opf_ctx_create(int device, opf_ctx** out) {
    OPF_CHECK(out, OPF_ERR_ARG, "opf_ctx_create: null out");
    int n = 0;
    OPF_HIP(hipGetDeviceCount(&n));
    OPF_CHECK(device >= 0 && device < n, OPF_ERR_ARG, "opf_ctx_create: device %d of %d", device, n);
    OPF_HIP(hipSetDevice(device));
    auto* c = new opf_ctx();
    c->device = device;
    // a blocking stream: measured on MI355X, kernels on a hipStreamNonBlocking
    // stream ran ~12% slower than the same kernels on a default-flag stream
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
        delete c;
        set_error("hipStreamCreate: %s", hipGetErrorString(e));
        return OPF_ERR_HIP;
    }
    c->own_stream = c->stream;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        c->cus = cus;
    if (int rc = scratch_setup(device, &c->noalloc_scratch); rc != OPF_OK) {
        hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    c->red_cap = 4096;
    OPF_HIP(hipMalloc(&c->red_buf, c->red_cap * sizeof(double)));
    OPF_HIP(hipHostMalloc(&c->red_host, 16 * sizeof(double), hipHostMallocDefault));
    *out = c;
    return OPF_OK;
}

extern "C" int opf_ctx_destroy(opf_ctx* c) {
    if (!c) return OPF_OK;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (auto& p : c->progs) hipFree(p.dev);
    if (c->red_buf) hipFree(c->red_buf);
    if (c->red_host) hipHostFree(c->red_host);
    if (c->stream != c->own_stream) hipStreamSynchronize(c->own_stream);
    hipStreamDestroy(c->own_stream);
    delete c;
    return OPF_OK;
}

extern "C" int opf_ctx_set_stream(opf_ctx* c, void* stream) {
    OPF_CHECK(c, OPF_ERR_ARG, "null ctx");
    hipStream_t next = stream ? (hipStream_t)stream : c->own_stream;
    if (next == c->stream) return OPF_OK;
    // keep the order: work already enqueued on the previous stream precedes what follows
    hipEvent_t ev;
    OPF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    OPF_HIP(hipEventRecord(ev, c->stream));
    OPF_HIP(hipStreamWaitEvent(next, ev, 0));
    OPF_HIP(hipEventDestroy(ev));
    c->stream = next;
    return OPF_OK;
}

extern "C" int opf_ctx_sync(opf_ctx* c) {
    OPF_CHECK(c, OPF_ERR_ARG, "null ctx");
    OPF_HIP(hipStreamSynchronize(c->stream));
    return OPF_OK;
}

extern "C" void* opf_ctx_stream(opf_ctx* c) { return c ? (void*)c->stream : nullptr; }

extern "C" int opf_ctx_set_jit(opf_ctx* c, int mode) {
    OPF_CHECK(c && mode >= 0 && mode <= 4, OPF_ERR_ARG, "opf_ctx_set_jit: null ctx or mode outside 0..4");
    c->jit_mode = mode;
    return OPF_OK;
}

namespace opf {
static uint64_t fnv1a(const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h ^ (uint64_t)n;
}

int upload_program(opf_ctx* ctx, const void* bytes, size_t n, void** dev, bool evictable) {
    uint64_t h = fnv1a(bytes, n);
    for (auto& p : ctx->progs)
        if (p.hash == h && p.bytes == n && p.evictable == evictable && std::memcmp(p.host.data(), bytes, n) == 0) {
            p.used = ++ctx->prog_tick;
            *dev = p.dev;
            return OPF_OK;
        }
    if (evictable) {
        size_t ne = 0;
        for (auto& p : ctx->progs) ne += p.evictable;
        if (ne >= opf_ctx::kMaxPrograms) {
            // least recently used first, after the context's queued work has drained
            auto victim = ctx->progs.end();
            for (auto it = ctx->progs.begin(); it != ctx->progs.end(); ++it)
                if (it->evictable && (victim == ctx->progs.end() || it->used < victim->used)) victim = it;
            OPF_HIP(hipStreamSynchronize(ctx->stream));
            OPF_HIP(hipFree(victim->dev));
            ctx->progs.erase(victim);
        }
    }
    void* d = nullptr;
    OPF_HIP(hipMalloc(&d, n));
    OPF_HIP(hipMemcpyAsync(d, bytes, n, hipMemcpyHostToDevice, ctx->stream));
    OPF_HIP(hipStreamSynchronize(ctx->stream));
    const unsigned char* b = (const unsigned char*)bytes;
    ctx->progs.push_back({h, d, n, ++ctx->prog_tick, evictable, std::vector<unsigned char>(b, b + n)});
    ++ctx->prog_uploads;
    *dev = d;
    return OPF_OK;
}
}  // namespace opf

// ---------------------------------------------------------------- raw buffers
extern "C" int opf_buf_create(opf_ctx* c, size_t nbytes, void** dptr) {
    OPF_CHECK(c && dptr, OPF_ERR_ARG, "opf_buf_create: null arg");
    hipError_t e = hipMalloc(dptr, nbytes ? nbytes : 1);
    OPF_CHECK(e == hipSuccess, OPF_ERR_NOMEM, "opf_buf_create: %s", hipGetErrorString(e));
    OPF_HIP(hipMemsetAsync(*dptr, 0, nbytes, c->stream));
    return OPF_OK;
}
extern "C" int opf_buf_destroy(opf_ctx* c, void* dptr) {
    OPF_CHECK(c, OPF_ERR_ARG, "null ctx");
    OPF_HIP(hipStreamSynchronize(c->stream));
    if (dptr) OPF_HIP(hipFree(dptr));
    return OPF_OK;
}
extern "C" int opf_buf_upload(opf_ctx* c, void* dptr, const void* host, size_t nbytes) {
    OPF_CHECK(c && dptr && host, OPF_ERR_ARG, "opf_buf_upload: null arg");
    OPF_HIP(hipMemcpyAsync(dptr, host, nbytes, hipMemcpyHostToDevice, c->stream));
    OPF_HIP(hipStreamSynchronize(c->stream));
    return OPF_OK;
}
extern "C" int opf_buf_download(opf_ctx* c, const void* dptr, void* host, size_t nbytes) {
    OPF_CHECK(c && dptr && host, OPF_ERR_ARG, "opf_buf_download: null arg");
    OPF_HIP(hipMemcpyAsync(host, dptr, nbytes, hipMemcpyDeviceToHost, c->stream));
    OPF_HIP(hipStreamSynchronize(c->stream));
    return OPF_OK;
}

extern "C" int opf_buf_zero(opf_ctx* c, void* dptr, size_t nbytes) {
    OPF_CHECK(c && dptr, OPF_ERR_ARG, "opf_buf_zero: null arg");
    OPF_HIP(hipMemsetAsync(dptr, 0, nbytes, c->stream));
    return OPF_OK;
}

// ---------------------------------------------------------------- mesh
extern "C" int opf_mesh_uniform_arrays(int nnodes, double lo, double hi, int pad, double* x, double* dx) {
    // MeshBuilder::set1DMesh(min, max, k) + setExtMesh(Symm) with range start 0
    // (CartesianMesh.hpp:296-303, 235-248, 279-287).
    OPF_CHECK(nnodes >= 2 && pad >= 0 && x && dx, OPF_ERR_ARG, "opf_mesh_uniform_arrays: bad args");
    const int ntot = nnodes + 2 * pad;
    for (int i = 0; i < nnodes; ++i) x[i + pad] = (hi - lo) / (nnodes - 1) * i + lo;
    for (int j = 0; j < nnodes - 1; ++j) dx[j + pad] = (hi - lo) / (nnodes - 1);
    // Symm extension of dx: ghost spacing mirrors the interior spacing
    for (int i = -pad; i < 0; ++i) dx[i + pad] = dx[2 * 0 - 1 - i + pad];
    for (int i = nnodes - 1; i < nnodes + pad - 1; ++i) dx[i + pad] = dx[2 * nnodes - 3 - i + pad];
    for (int i = -1; i >= -pad; --i) x[i + pad] = x[i + 1 + pad] - dx[i + pad];
    for (int i = nnodes; i < nnodes + pad; ++i) x[i + pad] = x[i - 1 + pad] + dx[i - 1 + pad];
    (void)ntot;
    return OPF_OK;
}

extern "C" int opf_mesh_create(opf_ctx* ctx, int dim, const int start[3], const int nnodes[3], int pad,
                               const double* const x[3], const double* const dx[3], opf_mesh** out) {
    OPF_CHECK(ctx && out && dim >= 1 && dim <= 3 && pad >= 0, OPF_ERR_ARG, "opf_mesh_create: bad args");
    auto* m = new opf_mesh();
    m->ctx = ctx;
    m->dim = dim;
    m->pad = pad;
    for (int d = 0; d < 3; ++d) {
        if (d >= dim) {
            m->start[d] = 0;
            m->nnodes[d] = 1;
            continue;
        }
        m->start[d] = start ? start[d] : 0;
        m->nnodes[d] = nnodes[d];
        if (nnodes[d] < 2 || !x[d] || !dx[d]) {
            delete m;
            set_error("opf_mesh_create: dim %d needs >= 2 nodes and arrays", d);
            return OPF_ERR_ARG;
        }
        int nx = nnodes[d] + 2 * pad;
        m->hx[d].assign(x[d], x[d] + nx);
        m->hdx[d].assign(dx[d], dx[d] + nx - 1);
        bool uni = true;
        for (int i = 1; i < nx - 1; ++i) uni &= (m->hdx[d][i] == m->hdx[d][0]);
        m->uniform[d] = uni;
        m->h[d] = m->hdx[d][0];
        OPF_HIP(hipMalloc(&m->x_dev[d], nx * sizeof(double)));
        OPF_HIP(hipMalloc(&m->dx_dev[d], (nx - 1) * sizeof(double)));
        OPF_HIP(hipMemcpy(m->x_dev[d], m->hx[d].data(), nx * sizeof(double), hipMemcpyHostToDevice));
        OPF_HIP(hipMemcpy(m->dx_dev[d], m->hdx[d].data(), (nx - 1) * sizeof(double), hipMemcpyHostToDevice));
    }
    *out = m;
    return OPF_OK;
}

extern "C" int opf_mesh_destroy(opf_mesh* m) {
    if (!m) return OPF_OK;
    for (int d = 0; d < 3; ++d) {
        if (m->x_dev[d]) hipFree(m->x_dev[d]);
        if (m->dx_dev[d]) hipFree(m->dx_dev[d]);
        if (m->x_f32[d]) hipFree(m->x_f32[d]);
        if (m->dx_f32[d]) hipFree(m->dx_f32[d]);
    }
    delete m;
    return OPF_OK;
}

// ---------------------------------------------------------------- field descriptor
extern "C" int opf_field_desc_finalize(opf_field_desc* f, const int mesh_start[3], const int mesh_nnodes[3]) {
    // ExprBuilder<CartesianField>::calculateRanges / validateRanges / build
    // (CartesianField.hpp:929-1029), single-process (no split strategy).
    OPF_CHECK(f && f->dim >= 1 && f->dim <= 3, OPF_ERR_ARG, "opf_field_desc_finalize: bad dim");
    OPF_CHECK(f->dtype == OPF_F64 || f->dtype == OPF_F32, OPF_ERR_ARG, "bad dtype");
    for (int d = 0; d < f->dim; ++d)
        for (int s = 0; s < 2; ++s) {
            f->padding = std::max(f->padding, f->ext[d][s]);
            int t = f->bc_type[d][s];
            OPF_CHECK(t == OPF_BC_UNDEFINED || t == OPF_BC_DIRC || t == OPF_BC_NEUM || t == OPF_BC_PERIODIC ||
                              t == OPF_BC_SYMM || t == OPF_BC_ASYMM,
                      OPF_ERR_UNSUPPORTED, "bc type %d not supported (dim %d side %d)", t, d, s);
        }
    for (int d = 0; d < 3; ++d) {
        if (d >= f->dim) {
            f->mesh_start[d] = 0;
            f->mesh_end[d] = 1;
            f->local_start[d] = f->assign_start[d] = f->access_start[d] = f->logical_start[d] = 0;
            f->local_end[d] = f->assign_end[d] = f->access_end[d] = f->logical_end[d] = 1;
            f->offset[d] = 0;
            f->extent[d] = 1;
            continue;
        }
        int ms = mesh_start ? mesh_start[d] : 0;
        int me = ms + mesh_nnodes[d];
        f->mesh_start[d] = ms;
        f->mesh_end[d] = me;
        int as = ms, ae = me, cs = ms, ce = me;  // assignable, accessible
        bool corner = f->loc[d] == OPF_LOC_CORNER;
        switch (f->bc_type[d][0]) {
            case OPF_BC_DIRC:
                if (corner) as++;
                break;
            default:
                break;
        }
        switch (f->bc_type[d][1]) {
            case OPF_BC_DIRC:
                if (corner) {
                    ae--;
                } else {
                    ce--;
                    ae--;
                }
                break;
            case OPF_BC_NEUM:
            case OPF_BC_UNDEFINED:
            case OPF_BC_SYMM:
            case OPF_BC_ASYMM:
                if (!corner) {
                    ce--;
                    ae--;
                }
                break;
            case OPF_BC_PERIODIC:
                ce--;
                ae--;
                break;
        }
        int gs = cs - f->ext[d][0], ge = ce + f->ext[d][1];
        // validateRanges
        cs = std::max(cs, gs);
        ce = std::min(ce, ge);
        int ls = std::max(cs, gs), le = std::min(ce, ge);
        as = std::max(as, cs);
        ae = std::min(ae, ce);
        f->access_start[d] = cs;
        f->access_end[d] = ce;
        f->assign_start[d] = as;
        f->assign_end[d] = ae;
        f->logical_start[d] = gs;
        f->logical_end[d] = ge;
        f->local_start[d] = ls;
        f->local_end[d] = le;
        f->offset[d] = ls - f->padding;
        f->extent[d] = le - ls + 2 * f->padding;
        OPF_CHECK(le > ls, OPF_ERR_ARG, "empty local range in dim %d", d);
    }
    return OPF_OK;
}

// ---------------------------------------------------------------- ghost info
namespace opf {

int build_ghost_info(const opf_field* f, GhostInfo* gi) {
    const opf_field_desc& d = f->d;
    const opf_mesh* m = f->mesh;
    memset(gi, 0, sizeof(*gi));
    gi->dim = d.dim;
    for (int i = 0; i < 3; ++i) {
        gi->ls[i] = d.local_start[i];
        gi->le[i] = d.local_end[i];
        gi->gs[i] = i < d.dim ? std::min(d.logical_start[i], d.local_start[i]) : 0;
        gi->ge[i] = i < d.dim ? std::max(d.logical_end[i], d.local_end[i]) : 1;
        gi->off[i] = d.offset[i];
        gi->ext[i] = d.extent[i];
        // ghost box must be inside storage
        OPF_CHECK(gi->gs[i] >= d.offset[i] && gi->ge[i] <= d.offset[i] + d.extent[i], OPF_ERR_RANGE,
                  "logical range exceeds storage in dim %d (padding < ext)", i);
    }
    for (int i = 0; i < d.dim; ++i) {
        bool corner = d.loc[i] == OPF_LOC_CORNER;
        gi->per[i] = d.bc_type[i][0] == OPF_BC_PERIODIC;
        gi->corner[i] = corner ? 1 : 0;
        for (int s = 0; s < 2; ++s) {
            int nl = s == 0 ? d.local_start[i] - gi->gs[i] : gi->ge[i] - d.local_end[i];
            OPF_CHECK(nl <= 4, OPF_ERR_UNSUPPORTED, "more than 4 ghost layers");
            // every mirror / wrap source must be an interior index
            const int L = d.local_end[i] - d.local_start[i];
            OPF_CHECK(nl <= (corner && !gi->per[i] ? L - 1 : L), OPF_ERR_UNSUPPORTED,
                      "ghost width %d too large for %d interior points in dim %d", nl, L, i);
            gi->ngl[i][s] = nl;
            int t = d.bc_type[i][s];
            // step-1 fills only when localRange touches accessibleRange and bc is not periodic
            bool touches = s == 0 ? d.local_start[i] == d.access_start[i] : d.local_end[i] == d.access_end[i];
            if (t == OPF_BC_UNDEFINED || !touches) {
                gi->nobc[i][s] = gi->per[i] ? 0 : 1;
            }
            for (int l = 0; l < nl; ++l) {
                GhostRule& r = gi->rule[i][s][l];
                int c = s == 0 ? d.local_start[i] - 1 - l : d.local_end[i] + l;  // ghost coordinate
                int mi;                                                           // mirror coordinate
                if (s == 0)
                    mi = corner ? 2 * d.local_start[i] - c : 2 * d.local_start[i] - 1 - c;
                else
                    mi = corner ? 2 * d.local_end[i] - 2 - c : 2 * d.local_end[i] - 1 - c;
                double g = d.bc_value[i][s];
                if (gi->per[i]) {
                    r.type = 5;
                    continue;
                }
                if (gi->nobc[i][s]) {
                    r.type = 0;
                    continue;
                }
                OPF_CHECK(m->has_node(i, c) && m->has_node(i, mi) && m->has_dx(i, c) && m->has_dx(i, mi),
                          OPF_ERR_RANGE, "mesh padding too small for ghost layer %d of dim %d", l, i);
                switch (t) {
                    case OPF_BC_DIRC: {
                        // Interpolator1D::intp(x1, y1=bc, x2, y2=u_m, x) (Interpolator.hpp:21-40)
                        double x1, x2, x;
                        if (corner) {
                            x1 = s == 0 ? m->x(i, d.local_start[i]) : m->x(i, d.local_end[i] - 1);
                            x2 = m->x(i, mi);
                            x = m->x(i, c);
                        } else {
                            x1 = s == 0 ? m->x(i, d.local_start[i]) : m->x(i, d.local_end[i]);
                            x2 = m->x(i, mi) + m->dx(i, mi) / 2.;
                            x = m->x(i, c) + m->dx(i, c) / 2.;
                        }
                        double a = x1 - x, b = x2 - x;
                        r.type = 1;
                        r.c1 = a;
                        r.c2 = b * g;
                        r.c3 = a - b;
                    } break;
                    case OPF_BC_NEUM: {
                        double k = corner ? (m->x(i, c) - m->x(i, mi))
                                          : (m->x(i, c) + m->dx(i, c) / 2. - m->x(i, mi) - m->dx(i, mi) / 2.);
                        r.type = 2;
                        r.c1 = g * k;
                    } break;
                    case OPF_BC_SYMM:
                        r.type = 3;
                        break;
                    case OPF_BC_ASYMM:
                        r.type = 4;
                        break;
                    default:
                        set_error("bc type %d cannot fill ghosts", t);
                        return OPF_ERR_UNSUPPORTED;
                }
            }
        }
    }
    return OPF_OK;
}

template <typename T>
__device__ __forceinline__ T apply_rule(const GhostRule& r, T v) {
    switch (r.type) {
        case 1: return ((T)r.c1 * v - (T)r.c2) / (T)r.c3;
        case 2: return v + (T)r.c1;
        case 3: return v;
        case 4: return -v;
        default: return v;
    }
}

// Closed form of updatePaddingImpl_final's ordered passes (CartesianField.hpp:365-629):
// a ghost position p takes the value obtained by wrapping its periodic ghost
// coordinates, mirroring the others into the interior, and applying the
// non-periodic BC rules of its ghost dims in ascending dim order. Positions are
// enumerated as a disjoint union of slabs: slab k has coords j<k interior,
// coord k ghost, coords j>k anywhere in the ghost box.
template <typename T, typename I>
__global__ void ghost_kernel(T* __restrict__ data, GhostInfo gi, I n0, I n1, I n2) {
    const I t = (I)blockIdx.x * (I)blockDim.x + (I)threadIdx.x;
    int p[3], src[3], mir[3], lay[3], side[3];
    int act = ghost_locate(gi, t, n0, n1, n2, p, src, mir, lay, side);
    if (act == GHOST_SKIP) return;
    if (act == GHOST_COPY) {
        data[ghost_lin(gi, p)] = data[ghost_lin(gi, src)];
        return;
    }
    T v = data[ghost_lin(gi, mir)];
    for (int j = 0; j < 3; ++j)
        if (lay[j] >= 0) v = apply_rule<T>(gi.rule[j][side[j]][lay[j]], v);
    data[ghost_lin(gi, p)] = v;
}

// step 0 of updatePaddingImpl_final: Dirichlet boundary nodes of corner-located dims
template <typename T>
__global__ void dirc_node_kernel(T* __restrict__ data, GhostInfo gi, int dimk, int coord, T val) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int p[3];
    for (int j = 0; j < 3; ++j) {
        int len = j == dimk ? 1 : gi.le[j] - gi.ls[j];
        int r = (int)(t % len);
        t /= len;
        p[j] = j == dimk ? coord : gi.ls[j] + r;
    }
    if (t > 0) return;
    long long lp = (p[0] - gi.off[0]) + (long long)gi.ext[0] * ((p[1] - gi.off[1]) + (long long)gi.ext[1] * (p[2] - gi.off[2]));
    data[lp] = val;
}

static long long slab_count(const GhostInfo& gi, int k) {
    long long c = 1;
    for (int j = 0; j < 3; ++j) {
        if (j < k) c *= gi.le[j] - gi.ls[j];
        else if (j == k) c *= (gi.ls[j] - gi.gs[j]) + (gi.ge[j] - gi.le[j]);
        else c *= gi.ge[j] - gi.gs[j];
    }
    return c;
}

template <typename T>
static int ghost_fill_t(const opf_field* f, void* data) {
    GhostInfo gi;
    OPF_TRY(build_ghost_info(f, &gi));
    hipStream_t st = f->ctx->stream;
    const opf_field_desc& d = f->d;
    // step 0
    for (int i = 0; i < d.dim; ++i) {
        if (d.loc[i] != OPF_LOC_CORNER) continue;
        for (int s = 0; s < 2; ++s) {
            bool touches = s == 0 ? d.local_start[i] == d.access_start[i] : d.local_end[i] == d.access_end[i];
            if (!touches || d.bc_type[i][s] != OPF_BC_DIRC) continue;
            long long n = 1;
            for (int j = 0; j < 3; ++j) n *= j == i ? 1 : gi.le[j] - gi.ls[j];
            int coord = s == 0 ? d.local_start[i] : d.local_end[i] - 1;
            hipLaunchKernelGGL(dirc_node_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (T*)data, gi,
                               i, coord, (T)d.bc_value[i][s]);
        }
    }
    long long n0 = slab_count(gi, 0), n1 = d.dim > 1 ? slab_count(gi, 1) : 0, n2 = d.dim > 2 ? slab_count(gi, 2) : 0;
    long long n = n0 + n1 + n2;
    if (n > 0 && n + 256 < (1LL << 32))  // 32-bit thread indices
        hipLaunchKernelGGL((ghost_kernel<T, uint32_t>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (T*)data,
                           gi, (uint32_t)n0, (uint32_t)n1, (uint32_t)n2);
    else if (n > 0)
        hipLaunchKernelGGL((ghost_kernel<T, long long>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (T*)data,
                           gi, n0, n1, n2);
    OPF_HIP(hipGetLastError());
    return OPF_OK;
}

int launch_ghost_fill(const opf_field* f, void* data) {
    return f->d.dtype == OPF_F64 ? ghost_fill_t<double>(f, data) : ghost_fill_t<float>(f, data);
}

}  // namespace opf

// ---------------------------------------------------------------- fields
extern "C" int opf_field_create(opf_mesh* mesh, const opf_field_desc* desc, opf_field** out) {
    OPF_CHECK(mesh && desc && out, OPF_ERR_ARG, "opf_field_create: null arg");
    OPF_CHECK(desc->dim == mesh->dim, OPF_ERR_ARG, "field dim %d != mesh dim %d", desc->dim, mesh->dim);
    for (int d = 0; d < desc->dim; ++d)
        OPF_CHECK(desc->mesh_start[d] == mesh->start[d] && desc->mesh_end[d] == mesh->start[d] + mesh->nnodes[d],
                  OPF_ERR_ARG, "descriptor not finalized for this mesh (dim %d)", d);
    auto* f = new opf_field();
    f->ctx = mesh->ctx;
    f->mesh = mesh;
    f->d = *desc;
    f->count = (size_t)desc->extent[0] * desc->extent[1] * desc->extent[2];
    f->nbytes = f->count * dtype_size(desc->dtype);
    hipError_t e = hipMalloc(&f->data, f->nbytes);
    if (e != hipSuccess) {
        delete f;
        set_error("opf_field_create: hipMalloc(%zu) %s", (size_t)0, hipGetErrorString(e));
        return OPF_ERR_NOMEM;
    }
    OPF_HIP(hipMemsetAsync(f->data, 0, f->nbytes, f->ctx->stream));
    int rc = launch_ghost_fill(f, f->data);
    if (rc != OPF_OK) {
        hipFree(f->data);
        delete f;
        return rc;
    }
    *out = f;
    return OPF_OK;
}

extern "C" int opf_field_destroy(opf_field* f) {
    if (!f) return OPF_OK;
    hipStreamSynchronize(f->ctx->stream);
    if (f->data) hipFree(f->data);
    if (f->spare) hipFree(f->spare);
    delete f;
    return OPF_OK;
}

extern "C" const opf_field_desc* opf_field_get_desc(const opf_field* f) { return f ? &f->d : nullptr; }
extern "C" size_t opf_field_nbytes(const opf_field* f) { return f ? f->nbytes : 0; }
extern "C" void* opf_field_device_ptr(opf_field* f) { return f ? f->data : nullptr; }

extern "C" int opf_field_upload(opf_field* f, const void* host, size_t nbytes) {
    OPF_CHECK(f && host && nbytes == f->nbytes, OPF_ERR_ARG, "opf_field_upload: size %zu != %zu", nbytes,
              f ? f->nbytes : 0);
    OPF_HIP(hipMemcpyAsync(f->data, host, nbytes, hipMemcpyHostToDevice, f->ctx->stream));
    OPF_HIP(hipStreamSynchronize(f->ctx->stream));
    f->spare_valid = false;
    return OPF_OK;
}

extern "C" int opf_field_download(const opf_field* f, void* host, size_t nbytes) {
    OPF_CHECK(f && host && nbytes == f->nbytes, OPF_ERR_ARG, "opf_field_download: size mismatch");
    OPF_HIP(hipMemcpyAsync(host, f->data, nbytes, hipMemcpyDeviceToHost, f->ctx->stream));
    OPF_HIP(hipStreamSynchronize(f->ctx->stream));
    return OPF_OK;
}

extern "C" int opf_field_update_padding(opf_field* f) {
    OPF_CHECK(f, OPF_ERR_ARG, "null field");
    return launch_ghost_fill(f, f->data);
}

static bool same_shape(const opf_field_desc& a, const opf_field_desc& b) {
    if (a.dim != b.dim || a.dtype != b.dtype) return false;
    for (int i = 0; i < 3; ++i)
        if (a.extent[i] != b.extent[i] || a.offset[i] != b.offset[i]) return false;
    return true;
}

extern "C" int opf_field_copy(opf_field* dst, const opf_field* src) {
    OPF_CHECK(dst && src && same_shape(dst->d, src->d), OPF_ERR_ARG, "opf_field_copy: shape mismatch");
    if (dst == src) return OPF_OK;
    OPF_HIP(hipMemcpyAsync(dst->data, src->data, dst->nbytes, hipMemcpyDeviceToDevice, dst->ctx->stream));
    dst->spare_valid = false;
    return OPF_OK;
}

extern "C" int opf_field_swap(opf_field* a, opf_field* b) {
    OPF_CHECK(a && b && same_shape(a->d, b->d), OPF_ERR_ARG, "opf_field_swap: shape mismatch");
    std::swap(a->data, b->data);
    std::swap(a->spare, b->spare);
    std::swap(a->spare_valid, b->spare_valid);
    return OPF_OK;
}

template <typename T>
__global__ void assign_const_kernel(T* __restrict__ data, int op, T c, int s0, int s1, int s2, int n0, int n1, int n2,
                                    int o0, int o1, int o2, int e0, int e1) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long tot = (long long)n0 * n1 * n2;
    if (t >= tot) return;
    int i = (int)(t % n0) + s0;
    long long r = t / n0;
    int j = (int)(r % n1) + s1;
    int k = (int)(r / n1) + s2;
    long long l = (i - o0) + (long long)e0 * ((j - o1) + (long long)e1 * (k - o2));
    T v = data[l];
    switch (op) {
        case OPF_OP_EQ: v = c; break;
        case OPF_OP_ADD: v += c; break;
        case OPF_OP_MINUS: v -= c; break;
        case OPF_OP_MUL: v *= c; break;
        case OPF_OP_DIV: v /= c; break;
    }
    data[l] = v;
}

extern "C" int opf_field_assign_const(opf_field* f, int op, double c) {
    OPF_CHECK(f && op >= OPF_OP_EQ && op <= OPF_OP_DIV, OPF_ERR_ARG, "opf_field_assign_const: bad args");
    const opf_field_desc& d = f->d;
    int s[3], n[3];
    for (int i = 0; i < 3; ++i) {
        s[i] = std::max(d.assign_start[i], d.local_start[i]);
        n[i] = std::min(d.assign_end[i], d.local_end[i]) - s[i];
        if (n[i] <= 0) return launch_ghost_fill(f, f->data);
    }
    long long tot = (long long)n[0] * n[1] * n[2];
    if (d.dtype == OPF_F64)
        hipLaunchKernelGGL(assign_const_kernel<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                           f->ctx->stream, (double*)f->data, op, c, s[0], s[1], s[2], n[0], n[1], n[2], d.offset[0],
                           d.offset[1], d.offset[2], d.extent[0], d.extent[1]);
    else
        hipLaunchKernelGGL(assign_const_kernel<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                           f->ctx->stream, (float*)f->data, op, (float)c, s[0], s[1], s[2], n[0], n[1], n[2],
                           d.offset[0], d.offset[1], d.offset[2], d.extent[0], d.extent[1]);
    OPF_HIP(hipGetLastError());
    return launch_ghost_fill(f, f->data);
}

// ---------------------------------------------------------------- reductions
namespace opf {

__device__ __forceinline__ double red_combine(int red, double a, double b) {
    switch (red) {
        case OPF_RED_MAX:
        case OPF_RED_MAXABS: return a > b ? a : b;
        case OPF_RED_MIN: return a < b ? a : b;
        default: return a + b;
    }
}
__device__ __forceinline__ double red_identity(int red) {
    switch (red) {
        case OPF_RED_MAX: return -INFINITY;
        case OPF_RED_MIN: return INFINITY;
        default: return 0.0;
    }
}

// Block tree reduction in a fixed order (wave64 butterfly then across waves).
__device__ double block_reduce(double v, int red) {
    __shared__ double sh[16];
    for (int o = 32; o > 0; o >>= 1) v = red_combine(red, v, __shfl_xor(v, o, 64));
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    int nw = (blockDim.x + 63) >> 6;
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int i = 1; i < nw; ++i) r = red_combine(red, r, sh[i]);
        sh[0] = r;
    }
    __syncthreads();
    double r = sh[0];
    __syncthreads();
    return r;
}

template <typename T>
__global__ void reduce_pass1(const T* __restrict__ data, int red, int s0, int s1, int s2, int n0, int n1, int n2,
                             int o0, int o1, int o2, int e0, int e1, long long chunk, double* __restrict__ part) {
    long long tot = (long long)n0 * n1 * n2;
    long long beg = (long long)blockIdx.x * chunk;
    long long end = beg + chunk < tot ? beg + chunk : tot;
    double acc = red_identity(red);
    for (long long t = beg + threadIdx.x; t < end; t += blockDim.x) {
        int i = (int)(t % n0) + s0;
        long long r = t / n0;
        int j = (int)(r % n1) + s1;
        int k = (int)(r / n1) + s2;
        double v = (double)data[(i - o0) + (long long)e0 * ((j - o1) + (long long)e1 * (k - o2))];
        if (red == OPF_RED_SUMSQ) v = v * v;
        if (red == OPF_RED_MAXABS) v = fabs(v);
        acc = red_combine(red, acc, v);
    }
    acc = block_reduce(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void reduce_pass2(const double* __restrict__ part, int n, int red, double* __restrict__ out) {
    double acc = red_identity(red);
    for (int t = threadIdx.x; t < n; t += blockDim.x) acc = red_combine(red, acc, part[t]);
    acc = block_reduce(acc, red);
    if (threadIdx.x == 0) out[0] = acc;
}

}  // namespace opf

extern "C" int opf_reduce(const opf_field* f, int red, const int rs[3], const int re[3], double* out) {
    OPF_CHECK(f && rs && re && out && red >= 0 && red <= OPF_RED_MAXABS, OPF_ERR_ARG, "opf_reduce: bad args");
    const opf_field_desc& d = f->d;
    int s[3], n[3];
    for (int i = 0; i < 3; ++i) {
        s[i] = i < d.dim ? rs[i] : 0;
        n[i] = i < d.dim ? re[i] - rs[i] : 1;
        OPF_CHECK(s[i] >= d.offset[i] && s[i] + n[i] <= d.offset[i] + d.extent[i], OPF_ERR_RANGE,
                  "opf_reduce: range outside storage in dim %d", i);
        if (n[i] <= 0) {
            *out = red == OPF_RED_MAX ? -INFINITY : red == OPF_RED_MIN ? INFINITY : 0.0;
            return OPF_OK;
        }
    }
    opf_ctx* c = f->ctx;
    long long tot = (long long)n[0] * n[1] * n[2];
    long long nb = std::min<long long>(2048, (tot + 255) / 256);
    long long chunk = (tot + nb - 1) / nb;
    nb = (tot + chunk - 1) / chunk;
    if (d.dtype == OPF_F64)
        hipLaunchKernelGGL(reduce_pass1<double>, dim3((unsigned)nb), dim3(256), 0, c->stream, (const double*)f->data,
                           red, s[0], s[1], s[2], n[0], n[1], n[2], d.offset[0], d.offset[1], d.offset[2],
                           d.extent[0], d.extent[1], chunk, c->red_buf);
    else
        hipLaunchKernelGGL(reduce_pass1<float>, dim3((unsigned)nb), dim3(256), 0, c->stream, (const float*)f->data,
                           red, s[0], s[1], s[2], n[0], n[1], n[2], d.offset[0], d.offset[1], d.offset[2],
                           d.extent[0], d.extent[1], chunk, c->red_buf);
    hipLaunchKernelGGL(reduce_pass2, dim3(1), dim3(256), 0, c->stream, c->red_buf, (int)nb, red,
                       c->red_buf + c->red_cap - 1);
    OPF_HIP(hipGetLastError());
    OPF_HIP(hipMemcpyAsync(c->red_host, c->red_buf + c->red_cap - 1, sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    OPF_HIP(hipStreamSynchronize(c->stream));
    *out = c->red_host[0];
    return OPF_OK;
}

// ---------------------------------------------------------------- synthetic inputs
namespace {
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double normal() {
        // Box-Muller, first output only
        double u1 = uniform(), u2 = uniform();
        if (u1 <= 0.0) u1 = 0x1.0p-53;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};
}  // namespace

extern "C" int opf_gen_uniform(uint64_t seed, size_t n, double lo, double hi, double* out) {
    OPF_CHECK(out || n == 0, OPF_ERR_ARG, "null out");
    SplitMix64 g{seed};
    for (size_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * g.uniform();
    return OPF_OK;
}

namespace opf {
// bumps on a unit-cube cell-centred grid (x-fastest, no ghosts); the same
// generator and coordinates (h*i + 0.5*h) as opf_gen_bumps on a uniform mesh
void gen_bumps_grid(int dim, const int n[3], uint64_t seed, int nbumps, double noise, double* out) {
    SplitMix64 g{seed};
    double c[64][3], w[64], A[64];
    nbumps = std::min(std::max(nbumps, 0), 64);
    for (int b = 0; b < nbumps; ++b) {
        for (int k = 0; k < 3; ++k) c[b][k] = k < dim ? 0.2 + 0.6 * g.uniform() : 0.0;
        w[b] = 0.03 + 0.07 * g.uniform();
        A[b] = 0.5 + 0.5 * g.uniform();
    }
    int nn[3] = {n[0], dim > 1 ? n[1] : 1, dim > 2 ? n[2] : 1};
    double h[3];
    for (int q = 0; q < 3; ++q) h[q] = 1.0 / nn[q];
    size_t l = 0;
    for (int k = 0; k < nn[2]; ++k)
        for (int j = 0; j < nn[1]; ++j)
            for (int i = 0; i < nn[0]; ++i, ++l) {
                int idx[3] = {i, j, k};
                double xc[3];
                for (int q = 0; q < 3; ++q) xc[q] = h[q] * idx[q] + .5 * h[q];
                double v = 0.0;
                for (int b = 0; b < nbumps; ++b) {
                    double r2 = 0.0;
                    for (int q = 0; q < dim; ++q) r2 += (xc[q] - c[b][q]) * (xc[q] - c[b][q]);
                    v += A[b] * std::exp(-r2 / (2.0 * w[b] * w[b]));
                }
                if (noise != 0.0) v += noise * g.normal();
                out[l] = v;
            }
}

// Device version of gen_bumps_grid: the same bump parameters (drawn on the
// host from the same SplitMix64 stream) and the same per-cell arithmetic; the
// noise of cell l uses draws P + 2l + 1 and P + 2l + 2 of the stream (P = the
// parameter draws), which is where the sequential host generator takes them.
// exp/log/cos are the device's, so values may differ from the host's in the
// last bit.
namespace {
constexpr int DEV_MAX_BUMPS = 32;
struct DevBumps {
    int dim, nb;
    int n[3];
    double h[3];
    double c[DEV_MAX_BUMPS][3], w[DEV_MAX_BUMPS], A[DEV_MAX_BUMPS];
    double noise;
    uint64_t seed, p;
};

__device__ __forceinline__ double smx_uniform(uint64_t seed, uint64_t k) {
    uint64_t z = seed + k * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

__global__ void gen_bumps_kernel(DevBumps a, double* __restrict__ out, long long N) {
    for (long long l = blockIdx.x * (long long)blockDim.x + threadIdx.x; l < N;
         l += (long long)gridDim.x * blockDim.x) {
        const int idx[3] = {(int)(l % a.n[0]), (int)((l / a.n[0]) % a.n[1]), (int)(l / ((long long)a.n[0] * a.n[1]))};
        double xc[3];
        for (int q = 0; q < 3; ++q) xc[q] = a.h[q] * idx[q] + .5 * a.h[q];
        double v = 0.0;
        for (int b = 0; b < a.nb; ++b) {
            double r2 = 0.0;
            for (int q = 0; q < a.dim; ++q) r2 += (xc[q] - a.c[b][q]) * (xc[q] - a.c[b][q]);
            v += a.A[b] * exp(-r2 / (2.0 * a.w[b] * a.w[b]));
        }
        if (a.noise != 0.0) {
            double u1 = smx_uniform(a.seed, a.p + 2 * (uint64_t)l + 1), u2 = smx_uniform(a.seed, a.p + 2 * (uint64_t)l + 2);
            if (u1 <= 0.0) u1 = 0x1.0p-53;
            v += a.noise * (sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
        }
        out[l] = v;
    }
}
}  // namespace

int gen_bumps_grid_dev(hipStream_t st, int dim, const int n[3], uint64_t seed, int nbumps, double noise, double* out) {
    OPF_CHECK(nbumps >= 0 && nbumps <= DEV_MAX_BUMPS, OPF_ERR_ARG, "gen_bumps_grid_dev: 0..%d bumps", DEV_MAX_BUMPS);
    DevBumps a{};
    SplitMix64 g{seed};
    a.dim = dim;
    a.nb = nbumps;
    for (int b = 0; b < nbumps; ++b) {
        for (int k = 0; k < 3; ++k) a.c[b][k] = k < dim ? 0.2 + 0.6 * g.uniform() : 0.0;
        a.w[b] = 0.03 + 0.07 * g.uniform();
        a.A[b] = 0.5 + 0.5 * g.uniform();
    }
    for (int q = 0; q < 3; ++q) {
        a.n[q] = q < dim ? n[q] : 1;
        a.h[q] = 1.0 / a.n[q];
    }
    a.noise = noise;
    a.seed = seed;
    a.p = (uint64_t)nbumps * (dim + 2);
    const long long N = (long long)a.n[0] * a.n[1] * a.n[2];
    const int blocks = (int)std::min<long long>((N + 255) / 256, 8192);
    hipLaunchKernelGGL(gen_bumps_kernel, dim3(std::max(blocks, 1)), dim3(256), 0, st, a, out, N);
    OPF_HIP(hipGetLastError());
    return OPF_OK;
}
}  // namespace opf

extern "C" int opf_gen_bumps_grid_dev(opf_ctx* ctx, const opf_poisson_desc* d, uint64_t seed, int nbumps, double noise,
                                      double* dev_out) {
    OPF_CHECK(ctx && d && dev_out && d->dim >= 1 && d->dim <= 3, OPF_ERR_ARG, "opf_gen_bumps_grid_dev: bad args");
    return gen_bumps_grid_dev(ctx->stream, d->dim, d->n, seed, nbumps, noise, dev_out);
}

extern "C" int opf_gen_bumps(const opf_mesh* m, const opf_field_desc* d, uint64_t seed, int nbumps, double noise,
                             double* out) {
    OPF_CHECK(m && d && out && nbumps >= 0 && nbumps <= 64, OPF_ERR_ARG, "opf_gen_bumps: bad args");
    SplitMix64 g{seed};
    double c[64][3], w[64], A[64];
    for (int b = 0; b < nbumps; ++b) {
        for (int k = 0; k < 3; ++k) c[b][k] = k < d->dim ? 0.2 + 0.6 * g.uniform() : 0.0;
        w[b] = 0.03 + 0.07 * g.uniform();
        A[b] = 0.5 + 0.5 * g.uniform();
    }
    size_t tot = (size_t)d->extent[0] * d->extent[1] * d->extent[2];
    for (size_t i = 0; i < tot; ++i) out[i] = 0.0;
    for (int k = d->local_start[2]; k < d->local_end[2]; ++k)
        for (int j = d->local_start[1]; j < d->local_end[1]; ++j)
            for (int i = d->local_start[0]; i < d->local_end[0]; ++i) {
                int idx[3] = {i, j, k};
                double xc[3] = {0, 0, 0};
                for (int q = 0; q < d->dim; ++q)
                    xc[q] = d->loc[q] == OPF_LOC_CORNER ? m->x(q, idx[q]) : m->x(q, idx[q]) + .5 * m->dx(q, idx[q]);
                double v = 0.0;
                for (int b = 0; b < nbumps; ++b) {
                    double r2 = 0.0;
                    for (int q = 0; q < d->dim; ++q) r2 += (xc[q] - c[b][q]) * (xc[q] - c[b][q]);
                    v += A[b] * std::exp(-r2 / (2.0 * w[b] * w[b]));
                }
                if (noise != 0.0) v += noise * g.normal();
                size_t l = (size_t)(i - d->offset[0]) +
                           (size_t)d->extent[0] * ((size_t)(j - d->offset[1]) + (size_t)d->extent[1] * (k - d->offset[2]));
                out[l] = v;
            }
    return OPF_OK;
}
