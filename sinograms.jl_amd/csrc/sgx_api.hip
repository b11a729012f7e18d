// C ABI of libsgx: contexts, argument checking, staging of host arrays,
// per-call geometry tables and the fused plan (weight -> filter -> back).
//
// The argument checks reproduce the reference's error paths:
//   throw("size mismatch")  src/fbp2/back-par.jl:220, back-fan.jl:232,
//                           src/fbp3/cbct-back.jl:116        -> SGX_ERR_SIZE
//   error("nb=... + extra=... > npad=...")  src/fbp/filter.jl:131 -> SGX_ERR_VALUE

#include "sgx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

using sgx::Error;

// Device scratch of sgx_cbct_back_noalloc: the call's geometry tables and
// the staged backprojector's per-(block, view) window table, in one
// file-scope array of kScratchBytes on every GPU the library runs on (held
// until the process ends).  Host code never passes `scratch` itself to a
// kernel or a copy: it uses the per-device address that
// hipGetSymbolAddress returns, taken once per device when a context is
// created (scratch_ptr).
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// With `device` current: record the address of that device's `scratch`.
hipError_t scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[device] = p;
    return e;
}

unsigned char* scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices)
               ? static_cast<unsigned char*>(g_scratch_ptr[device])
               : nullptr;
}
}  // namespace

// Kernel shapes of one plan: the cone-beam backprojector (0 staged, 1
// gather); the 2-D backprojector's view grouping target and LDS view chunk.
struct PlanTune {
    int cbct = 0;
    int want_blocks = 2048;
    int vc_max = 64;
    int zform = 0;       // cone beam, staged kernel: block form (0 16 z, 1 mirror where the
                         // geometry allows it, 2 8 z, 3 8 z with wide windows)
};

// Largest view-grouping target and smallest LDS chunk the autotuner tries.
constexpr int kTuneMaxBlocks = 4096;
// View chunks of a cone-beam plan execute (SGX_OPT_VIEW_CHUNKS): the default,
// the most, and the fewest views a chunk may hold.  One chunk by default: on
// the 256^2 x 128 FDK workload (360 views) 2, 3 and 4 chunks took 1.27-1.30 ms
// per FP32 execute against 1.17 ms (the scaled-FP16 form 0.97-1.02 against
// 0.98 ms): the overlapped filters do not make up for the extra
// backprojection launches.
constexpr int kViewChunks = 1;
constexpr int kMaxViewChunks = 8;
constexpr int kMinChunkViews = 32;
constexpr int kTuneMinChunk = 16;

struct sgx_ctx {
    int device = 0;
    std::map<std::string, PlanTune> tune_cache;   // autotuned shapes per size bucket
    int cbct_kernel = 0;          // SGX_OPT_CBCT_KERNEL
    int cbct_mirror = 1;          // SGX_OPT_CBCT_MIRROR
    int view_chunks = kViewChunks;  // SGX_OPT_VIEW_CHUNKS
    std::string scratch_err;      // why the device scratch is unusable (empty: usable)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    struct Buf {
        void* p = nullptr;
        size_t n = 0;
    };
    Buf ws[12];

    // Device filter kernels of sgx_sino_filter (g, then the FFT tables), kept
    // per (Hk contents, nb, precision) so that repeated calls with one filter
    // skip the O(npad^2) host transform and the upload.  Least recently used
    // entries are dropped beyond kFilterCache.
    struct FilterEntry {
        std::vector<float> hk;
        int64_t nb = 0;
        int prec = 0;
        void* dev = nullptr;
        int64_t fft_n = 0;
        uint64_t last = 0;
    };
    static constexpr size_t kFilterCache = 8;
    std::vector<FilterEntry> filters;
    uint64_t filter_clock = 0;

    // Grow-only device workspace, one slot per role.  Reuse is ordered by the
    // context's stream; a grow waits for the stream before freeing.
    void* get(int slot, size_t bytes)
    {
        Buf& b = ws[slot];
        if (bytes == 0) return nullptr;
        if (b.n < bytes) {
            if (b.p) {
                SGX_HIP(hipStreamSynchronize(stream));
                SGX_HIP(hipFree(b.p));
                b.p = nullptr;
                b.n = 0;
            }
            SGX_HIP(hipMalloc(&b.p, bytes));
            b.n = bytes;
        }
        return b.p;
    }
    ~sgx_ctx()
    {
        for (Buf& b : ws)
            if (b.p) (void)hipFree(b.p);
        for (FilterEntry& e : filters)
            if (e.dev) (void)hipFree(e.dev);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

enum Slot {
    W_TABLES = 0,   // per-call view / axis tables
    W_PARTIAL,      // 2-D view-group partial images
    W_IN,           // staged input
    W_OUT,          // staged output
    W_MASK,         // staged mask
    W_WIN,          // cone-beam staged windows
    W_EXTRA,
};

template <class F>
int guard(sgx_ctx* ctx, F&& f)
{
    if (!ctx) return SGX_ERR_ARG;
    try {
        SGX_HIP(hipSetDevice(ctx->device));
        f();
        ctx->err.clear();
        return SGX_OK;
    } catch (const Error& e) {
        ctx->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return SGX_ERR_DEVICE;
    }
}

void require(bool ok, int code, const char* msg)
{
    if (!ok) throw Error(code, msg);
}

// Host -> device copy of a small host table into a region of a workspace slot.
struct TablePack {
    std::vector<unsigned char> host;
    size_t add(const void* src, size_t bytes)
    {
        const size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + bytes);
        std::memcpy(host.data() + off, src, bytes);
        return off;
    }
    unsigned char* upload(sgx_ctx* ctx, int slot)
    {
        auto* d = static_cast<unsigned char*>(ctx->get(slot, host.size()));
        SGX_HIP(hipMemcpyAsync(d, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream));
        // the pack's host memory must outlive the copy
        SGX_HIP(hipStreamSynchronize(ctx->stream));
        return d;
    }
};

std::vector<float> axis_over(const double* a, int64_t n, double ds)
{
    std::vector<float> v(n);
    for (int64_t i = 0; i < n; ++i) v[i] = (float)((float)a[i] / (float)ds);
    return v;
}

// Staging of a large array: device pointer as is, or a copy of host memory.
const void* stage_in(sgx_ctx* ctx, int slot, const void* p, size_t bytes, uint32_t flags)
{
    if (!(flags & SGX_HOST) || p == nullptr) return p;
    void* d = ctx->get(slot, bytes);
    SGX_HIP(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, ctx->stream));
    return d;
}

void finish(sgx_ctx* ctx, void* host_out, const void* dev_out, size_t bytes, uint32_t flags)
{
    if (flags & SGX_HOST) {
        SGX_HIP(hipMemcpyAsync(host_out, dev_out, bytes, hipMemcpyDeviceToHost, ctx->stream));
        SGX_HIP(hipStreamSynchronize(ctx->stream));
    } else if (!(flags & SGX_ASYNC)) {
        SGX_HIP(hipStreamSynchronize(ctx->stream));
    }
}

std::vector<double> filter_kernel_g(const float* Hk, int64_t npad)
{
    // g = real(ifft(Hk)): g[k] = (1/npad) sum_j Hk[j] cos(2 pi j k / npad)
    std::vector<double> cs(npad), g(npad);
    const double w = 2.0 * M_PI / (double)npad;
    for (int64_t j = 0; j < npad; ++j) cs[j] = std::cos(w * (double)j);
    for (int64_t k = 0; k < npad; ++k) {
        double acc = 0;
        int64_t ph = 0;
        for (int64_t j = 0; j < npad; ++j) {
            acc += (double)Hk[j] * cs[ph];
            ph += k;
            if (ph >= npad) ph -= npad;
        }
        g[k] = acc / (double)npad;
    }
    return g;
}

// The context's device filter kernel for (Hk, nb, prec): g = real(ifft(Hk))
// in FP32, then (FP32 at an FFT size) the tables of the FFT form.  Built and
// uploaded on a miss; a hit is a compare of npad floats.
const sgx_ctx::FilterEntry& filter_entry(sgx_ctx* ctx, const float* Hk, int64_t npad,
                                         int64_t nb, int prec)
{
    auto& fs = ctx->filters;
    const uint64_t now = ++ctx->filter_clock;
    for (auto& e : fs) {
        if (e.nb == nb && e.prec == prec && (int64_t)e.hk.size() == npad &&
            std::memcmp(e.hk.data(), Hk, sizeof(float) * npad) == 0) {
            e.last = now;
            return e;
        }
    }
    const auto gd = filter_kernel_g(Hk, npad);
    std::vector<float> gf(gd.begin(), gd.end());
    const auto ft = prec == SGX_PREC_FP32
                        ? sgx::fft_filter_tables(Hk, npad, nb, nullptr, nullptr, 0, 0.0)
                        : std::vector<float>();
    gf.insert(gf.end(), ft.begin(), ft.end());
    if (fs.size() >= sgx_ctx::kFilterCache) {
        // drop the least recently used entry (kernels queued on the stream
        // may still read it)
        auto old = std::min_element(fs.begin(), fs.end(), [](const auto& a, const auto& b) {
            return a.last < b.last;
        });
        SGX_HIP(hipStreamSynchronize(ctx->stream));
        SGX_HIP(hipFree(old->dev));
        fs.erase(old);
    }
    sgx_ctx::FilterEntry e;
    e.hk.assign(Hk, Hk + npad);
    e.nb = nb;
    e.prec = prec;
    e.fft_n = (int64_t)(ft.size() / 5);
    e.last = now;
    SGX_HIP(hipMalloc(&e.dev, 4 * gf.size()));
    try {
        SGX_HIP(hipMemcpyAsync(e.dev, gf.data(), 4 * gf.size(), hipMemcpyHostToDevice,
                               ctx->stream));
        SGX_HIP(hipStreamSynchronize(ctx->stream));   // gf is a local
    } catch (...) {
        (void)hipFree(e.dev);
        throw;
    }
    fs.push_back(std::move(e));
    return fs.back();
}

void back2d_common(sgx_ctx* ctx, int geom, const float* sino, int64_t nb, int64_t na,
                   int64_t ia_skip, const double* sinb, const double* cosb, double dsd,
                   double dso, double source_offset, double ds, double offset, const double* xc,
                   int64_t nx, const double* yc, int64_t ny, const uint8_t* mask, float* image,
                   uint32_t flags)
{
    require(sino && sinb && cosb && xc && yc && image, SGX_ERR_ARG, "null pointer argument");
    require(nb > 1 && na > 0 && nx > 0 && ny > 0 && ia_skip >= 1, SGX_ERR_ARG, "bad size");
    require(nb < (1 << 30) && nx * ny < (int64_t(1) << 31), SGX_ERR_UNSUPPORTED, "too large");
    require(ds != 0.0, SGX_ERR_ARG, "ds must be nonzero");
    const int64_t nv = (na + ia_skip - 1) / ia_skip;

    TablePack pack;
    std::vector<sgx::ViewCS> cs(nv);
    for (int64_t i = 0; i < nv; ++i) cs[i] = {(float)cosb[i], (float)sinb[i]};
    const size_t o_cs = pack.add(cs.data(), cs.size() * sizeof(sgx::ViewCS));
    const auto xv = axis_over(xc, nx, ds), yv = axis_over(yc, ny, ds);
    const size_t o_x = pack.add(xv.data(), xv.size() * 4);
    const size_t o_y = pack.add(yv.data(), yv.size() * 4);
    unsigned char* d = pack.upload(ctx, W_TABLES);

    const float* q = static_cast<const float*>(
        stage_in(ctx, W_IN, sino, sizeof(float) * nb * na, flags));
    const uint8_t* m = static_cast<const uint8_t*>(stage_in(ctx, W_MASK, mask, nx * ny, flags));
    float* img = (flags & SGX_HOST) ? static_cast<float*>(ctx->get(W_OUT, 4 * nx * ny)) : image;

    sgx::Back2dArgs a{};
    a.q = q;
    a.nb = (int)nb;
    a.nviews = (int)nv;
    a.view_stride = nb * ia_skip;
    a.cs = reinterpret_cast<const sgx::ViewCS*>(d + o_cs);
    a.xds = reinterpret_cast<const float*>(d + o_x);
    a.yds = reinterpret_cast<const float*>(d + o_y);
    a.nx = (int)nx;
    a.ny = (int)ny;
    a.mask = m;
    a.image = img;
    a.geom = geom;
    a.wb = (float)((nb + 1) / 2.0 + (double)(float)offset);
    a.dsd_ds = (float)dsd / (float)ds;
    a.dso_ds = (float)dso / (float)ds;
    a.so_ds = (float)source_offset / (float)ds;
    const int64_t pf = sgx::back2d_partial_floats(a.nb, a.nviews, a.nx, a.ny, a.want_blocks,
                                                  a.vc_max);
    a.partial = static_cast<float*>(ctx->get(W_PARTIAL, 4 * pf));
    a.partial_cap = pf;
    if (geom == 1)
        sgx::back2d_rel_bounds(xc, nx, yc, ny, (flags & SGX_HOST) ? mask : nullptr, ds,
                               a.rel_rmax, a.rel_h);
    sgx::launch_back2d(a, ctx->stream);
    finish(ctx, image, img, 4 * nx * ny, flags);
}

}  // namespace

extern "C" {

int sgx_version(void) { return 100; }

int sgx_ctx_create(int device, sgx_ctx** out)
{
    if (!out) return SGX_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return SGX_ERR_DEVICE;
    auto* c = new (std::nothrow) sgx_ctx();
    if (!c) return SGX_ERR_DEVICE;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return SGX_ERR_DEVICE;
    }
    c->own_stream = true;
    // the address of this device's scratch (sgx_cbct_back_noalloc); a failure
    // here leaves that entry point unusable on the device, nothing else
    const hipError_t se = scratch_setup(device);
    c->scratch_err = se == hipSuccess ? std::string()
                                      : std::string("hipGetSymbolAddress(scratch): ") +
                                            hipGetErrorString(se);
    *out = c;
    return SGX_OK;
}

void sgx_ctx_destroy(sgx_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
}

const char* sgx_last_error(const sgx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sgx_ctx_set_stream(sgx_ctx* ctx, void* hip_stream)
{
    return guard(ctx, [&] {
        SGX_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->own_stream && ctx->stream) SGX_HIP(hipStreamDestroy(ctx->stream));
        ctx->stream = static_cast<hipStream_t>(hip_stream);
        ctx->own_stream = false;
    });
}

int sgx_ctx_synchronize(sgx_ctx* ctx)
{
    return guard(ctx, [&] { SGX_HIP(hipStreamSynchronize(ctx->stream)); });
}

int sgx_ctx_set_option(sgx_ctx* ctx, int option, int64_t value)
{
    return guard(ctx, [&] {
        require(option == SGX_OPT_CBCT_KERNEL || option == SGX_OPT_CBCT_MIRROR ||
                    option == SGX_OPT_VIEW_CHUNKS,
                SGX_ERR_ARG, "unknown option");
        if (option == SGX_OPT_VIEW_CHUNKS) {
            require(value >= 1 && value <= kMaxViewChunks, SGX_ERR_ARG,
                    "SGX_OPT_VIEW_CHUNKS is 1 to 8");
            ctx->view_chunks = (int)value;
            return;
        }
        if (option == SGX_OPT_CBCT_MIRROR) {
            require(value >= 0 && value <= 2, SGX_ERR_ARG, "SGX_OPT_CBCT_MIRROR is 0, 1 or 2");
            ctx->cbct_mirror = (int)value;
            return;
        }
        require(value >= 0 && value <= 4, SGX_ERR_ARG, "SGX_OPT_CBCT_KERNEL is 0 to 4");
        ctx->cbct_kernel = (int)value;
    });
}

int sgx_ctx_get_option(const sgx_ctx* ctx, int option, int64_t* value)
{
    if (!ctx || !value ||
        (option != SGX_OPT_CBCT_KERNEL && option != SGX_OPT_CBCT_MIRROR &&
         option != SGX_OPT_VIEW_CHUNKS))
        return SGX_ERR_ARG;
    *value = option == SGX_OPT_CBCT_MIRROR  ? ctx->cbct_mirror
             : option == SGX_OPT_VIEW_CHUNKS ? ctx->view_chunks
                                             : ctx->cbct_kernel;
    return SGX_OK;
}

int sgx_back_par(sgx_ctx* ctx, const float* sino, int64_t nb, int64_t na, int64_t ia_skip,
                 const double* sinphi, const double* cosphi, double ds, double offset,
                 const double* xc, int64_t nx, const double* yc, int64_t ny,
                 const uint8_t* mask, float* image, uint32_t flags)
{
    return guard(ctx, [&] {
        back2d_common(ctx, 0, sino, nb, na, ia_skip, sinphi, cosphi, 0, 0, 0, ds, offset, xc,
                      nx, yc, ny, mask, image, flags);
    });
}

int sgx_back_fan(sgx_ctx* ctx, const float* sino, int64_t nb, int64_t na, int64_t ia_skip,
                 const double* sinb, const double* cosb, double dsd, double dso,
                 double source_offset, int is_arc, double ds, double offset,
                 const double* xc, int64_t nx, const double* yc, int64_t ny,
                 const uint8_t* mask, float* image, uint32_t flags)
{
    return guard(ctx, [&] {
        back2d_common(ctx, is_arc ? 1 : 2, sino, nb, na, ia_skip, sinb, cosb, dsd, dso,
                      source_offset, ds, offset, xc, nx, yc, ny, mask, image, flags);
    });
}

}  // extern "C"

namespace {

// Host tables of one cone-beam call: the used views' angles, the axes / ds,
// room for the per-view cos/sin (formed on the device) and the source z.
struct CbctCall {
    TablePack pack;
    size_t o_b = 0, o_x = 0, o_y = 0, o_z = 0, o_cs = 0, o_zs = 0;
    int64_t nv = 0;
    bool has_zs = false;
    bool z_sym = false;     // the z axis (/ ds, FP32) is symmetric: z[nz-1-k] = -z[k]
};

// z[n-1-k] == -z[k] for every k, exactly in FP32
bool axis_symmetric(const std::vector<float>& z)
{
    const size_t n = z.size();
    for (size_t k = 0; k < n; ++k)
        if (z[n - 1 - k] != -z[k]) return false;
    return true;
}

void cbct_check(const void* proj, int64_t ns, int64_t nt, int64_t na, int64_t ia_skip,
                const double* betas, double ds, double dt, const double* xc, int64_t nx,
                const double* yc, int64_t ny, const double* zc, int64_t nz, const float* image,
                uint32_t flags)
{
    require(proj && betas && xc && yc && zc && image, SGX_ERR_ARG, "null pointer argument");
    require(ns > 1 && nt > 1 && na > 0 && nx > 0 && ny > 0 && nz > 0 && ia_skip >= 1,
            SGX_ERR_ARG, "bad size");
    require(ds != 0.0 && dt != 0.0, SGX_ERR_ARG, "ds and dt must be nonzero");
    require(!((flags & SGX_Q_BF16) && (flags & SGX_HOST)), SGX_ERR_UNSUPPORTED,
            "bf16 projections must be device memory");
    require(nx * ny < (int64_t(1) << 31) / std::max<int64_t>(nz, 1), SGX_ERR_UNSUPPORTED,
            "volume too large");
}

CbctCall cbct_call_tables(int64_t na, int64_t ia_skip, const double* betas, double ds,
                          const double* xc, int64_t nx, const double* yc, int64_t ny,
                          const double* zc, int64_t nz, const sgx_cbct_opts* opts)
{
    CbctCall T;
    T.nv = (na + ia_skip - 1) / ia_skip;
    std::vector<double> b(T.nv);
    for (int64_t i = 0; i < T.nv; ++i) b[i] = betas[i * ia_skip];
    T.o_b = T.pack.add(b.data(), b.size() * 8);
    const auto xv = axis_over(xc, nx, ds), yv = axis_over(yc, ny, ds), zv = axis_over(zc, nz, ds);
    T.o_x = T.pack.add(xv.data(), xv.size() * 4);
    T.o_y = T.pack.add(yv.data(), yv.size() * 4);
    T.o_z = T.pack.add(zv.data(), zv.size() * 4);
    T.z_sym = axis_symmetric(zv);
    T.o_cs = T.pack.add(b.data(), T.nv * sizeof(sgx::ViewCS));   // space only
    T.has_zs = opts && opts->source_zs;
    if (T.has_zs) {
        std::vector<float> zsv(T.nv);
        for (int64_t i = 0; i < T.nv; ++i)
            zsv[i] = (float)((float)opts->source_zs[i * ia_skip] / (float)ds);
        T.o_zs = T.pack.add(zsv.data(), T.nv * 4);
    }
    return T;
}

// The kernels' arguments for tables uploaded at device address d (the
// per-view cos/sin still to be formed by launch_view_trig).
sgx::CbctArgs cbct_call_args(const CbctCall& T, unsigned char* d, const void* q, bool bf16,
                             int64_t ns, int64_t nt, int64_t pitch, int64_t ia_skip, double dsd,
                             double dso, double ds, double dt, double offset_s, double offset_t,
                             int is_arc, int64_t nx, int64_t ny, int64_t nz, const uint8_t* m,
                             float* img, const sgx_cbct_opts* opts)
{
    sgx::CbctArgs a{};
    a.q = q;
    a.fmt = bf16 ? sgx::SGX_Q_FMT_BF16 : sgx::SGX_Q_FMT_F32;
    a.vscale = nullptr;
    a.ns = (int)ns;
    a.nt = (int)nt;
    a.pitch = (int)pitch;
    a.nviews = (int)T.nv;
    a.view_stride = pitch * nt * ia_skip;
    a.cs = reinterpret_cast<sgx::ViewCS*>(d + T.o_cs);
    a.xds = reinterpret_cast<const float*>(d + T.o_x);
    a.yds = reinterpret_cast<const float*>(d + T.o_y);
    a.zds = reinterpret_cast<const float*>(d + T.o_z);
    a.nx = (int)nx;
    a.ny = (int)ny;
    a.nz = (int)nz;
    a.mask = m;
    a.image = img;
    a.dsd_ds = (float)dsd / (float)ds;
    a.dso_ds = (float)dso / (float)ds;
    a.ws = (float)((ns + 1) / 2.0 + (double)(float)offset_s);
    a.wt = (float)((nt + 1) / 2.0 + (double)(float)offset_t);
    a.inv_dt_ds = 1.0f / ((float)dt / (float)ds);
    a.is_arc = is_arc != 0;
    a.zs = T.has_zs ? reinterpret_cast<const float*>(d + T.o_zs) : nullptr;
    a.arc_t_exact = opts && opts->arc_t_exact;
    a.so_ds = opts ? (float)opts->source_offset / (float)ds : 0.f;
    return a;
}

void cbct_back_impl(sgx_ctx* ctx, const void* proj, int64_t ns, int64_t nt, int64_t na,
                    int64_t ia_skip, const double* betas, double dsd, double dso,
                    double ds, double dt, double offset_s, double offset_t, int is_arc,
                    const double* xc, int64_t nx, const double* yc, int64_t ny,
                    const double* zc, int64_t nz, const uint8_t* mask, float* image,
                    uint32_t flags, const sgx_cbct_opts* opts)
{
    cbct_check(proj, ns, nt, na, ia_skip, betas, ds, dt, xc, nx, yc, ny, zc, nz, image, flags);
    CbctCall T = cbct_call_tables(na, ia_skip, betas, ds, xc, nx, yc, ny, zc, nz, opts);
    unsigned char* d = T.pack.upload(ctx, W_TABLES);
    sgx::launch_view_trig(reinterpret_cast<const double*>(d + T.o_b), T.nv, 1,
                          reinterpret_cast<sgx::ViewCS*>(d + T.o_cs), ctx->stream);

    // The kernels take rows padded to a multiple of 4 bins (16-byte aligned)
    // and, for 16-bit data, readable bytes after the last view: device FP32
    // input that already has that shape is used in place, anything else is
    // copied into a padded buffer (pad bins are never read as taps).
    const size_t esz = (flags & SGX_Q_BF16) ? 2 : 4;
    const int64_t pitch = (ns + 3) / 4 * 4;
    const void* q = proj;
    if ((flags & (SGX_HOST | SGX_Q_BF16)) || pitch != ns ||
        reinterpret_cast<uintptr_t>(proj) % 16 != 0) {
        const size_t body = esz * pitch * nt * na;
        void* dq = ctx->get(W_IN, body + sgx::kQ16Pad);
        SGX_HIP(hipMemsetAsync(static_cast<char*>(dq) + body, 0, sgx::kQ16Pad, ctx->stream));
        SGX_HIP(hipMemcpy2DAsync(dq, esz * pitch, proj, esz * ns, esz * ns, nt * na,
                                 (flags & SGX_HOST) ? hipMemcpyHostToDevice
                                                    : hipMemcpyDeviceToDevice,
                                 ctx->stream));
        q = dq;
    }
    const int64_t nvox = nx * ny * nz;
    const uint8_t* m = static_cast<const uint8_t*>(stage_in(ctx, W_MASK, mask, nvox, flags));
    float* img = (flags & SGX_HOST) ? static_cast<float*>(ctx->get(W_OUT, 4 * nvox)) : image;
    sgx::CbctArgs a = cbct_call_args(T, d, q, (flags & SGX_Q_BF16) != 0, ns, nt, pitch, ia_skip,
                                     dsd, dso, ds, dt, offset_s, offset_t, is_arc, nx, ny, nz, m,
                                     img, opts);
    // the FP32 and BF16 data of this entry point take the mirror form only when
    // it is asked for everywhere (it measured slower for them)
    a.zform = ctx->cbct_mirror == 2 && sgx::cbct_mirror_ok(a, T.z_sym) ? 1
              : ctx->cbct_kernel == 3                                  ? 2
              : ctx->cbct_kernel == 4                                  ? 3
                                                                       : 0;
    if (ctx->cbct_kernel != 1) {
        int4* win = static_cast<int4*>(
            ctx->get(W_WIN, sizeof(int4) * sgx::cbct_window_entries(a)));
        sgx::launch_cbct_windows(a, win, ctx->stream);
        a.win = win;
    }
    sgx::launch_cbct_back(a, ctx->stream);
    finish(ctx, image, img, 4 * nvox, flags);
}
}  // namespace

extern "C" {

int sgx_cbct_back(sgx_ctx* ctx, const void* proj, int64_t ns, int64_t nt, int64_t na,
                  int64_t ia_skip, const double* betas, double dsd, double dso,
                  double ds, double dt, double offset_s, double offset_t, int is_arc,
                  const double* xc, int64_t nx, const double* yc, int64_t ny,
                  const double* zc, int64_t nz, const uint8_t* mask, float* image,
                  uint32_t flags)
{
    return guard(ctx, [&] {
        cbct_back_impl(ctx, proj, ns, nt, na, ia_skip, betas, dsd, dso, ds, dt, offset_s,
                       offset_t, is_arc, xc, nx, yc, ny, zc, nz, mask, image, flags, nullptr);
    });
}

int sgx_cbct_back_opts(sgx_ctx* ctx, const void* proj, int64_t ns, int64_t nt, int64_t na,
                       int64_t ia_skip, const double* betas, double dsd, double dso,
                       double ds, double dt, double offset_s, double offset_t, int is_arc,
                       const double* xc, int64_t nx, const double* yc, int64_t ny,
                       const double* zc, int64_t nz, const uint8_t* mask, float* image,
                       uint32_t flags, const sgx_cbct_opts* opts)
{
    return guard(ctx, [&] {
        cbct_back_impl(ctx, proj, ns, nt, na, ia_skip, betas, dsd, dso, ds, dt, offset_s,
                       offset_t, is_arc, xc, nx, yc, ny, zc, nz, mask, image, flags, opts);
    });
}

int sgx_cbct_back_noalloc(sgx_ctx* ctx, const float* proj, int64_t ns, int64_t nt, int64_t na,
                          int64_t ia_skip, const double* betas, double dsd, double dso,
                          double ds, double dt, double offset_s, double offset_t, int is_arc,
                          const double* xc, int64_t nx, const double* yc, int64_t ny,
                          const double* zc, int64_t nz, const uint8_t* mask, float* image,
                          uint32_t flags)
{
    return guard(ctx, [&] {
        require((flags & ~(uint32_t)SGX_ASYNC) == 0, SGX_ERR_UNSUPPORTED,
                "sgx_cbct_back_noalloc: device FP32 data only (flags: SGX_ASYNC)");
        cbct_check(proj, ns, nt, na, ia_skip, betas, ds, dt, xc, nx, yc, ny, zc, nz, image,
                   flags);
        require(ns % 4 == 0 && reinterpret_cast<uintptr_t>(proj) % 16 == 0,
                SGX_ERR_UNSUPPORTED,
                "sgx_cbct_back_noalloc: ns must be a multiple of 4 and proj 16-byte aligned "
                "(other shapes need a padded copy: use sgx_cbct_back)");
        require(ctx->scratch_err.empty(), SGX_ERR_DEVICE, ctx->scratch_err.c_str());
        unsigned char* base = scratch_ptr(ctx->device);
        require(base != nullptr, SGX_ERR_DEVICE, "device scratch not set up");
        CbctCall T = cbct_call_tables(na, ia_skip, betas, ds, xc, nx, yc, ny, zc, nz, nullptr);
        const size_t tbytes = (T.pack.host.size() + 255) & ~size_t(255);
        sgx::CbctArgs a = cbct_call_args(T, base, proj, false, ns, nt, ns, ia_skip, dsd, dso, ds,
                                         dt, offset_s, offset_t, is_arc, nx, ny, nz, mask, image,
                                         nullptr);
        a.zform = ctx->cbct_mirror == 2 && sgx::cbct_mirror_ok(a, T.z_sym) ? 1
                  : ctx->cbct_kernel == 3                                  ? 2
                  : ctx->cbct_kernel == 4                                  ? 3
                                                                           : 0;
        const bool staged = ctx->cbct_kernel != 1;
        const size_t need =
            tbytes + (staged ? sizeof(int4) * (size_t)sgx::cbct_window_entries(a) : 0);
        if (need > kScratchBytes) {
            std::ostringstream os;
            os << "sgx_cbct_back_noalloc: this call needs " << need << " bytes of scratch, "
               << "more than its " << kScratchBytes << " (use sgx_cbct_back)";
            throw Error(SGX_ERR_UNSUPPORTED, os.str());
        }
        SGX_HIP(hipMemcpyAsync(base, T.pack.host.data(), T.pack.host.size(),
                               hipMemcpyHostToDevice, ctx->stream));
        // the tables' host memory must outlive the copy
        SGX_HIP(hipStreamSynchronize(ctx->stream));
        sgx::launch_view_trig(reinterpret_cast<const double*>(base + T.o_b), T.nv, 1,
                              reinterpret_cast<sgx::ViewCS*>(base + T.o_cs), ctx->stream);
        if (staged) {
            int4* win = reinterpret_cast<int4*>(base + tbytes);
            sgx::launch_cbct_windows(a, win, ctx->stream);
            a.win = win;
        }
        sgx::launch_cbct_back(a, ctx->stream);
        if (!(flags & SGX_ASYNC)) SGX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int sgx_radon_ellipses(sgx_ctx* ctx, int kind, const double* s, int64_t nb,
                       const double* betas, int64_t na, double dsd, double dso,
                       const double* ell, int64_t nell, double* out, uint32_t flags)
{
    return guard(ctx, [&] {
        require(s && betas && out && (ell || nell == 0), SGX_ERR_ARG, "null pointer argument");
        require(kind == SGX_SINO_PAR || kind == SGX_SINO_FAN_ARC || kind == SGX_SINO_FAN_FLAT,
                SGX_ERR_ARG, "radon_ellipses: 2-D geometry kind");
        require(nb > 0 && na > 0 && nell >= 0 && nb * na < (int64_t(1) << 40), SGX_ERR_ARG,
                "bad size");
        require(kind == SGX_SINO_PAR || dsd != 0.0, SGX_ERR_ARG, "dsd must be nonzero");
        TablePack pack;
        const size_t o_s = pack.add(s, 8 * nb), o_b = pack.add(betas, 8 * na);
        const size_t o_e = nell ? pack.add(ell, 8 * 6 * nell) : 0;
        unsigned char* d = pack.upload(ctx, W_TABLES);
        double* dout = (flags & SGX_HOST)
                           ? static_cast<double*>(ctx->get(W_OUT, 8 * nb * na))
                           : out;
        sgx::launch_radon_ellipses(kind, reinterpret_cast<const double*>(d + o_s), (int)nb,
                                   reinterpret_cast<const double*>(d + o_b), (int)na, dsd, dso,
                                   reinterpret_cast<const double*>(d + o_e), (int)nell, dout,
                                   ctx->stream);
        finish(ctx, out, dout, 8 * nb * na, flags);
    });
}

int sgx_radon_ellipsoids(sgx_ctx* ctx, int kind, const double* s, int64_t ns,
                         const double* t, int64_t nt, const double* betas,
                         const double* source_zs, int64_t na, double dsd, double dso,
                         const double* ell, int64_t nell, double* out, uint32_t flags)
{
    return guard(ctx, [&] {
        require(s && t && betas && out && (ell || nell == 0), SGX_ERR_ARG,
                "null pointer argument");
        require(kind == SGX_CT_FAN_ARC || kind == SGX_CT_FAN_FLAT, SGX_ERR_ARG,
                "radon_ellipsoids: cone-beam geometry kind");
        require(ns > 0 && nt > 0 && na > 0 && nell >= 0 && ns * nt * na < (int64_t(1) << 40),
                SGX_ERR_ARG, "bad size");
        require(dsd != 0.0, SGX_ERR_ARG, "dsd must be nonzero");
        // the unrotated detector point of each bin (the cone rays' geometry)
        std::vector<double> lx(ns), ly(ns);
        for (int64_t i = 0; i < ns; ++i) {
            if (kind == SGX_CT_FAN_ARC) {
                const double g = s[i] / dsd;
                lx[i] = dsd * std::sin(g);
                ly[i] = dso - dsd * std::cos(g);
            } else {
                lx[i] = s[i];
                ly[i] = dso - dsd;
            }
        }
        TablePack pack;
        const size_t o_x = pack.add(lx.data(), 8 * ns), o_y = pack.add(ly.data(), 8 * ns);
        const size_t o_t = pack.add(t, 8 * nt), o_b = pack.add(betas, 8 * na);
        const size_t o_z = source_zs ? pack.add(source_zs, 8 * na) : 0;
        const size_t o_e = nell ? pack.add(ell, 8 * 7 * nell) : 0;
        unsigned char* d = pack.upload(ctx, W_TABLES);
        const int64_t n = ns * nt * na;
        double* dout = (flags & SGX_HOST) ? static_cast<double*>(ctx->get(W_OUT, 8 * n)) : out;
        auto P = [&](size_t o) { return reinterpret_cast<const double*>(d + o); };
        sgx::launch_radon_ellipsoids(P(o_x), P(o_y), (int)ns, P(o_t), (int)nt, P(o_b),
                                     source_zs ? P(o_z) : nullptr, (int)na, dso, P(o_e),
                                     (int)nell, dout, ctx->stream);
        finish(ctx, out, dout, 8 * n, flags);
    });
}

int sgx_sino_filter(sgx_ctx* ctx, const float* in, int64_t nb, int64_t ncols,
                    const float* Hk, int64_t npad, int64_t extra, int prec,
                    float* out, uint32_t flags)
{
    return guard(ctx, [&] {
        require(in && Hk && out, SGX_ERR_ARG, "null pointer argument");
        require(nb > 0 && ncols >= 0 && npad > 0 && extra >= 0, SGX_ERR_ARG, "bad size");
        require(prec == SGX_PREC_FP32 || prec == SGX_PREC_BF16, SGX_ERR_ARG, "bad precision");
        if (nb + extra > npad) {
            std::ostringstream os;
            os << "nb=" << nb << " + extra=" << extra << " > npad=" << npad;
            throw Error(SGX_ERR_VALUE, os.str());
        }
        const sgx_ctx::FilterEntry& fe = filter_entry(ctx, Hk, npad, nb, prec);
        float* g = static_cast<float*>(fe.dev);
        const int64_t nout = (nb + extra) * ncols;
        const float* din = static_cast<const float*>(
            stage_in(ctx, W_IN, in, 4 * nb * ncols, flags));
        float* dout = (flags & SGX_HOST) ? static_cast<float*>(ctx->get(W_OUT, 4 * nout)) : out;

        sgx::FilterArgs a{};
        a.in = din;
        a.nb = (int)nb;
        a.ncols = ncols;
        a.nt = 1;
        a.g = g;
        a.npad = (int)npad;
        a.extra = (int)extra;
        a.out = dout;
        a.out_bf16 = false;
        a.split_bf16 = prec == SGX_PREC_BF16;
        a.scale = 1.0f;
        if (fe.fft_n > 0) {
            a.fft = g + npad;
            a.fft_n = (int)fe.fft_n;
        }
        sgx::launch_sino_filter(a, ctx->stream);
        finish(ctx, out, dout, 4 * nout, flags);
    });
}

int sgx_parker_fan_short(sgx_ctx* ctx, int64_t nb, int64_t na, const double* ar,
                         const double* gam, double gamma_max, float* out, uint32_t flags)
{
    return guard(ctx, [&] {
        require(ar && gam && out, SGX_ERR_ARG, "null pointer argument");
        require(nb > 0 && na > 1, SGX_ERR_ARG, "bad size");
        // equal spacing check (parker.jl:73-74)
        const double da0 = ar[1] - ar[0];
        for (int64_t i = 1; i + 1 < na; ++i)
            require(std::fabs((ar[i + 1] - ar[i]) - da0) <= 1e-4 * std::fabs(da0) + 1e-12,
                    SGX_ERR_VALUE, "Unequal angles not yet supported");
        double amin = ar[0];
        for (int64_t i = 1; i < na; ++i) amin = std::min(amin, ar[i]);
        TablePack pack;
        std::vector<double> bet(na);
        for (int64_t i = 0; i < na; ++i) bet[i] = std::max(ar[i] - amin, 0.0);
        const size_t o_g = pack.add(gam, nb * 8);
        const size_t o_b = pack.add(bet.data(), na * 8);
        unsigned char* d = pack.upload(ctx, W_TABLES);
        float* o = (flags & SGX_HOST) ? static_cast<float*>(ctx->get(W_OUT, 4 * nb * na)) : out;
        sgx::launch_parker(reinterpret_cast<const double*>(d + o_g),
                           reinterpret_cast<const double*>(d + o_b), (int)nb, (int)na, gamma_max,
                           o, ctx->stream);
        finish(ctx, out, o, 4 * nb * na, flags);
    });
}

int sgx_project_bdd(sgx_ctx* ctx, const void* phantom, int64_t nPix, int64_t nDet,
                    const double* angle, int64_t nangle, double DSD, double DS0, double pSize,
                    double dSize, double* sino, uint32_t flags)
{
    return guard(ctx, [&] {
        require(phantom && angle && sino, SGX_ERR_ARG, "null pointer argument");
        require(nPix > 0 && nDet > 0 && nangle > 0, SGX_ERR_ARG, "bad size");
        require(nPix * nPix < (int64_t(1) << 31), SGX_ERR_UNSUPPORTED, "image too large");
        const bool f32 = (flags & SGX_IN_F32) != 0;
        const size_t isz = f32 ? 4 : 8;
        TablePack pack;
        const size_t o_a = pack.add(angle, 8 * nangle);
        unsigned char* d = pack.upload(ctx, W_TABLES);
        const void* img = stage_in(ctx, W_IN, phantom, isz * nPix * nPix, flags);
        const int64_t nout = nangle * nDet;
        double* out = (flags & SGX_HOST) ? static_cast<double*>(ctx->get(W_OUT, 8 * nout)) : sino;
        double* partial = static_cast<double*>(
            ctx->get(W_PARTIAL, 8 * sgx::bdd_project_partial_doubles((int)nPix, (int)nDet,
                                                                    (int)nangle)));
        sgx::launch_bdd_project(img, f32, (int)nPix, (int)nDet,
                                reinterpret_cast<const double*>(d + o_a), (int)nangle, DSD, DS0,
                                pSize, dSize, partial, out, ctx->stream);
        finish(ctx, sino, out, 8 * nout, flags);
    });
}

int sgx_backproject_bdd(sgx_ctx* ctx, const double* sino, int64_t nPix, int64_t nDet,
                        const double* angle, int64_t nangle, double DSD, double DS0,
                        double pSize, double dSize, int ref_alias, double* image,
                        uint32_t flags)
{
    return guard(ctx, [&] {
        require(sino && angle && image, SGX_ERR_ARG, "null pointer argument");
        require(nPix > 0 && nDet > 1 && nangle > 0, SGX_ERR_ARG, "bad size");
        require(nPix * nPix < (int64_t(1) << 31), SGX_ERR_UNSUPPORTED, "image too large");
        TablePack pack;
        const size_t o_a = pack.add(angle, 8 * nangle);
        unsigned char* d = pack.upload(ctx, W_TABLES);
        const double* sn = static_cast<const double*>(
            stage_in(ctx, W_IN, sino, 8 * nangle * nDet, flags));
        const int64_t nout = nPix * nPix;
        double* out = (flags & SGX_HOST) ? static_cast<double*>(ctx->get(W_OUT, 8 * nout)) : image;
        void* work = ctx->get(W_PARTIAL, (size_t)sgx::bdd_back_work_bytes(
                                             (int)nPix, (int)nDet, (int)nangle, ref_alias != 0));
        sgx::launch_bdd_backproject(sn, (int)nPix, (int)nDet,
                                    reinterpret_cast<const double*>(d + o_a), (int)nangle, DSD,
                                    DS0, pSize, dSize, ref_alias != 0, out, work, ctx->stream);
        finish(ctx, image, out, 8 * nout, flags);
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------

struct sgx_plan {
    sgx_ctx* ctx = nullptr;
    PlanTune tune;
    bool tuned = false;
    sgx_geom geom{};
    int prec = SGX_PREC_FP32;
    int64_t nx = 0, ny = 0, nz = 1;
    int64_t npad = 0;
    void* tables = nullptr;       // device: cs, gam, bet, axes, g, wbin, wview, s/t axes
    uint8_t* mask = nullptr;      // device (or null)
    void* q = nullptr;            // filtered data (pitch, nt, na), column-major
    int64_t pitch = 0;            // elements per detector row of q (>= ns)
    int4* win = nullptr;          // cone-beam: per-(block, view) staged windows
    unsigned int* vbound = nullptr; // per-view max |weighted input| (F16S form)
    float* vscale = nullptr;        // per-view FP16 scale (F16S form)
    float g1 = 0.f;                 // sum |g| (F16S output bound)
    float wmax = 0.f;               // bound of |bin x cylinder weight| (F16S, FFT form)
    float gscale = 1.f;             // power of two with max|g| * gscale <= 2^14 (split FP16)
    float* partial = nullptr;
    int64_t partial_cap = 0;
    float rel_rmax = -1.f, rel_h = -1.f;  // arc fan: bounds of the tile-relative form
    // table offsets
    sgx::ViewCS* cs = nullptr;
    double* gam = nullptr;
    double* bet = nullptr;
    float *xds = nullptr, *yds = nullptr, *zds = nullptr, *g = nullptr, *wbin = nullptr,
          *wview = nullptr, *s_axis = nullptr, *t_axis = nullptr;
    float* wview_owned = nullptr;
    float* ptab = nullptr;          // short-scan Parker weights (ns, na), formed once in FP64
    float* fft = nullptr;           // FFT-form filter table (npad 512-2048), in `tables`
    int fft_n = 0;                  // its transform size
    float* wtab = nullptr;          // cone beam on the FFT form: (ns, nt) weights, in `tables`
    float* zs = nullptr;            // cone-beam option: per-view source z / ds (or null)
    bool arc_t_exact = false;       // cone-beam option: exact arc-detector t magnification
    float so_ds = 0.f;              // cone-beam option: source offset / ds
    bool z_sym = false;             // cone-beam: the z axis is symmetric about 0
    bool mirror_allowed = true;     // cone-beam: the autotuner may try the mirror form
    int win_zform = 0;              // the block form `win` was computed for
    int64_t win_cap = 0;            // entries `win` holds
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    bool timed = false;
    // cone beam: view chunks of an execute (SGX_OPT_VIEW_CHUNKS at creation),
    // and the side stream and events that run later chunks' filters beside
    // earlier chunks' backprojections (made on first use)
    int view_chunks = 1;
    hipStream_t side = nullptr;
    hipEvent_t pev[kMaxViewChunks] = {};
    // Bumped by every change that a captured batch graph would not see
    // (options, view weights, kernel shapes): batches recapture on a change.
    uint64_t gen = 0;

    ~sgx_plan()
    {
        for (hipEvent_t& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t& e : pev)
            if (e) (void)hipEventDestroy(e);
        if (side) (void)hipStreamDestroy(side);
        if (wview_owned) (void)hipFree(wview_owned);
        if (ptab) (void)hipFree(ptab);
        if (zs) (void)hipFree(zs);
        if (tables) (void)hipFree(tables);
        if (mask) (void)hipFree(mask);
        if (q) (void)hipFree(q);
        if (win) (void)hipFree(win);
        if (vbound) (void)hipFree(vbound);
        if (vscale) (void)hipFree(vscale);
        if (partial) (void)hipFree(partial);
    }
    bool is3d() const { return geom.kind == SGX_CT_FAN_ARC || geom.kind == SGX_CT_FAN_FLAT; }
    bool is_arc() const { return geom.kind == SGX_SINO_FAN_ARC || geom.kind == SGX_CT_FAN_ARC; }

    sgx::FilterArgs filter_args(const float* in, void* out) const
    {
        const sgx_geom& G = geom;
        sgx::FilterArgs a{};
        a.in = in;
        a.nb = (int)G.ns;
        a.nt = (int)G.nt;
        a.ncols = G.nt * G.na;
        a.g = g;
        a.npad = (int)npad;
        a.extra = (int)(pitch - G.ns);       // zero bins that pad each row to the pitch
        a.out = out;
        a.out_bf16 = prec == SGX_PREC_BF16;
        a.split_bf16 = prec == SGX_PREC_BF16;
        if (prec == SGX_PREC_F16S) {
            a.out_f16s = true;
            a.split_f16 = true;
            a.gscale = gscale;
            a.vbound = vbound;
            a.g1 = g1;
            a.wmax = fft ? wmax : 0.f;
            a.vscale = vscale;
        }
        a.scale = (float)G.angle_weight;
        a.wbin = wbin;
        a.wview = wview;
        if (is3d()) {
            a.cyl = is_arc() ? 2 : 1;
            a.cyl_dsd = (float)G.dsd;
            a.cyl_dso = (float)G.dso;
            a.s_axis = s_axis;
            a.t_axis = t_axis;
        }
        a.ptab = ptab;
        a.fft = fft;
        a.fft_n = fft_n;
        a.wtab = wtab;
        return a;
    }
};

sgx::CbctArgs cbct_args_of(const sgx_plan& P, float* img)
{
    const sgx_geom& G = P.geom;
    sgx::CbctArgs a{};
    a.q = P.q;
    a.fmt = P.prec == SGX_PREC_BF16   ? sgx::SGX_Q_FMT_BF16
            : P.prec == SGX_PREC_F16S ? sgx::SGX_Q_FMT_F16S
                                      : sgx::SGX_Q_FMT_F32;
    a.vscale = P.vscale;
    a.ns = (int)G.ns;
    a.nt = (int)G.nt;
    a.pitch = (int)P.pitch;
    a.nviews = (int)G.na;
    a.view_stride = P.pitch * G.nt;
    a.cs = P.cs;
    a.xds = P.xds;
    a.yds = P.yds;
    a.zds = P.zds;
    a.nx = (int)P.nx;
    a.ny = (int)P.ny;
    a.nz = (int)P.nz;
    a.mask = P.mask;
    a.image = img;
    a.dsd_ds = (float)G.dsd / (float)G.ds;
    a.dso_ds = (float)G.dso / (float)G.ds;
    a.ws = (float)((G.ns + 1) / 2.0 + (double)(float)G.offset_s);
    a.wt = (float)((G.nt + 1) / 2.0 + (double)(float)G.offset_t);
    a.inv_dt_ds = 1.0f / ((float)G.dt / (float)G.ds);
    a.is_arc = P.is_arc();
    a.win = P.win;
    a.zs = P.zs;
    a.arc_t_exact = P.arc_t_exact;
    a.so_ds = P.so_ds;
    a.f16_packed = P.tune.cbct == 2;
    a.zform = P.tune.zform == 1 ? (sgx::cbct_mirror_ok(a, P.z_sym) ? 1 : 0) : P.tune.zform;
    return a;
}

// The staged backprojector's window table of a cone-beam plan, for its
// current form (the mirror form's blocks have their own windows).
void plan_windows(sgx_plan& P, hipStream_t st)
{
    const sgx::CbctArgs a = cbct_args_of(P, nullptr);
    const int64_t n = sgx::cbct_window_entries(a);
    if (n > P.win_cap) {
        // a form with more blocks (Z8): the old table may still be in use
        SGX_HIP(hipStreamSynchronize(st));
        if (P.win) SGX_HIP(hipFree(P.win));
        P.win = nullptr;
        P.win_cap = 0;
        SGX_HIP(hipMalloc(&P.win, sizeof(int4) * n));
        P.win_cap = n;
    }
    sgx::launch_cbct_windows(a, P.win, st);
    P.win_zform = a.zform;
}

// Recompute the window table when the plan's form changed since (a tuning
// choice): a no-op otherwise.
void ensure_windows(sgx_plan& P, hipStream_t st)
{
    if (P.is3d() && P.win && cbct_args_of(P, nullptr).zform != P.win_zform)
        plan_windows(P, st);
}

// One plan execute on stream `st` (device pointers): weights + filter, then
// backprojection, with the plan's kernel shapes.  `timed` brackets the two
// phases with the plan's events (left out under graph capture).
// The weights + filter phase of a plan execute (into the plan's q).
void plan_filter_phase(sgx_plan& P, const float* in, hipStream_t st)
{
    const sgx::FilterArgs fa = P.filter_args(in, P.q);
    if (P.prec == SGX_PREC_F16S) sgx::launch_filter_bound(fa, P.vbound, st);
    sgx::launch_sino_filter(fa, st);
}

// Backprojection arguments of a 2-D plan execute into img.
sgx::Back2dArgs back2d_run_args(const sgx_plan& P, float* img)
{
    const sgx_geom& G = P.geom;
    sgx::Back2dArgs a{};
    a.q = static_cast<const float*>(P.q);
    a.nb = (int)G.ns;
    a.nviews = (int)G.na;
    a.view_stride = G.ns;
    a.cs = P.cs;
    a.xds = P.xds;
    a.yds = P.yds;
    a.nx = (int)P.nx;
    a.ny = (int)P.ny;
    a.mask = P.mask;
    a.image = img;
    a.geom = (G.kind == SGX_SINO_PAR || G.kind == SGX_SINO_MOJ) ? 0
                                                              : (G.kind == SGX_SINO_FAN_ARC ? 1 : 2);
    a.wb = (float)((G.ns + 1) / 2.0 + (double)(float)G.offset_s);
    a.dsd_ds = (float)G.dsd / (float)G.ds;
    a.dso_ds = (float)G.dso / (float)G.ds;
    a.so_ds = (float)G.source_offset / (float)G.ds;
    a.partial = P.partial;
    a.partial_cap = P.partial_cap;
    a.want_blocks = P.tune.want_blocks;
    a.vc_max = P.tune.vc_max;
    a.rel_rmax = P.rel_rmax;
    a.rel_h = P.rel_h;
    return a;
}

// The views [v0, v1) of a filter launch (pointers offset to view v0).
sgx::FilterArgs filter_chunk(const sgx_plan& P, sgx::FilterArgs f, int64_t v0, int64_t v1)
{
    const int64_t nt = P.geom.nt, nb = P.geom.ns;
    const size_t esz = P.prec == SGX_PREC_FP32 ? 4 : 2;
    f.in += v0 * nt * nb;
    f.out = static_cast<char*>(f.out) + v0 * nt * (int64_t)P.pitch * esz;
    f.ncols = nt * (v1 - v0);
    if (f.wview) f.wview += v0;
    if (f.ptab) f.ptab += v0 * nb;
    if (f.vscale) f.vscale += v0;
    if (f.vbound) f.vbound += v0;
    return f;
}

// Number of view chunks of a cone-beam execute: the plan's setting, for the
// staged backprojector's FP32-interpolation forms, at least kMinChunkViews
// views a chunk; 1 otherwise.
int view_chunks_of(const sgx_plan& P, const sgx::CbctArgs& a)
{
    if (!P.is3d() || P.tune.cbct != 0 || !a.win || a.f16_packed || P.view_chunks <= 1) return 1;
    return (int)std::max<int64_t>(1, std::min<int64_t>(P.view_chunks, P.geom.na / kMinChunkViews));
}

// Cone-beam execute in view chunks: chunk k's views are filtered, then
// backprojected, the first backprojection storing and the others adding to
// the image (the same sums in the same order whatever runs beside them).
// Outside stream capture, the filters of chunks 1.. run on the plan's side
// stream, each backprojection waiting for its own chunk's filter, so that
// they overlap the backprojections of earlier chunks; the side stream starts
// after the work queued on `st` before the call (which includes earlier
// executes' reads of the plan's buffers) and `st` has joined it at the end.
// Under capture every launch goes on `st` in chunk order.
void plan_run_chunked(sgx_plan& P, const float* in, float* img, hipStream_t st, bool timed,
                      int K)
{
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    SGX_HIP(hipStreamIsCapturing(st, &cap));
    const bool overlap = cap == hipStreamCaptureStatusNone;
    if (overlap && !P.side) {
        SGX_HIP(hipStreamCreateWithFlags(&P.side, hipStreamNonBlocking));
        for (hipEvent_t& e : P.pev) SGX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    const int64_t na = P.geom.na;
    auto v_at = [&](int k) { return na * k / K; };
    const sgx::FilterArgs fa = P.filter_args(in, P.q);
    auto filter = [&](int k, hipStream_t s) {
        const sgx::FilterArgs f = filter_chunk(P, fa, v_at(k), v_at(k + 1));
        if (P.prec == SGX_PREC_F16S) sgx::launch_filter_bound(f, P.vbound + v_at(k), s);
        sgx::launch_sino_filter(f, s);
    };
    if (timed) SGX_HIP(hipEventRecord(P.ev[0], st));
    if (overlap) {
        SGX_HIP(hipEventRecord(P.pev[0], st));
        SGX_HIP(hipStreamWaitEvent(P.side, P.pev[0], 0));
    }
    filter(0, st);
    if (timed) SGX_HIP(hipEventRecord(P.ev[1], st));
    if (overlap) {
        for (int k = 1; k < K; ++k) {
            filter(k, P.side);
            SGX_HIP(hipEventRecord(P.pev[k], P.side));
        }
    }
    const sgx::CbctArgs a = cbct_args_of(P, img);
    const size_t esz = P.prec == SGX_PREC_FP32 ? 4 : 2;
    for (int k = 0; k < K; ++k) {
        if (k > 0) {
            if (overlap) SGX_HIP(hipStreamWaitEvent(st, P.pev[k], 0));
            else filter(k, st);
        }
        const int64_t v0 = v_at(k);
        sgx::CbctArgs c = a;
        c.q = static_cast<const char*>(a.q) + v0 * a.view_stride * (int64_t)esz;
        c.cs += v0;
        if (c.vscale) c.vscale += v0;
        if (c.zs) c.zs += v0;
        c.win += v0;
        c.win_stride = (int)na;
        c.nviews = (int)(v_at(k + 1) - v0);
        c.accumulate = k > 0;
        sgx::launch_cbct_back(c, st);
    }
    if (timed) {
        SGX_HIP(hipEventRecord(P.ev[2], st));
        P.timed = true;
    }
}

void plan_run(sgx_plan& P, const float* in, float* img, hipStream_t st, bool timed)
{
    if (P.is3d()) {
        ensure_windows(P, st);
        const int K = view_chunks_of(P, cbct_args_of(P, img));
        if (K > 1) {
            plan_run_chunked(P, in, img, st, timed, K);
            return;
        }
    }
    if (timed) SGX_HIP(hipEventRecord(P.ev[0], st));
    plan_filter_phase(P, in, st);
    if (timed) SGX_HIP(hipEventRecord(P.ev[1], st));
    if (P.is3d()) {
        ensure_windows(P, st);
        sgx::CbctArgs a = cbct_args_of(P, img);
        if (P.tune.cbct == 1) a.win = nullptr;
        sgx::launch_cbct_back(a, st);
    } else {
        sgx::launch_back2d(back2d_run_args(P, img), st);
    }
    if (timed) {
        SGX_HIP(hipEventRecord(P.ev[2], st));
        P.timed = true;
    }
}

// Candidate shapes of the autotuner.
std::vector<PlanTune> tune_candidates(const sgx_plan& P)
{
    std::vector<PlanTune> c;
    if (P.is3d()) {
        for (int k : {0, 1}) c.push_back(PlanTune{k, 2048, 64});
        sgx_plan& Q = const_cast<sgx_plan&>(P);
        const int m = Q.tune.zform;
        Q.tune.zform = 1;
        const bool can_mirror = P.mirror_allowed && cbct_args_of(Q, nullptr).zform == 1;
        Q.tune.zform = m;
        if (can_mirror) c.push_back(PlanTune{0, 2048, 64, 1});
        c.push_back(PlanTune{0, 2048, 64, 2});
        c.push_back(PlanTune{0, 2048, 64, 3});
        if (P.prec == SGX_PREC_F16S) c.push_back(PlanTune{2, 2048, 64});
    } else {
        for (int want : {1024, 2048, kTuneMaxBlocks})
            for (int vc : {kTuneMinChunk, 32, 64}) c.push_back(PlanTune{0, want, vc});
    }
    return c;
}

uint64_t ceil_log2(int64_t v)
{
    uint64_t k = 0;
    while ((int64_t(1) << k) < v) ++k;
    return k;
}

// Size bucket of a plan: kind, precision and the power-of-two ceilings of
// the image and detector sizes and the view count.
std::string tune_bucket(const sgx_plan& P)
{
    const sgx_geom& G = P.geom;
    std::ostringstream os;
    os << G.kind << '/' << P.prec << '/' << ceil_log2(P.nx) << '/' << ceil_log2(P.ny) << '/'
       << ceil_log2(P.nz) << '/' << ceil_log2(G.ns) << '/' << ceil_log2(G.nt) << '/'
       << ceil_log2(G.na);
    return os.str();
}

// Time every candidate shape on (in, img) and keep the fastest in the plan
// and in the context's cache for the plan's size bucket.
void autotune(sgx_plan& P, const float* in, float* img, int reps, bool use_cache)
{
    sgx_ctx* ctx = P.ctx;
    const std::string key = tune_bucket(P);
    if (use_cache) {
        auto it = ctx->tune_cache.find(key);
        if (it != ctx->tune_cache.end()) {
            P.tune = it->second;
            P.tuned = true;
            ++P.gen;
            plan_run(P, in, img, ctx->stream, false);
            return;
        }
    }
    hipStream_t st = ctx->stream;
    hipEvent_t e0, e1;
    SGX_HIP(hipEventCreate(&e0));
    SGX_HIP(hipEventCreate(&e1));
    float best = 3.4e38f;
    PlanTune best_t = P.tune;
    try {
        for (const PlanTune& t : tune_candidates(P)) {
            P.tune = t;
            plan_run(P, in, img, st, false);                      // warm
            SGX_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < reps; ++r) plan_run(P, in, img, st, false);
            SGX_HIP(hipEventRecord(e1, st));
            SGX_HIP(hipEventSynchronize(e1));
            float ms = 0;
            SGX_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                best_t = t;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    P.tune = best_t;
    P.tuned = true;
    ++P.gen;
    ctx->tune_cache[key] = best_t;
    plan_run(P, in, img, st, false);          // leave the chosen shape's result in img
}

extern "C" {

int sgx_plan_create(sgx_ctx* ctx, const sgx_geom* geom, const double* betas,
                    const double* xc, int64_t nx, const double* yc, int64_t ny,
                    const double* zc, int64_t nz, const uint8_t* mask,
                    const float* Hk, int64_t npad, int prec, sgx_plan** out)
{
    return guard(ctx, [&] {
        require(geom && betas && xc && yc && Hk && out, SGX_ERR_ARG, "null pointer argument");
        *out = nullptr;
        const sgx_geom& G = *geom;
        require(G.kind >= SGX_SINO_PAR && G.kind <= SGX_SINO_MOJ, SGX_ERR_ARG, "bad kind");
        const bool is3d = G.kind == SGX_CT_FAN_ARC || G.kind == SGX_CT_FAN_FLAT;
        const bool moj = G.kind == SGX_SINO_MOJ;
        require(!moj || !G.parker, SGX_ERR_ARG, "no Parker weighting for a mojette geometry");
        require(G.ns > 1 && G.na > 0 && nx > 0 && ny > 0, SGX_ERR_ARG, "bad size");
        require(!is3d || (zc && nz > 0 && G.nt > 1), SGX_ERR_ARG, "3-D plan needs zc, nz, nt");
        require(is3d || G.nt == 1, SGX_ERR_ARG, "2-D plan needs nt == 1");
        require(prec == SGX_PREC_FP32 || prec == SGX_PREC_BF16 || prec == SGX_PREC_F16S,
                SGX_ERR_ARG, "bad precision");
        require(prec == SGX_PREC_FP32 || is3d, SGX_ERR_UNSUPPORTED,
                "the 16-bit storage forms are provided for cone-beam plans");
        require(G.ds != 0.0 && (!is3d || G.dt != 0.0), SGX_ERR_ARG, "ds / dt must be nonzero");
        if (G.ns > npad) {
            std::ostringstream os;
            os << "nb=" << G.ns << " + extra=0 > npad=" << npad;
            throw Error(SGX_ERR_VALUE, os.str());
        }
        auto P = std::make_unique<sgx_plan>();
        P->ctx = ctx;
        P->geom = G;
        P->prec = prec;
        P->nx = nx;
        P->ny = ny;
        P->nz = is3d ? nz : 1;
        P->npad = npad;
        P->tune.cbct = ctx->cbct_kernel >= 3 ? 0 : ctx->cbct_kernel;
        // the mirror form: by default for the scaled-FP16 form, where it
        // measured faster; the autotuner times it for the others
        P->mirror_allowed = ctx->cbct_mirror != 0;
        P->view_chunks = ctx->view_chunks;
        P->tune.zform =
            ctx->cbct_mirror == 2 || (ctx->cbct_mirror == 1 && prec == SGX_PREC_F16S) ? 1
            : ctx->cbct_kernel == 3                                                     ? 2
            : ctx->cbct_kernel == 4                                                     ? 3
                                                                                        : 0;

        const int64_t ns = G.ns, nt = G.nt, na = G.na;
        TablePack pack;
        std::vector<double> b(betas, betas + na);
        const size_t o_b = pack.add(b.data(), na * 8);
        const size_t o_cs = pack.add(b.data(), na * sizeof(sgx::ViewCS));
        const auto xv = axis_over(xc, nx, G.ds), yv = axis_over(yc, ny, G.ds);
        const size_t o_x = pack.add(xv.data(), nx * 4);
        const size_t o_y = pack.add(yv.data(), ny * 4);
        size_t o_z = 0;
        if (is3d) {
            const auto zv = axis_over(zc, nz, G.ds);
            o_z = pack.add(zv.data(), nz * 4);
            P->z_sym = axis_symmetric(zv);
        }
        const auto gd = filter_kernel_g(Hk, npad);
        std::vector<float> gf(gd.begin(), gd.end());
        const size_t o_g = pack.add(gf.data(), npad * 4);
        double g1 = 0.0;
        for (float v : gf) g1 += std::fabs((double)v);
        P->g1 = (float)g1;   // |q| <= g1 max|input|; FP16 keeps 2x headroom above the bound
        {
            float gmax = 0.f;
            for (float v : gf) gmax = std::max(gmax, std::fabs(v));
            int e = 0;
            if (gmax > 0.f) (void)std::frexp(gmax, &e);       // gmax <= 2^e
            P->gscale = std::ldexp(1.f, 14 - e);
        }
        // detector axis s (and t), fan angles gamma
        std::vector<double> s(ns), gamv(ns);
        for (int64_t i = 0; i < ns; ++i)
            s[i] = G.ds * ((double)i - ((ns - 1) / 2.0 + (double)(float)G.offset_s));
        for (int64_t i = 0; i < ns; ++i) gamv[i] = (G.kind == SGX_SINO_FAN_ARC || G.kind == SGX_CT_FAN_ARC)
                                                       ? s[i] / G.dsd : std::atan(s[i] / G.dsd);
        size_t o_wbin = 0, o_gam = 0, o_bet = 0, o_s = 0, o_t = 0;
        std::vector<float> wbin_host;
        if (G.kind == SGX_SINO_FAN_ARC || G.kind == SGX_SINO_FAN_FLAT) {
            // fbp_sino_weight (src/fbp2/sino-weight.jl:12-16)
            std::vector<float>& w = wbin_host;
            w.resize(ns);
            for (int64_t i = 0; i < ns; ++i)
                w[i] = (float)(std::fabs(G.dso * std::cos(gamv[i]) -
                                         G.source_offset * std::sin(gamv[i])) / G.dsd);
            o_wbin = pack.add(w.data(), ns * 4);
        }
        if (G.parker) {
            require(G.kind != SGX_SINO_PAR && G.kind != SGX_SINO_MOJ, SGX_ERR_ARG,
                    "parallel-beam Parker weights are passed per view");
            double amin = betas[0];
            for (int64_t i = 1; i < na; ++i) amin = std::min(amin, betas[i]);
            std::vector<double> bet(na);
            for (int64_t i = 0; i < na; ++i) bet[i] = std::max(betas[i] - amin, 0.0);
            o_gam = pack.add(gamv.data(), ns * 8);
            o_bet = pack.add(bet.data(), na * 8);
        }
        if (is3d) {
            std::vector<float> sf(ns), tf(nt);
            for (int64_t i = 0; i < ns; ++i) sf[i] = (float)s[i];
            for (int64_t i = 0; i < nt; ++i)
                tf[i] = (float)(G.dt * ((double)i - ((nt - 1) / 2.0 + (double)(float)G.offset_t)));
            o_s = pack.add(sf.data(), ns * 4);
            o_t = pack.add(tf.data(), nt * 4);
        }
        // FFT-form filter tables at npad = 1024 (with the per-bin weight factors)
        const auto ft = sgx::fft_filter_tables(
            Hk, npad, ns, wbin_host.empty() ? nullptr : wbin_host.data(), s.data(),
            is3d ? ((G.kind == SGX_CT_FAN_ARC) ? 2 : 1) : 0, G.dsd);
        const size_t o_fft = ft.empty() ? 0 : pack.add(ft.data(), ft.size() * 4);
        // cone beam on the FFT form: wbin x cylinder weight per (bin, row), in
        // FP64 then rounded (fdk_weight_cyl, src/fbp3/fdk.jl:21-45)
        size_t o_wtab = 0;
        double wtmax = 0.0;
        if (!ft.empty() && is3d) {
            std::vector<float> wt((size_t)ns * nt);
            const bool arc = G.kind == SGX_CT_FAN_ARC;
            for (int64_t it = 0; it < nt; ++it) {
                const double tt = G.dt * ((double)it - ((nt - 1) / 2.0 + (double)(float)G.offset_t));
                for (int64_t i = 0; i < ns; ++i) {
                    const double w = arc ? (G.dso / G.dsd) * std::cos(s[i] / G.dsd) /
                                               std::sqrt((tt / G.dsd) * (tt / G.dsd) + 1.0)
                                         : G.dso / std::sqrt(s[i] * s[i] + tt * tt + G.dsd * G.dsd);
                    wt[it * ns + i] = (float)((wbin_host.empty() ? 1.0 : (double)wbin_host[i]) * w);
                    wtmax = std::max(wtmax, std::fabs((double)wt[it * ns + i]));
                }
            }
            // the streaming range pass multiplies max |input| by this bound
            // (rounded up: every table entry is at most wtmax)
            P->wmax = (float)(wtmax * (1.0 + 1e-6));
            o_wtab = pack.add(wt.data(), wt.size() * 4);
        }
        SGX_HIP(hipMalloc(&P->tables, pack.host.size()));
        SGX_HIP(hipMemcpy(P->tables, pack.host.data(), pack.host.size(), hipMemcpyHostToDevice));
        auto* d = static_cast<unsigned char*>(P->tables);
        P->cs = reinterpret_cast<sgx::ViewCS*>(d + o_cs);
        P->xds = reinterpret_cast<float*>(d + o_x);
        P->yds = reinterpret_cast<float*>(d + o_y);
        if (is3d) P->zds = reinterpret_cast<float*>(d + o_z);
        P->g = reinterpret_cast<float*>(d + o_g);
        if (!ft.empty()) {
            P->fft = reinterpret_cast<float*>(d + o_fft);
            P->fft_n = (int)(ft.size() / 5);
            if (o_wtab) P->wtab = reinterpret_cast<float*>(d + o_wtab);
        }
        if (o_wbin) P->wbin = reinterpret_cast<float*>(d + o_wbin);
        if (G.parker) {
            P->gam = reinterpret_cast<double*>(d + o_gam);
            P->bet = reinterpret_cast<double*>(d + o_bet);
        }
        if (is3d) {
            P->s_axis = reinterpret_cast<float*>(d + o_s);
            P->t_axis = reinterpret_cast<float*>(d + o_t);
        }
        sgx::launch_view_trig(reinterpret_cast<const double*>(d + o_b), na, 1, P->cs,
                              ctx->stream, moj);
        if (moj) {
            // the filter for the view's radial spacing d_moj = d f is Hk(d) / f
            // (the ramp scales as 1 / spacing): a per-view weight 1 / f
            std::vector<float> w(na);
            for (int64_t i = 0; i < na; ++i)
                w[i] = (float)(1.0 / std::max(std::fabs(std::cos(betas[i])),
                                              std::fabs(std::sin(betas[i]))));
            SGX_HIP(hipMalloc(&P->wview_owned, 4 * na));
            SGX_HIP(hipMemcpy(P->wview_owned, w.data(), 4 * na, hipMemcpyHostToDevice));
            P->wview = P->wview_owned;
        }
        if (G.parker) {
            // parker_weight_fan_short (parker.jl:96-106) in FP64, cast to Float32
            // once per plan; the filter multiplies it into its operand load
            SGX_HIP(hipMalloc(&P->ptab, 4 * ns * na));
            sgx::launch_parker(P->gam, P->bet, (int)ns, (int)na, G.gamma_max, P->ptab,
                               ctx->stream);
        }
        const int64_t nvox = nx * ny * P->nz;
        if (mask) {
            SGX_HIP(hipMalloc(&P->mask, nvox));
            SGX_HIP(hipMemcpy(P->mask, mask, nvox, hipMemcpyHostToDevice));
        }
        // filtered data: column-major; cone-beam rows padded to a multiple of
        // 4 bins (aligned 16-byte staging loads), padding bins written as 0
        P->pitch = is3d ? (ns + 3) / 4 * 4 : ns;
        const size_t esz = prec == SGX_PREC_FP32 ? 4 : 2;
        const size_t qbytes = esz * P->pitch * nt * na + sgx::kQ16Pad;
        SGX_HIP(hipMalloc(&P->q, qbytes));
        SGX_HIP(hipMemset(P->q, 0, qbytes));
        if (prec == SGX_PREC_F16S) {
            SGX_HIP(hipMalloc(&P->vbound, 4 * na));
            SGX_HIP(hipMalloc(&P->vscale, 4 * na));
        }
        if (!is3d) {
            // sized for the largest shape the autotuner tries
            P->partial_cap = sgx::back2d_partial_floats((int)ns, (int)na, (int)nx, (int)ny,
                                                        kTuneMaxBlocks, kTuneMinChunk);
            if (P->partial_cap) SGX_HIP(hipMalloc(&P->partial, 4 * P->partial_cap));
            if (G.kind == SGX_SINO_FAN_ARC)
                sgx::back2d_rel_bounds(xc, nx, yc, ny, mask, G.ds, P->rel_rmax, P->rel_h);
        }
        if (is3d) {
            // the staged backprojector's detector windows, fixed by the geometry
            plan_windows(*P, ctx->stream);
        }
        for (hipEvent_t& e : P->ev) SGX_HIP(hipEventCreate(&e));
        SGX_HIP(hipStreamSynchronize(ctx->stream));
        *out = P.release();
    });
}

void sgx_plan_destroy(sgx_plan* plan)
{
    if (!plan) return;
    (void)hipSetDevice(plan->ctx->device);
    (void)hipStreamSynchronize(plan->ctx->stream);
    delete plan;
}

int sgx_plan_set_view_weight(sgx_plan* plan, const float* wview)
{
    if (!plan) return SGX_ERR_ARG;
    return guard(plan->ctx, [&] {
        sgx_plan& P = *plan;
        const int64_t na = P.geom.na;
        require(P.geom.kind != SGX_SINO_MOJ, SGX_ERR_ARG,
                "a mojette plan's per-view weight is its radial-spacing factor");
        ++P.gen;
        if (!wview) {
            P.wview = nullptr;
            return;
        }
        if (!P.wview_owned) {
            void* p = nullptr;
            SGX_HIP(hipMalloc(&p, 4 * na));
            P.wview_owned = static_cast<float*>(p);
        }
        SGX_HIP(hipStreamSynchronize(plan->ctx->stream));
        SGX_HIP(hipMemcpy(P.wview_owned, wview, 4 * na, hipMemcpyHostToDevice));
        P.wview = P.wview_owned;
    });
}

int sgx_plan_set_cbct_opts(sgx_plan* plan, const sgx_cbct_opts* opts)
{
    if (!plan) return SGX_ERR_ARG;
    return guard(plan->ctx, [&] {
        sgx_plan& P = *plan;
        require(P.is3d(), SGX_ERR_ARG, "cone-beam options need a cone-beam plan");
        const int64_t na = P.geom.na;
        SGX_HIP(hipStreamSynchronize(plan->ctx->stream));
        if (opts && opts->source_zs) {
            std::vector<float> z(na);
            for (int64_t i = 0; i < na; ++i)
                z[i] = (float)((float)opts->source_zs[i] / (float)P.geom.ds);
            if (!P.zs) {
                void* p = nullptr;
                SGX_HIP(hipMalloc(&p, 4 * na));
                P.zs = static_cast<float*>(p);
            }
            SGX_HIP(hipMemcpy(P.zs, z.data(), 4 * na, hipMemcpyHostToDevice));
        } else if (P.zs) {
            SGX_HIP(hipFree(P.zs));
            P.zs = nullptr;
        }
        P.arc_t_exact = opts && opts->arc_t_exact;
        P.so_ds = opts ? (float)opts->source_offset / (float)P.geom.ds : 0.f;
        ++P.gen;
        // the staged backprojector's windows follow the t mapping
        plan_windows(P, plan->ctx->stream);
        SGX_HIP(hipStreamSynchronize(plan->ctx->stream));
    });
}

int sgx_plan_execute(sgx_plan* plan, const float* sino, float* image, uint32_t flags)
{
    if (!plan) return SGX_ERR_ARG;
    sgx_ctx* ctx = plan->ctx;
    return guard(ctx, [&] {
        require(sino && image, SGX_ERR_ARG, "null pointer argument");
        sgx_plan& P = *plan;
        const sgx_geom& G = P.geom;
        const int64_t nin = G.ns * G.nt * G.na, nvox = P.nx * P.ny * P.nz;
        const float* in = static_cast<const float*>(stage_in(ctx, W_IN, sino, 4 * nin, flags));
        float* img = (flags & SGX_HOST) ? static_cast<float*>(ctx->get(W_OUT, 4 * nvox)) : image;
        plan_run(P, in, img, ctx->stream, true);
        finish(ctx, image, img, 4 * nvox, flags);
    });
}

const void* sgx_plan_filtered(const sgx_plan* plan) { return plan ? plan->q : nullptr; }

int64_t sgx_plan_pitch(const sgx_plan* plan) { return plan ? plan->pitch : 0; }

int sgx_plan_read_filtered(sgx_plan* plan, void* host, float* view_scales)
{
    if (!plan) return SGX_ERR_ARG;
    sgx_ctx* ctx = plan->ctx;
    return guard(ctx, [&] {
        require(host, SGX_ERR_ARG, "null pointer argument");
        const sgx_geom& G = plan->geom;
        const size_t esz = plan->prec == SGX_PREC_FP32 ? 4 : 2;
        SGX_HIP(hipMemcpy2DAsync(host, esz * G.ns, plan->q, esz * plan->pitch, esz * G.ns,
                                 G.nt * G.na, hipMemcpyDeviceToHost, ctx->stream));
        if (view_scales) {
            if (plan->prec == SGX_PREC_F16S) {
                SGX_HIP(hipMemcpyAsync(view_scales, plan->vscale, 4 * G.na,
                                       hipMemcpyDeviceToHost, ctx->stream));
            } else {
                for (int64_t i = 0; i < G.na; ++i) view_scales[i] = 1.0f;
            }
        }
        SGX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int64_t sgx_plan_stored_bytes(const sgx_plan* plan)
{
    if (!plan) return 0;
    const sgx_geom& G = plan->geom;
    const int64_t esz = plan->prec == SGX_PREC_FP32 ? 4 : 2;
    return esz * plan->pitch * G.nt * G.na;
}

int sgx_plan_read_stored(sgx_plan* plan, void* host, int64_t bytes)
{
    if (!plan) return SGX_ERR_ARG;
    sgx_ctx* ctx = plan->ctx;
    return guard(ctx, [&] {
        require(host, SGX_ERR_ARG, "null pointer argument");
        require(bytes == sgx_plan_stored_bytes(plan), SGX_ERR_SIZE, "size mismatch");
        SGX_HIP(hipMemcpyAsync(host, plan->q, bytes, hipMemcpyDeviceToHost, ctx->stream));
        SGX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int sgx_plan_view_weight(sgx_plan* plan, float* out)
{
    if (!plan) return SGX_ERR_ARG;
    sgx_ctx* ctx = plan->ctx;
    return guard(ctx, [&] {
        require(out, SGX_ERR_ARG, "null pointer argument");
        const sgx_geom& G = plan->geom;
        const int64_t n = G.ns * G.nt * G.na;
        float* d = static_cast<float*>(ctx->get(W_OUT, 4 * n));
        sgx::launch_filter_weights(plan->filter_args(nullptr, nullptr), d, ctx->stream);
        SGX_HIP(hipMemcpyAsync(out, d, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
        SGX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int sgx_plan_last_times(const sgx_plan* plan, float* filter_ms, float* back_ms)
{
    if (!plan) return SGX_ERR_ARG;
    return guard(plan->ctx, [&] {
        require(plan->timed, SGX_ERR_ARG, "no execute recorded yet");
        SGX_HIP(hipEventSynchronize(plan->ev[2]));
        float a = 0, b = 0;
        SGX_HIP(hipEventElapsedTime(&a, plan->ev[0], plan->ev[1]));
        SGX_HIP(hipEventElapsedTime(&b, plan->ev[1], plan->ev[2]));
        if (filter_ms) *filter_ms = a;
        if (back_ms) *back_ms = b;
    });
}

int sgx_plan_autotune(sgx_plan* plan, const float* sino, float* image, int reps, int use_cache)
{
    if (!plan) return SGX_ERR_ARG;
    return guard(plan->ctx, [&] {
        require(sino && image, SGX_ERR_ARG, "null pointer argument");
        require(reps >= 1, SGX_ERR_ARG, "reps must be >= 1");
        autotune(*plan, sino, image, reps, use_cache != 0);
        SGX_HIP(hipStreamSynchronize(plan->ctx->stream));
    });
}

int sgx_plan_tuning(const sgx_plan* plan, int64_t values[4])
{
    if (!plan || !values) return SGX_ERR_ARG;
    values[0] = plan->tune.cbct != 0 ? plan->tune.cbct
                : plan->win_zform == 1 ? 3 : plan->win_zform == 2 ? 4
                : plan->win_zform == 3 ? 5 : 0;
    values[1] = plan->tune.want_blocks;
    values[2] = plan->tune.vc_max;
    values[3] = plan->tuned ? 1 : 0;
    return SGX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------

struct sgx_batch {
    struct Item {
        sgx_plan* plan;
        const float* in;
        float* out;
    };
    // A backprojection queue: kind 2 (2-D) or 3 (cone beam), the kernel
    // variant, and its items and task table in the batch's device buffer.
    struct Queue {
        int kind = 0, variant = 0;
        std::vector<int> items;              // batch item indices, queue order
        size_t items_off = 0, tasks_off = 0;
        int ntasks = 0;
        size_t lds = 0;
    };
    sgx_ctx* ctx = nullptr;
    std::vector<Item> items;
    std::vector<uint64_t> gens;              // each item's plan generation at capture
    std::vector<hipStream_t> branches;       // capture streams (default priority)
    std::vector<hipEvent_t> joins;
    std::vector<hipEvent_t> filtered;        // per (queue, branch): that queue's items on
                                             // the branch are filtered
    hipStream_t origin = nullptr;            // capture origin
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;
    bool launched = false;
    bool queued = true;                      // the captured graph uses queues
    int mode = -1;                           // requested: 0 no queues, 1 queues, 2 timed choice
    float ms_queues = -1.f, ms_launches = -1.f;   // the timed choice's replays
    int64_t nbranch = 0;
    std::vector<Queue> queues;
    bool gather_queues = false;              // queue gather-tuned cone-beam items too
    void* qbuf = nullptr;                    // queue items and task tables
    size_t qcap = 0;

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~sgx_batch()
    {
        drop_graph();
        for (hipEvent_t e : joins) (void)hipEventDestroy(e);
        for (hipEvent_t e : filtered) (void)hipEventDestroy(e);
        for (hipStream_t s : branches) (void)hipStreamDestroy(s);
        for (hipEvent_t e : {fork, t0, t1})
            if (e) (void)hipEventDestroy(e);
        if (origin) (void)hipStreamDestroy(origin);
        if (qbuf) (void)hipFree(qbuf);
    }

    void ensure_streams(size_t n)
    {
        while (branches.size() < n) {
            hipStream_t s;
            SGX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            branches.push_back(s);
            hipEvent_t e;
            SGX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            joins.push_back(e);
        }
    }
    void ensure_filter_events(size_t n)
    {
        while (filtered.size() < n) {
            hipEvent_t f;
            SGX_HIP(hipEventCreateWithFlags(&f, hipEventDisableTiming));
            filtered.push_back(f);
        }
    }

    // Sort the items into backprojection queues (use_queues) or
    // none.  An item goes to a queue when its plan occurs once in the batch
    // (the filter phases of all queued items run before any queue, and a
    // plan's items share its filtered buffer); the others run as whole plan
    // executes.  A queue orders its items by the cost of one task
    // (views), largest first, and each item's tasks in block order.
    void build_queues(bool use_queues)
    {
        queues.clear();
        if (!use_queues) return;
        std::map<sgx_plan*, int> count;
        for (const Item& it : items) count[it.plan] += 1;
        std::map<std::pair<int, int>, size_t> qindex;
        for (int i = 0; i < (int)items.size(); ++i) {
            const Item& it = items[i];
            sgx_plan& P = *it.plan;
            if (count[it.plan] > 1) continue;
            int kind, variant;
            if (P.is3d()) {
                // gather-tuned items (coarse voxels, the LDS windows miss)
                // stay whole executes: their filter and backprojection overlap
                // the other branches' work better than a queue that waits
                // for every filter of its chunk
                if (P.tune.cbct == 1 && !gather_queues) continue;
                sgx::CbctArgs a = cbct_args_of(P, it.out);
                if (P.tune.cbct == 1) a.win = nullptr;
                kind = 3;
                variant = sgx::cbct_variant(a);
            } else {
                sgx::Back2dArgs a = back2d_run_args(P, it.out);
                a.vc_max = 64;
                kind = 2;
                variant = sgx::back2d_variant(a);
            }
            auto key = std::make_pair(kind, variant);
            if (!qindex.count(key)) {
                qindex[key] = queues.size();
                queues.push_back(Queue{});
                queues.back().kind = kind;
                queues.back().variant = variant;
            }
            queues[qindex[key]].items.push_back(i);
        }
        // host images of the device buffer: per queue items then tasks, then
        // one counter per queue
        std::vector<unsigned char> host;
        auto put = [&](const void* p, size_t n) {
            const size_t off = (host.size() + 15) & ~size_t(15);
            host.resize(off + n);
            std::memcpy(host.data() + off, p, n);
            return off;
        };
        // Items ordered by task cost (views), largest first; then each
        // variant's items are cut into chunks of kChunkItems, one queue
        // launch each, so that a chunk's backprojection overlaps the filter
        // phases of the next.  2-D queues first: their filters are short.
        constexpr size_t kChunkItems = 4;
        auto cost = [&](int i) {
            const sgx_plan& P = *items[i].plan;
            return std::make_pair(P.geom.na, P.nx * P.ny * P.nz);
        };
        {
            std::vector<Queue> chunked;
            for (Queue& Q : queues) {
                std::stable_sort(Q.items.begin(), Q.items.end(),
                                 [&](int x, int y) { return cost(x) > cost(y); });
                for (size_t c = 0; c < Q.items.size(); c += kChunkItems) {
                    Queue C;
                    C.kind = Q.kind;
                    C.variant = Q.variant;
                    C.items.assign(Q.items.begin() + c,
                                   Q.items.begin() + std::min(Q.items.size(), c + kChunkItems));
                    chunked.push_back(std::move(C));
                }
            }
            queues.swap(chunked);
        }
        std::stable_sort(queues.begin(), queues.end(),
                         [](const Queue& a, const Queue& b) { return a.kind < b.kind; });
        for (Queue& Q : queues) {
            // the items' tasks one item after another, each in block order
            // (as a launch of the item alone would dispatch them)
            std::vector<int2> tasks;
            if (Q.kind == 3) {
                std::vector<sgx::CbctItem> its;
                for (int k = 0; k < (int)Q.items.size(); ++k) {
                    const Item& it = items[Q.items[k]];
                    sgx::CbctArgs a = cbct_args_of(*it.plan, it.out);
                    if (it.plan->tune.cbct == 1) a.win = nullptr;
                    its.push_back(sgx::cbct_item(a));
                    const int64_t nb = sgx::cbct_item_blocks(a);
                    for (int64_t b = 0; b < nb; ++b) tasks.push_back(make_int2(k, (int)b));
                }
                Q.items_off = put(its.data(), its.size() * sizeof(sgx::CbctItem));
            } else {
                std::vector<sgx::Back2dItem> its;
                for (int k = 0; k < (int)Q.items.size(); ++k) {
                    const Item& it = items[Q.items[k]];
                    its.push_back(sgx::back2d_item(back2d_run_args(*it.plan, it.out)));
                    Q.lds = std::max(Q.lds, sgx::back2d_item_lds(its.back()));
                    const int nt = sgx::back2d_item_tiles(its.back());
                    for (int b = 0; b < nt; ++b) tasks.push_back(make_int2(k, b));
                }
                Q.items_off = put(its.data(), its.size() * sizeof(sgx::Back2dItem));
            }
            require(tasks.size() < (size_t(1) << 31), SGX_ERR_UNSUPPORTED, "batch: too many tasks");
            Q.ntasks = (int)tasks.size();
            Q.tasks_off = put(tasks.data(), tasks.size() * sizeof(int2));
        }
        if (qcap < host.size()) {
            SGX_HIP(hipStreamSynchronize(ctx->stream));
            if (qbuf) SGX_HIP(hipFree(qbuf));
            qbuf = nullptr;
            qcap = 0;
            SGX_HIP(hipMalloc(&qbuf, host.size()));
            qcap = host.size();
        }
        if (queues.empty()) return;
        SGX_HIP(hipMemcpy(qbuf, host.data(), host.size(), hipMemcpyHostToDevice));
    }

    // Capture every item.  Branches (at most kBranches) take the plans,
    // largest work first onto the least loaded branch; the items of one plan
    // stay on one branch, in the order they were added.  On each branch the
    // filter phases of its queued items run first, queue by queue, then its
    // whole plan executes.  Each queue runs as one launch (one block per
    // task of its table) on a stream of its own, once its own items are
    // filtered.
    void capture(bool use_queues)
    {
        constexpr int kBranches = 4;
        // every cone-beam plan's window table matches its form before capture
        for (const Item& it : items) ensure_windows(*it.plan, ctx->stream);
        build_queues(use_queues);
        std::vector<bool> in_queue(items.size(), false);
        for (const Queue& Q : queues)
            for (int i : Q.items) in_queue[i] = true;
        std::vector<sgx_plan*> plans;
        std::map<sgx_plan*, double> work;
        for (const Item& it : items) {
            if (!work.count(it.plan)) plans.push_back(it.plan);
            const sgx_plan& P = *it.plan;
            work[it.plan] += (double)P.nx * P.ny * P.nz * P.geom.na;
        }
        std::sort(plans.begin(), plans.end(),
                  [&](sgx_plan* a, sgx_plan* b) { return work[a] > work[b]; });
        const int nb = (int)std::min<size_t>(kBranches, plans.size());
        std::vector<double> load(nb, 0.0);
        std::map<sgx_plan*, int> branch_of;
        for (sgx_plan* p : plans) {
            const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            branch_of[p] = b;
            load[b] += work[p];
        }
        const int nq = (int)queues.size();
        ensure_streams((size_t)(nb + nq));
        ensure_filter_events((size_t)(nb * nq));
        drop_graph();
        SGX_HIP(hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal));
        try {
            SGX_HIP(hipEventRecord(fork, origin));
            for (int b = 0; b < nb; ++b) SGX_HIP(hipStreamWaitEvent(branches[b], fork, 0));
            std::vector<std::vector<bool>> used(nq, std::vector<bool>(nb, false));
            for (int j = 0; j < nq; ++j) {
                for (int i : queues[j].items) {
                    const int b = branch_of[items[i].plan];
                    plan_filter_phase(*items[i].plan, items[i].in, branches[b]);
                    used[j][b] = true;
                }
                for (int b = 0; b < nb; ++b)
                    if (used[j][b]) SGX_HIP(hipEventRecord(filtered[j * nb + b], branches[b]));
            }
            for (int i = 0; i < (int)items.size(); ++i)
                if (!in_queue[i])
                    plan_run(*items[i].plan, items[i].in, items[i].out,
                             branches[branch_of[items[i].plan]], false);
            for (int j = 0; j < nq; ++j) {
                const Queue& Q = queues[j];
                hipStream_t st = branches[nb + j];
                for (int b = 0; b < nb; ++b)
                    if (used[j][b]) SGX_HIP(hipStreamWaitEvent(st, filtered[j * nb + b], 0));
                const char* base = static_cast<const char*>(qbuf);
                const int2* tasks = reinterpret_cast<const int2*>(base + Q.tasks_off);
                if (Q.kind == 3)
                    sgx::launch_cbct_queue(
                        Q.variant, reinterpret_cast<const sgx::CbctItem*>(base + Q.items_off),
                        tasks, Q.ntasks, st);
                else
                    sgx::launch_back2d_queue(
                        Q.variant, reinterpret_cast<const sgx::Back2dItem*>(base + Q.items_off),
                        tasks, Q.ntasks, Q.lds, st);
            }
            for (int b = 0; b < nb + nq; ++b) {
                SGX_HIP(hipEventRecord(joins[b], branches[b]));
                SGX_HIP(hipStreamWaitEvent(origin, joins[b], 0));
            }
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(origin, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        SGX_HIP(hipStreamEndCapture(origin, &graph));
        SGX_HIP(hipGraphInstantiateWithFlags(&exec, graph, 0));
        gens.clear();
        for (const Item& it : items) gens.push_back(it.plan->gen);
        nbranch = nb;
        queued = use_queues;
        dirty = false;
    }

    // True when some item's plan changed since the capture.
    bool stale() const
    {
        if (gens.size() != items.size()) return true;
        for (size_t i = 0; i < items.size(); ++i)
            if (items[i].plan->gen != gens[i]) return true;
        return false;
    }
};

extern "C" {

int sgx_batch_create(sgx_ctx* ctx, sgx_batch** out)
{
    return guard(ctx, [&] {
        require(out, SGX_ERR_ARG, "null pointer argument");
        *out = nullptr;
        auto B = std::make_unique<sgx_batch>();
        B->ctx = ctx;
        SGX_HIP(hipStreamCreateWithFlags(&B->origin, hipStreamNonBlocking));
        SGX_HIP(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
        SGX_HIP(hipEventCreate(&B->t0));
        SGX_HIP(hipEventCreate(&B->t1));
        *out = B.release();
    });
}

void sgx_batch_destroy(sgx_batch* batch)
{
    if (!batch) return;
    (void)hipSetDevice(batch->ctx->device);
    (void)hipStreamSynchronize(batch->ctx->stream);
    delete batch;
}

int sgx_batch_add(sgx_batch* batch, sgx_plan* plan, const float* sino, float* image)
{
    if (!batch) return SGX_ERR_ARG;
    return guard(batch->ctx, [&] {
        require(plan && sino && image, SGX_ERR_ARG, "null pointer argument");
        require(plan->ctx == batch->ctx, SGX_ERR_ARG, "the plan belongs to another context");
        batch->items.push_back({plan, sino, image});
        batch->dirty = true;
    });
}

int sgx_batch_execute(sgx_batch* batch, uint32_t flags)
{
    if (!batch) return SGX_ERR_ARG;
    sgx_ctx* ctx = batch->ctx;
    return guard(ctx, [&] {
        require(!batch->items.empty(), SGX_ERR_ARG, "empty batch");
        require(!(flags & SGX_HOST), SGX_ERR_ARG, "batches take device pointers");
        const int mode = (flags & SGX_BATCH_NO_QUEUE) ? 0 : (flags & SGX_BATCH_QUEUE) ? 1 : 2;
        if (!batch->dirty && (batch->stale() || batch->mode != mode)) batch->dirty = true;
        if (batch->dirty) {
            if (!(flags & SGX_BATCH_NO_TUNE)) {
                for (auto& it : batch->items)
                    if (!it.plan->tuned) autotune(*it.plan, it.in, it.out, 3, true);
            }
            SGX_HIP(hipStreamSynchronize(ctx->stream));
            if (mode == 2) {
                // time a replay of each form (after one untimed replay) and
                // keep the faster
                auto timed = [&]() {
                    SGX_HIP(hipGraphLaunch(batch->exec, ctx->stream));
                    SGX_HIP(hipEventRecord(batch->t0, ctx->stream));
                    SGX_HIP(hipGraphLaunch(batch->exec, ctx->stream));
                    SGX_HIP(hipEventRecord(batch->t1, ctx->stream));
                    SGX_HIP(hipEventSynchronize(batch->t1));
                    float ms = 0.f;
                    SGX_HIP(hipEventElapsedTime(&ms, batch->t0, batch->t1));
                    return ms;
                };
                batch->capture(true);
                batch->ms_queues = batch->queues.empty() ? -1.f : timed();
                batch->capture(false);
                batch->ms_launches = timed();
                if (batch->ms_queues >= 0.f && batch->ms_queues < batch->ms_launches)
                    batch->capture(true);
            } else {
                batch->capture(mode == 1);
                batch->ms_queues = batch->ms_launches = -1.f;
            }
            batch->mode = mode;
        }
        SGX_HIP(hipEventRecord(batch->t0, ctx->stream));
        SGX_HIP(hipGraphLaunch(batch->exec, ctx->stream));
        SGX_HIP(hipEventRecord(batch->t1, ctx->stream));
        batch->launched = true;
        if (!(flags & SGX_ASYNC)) SGX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int sgx_batch_last_ms(sgx_batch* batch, float* ms)
{
    if (!batch || !ms) return SGX_ERR_ARG;
    return guard(batch->ctx, [&] {
        require(batch->launched, SGX_ERR_ARG, "no execute recorded yet");
        SGX_HIP(hipEventSynchronize(batch->t1));
        SGX_HIP(hipEventElapsedTime(ms, batch->t0, batch->t1));
    });
}

int sgx_batch_info(const sgx_batch* batch, int64_t* items, int64_t* branches)
{
    if (!batch) return SGX_ERR_ARG;
    if (items) *items = (int64_t)batch->items.size();
    if (branches) *branches = batch->nbranch;
    return SGX_OK;
}

int sgx_batch_choice(const sgx_batch* batch, float* ms_queues, float* ms_launches)
{
    if (!batch) return SGX_ERR_ARG;
    if (ms_queues) *ms_queues = batch->ms_queues;
    if (ms_launches) *ms_launches = batch->ms_launches;
    return SGX_OK;
}

int sgx_batch_queues(const sgx_batch* batch, int64_t* queues, int64_t* queued_items,
                     int64_t* tasks)
{
    if (!batch) return SGX_ERR_ARG;
    int64_t n = 0, t = 0;
    for (const auto& Q : batch->queues) {
        n += (int64_t)Q.items.size();
        t += Q.ntasks;
    }
    if (queues) *queues = (int64_t)batch->queues.size();
    if (queued_items) *queued_items = n;
    if (tasks) *tasks = t;
    return SGX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// K8: plan pieces for training (filter only, backprojection of a given q and
// its exact transpose) and the learned correction network
// ---------------------------------------------------------------------------

namespace {

sgx::Back2dArgs back2d_args_of(const sgx_plan& P, const float* q, float* img)
{
    const sgx_geom& G = P.geom;
    sgx::Back2dArgs a{};
    a.q = q;
    a.nb = (int)G.ns;
    a.nviews = (int)G.na;
    a.view_stride = G.ns;
    a.cs = P.cs;
    a.xds = P.xds;
    a.yds = P.yds;
    a.nx = (int)P.nx;
    a.ny = (int)P.ny;
    a.mask = P.mask;
    a.image = img;
    a.geom = (G.kind == SGX_SINO_PAR || G.kind == SGX_SINO_MOJ) ? 0
                                                                  : (G.kind == SGX_SINO_FAN_ARC ? 1 : 2);
    a.wb = (float)((G.ns + 1) / 2.0 + (double)(float)G.offset_s);
    a.dsd_ds = (float)G.dsd / (float)G.ds;
    a.dso_ds = (float)G.dso / (float)G.ds;
    a.so_ds = (float)G.source_offset / (float)G.ds;
    a.partial = P.partial;
    a.partial_cap = P.partial_cap;
    a.want_blocks = P.tune.want_blocks;
    a.vc_max = P.tune.vc_max;
    // the per-pixel arc form: sgx_plan_back and sgx_plan_back_t are an exact
    // transpose pair
    a.pixel_form = true;
    return a;
}

}  // namespace

struct sgx_net {
    sgx_ctx* ctx = nullptr;
    int layers = 0, channels = 0;
    std::vector<int> cin, cout;
    std::vector<int64_t> woff, boff;       // offsets of each layer's weights / biases
    int64_t nparams = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    int step = 0;
    // training state of the last kept forward
    const float* q = nullptr;
    int64_t rows = 0, n = 0;
    float* z = nullptr;                    // pre-activations of layers 1 .. L-1
    int64_t z_cap = 0;
    float* d = nullptr;                    // two gradient buffers
    int64_t d_cap = 0;
    float* part = nullptr;
    int64_t part_cap = 0;
    uint16_t* wbf = nullptr;               // BF16 weights of the MFMA layers
    int64_t wver = 1;                      // bumped by every parameter write
    int64_t wbf_ver = 0;                   // parameters the fused form's wbf holds
    bool kept = false;

    ~sgx_net()
    {
        for (float* p : {w, g, m, v, z, d, part})
            if (p) (void)hipFree(p);
        if (wbf) (void)hipFree(wbf);
    }
    void grow(float*& p, int64_t& cap, int64_t need)
    {
        if (cap >= need) return;
        SGX_HIP(hipStreamSynchronize(ctx->stream));
        if (p) SGX_HIP(hipFree(p));
        p = nullptr;
        cap = 0;
        SGX_HIP(hipMalloc(&p, sizeof(float) * need));
        cap = need;
    }
};

extern "C" {

int sgx_plan_filter(sgx_plan* plan, const float* sino, float* q, uint32_t flags)
{
    if (!plan) return SGX_ERR_ARG;
    sgx_ctx* ctx = plan->ctx;
    return guard(ctx, [&] {
        require(sino && q, SGX_ERR_ARG, "null pointer argument");
        require(!(flags & SGX_HOST), SGX_ERR_ARG, "device pointers only");
        require(!plan->is3d() && plan->prec == SGX_PREC_FP32, SGX_ERR_UNSUPPORTED,
                "sgx_plan_filter is provided for 2-D FP32 plans");
        sgx::launch_sino_filter(plan->filter_args(sino, q), ctx->stream);
        finish(ctx, nullptr, nullptr, 0, flags);
    });
}

int sgx_plan_back(sgx_plan* plan, const float* q, float* image, uint32_t flags)
{
    if (!plan) return SGX_ERR_ARG;
    sgx_ctx* ctx = plan->ctx;
    return guard(ctx, [&] {
        require(q && image, SGX_ERR_ARG, "null pointer argument");
        require(!(flags & SGX_HOST), SGX_ERR_ARG, "device pointers only");
        require(!plan->is3d(), SGX_ERR_UNSUPPORTED, "sgx_plan_back is provided for 2-D plans");
        sgx::launch_back2d(back2d_args_of(*plan, q, image), ctx->stream);
        finish(ctx, nullptr, nullptr, 0, flags);
    });
}

int sgx_plan_back_t(sgx_plan* plan, const float* image, float* q, uint32_t flags)
{
    if (!plan) return SGX_ERR_ARG;
    sgx_ctx* ctx = plan->ctx;
    return guard(ctx, [&] {
        require(q && image, SGX_ERR_ARG, "null pointer argument");
        require(!(flags & SGX_HOST), SGX_ERR_ARG, "device pointers only");
        require(!plan->is3d(), SGX_ERR_UNSUPPORTED, "sgx_plan_back_t is provided for 2-D plans");
        sgx::launch_back2d_t(back2d_args_of(*plan, nullptr, nullptr), image, q, ctx->stream);
        finish(ctx, nullptr, nullptr, 0, flags);
    });
}

int sgx_net_create(sgx_ctx* ctx, int layers, int channels, uint64_t seed, sgx_net** out)
{
    return guard(ctx, [&] {
        require(out, SGX_ERR_ARG, "null pointer argument");
        require(layers >= 2 && layers <= 16 && channels >= 1 && channels <= 32, SGX_ERR_ARG,
                "layers 2..16, channels 1..32");
        *out = nullptr;
        auto N = std::make_unique<sgx_net>();
        N->ctx = ctx;
        N->layers = layers;
        N->channels = channels;
        for (int l = 0; l < layers; ++l) {
            N->cin.push_back(l == 0 ? 1 : channels);
            N->cout.push_back(l == layers - 1 ? 1 : channels);
            N->woff.push_back(N->nparams);
            N->nparams += (int64_t)N->cout[l] * N->cin[l] * 9;
            N->boff.push_back(N->nparams);
            N->nparams += N->cout[l];
        }
        // uniform(-a, a) with a = sqrt(6 / fan_in) (GELU layers), the last
        // layer and every bias 0: the initial correction is exactly zero
        std::vector<float> h(N->nparams, 0.f);
        uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
        auto uni = [&]() {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            return (double)(st >> 11) * (1.0 / 9007199254740992.0);
        };
        for (int l = 0; l + 1 < layers; ++l) {
            const double a = std::sqrt(6.0 / (N->cin[l] * 9.0));
            for (int64_t e = 0; e < (int64_t)N->cout[l] * N->cin[l] * 9; ++e)
                h[N->woff[l] + e] = (float)((2.0 * uni() - 1.0) * a);
        }
        for (float** p : {&N->w, &N->g, &N->m, &N->v}) SGX_HIP(hipMalloc(p, 4 * N->nparams));
        SGX_HIP(hipMalloc(&N->wbf, 2 * sgx::nn_wprep_halves() * std::max(1, layers - 2)));
        SGX_HIP(hipMemcpy(N->w, h.data(), 4 * N->nparams, hipMemcpyHostToDevice));
        for (float* p : {N->g, N->m, N->v}) sgx::launch_fill(p, N->nparams, 0.f, ctx->stream);
        SGX_HIP(hipStreamSynchronize(ctx->stream));
        *out = N.release();
    });
}

void sgx_net_destroy(sgx_net* net)
{
    if (!net) return;
    (void)hipSetDevice(net->ctx->device);
    (void)hipStreamSynchronize(net->ctx->stream);
    delete net;
}

int64_t sgx_net_num_params(const sgx_net* net) { return net ? net->nparams : 0; }

int sgx_net_params(sgx_net* net, float* host, int set)
{
    if (!net) return SGX_ERR_ARG;
    return guard(net->ctx, [&] {
        require(host, SGX_ERR_ARG, "null pointer argument");
        SGX_HIP(hipStreamSynchronize(net->ctx->stream));
        if (set) {
            SGX_HIP(hipMemcpy(net->w, host, 4 * net->nparams, hipMemcpyHostToDevice));
            net->wver += 1;
        }
        else SGX_HIP(hipMemcpy(host, net->w, 4 * net->nparams, hipMemcpyDeviceToHost));
    });
}

int sgx_net_grads(sgx_net* net, float* host)
{
    if (!net) return SGX_ERR_ARG;
    return guard(net->ctx, [&] {
        require(host, SGX_ERR_ARG, "null pointer argument");
        SGX_HIP(hipStreamSynchronize(net->ctx->stream));
        SGX_HIP(hipMemcpy(host, net->g, 4 * net->nparams, hipMemcpyDeviceToHost));
    });
}

int sgx_net_zero_grad(sgx_net* net)
{
    if (!net) return SGX_ERR_ARG;
    return guard(net->ctx, [&] { sgx::launch_fill(net->g, net->nparams, 0.f, net->ctx->stream); });
}

int sgx_net_forward(sgx_net* net, const float* q, int64_t rows, int64_t n, float* out, int keep)
{
    if (!net) return SGX_ERR_ARG;
    sgx_ctx* ctx = net->ctx;
    return guard(ctx, [&] {
        require(q && out, SGX_ERR_ARG, "null pointer argument");
        require(rows > 0 && n > 0 && rows < 65536, SGX_ERR_ARG, "rows must be 1..65535");
        sgx_net& N = *net;
        const int L = N.layers;
        const int64_t per = rows * N.channels * n;          // one hidden activation
        if (!keep && sgx::net_fused_ok(L, N.channels)) {
            // inference: one fused launch; the BF16 weights are converted
            // once per parameter update
            if (N.wbf_ver != N.wver) {
                sgx::launch_net_prep(N.w, L, N.woff.data(), N.wbf, ctx->stream);
                N.wbf_ver = N.wver;
            }
            sgx::launch_net_infer(q, (int)rows, (int)n, N.w, L, N.woff.data(), N.boff.data(),
                                  N.wbf, out, ctx->stream);
            N.kept = false;
            return;
        }
        N.wbf_ver = 0;                       // the per-layer kernels reuse wbf
        // hidden pre-activations: all L-1 kept for training, else two in turn
        N.grow(N.z, N.z_cap, per * (keep ? (L - 1) : std::min(L - 1, 2)));
        const float* in = q;
        for (int l = 0; l < L; ++l) {
            const bool last = l == L - 1;
            float* dst = last ? out : N.z + per * (keep ? l : (l % 2));
            sgx::launch_conv(in, (int)rows, N.cin[l], N.cout[l], (int)n, N.w + N.woff[l],
                             N.w + N.boff[l], last ? q : nullptr, l > 0, dst, N.wbf,
                             ctx->stream);
            in = dst;
        }
        N.q = q;
        N.rows = rows;
        N.n = n;
        N.kept = keep != 0;
    });
}

int sgx_net_backward(sgx_net* net, const float* gout)
{
    if (!net) return SGX_ERR_ARG;
    sgx_ctx* ctx = net->ctx;
    return guard(ctx, [&] {
        require(gout, SGX_ERR_ARG, "null pointer argument");
        sgx_net& N = *net;
        require(N.kept, SGX_ERR_ARG, "sgx_net_backward needs a kept forward");
        N.wbf_ver = 0;                       // the data-gradient kernels reuse wbf
        const int L = N.layers;
        const int64_t per = N.rows * N.channels * N.n;
        N.grow(N.d, N.d_cap, 2 * per);
        const int blocks = sgx::nn_wgrad_blocks((int)N.rows, (int)N.n);
        N.grow(N.part, N.part_cap, (int64_t)blocks * (N.channels * N.channels * 9 + N.channels));
        const float* delta = gout;               // dL / dz_L (the output is q + z_L)
        for (int l = L - 1; l >= 0; --l) {
            const float* x = l == 0 ? N.q : N.z + per * (l - 1);
            sgx::launch_conv_wgrad(delta, x, l > 0, (int)N.rows, N.cin[l], N.cout[l], (int)N.n,
                                   N.part, N.g + N.woff[l], N.g + N.boff[l], ctx->stream);
            if (l == 0) break;
            float* din = N.d + per * (l % 2);
            sgx::launch_conv_dgrad(delta, (int)N.rows, N.cin[l], N.cout[l], (int)N.n,
                                   N.w + N.woff[l], N.z + per * (l - 1), din, N.wbf,
                                   ctx->stream);
            delta = din;
        }
    });
}

int sgx_net_adam(sgx_net* net, float lr, float beta1, float beta2, float eps)
{
    if (!net) return SGX_ERR_ARG;
    return guard(net->ctx, [&] {
        net->step += 1;
        net->wver += 1;
        sgx::launch_adam(net->w, net->g, net->m, net->v, net->nparams, lr, beta1, beta2, eps,
                         net->step, net->ctx->stream);
    });
}

int sgx_mse(sgx_ctx* ctx, const float* x, const float* t, int64_t n, float scale, float* grad,
            double* sum_sq)
{
    return guard(ctx, [&] {
        require(x && t && n > 0, SGX_ERR_ARG, "bad argument");
        const int blocks = sgx::nn_mse_blocks(n);
        float* part = static_cast<float*>(ctx->get(W_EXTRA, 4 * blocks));
        sgx::launch_mse(x, t, n, scale, grad, part, ctx->stream);
        if (sum_sq) {
            std::vector<float> h(blocks);
            SGX_HIP(hipMemcpyAsync(h.data(), part, 4 * blocks, hipMemcpyDeviceToHost,
                                   ctx->stream));
            SGX_HIP(hipStreamSynchronize(ctx->stream));
            double s = 0;
            for (float v : h) s += v;
            *sum_sq = s;
        }
    });
}

}  // extern "C"
