// K3: voxel-driven cone-beam backprojection for FDK (gfx950).
//
// Replaces cbct_back_fan! / cbct_back_fan_voxel (src/fbp3/cbct-back.jl:95-213).
// Per voxel (x, y, z) (all divided by ds) and per view beta:
//   x_beta = x cos + y sin,   y_beta = dso - (-x sin + y cos)
//   flat: mag = dsd / y_beta, s' = mag x_beta,          w2 = mag^2
//   arc:  mag = dsd / y_beta, s' = dsd atan(x_beta, y_beta),
//         w2 = dsd^2 / (x_beta^2 + y_beta^2)
//   t' = mag z,  bs = s' + ws,  bt = t' / (dt/ds) + wt
//   if 1 <= floor(bs) < ns and 1 <= floor(bt) < nt: add w2 * bilinear(q)
//
// Mapping.  One lane owns one (x, y) column and ZB voxels along z; a
// wavefront owns an 8 x 8 patch of columns, so the 64 lanes of every gather
// touch a ~17-bin x 2-row window of one view (a few cache lines).  All
// z-independent work (rotation, the one divide, the arc atan2, w2, bs, the s
// weights and the row base address) is done once per column and view and
// amortised over ZB voxels; per voxel the loop is one FMA for bt, the floor /
// bounds logic, one row-pair load (taps is and is+1 of rows it and it+1) and
// four FMAs.  Per-view cos/sin are wave-uniform and read through the scalar
// cache.  Accumulation is in registers (FP32); each voxel is written once, so
// no atomics.
//
// Storage.  q is FP32, bfloat16, or FP16 with a per-view scale (the 16-bit
// forms), either column-major (the direct entry point) or in the row-pair
// difference layout the plan's filter writes: entry (is, it) of a view holds
// (q[is, it], q[is, it+1] - q[is, it]), so the four taps of an update are one
// 16-byte (FP32) or 8-byte (16-bit) load and the t interpolation is one FMA
// per column.

#include "sgx_internal.h"
#include "sgx.h"

#include <hip/hip_fp16.h>

#include <cstdlib>

namespace sgx {
namespace {

constexpr int kWaves = 4;      // waves per block, each on its own z-slab
constexpr int kPatch = 8;      // 8 x 8 columns per wave

// Storage formats of the filtered projections.
enum { F_F32 = 0, F_BF16 = 1, F_F16S = 2 };

// Raw buffer resource over one view: the 32-bit offsets of the gathers are
// VGPR byte offsets and the second detector row is reached through the
// wave-uniform SGPR offset, so both taps of a row pair share one address.
// Loads at offsets beyond num_records return 0 instead of faulting.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t view_rsrc(const char* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(p), 0, (int)bytes, 0x00020000);
}

// 16-bit storage: the pair (is, is+1) starts at a 2-byte-aligned offset, so
// it is taken from the dword-aligned 8-byte window that contains it (one
// aligned load plus one v_alignbit).  The window can reach 4 bytes past the
// pair: every 16-bit buffer handed to this kernel has 4 bytes of padding
// after its last view (plan buffers are allocated so; the direct entry point
// stages into a padded copy) and the views' resource covers them.  When the
// row pitch is a multiple of 4 bytes the second row shares the first row's
// window offset (through the SGPR offset); otherwise its offset is formed
// separately.
__device__ __forceinline__ uint32_t pair16(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                           uint32_t soff)
{
    const auto u = __builtin_amdgcn_raw_buffer_load_b64(r, voff & ~3u, soff, 0);
    return __builtin_amdgcn_alignbit(u[1], u[0], (voff & 2u) << 3);
}

__device__ __forceinline__ float2 dec_bf16(uint32_t u)
{
    return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u));
}

__device__ __forceinline__ float2 dec_f16(uint32_t u)
{
    return make_float2(__half2float(__ushort_as_half((unsigned short)(u & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short)(u >> 16))));
}

// The four taps of one update as (a0, d0, a1, d1): a = q[is, it] and
// d = q[is, it+1] - q[is, it] for the two columns is (0) and is+1 (1).
// Row-pair layout: entry (is, it) of a view already holds (a, d) for column
// is, so the two entries (is, it), (is+1, it) are one 16-byte (FP32) or
// 8-byte (16-bit) load.  Column-major layout: two row loads (row it+1
// through the SGPR offset) and the differences formed here.
template <int FMT, bool PITCH4, bool RP>
__device__ __forceinline__ float4 taps(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t rowb)
{
    if (RP) {
        if (FMT == F_F32) {
            const auto u = __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0);
            return make_float4(__uint_as_float(u[0]), __uint_as_float(u[1]),
                               __uint_as_float(u[2]), __uint_as_float(u[3]));
        } else {
            const auto u = __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0);
            const float2 e0 = FMT == F_BF16 ? dec_bf16(u[0]) : dec_f16(u[0]);
            const float2 e1 = FMT == F_BF16 ? dec_bf16(u[1]) : dec_f16(u[1]);
            return make_float4(e0.x, e0.y, e1.x, e1.y);
        }
    }
    float2 lo, hi;
    if (FMT == F_F32) {
        const auto a = __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0);
        const auto b = __builtin_amdgcn_raw_buffer_load_b64(r, voff, rowb, 0);
        lo = make_float2(__uint_as_float(a[0]), __uint_as_float(a[1]));
        hi = make_float2(__uint_as_float(b[0]), __uint_as_float(b[1]));
    } else {
        const uint32_t a = pair16(r, voff, 0);
        const uint32_t b = PITCH4 ? pair16(r, voff, rowb) : pair16(r, voff + rowb, 0);
        lo = FMT == F_BF16 ? dec_bf16(a) : dec_f16(a);
        hi = FMT == F_BF16 ? dec_bf16(b) : dec_f16(b);
    }
    return make_float4(lo.x, hi.x - lo.x, lo.y, hi.y - lo.y);
}

template <int FMT, bool RP> struct Tap {
    // bytes per stored sample (column-major) or per row-pair entry
    static constexpr uint32_t bytes = (FMT == F_F32 ? 4 : 2) * (RP ? 2 : 1);
    // readable padding the gathers may touch past the last view
    static constexpr uint32_t pad = (FMT != F_F32 && !RP) ? 4 : 0;
};

// Byte offset that no view's buffer resource covers (every view is below
// 2 GiB): a buffer load there returns zeros without touching memory.
constexpr uint32_t kMissOffset = 0x80000000u;

// Block order.  Workgroups are handed to the eight XCDs round-robin, and
// each XCD has its own L2.  The physical block index is renumbered so that
// XCD k runs one contiguous run of logical blocks (a compact z-slab and
// y-band of the volume), whose footprint on each view is a fraction of the
// detector, instead of every eighth tile of the whole volume.
__device__ __forceinline__ int xcd_order(int phys, int nblocks)
{
    const int q = nblocks >> 3, r = nblocks & 7;
    const int k = phys & 7, slot = phys >> 3;
    return k * q + min(k, r) + slot;
}

// Per update the loop below costs one FMA (bt), a truncating convert, a
// fract, one 24-bit multiply-add for the byte offset, one buffer load (row
// pair layout) and four FMAs: p0 = a0 + wu d0, p1 = a1 + wu d1, then
// acc += wa p0 + wb p1 with the s weights and w2 folded per view
// (wa = w2 (1 - wr), wb = w2 wr).
// The bounds tests are hoisted: bt is linear in z, so when bt at the lane's
// lowest and highest z are both in [1, nt) every voxel of the lane is in
// range; only lanes at the detector's top or bottom edge take the checked
// path, where a voxel out of range loads zeros (offset kMissOffset).
template <int ZB, bool ARC, int FMT, bool PITCH4, bool RP>
__global__ __launch_bounds__(64 * kWaves) void cbct_back_kernel(
    const void* __restrict__ q, int ns, int nt, int nviews, int64_t view_stride,
    const ViewCS* __restrict__ cs, const float* __restrict__ vscale,
    const float* __restrict__ xds, const float* __restrict__ yds,
    const float* __restrict__ zds, int nx, int ny, int nz, const uint8_t* __restrict__ mask,
    float* __restrict__ image, float dsd_ds, float dso_ds, float ws, float wt,
    float inv_dt_ds, int remap)
{
    using T = Tap<FMT, RP>;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tiles_x = (nx + kPatch - 1) / kPatch;
    const int tiles = tiles_x * ((ny + kPatch - 1) / kPatch);
    const int blk = remap ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x;
    const int tile = blk % tiles, zblk = blk / tiles;
    const int ix = (tile % tiles_x) * kPatch + (lane & (kPatch - 1));
    const int iy = (tile / tiles_x) * kPatch + (lane / kPatch);
    const int iz0 = (zblk * kWaves + wave) * ZB;
    if (ix >= nx || iy >= ny || iz0 >= nz) return;

    const int64_t nxy = (int64_t)nx * ny;
    const int64_t col = ix + (int64_t)nx * iy;

    float z[ZB], acc[ZB];
    bool any = false;
#pragma unroll
    for (int j = 0; j < ZB; ++j) {
        const int iz = iz0 + j;
        const bool in = iz < nz && (mask == nullptr || mask[col + nxy * iz]);
        any |= in;
        z[j] = zds[min(iz, nz - 1)];
        acc[j] = 0.f;
    }
    if (!any) {
#pragma unroll
        for (int j = 0; j < ZB; ++j)
            if (iz0 + j < nz) image[col + nxy * (iz0 + j)] = 0.f;
        return;
    }
    float zlo = z[0], zhi = z[0];
#pragma unroll
    for (int j = 1; j < ZB; ++j) {
        zlo = fminf(zlo, z[j]);
        zhi = fmaxf(zhi, z[j]);
    }

    const float x = xds[ix], y = yds[iy];
    const float dsd2 = dsd_ds * dsd_ds;
    const float nsf = (float)ns, ntf = (float)nt;
    const uint32_t rowb = (uint32_t)ns * T::bytes;           // bytes per detector row
    // bytes per view (+ padding reachable by the 16-bit window loads)
    const uint32_t vbytes = (uint32_t)nt * rowb + T::pad;
    const int64_t viewb = view_stride * (int64_t)(FMT == F_F32 ? 4 : 2) * (RP ? 2 : 1);

    const char* v0 = static_cast<const char*>(q);
    for (int ia = 0; ia < nviews; ++ia, v0 += viewb) {
        const ViewCS v = cs[ia];
        const float xb = x * v.c + y * v.s;
        const float yb = dso_ds - (-x * v.s + y * v.c);
        const float mag = dsd_ds * __builtin_amdgcn_rcpf(yb);
        float sp, w2;
        if (ARC) {
            sp = dsd_ds * atan2f(xb, yb);
            w2 = dsd2 * __builtin_amdgcn_rcpf(xb * xb + yb * yb);
        } else {
            sp = mag * xb;
            w2 = mag * mag;
        }
        if (FMT == F_F16S) w2 *= vscale[ia];
        const float bs = sp + ws;
        // 1 <= floor(bs) < ns  <=>  1 <= bs < ns (also false for NaN)
        if (!(bs >= 1.f && bs < nsf)) continue;          // the column misses the detector
        const float isf = floorf(bs);
        const float wb = w2 * (bs - isf);
        const float wa = w2 - wb;
        // byte offset of (is-1, it-1) is it*rowb + (is-1)*bytes - rowb
        const uint32_t colb = ((uint32_t)isf - 1u) * T::bytes - rowb;
        const float mz = mag * inv_dt_ds;
        const __amdgpu_buffer_rsrc_t rs = view_rsrc(v0, vbytes);

        float4 tp[ZB];
        float wu[ZB];
        const float blo = fmaf(mz, zlo, wt), bhi = fmaf(mz, zhi, wt);
        if (blo >= 1.f && blo < ntf && bhi >= 1.f && bhi < ntf) {
#pragma unroll
            for (int j = 0; j < ZB; ++j) {
                const float bt = fmaf(mz, z[j], wt);
                const uint32_t it = (uint32_t)bt;               // floor, bt >= 1
                wu[j] = __builtin_amdgcn_fractf(bt);
                tp[j] = taps<FMT, PITCH4, RP>(rs, __umul24(it, rowb) + colb, rowb);
            }
        } else {
#pragma unroll
            for (int j = 0; j < ZB; ++j) {
                const float bt = fmaf(mz, z[j], wt);
                const bool ok = bt >= 1.f && bt < ntf;          // 1 <= floor(bt) < nt
                const uint32_t it = (uint32_t)(ok ? bt : 1.f);
                wu[j] = ok ? __builtin_amdgcn_fractf(bt) : 0.f;
                const uint32_t boff = ok ? __umul24(it, rowb) + colb : kMissOffset;
                tp[j] = taps<FMT, PITCH4, RP>(rs, boff, rowb);
            }
        }
#pragma unroll
        for (int j = 0; j < ZB; ++j) {
            const float p0 = fmaf(wu[j], tp[j].y, tp[j].x);
            const float p1 = fmaf(wu[j], tp[j].w, tp[j].z);
            acc[j] = fmaf(wb, p1, fmaf(wa, p0, acc[j]));
        }
    }

#pragma unroll
    for (int j = 0; j < ZB; ++j) {
        const int iz = iz0 + j;
        if (iz < nz) {
            const bool in = mask == nullptr || mask[col + nxy * iz];
            image[col + nxy * iz] = in ? acc[j] : 0.f;
        }
    }
}

template <int ZB, bool ARC, int FMT, bool PITCH4, bool RP>
void launch(const CbctArgs& a, hipStream_t st)
{
    const int64_t tiles = (int64_t)((a.nx + kPatch - 1) / kPatch) * ((a.ny + kPatch - 1) / kPatch);
    const int64_t zblocks = (a.nz + ZB * kWaves - 1) / (ZB * kWaves);
    if (tiles * zblocks >= (int64_t(1) << 31))
        throw Error(SGX_ERR_UNSUPPORTED, "cbct_back: volume too large for one launch");
    dim3 grid((unsigned)(tiles * zblocks)), block(64 * kWaves);
    hipLaunchKernelGGL((cbct_back_kernel<ZB, ARC, FMT, PITCH4, RP>), grid, block, 0, st, a.q, a.ns,
                       a.nt, a.nviews, a.view_stride, a.cs, a.vscale, a.xds, a.yds, a.zds, a.nx,
                       a.ny, a.nz, a.mask, a.image, a.dsd_ds, a.dso_ds, a.ws, a.wt,
                       a.inv_dt_ds, a.xcd_order ? 1 : 0);
    SGX_HIP(hipGetLastError());
}

__global__ void view_trig_kernel(const double* __restrict__ betas, int64_t n, int64_t stride,
                                 ViewCS* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) {
        double s, c;
        sincos(betas[i * stride], &s, &c);
        out[i].c = (float)c;
        out[i].s = (float)s;
    }
}

}  // namespace

void launch_view_trig(const double* betas, int64_t n, int64_t stride, ViewCS* out,
                      hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(view_trig_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       betas, n, stride, out);
    SGX_HIP(hipGetLastError());
}

void launch_cbct_back(const CbctArgs& args, hipStream_t st)
{
    constexpr int ZB = 8;
    // SGX_CBCT_XCD_ORDER=1 selects the XCD-grouped block order (a tuning
    // switch; on the 256^2 x 128 FDK workload the plain order measured faster).
    static const bool env_xcd = [] {
        const char* e = std::getenv("SGX_CBCT_XCD_ORDER");
        return e && e[0] == '1';
    }();
    CbctArgs a = args;
    if (env_xcd) a.xcd_order = true;
    // Byte offsets within one view are 32-bit.
    if ((int64_t)a.ns * a.nt * 8 >= (int64_t(1) << 31))
        throw Error(SGX_ERR_UNSUPPORTED, "cbct_back: one view must be below 2 GiB");
    const bool pitch4 = (a.ns % 2) == 0;       // 16-bit rows a multiple of 4 bytes
    if (a.row_pair) {
        if (a.fmt == SGX_Q_FMT_BF16) {
            if (a.is_arc) launch<ZB, true, F_BF16, true, true>(a, st);
            else launch<ZB, false, F_BF16, true, true>(a, st);
        } else if (a.fmt == SGX_Q_FMT_F16S) {
            if (a.is_arc) launch<ZB, true, F_F16S, true, true>(a, st);
            else launch<ZB, false, F_F16S, true, true>(a, st);
        } else {
            if (a.is_arc) launch<ZB, true, F_F32, true, true>(a, st);
            else launch<ZB, false, F_F32, true, true>(a, st);
        }
    } else if (a.fmt == SGX_Q_FMT_BF16) {
        if (pitch4) {
            if (a.is_arc) launch<ZB, true, F_BF16, true, false>(a, st);
            else launch<ZB, false, F_BF16, true, false>(a, st);
        } else {
            if (a.is_arc) launch<ZB, true, F_BF16, false, false>(a, st);
            else launch<ZB, false, F_BF16, false, false>(a, st);
        }
    } else if (a.fmt == SGX_Q_FMT_F16S) {
        throw Error(SGX_ERR_UNSUPPORTED, "cbct_back: scaled FP16 data come in the row-pair layout");
    } else {
        if (a.is_arc) launch<ZB, true, F_F32, true, false>(a, st);
        else launch<ZB, false, F_F32, true, false>(a, st);
    }
}

}  // namespace sgx
