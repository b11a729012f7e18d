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
// bounds logic, two paired loads (taps is and is+1 of rows it and it+1) and
// the bilinear blend.  Per-view cos/sin are wave-uniform and read through the
// scalar cache.  Accumulation is in registers (FP32); each voxel is written
// once, so no atomics.
//
// Storage.  q is FP32, or bfloat16 (the precision form): a pair of adjacent
// bins is one 4-byte (bf16) or 8-byte (FP32) load.

#include "sgx_internal.h"
#include "sgx.h"

namespace sgx {
namespace {

constexpr int kWaves = 4;      // waves per block, each on its own z-slab
constexpr int kPatch = 8;      // 8 x 8 columns per wave

__device__ __forceinline__ float2 load_pair(const float* p)
{
    float2 v;
    __builtin_memcpy(&v, p, sizeof(float2));
    return v;
}

__device__ __forceinline__ float2 load_pair(const __hip_bfloat16* p)
{
    uint32_t u;
    __builtin_memcpy(&u, p, sizeof(uint32_t));
    float2 v;
    v.x = __uint_as_float(u << 16);
    v.y = __uint_as_float(u & 0xffff0000u);
    return v;
}

template <int ZB, bool ARC, typename QT>
__global__ __launch_bounds__(64 * kWaves) void cbct_back_kernel(
    const QT* __restrict__ q, int ns, int nt, int nviews, int64_t view_stride,
    const ViewCS* __restrict__ cs, const float* __restrict__ xds,
    const float* __restrict__ yds, const float* __restrict__ zds, int nx, int ny, int nz,
    const uint8_t* __restrict__ mask, float* __restrict__ image, float dsd_ds, float dso_ds,
    float ws, float wt, float inv_dt_ds)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tiles_x = (nx + kPatch - 1) / kPatch;
    const int ix = (blockIdx.x % tiles_x) * kPatch + (lane & (kPatch - 1));
    const int iy = (blockIdx.x / tiles_x) * kPatch + (lane / kPatch);
    const int iz0 = (blockIdx.y * kWaves + wave) * ZB;
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
        z[j] = iz < nz ? zds[iz] : 0.f;
        acc[j] = 0.f;
    }
    if (!any) {
#pragma unroll
        for (int j = 0; j < ZB; ++j)
            if (iz0 + j < nz) image[col + nxy * (iz0 + j)] = 0.f;
        return;
    }

    const float x = xds[ix], y = yds[iy];
    const float dsd2 = dsd_ds * dsd_ds;
    const float nsf = (float)ns, ntf = (float)nt;

    const QT* qv = q;
    for (int ia = 0; ia < nviews; ++ia, qv += view_stride) {
        const ViewCS v = cs[ia];
        const float xb = x * v.c + y * v.s;
        const float yb = dso_ds - (-x * v.s + y * v.c);
        const float mag = dsd_ds / yb;
        float sp, w2;
        if (ARC) {
            sp = dsd_ds * atan2f(xb, yb);
            w2 = dsd2 / (xb * xb + yb * yb);
        } else {
            sp = mag * xb;
            w2 = mag * mag;
        }
        const float bs = sp + ws;
        // 1 <= floor(bs) < ns  <=>  1 <= bs < ns (also false for NaN)
        if (!(bs >= 1.f && bs < nsf)) continue;  // the column misses the detector
        const float isf = floorf(bs);
        const float wr = bs - isf;
        const int i0 = (int)isf - 1;
        const float mz = mag * inv_dt_ds;
        const QT* base = qv + i0;
        // Branch-free over z so that all 2*ZB paired loads issue back to back:
        // an out-of-range row is clamped to a valid address and weighted 0.
        float2 lo[ZB], hi[ZB];
        float wu[ZB], wz[ZB];
#pragma unroll
        for (int j = 0; j < ZB; ++j) {
            const float bt = fmaf(mz, z[j], wt);
            const bool ok = bt >= 1.f && bt < ntf;          // 1 <= floor(bt) < nt
            wz[j] = ok ? w2 : 0.f;
            const float btc = ok ? bt : 1.f;
            const float itf = floorf(btc);
            wu[j] = btc - itf;
            const QT* r0 = base + (int64_t)((int)itf - 1) * ns;
            lo[j] = load_pair(r0);
            hi[j] = load_pair(r0 + ns);
        }
#pragma unroll
        for (int j = 0; j < ZB; ++j) {
            const float p1 = fmaf(wr, lo[j].y - lo[j].x, lo[j].x);
            const float p2 = fmaf(wr, hi[j].y - hi[j].x, hi[j].x);
            acc[j] = fmaf(wz[j], fmaf(wu[j], p2 - p1, p1), acc[j]);
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

template <int ZB, bool ARC, typename QT>
void launch(const CbctArgs& a, hipStream_t st)
{
    const int tiles = ((a.nx + kPatch - 1) / kPatch) * ((a.ny + kPatch - 1) / kPatch);
    const int zblocks = (a.nz + ZB * kWaves - 1) / (ZB * kWaves);
    dim3 grid(tiles, zblocks), block(64 * kWaves);
    hipLaunchKernelGGL((cbct_back_kernel<ZB, ARC, QT>), grid, block, 0, st,
                       static_cast<const QT*>(a.q), a.ns, a.nt, a.nviews, a.view_stride, a.cs,
                       a.xds, a.yds, a.zds, a.nx, a.ny, a.nz, a.mask, a.image, a.dsd_ds,
                       a.dso_ds, a.ws, a.wt, a.inv_dt_ds);
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

void launch_cbct_back(const CbctArgs& a, hipStream_t st)
{
    constexpr int ZB = 8;
    if (a.q_bf16) {
        if (a.is_arc) launch<ZB, true, __hip_bfloat16>(a, st);
        else launch<ZB, false, __hip_bfloat16>(a, st);
    } else {
        if (a.is_arc) launch<ZB, true, float>(a, st);
        else launch<ZB, false, float>(a, st);
    }
}

}  // namespace sgx
