// K1 / K2: pixel-driven 2-D backprojection, parallel beam and fan beam
// (arc and flat detectors), for gfx950.
//
// Replaces fbp_back_par_xy (src/fbp2/back-par.jl:276-315) and
// fbp_back_fan_xy (src/fbp2/back-fan.jl:280-336).  Per pixel (x, y) / ds
// and per view:
//   parallel: bb = x cos + y sin + wb,                     w2 = 1
//   fan:      d = dso + x sin - y cos,  r = x cos + y sin - source_offset
//             arc:  bb = dsd atan(r, d) + wb,  w2 = dsd^2 / (d^2 + r^2)
//                   (atan2 by a minimax polynomial, w2 by a reciprocal)
//             flat: m = dsd / d, bb = m r + wb, w2 = m^2
//   il = floor(bb); if 1 <= il and il + 1 <= nb: add w2 * lerp(q[il], q[il+1])
//
// Layout.  A block owns a 16 x 16 pixel tile (one pixel per lane) and a
// group of views, processed in chunks of views.  For each view of a chunk one
// thread computes the tile's detector window from the tile's corners (bb is
// monotone along each axis of the pixel box while the source is outside it),
// with a bin of margin; only the windows are staged in LDS, several loads in
// flight per thread.  The two taps of each update are then LDS reads; a lane
// whose taps fall outside its view's window (windows wider than the LDS
// slot, pixels at or behind the source) reads them from global memory, so
// the result never depends on the window.  Per-view cos/sin are wave-uniform
// scalar loads and the sum is kept in a register, views in order.  When there
// are too few pixel tiles to fill the GPU the views are split over G groups;
// each group writes a partial image and a second kernel adds the G partials
// in a fixed order (deterministic, no atomics).

#include "sgx_internal.h"
#include "sgx.h"

namespace sgx {
namespace {

constexpr int kTile = 16;
constexpr int kThreads = kTile * kTile;
constexpr int kLdsBytes = 48 * 1024;
constexpr int kU = 8;           // views per unrolled step of the backprojector
constexpr int kWinMax = 128;    // staged window slot per view (bins), a power of two
constexpr int kStage = 8;       // staging loads in flight per thread

// Detector position bb (1-based bins) and weight w2 of pixel (x, y) (/ ds) in
// one view; shared by the backprojector and its transpose.
template <int GEOM>
__device__ __forceinline__ void pixel_bin(float x, float y, ViewCS t, float wb, float dsd_ds,
                                          float dso_ds, float so_ds, float& bb, float& w2)
{
    w2 = 1.f;
    if (GEOM == 0) {
        bb = x * t.c + y * t.s + wb;
    } else {
        const float d = dso_ds + x * t.s - y * t.c;
        const float r = x * t.c + y * t.s - so_ds;
        if (GEOM == 1 || GEOM == 3) {
            bb = dsd_ds * (GEOM == 3 ? atan2_narrow(r, d) : atan2_fast(r, d)) + wb;
            w2 = dsd_ds * dsd_ds * __builtin_amdgcn_rcpf(d * d + r * r);
        } else {
            const float m = dsd_ds / d;
            bb = m * r + wb;
            w2 = m * m;
        }
    }
}

// Window slot per view: the smallest power of two >= nb, at most kWinMax.
int window_bins(int nb)
{
    int w = 4;
    while (w < nb && w < kWinMax) w *= 2;
    return w;
}

int chunk_views(int nb, int vc_max)
{
    return std::max(1, std::min(vc_max, kLdsBytes / (4 * window_bins(nb) + 8)));
}

// GEOM: 0 parallel, 1 fan arc, 2 fan flat, 3 fan arc whose detector spans
// less than +-45 degrees (atan2_narrow)
template <int GEOM>
__global__ __launch_bounds__(kThreads) void back2d_kernel(
    const float* __restrict__ q, int nb, int nviews, int64_t view_stride,
    const ViewCS* __restrict__ cs, const float* __restrict__ xds,
    const float* __restrict__ yds, int nx, int ny, const uint8_t* __restrict__ mask,
    float* __restrict__ out, int64_t out_group_stride, int views_per_group, int vc, int wshift,
    float wb, float dsd_ds, float dso_ds, float so_ds)
{
    // vc window slots of wm bins, then per view int2 {lo + 1, n - 1}: the
    // window holds 0-based bins lo .. lo + n - 1 (n = 0: nothing staged)
    extern __shared__ float lds[];
    const int wm = 1 << wshift;
    int2* win = reinterpret_cast<int2*>(lds + vc * wm);
    const int tid = threadIdx.x;
    const int tiles_x = (nx + kTile - 1) / kTile;
    const int tx0 = (blockIdx.x % tiles_x) * kTile, ty0 = (blockIdx.x / tiles_x) * kTile;
    const int ix = tx0 + (tid % kTile);
    const int iy = ty0 + (tid / kTile);
    const int64_t pix = ix + (int64_t)nx * iy;
    const bool inside = ix < nx && iy < ny;
    const bool active = inside && (mask == nullptr || mask[pix]);
    float* dst = out + out_group_stride * blockIdx.y;

    if (!__syncthreads_or(active)) {          // whole tile outside the mask
        if (inside) dst[pix] = 0.f;
        return;
    }

    const int v_begin = blockIdx.y * views_per_group;
    const int v_end = min(nviews, v_begin + views_per_group);
    const float x = inside ? xds[ix] : 0.f;
    const float y = inside ? yds[iy] : 0.f;
    const float nb1 = (float)nb;               // valid: 1 <= bb < nb
    // the tile's box of pixel centres
    float bx0 = xds[tx0], bx1 = bx0, by0 = yds[ty0], by1 = by0;
    for (int k = 1; k < kTile; ++k) {
        const float xv = xds[min(tx0 + k, nx - 1)], yv = yds[min(ty0 + k, ny - 1)];
        bx0 = fminf(bx0, xv); bx1 = fmaxf(bx1, xv);
        by0 = fminf(by0, yv); by1 = fmaxf(by1, yv);
    }
    float acc = 0.f;

    for (int c0 = v_begin; c0 < v_end; c0 += vc) {
        const int nv = min(vc, v_end - c0);
        __syncthreads();                       // the previous chunk is consumed
        if (tid < nv) {
            const ViewCS t = cs[c0 + tid];
            float bmin = 3.4e38f, bmax = -3.4e38f;
            bool ok = true;
            for (int k = 0; k < 4; ++k) {
                const float cx = (k & 1) ? bx1 : bx0, cy = (k & 2) ? by1 : by0;
                if (GEOM != 0) ok = ok && dso_ds + cx * t.s - cy * t.c > 0.f;   // d > 0
                float bb, w2;
                pixel_bin<GEOM>(cx, cy, t, wb, dsd_ds, dso_ds, so_ds, bb, w2);
                bmin = fminf(bmin, bb);
                bmax = fmaxf(bmax, bb);
            }
            int lo = 0, n = 0;
            if (ok && bmax >= 1.f && bmin < nb1) {      // false for NaN
                lo = max((int)floorf(fmaxf(bmin, 0.f)) - 2, 0);
                const int hi = min((int)floorf(fminf(bmax, nb1)) + 1, nb - 1);
                n = hi - lo + 1;
                if (n > wm) n = 0;                       // too wide: taps from global memory
            }
            win[tid] = make_int2(lo + 1, n - 1);
        }
        __syncthreads();
        const int total = nv << wshift;
        for (int i0 = tid; i0 < total; i0 += kThreads * kStage) {
            float t[kStage];
#pragma unroll
            for (int u = 0; u < kStage; ++u) {
                const int i = i0 + u * kThreads;
                const int v = i >> wshift, k = i & (wm - 1);
                t[u] = 0.f;
                if (i < total) {
                    const int2 w = win[v];
                    if (k <= w.y) t[u] = q[view_stride * (int64_t)(c0 + v) + (w.x - 1) + k];
                }
            }
#pragma unroll
            for (int u = 0; u < kStage; ++u) {
                const int i = i0 + u * kThreads;
                if (i < total) lds[i] = t[u];
            }
        }
        __syncthreads();
        if (active) {
            // kU views at a time, their LDS reads in flight together; the
            // last batch of a chunk repeats view nv - 1 with weight 0.  Out of
            // range (1 <= floor(bb) <= nb - 1 fails): bin 1, weight 0.  Taps
            // outside a staged window are reloaded from global memory in a
            // branch taken only when some lane of the wave needs it.  The sum
            // keeps the view order.
            for (int v = 0; v < nv; v += kU) {
                float w[kU], wr[kU], a[kU], b[kU];
                int il[kU];
                uint32_t miss = 0;
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int vu = min(v + u, nv - 1);
                    float bb, w2;
                    pixel_bin<GEOM>(x, y, cs[c0 + vu], wb, dsd_ds, dso_ds, so_ds, bb, w2);
                    const bool ok = bb >= 1.f && bb < nb1;
                    w[u] = (ok && v + u < nv) ? w2 : 0.f;
                    const float bbc = ok ? bb : 1.f;
                    const float ilf = floorf(bbc);
                    wr[u] = bbc - ilf;
                    il[u] = (int)ilf;                     // taps il - 1 and il (0-based)
                    const int2 wv = win[vu];
                    const int local = il[u] - wv.x;      // il - 1 - lo
                    const bool in = local >= 0 && local < wv.y;
                    miss |= (uint32_t)(!in) << u;
                    const float* p = lds + (vu << wshift) + (in ? local : 0);
                    a[u] = p[0];
                    b[u] = p[1];
                }
                if (__any(miss != 0u)) {
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        if ((miss >> u) & 1u) {
                            const float* qv = q + view_stride * (int64_t)(c0 + min(v + u, nv - 1));
                            const uint32_t o = (uint32_t)(il[u] - 1);
                            a[u] = qv[o];
                            b[u] = qv[o + 1u];
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < kU; ++u) acc = fmaf(w[u], fmaf(wr[u], b[u] - a[u], a[u]), acc);
            }
        }
    }
    if (inside) dst[pix] = active ? acc : 0.f;
}

// Exact transpose of back2d_kernel:
//   qt[b, v] = sum over active pixels p of img[p] w2 ((1 - wr) [b = il - 1] + wr [b = il])
// Block (view v, pixel group g): bins accumulate in LDS (atomics), then are
// added into qt[:, v] (global atomics when a view has several groups; qt is
// zeroed first).  Float atomics: the sum order varies from run to run.
template <int GEOM>
__global__ __launch_bounds__(kThreads) void back2d_t_kernel(
    const float* __restrict__ img, int nb, int nviews, const ViewCS* __restrict__ cs,
    const float* __restrict__ xds, const float* __restrict__ yds, int nx, int ny,
    const uint8_t* __restrict__ mask, float* __restrict__ qt, float wb, float dsd_ds,
    float dso_ds, float so_ds)
{
    extern __shared__ float bins[];
    const int v = blockIdx.x;
    const int64_t npix = (int64_t)nx * ny;
    for (int b = threadIdx.x; b < nb; b += kThreads) bins[b] = 0.f;
    __syncthreads();
    const ViewCS t = cs[v];
    const float nb1 = (float)nb;
    for (int64_t p = blockIdx.y * (int64_t)kThreads + threadIdx.x; p < npix;
         p += (int64_t)gridDim.y * kThreads) {
        if (mask != nullptr && !mask[p]) continue;
        const float g = img[p];
        if (g == 0.f) continue;
        const int ix = (int)(p % nx), iy = (int)(p / nx);
        float bb, w2;
        pixel_bin<GEOM>(xds[ix], yds[iy], t, wb, dsd_ds, dso_ds, so_ds, bb, w2);
        if (!(bb >= 1.f && bb < nb1)) continue;        // 1 <= floor(bb) <= nb - 1
        const float ilf = floorf(bb);
        const float wr = bb - ilf;
        const int il = (int)ilf;
        const float gw = g * w2;
        atomicAdd(&bins[il - 1], gw * (1.f - wr));
        atomicAdd(&bins[il], gw * wr);
    }
    __syncthreads();
    float* dst = qt + (int64_t)nb * v;
    for (int b = threadIdx.x; b < nb; b += kThreads) {
        if (gridDim.y == 1) dst[b] = bins[b];
        else if (bins[b] != 0.f) atomicAdd(&dst[b], bins[b]);
    }
}

__global__ void zero_f32_kernel(float* p, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

__global__ void sum_partials_kernel(const float* __restrict__ part, int groups, int64_t n,
                                    const uint8_t* __restrict__ mask, float* __restrict__ image)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int g = 0; g < groups; ++g) s += part[g * n + i];
    image[i] = (mask == nullptr || mask[i]) ? s : 0.f;
}

int groups_for(int nb, int nviews, int nx, int ny, int want_blocks, int vc_max)
{
    const int tiles = ((nx + kTile - 1) / kTile) * ((ny + kTile - 1) / kTile);
    const int vc = chunk_views(nb, vc_max);
    const int max_groups = (nviews + vc - 1) / vc;
    const int want = (want_blocks + tiles - 1) / tiles;
    return std::max(1, std::min(want, max_groups));
}

}  // namespace

int64_t back2d_partial_floats(int nb, int nviews, int nx, int ny, int want_blocks, int vc_max)
{
    const int g = groups_for(nb, nviews, nx, ny, want_blocks, vc_max);
    return g > 1 ? (int64_t)g * nx * ny : 0;
}

void launch_back2d(const Back2dArgs& a, hipStream_t st)
{
    if (a.want_blocks < 1 || a.vc_max < 1) throw Error(SGX_ERR_ARG, "back2d: bad shape");
    const int vc = chunk_views(a.nb, a.vc_max);
    int groups = groups_for(a.nb, a.nviews, a.nx, a.ny, a.want_blocks, a.vc_max);
    const int64_t npix = (int64_t)a.nx * a.ny;
    if (groups > 1 && (a.partial == nullptr || a.partial_cap < groups * npix)) groups = 1;
    // views per group: a whole number of LDS chunks
    const int chunks = (a.nviews + vc - 1) / vc;
    const int vpg = ((chunks + groups - 1) / groups) * vc;
    groups = (a.nviews + vpg - 1) / vpg;
    const int tiles = ((a.nx + kTile - 1) / kTile) * ((a.ny + kTile - 1) / kTile);
    float* out = groups > 1 ? a.partial : a.image;
    dim3 grid(tiles, groups), block(kThreads);
    const int wm = window_bins(a.nb);
    int wshift = 0;
    while ((1 << wshift) < wm) ++wshift;
    const size_t lds = (size_t)vc * wm * sizeof(float) + (size_t)vc * sizeof(int2);
    switch (a.geom) {
    case 0:
        hipLaunchKernelGGL(back2d_kernel<0>, grid, block, lds, st, a.q, a.nb, a.nviews,
                           a.view_stride, a.cs, a.xds, a.yds, a.nx, a.ny, a.mask, out, npix, vpg,
                           vc, wshift, a.wb, a.dsd_ds, a.dso_ds, a.so_ds);
        break;
    case 1:
        // bins 1 <= bb < nb are at |gamma| = |bb - wb| / dsd_ds; below 0.75 rad
        // (43 degrees) on both sides the narrow arctangent serves
        if (std::max(std::fabs(1.f - a.wb), std::fabs((float)a.nb - a.wb)) < 0.75f * a.dsd_ds)
            hipLaunchKernelGGL(back2d_kernel<3>, grid, block, lds, st, a.q, a.nb, a.nviews,
                               a.view_stride, a.cs, a.xds, a.yds, a.nx, a.ny, a.mask, out, npix,
                               vpg, vc, wshift, a.wb, a.dsd_ds, a.dso_ds, a.so_ds);
        else
            hipLaunchKernelGGL(back2d_kernel<1>, grid, block, lds, st, a.q, a.nb, a.nviews,
                               a.view_stride, a.cs, a.xds, a.yds, a.nx, a.ny, a.mask, out, npix,
                               vpg, vc, wshift, a.wb, a.dsd_ds, a.dso_ds, a.so_ds);
        break;
    default:
        hipLaunchKernelGGL(back2d_kernel<2>, grid, block, lds, st, a.q, a.nb, a.nviews,
                           a.view_stride, a.cs, a.xds, a.yds, a.nx, a.ny, a.mask, out, npix, vpg,
                           vc, wshift, a.wb, a.dsd_ds, a.dso_ds, a.so_ds);
        break;
    }
    SGX_HIP(hipGetLastError());
    if (groups > 1) {
        hipLaunchKernelGGL(sum_partials_kernel, dim3((unsigned)((npix + 255) / 256)), dim3(256),
                           0, st, a.partial, groups, npix, a.mask, a.image);
        SGX_HIP(hipGetLastError());
    }
}

void launch_back2d_t(const Back2dArgs& a, const float* img, float* qt, hipStream_t st)
{
    if ((int64_t)a.nb * 4 > kLdsBytes)
        throw Error(SGX_ERR_UNSUPPORTED, "back2d_t: nb > 12288 bins is not supported");
    const int64_t npix = (int64_t)a.nx * a.ny;
    // pixel groups per view so that views x groups fill the GPU
    const int groups = (int)std::max<int64_t>(1, std::min<int64_t>((2048 + a.nviews - 1) / a.nviews,
                                                                   (npix + 4095) / 4096));
    if (groups > 1) {
        const int64_t n = (int64_t)a.nb * a.nviews;
        hipLaunchKernelGGL(zero_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           qt, n);
        SGX_HIP(hipGetLastError());
    }
    dim3 grid((unsigned)a.nviews, (unsigned)groups), block(kThreads);
    const size_t lds = (size_t)a.nb * sizeof(float);
    switch (a.geom) {
    case 0:
        hipLaunchKernelGGL(back2d_t_kernel<0>, grid, block, lds, st, img, a.nb, a.nviews, a.cs,
                           a.xds, a.yds, a.nx, a.ny, a.mask, qt, a.wb, a.dsd_ds, a.dso_ds, a.so_ds);
        break;
    case 1:
        hipLaunchKernelGGL(back2d_t_kernel<1>, grid, block, lds, st, img, a.nb, a.nviews, a.cs,
                           a.xds, a.yds, a.nx, a.ny, a.mask, qt, a.wb, a.dsd_ds, a.dso_ds, a.so_ds);
        break;
    default:
        hipLaunchKernelGGL(back2d_t_kernel<2>, grid, block, lds, st, img, a.nb, a.nviews, a.cs,
                           a.xds, a.yds, a.nx, a.ny, a.mask, qt, a.wb, a.dsd_ds, a.dso_ds, a.so_ds);
        break;
    }
    SGX_HIP(hipGetLastError());
}

}  // namespace sgx
