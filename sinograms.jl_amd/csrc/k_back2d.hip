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
// with a bin of margin, its start rounded down to 4 bins; only the windows
// are staged in LDS, in runs of 32 bins per thread (32 loads in flight, runs
// past a window's end skipped).  The two taps of each update are then LDS reads; a lane
// whose taps fall outside its view's window (windows wider than the LDS
// slot, pixels at or behind the source) reads them from global memory, so
// the result never depends on the window.  Per-view cos/sin are wave-uniform
// scalar loads and the sum is kept in a register, views in order.  When there
// are too few pixel tiles to fill the GPU the views are split over G groups;
// each group writes a partial image and a second kernel adds the G partials
// in a fixed order (deterministic, no atomics).
//
// Arc fan, tile-relative form (GEOM 4).  The bin coordinate of an arc fan is
// bb = dsd theta + wb with theta = atan2(r, d): about 440 bins at the edge of a
// 888-bin detector, where FP32 resolves only 3e-5 of a bin.  Per (tile, view)
// one thread forms, in FP64, the tile centre's (d0, r0), theta0 =
// atan2(r0, d0) and K = dsd theta0 + wb = I0 + f0 (I0 an integer).  A pixel at
// (dx, dy) from the centre has, exactly,
//   t = tan(theta - theta0) = (r d0 - r0 d) / (d d0 + r r0)
//     = (dx A + dy B) / (R2 + E),  E = dx B - dy A,
//   A = c d0 - s r0,  B = s d0 + c r0,  R2 = d0^2 + r0^2,
// so bb = I0 + (f0 + dsd atan(t)), with |t| small (a tile seen from the
// source) and atan(t) its odd Taylor series of NT terms (NT chosen on the host
// from a bound on |t|).  The fractional coordinate f0 + dsd atan(t) is a few
// bins, so FP32 keeps it to about 1e-6 of a bin.  The weight
// w2 = dsd^2 / (d^2 + r^2) comes without a second reciprocal: by Lagrange's
// identity num^2 + den^2 = (d^2 + r^2) R2, so w2 = dsd^2 R2 / den^2 / (1 + t^2),
// with 1 / (1 + t^2) its series of NT terms (KR folded into its
// coefficients).  The inner loop issues 20 VALU instructions per update (NT =
// 3) against about 36 for the per-pixel atan2 form; the per-pixel form stays
// for geometries where |t| is not small, and for sgx_plan_back /
// sgx_plan_back_t, whose pair is an exact transpose.
//
// Per-update LDS traffic: the two taps (one ds_read2) and the view's
// constants (one 16-byte and one 8-byte read shared by the wave).  Tap
// addresses are absolute LDS addresses (no add of the dynamic LDS base per
// update; the block reductions use words of the dynamic LDS, not the
// runtime's static buffer).

#include "sgx_internal.h"
#include "sgx.h"

#include <algorithm>
#include <cmath>

namespace sgx {
namespace {

constexpr int kTile = 16;
constexpr int kThreads = kTile * kTile;
constexpr int kLdsBytes = 48 * 1024;
// Views per unrolled step of the backprojector, and a minimum of waves per
// SIMD for the launch bound (1: none).  Measured on C1 (888 x 984 -> 512^2):
// 8 views per step take 160 VGPRs (3 waves per SIMD), 0.188 ms; 4 views 125
// VGPRs (4 waves, the most the 36 KB of LDS per block allows), 0.162 ms; 4
// views held to 5 waves 96 VGPRs with 10 spilled, 0.166 ms; 2 views at 5
// waves 0.176 ms.
#ifndef SGX_B2D_U
#define SGX_B2D_U 4
#endif
#ifndef SGX_B2D_WAVES
#define SGX_B2D_WAVES 1
#endif
constexpr int kU = SGX_B2D_U;   // views per unrolled step of the backprojector
constexpr int kWinMax = 128;    // staged window slot per view (bins), a power of two
constexpr int kRunBins = 32;    // staging: bins per thread and run (loads in flight together)

// Raw buffer resource over `bytes` bytes at p: loads past the end return 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t sino_rsrc(const float* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), 0, (int)bytes, 0x00020000);
}

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

// Window slot per view: the smallest power of two >= nb, at least kRunBins
// (one staging run) and at most kWinMax.
int window_bins(int nb)
{
    int w = kRunBins;
    while (w < nb && w < kWinMax) w *= 2;
    return w;
}

// Per-(tile, view) constants of the tile-relative arc form (see above), 32
// bytes: A, B, R2, f0; KR = dsd^2 R2; the in-detector range of the fractional
// coordinate [blo, bhi) = [1 - I0, nb - I0) (so I0 = 1 - blo); base = the LDS
// byte address of tap il - 1 when floor(f0 + dsd atan(t)) = 0.
struct alignas(16) RelView {
    float A, B, R2, f0, KR, blo, bhi;
    int base;
};
static_assert(sizeof(RelView) == 32, "RelView");

// Block-wide reduction words after the per-view tables of the dynamic LDS
// (the runtime's reductions would take static LDS ahead of the dynamic block,
// and every LDS address of the inner loop would carry that offset).
constexpr int kFlagBytes = 16;
// Window slots are wm + kSlotPad floats apart: the staging writes of lanes
// that move the same run of consecutive slots then start 16 bytes apart in
// the LDS banks (no conflicts), where a power-of-two pitch put them all on
// one bank.
constexpr int kSlotPad = 4;

// Dynamic LDS of a chunk of vc views with slots of wm bins.
size_t chunk_lds(int vc, int wm)
{
    return (size_t)vc * ((wm + kSlotPad) * sizeof(float) + sizeof(int2) + sizeof(RelView)) +
           kFlagBytes;
}

// Views of one LDS chunk: at most vc_max and what fits in kLdsBytes with the
// per-view tables; a multiple of kU once it reaches kU.
int chunk_views(int nb, int vc_max)
{
    const int per = 4 * (window_bins(nb) + kSlotPad) + 8 + (int)sizeof(RelView);
    int vc = std::max(1, std::min(vc_max, (kLdsBytes - kFlagBytes) / per));
    if (vc >= kU) vc -= vc % kU;
    return vc;
}

// Coefficients of dsd atan(t) = dsd (t - t^3/3 + t^5/5 - ...), NT terms.
template <int NT> struct RelPoly {
    float c[NT];
    __device__ explicit RelPoly(float dsd_ds)
    {
#pragma unroll
        for (int k = 0; k < NT; ++k) c[k] = dsd_ds / (float)(2 * k + 1) * ((k & 1) ? -1.f : 1.f);
    }
    // f0 + dsd atan(t) (Horner in t^2)
    __device__ __forceinline__ float bin(float t, float t2, float f0) const
    {
        float p = c[NT - 1];
#pragma unroll
        for (int k = NT - 2; k >= 0; --k) p = fmaf(t2, p, c[k]);
        return fmaf(t, p, f0);
    }
    // K / (1 + t^2) = K (1 - t^2 + t^4 - ...), NT terms (the signs are
    // source modifiers of the FMAs)
    __device__ __forceinline__ static float inv1p(float t2, float K)
    {
        float p = (NT & 1) ? K : -K;
#pragma unroll
        for (int k = NT - 2; k >= 0; --k) p = fmaf(t2, p, (k & 1) ? -K : K);
        return p;
    }
};

// LDS byte address of a pointer into the block's LDS, and a float pointer
// in the LDS address space from one: the inner loop forms tap addresses from
// absolute LDS addresses, so no add of the dynamic LDS base per update.
using LdsFloat = __attribute__((address_space(3))) const float;
__device__ __forceinline__ int lds_byte(const void* p)
{
    return (int)(uint32_t)reinterpret_cast<uintptr_t>(p);
}
__device__ __forceinline__ LdsFloat* lds_ptr(int byte)
{
    return reinterpret_cast<LdsFloat*>((uint32_t)byte);
}

// floor(x) as an integer in one instruction (v_cvt_flr_i32_f32) for the
// small, finite coordinates of the tile-relative form
__device__ __forceinline__ int cvt_floor(float x)
{
    int r;
    asm volatile("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// GEOM: 0 parallel, 1 fan arc, 2 fan flat, 3 fan arc whose detector spans
// less than +-45 degrees (atan2_narrow), 4 fan arc in the tile-relative form
// with NT series terms (NT is used by GEOM 4 only)
// One block's work: pixel tile `tile` over view group `group` (views
// group * views_per_group ..), with the block's dynamic LDS `lds`.
template <int GEOM, int NT>
__device__ __forceinline__ void back2d_tile(
    const float* __restrict__ q, int nb, int nviews, int64_t view_stride,
    const ViewCS* __restrict__ cs, const float* __restrict__ xds,
    const float* __restrict__ yds, int nx, int ny, const uint8_t* __restrict__ mask,
    float* __restrict__ out, int64_t out_group_stride, int views_per_group, int vc, int wshift,
    float wb, float dsd_ds, float dso_ds, float so_ds, int tile, int group, float* lds)
{
    // vc window slots of wm bins (sp floats apart), then per view int2 {lo + 1, n - 1}: the
    // window holds 0-based bins lo .. lo + n - 1 (n = 0: nothing staged),
    // then (GEOM 4) per view a RelView
    constexpr bool REL = GEOM == 4;
    const int wm = 1 << wshift, sp = wm + kSlotPad;
    int2* win = reinterpret_cast<int2*>(lds + vc * sp);
    RelView* rel = reinterpret_cast<RelView*>(win + vc);
    int* flags = reinterpret_cast<int*>(rel + vc);   // [0]: some pixel active, [1]: chunk interior
    const int tid = threadIdx.x;
    const int tiles_x = (nx + kTile - 1) / kTile;
    const int tx0 = (tile % tiles_x) * kTile, ty0 = (tile / tiles_x) * kTile;
    const int ix = tx0 + (tid % kTile);
    const int iy = ty0 + (tid / kTile);
    const int64_t pix = ix + (int64_t)nx * iy;
    const bool inside = ix < nx && iy < ny;
    const bool active = inside && (mask == nullptr || mask[pix]);
    float* dst = out + out_group_stride * group;

    // whole tile outside the mask: nothing to add
    if (tid == 0) flags[0] = 0;
    __syncthreads();
    if (active) flags[0] = 1;
    __syncthreads();
    if (!flags[0]) {
        if (inside) dst[pix] = 0.f;
        return;
    }

    const int v_begin = group * views_per_group;
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
    // GEOM 4: the pixel relative to the tile centre
    const float xt = 0.5f * (bx0 + bx1), yt = 0.5f * (by0 + by1);
    const float dx = x - xt, dy = y - yt;
    const RelPoly<NT> poly(dsd_ds);
    float acc = 0.f;
    const float ndy = -dy;
    const int lds0 = lds_byte(lds);

    for (int c0 = v_begin; c0 < v_end; c0 += vc) {
        const int nv = min(vc, v_end - c0);
        // GEOM 4 pads the chunk to a multiple of kU views of zero data and weight
        const int nvp = REL ? min(vc, (nv + kU - 1) / kU * kU) : nv;
        __syncthreads();                       // the previous chunk is consumed
        bool interior = true;                  // GEOM 4: this view needs no per-lane tests
        if (tid < nvp) {
            int lo = 0, n = 0;
            float bmin = 3.4e38f, bmax = -3.4e38f;
            const bool real = tid < nv;
            const ViewCS t = cs[c0 + min(tid, nv - 1)];
            bool ok = real;
            for (int k = 0; k < 4; ++k) {
                const float cx = (k & 1) ? bx1 : bx0, cy = (k & 2) ? by1 : by0;
                if (GEOM != 0) ok = ok && dso_ds + cx * t.s - cy * t.c > 0.f;   // d > 0
                float bb, w2;
                pixel_bin<REL ? 1 : GEOM>(cx, cy, t, wb, dsd_ds, dso_ds, so_ds, bb, w2);
                bmin = fminf(bmin, bb);
                bmax = fmaxf(bmax, bb);
            }
            if (ok && bmax >= 1.f && bmin < nb1) {      // false for NaN
                // the window starts at a multiple of 4 bins (16-byte staging loads)
                lo = max((int)floorf(fmaxf(bmin, 0.f)) - 2, 0) & ~3;
                const int hi = min((int)floorf(fminf(bmax, nb1)) + 1, nb - 1);
                n = hi - lo + 1;
                if (n > wm) n = 0;                       // too wide: taps from global memory
            }
            win[tid] = make_int2(lo + 1, n - 1);
            if (REL) {
                RelView r;
                if (real) {
                    // the tile centre in FP64 from the FP32 inputs
                    const double c = t.c, sn = t.s;
                    const double d0 = (double)dso_ds + (double)xt * sn - (double)yt * c;
                    const double r0 = (double)xt * c + (double)yt * sn - (double)so_ds;
                    const double K = (double)dsd_ds * atan2(r0, d0) + (double)wb;
                    const double I0 = floor(K);
                    r.A = (float)(c * d0 - sn * r0);
                    r.B = (float)(sn * d0 + c * r0);
                    r.R2 = (float)(d0 * d0 + r0 * r0);
                    r.KR = (float)((double)dsd_ds * (double)dsd_ds * (d0 * d0 + r0 * r0));
                    r.f0 = (float)(K - I0);
                    r.blo = (float)(1.0 - I0);
                    r.bhi = (float)((double)nb - I0);
                    r.base = lds_byte(lds) + 4 * (tid * sp + ((int)I0 - 1 - lo));
                    // every pixel's bb is inside [bmin, bmax] up to FP32 rounding
                    // of the corners; one bin of margin on each side
                    interior = ok && n > 0 && bmin >= 2.f && bmax < nb1 - 1.f &&
                               (int)floorf(bmin) - 2 >= 0 && (int)floorf(bmax) + 1 <= nb - 1;
                } else {
                    // padding view: zero data (n = 0), zero weight
                    r.A = r.B = 0.f;
                    r.R2 = 1e30f;
                    r.KR = 0.f;
                    r.f0 = 0.5f;
                    r.blo = r.bhi = 0.f;      // I0 = 1
                    r.base = lds_byte(lds) + 4 * (tid * sp);
                }
                rel[tid] = r;
            }
        }
        // GEOM 4: the chunk is interior when all its views are (a reduction
        // barrier only where it is used)
        bool chunk_interior = false;
        if constexpr (REL) {
            // flags[1] was last read before the barrier that opened this chunk
            if (tid == 0) flags[1] = 1;
            __syncthreads();
            if (!interior) flags[1] = 0;
            __syncthreads();
            chunk_interior = __builtin_amdgcn_readfirstlane(flags[1]) != 0;
        } else {
            __syncthreads();
        }
        // Staging: slot v holds bins lo .. lo + wm - 1 of view c0 + v (lo a
        // multiple of 4, wm a multiple of kRunBins).  A thread moves one run
        // of kRunBins bins of a slot: dword loads at immediate offsets, all in
        // flight together (rows need no alignment), then 16-byte LDS writes.
        // Runs past a window's end are skipped; bins past a window's end
        // within a run are data of the row or of the next view (0 past the
        // sinogram's end), which no tap reads.  The padding views of the
        // tile-relative form read bins 0 and 1 of their slot with weight 0:
        // those are zeroed.
        {
            const int rshift = wshift - 5;                       // runs per slot: 2^rshift
            int vbits = 0;                                       // slots: at most 2^vbits
            while ((1 << vbits) < nvp) ++vbits;
            const int64_t left = (int64_t)(nviews - 1 - c0) * view_stride + nb;  // floats
            const __amdgpu_buffer_rsrc_t rs =
                sino_rsrc(q + view_stride * (int64_t)c0,
                          (uint32_t)std::min<int64_t>(left * 4, 0x7fffffff));
            // consecutive lanes take the same run of consecutive slots
            for (int rr = tid; rr < (1 << (vbits + rshift)); rr += kThreads) {
                const int v = rr & ((1 << vbits) - 1);
                const int k0 = (rr >> vbits) * kRunBins;
                if (v >= nvp) continue;
                float* dst = lds + v * sp + k0;
                if (REL && v >= nv) {
                    if (k0 == 0) *reinterpret_cast<float4*>(dst) = make_float4(0.f, 0.f, 0.f, 0.f);
                    continue;
                }
                const int2 w = win[v];
                if (k0 > w.y) continue;                          // past the window (or none)
                const uint32_t off =
                    (uint32_t)((view_stride * (int64_t)v + (w.x - 1) + k0) * 4);
                float t[kRunBins];
#pragma unroll
                for (int j = 0; j < kRunBins; ++j)
                    t[j] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off + 4 * j, 0, 0));
#pragma unroll
                for (int j = 0; j < kRunBins; j += 4)
                    *reinterpret_cast<float4*>(dst + j) = make_float4(t[j], t[j + 1], t[j + 2], t[j + 3]);
            }
        }
        __syncthreads();
        if (!active) continue;
        if (REL && chunk_interior) {
            // every tap of every lane is in its view's window and on the detector
            for (int v = 0; v < nv; v += kU) {
                float w[kU], wr[kU], a[kU], b[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const RelView r = rel[v + u];
                    const float den = fmaf(dx, r.B, fmaf(ndy, r.A, r.R2));   // R2 + E
                    const float num = fmaf(dx, r.A, dy * r.B);
                    const float rt = __builtin_amdgcn_rcpf(den);
                    const float tt = num * rt;
                    const float t2 = tt * tt;
                    const float bbf = poly.bin(tt, t2, r.f0);
                    wr[u] = __builtin_amdgcn_fractf(bbf);
                    LdsFloat* p = lds_ptr(r.base + 4 * cvt_floor(bbf));
                    a[u] = p[0];
                    b[u] = p[1];
                    w[u] = RelPoly<NT>::inv1p(t2, r.KR) * rt * rt;
                }
#pragma unroll
                for (int u = 0; u < kU; ++u) acc = fmaf(w[u], fmaf(wr[u], b[u] - a[u], a[u]), acc);
            }
        } else if (REL) {
            // per-lane tests: bins off the detector weigh 0, taps outside the
            // window come from global memory
            for (int v = 0; v < nv; v += kU) {
                float w[kU], wr[kU], a[kU], b[kU];
                int il[kU];
                uint32_t miss = 0;
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const RelView r = rel[v + u];
                    const float den = fmaf(dx, r.B, fmaf(ndy, r.A, r.R2));   // R2 + E
                    const float num = fmaf(dx, r.A, dy * r.B);
                    const float rt = __builtin_amdgcn_rcpf(den);
                    const float tt = num * rt;
                    const float t2 = tt * tt;
                    const float bbf = poly.bin(tt, t2, r.f0);
                    const bool ok = bbf >= r.blo && bbf < r.bhi;   // 1 <= floor(bb) < nb
                    w[u] = ok ? RelPoly<NT>::inv1p(t2, r.KR) * rt * rt : 0.f;
                    const float bbc = ok ? bbf : r.blo;
                    const int fi = (int)floorf(bbc);
                    wr[u] = __builtin_amdgcn_fractf(bbc);
                    il[u] = (int)(1.f - r.blo) + fi;              // taps il - 1 and il (0-based)
                    const int local = ((r.base - lds0) >> 2) - (v + u) * sp + fi;
                    const int2 wv = win[v + u];
                    const bool in = local >= 0 && local < wv.y;
                    miss |= (uint32_t)(!in) << u;
                    const float* p = lds + (v + u) * sp + (in ? local : 0);
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
        } else {
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
                    const float* p = lds + vu * sp + (in ? local : 0);
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

template <int GEOM, int NT = 1>
__global__ __launch_bounds__(kThreads, SGX_B2D_WAVES) void back2d_kernel(
    const float* __restrict__ q, int nb, int nviews, int64_t view_stride,
    const ViewCS* __restrict__ cs, const float* __restrict__ xds,
    const float* __restrict__ yds, int nx, int ny, const uint8_t* __restrict__ mask,
    float* __restrict__ out, int64_t out_group_stride, int views_per_group, int vc, int wshift,
    float wb, float dsd_ds, float dso_ds, float so_ds)
{
    extern __shared__ float lds[];
    back2d_tile<GEOM, NT>(q, nb, nviews, view_stride, cs, xds, yds, nx, ny, mask, out,
                          out_group_stride, views_per_group, vc, wshift, wb, dsd_ds, dso_ds,
                          so_ds, blockIdx.x, blockIdx.y, lds);
}

// Batch form: one launch over the tasks (item, pixel tile) of many
// reconstructions; block b does task b of the table (ordered by the host,
// largest tasks first, interleaved over the XCDs); a task backprojects
// every view of its tile (no view groups) into the item's image.
template <int GEOM, int NT>
__global__ __launch_bounds__(kThreads, SGX_B2D_WAVES) void back2d_queue_kernel(
    const Back2dItem* __restrict__ items, const int2* __restrict__ tasks)
{
    extern __shared__ float lds[];
    const int2 tk = tasks[blockIdx.x];
    using CItem = const __attribute__((address_space(4))) Back2dItem;
    const CItem& it = ((CItem*)(items))[tk.x];
    back2d_tile<GEOM, NT>(it.q, it.nb, it.nviews, it.view_stride, it.cs, it.xds, it.yds, it.nx,
                          it.ny, it.mask, it.image, 0, it.nviews, it.vc, it.wshift, it.wb,
                          it.dsd_ds, it.dso_ds, it.so_ds, tk.y, 0, lds);
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

int back2d_rel_terms(const Back2dArgs& a)
{
    if (a.geom != 1 || a.pixel_form || a.rel_rmax < 0.f || a.rel_h < 0.f) return 0;
    if (chunk_views(a.nb, a.vc_max) % kU != 0) return 0;
    // The source is at distance >= dso from the origin in every view, so a
    // tile centre is at least dso - rmax from it and its pixels (within
    // rel_h of the centre) subtend |theta - theta0| <= asin(z); t = tan of it.
    const double rho = (double)a.dso_ds - (double)a.rel_rmax;
    if (!(rho > 0.0)) return 0;
    const double z = (double)a.rel_h / rho;
    if (!(z <= 0.25)) return 0;
    const double T = z / std::sqrt(1.0 - z * z);
    // the first omitted terms: of dsd atan(t) below 1e-6 of a bin, of
    // 1 / (1 + t^2) below 1e-6 relative
    for (int nt : {3, 4, 5, 6, 8})
        if ((double)a.dsd_ds * std::pow(T, 2 * nt + 1) / (2 * nt + 1) <= 1e-6 &&
            std::pow(T, 2 * nt) <= 1e-6)
            return nt;
    return 0;
}

// Kernel variant of a 2-D backprojection: GEOM * 16 + NT.
int back2d_variant(const Back2dArgs& a)
{
    switch (a.geom) {
    case 0:
        return 0 * 16 + 1;
    case 1: {
        const int nt = back2d_rel_terms(a);
        if (nt > 0) return 4 * 16 + nt;
        // bins 1 <= bb < nb are at |gamma| = |bb - wb| / dsd_ds; below 0.75 rad
        // (43 degrees) on both sides the narrow arctangent serves
        if (std::max(std::fabs(1.f - a.wb), std::fabs((float)a.nb - a.wb)) < 0.75f * a.dsd_ds)
            return 3 * 16 + 1;
        return 1 * 16 + 1;
    }
    default:
        return 2 * 16 + 1;
    }
}

Back2dItem back2d_item(const Back2dArgs& a)
{
    Back2dItem it{};
    it.q = a.q;
    it.cs = a.cs;
    it.xds = a.xds;
    it.yds = a.yds;
    it.mask = a.mask;
    it.image = a.image;
    it.view_stride = a.view_stride;
    it.nb = a.nb;
    it.nviews = a.nviews;
    it.nx = a.nx;
    it.ny = a.ny;
    it.vc = chunk_views(a.nb, 64);
    const int wm = window_bins(a.nb);
    int wshift = 0;
    while ((1 << wshift) < wm) ++wshift;
    it.wshift = wshift;
    it.wb = a.wb;
    it.dsd_ds = a.dsd_ds;
    it.dso_ds = a.dso_ds;
    it.so_ds = a.so_ds;
    return it;
}

int back2d_item_tiles_x(const Back2dItem& it) { return (it.nx + kTile - 1) / kTile; }

int back2d_item_tiles(const Back2dItem& it)
{
    return ((it.nx + kTile - 1) / kTile) * ((it.ny + kTile - 1) / kTile);
}

size_t back2d_item_lds(const Back2dItem& it)
{
    return chunk_lds(it.vc, 1 << it.wshift);
}

template <int GEOM, int NT>
static const void* b2q_fn()
{
    return reinterpret_cast<const void*>(&back2d_queue_kernel<GEOM, NT>);
}

static const void* back2d_queue_fn(int variant)
{
    switch (variant) {
    case 0 * 16 + 1: return b2q_fn<0, 1>();
    case 1 * 16 + 1: return b2q_fn<1, 1>();
    case 2 * 16 + 1: return b2q_fn<2, 1>();
    case 3 * 16 + 1: return b2q_fn<3, 1>();
    case 4 * 16 + 3: return b2q_fn<4, 3>();
    case 4 * 16 + 4: return b2q_fn<4, 4>();
    case 4 * 16 + 5: return b2q_fn<4, 5>();
    case 4 * 16 + 6: return b2q_fn<4, 6>();
    case 4 * 16 + 8: return b2q_fn<4, 8>();
    default: throw Error(SGX_ERR_ARG, "back2d: unknown kernel variant");
    }
}

void launch_back2d_queue(int variant, const Back2dItem* items, const int2* tasks, int ntasks,
                         size_t lds, hipStream_t st)
{
    if (ntasks <= 0) return;
    if (lds > (size_t)kLdsBytes) throw Error(SGX_ERR_ARG, "back2d queue: LDS");
    void* args[] = {&items, &tasks};
    SGX_HIP(hipLaunchKernel(back2d_queue_fn(variant), dim3((unsigned)ntasks), dim3(kThreads), args,
                            lds, st));
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
    const size_t lds = chunk_lds(vc, wm);
#define SGX_B2D(G, NT)                                                                          \
    hipLaunchKernelGGL((back2d_kernel<G, NT>), grid, block, lds, st, a.q, a.nb, a.nviews,       \
                       a.view_stride, a.cs, a.xds, a.yds, a.nx, a.ny, a.mask, out, npix, vpg, vc, \
                       wshift, a.wb, a.dsd_ds, a.dso_ds, a.so_ds)
    switch (back2d_variant(a)) {
    case 0 * 16 + 1: SGX_B2D(0, 1); break;
    case 1 * 16 + 1: SGX_B2D(1, 1); break;
    case 2 * 16 + 1: SGX_B2D(2, 1); break;
    case 3 * 16 + 1: SGX_B2D(3, 1); break;
    case 4 * 16 + 3: SGX_B2D(4, 3); break;
    case 4 * 16 + 4: SGX_B2D(4, 4); break;
    case 4 * 16 + 5: SGX_B2D(4, 5); break;
    case 4 * 16 + 6: SGX_B2D(4, 6); break;
    case 4 * 16 + 8: SGX_B2D(4, 8); break;
    default: throw Error(SGX_ERR_ARG, "back2d: unknown kernel variant");
    }
#undef SGX_B2D
    SGX_HIP(hipGetLastError());
    if (groups > 1) {
        hipLaunchKernelGGL(sum_partials_kernel, dim3((unsigned)((npix + 255) / 256)), dim3(256),
                           0, st, a.partial, groups, npix, a.mask, a.image);
        SGX_HIP(hipGetLastError());
    }
}

void back2d_rel_bounds(const double* xc, int64_t nx, const double* yc, int64_t ny,
                       const uint8_t* mask, double ds, float& rmax, float& h)
{
    // rmax: the largest |(x, y)| / ds over the pixels that are backprojected
    double r2 = 0.0;
    for (int64_t j = 0; j < ny; ++j) {
        const double y = yc[j] / ds;
        for (int64_t i = 0; i < nx; ++i) {
            if (mask && !mask[i + nx * j]) continue;
            const double x = xc[i] / ds;
            r2 = std::max(r2, x * x + y * y);
        }
    }
    // h: the largest half-diagonal of a 16 x 16 tile's box of pixel centres
    auto half = [&](const double* c, int64_t n) {
        double m = 0.0;
        for (int64_t t0 = 0; t0 < n; t0 += kTile) {
            double lo = c[t0], hi = c[t0];
            for (int64_t k = t0; k < std::min<int64_t>(n, t0 + kTile); ++k) {
                lo = std::min(lo, c[k]);
                hi = std::max(hi, c[k]);
            }
            m = std::max(m, 0.5 * (hi - lo));
        }
        return m / std::fabs(ds);
    };
    const double hx = half(xc, nx), hy = half(yc, ny);
    // pixel centres are rounded to FP32 in the kernel: a small relative margin
    rmax = (float)(std::sqrt(r2) * (1.0 + 1e-6));
    h = (float)(std::sqrt(hx * hx + hy * hy) * (1.0 + 1e-6));
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
