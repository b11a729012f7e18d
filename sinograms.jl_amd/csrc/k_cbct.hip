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
// Options (off by default, commented out in the reference at :36-59, :157,
// :178-180, :188): t' = mag_t (z - zs[view]) with a per-view source z; for
// the arc detector mag_t = dsd / sqrt(r^2 + y_beta^2) (the ray meets the
// detector cylinder) instead of dsd / y_beta; and a lateral source offset,
// r = x_beta - source_offset / ds in place of x_beta in s' and w2 (as the
// commented arc branch has it, and as fbp_back_fan_xy applies it to both
// detectors, src/fbp2/back-fan.jl:303-311).  The t options enter as per-view,
// z-independent terms: bt = mz z + wtv with mz = mag_t / (dt/ds) and
// wtv = wt - mz zs (= wt exactly when zs = 0).
//
// Storage.  q is column-major (ns, nt, na) with a row pitch of `pitch`
// elements (>= ns), in FP32, bfloat16, or FP16 with a per-view scale folded
// into w2 (the 16-bit forms).
//
// Two kernels share the per-voxel arithmetic:
//
// * cbct_staged_kernel (default).  A block (4 waves) owns a 16 x 16 patch of
//   (x, y) columns and 16 voxels along z; every lane owns one column and the
//   block's 16 z.  For every view the block's footprint on the detector (a
//   window of at most kWmax bins x kHmax rows, taken from a per-(block, view)
//   table that cbct_window_kernel computes from the block's corners) is
//   staged in LDS as FP32 pairs (q[is, it], q[is, it+1] - q[is, it]), with
//   16-byte coalesced loads and the next view's window loaded while the
//   current view is backprojected (two LDS buffers, one barrier per view).
//   An update is then one ds_read2_b64 (taps is and is+1 of rows it and
//   it+1), an FMA per column for the t interpolation and two FMAs into the
//   accumulator with the s weights and w2 folded per view.  The texture
//   addresser sees about 12 wide loads per block and view instead of one
//   64-lane gather per 64 updates, which is what bounds the gather kernel.
//   A lane whose taps fall outside the staged window (which the corner bound
//   rules out up to rounding, or when the window is larger than the LDS
//   buffer) and lanes at the detector's top or bottom edge take the gather
//   path below for that view, so the result never depends on the window.
//
// * cbct_back_kernel (gather).  Each wave owns an 8 x 8 patch of columns and
//   8 voxels along z; taps are gathered from global memory, two row loads
//   per update through one buffer resource per view (the second row through
//   the SGPR offset).
//
// Both accumulate in FP32 registers and write each voxel once (no atomics).

#include "sgx_internal.h"
#include "sgx.h"

#include <hip/hip_fp16.h>

#include <cstdlib>

namespace sgx {
namespace {

// Storage formats of the filtered projections.
enum { F_F32 = 0, F_BF16 = 1, F_F16S = 2 };

template <int FMT> struct Fmt {
    static constexpr uint32_t esz = FMT == F_F32 ? 4 : 2;    // bytes per sample
    static constexpr uint32_t pad = FMT == F_F32 ? 0 : 4;    // readable bytes past a view
};

// Raw buffer resource over one view.  Loads at offsets at or beyond
// num_records return 0 without touching memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t view_rsrc(const char* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(p), 0, (int)bytes, 0x00020000);
}

// Byte offset that no view's buffer resource covers (every view is below
// 2 GiB): a buffer load there returns zeros.
constexpr uint32_t kMissOffset = 0x80000000u;

// 16-bit storage: the pair (is, is+1) starts at a 2-byte-aligned offset, so
// it is taken from the dword-aligned 8-byte window that contains it (one
// aligned load plus one v_alignbit).  The window can reach 4 bytes past the
// pair: every 16-bit buffer handed to these kernels has kQ16Pad bytes of
// padding after its last view and the views' resource covers 4 of them.
__device__ __forceinline__ uint32_t pair16(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                           uint32_t soff)
{
    const auto u = __builtin_amdgcn_raw_buffer_load_b64(r, voff & ~3u, soff, 0);
    return __builtin_amdgcn_alignbit(u[1], u[0], (voff & 2u) << 3);
}

__device__ __forceinline__ float bf16f(uint32_t h) { return __uint_as_float(h << 16); }
__device__ __forceinline__ float f16f(uint32_t h)
{
    return __half2float(__ushort_as_half((unsigned short)h));
}

template <int FMT> __device__ __forceinline__ float2 dec2(uint32_t u)
{
    if (FMT == F_BF16) return make_float2(bf16f(u & 0xffffu), bf16f(u >> 16));
    return make_float2(f16f(u & 0xffffu), f16f(u >> 16));
}

// Bin-pair difference a - b of two FP16 pairs (v_pk_add_f16 with a negated
// operand; each half rounded once).
__device__ __forceinline__ uint32_t half2_sub(uint32_t a, uint32_t b)
{
    const __half2 r = __hsub2(__builtin_bit_cast(__half2, a), __builtin_bit_cast(__half2, b));
    return __builtin_bit_cast(uint32_t, r);
}

// The four taps of one update from global memory as (a0, d0, a1, d1):
// a = q[is, it] and d = q[is, it+1] - q[is, it] for the columns is (0) and
// is+1 (1); voff is the byte offset of (is, it) in the view, rowb the row
// pitch in bytes (a multiple of 8 for the 16-bit forms).
template <int FMT>
__device__ __forceinline__ float4 taps(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t rowb)
{
    float2 lo, hi;
    if (FMT == F_F32) {
        const auto a = __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0);
        const auto b = __builtin_amdgcn_raw_buffer_load_b64(r, voff, rowb, 0);
        lo = make_float2(__uint_as_float(a[0]), __uint_as_float(a[1]));
        hi = make_float2(__uint_as_float(b[0]), __uint_as_float(b[1]));
    } else if (FMT == F_F16S) {
        // the row difference rounded in FP16, as the staged LDS entries hold
        // it: a voxel gets the same value whichever path serves its view
        const uint32_t a = pair16(r, voff, 0), d = half2_sub(pair16(r, voff, rowb), a);
        const float2 av = dec2<FMT>(a), dv = dec2<FMT>(d);
        return make_float4(av.x, dv.x, av.y, dv.y);
    } else {
        lo = dec2<FMT>(pair16(r, voff, 0));
        hi = dec2<FMT>(pair16(r, voff, rowb));
    }
    return make_float4(lo.x, hi.x - lo.x, lo.y, hi.y - lo.y);
}

// Per-view, z-independent part of one column's update.
struct ColView {
    bool hit;        // 1 <= floor(bs) < ns
    float isf;       // floor(bs)
    float wa, wb;    // w2 (1 - wr), w2 wr
    float mz;        // dbt / dz
    float wt;        // bt at z = 0: wt - mz zs
};

// mz and the t offset of one column and view (see the options above); OPT
// false compiles the options out (the reference's geometry).
template <bool ARC, bool OPT>
__device__ __forceinline__ void t_map(float rb, float yb, float mag, float dsd_ds,
                                      float inv_dt_ds, float wt, const float* zs, int ia,
                                      int arc_t_exact, float& mz, float& wtv)
{
    float m = mag;
    if (ARC && OPT && arc_t_exact) m = dsd_ds * __builtin_amdgcn_rsqf(rb * rb + yb * yb);
    mz = m * inv_dt_ds;
    wtv = wt;
    if (OPT && zs) wtv = fmaf(-mz, zs[ia], wt);
}

template <bool ARC, int FMT, bool OPT>
__device__ __forceinline__ ColView col_view(float x, float y, ViewCS v, float dsd_ds,
                                            float dso_ds, float ws, float nsf, float inv_dt_ds,
                                            float vs, float wt, const float* zs, int ia,
                                            int arc_t_exact, int narrow, float so_ds)
{
    ColView r;
    const float xb0 = x * v.c + y * v.s;
    const float xb = OPT ? xb0 - so_ds : xb0;          // r_loop (source offset)
    const float yb = dso_ds - (-x * v.s + y * v.c);
    const float mag = dsd_ds * __builtin_amdgcn_rcpf(yb);
    float sp, w2;
    if (ARC) {
        sp = dsd_ds * (narrow ? atan2_narrow(xb, yb) : atan2_fast(xb, yb));
        w2 = dsd_ds * dsd_ds * __builtin_amdgcn_rcpf(xb * xb + yb * yb);
    } else {
        sp = mag * xb;
        w2 = mag * mag;
    }
    if (FMT == F_F16S) w2 *= vs;
    const float bs = sp + ws;
    r.hit = bs >= 1.f && bs < nsf;          // 1 <= floor(bs) < ns (false for NaN)
    r.isf = floorf(bs);
    r.wb = w2 * (bs - r.isf);
    r.wa = w2 - r.wb;
    t_map<ARC, OPT>(xb, yb, mag, dsd_ds, inv_dt_ds, wt, zs, ia, arc_t_exact, r.mz, r.wt);
    return r;
}

// One view of ZB voxels from global memory.  `fast`: every voxel's bt is in
// [1, nt); otherwise voxels out of range load zeros (offset kMissOffset)
// and weigh 0.
template <int ZB, int FMT>
__device__ __forceinline__ void gather_view(const ColView& cv, const float* z, float wt,
                                            float ntf, bool fast, __amdgpu_buffer_rsrc_t rs,
                                            uint32_t rowb, float* acc)
{
    // byte offset of (is-1, it-1) is it*rowb + (is-1)*esz - rowb
    const uint32_t colb = ((uint32_t)cv.isf - 1u) * Fmt<FMT>::esz - rowb;
    constexpr int C = ZB < 8 ? ZB : 8;                 // voxels per batch of loads
#pragma unroll
    for (int h = 0; h < ZB; h += C) {
        float4 tp[C];
        float wu[C];
        if (fast) {
#pragma unroll
            for (int j = 0; j < C; ++j) {
                const float bt = fmaf(cv.mz, z[h + j], wt);
                wu[j] = __builtin_amdgcn_fractf(bt);
                tp[j] = taps<FMT>(rs, __umul24((uint32_t)bt, rowb) + colb, rowb);
            }
        } else {
#pragma unroll
            for (int j = 0; j < C; ++j) {
                const float bt = fmaf(cv.mz, z[h + j], wt);
                const bool ok = bt >= 1.f && bt < ntf;       // 1 <= floor(bt) < nt
                wu[j] = ok ? __builtin_amdgcn_fractf(bt) : 0.f;
                const uint32_t boff = ok ? __umul24((uint32_t)bt, rowb) + colb : kMissOffset;
                tp[j] = taps<FMT>(rs, boff, rowb);
            }
        }
#pragma unroll
        for (int j = 0; j < C; ++j) {
            const float p0 = fmaf(wu[j], tp[j].y, tp[j].x);
            const float p1 = fmaf(wu[j], tp[j].w, tp[j].z);
            acc[h + j] = fmaf(cv.wb, p1, fmaf(cv.wa, p0, acc[h + j]));
        }
    }
}

// ---------------------------------------------------------------------------
// gather kernel
// ---------------------------------------------------------------------------

constexpr int kWaves = 4;      // waves per block, each on its own z-slab
constexpr int kGatherVariant = 16;   // cbct_variant: the gather kernel
constexpr int kPackedVariant = 32;   // cbct_variant: the packed-FP16 kernel
constexpr int kMirrorVariant = 64;   // cbct_variant: the mirror form of the staged kernel
constexpr int kZ8Variant = 128;      // cbct_variant: the Z8 form of the staged kernel
constexpr int kZ8WVariant = 256;     // cbct_variant: the wide Z8 form of the staged kernel
constexpr int kPatch = 8;      // 8 x 8 columns per wave

// One block of the gather kernel: block `blk` (8 x 8 columns x 4 waves of
// ZB z).
template <int ZB, bool ARC, int FMT, bool OPT>
__device__ __forceinline__ void cbct_gather_block(
    const void* __restrict__ q, int ns, int nt, int pitch, int nviews, int64_t view_stride,
    const ViewCS* __restrict__ cs, const float* __restrict__ vscale,
    const float* __restrict__ xds, const float* __restrict__ yds,
    const float* __restrict__ zds, int nx, int ny, int nz, const uint8_t* __restrict__ mask,
    float* __restrict__ image, float dsd_ds, float dso_ds, float ws, float wt,
    float inv_dt_ds, const float* __restrict__ zs, int arc_t_exact, int narrow, float so_ds,
    int blk)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tiles_x = (nx + kPatch - 1) / kPatch;
    const int tiles = tiles_x * ((ny + kPatch - 1) / kPatch);
    const int tile = blk % tiles, zblk = blk / tiles;
    const int ix = (tile % tiles_x) * kPatch + (lane & (kPatch - 1));
    const int iy = (tile / tiles_x) * kPatch + (lane / kPatch);
    const int iz0 = (zblk * kWaves + wave) * ZB;
    if (ix >= nx || iy >= ny || iz0 >= nz) return;

    const int64_t nxy = (int64_t)nx * ny;
    const int64_t col = ix + (int64_t)nx * iy;

    float z[ZB], acc[ZB];
    uint32_t inbits = 0;           // bit j: voxel iz0 + j exists and is inside the mask
#pragma unroll
    for (int j = 0; j < ZB; ++j) {
        const int iz = iz0 + j;
        const bool in = iz < nz && (mask == nullptr || mask[col + nxy * iz]);
        inbits |= (uint32_t)in << j;
        z[j] = zds[min(iz, nz - 1)];
        acc[j] = 0.f;
    }
    if (inbits != 0) {
        float zlo = z[0], zhi = z[0];
#pragma unroll
        for (int j = 1; j < ZB; ++j) {
            zlo = fminf(zlo, z[j]);
            zhi = fmaxf(zhi, z[j]);
        }
        const float x = xds[ix], y = yds[iy];
        const float nsf = (float)ns, ntf = (float)nt;
        const uint32_t rowb = (uint32_t)pitch * Fmt<FMT>::esz;
        const uint32_t vbytes = (uint32_t)nt * rowb + Fmt<FMT>::pad;
        const int64_t viewb = view_stride * (int64_t)Fmt<FMT>::esz;
        const char* v0 = static_cast<const char*>(q);
        for (int ia = 0; ia < nviews; ++ia, v0 += viewb) {
            const ColView cv = col_view<ARC, FMT, OPT>(x, y, cs[ia], dsd_ds, dso_ds, ws, nsf,
                                                       inv_dt_ds,
                                                       FMT == F_F16S ? vscale[ia] : 1.f, wt, zs,
                                                       ia, arc_t_exact, narrow, so_ds);
            if (!cv.hit) continue;                        // the column misses the detector
            const float blo = fmaf(cv.mz, zlo, cv.wt), bhi = fmaf(cv.mz, zhi, cv.wt);
            const bool fast = blo >= 1.f && blo < ntf && bhi >= 1.f && bhi < ntf;
            gather_view<ZB, FMT>(cv, z, cv.wt, ntf, fast, view_rsrc(v0, vbytes), rowb, acc);
        }
    }
#pragma unroll
    for (int j = 0; j < ZB; ++j)
        if (iz0 + j < nz) image[col + nxy * (iz0 + j)] = ((inbits >> j) & 1u) ? acc[j] : 0.f;
}

template <int ZB, bool ARC, int FMT, bool OPT>
__global__ __launch_bounds__(64 * kWaves) void cbct_back_kernel(
    const void* __restrict__ q, int ns, int nt, int pitch, int nviews, int64_t view_stride,
    const ViewCS* __restrict__ cs, const float* __restrict__ vscale,
    const float* __restrict__ xds, const float* __restrict__ yds,
    const float* __restrict__ zds, int nx, int ny, int nz, const uint8_t* __restrict__ mask,
    float* __restrict__ image, float dsd_ds, float dso_ds, float ws, float wt,
    float inv_dt_ds, const float* __restrict__ zs, int arc_t_exact, int narrow, float so_ds)
{
    cbct_gather_block<ZB, ARC, FMT, OPT>(q, ns, nt, pitch, nviews, view_stride, cs, vscale, xds,
                                         yds, zds, nx, ny, nz, mask, image, dsd_ds, dso_ds, ws,
                                         wt, inv_dt_ds, zs, arc_t_exact, narrow, so_ds,
                                         blockIdx.x);
}

// ---------------------------------------------------------------------------
// staged kernel
// ---------------------------------------------------------------------------

// Build-time tuning of the staged kernel (defaults measured best on the
// 256^2 x 128 FDK workload): LDS loads in flight per batch, minimum waves
// per SIMD of the flat-detector forms (caps the VGPRs), LDS row pitch in
// entries.  More waves per SIMD hide the staging and LDS latency that the
// per-view barrier exposes, though the VGPR cap spills some registers (to
// scratch, outside the inner loop's working set).  Against 3 waves and
// batches of 8 (119 VGPRs, 4 waves): FP32 1.096 -> 0.986 ms at 5 waves (96
// VGPRs, 27 spilled; the LDS holds 5 blocks of its 16-deep form per CU),
// scaled FP16 (mirror form) 0.922 -> 0.878 ms at 5 and 0.737 ms at 6 (80
// VGPRs, 12 spilled; 7 waves: 0.732 ms with 25 spilled), BF16 1.071 -> 1.052
// ms.  Squeezing the FP32 blocks to 6 per CU measured slower: pitch 52,
// 80 VGPRs with 59 spilled, 0.995 ms; 28-row windows (SGX_STAGED_ROWS, 0.3 %
// of the workload's windows then overflow), 80 VGPRs, 1.245 ms.  The arc
// detector's forms (the same volume from a 512 x 384 arc panel): FP32
// 1.110 -> 1.083 ms at 5 waves (from 3), scaled FP16 1.018 -> 0.815 ms at 6
// (from 4; uncapped they take 131 VGPRs and run at 3 waves), BF16 1.13 ->
// 1.11 ms.
#ifndef SGX_STAGED_BATCH
#define SGX_STAGED_BATCH 4
#endif
#ifndef SGX_STAGED_WAVES
#define SGX_STAGED_WAVES 6
#endif
#ifndef SGX_STAGED_WAVES_ARC16
#define SGX_STAGED_WAVES_ARC16 6
#endif
#ifndef SGX_STAGED_WAVES_ARC32
#define SGX_STAGED_WAVES_ARC32 5
#endif

#ifndef SGX_STAGED_PITCH16
#define SGX_STAGED_PITCH16 52
#endif
// minimum waves per SIMD of the packed-FP16 kernel (1.079 ms at 3, 1.016 at
// 5, 0.943 at 6 on the 256^2 x 128 workload)
#ifndef SGX_PACKED_WAVES
#define SGX_PACKED_WAVES 6
#endif
#ifndef SGX_STAGED_PITCH
#define SGX_STAGED_PITCH 58
#endif
// the wide 8-deep form's LDS row pitch (entries): 16 bytes of bank stagger
// per row for both entry sizes
#ifndef SGX_STAGED_PITCH_W
#define SGX_STAGED_PITCH_W 66
#endif
#ifndef SGX_STAGED_PITCH16_W
#define SGX_STAGED_PITCH16_W 68
#endif
#ifndef SGX_STAGED_SCHED
#define SGX_STAGED_SCHED 1
#endif
constexpr int kSP = 16;               // block patch: 16 x 16 columns (2 x 2 waves of 8 x 8)
constexpr int kSZ = 16;               // voxels per lane along z
#ifndef SGX_STAGED_WMAX
#define SGX_STAGED_WMAX 48
#endif
constexpr int kWmax = SGX_STAGED_WMAX; // staged window: bins (quads of 4) ...
constexpr int kHmax = 32;             // ... x rows (16 groups of 2)
// rows of the 16-deep form's windows (at most kHmax; the packed kernel
// shares its window table)
#ifndef SGX_STAGED_ROWS
#define SGX_STAGED_ROWS 32
#endif
constexpr int kH16 = SGX_STAGED_ROWS;
static_assert(kH16 <= kHmax && kH16 % 2 == 0, "16-deep window rows");
constexpr int kSB = SGX_STAGED_BATCH;

// LDS entry of the staged window: (q[is, it], q[is, it+1] - q[is, it]) as two
// FP32 (8 bytes), or for the scaled-FP16 form as two FP16 (4 bytes; the
// difference of two FP16 values is exact in FP16 whenever they are within a
// factor 2 of each other, and otherwise rounded once).  The pitch keeps rows
// 16-byte aligned and staggers the banks of neighbouring rows.
//
// Mirror form (MIR).  When the source stays in the plane z = 0 (no per-view
// source z), the detector's t centre is its middle row (offset_t = 0) and
// the z axis is symmetric (z[nz-1-k] = -z[k]), the voxel at -z lands on the
// detector row mirrored from the one z lands on: bt(-z) = nt + 1 - bt(z),
// with the same s position and weights.  A mirror block owns 8 voxels z > 0
// per lane and their 8 mirror images; its window (at most kHm rows) is the
// one of the upper 8, and the lower voxels read a second window holding the
// mirrored rows in reverse order, entries (q~[r], q~[r+1] - q~[r]) with
// q~[r] = q[nt - 1 - r], at the same entry row and column.  The two windows
// share each LDS row (the mirror one at entry column `half` of the row), so
// the second read takes the first one's address with an immediate offset.
// One position (FMA, floor, fraction, address) then serves two voxels: 6
// VALU per update in the inner loop instead of 8.
//
// Z8 form: blocks of 16 x 16 columns x 8 z (twice as many blocks, windows of
// at most kHm rows), for volumes too small to fill the GPU with 16-deep
// blocks; one of the autotuner's candidates.
//
// Wide Z8 form (Z8W): 8-deep blocks with windows of up to kWmaxW bins x
// kHmax rows, for coarse voxels (a 16 x 16 x 8 block of 2 mm voxels seen at
// magnification 2 spans about 50 bins and 27 rows of 1.2 mm, past the other
// forms' windows, whose blocks would then gather).  Every thread stages.
constexpr int kSZm = 8;               // mirror / Z8 forms: voxels per lane positioned
constexpr int kHm = 20;               // mirror / Z8 forms: window rows (each window)
constexpr int kWmaxW = 64;            // Z8W form: window bins
constexpr int kZ16 = 0, kZMirror = 1, kZ8 = 2, kZ8W = 3;   // block forms of the staged kernel (ZF)
template <int FMT, int ZF = kZ16> struct Lds {
    static constexpr bool half = FMT == F_F16S;
    static constexpr bool mir = ZF == kZMirror;
    static constexpr bool wide = ZF == kZ8W;
    static constexpr uint32_t esz = half ? 4 : 8;
    static constexpr int wmax = wide ? kWmaxW : kWmax;        // window bins
    static constexpr int wpitch = wide ? (half ? SGX_STAGED_PITCH16_W : SGX_STAGED_PITCH_W)
                                       : (half ? SGX_STAGED_PITCH16 : SGX_STAGED_PITCH);
    static constexpr int pitch = mir ? 2 * wpitch : wpitch;   // entries per LDS row
    static constexpr int rows = ZF == kZ16 ? kH16 : wide ? kHmax : kHm;   // rows per buffer
    static constexpr int buf = rows * pitch;                  // entries per buffer
    // one stager per quad and row pair of each window
    static constexpr int stagers = (wmax / 4) * (rows / 2) * (mir ? 2 : 1);
    static_assert(wpitch >= wmax + 2 && (wpitch * esz) % 16 == 0, "pitch");
    static_assert(stagers <= 256, "one staging thread per quad and row pair");
    // the second read of a mirror pair: an immediate offset of ds_read2
    static_assert(!mir || wpitch + 1 <= 255, "ds_read2 offset");
};

static_assert(kSZ % kSB == 0 && kSZm % (kSB / 2) == 0, "batch divides the voxels per lane");

// Minimum waves per SIMD of a staged kernel instance: SGX_STAGED_WAVES for
// the flat detector, capped by the blocks of 4 waves that the 160 KiB of LDS
// per CU holds (4 for the FP32 and BF16 mirror and wide forms).
template <bool ARC, int FMT, int ZF> constexpr int staged_waves()
{
    constexpr int by_lds = (160 * 1024) / (2 * Lds<FMT, ZF>::buf * (int)Lds<FMT, ZF>::esz);
    constexpr int want = ARC ? (FMT != F_F32 ? SGX_STAGED_WAVES_ARC16 : SGX_STAGED_WAVES_ARC32)
                             : SGX_STAGED_WAVES;
    return want < by_lds ? want : by_lds;
}

// Window of one (block, view): entries (clo + c, rlo + r), c < W, r < H,
// stored as int4 {clo | interior, rlo, W, H}.  clo is a multiple of 4;
// W = H = 0 when nothing is staged.  `interior` (bit 0) marks a window that
// was not clamped and whose corner bound keeps one bin and one row of margin
// inside the detector on every side: every lane of the block then has
// 1 <= floor(bs) < ns - 1, 1 <= floor(bt) < nt - 1 and its taps inside the
// window, so the kernel skips the per-lane window tests.
__global__ void cbct_window_kernel(int nblocks, int nviews, const ViewCS* __restrict__ cs,
                                   const float* __restrict__ xds, const float* __restrict__ yds,
                                   const float* __restrict__ zds, int nx, int ny, int nz, int ns,
                                   int nt, float dsd_ds, float dso_ds, float ws, float wt,
                                   float inv_dt_ds, int is_arc, const float* __restrict__ zs,
                                   int arc_t_exact, float so_ds, int zform,
                                   int4* __restrict__ win)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)nblocks * nviews) return;
    const int blk = (int)(i / nviews), ia = (int)(i % nviews);
    const int tiles_x = (nx + kSP - 1) / kSP;
    const int tiles = tiles_x * ((ny + kSP - 1) / kSP);
    const int tile = blk % tiles, zblk = blk / tiles;
    // a mirror block's window is the one of its voxels z > 0
    const int zn = zform == kZ16 ? kSZ : kSZm;
    const int hmax = zform == kZ16 ? kH16 : zform == kZ8W ? kHmax : kHm;
    const int wmax = zform == kZ8W ? kWmaxW : kWmax;
    const int ix0 = (tile % tiles_x) * kSP, iy0 = (tile / tiles_x) * kSP;
    const int iz0 = zform == kZMirror ? nz / 2 + zblk * kSZm : zblk * zn;
    // extent of the block's voxel centres (any axis order)
    float x0 = xds[ix0], x1 = x0, y0 = yds[iy0], y1 = y0, z0 = zds[min(iz0, nz - 1)], z1 = z0;
    for (int k = 1; k < kSP; ++k) {
        const float xv = xds[min(ix0 + k, nx - 1)], yv = yds[min(iy0 + k, ny - 1)];
        x0 = fminf(x0, xv); x1 = fmaxf(x1, xv);
        y0 = fminf(y0, yv); y1 = fmaxf(y1, yv);
    }
    for (int k = 1; k < zn; ++k) {
        const float zv = zds[min(iz0 + k, nz - 1)];
        z0 = fminf(z0, zv); z1 = fmaxf(z1, zv);
    }
    // bs and bt are quasi-linear in (x, y) while y_beta > 0 (a ratio of
    // linear functions), so their extremes over the box are at its corners.
    // With the exact arc t magnification dsd / |(x_beta, y_beta)| that does
    // not hold for bt; it is bounded by the magnifications dsd / max |.| and
    // dsd / min y_beta over the corners (|.| >= y_beta; |.| is convex and
    // y_beta linear, so both extremes are at corners).
    const ViewCS v = cs[ia];
    const float zsv = zs ? zs[ia] : 0.f;
    const bool texact = is_arc && arc_t_exact;
    float smin = 3.4e38f, smax = -3.4e38f, tmin = 3.4e38f, tmax = -3.4e38f;
    float ybmin = 3.4e38f, dmax = 0.f;
    bool ok = true;
    for (int k = 0; k < 4; ++k) {
        const float x = (k & 1) ? x1 : x0, y = (k & 2) ? y1 : y0;
        const float xb = x * v.c + y * v.s - so_ds;
        const float yb = dso_ds - (-x * v.s + y * v.c);
        ok = ok && yb > 0.f;
        const float mag = dsd_ds * __builtin_amdgcn_rcpf(yb);
        const float sp = is_arc ? dsd_ds * atan2_fast(xb, yb) : mag * xb;
        const float bs = sp + ws;
        smin = fminf(smin, bs); smax = fmaxf(smax, bs);
        ybmin = fminf(ybmin, yb);
        dmax = fmaxf(dmax, sqrtf(xb * xb + yb * yb));
        if (!texact) {
            const float mz = mag * inv_dt_ds;
            const float ta = fmaf(mz, z0 - zsv, wt), tb = fmaf(mz, z1 - zsv, wt);
            tmin = fminf(tmin, fminf(ta, tb)); tmax = fmaxf(tmax, fmaxf(ta, tb));
        }
    }
    if (texact) {
        const float mzs[2] = {dsd_ds / dmax * inv_dt_ds, dsd_ds / ybmin * inv_dt_ds};
        for (int k = 0; k < 2; ++k) {
            const float ta = fmaf(mzs[k], z0 - zsv, wt), tb = fmaf(mzs[k], z1 - zsv, wt);
            tmin = fminf(tmin, fminf(ta, tb)); tmax = fmaxf(tmax, fmaxf(ta, tb));
        }
    }
    int4 w = make_int4(0, 0, 0, 0);
    // entry columns is-1, is for 1 <= is = floor(bs) < ns, one bin of margin
    // each side; entry rows it-1 for 1 <= it = floor(bt) < nt, likewise
    if (ok && smax >= 1.f && smin < (float)ns && tmax >= 1.f && tmin < (float)nt) {
        const int clo = max((int)floorf(fmaxf(smin, 0.f)) - 2, 0) & ~3;
        const int chi = min((int)floorf(fminf(smax, (float)ns)) + 1, ns - 1);
        const int rlo = max((int)floorf(fmaxf(tmin, 0.f)) - 2, 0);
        const int rhi = min((int)floorf(fminf(tmax, (float)nt)), nt - 2);
        if (chi >= clo && rhi >= rlo) {
            const int W = chi - clo + 1, H = rhi - rlo + 1;
            const bool interior = smin >= 2.f && smax < (float)(ns - 1) && tmin >= 2.f &&
                                  tmax < (float)(nt - 2) && W <= wmax && H <= hmax;
            w = make_int4(clo | (interior ? 1 : 0), rlo, min(W, wmax), min(H, hmax));
        }
    }
    win[i] = w;
}

// One block's work: voxel block `blk` (16 x 16 columns x 16 z; for the
// mirror form 16 x 16 columns x 8 z > 0 and their 8 images; for the Z8 form
// 16 x 16 x 8), all views, with the block's LDS `lds` (2 * Lds<FMT, ZF>::buf
// entries).
template <bool ARC, int FMT, bool OPT, int ZF>
__device__ __forceinline__ void cbct_staged_block(
    const void* __restrict__ q, int ns, int nt, int pitch, int nviews, int64_t view_stride,
    const ViewCS* __restrict__ cs, const float* __restrict__ vscale,
    const float* __restrict__ xds, const float* __restrict__ yds,
    const float* __restrict__ zds, int nx, int ny, int nz, const uint8_t* __restrict__ mask,
    float* __restrict__ image, float dsd_ds, float dso_ds, float ws, float wt,
    float inv_dt_ds, const int4* __restrict__ win, const float* __restrict__ zs,
    int arc_t_exact, int narrow, float so_ds, int blk, char* lds, int win_stride, int accumulate)
{
    using L = Lds<FMT, ZF>;
    constexpr bool MIR = ZF == kZMirror;
    constexpr int kLP = L::pitch;
    constexpr int NZ = ZF == kZ16 ? kSZ : kSZm; // voxels per lane whose position is computed
    constexpr int NV = MIR ? 2 * kSZm : NZ;     // voxels per lane
    constexpr int SB = MIR ? kSB / 2 : kSB;     // positions per batch (kSB LDS reads)
    // byte offset of the mirror window's entries within an LDS row
    constexpr uint32_t kMirOff = (uint32_t)L::wpitch * L::esz;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = (nx + kSP - 1) / kSP;
    const int tiles = tiles_x * ((ny + kSP - 1) / kSP);
    const int tile = blk % tiles, zblk = blk / tiles;
    const int ix = (tile % tiles_x) * kSP + (wave & 1) * 8 + (lane & 7);
    const int iy = (tile / tiles_x) * kSP + (wave >> 1) * 8 + (lane >> 3);
    // voxel j of the lane: izb + j; for the mirror form j < kSZm are z > 0
    // and kSZm + j the image of j
    const int izb = MIR ? nz / 2 + zblk * kSZm : zblk * NZ;
    auto vox = [&](int j) { return (MIR && j >= kSZm) ? nz - 1 - (izb + j - kSZm) : izb + j; };
    const bool inxy = ix < nx && iy < ny;
    const int64_t nxy = (int64_t)nx * ny;
    const int64_t col = inxy ? ix + (int64_t)nx * iy : 0;

    // Every thread stays to the end: all of them stage and meet the barriers.
    // Bit j of inbits: voxel j exists and is inside the mask.
    float z[NV], acc[NV];
    uint32_t inbits = 0;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const int iz = vox(j);
        const bool in = inxy && iz < nz && (mask == nullptr || mask[col + nxy * iz]);
        inbits |= (uint32_t)in << j;
        z[j] = zds[min(iz, nz - 1)];
        acc[j] = 0.f;
    }
    const bool any = inbits != 0;
    // z range of the voxels whose position is computed
    float zlo = z[0], zhi = z[0];
#pragma unroll
    for (int j = 1; j < NZ; ++j) {
        zlo = fminf(zlo, z[j]);
        zhi = fmaxf(zhi, z[j]);
    }
    const float x = inxy ? xds[ix] : 0.f, y = inxy ? yds[iy] : 0.f;
    const float nsf = (float)ns, ntf = (float)nt;
    const uint32_t rowb = (uint32_t)pitch * Fmt<FMT>::esz;
    const uint32_t vbytes = (uint32_t)nt * rowb + Fmt<FMT>::pad;
    const int64_t viewb = view_stride * (int64_t)Fmt<FMT>::esz;
    const char* qb = static_cast<const char*>(q);
    // the block's row of the window table (win_stride views per block; a
    // launch over a chunk of the views starts at its first view's entry)
    const int4* bwin = win + (int64_t)blk * win_stride;

    // Staging role: thread (quad sq, group sg) moves bins clo+4sq .. +3 of
    // entry rows 2sg, 2sg+1 of its window, i.e. data rows 2sg .. 2sg+2 of it.
    // Mirror form: groups sg >= kHm / 2 stage the mirror window, whose entry
    // row r reads the data rows nt - 1 - (rlo + r) downwards.
    const bool stager = tid < L::stagers;
    constexpr int QD = L::wmax / 4;                   // quads per window row
    const int sq = tid % QD, sg = tid / QD;
    const bool mst = MIR && sg >= kHm / 2;
    const int sgw = mst ? sg - kHm / 2 : sg;          // row pair within the window
    const uint32_t toff = ((uint32_t)(2 * sgw) * (uint32_t)pitch + 4u * sq) * Fmt<FMT>::esz;
    char* const sdst = lds + ((2 * sgw) * kLP + (mst ? L::wpitch : 0) + 4 * sq) * L::esz;
    float sv[3][4];          // FP32 / BF16: the window's values, widened
    uint2 hv[3];             // F16S: the raw FP16 values (4 bins per row)
    auto stage_load = [&](int ia, int4 w) {
        if (!stager || 4 * sq >= w.z || 2 * sgw >= w.w) return;
        const __amdgpu_buffer_rsrc_t rs = view_rsrc(qb + ia * viewb, vbytes);
        const uint32_t cb = (uint32_t)(w.x & ~3) * Fmt<FMT>::esz;
        // offset of data row 2sgw (+ k rows in the window's direction)
        const uint32_t off = mst ? ((uint32_t)(nt - 1 - w.y - 2 * sgw) * (uint32_t)pitch + 4u * sq) *
                                           Fmt<FMT>::esz + cb
                                 : toff + (uint32_t)w.y * (uint32_t)pitch * Fmt<FMT>::esz + cb;
        const uint32_t step = mst ? 0u - rowb : rowb;
        const bool third = 2 * sgw + 2 <= w.w;             // data row 2sgw+2 needed
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (k == 2 && !third) break;
            const uint32_t o = off + (uint32_t)k * step;
            if (FMT == F_F32) {
                const auto u = __builtin_amdgcn_raw_buffer_load_b128(rs, o, 0, 0);
#pragma unroll
                for (int e = 0; e < 4; ++e) sv[k][e] = __uint_as_float(u[e]);
            } else if (FMT == F_BF16) {
                const auto u = __builtin_amdgcn_raw_buffer_load_b64(rs, o, 0, 0);
                const float2 e01 = dec2<FMT>(u[0]), e23 = dec2<FMT>(u[1]);
                sv[k][0] = e01.x; sv[k][1] = e01.y; sv[k][2] = e23.x; sv[k][3] = e23.y;
            } else {
                const auto u = __builtin_amdgcn_raw_buffer_load_b64(rs, o, 0, 0);
                hv[k] = make_uint2(u[0], u[1]);
            }
        }
    };
    auto stage_store = [&](int buf, int4 w) {
        if (!stager || 4 * sq >= w.z || 2 * sgw >= w.w) return;
        char* d = sdst + buf * (L::buf * L::esz);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (2 * sgw + k >= w.w) break;
            if (L::half) {
                // row differences of bin pairs in FP16, then entries (a, d) per bin
                const uint32_t d01 = half2_sub(hv[k + 1].x, hv[k].x);
                const uint32_t d23 = half2_sub(hv[k + 1].y, hv[k].y);
                const uint4 e = make_uint4(__builtin_amdgcn_perm(d01, hv[k].x, 0x05040100u),
                                           __builtin_amdgcn_perm(d01, hv[k].x, 0x07060302u),
                                           __builtin_amdgcn_perm(d23, hv[k].y, 0x05040100u),
                                           __builtin_amdgcn_perm(d23, hv[k].y, 0x07060302u));
                *reinterpret_cast<uint4*>(d + k * kLP * L::esz) = e;
            } else {
                const float4 lo = make_float4(sv[k][0], sv[k + 1][0] - sv[k][0], sv[k][1],
                                              sv[k + 1][1] - sv[k][1]);
                const float4 hi = make_float4(sv[k][2], sv[k + 1][2] - sv[k][2], sv[k][3],
                                              sv[k + 1][3] - sv[k][3]);
                *reinterpret_cast<float4*>(d + k * kLP * L::esz) = lo;
                *reinterpret_cast<float4*>(d + k * kLP * L::esz + 16) = hi;
            }
        }
    };
    const char* const lds_bytes = reinterpret_cast<const char*>(lds);

    // Per-view lane geometry: bs, w2, mz and the lane's window coordinates.
    struct LaneView {
        float bs, w2, mz, wt, blo, bhi;
        bool inwin;
        int ce;
    };
    // check = false: the window is interior (all lanes inside, see above).
    auto lane_view = [&](int ia, int4 w, bool check) {
        LaneView r;
        const ViewCS v = cs[ia];
        const float xb0 = x * v.c + y * v.s;
        const float xb = OPT ? xb0 - so_ds : xb0;      // r_loop (source offset)
        const float yb = dso_ds - (-x * v.s + y * v.c);
        const float mag = dsd_ds * __builtin_amdgcn_rcpf(yb);
        float sp;
        if (ARC) {
            sp = dsd_ds * (narrow ? atan2_narrow(xb, yb) : atan2_fast(xb, yb));
            r.w2 = dsd_ds * dsd_ds * __builtin_amdgcn_rcpf(xb * xb + yb * yb);
        } else {
            sp = mag * xb;
            r.w2 = mag * mag;
        }
        if (FMT == F_F16S) r.w2 *= vscale[ia];
        r.bs = sp + ws;
        t_map<ARC, OPT>(xb, yb, mag, dsd_ds, inv_dt_ds, wt, zs, ia, arc_t_exact, r.mz, r.wt);
        // Window coordinates of the lane's taps: entry columns ce, ce+1 and
        // entry rows (int)bt - 1 - rlo for bt from blo to bhi.  The test
        // implies bs, blo, bhi >= 1 (truncation is floor there) and the
        // detector bounds 1 <= floor(bs) < ns, 1 <= floor(bt) < nt.  The
        // mirror images sit at nt + 1 - bt: they need bt > 1 (their floor
        // below nt), which an interior window gives with a row to spare.
        r.ce = (int)r.bs - 1 - (w.x & ~3);
        r.inwin = true;
        r.blo = r.bhi = 0.f;
        if (check) {
            r.blo = fmaf(r.mz, zlo, r.wt);
            r.bhi = fmaf(r.mz, zhi, r.wt);
            const int rl = (int)r.blo - 1 - w.y, rh = (int)r.bhi - 1 - w.y;
            const uint32_t wlim = w.z > 1 ? (uint32_t)(w.z - 1) : 0u;
            r.inwin = (uint32_t)r.ce < wlim && (uint32_t)rl < (uint32_t)w.w &&
                      (uint32_t)rh < (uint32_t)w.w && (!MIR || fminf(r.blo, r.bhi) > 1.f);
        }
        return r;
    };

    bool missed = false;       // some view's taps were outside the staged window
    int4 wcur = bwin[0];
    stage_load(0, wcur);
    stage_store(0, wcur);
    __syncthreads();
    for (int ia = 0; ia < nviews; ++ia) {
        const int buf = ia & 1;
        const int4 wnext = ia + 1 < nviews ? bwin[ia + 1] : make_int4(0, 0, 0, 0);
        stage_load(ia + 1, wnext);                          // in flight during this view
        if (any) {
            const LaneView lv = lane_view(ia, wcur, (wcur.x & 1) == 0);
            if (lv.inwin) {
                const float wb = lv.w2 * __builtin_amdgcn_fractf(lv.bs);
                const float wa = lv.w2 - wb;
                // byte offset of the LDS entry (ce, it - 1 - rlo) is it * kLP * esz + lb
                const uint32_t lb = ((uint32_t)lv.ce +
                                     (uint32_t)(buf * L::buf - (wcur.y + 1) * kLP)) * L::esz;
#pragma unroll
                for (int h = 0; h < NZ; h += SB) {
                    float4 tp[SB], tm[SB];    // (a0, d0, a1, d1); tm: the mirror window
                    float wu[SB];
#pragma unroll
                    for (int j = 0; j < SB; ++j) {
                        const float bt = fmaf(lv.mz, z[h + j], lv.wt);
                        wu[j] = __builtin_amdgcn_fractf(bt);
                        // one in-bounds offset (the 32-bit sum wraps to the entry's offset)
                        const uint32_t o =
                            __umul24((uint32_t)bt, (uint32_t)(kLP * L::esz)) + lb;
                        const char* p = lds_bytes + (int)o;
                        if (L::half) {
                            const uint32_t* u = reinterpret_cast<const uint32_t*>(p);
                            const uint32_t u0 = u[0], u1 = u[1];
                            tp[j] = make_float4(f16f(u0 & 0xffffu), f16f(u0 >> 16),
                                                f16f(u1 & 0xffffu), f16f(u1 >> 16));
                            if (MIR) {
                                const uint32_t* m = reinterpret_cast<const uint32_t*>(p + kMirOff);
                                const uint32_t m0 = m[0], m1 = m[1];
                                tm[j] = make_float4(f16f(m0 & 0xffffu), f16f(m0 >> 16),
                                                    f16f(m1 & 0xffffu), f16f(m1 >> 16));
                            }
                        } else {
                            const float2* f = reinterpret_cast<const float2*>(p);
                            const float2 e0 = f[0], e1 = f[1];
                            tp[j] = make_float4(e0.x, e0.y, e1.x, e1.y);
                            if (MIR) {
                                const float2* m = reinterpret_cast<const float2*>(p + kMirOff);
                                const float2 m0 = m[0], m1 = m[1];
                                tm[j] = make_float4(m0.x, m0.y, m1.x, m1.y);
                            }
                        }
                    }
                    // every LDS read of the batch is issued before its
                    // arithmetic (the scheduler otherwise keeps two in flight)
                    if (SGX_STAGED_SCHED) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int j = 0; j < SB; ++j) {
                        const float p0 = fmaf(wu[j], tp[j].y, tp[j].x);
                        const float p1 = fmaf(wu[j], tp[j].w, tp[j].z);
                        acc[h + j] = fmaf(wb, p1, fmaf(wa, p0, acc[h + j]));
                        if (MIR) {
                            const float m0 = fmaf(wu[j], tm[j].y, tm[j].x);
                            const float m1 = fmaf(wu[j], tm[j].w, tm[j].z);
                            acc[kSZm + h + j] = fmaf(wb, m1, fmaf(wa, m0, acc[kSZm + h + j]));
                        }
                    }
                }
            } else {
                missed = true;
            }
        }
        stage_store(buf ^ 1, wnext);
        __syncthreads();
        wcur = wnext;
    }

    // Views whose taps missed the staged window (detector edges in t, windows
    // larger than the LDS buffer): gathered from global memory, lane by lane,
    // each voxel at its own z (the mirror images too).
    if (missed) {
        float zlo2 = z[NV - 1], zhi2 = z[NV - 1];
        if (MIR) {
#pragma unroll
            for (int j = kSZm; j < NV; ++j) {
                zlo2 = fminf(zlo2, z[j]);
                zhi2 = fmaxf(zhi2, z[j]);
            }
        }
        for (int ia = 0; ia < nviews; ++ia) {
            const LaneView lv = lane_view(ia, bwin[ia], true);
            if (lv.inwin || !(lv.bs >= 1.f && lv.bs < nsf)) continue;    // 1 <= floor(bs) < ns
            ColView cv;
            cv.isf = floorf(lv.bs);
            cv.wb = lv.w2 * (lv.bs - cv.isf);
            cv.wa = lv.w2 - cv.wb;
            cv.mz = lv.mz;
            cv.wt = lv.wt;
            const __amdgpu_buffer_rsrc_t rs = view_rsrc(qb + ia * viewb, vbytes);
            const float tlo = fminf(lv.blo, lv.bhi), thi = fmaxf(lv.blo, lv.bhi);
            gather_view<NZ, FMT>(cv, z, lv.wt, ntf, tlo >= 1.f && thi < ntf, rs, rowb, acc);
            if (MIR) {
                const float b0 = fmaf(lv.mz, zlo2, lv.wt), b1 = fmaf(lv.mz, zhi2, lv.wt);
                const bool fast = fminf(b0, b1) >= 1.f && fmaxf(b0, b1) < ntf;
                gather_view<NZ, FMT>(cv, z + kSZm, lv.wt, ntf, fast, rs, rowb, acc + kSZm);
            }
        }
    }

    // Masked-out voxels are written as 0 (a select on the prologue's bits:
    // no second mask read inside a branch).  A launch over a later chunk of
    // the views adds its sum to the image.
    if (inxy) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int iz = vox(j);
            if (iz >= nz) continue;
            float* o = image + col + nxy * iz;
            const float v = accumulate ? *o + acc[j] : acc[j];
            *o = ((inbits >> j) & 1u) ? v : 0.f;
        }
    }
}

template <bool ARC, int FMT, bool OPT, int ZF>
__global__ __launch_bounds__(256, (staged_waves<ARC, FMT, ZF>())) void
cbct_staged_kernel(
    const void* __restrict__ q, int ns, int nt, int pitch, int nviews, int64_t view_stride,
    const ViewCS* __restrict__ cs, const float* __restrict__ vscale,
    const float* __restrict__ xds, const float* __restrict__ yds,
    const float* __restrict__ zds, int nx, int ny, int nz, const uint8_t* __restrict__ mask,
    float* __restrict__ image, float dsd_ds, float dso_ds, float ws, float wt,
    float inv_dt_ds, const int4* __restrict__ win, const float* __restrict__ zs,
    int arc_t_exact, int narrow, float so_ds, int win_stride, int accumulate)
{
    using L = Lds<FMT, ZF>;
    __shared__ __attribute__((aligned(16))) char lds[2 * L::buf * L::esz];
    cbct_staged_block<ARC, FMT, OPT, ZF>(q, ns, nt, pitch, nviews, view_stride, cs, vscale, xds,
                                          yds, zds, nx, ny, nz, mask, image, dsd_ds, dso_ds, ws,
                                          wt, inv_dt_ds, win, zs, arc_t_exact, narrow, so_ds,
                                          blockIdx.x, lds, win_stride, accumulate);
}

// Batch form: one launch over the tasks (item, voxel block) of many
// reconstructions; block b does task b of the table, which the host orders
// (largest tasks first, interleaved so that block b % 8 keeps to one XCD's
// share).  The hardware dispatches blocks in index order, so the launch
// walks the table in order and backfills the GPU as blocks finish.
template <bool ARC, int FMT, bool OPT, int ZF>
__global__ __launch_bounds__(256, (staged_waves<ARC, FMT, ZF>())) void
cbct_queue_kernel(const CbctItem* __restrict__ items, const int2* __restrict__ tasks)
{
    using L = Lds<FMT, ZF>;
    __shared__ __attribute__((aligned(16))) char lds[2 * L::buf * L::esz];
    const int2 tk = tasks[blockIdx.x];
    // through the constant address space: invariant loads, as for kernel
    // arguments
    using CItem = const __attribute__((address_space(4))) CbctItem;
    const CItem& it = ((CItem*)(items))[tk.x];
    cbct_staged_block<ARC, FMT, OPT, ZF>(it.q, it.ns, it.nt, it.pitch, it.nviews,
                                          it.view_stride, it.cs, it.vscale, it.xds, it.yds,
                                          it.zds, it.nx, it.ny, it.nz, it.mask, it.image,
                                          it.dsd_ds, it.dso_ds, it.ws, it.wt, it.inv_dt_ds,
                                          it.win, it.zs, it.arc_t_exact, it.narrow, it.so_ds,
                                          tk.y, lds, it.nviews, 0);
}

// ---------------------------------------------------------------------------
// packed-FP16 kernel (scaled-FP16 storage)
// ---------------------------------------------------------------------------
//
// The blocks, window table and view loop of the staged kernel, with the
// interpolation of the scaled-FP16 form in packed FP16 arithmetic (selected
// by SGX_OPT_CBCT_KERNEL = 2):
//
// * LDS entry (c, r): {(q[c, r], q[c+1, r]), (d[c, r], d[c+1, r])}, two FP16
//   pairs, d = fp16(q[., r+1] - q[., r]), rows kPkRow bytes apart (a pitch
//   that staggers the banks of neighbouring rows).
// * Position word: Y = fma(1024 mz, z, 1024 (wt - rlo - 1) + 2^23) in FP32.
//   While 0 <= bt - rlo - 1 < 8192 its bits [22:10] are the entry row of the
//   position rounded to 1/1024 of a row and bits [9:0] the fraction f in
//   units of 1/1024: the row is one bit-field extract, and
//   ((Y & 0x3ff) | 0x3c00) - 1 is f as FP16, exact, formed for two voxels at
//   once.
// * Update: (q0 + f d0, q1 + f d1) in one v_pk_fma_f16 (each rounded once to
//   FP16), then acc += W . v in one v_dot2_f32_f16, W = FP16 ((1 - fs) w2,
//   fs w2) per lane and view, with every view's scale taken relative to the
//   largest so that W stays in FP16's normal range; the result is multiplied
//   back by the largest scale.  6.5 VALU per update in the inner loop
//   against 8 for the FP32 interpolation.
// * Measured on the 256^2 x 128 FDK workload (360 views) it does not pay:
//   the 8-byte entries repeat every bin (twice the staging work and LDS
//   writes of the FP32 interpolation's 4-byte entries), so the whole kernel
//   issues 9.7 VALU per update against 10.1 and runs 1.09 ms against
//   0.98 ms for the 16-deep FP32-interpolation blocks, both at 3 waves per
//   SIMD; at 6 waves 0.94 ms against 0.74 ms for the mirror form.  The rounding of the position to 1/1024 row, of v and of W each
//   add about as much error as the FP16 storage itself: 6.7e-5 from Float64
//   against 3.4e-5.
// * Lanes whose taps leave the window, or whose w2 would leave FP16's range,
//   take the FP32 gather path after the view loop, as in the staged kernel.

#ifndef SGX_PACKED_PITCH
#define SGX_PACKED_PITCH 52
#endif
constexpr uint32_t kPkRow = 8 * SGX_PACKED_PITCH;  // LDS row pitch (bytes); staggers the banks
constexpr uint32_t kPkBuf = kHmax * kPkRow;        // bytes per buffer
static_assert(SGX_PACKED_PITCH >= kWmax + 2, "a window row fits an LDS row");
constexpr int kPkPairs = kWmax / 2;                // staging threads per row group
constexpr int kPkRowsPer = 4;                      // entry rows per staging thread
constexpr int kPkStagers = kPkPairs * (kHmax / kPkRowsPer);
static_assert(kPkStagers <= 256, "one staging thread per entry pair and row group");
static_assert((kHmax & (kHmax - 1)) == 0, "the row field needs a power-of-two row count");

typedef _Float16 hf2 __attribute__((ext_vector_type(2)));
typedef __fp16 hp2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ hf2 as_hf2(uint32_t u) { return __builtin_bit_cast(hf2, u); }

// (a & m) | c in one v_and_or_b32.  A VOP3 instruction takes no literal here
// and at most one scalar operand, so c comes in a VGPR (the compiler would
// otherwise split a pair of literal masks into v_and + v_or).
__device__ __forceinline__ uint32_t and_or(uint32_t a, uint32_t m, uint32_t c)
{
    uint32_t r;
    asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(m), "v"(c));
    return r;
}

// Per-view lane geometry of the staged kernels (bs, w2 with the view's
// scale, the t map, the lane's entry column and, for windows that are not
// interior, whether every tap of the lane is inside the window).  `tmargin`
// widens the row test for positions that are rounded after it.
struct LaneGeom {
    float bs, w2, mz, wt, blo, bhi;
    bool inwin;
    int ce;
};
template <bool ARC, bool OPT>
__device__ __forceinline__ LaneGeom lane_geom(float x, float y, ViewCS v, int4 w, bool check,
                                              float zlo, float zhi, float dsd_ds, float dso_ds,
                                              float ws, float wt, float inv_dt_ds, float vs,
                                              const float* zs, int ia, int arc_t_exact,
                                              int narrow, float so_ds, float tmargin)
{
    LaneGeom r;
    const float xb0 = x * v.c + y * v.s;
    const float xb = OPT ? xb0 - so_ds : xb0;          // r_loop (source offset)
    const float yb = dso_ds - (-x * v.s + y * v.c);
    const float mag = dsd_ds * __builtin_amdgcn_rcpf(yb);
    float sp;
    if (ARC) {
        sp = dsd_ds * (narrow ? atan2_narrow(xb, yb) : atan2_fast(xb, yb));
        r.w2 = dsd_ds * dsd_ds * __builtin_amdgcn_rcpf(xb * xb + yb * yb);
    } else {
        sp = mag * xb;
        r.w2 = mag * mag;
    }
    r.w2 *= vs;
    r.bs = sp + ws;
    t_map<ARC, OPT>(xb, yb, mag, dsd_ds, inv_dt_ds, wt, zs, ia, arc_t_exact, r.mz, r.wt);
    r.ce = (int)r.bs - 1 - (w.x & ~3);
    r.inwin = true;
    r.blo = fmaf(r.mz, zlo, r.wt);
    r.bhi = fmaf(r.mz, zhi, r.wt);
    if (check) {
        const float tl = fminf(r.blo, r.bhi) - tmargin, th = fmaxf(r.blo, r.bhi) + tmargin;
        const int rl = (int)tl - 1 - w.y, rh = (int)th - 1 - w.y;
        const uint32_t wlim = w.z > 1 ? (uint32_t)(w.z - 1) : 0u;
        r.inwin = tl >= 1.f && (uint32_t)r.ce < wlim && (uint32_t)rl < (uint32_t)w.w &&
                  (uint32_t)rh < (uint32_t)w.w;
    }
    return r;
}

// One block's work (as cbct_staged_block) with the packed-FP16 update; `lds`
// holds kHmax rows of kPkRow bytes.
template <bool ARC, bool OPT>
__device__ __forceinline__ void cbct_packed_block(
    const void* __restrict__ q, int ns, int nt, int pitch, int nviews, int64_t view_stride,
    const ViewCS* __restrict__ cs, const float* __restrict__ vscale,
    const float* __restrict__ xds, const float* __restrict__ yds,
    const float* __restrict__ zds, int nx, int ny, int nz, const uint8_t* __restrict__ mask,
    float* __restrict__ image, float dsd_ds, float dso_ds, float ws, float wt,
    float inv_dt_ds, const int4* __restrict__ win, const float* __restrict__ zs,
    int arc_t_exact, int narrow, float so_ds, int blk, char* lds)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = (nx + kSP - 1) / kSP;
    const int tiles = tiles_x * ((ny + kSP - 1) / kSP);
    const int tile = blk % tiles, zblk = blk / tiles;
    const int ix = (tile % tiles_x) * kSP + (wave & 1) * 8 + (lane & 7);
    const int iy = (tile / tiles_x) * kSP + (wave >> 1) * 8 + (lane >> 3);
    const int iz0 = zblk * kSZ;
    const bool inxy = ix < nx && iy < ny;
    const int64_t nxy = (int64_t)nx * ny;
    const int64_t col = inxy ? ix + (int64_t)nx * iy : 0;

    float z[kSZ], acc[kSZ];
    uint32_t inbits = 0;
#pragma unroll
    for (int j = 0; j < kSZ; ++j) {
        const int iz = iz0 + j;
        const bool in = inxy && iz < nz && (mask == nullptr || mask[col + nxy * iz]);
        inbits |= (uint32_t)in << j;
        z[j] = zds[min(iz, nz - 1)];
        acc[j] = 0.f;
    }
    const bool any = inbits != 0;
    float zlo = z[0], zhi = z[0];
#pragma unroll
    for (int j = 1; j < kSZ; ++j) {
        zlo = fminf(zlo, z[j]);
        zhi = fmaxf(zhi, z[j]);
    }
    // The largest view scale (every wave forms it itself: no barrier).
    float vmax = 0.f;
    for (int i = lane; i < nviews; i += 64) vmax = fmaxf(vmax, vscale[i]);
    for (int o = 32; o >= 1; o >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, o));
    if (!(vmax > 0.f)) vmax = 1.f;
    const float rvmax = 1.f / vmax;

    const float x = inxy ? xds[ix] : 0.f, y = inxy ? yds[iy] : 0.f;
    const float nsf = (float)ns, ntf = (float)nt;
    const uint32_t rowb = (uint32_t)pitch * 2u;
    const uint32_t vbytes = (uint32_t)nt * rowb + Fmt<F_F16S>::pad;
    const int64_t viewb = view_stride * 2;
    const char* qb = static_cast<const char*>(q);
    const int4* bwin = win + (int64_t)blk * nviews;

    // Staging role: thread (pair sp, group sg) builds entries 2sp, 2sp+1 of
    // entry rows 4sg .. 4sg+3 from bins 2sp .. 2sp+3 of data rows 4sg ..
    // 4sg+4, and stores each entry row's two entries as one 16-byte write:
    // the lanes of a row write consecutive 16-byte slots (no bank conflicts).
    const bool stager = tid < kPkStagers;
    const int sp = tid % kPkPairs, sg = tid / kPkPairs;
    const uint32_t toff = ((uint32_t)(kPkRowsPer * sg) * (uint32_t)pitch + 2u * sp) * 2u;
    char* const sdst = lds + (kPkRowsPer * sg) * kPkRow + 16 * sp;
    uint2 hv[kPkRowsPer + 1];     // bins (2sp, 2sp+1), (2sp+2, 2sp+3) of the data rows
    auto stage_load = [&](int ia, int4 w) {
        if (!stager || 2 * sp >= w.z || kPkRowsPer * sg >= w.w) return;
        const __amdgpu_buffer_rsrc_t rs = view_rsrc(qb + ia * viewb, vbytes);
        const uint32_t off = toff + ((uint32_t)w.y * (uint32_t)pitch + (uint32_t)(w.x & ~3)) * 2u;
#pragma unroll
        for (int k = 0; k <= kPkRowsPer; ++k) {
            if (k > 0 && kPkRowsPer * sg + k > w.w) break;     // data row needed up to H
            hv[k].x = __builtin_amdgcn_raw_buffer_load_b32(rs, off + k * rowb, 0, 0);
            hv[k].y = __builtin_amdgcn_raw_buffer_load_b32(rs, off + k * rowb + 4u, 0, 0);
        }
    };
    auto stage_store = [&](int buf, int4 w) {
        if (!stager || 2 * sp >= w.z || kPkRowsPer * sg >= w.w) return;
        char* d = sdst + buf * kPkBuf;
#pragma unroll
        for (int k = 0; k < kPkRowsPer; ++k) {
            if (kPkRowsPer * sg + k >= w.w) break;
            const uint32_t q01 = hv[k].x, q23 = hv[k].y;
            const uint32_t d01 = half2_sub(hv[k + 1].x, q01);
            const uint32_t d23 = half2_sub(hv[k + 1].y, q23);
            // the pair starting at the odd bin: the upper half of one pair and
            // the lower half of the next
            const uint32_t q12 = __builtin_amdgcn_alignbit(q23, q01, 16);
            const uint32_t d12 = __builtin_amdgcn_alignbit(d23, d01, 16);
            *reinterpret_cast<uint4*>(d + k * kPkRow) = make_uint4(q01, d01, q12, d12);
        }
    };

    bool missed = false;
    int4 wcur = bwin[0];
    stage_load(0, wcur);
    stage_store(0, wcur);
    __syncthreads();
    for (int ia = 0; ia < nviews; ++ia) {
        const int buf = ia & 1;
        const int4 wnext = ia + 1 < nviews ? bwin[ia + 1] : make_int4(0, 0, 0, 0);
        stage_load(ia + 1, wnext);
        if (any) {
            const LaneGeom lv = lane_geom<ARC, OPT>(x, y, cs[ia], wcur, (wcur.x & 1) == 0, zlo,
                                                    zhi, dsd_ds, dso_ds, ws, wt, inv_dt_ds,
                                                    vscale[ia] * rvmax, zs, ia, arc_t_exact,
                                                    narrow, so_ds, 1.f / 1024.f);
            if (lv.inwin && lv.w2 < 32768.f) {
                const float wb = lv.w2 * __builtin_amdgcn_fractf(lv.bs);
                const float wa = lv.w2 - wb;
                const hp2 W = {(__fp16)wa, (__fp16)wb};
                const float mzs = lv.mz * 1024.f;
                const float wts = fmaf(lv.wt - (float)(wcur.y + 1), 1024.f, 8388608.f);
                // (the clamp keeps every address inside the LDS block whatever ce is)
                const uint32_t colb = 8u * min((uint32_t)lv.ce, (uint32_t)SGX_PACKED_PITCH - 1u) +
                                      kPkBuf * buf;
                const uint32_t kOnes = 0x3c003c00u;                 // (1.0, 1.0) in FP16
#pragma unroll
                for (int h = 0; h < kSZ; h += kSB) {
                    uint32_t Y[kSB];
                    uint2 e[kSB];
#pragma unroll
                    for (int j = 0; j < kSB; ++j) {
                        Y[j] = __float_as_uint(fmaf(mzs, z[h + j], wts));
                        const uint32_t row = __builtin_amdgcn_ubfe(Y[j], 10, 5);
                        e[j] = *reinterpret_cast<const uint2*>(lds + __umul24(row, kPkRow) + colb);
                    }
                    if (SGX_STAGED_SCHED) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int j = 0; j < kSB; j += 2) {
                        const uint32_t P = __builtin_amdgcn_perm(Y[j + 1], Y[j], 0x05040100u);
                        const hf2 f = as_hf2(and_or(P, 0x03ff03ffu, kOnes)) - (hf2)(1.0f16);
                        const hf2 v0 = __builtin_elementwise_fma(f.xx, as_hf2(e[j].y),
                                                                 as_hf2(e[j].x));
                        const hf2 v1 = __builtin_elementwise_fma(f.yy, as_hf2(e[j + 1].y),
                                                                 as_hf2(e[j + 1].x));
                        acc[h + j] = __builtin_amdgcn_fdot2(__builtin_bit_cast(hp2, v0), W,
                                                            acc[h + j], false);
                        acc[h + j + 1] = __builtin_amdgcn_fdot2(__builtin_bit_cast(hp2, v1), W,
                                                                acc[h + j + 1], false);
                    }
                }
            } else {
                missed = true;
            }
        }
        stage_store(buf ^ 1, wnext);
        __syncthreads();
        wcur = wnext;
    }

    // Views whose taps missed the staged window: gathered from global memory
    // in FP32 arithmetic, with the same relative scale.
    if (missed) {
        for (int ia = 0; ia < nviews; ++ia) {
            // the view loop's decision, recomputed from the same inputs
            const int4 w = bwin[ia];
            const LaneGeom lv = lane_geom<ARC, OPT>(x, y, cs[ia], w, (w.x & 1) == 0, zlo, zhi,
                                                    dsd_ds, dso_ds, ws, wt, inv_dt_ds,
                                                    vscale[ia] * rvmax, zs, ia, arc_t_exact,
                                                    narrow, so_ds, 1.f / 1024.f);
            if ((lv.inwin && lv.w2 < 32768.f) || !(lv.bs >= 1.f && lv.bs < nsf)) continue;
            ColView cv;
            cv.isf = floorf(lv.bs);
            cv.wb = lv.w2 * (lv.bs - cv.isf);
            cv.wa = lv.w2 - cv.wb;
            cv.mz = lv.mz;
            cv.wt = lv.wt;
            const float tlo = fminf(lv.blo, lv.bhi), thi = fmaxf(lv.blo, lv.bhi);
            const bool fast = tlo >= 1.f && thi < ntf;
            gather_view<kSZ, F_F16S>(cv, z, lv.wt, ntf, fast, view_rsrc(qb + ia * viewb, vbytes),
                                     rowb, acc);
        }
    }

    if (inxy) {
#pragma unroll
        for (int j = 0; j < kSZ; ++j)
            if (iz0 + j < nz)
                image[col + nxy * (iz0 + j)] = ((inbits >> j) & 1u) ? acc[j] * vmax : 0.f;
    }
}

template <bool ARC, bool OPT>
__global__ __launch_bounds__(256, SGX_PACKED_WAVES) void cbct_packed_kernel(
    const void* __restrict__ q, int ns, int nt, int pitch, int nviews, int64_t view_stride,
    const ViewCS* __restrict__ cs, const float* __restrict__ vscale,
    const float* __restrict__ xds, const float* __restrict__ yds,
    const float* __restrict__ zds, int nx, int ny, int nz, const uint8_t* __restrict__ mask,
    float* __restrict__ image, float dsd_ds, float dso_ds, float ws, float wt,
    float inv_dt_ds, const int4* __restrict__ win, const float* __restrict__ zs,
    int arc_t_exact, int narrow, float so_ds)
{
    __shared__ __attribute__((aligned(16))) char lds[2 * kPkBuf];
    cbct_packed_block<ARC, OPT>(q, ns, nt, pitch, nviews, view_stride, cs, vscale, xds, yds, zds,
                                nx, ny, nz, mask, image, dsd_ds, dso_ds, ws, wt, inv_dt_ds, win,
                                zs, arc_t_exact, narrow, so_ds, blockIdx.x, lds);
}

// Batch form (as cbct_queue_kernel).
template <bool ARC, bool OPT>
__global__ __launch_bounds__(256, SGX_PACKED_WAVES) void cbct_packed_queue_kernel(
    const CbctItem* __restrict__ items, const int2* __restrict__ tasks)
{
    __shared__ __attribute__((aligned(16))) char lds[2 * kPkBuf];
    const int2 tk = tasks[blockIdx.x];
    using CItem = const __attribute__((address_space(4))) CbctItem;
    const CItem& it = ((CItem*)(items))[tk.x];
    cbct_packed_block<ARC, OPT>(it.q, it.ns, it.nt, it.pitch, it.nviews, it.view_stride, it.cs,
                                it.vscale, it.xds, it.yds, it.zds, it.nx, it.ny, it.nz, it.mask,
                                it.image, it.dsd_ds, it.dso_ds, it.ws, it.wt, it.inv_dt_ds,
                                it.win, it.zs, it.arc_t_exact, it.narrow, it.so_ds, tk.y, lds);
}

// An arc detector whose bins (1 <= bs < ns, |gamma| = |bs - ws| / dsd_ds)
// span less than 0.75 rad (43 degrees) on both sides takes atan2_narrow.
int arc_narrow(const CbctArgs& a)
{
    return a.is_arc && std::max(std::fabs(1.f - a.ws), std::fabs((float)a.ns - a.ws)) <
                           0.75f * a.dsd_ds;
}

// The staged kernel's block form for `a` (the packed kernel has the 16-deep
// blocks).
int zform_of(const CbctArgs& a)
{
    return (a.fmt == SGX_Q_FMT_F16S && a.f16_packed) ? kZ16 : a.zform;
}

// Blocks of the staged (and packed) kernels: every 16 x 16 column tile
// times the z blocks (16 voxels; the mirror form: 8 voxels z > 0 and their
// images, the same count for its even nz / 16; the Z8 form: 8 voxels).
int64_t staged_blocks(const CbctArgs& a)
{
    const int64_t tiles = (int64_t)((a.nx + kSP - 1) / kSP) * ((a.ny + kSP - 1) / kSP);
    const int zf = zform_of(a);
    const int zb = (zf == kZ8 || zf == kZ8W) ? kSZm : kSZ;
    return tiles * ((a.nz + zb - 1) / zb);
}

template <int ZB, bool ARC, int FMT, bool OPT>
void launch_gather(const CbctArgs& a, hipStream_t st)
{
    const int64_t tiles = (int64_t)((a.nx + kPatch - 1) / kPatch) * ((a.ny + kPatch - 1) / kPatch);
    const int64_t zblocks = (a.nz + ZB * kWaves - 1) / (ZB * kWaves);
    if (tiles * zblocks >= (int64_t(1) << 31))
        throw Error(SGX_ERR_UNSUPPORTED, "cbct_back: volume too large for one launch");
    hipLaunchKernelGGL((cbct_back_kernel<ZB, ARC, FMT, OPT>), dim3((unsigned)(tiles * zblocks)),
                       dim3(64 * kWaves), 0, st, a.q, a.ns, a.nt, a.pitch, a.nviews,
                       a.view_stride, a.cs, a.vscale, a.xds, a.yds, a.zds, a.nx, a.ny, a.nz,
                       a.mask, a.image, a.dsd_ds, a.dso_ds, a.ws, a.wt, a.inv_dt_ds, a.zs,
                       a.arc_t_exact ? 1 : 0, arc_narrow(a), a.so_ds);
    SGX_HIP(hipGetLastError());
}

template <bool ARC, int FMT, bool OPT, int ZF>
void launch_staged(const CbctArgs& a, hipStream_t st)
{
    const int64_t nb = staged_blocks(a);
    if (nb >= (int64_t(1) << 31))
        throw Error(SGX_ERR_UNSUPPORTED, "cbct_back: volume too large for one launch");
    hipLaunchKernelGGL((cbct_staged_kernel<ARC, FMT, OPT, ZF>), dim3((unsigned)nb), dim3(256), 0,
                       st,
                       a.q, a.ns, a.nt, a.pitch, a.nviews, a.view_stride, a.cs, a.vscale, a.xds,
                       a.yds, a.zds, a.nx, a.ny, a.nz, a.mask, a.image, a.dsd_ds, a.dso_ds,
                       a.ws, a.wt, a.inv_dt_ds, a.win, a.zs, a.arc_t_exact ? 1 : 0,
                       arc_narrow(a), a.so_ds, a.win_stride > 0 ? a.win_stride : a.nviews,
                       a.accumulate ? 1 : 0);
    SGX_HIP(hipGetLastError());
}

template <bool ARC, bool OPT>
void launch_packed(const CbctArgs& a, hipStream_t st)
{
    const int64_t nb = staged_blocks(a);
    if (nb >= (int64_t(1) << 31))
        throw Error(SGX_ERR_UNSUPPORTED, "cbct_back: volume too large for one launch");
    hipLaunchKernelGGL((cbct_packed_kernel<ARC, OPT>), dim3((unsigned)nb), dim3(256), 0, st,
                       a.q, a.ns, a.nt, a.pitch, a.nviews, a.view_stride, a.cs, a.vscale, a.xds,
                       a.yds, a.zds, a.nx, a.ny, a.nz, a.mask, a.image, a.dsd_ds, a.dso_ds,
                       a.ws, a.wt, a.inv_dt_ds, a.win, a.zs, a.arc_t_exact ? 1 : 0,
                       arc_narrow(a), a.so_ds);
    SGX_HIP(hipGetLastError());
}

}  // namespace

bool cbct_mirror_ok(const CbctArgs& a, bool z_symmetric)
{
    // bt(-z) = nt + 1 - bt(z) needs the source in the plane z = 0 and the t
    // centre on the middle row: wt = (nt + 1) / 2 exactly
    return z_symmetric && a.zs == nullptr && a.nz % (2 * kSZm) == 0 &&
           a.wt == (float)((a.nt + 1) / 2.0) && !(a.fmt == SGX_Q_FMT_F16S && a.f16_packed);
}

namespace {

// The scaled-FP16 form takes the packed kernel when asked for it.
bool packed_form(const CbctArgs& a)
{
    return a.win && a.fmt == SGX_Q_FMT_F16S && a.f16_packed;
}

template <int FMT, bool OPT> void launch_opt(const CbctArgs& a, hipStream_t st)
{
    if (packed_form(a)) {
        if (a.is_arc) launch_packed<true, OPT>(a, st);
        else launch_packed<false, OPT>(a, st);
    } else if (a.win && a.zform == kZMirror) {
        if (a.is_arc) launch_staged<true, FMT, OPT, kZMirror>(a, st);
        else launch_staged<false, FMT, OPT, kZMirror>(a, st);
    } else if (a.win && a.zform == kZ8) {
        if (a.is_arc) launch_staged<true, FMT, OPT, kZ8>(a, st);
        else launch_staged<false, FMT, OPT, kZ8>(a, st);
    } else if (a.win && a.zform == kZ8W) {
        if (a.is_arc) launch_staged<true, FMT, OPT, kZ8W>(a, st);
        else launch_staged<false, FMT, OPT, kZ8W>(a, st);
    } else if (a.win) {
        if (a.is_arc) launch_staged<true, FMT, OPT, kZ16>(a, st);
        else launch_staged<false, FMT, OPT, kZ16>(a, st);
    } else {
        if (a.is_arc) launch_gather<8, true, FMT, OPT>(a, st);
        else launch_gather<8, false, FMT, OPT>(a, st);
    }
}

template <int FMT> void launch_fmt(const CbctArgs& a, hipStream_t st)
{
    if (a.zs || a.arc_t_exact || a.so_ds != 0.f) launch_opt<FMT, true>(a, st);
    else launch_opt<FMT, false>(a, st);
}

__global__ void view_trig_kernel(const double* __restrict__ betas, int64_t n, int64_t stride,
                                 int mojette, ViewCS* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) {
        double s, c;
        sincos(betas[i * stride], &s, &c);
        if (mojette) {
            // the view's radial spacing over d: max(|cos|, |sin|) (_d_moj)
            const double f = fmax(fabs(c), fabs(s));
            c /= f;
            s /= f;
        }
        out[i].c = (float)c;
        out[i].s = (float)s;
    }
}

}  // namespace

void launch_view_trig(const double* betas, int64_t n, int64_t stride, ViewCS* out,
                      hipStream_t st, bool mojette)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(view_trig_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       betas, n, stride, mojette ? 1 : 0, out);
    SGX_HIP(hipGetLastError());
}

int64_t cbct_window_entries(const CbctArgs& a) { return staged_blocks(a) * a.nviews; }

void launch_cbct_windows(const CbctArgs& a, int4* win, hipStream_t st)
{
    const int64_t nb = staged_blocks(a), n = nb * a.nviews;
    if (n <= 0) return;
    hipLaunchKernelGGL(cbct_window_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       (int)nb, a.nviews, a.cs, a.xds, a.yds, a.zds, a.nx, a.ny, a.nz, a.ns, a.nt,
                       a.dsd_ds, a.dso_ds, a.ws, a.wt, a.inv_dt_ds, a.is_arc ? 1 : 0, a.zs,
                       a.arc_t_exact ? 1 : 0, a.so_ds, zform_of(a), win);
    SGX_HIP(hipGetLastError());
}

int cbct_variant(const CbctArgs& a)
{
    const int fmt = a.fmt == SGX_Q_FMT_BF16 ? F_BF16 : a.fmt == SGX_Q_FMT_F16S ? F_F16S : F_F32;
    const bool opt = a.zs || a.arc_t_exact || a.so_ds != 0.f;
    return (a.is_arc ? 1 : 0) + (opt ? 2 : 0) + 4 * fmt + (a.win ? 0 : kGatherVariant) +
           (packed_form(a) ? kPackedVariant : 0) +
           (a.win && zform_of(a) == kZMirror ? kMirrorVariant : 0) +
           (a.win && zform_of(a) == kZ8 ? kZ8Variant : 0) +
           (a.win && zform_of(a) == kZ8W ? kZ8WVariant : 0);
}

CbctItem cbct_item(const CbctArgs& a)
{
    CbctItem it{};
    it.q = a.q;
    it.cs = a.cs;
    it.vscale = a.vscale;
    it.xds = a.xds;
    it.yds = a.yds;
    it.zds = a.zds;
    it.mask = a.mask;
    it.image = a.image;
    it.win = a.win;
    it.zs = a.zs;
    it.view_stride = a.view_stride;
    it.ns = a.ns;
    it.nt = a.nt;
    it.pitch = a.pitch;
    it.nviews = a.nviews;
    it.nx = a.nx;
    it.ny = a.ny;
    it.nz = a.nz;
    it.arc_t_exact = a.arc_t_exact ? 1 : 0;
    it.narrow = arc_narrow(a);
    it.dsd_ds = a.dsd_ds;
    it.dso_ds = a.dso_ds;
    it.ws = a.ws;
    it.wt = a.wt;
    it.inv_dt_ds = a.inv_dt_ds;
    it.so_ds = a.so_ds;
    return it;
}

int64_t cbct_item_blocks(const CbctArgs& a)
{
    if (a.win) return staged_blocks(a);
    const int64_t tiles = (int64_t)((a.nx + kPatch - 1) / kPatch) * ((a.ny + kPatch - 1) / kPatch);
    return tiles * ((a.nz + 8 * kWaves - 1) / (8 * kWaves));
}


// Batch form of the gather kernel (as cbct_queue_kernel).
template <bool ARC, int FMT, bool OPT>
__global__ __launch_bounds__(64 * kWaves) void cbct_gather_queue_kernel(
    const CbctItem* __restrict__ items, const int2* __restrict__ tasks)
{
    const int2 tk = tasks[blockIdx.x];
    using CItem = const __attribute__((address_space(4))) CbctItem;
    const CItem& it = ((CItem*)(items))[tk.x];
    cbct_gather_block<8, ARC, FMT, OPT>(it.q, it.ns, it.nt, it.pitch, it.nviews, it.view_stride,
                                        it.cs, it.vscale, it.xds, it.yds, it.zds, it.nx, it.ny,
                                        it.nz, it.mask, it.image, it.dsd_ds, it.dso_ds, it.ws,
                                        it.wt, it.inv_dt_ds, it.zs, it.arc_t_exact, it.narrow,
                                        it.so_ds, tk.y);
}

template <bool ARC, int FMT, bool OPT, int ZF>
static const void* cbq_fn()
{
    return reinterpret_cast<const void*>(&cbct_queue_kernel<ARC, FMT, OPT, ZF>);
}

template <bool ARC, int FMT, bool OPT>
static const void* cbg_fn()
{
    return reinterpret_cast<const void*>(&cbct_gather_queue_kernel<ARC, FMT, OPT>);
}

template <bool ARC, bool OPT>
static const void* cbp_fn()
{
    return reinterpret_cast<const void*>(&cbct_packed_queue_kernel<ARC, OPT>);
}

static const void* cbct_queue_fn(int variant)
{
    const bool arc = variant & 1, opt = variant & 2;
    const int fmt = (variant >> 2) & 3;
    if (variant & kPackedVariant)
        return arc ? (opt ? cbp_fn<true, true>() : cbp_fn<true, false>())
                   : (opt ? cbp_fn<false, true>() : cbp_fn<false, false>());
    if (variant & kGatherVariant) {
#define SGX_CBG(F)                                                                   \
    return arc ? (opt ? cbg_fn<true, F, true>() : cbg_fn<true, F, false>())          \
               : (opt ? cbg_fn<false, F, true>() : cbg_fn<false, F, false>())
        if (fmt == F_BF16) SGX_CBG(F_BF16);
        if (fmt == F_F16S) SGX_CBG(F_F16S);
        SGX_CBG(F_F32);
#undef SGX_CBG
    }
#define SGX_CBQ2(F, M)                                                               \
    return arc ? (opt ? cbq_fn<true, F, true, M>() : cbq_fn<true, F, false, M>())    \
               : (opt ? cbq_fn<false, F, true, M>() : cbq_fn<false, F, false, M>())
#define SGX_CBQ(F)                                                                   \
    do {                                                                             \
        if (variant & kMirrorVariant) SGX_CBQ2(F, kZMirror);                         \
        if (variant & kZ8Variant) SGX_CBQ2(F, kZ8);                                  \
        if (variant & kZ8WVariant) SGX_CBQ2(F, kZ8W);                                \
        SGX_CBQ2(F, kZ16);                                                           \
    } while (0)
    if (fmt == F_BF16) SGX_CBQ(F_BF16);
    if (fmt == F_F16S) SGX_CBQ(F_F16S);
    SGX_CBQ(F_F32);
#undef SGX_CBQ
#undef SGX_CBQ2
}

void launch_cbct_queue(int variant, const CbctItem* items, const int2* tasks, int ntasks,
                       hipStream_t st)
{
    if (ntasks <= 0) return;
    void* args[] = {&items, &tasks};
    SGX_HIP(hipLaunchKernel(cbct_queue_fn(variant), dim3((unsigned)ntasks), dim3(64 * kWaves),
                            args, 0, st));
}

void launch_cbct_back(const CbctArgs& a, hipStream_t st)
{
    // Byte offsets within one view are 32-bit.
    if ((int64_t)a.pitch * a.nt * 4 + 64 >= (int64_t(1) << 31))
        throw Error(SGX_ERR_UNSUPPORTED, "cbct_back: one view must be below 2 GiB");
    if (a.pitch < a.ns || a.pitch % 4 != 0)
        throw Error(SGX_ERR_ARG, "cbct_back: the row pitch must be >= ns and a multiple of 4");
    if (a.fmt == SGX_Q_FMT_BF16) launch_fmt<F_BF16>(a, st);
    else if (a.fmt == SGX_Q_FMT_F16S) launch_fmt<F_F16S>(a, st);
    else launch_fmt<F_F32>(a, st);
}

}  // namespace sgx
