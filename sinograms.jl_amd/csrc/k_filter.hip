// K4 + K5: ramp filtering of sinogram / projection rows, with the pre-weights
// (fan sinogram weight, FDK cylinder weight, Parker short-scan weight in FP64,
// angular weight) fused into the operand load.  For gfx950.  Two forms: an
// FFT on registers at npad = 512, 1024 or 2048 (the plans' usual case, below the
// Toeplitz form), and for any other npad a batched Toeplitz contraction on
// MFMA.
//
// Replaces fbp_sino_filter (src/fbp/filter.jl:122-147): the reference zero-pads
// every column to npad, multiplies its FFT by the real, even Hk and keeps the
// real part of rows 1..nb.  That is exactly the circular convolution
//   out[n, c] = sum_{m < nb} g[(n - m) mod npad] in[m, c],   g = real(ifft(Hk)),
// i.e. out = T * B with the nb x nb Toeplitz matrix T[n, m] = g[(n-m) mod npad]
// and B the (weighted) input, nb x ncols, column-major.
//
// Tiling.  A block (4 waves) owns 32 columns and a band of row tiles of 32;
// each wave keeps RT 32x32 accumulators.  The input is streamed in chunks of
// 64 rows (m): one coalesced pass per chunk from HBM, weighted, then staged in
// LDS (FP32, or pre-split into bf16 hi / lo halves).  T is never stored: the
// block keeps the band of g it needs in LDS and every A fragment is read from
// it at index n - m.
//   FP32 form:  v_mfma_f32_32x32x2_f32.
//   BF16 form:  x = hi + lo with hi = bf16 truncation, lo = bf16(x - hi);
//               T*B ~ Th*Bh + Th*Bl + Tl*Bh on v_mfma_f32_32x32x16_bf16 with
//               FP32 accumulation (split-BF16x3).
// Output rows n >= nb (the reference's `extra` rows) are written as 0.
// Outputs: FP32, BF16 (RNE), or FP16 scaled by a per-view power of two
// (F16S: the scale bounds the view's output, from a pre-pass over its
// weighted input), column-major (nb + extra, ncols).

#include "sgx_internal.h"
#include "sgx.h"

#include <hip/hip_fp16.h>

#include <cmath>
#include <cstdlib>

namespace sgx {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kKC = 64;            // input rows staged per chunk
// output element types, column-major (nb + extra, ncols)
enum { OUT_F32 = 0, OUT_BF16 = 1, OUT_F16S = 2 };
constexpr int kBPadF = kKC + 1;    // FP32 staging stride (conflict-free column reads)
constexpr int kBPadH = kKC + 8;    // bf16 staging stride (16-byte aligned fragments)
constexpr int kStageHalves = 2 * 32 * kBPadH;   // one staging buffer: hi and lo parts
constexpr int kStageFloats = 32 * kBPadF;       // one FP32 staging buffer

struct Params {
    const float* in;
    int nb;
    int64_t ncols;
    int nt;
    const float* g;
    int npad;
    int nrows_out;          // nb + extra
    void* out;
    float scale;
    const float* wbin;
    const float* wview;
    int cyl;
    float cyl_dsd, cyl_dso;
    const float* s_axis;
    const float* t_axis;
    const float* ptab;            // (nb, na) short-scan Parker weights, or null
    const unsigned int* vbound;   // F16S: per-view max |weighted input| (float bits)
    float g1;                     // F16S: sum |g|
    float wmax;                   // F16S on the FFT form: bound of |bin x cylinder weight|
    float gscale;                 // split-FP16: power of two applied to g
    float* vscale;                // F16S: per-view output scale (written)
    int row_tiles_per_block;
    const float* fft;             // FFT form: its table, or null
    int fft_n;                    // FFT form: transform size (1024 or 2048)
    const float* wtab;            // FFT form, cone beam: (nb, nt) bin x cylinder weights
};

// F16S output scale of a view: the power of two s >= bound * g1 / 2^14, so
// that |q| / s <= 2^14 (|q| <= bound * g1) and differences of neighbours stay
// below 2^15, inside FP16's range.  Scaling by a power of two is exact.
__device__ __forceinline__ float f16s_scale(unsigned int bound_bits, float g1)
{
    const float b = __uint_as_float(bound_bits) * g1;
    if (!(b > 0.f)) return 1.f;
    int e;
    (void)frexpf(b, &e);                 // b <= 2^e
    return ldexpf(1.f, e - 14);
}

__device__ __forceinline__ double parker_w(double gg, double bb, double gm)
{
    // parker_weight_fan_short (src/fbp/parker.jl:96-106), regions in the
    // reference's order; 0 outside, and 0 at exactly bb == pi - 2 gg.
    const double pi = 3.14159265358979323846;
    double w = 0.0;
    if (bb < 2.0 * (gm - gg)) {
        const double s = sin(pi / 2.0 * (bb / (2.0 * (gm - gg))));
        w = s * s;
    } else if (bb < pi - 2.0 * gg) {
        w = 1.0;
    }
    if (pi - 2.0 * gg < bb && bb <= pi + 2.0 * gm) {
        const double s = sin(pi / 2.0 * ((pi + 2.0 * gm - bb) / (2.0 * (gm + gg))));
        w = s * s;
    }
    return w;
}

__device__ __forceinline__ float weight(const Params& p, int m, int64_t c)
{
    float w = p.scale;
    if (p.wbin) w *= p.wbin[m];
    if (p.cyl) {
        const int t = (int)(c % p.nt);
        const float s = p.s_axis[m], tt = p.t_axis[t];
        if (p.cyl == 1) {
            w *= p.cyl_dso / sqrtf(s * s + tt * tt + p.cyl_dsd * p.cyl_dsd);
        } else {
            const float r = tt / p.cyl_dsd;
            w *= (p.cyl_dso / p.cyl_dsd) * cosf(s / p.cyl_dsd) / sqrtf(r * r + 1.f);
        }
    }
    if (p.wview) w *= p.wview[c / p.nt];
    if (p.ptab) w *= p.ptab[(c / p.nt) * p.nb + m];
    return w;
}

__device__ __forceinline__ __bf16 bf16_trunc(float x)
{
    return __builtin_bit_cast(__bf16, (uint16_t)(__float_as_uint(x) >> 16));
}

__device__ __forceinline__ __bf16 bf16_rne(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __builtin_bit_cast(__bf16, (uint16_t)(u >> 16));
}

__device__ __forceinline__ float bf16_to_f(__bf16 h)
{
    return __uint_as_float((uint32_t)__builtin_bit_cast(uint16_t, h) << 16);
}

// Products: FP32 MFMA, split-BF16x3 or split-FP16x3 (x = hi + lo per operand,
// hi*hi + hi*lo + lo*hi accumulated in FP32).  FP16 operands are scaled by
// powers of two first (g by p.gscale, each view's input by 2^14 over the
// power of two above its bound) so that both parts stay in FP16's normal
// range; the epilogue divides the scales out exactly.
enum { M_F32 = 0, M_BF16X3 = 1, M_F16X3 = 2 };
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

// Detector widths: the common instance takes nb <= 2048, the wide one nb <=
// 8192 (its band of g and bin factors use about 154 KB of LDS, one block per
// CU).  The band of a block is 4 RT 32 + kp - 1 elements long.
constexpr int kNbSmall = 2048;
constexpr int kNbWide = 8192;
__host__ __device__ constexpr int band_max(int nbmax) { return 4 * 8 * 32 + nbmax; }

__device__ __forceinline__ float in_scale(const unsigned int* vbound, int view)
{
    const float b = __uint_as_float(vbound[view]);
    if (!(b > 0.f)) return 1.f;
    int e;
    (void)frexpf(b, &e);                 // b <= 2^e
    return ldexpf(1.f, 14 - e);
}

template <int MODE> __device__ __forceinline__ void split2(float x, uint16_t& hi, uint16_t& lo)
{
    if (MODE == M_BF16X3) {
        const __bf16 h = bf16_trunc(x);
        hi = __builtin_bit_cast(uint16_t, h);
        lo = __builtin_bit_cast(uint16_t, bf16_rne(x - bf16_to_f(h)));
    } else {
        const _Float16 h = (_Float16)x;
        hi = __builtin_bit_cast(uint16_t, h);
        lo = __builtin_bit_cast(uint16_t, (_Float16)(x - (float)h));
    }
}

template <int MODE, int OUT, int RT, int NBMAX>
__global__ __launch_bounds__(256) void toeplitz_kernel(Params p)
{
    constexpr bool SPLIT = MODE != M_F32;
    constexpr int kBandMax = band_max(NBMAX);            // longest band of g
    // FP32: the band of g as floats.  Split modes: hi / lo parts of the band,
    // reversed (R[i] = band[glen-1-i], so a fragment's 8 taps ascend), each
    // in two copies shifted by one element so every fragment starts on a
    // dword: band16[(2*part + copy) * kBandMax + i] = R[i + copy].
    __shared__ __attribute__((aligned(16))) unsigned char gbuf[SPLIT ? 4 * kBandMax * 2
                                                                     : (kBandMax + 64) * 4];
    // staging: two buffers (chunk c in buffer c & 1), one for RT = 1 (a
    // larger static LDS block measurably slows those small launches)
    __shared__ __attribute__((aligned(16))) unsigned char bstage[(RT == 1 ? 1 : 2) *
                                                                 (SPLIT ? 2 * 32 * kBPadH * 2
                                                                        : 32 * kBPadF * 4)];
    // per-bin factors of the fused weight: wbin (and the arc cylinder weight's
    // cos(s/dsd)), and s^2 for the flat cylinder weight
    __shared__ float bin_a[NBMAX], bin_s2[NBMAX];
    float* gl = reinterpret_cast<float*>(gbuf);
    uint16_t* band16 = reinterpret_cast<uint16_t*>(gbuf);
    float* bs_f = reinterpret_cast<float*>(bstage);
    uint16_t* bs_hi = reinterpret_cast<uint16_t*>(bstage);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t c0 = (int64_t)blockIdx.x * 32;
    const int tile0 = blockIdx.y * 4 * RT;              // first row tile of the block
    const int n_lo = tile0 * 32;
    const int kp = ((p.nb + kKC - 1) / kKC) * kKC;       // padded K
    const int glen = 4 * RT * 32 + kp - 1;

    // band: band[j] = g[(n_lo - (kp - 1) + j) mod npad];  T[n, m] = band[n - n_lo - m + kp - 1]
    for (int j = tid; j < glen; j += 256) {
        int k = (n_lo - (kp - 1) + j) % p.npad;
        if (k < 0) k += p.npad;
        if (SPLIT) {
            uint16_t hi, lo;
            split2<MODE>(MODE == M_F16X3 ? p.g[k] * p.gscale : p.g[k], hi, lo);
            const int i = glen - 1 - j;
            band16[0 * kBandMax + i] = hi;
            band16[2 * kBandMax + i] = lo;
            if (i >= 1) {
                band16[1 * kBandMax + i - 1] = hi;
                band16[3 * kBandMax + i - 1] = lo;
            }
        } else {
            gl[j] = p.g[k];
        }
    }

    for (int m = tid; m < p.nb; m += 256) {
        float a = p.wbin ? p.wbin[m] : 1.f;
        if (p.cyl == 2) a *= cosf(p.s_axis[m] / p.cyl_dsd);
        bin_a[m] = a;
        bin_s2[m] = p.cyl == 1 ? p.s_axis[m] * p.s_axis[m] : 0.f;
    }
    __syncthreads();          // the first chunk's weights read other threads' factors

    floatx16 acc[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[r][i] = 0.f;

    const int col = lane & 31, half = lane >> 5;
    const int scol = tid >> 3, skk = (tid & 7) * 8;      // staging: column, row offset

    // Per-thread staging column: its view / row index and every factor of
    // the fused weight that does not depend on the bin m, formed once.
    const int64_t sc = c0 + scol;
    const bool cvalid = sc < p.ncols;
    int sview = 0, st = 0;
    if (cvalid) {
        sview = (int)(sc / p.nt);
        st = (int)(sc - (int64_t)sview * p.nt);
    }
    float colw = p.scale;
    if (cvalid && p.wview) colw *= p.wview[sview];
    if (MODE == M_F16X3 && cvalid) colw *= in_scale(p.vbound, sview);
    // Cylinder weight: flat dso / sqrt(s^2 + t^2 + dsd^2) is
    // cyl_k * rsqrt(bin_s2[m] + cyl_c) with cyl_k = dso; arc (dso/dsd)
    // cos(s/dsd) / sqrt((t/dsd)^2 + 1) is bin_a[m] (the cosine) times cyl_k.
    // The weight's branches are uniform (launch parameters).
    const bool cyl_flat = p.cyl == 1;
    float cyl_c = 1.f, cyl_k = 1.f;
    if (cvalid && p.cyl == 1) {
        const float tt = p.t_axis[st];
        cyl_c = tt * tt + p.cyl_dsd * p.cyl_dsd;
        cyl_k = p.cyl_dso;
    } else if (cvalid && p.cyl == 2) {
        const float r = p.t_axis[st] / p.cyl_dsd;
        cyl_k = (p.cyl_dso / p.cyl_dsd) / sqrtf(r * r + 1.f);
    }
    const float* src = p.in + (int64_t)p.nb * (cvalid ? sc : 0);
    const float* prow = (p.ptab && cvalid) ? p.ptab + (int64_t)sview * p.nb : nullptr;

    // rows m0 + skk .. + 7 of this thread's column, and their Parker factors
    auto load8 = [&](int m0, float (&v)[8], float (&pw)[8]) {
        const int m = m0 + skk;
        if (cvalid && m + 8 <= p.nb) {
            __builtin_memcpy(v, src + m, sizeof(float) * 8);
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = (cvalid && m + e < p.nb) ? src[m + e] : 0.f;
        }
        if (!p.ptab) return;                // pw is read only with a table
        if (prow && m + 8 <= p.nb) {
            __builtin_memcpy(pw, prow + m, sizeof(float) * 8);
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) pw[e] = (prow && m + e < p.nb) ? prow[m + e] : 1.f;
        }
    };
    // the fused weight (rows past nb become 0)
    auto weigh = [&](int m0, float (&v)[8], const float (&pw)[8]) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int m = m0 + skk + e;
            const int mm = min(m, p.nb - 1);
            float w = colw * bin_a[mm];
            if (cyl_flat) w *= cyl_k * rsqrtf(bin_s2[mm] + cyl_c);
            else if (p.cyl) w *= cyl_k;
            if (p.ptab) w *= pw[e];
            v[e] = (cvalid && m < p.nb) ? v[e] * w : 0.f;
        }
    };
    auto put = [&](int buf, const float (&v)[8]) {
        if (SPLIT) {
            uint16_t hi[8], lo[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) split2<MODE>(v[e], hi[e], lo[e]);
            uint16_t* bh = bs_hi + buf * kStageHalves;
            __builtin_memcpy(bh + scol * kBPadH + skk, hi, sizeof(hi));   // 16-byte stores
            __builtin_memcpy(bh + 32 * kBPadH + scol * kBPadH + skk, lo, sizeof(lo));
        } else {
            float* bf = bs_f + buf * kStageFloats;
#pragma unroll
            for (int e = 0; e < 8; ++e) bf[scol * kBPadF + skk + e] = v[e];
        }
    };

    // Split modes: per lane, the start of its A fragments in the reversed
    // band and the copy that puts it on a dword (its parity is the same for
    // every row tile and K step).
    const int par = (glen - col - kp) & 1;
    const uint16_t* ahi = band16 + (0 + par) * kBandMax;
    const uint16_t* alo = band16 + (2 + par) * kBandMax;

    auto mma = [&](int buf, int m0) {
        if (SPLIT) {
            const uint16_t* bh = bs_hi + buf * kStageHalves;
            const uint16_t* bl = bh + 32 * kBPadH;
#pragma unroll
            for (int ks = 0; ks < kKC / 16; ++ks) {
                const int kb = ks * 16 + half * 8;
                const uint4 bh4 = *reinterpret_cast<const uint4*>(bh + col * kBPadH + kb);
                const uint4 bl4 = *reinterpret_cast<const uint4*>(bl + col * kBPadH + kb);
#pragma unroll
                for (int r = 0; r < RT; ++r) {
                    const int n = (tile0 + wave * RT + r) * 32 + col;
                    // A[n][m0 + kb + j] = R[s + j], s = glen - 1 - (n - n_lo - m0 - kb + kp - 1)
                    const int s = glen - (n - n_lo) + m0 + kb - kp;
                    const uint32_t* a32 = reinterpret_cast<const uint32_t*>(ahi + (s - par));
                    const uint32_t* l32 = reinterpret_cast<const uint32_t*>(alo + (s - par));
                    const uint4 ah4 = make_uint4(a32[0], a32[1], a32[2], a32[3]);
                    const uint4 al4 = make_uint4(l32[0], l32[1], l32[2], l32[3]);
                    if (MODE == M_BF16X3) {
                        const bf16x8 ah = __builtin_bit_cast(bf16x8, ah4);
                        const bf16x8 al = __builtin_bit_cast(bf16x8, al4);
                        const bf16x8 b_h = __builtin_bit_cast(bf16x8, bh4);
                        const bf16x8 b_l = __builtin_bit_cast(bf16x8, bl4);
                        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, b_h, acc[r], 0, 0, 0);
                        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, b_l, acc[r], 0, 0, 0);
                        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, b_h, acc[r], 0, 0, 0);
                    } else {
                        const halfx8 ah = __builtin_bit_cast(halfx8, ah4);
                        const halfx8 al = __builtin_bit_cast(halfx8, al4);
                        const halfx8 b_h = __builtin_bit_cast(halfx8, bh4);
                        const halfx8 b_l = __builtin_bit_cast(halfx8, bl4);
                        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, b_h, acc[r], 0, 0, 0);
                        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, b_l, acc[r], 0, 0, 0);
                        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, b_h, acc[r], 0, 0, 0);
                    }
                }
            }
        } else {
            const float* bf = bs_f + buf * kStageFloats;
#pragma unroll 4
            for (int ks = 0; ks < kKC / 2; ++ks) {
                const int k = ks * 2 + half;
                const float b = bf[col * kBPadF + k];
#pragma unroll
                for (int r = 0; r < RT; ++r) {
                    const int n = (tile0 + wave * RT + r) * 32 + col;
                    const float a = gl[(n - n_lo) - (m0 + k) + kp - 1];
                    acc[r] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[r], 0, 0, 0);
                }
            }
        }
    };

    if (RT == 1) {
        // Small launches (one block per CU): single-buffered chunks, the
        // next chunk's loads in flight while the current one is weighted,
        // staged and multiplied out (the shortest latency per chunk).
        float vn[8], pn[8];
        load8(0, vn, pn);
        for (int m0 = 0; m0 < p.nb; m0 += kKC) {
            float v[8], pw[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                v[e] = vn[e];
                pw[e] = pn[e];
            }
            if (m0 + kKC < p.nb) load8(m0 + kKC, vn, pn);
            weigh(m0, v, pw);
            __syncthreads();
            put(0, v);
            __syncthreads();
            mma(0, m0);
        }
    } else {
        // Double-buffered chunks: chunk c is multiplied out of buffer c & 1
        // while chunk c + 1 is weighted, split and written to the other
        // buffer, one barrier per chunk; the global loads run a chunk ahead
        // of the weighting.
        float vq[8], pq[8];
        {
            float v0[8], p0[8];
            load8(0, v0, p0);
            load8(kKC, vq, pq);                 // rows past nb load as 0
            weigh(0, v0, p0);
            put(0, v0);
        }
        __syncthreads();
        int buf = 0;
        for (int m0 = 0; m0 < p.nb; m0 += kKC) {
            float v[8], pw[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                v[e] = vq[e];
                pw[e] = pq[e];
            }
            load8(m0 + 2 * kKC, vq, pq);
            weigh(m0 + kKC, v, pw);
            mma(buf, m0);
            put(buf ^ 1, v);                    // past the last chunk: written, never read
            __syncthreads();
            buf ^= 1;
        }
    }

    // store: acc[r][i] holds row 8*(i/4) + 4*half + i%4 of tile r, column `col`
    const int64_t c = c0 + col;
    if (c < p.ncols) {
        // every factor below is a power of two: the rescaling is exact
        float inv = 1.f;
        const int view = (int)(c / p.nt);
        if (MODE == M_F16X3) inv = 1.f / (p.gscale * in_scale(p.vbound, view));
        if (OUT == OUT_F16S) {
            const float sc = f16s_scale(p.vbound[view], p.g1);
            inv /= sc;
            if (c % p.nt == 0 && wave == 0 && half == 0 && blockIdx.y == 0)
                p.vscale[view] = sc;
        }
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const int nbase = (tile0 + wave * RT + r) * 32;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int n0 = nbase + 8 * g + 4 * half;
                if (n0 >= p.nrows_out) continue;
                float v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) v[k] = n0 + k < p.nb ? acc[r][4 * g + k] * inv : 0.f;
                const int64_t o = n0 + (int64_t)p.nrows_out * c;
                const bool whole = n0 + 4 <= p.nrows_out;
                if (OUT == OUT_F32) {
                    float* d = static_cast<float*>(p.out) + o;
                    if (whole) {
                        __builtin_memcpy(d, v, sizeof(v));
                    } else {
                        for (int k = 0; k < 4 && n0 + k < p.nrows_out; ++k) d[k] = v[k];
                    }
                } else {
                    uint16_t h[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        h[k] = OUT == OUT_BF16 ? __builtin_bit_cast(uint16_t, bf16_rne(v[k]))
                                               : __half_as_ushort(__float2half_rn(v[k]));
                    uint16_t* d = static_cast<uint16_t*>(p.out) + o;
                    if (whole) {
                        __builtin_memcpy(d, h, sizeof(h));
                    } else {
                        for (int k = 0; k < 4 && n0 + k < p.nrows_out; ++k) d[k] = h[k];
                    }
                }
            }
        }
    }
}

template <int MODE, int OUT, int RT, int NBMAX>
void launch_rt(const Params& p, hipStream_t st)
{
    const int row_tiles = (p.nrows_out + 31) / 32;
    const int row_blocks = (row_tiles + 4 * RT - 1) / (4 * RT);
    dim3 grid((unsigned)((p.ncols + 31) / 32), (unsigned)row_blocks);
    hipLaunchKernelGGL((toeplitz_kernel<MODE, OUT, RT, NBMAX>), grid, dim3(256), 0, st, p);
    SGX_HIP(hipGetLastError());
}

template <int MODE, int OUT>
void launch_mode(const Params& p, hipStream_t st)
{
    const int row_tiles = (p.nrows_out + 31) / 32;
    const int64_t col_blocks = (p.ncols + 31) / 32;
    // Prefer one block per column band (the input is read and weighted once);
    // split the rows when the columns alone cannot fill the GPU.
    int rt = (row_tiles + 3) / 4;
    if (col_blocks < 1024) rt = 1;
    if (p.nb > kNbSmall) {            // wide detectors: 64+ row tiles, RT 1 or 8
        if (rt <= 1) launch_rt<MODE, OUT, 1, kNbWide>(p, st);
        else launch_rt<MODE, OUT, 8, kNbWide>(p, st);
    } else if (rt <= 1) {
        launch_rt<MODE, OUT, 1, kNbSmall>(p, st);
    } else if (rt <= 2) {
        launch_rt<MODE, OUT, 2, kNbSmall>(p, st);
    } else if (rt <= 4) {
        launch_rt<MODE, OUT, 4, kNbSmall>(p, st);
    } else {
        launch_rt<MODE, OUT, 8, kNbSmall>(p, st);
    }
}

// ---------------------------------------------------------------------------
// FFT form, npad = 512, 1024 or 2048 (512 as a 1024-point transform, see
// fft_filter_tables): the reference's own algorithm (zero-pad,
// fft, .* Hk, ifft, real part), on two columns at once.  z = x_a + i x_b (the
// weighted columns) is transformed; the response enters as
// He = (Hk[k] + Hk[-k]) / 2, the even part of Hk (real(ifft(fft(x) .* Hk)) =
// x circularly convolved with real(ifft(Hk)) = ifft(He)), and since He is real
// and even, ifft(fft(z) .* He) = y_a + i y_b exactly: the real part is column
// a's output and the imaginary part column b's.
//
// A half-wave (32 lanes) owns one transform of N = 32 M points (M = 32 or
// 64), held in registers (M complex per lane), by the four-step method:
//   n = n1 + 32 n2, k = M k1 + k2;   lane n1 holds z[n1 + 32 n2], n2 < M
//   A: M-point DFT over n2 in registers           Y[n1, k2]
//   B: Y[n1, k2] *= W_N^(n1 k2)                    (table in LDS)
//   C: transpose through LDS (wave-local)           lane l holds Y[., l + 32 h]
//   D: M/32 32-point DFTs over n1 in registers    X[M k1 + l + 32 h] at lane l
// The inverse is the same forward transform applied to the swapped spectrum:
// ifft(Y) = swap(fft(swap(Y))) / N with swap(a + ib) = b + ia, and 1/N is in
// the spectrum table.  Element M k1 + l + 32 h of lane l is element
// l + 32 n2 with n2 = (M/32) k1 + h: exactly the layout step A takes, so the
// inverse needs no extra data movement, and its output R[M j1 + 32 h + l]
// lands at lane l: rows stored coalesced.  Both steps are radix-2
// decimation in time, A with its inputs in bit-reversed registers, D on
// registers renamed so that its natural input gives outputs in bit-reversed
// registers: renamings resolved at compile time.  Each twiddled butterfly
// takes the scaled form (6 VALU instead of 8; see bfly).
//
// Bound: VALU (about 680 wave instructions per column at M = 32, both
// transforms and the weights) and HBM (each column read once and written once).

constexpr int kFftCols = 16;          // columns per block: 8 transforms of 2 columns
constexpr int kFftTP = 33;            // transpose pitch (floats): conflict-free rows and columns

template <int BITS>
__host__ __device__ constexpr int bitrev(int i)
{
    int r = 0;
    for (int b = 0; b < BITS; ++b) r |= ((i >> b) & 1) << (BITS - 1 - b);
    return r;
}
constexpr int log2i(int m) { return m <= 1 ? 0 : 1 + log2i(m / 2); }

// cos and sin of 2 pi t / 64, t = 0..31 (W_M^t = W_64^(t 64 / M))
__host__ __device__ constexpr float tw64_cos(int t)
{
    constexpr float c[32] = {
        1.0f, 9.951847267e-01f, 9.807852804e-01f, 9.569403357e-01f, 9.238795325e-01f,
        8.819212643e-01f, 8.314696123e-01f, 7.730104534e-01f, 7.071067812e-01f,
        6.343932842e-01f, 5.555702330e-01f, 4.713967368e-01f, 3.826834324e-01f,
        2.902846773e-01f, 1.950903220e-01f, 9.801714033e-02f, 0.0f, -9.801714033e-02f,
        -1.950903220e-01f, -2.902846773e-01f, -3.826834324e-01f, -4.713967368e-01f,
        -5.555702330e-01f, -6.343932842e-01f, -7.071067812e-01f, -7.730104534e-01f,
        -8.314696123e-01f, -8.819212643e-01f, -9.238795325e-01f, -9.569403357e-01f,
        -9.807852804e-01f, -9.951847267e-01f};
    return c[t];
}
__host__ __device__ constexpr float tw64_sin(int t) { return t <= 16 ? tw64_cos(16 - t) : tw64_cos(t - 16); }

// Radix-2 decimation-in-time butterfly (a, b) -> (a + W b, a - W b) with
// W = W_M^t = c - i s, c = cos(2 pi t / M), s = sin(2 pi t / M), t known at
// compile time.  A general twiddle takes the scaled (tangent) form: with
// |c| >= |s|, W b = c (b_r + r b_i, b_i - r b_r), r = s / c, and the
// outputs are a +- c (u, v): two FMAs for (u, v) and four for the outputs,
// 6 VALU instead of 8 for the product and the two sums (with |s| > |c| the
// same with s and r = c / s).
template <int M>
__device__ __forceinline__ void bfly(float& ar, float& ai, float& br, float& bi, int t)
{
    if (t == 0) {
        const float r = br, i = bi;
        br = ar - r;
        bi = ai - i;
        ar += r;
        ai += i;
    } else if (4 * t == M) {                  // W = -i: W b = (b_i, -b_r)
        const float r = br, i = bi;
        br = ar - i;
        bi = ai + r;
        ar += i;
        ai -= r;
    } else {
        const float c = tw64_cos(t * (64 / M)), sn = tw64_sin(t * (64 / M));
        const float r = br, i = bi;
        float u, v, k;
        if (fabsf(c) >= fabsf(sn)) {
            const float q = sn / c;
            u = fmaf(q, i, r);
            v = fmaf(-q, r, i);
            k = c;
        } else {
            const float q = c / sn;
            u = fmaf(q, r, i);
            v = fmaf(q, i, -r);
            k = sn;
        }
        br = fmaf(-k, u, ar);
        bi = fmaf(-k, v, ai);
        ar = fmaf(k, u, ar);
        ai = fmaf(k, v, ai);
    }
}

// 32-point forward DFT, radix-2 decimation in time on registers renamed at
// compile time: natural input (element n in register n), X[k] ends in
// register bitrev<5>(k) (position p of the textbook DIT lives in register
// bitrev<5>(p)).
__device__ __forceinline__ void dft32_dit_br(float* re, float* im)
{
#pragma unroll
    for (int span = 1; span < 32; span *= 2) {
#pragma unroll
        for (int g = 0; g < 32; g += 2 * span) {
#pragma unroll
            for (int j = 0; j < span; ++j) {
                const int a = bitrev<5>(g + j), b = bitrev<5>(g + j + span);
                bfly<32>(re[a], im[a], re[b], im[b], j * (32 / (2 * span)));
            }
        }
    }
}

// M-point forward DFT, radix-2 decimation in time, in place: element n in
// register bitrev(n), X[k] ends in register k.  HALF: elements M/2 .. M-1
// (the odd registers) are zero and not read.
template <int M, bool HALF>
__device__ __forceinline__ void dft_dit(float (&re)[M], float (&im)[M])
{
#pragma unroll
    for (int span = 1; span < M; span *= 2) {
#pragma unroll
        for (int g = 0; g < M; g += 2 * span) {
#pragma unroll
            for (int j = 0; j < span; ++j) {
                const int a = g + j, b = a + span;
                if (HALF && span == 1) {          // x_b = 0
                    re[b] = re[a];
                    im[b] = im[a];
                    continue;
                }
                bfly<M>(re[a], im[a], re[b], im[b], j * (M / (2 * span)));
            }
        }
    }
}

// Steps B and C: twiddle the step-A output (k2 in register k2, lane n1 = l)
// by W_N^(n1 k2) and transpose it through this half-wave's LDS rows, 32 rows
// of k2 at a time, so that lane l holds element n1 of k2 = l + 32 h in
// register 32 h + n1.  In place: a round's registers are dead once written.
template <int M>
__device__ __forceinline__ void twiddle_transpose(float (&re)[M], float (&im)[M],
                                                  const float* tw_re, const float* tw_im,
                                                  float* tb, int l)
{
#pragma unroll
    for (int k2 = 1; k2 < M; ++k2) {
        const float wr = tw_re[k2 * 32 + l], wi = tw_im[k2 * 32 + l];
        const float r = re[k2], i = im[k2];
        re[k2] = fmaf(r, wr, -(i * wi));
        im[k2] = fmaf(r, wi, i * wr);
    }
    // LDS operations of a wave complete in order; the wave barriers keep the
    // compiler from moving them across one another.
#pragma unroll
    for (int h = 0; h < M / 32; ++h) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            float* v = c ? im : re;
#pragma unroll
            for (int k = 0; k < 32; ++k) tb[k * kFftTP + l] = v[32 * h + k];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int n1 = 0; n1 < 32; ++n1) v[32 * h + n1] = tb[l * kFftTP + n1];
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// Buffer resource over a range: loads past `bytes` return 0 and stores past
// it are dropped, without touching memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t col_rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes,
                                             0x00020000);
}

// N = 32 M (M = 32: npad 512 or 1024, M = 64: npad 2048).  HALF: nb + extra <= N / 2,
// the upper half of every input is zero and of every output column unused.
template <int OUT, bool HALF, int M>
__global__ __launch_bounds__(256, M == 64 ? 2 : 1) void fft_filter_kernel(Params p)
{
    constexpr int N = 32 * M, H = M / 32, BITS = log2i(M);
    // the table: twiddles (re, im), spectrum, bin factors (a, s^2), 0 past nb
    __shared__ __attribute__((aligned(16))) float tab[5 * N];
    __shared__ float tbuf[8 * 32 * kFftTP];                      // a region per transform
    const float* tw_re = tab;                                    // tw[k2 * 32 + n1]
    const float* tw_im = tab + N;
    const float* spec = tab + 2 * N;
    const float* bin_a = tab + 3 * N;
    const float* bin_s2 = tab + 4 * N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    {
        constexpr int R = 5 * N / 1024;                          // float4 per thread
        float4 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = reinterpret_cast<const float4*>(p.fft)[tid + 256 * r];
#pragma unroll
        for (int r = 0; r < R; ++r) reinterpret_cast<float4*>(tab)[tid + 256 * r] = v[r];
    }
    __syncthreads();

    const int l = lane & 31, half = lane >> 5;
    const int f = wave * 2 + half;                        // transform within the block
    float* tb = tbuf + f * 32 * kFftTP;
    const int64_t ca = (int64_t)blockIdx.x * kFftCols + 2 * f, cb = ca + 1;
    const bool va = ca < p.ncols, vb = cb < p.ncols;
    // The wave's 4 columns (its two transforms) under one buffer resource,
    // which must be wave-uniform: reads past them return 0.
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int64_t cw = (int64_t)blockIdx.x * kFftCols + 4 * wv;
    const int64_t rem = p.ncols - cw;
    const int nw = rem <= 0 ? 0 : rem >= 4 ? 4 : (int)rem;

    // Step A input: lane n1 = l holds z[l + 32 j] in register bitrev(j).
    // Every load is issued before the first is used; loads past the wave's
    // columns return 0.
    constexpr int JN = HALF ? M / 2 : M;
    const __amdgpu_buffer_rsrc_t rin =
        col_rsrc(p.in + (int64_t)p.nb * (nw ? cw : 0), 4u * (uint32_t)p.nb * (uint32_t)nw);
    const uint32_t oa = 4u * (uint32_t)(2 * half * p.nb), ob = oa + 4u * (uint32_t)p.nb;
    float re[M], im[M];
#pragma unroll
    for (int j = 0; j < JN; ++j) {
        const uint32_t off = 4u * (uint32_t)(l + 32 * j);
        re[bitrev<BITS>(j)] =
            __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rin, oa + off, 0, 0));
        im[bitrev<BITS>(j)] =
            __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rin, ob + off, 0, 0));
    }
    // Column factors of the fused weight (as in the Toeplitz form): colw, the
    // cylinder constants and the Parker row.
    float colw[2] = {0.f, 0.f}, cyl_c[2] = {1.f, 1.f}, cyl_k[2] = {1.f, 1.f};
    const float* prow[2] = {nullptr, nullptr};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        // a column past ncols takes the last column's factors (its input reads as 0)
        const int64_t c = std::min(h ? cb : ca, p.ncols - 1);
        const int view = (int)(c / p.nt), t = (int)(c - (int64_t)view * p.nt);
        colw[h] = p.scale * (p.wview ? p.wview[view] : 1.f);
        if (p.cyl == 1) {
            const float tt = p.t_axis[t];
            cyl_c[h] = tt * tt + p.cyl_dsd * p.cyl_dsd;
            cyl_k[h] = p.cyl_dso;
        } else if (p.cyl == 2) {
            const float r = p.t_axis[t] / p.cyl_dsd;
            cyl_k[h] = (p.cyl_dso / p.cyl_dsd) / sqrtf(r * r + 1.f);
        }
        if (p.ptab) prow[h] = p.ptab + (int64_t)view * p.nb;
    }
    if (p.ptab) {
#pragma unroll
        for (int j = 0; j < JN; ++j) {
            const int mm = min(l + 32 * j, p.nb - 1);
            re[bitrev<BITS>(j)] *= prow[0][mm];
            im[bitrev<BITS>(j)] *= prow[1][mm];
        }
    }
    if (M == 32 && p.wtab) {
        // cone beam: the bin and cylinder factors from the plan's (nb, nt)
        // table, one coalesced load per sample (rows of the two columns).
        // Not at M = 64: those registers are already full (the loads in
        // flight would spill).
        const int64_t ta = std::min(ca, p.ncols - 1) % p.nt, tb2 = std::min(cb, p.ncols - 1) % p.nt;
        const __amdgpu_buffer_rsrc_t rw =
            col_rsrc(p.wtab, 4u * (uint32_t)p.nb * (uint32_t)p.nt);
        const uint32_t wa0 = 4u * (uint32_t)(ta * p.nb), wb0 = 4u * (uint32_t)(tb2 * p.nb);
#pragma unroll
        for (int j = 0; j < JN; ++j) {
            const int m = l + 32 * j, r = bitrev<BITS>(j);
            const uint32_t off = 4u * (uint32_t)m;
            const float fa = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rw, wa0 + off, 0, 0));
            const float fb = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rw, wb0 + off, 0, 0));
            re[r] = m < p.nb ? re[r] * (colw[0] * fa) : 0.f;
            im[r] = m < p.nb ? im[r] * (colw[1] * fb) : 0.f;
        }
    } else {
#pragma unroll
    for (int j = 0; j < JN; ++j) {
        const int m = l + 32 * j, r = bitrev<BITS>(j);     // m < N: bin tables are 0 past nb
        const float a = bin_a[m];
        float wa = colw[0] * a, wb = colw[1] * a;
        if (p.cyl == 1) {
            wa *= cyl_k[0] * rsqrtf(bin_s2[m] + cyl_c[0]);
            wb *= cyl_k[1] * rsqrtf(bin_s2[m] + cyl_c[1]);
        } else if (p.cyl) {
            wa *= cyl_k[0];
            wb *= cyl_k[1];
        }
        // samples past nb belong to the next column: weight 0, value 0
        re[r] = m < p.nb ? re[r] * wa : 0.f;
        im[r] = m < p.nb ? im[r] * wb : 0.f;
    }
    }

    dft_dit<M, HALF>(re, im);                                   // A
    twiddle_transpose<M>(re, im, tw_re, tw_im, tb, l);          // B, C
#pragma unroll
    for (int h = 0; h < H; ++h) dft32_dit_br(re + 32 * h, im + 32 * h);   // D
    // X[M k1 + l + 32 h] is in register 32 h + bitrev5(k1).  Spectrum (He / N)
    // and swap give the inverse's step-A input: element n2 = H k1 + h of lane
    // l, in register bitrev(n2).
    float ar[M], ai[M];
#pragma unroll
    for (int h = 0; h < H; ++h) {
#pragma unroll
        for (int k1 = 0; k1 < 32; ++k1) {
            const float sp = spec[M * k1 + l + 32 * h];
            const int src = 32 * h + bitrev<5>(k1), dst = bitrev<BITS>(H * k1 + h);
            ar[dst] = im[src] * sp;
            ai[dst] = re[src] * sp;
        }
    }
    dft_dit<M, false>(ar, ai);
    twiddle_transpose<M>(ar, ai, tw_re, tw_im, tb, l);
#pragma unroll
    for (int h = 0; h < H; ++h) dft32_dit_br(ar + 32 * h, ai + 32 * h);
    // R[M j1 + l + 32 h] in register 32 h + bitrev5(j1)

    // Output: column a is Im R, column b is Re R; rows nb .. nrows_out - 1 are 0.
    constexpr uint32_t esz = OUT == OUT_F32 ? 4u : 2u;
    const __amdgpu_buffer_rsrc_t ro =
        col_rsrc(static_cast<char*>(p.out) + (int64_t)p.nrows_out * (nw ? cw : 0) * esz,
                 esz * (uint32_t)p.nrows_out * (uint32_t)nw);
    constexpr int J1N = HALF ? 16 : 32;
#pragma unroll
    for (int c2 = 0; c2 < 2; ++c2) {
        const int64_t c = c2 ? cb : ca;
        if (!(c2 ? vb : va)) continue;
        float inv = 1.f;
        if (OUT == OUT_F16S) {
            const int view = (int)(c / p.nt);
            const float sc = f16s_scale(p.vbound[view], p.g1);
            inv = 1.f / sc;                      // a power of two: exact
            if (c % p.nt == 0 && l == 0) p.vscale[view] = sc;
        }
        const uint32_t oc = esz * (uint32_t)((2 * half + c2) * p.nrows_out);
#pragma unroll
        for (int h = 0; h < H; ++h) {
#pragma unroll
            for (int j1 = 0; j1 < J1N; ++j1) {
                const int n = M * j1 + 32 * h + l;
                if (n >= p.nrows_out) continue;          // the next column's rows
                const int r = 32 * h + bitrev<5>(j1);
                const float v = n < p.nb ? (c2 ? ar[r] : ai[r]) * inv : 0.f;
                if (OUT == OUT_F32) {
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), ro, oc + esz * n,
                                                          0, 0);
                } else {
                    const uint16_t u = OUT == OUT_BF16 ? __builtin_bit_cast(uint16_t, bf16_rne(v))
                                                       : __half_as_ushort(__float2half_rn(v));
                    __builtin_amdgcn_raw_buffer_store_b16(u, ro, oc + esz * n, 0, 0);
                }
            }
        }
    }
}

template <int OUT, int M>
void launch_fft_m(const Params& p, hipStream_t st, dim3 grid)
{
    if (p.nrows_out <= 16 * M)
        hipLaunchKernelGGL((fft_filter_kernel<OUT, true, M>), grid, dim3(256), 0, st, p);
    else
        hipLaunchKernelGGL((fft_filter_kernel<OUT, false, M>), grid, dim3(256), 0, st, p);
}

template <int OUT>
void launch_fft(const Params& p, hipStream_t st)
{
    const int64_t blocks = (p.ncols + kFftCols - 1) / kFftCols;
    if (blocks >= (int64_t(1) << 31))
        throw Error(SGX_ERR_UNSUPPORTED, "sino_filter: too many columns for one launch");
    const dim3 grid((unsigned)blocks);
    if (p.fft_n == 1024) launch_fft_m<OUT, 32>(p, st, grid);
    else launch_fft_m<OUT, 64>(p, st, grid);
    SGX_HIP(hipGetLastError());
}

// Per-view max |weighted input|.  Block (view, y) covers kBoundRows detector
// rows of one view, a contiguous run of rows x nb samples that it reads as
// one flat range: 16-byte loads when nb is a multiple of 4 and the input is
// 16-byte aligned (the 4 samples of a load then share a row), else scalar
// loads; kBoundLoads loads in flight per thread.  The bin factors of the
// fused weight (wbin, Parker, arc cos(s/dsd), s^2 of the flat cylinder
// weight) and the row factors sit in LDS.  One atomic per block.
constexpr int kBoundRows = 16;
constexpr int kBoundLoads = 8;
template <int kBoundMaxBins>
__global__ __launch_bounds__(256) void bound_kernel(Params p, unsigned int* __restrict__ vbound,
                                                    int vec)
{
    __shared__ __attribute__((aligned(16))) float bin_a[kBoundMaxBins];
    __shared__ __attribute__((aligned(16))) float bin_s2[kBoundMaxBins];
    __shared__ float row_c[kBoundRows], row_t2[kBoundRows];
    const int tid = threadIdx.x;
    const int view = blockIdx.x;
    const int t0 = blockIdx.y * kBoundRows;
    const int nrows = min(kBoundRows, p.nt - t0);
    for (int m = tid; m < p.nb; m += 256) {
        float a = p.wbin ? p.wbin[m] : 1.f;
        if (p.cyl == 2) a *= cosf(p.s_axis[m] / p.cyl_dsd);
        if (p.ptab) a *= p.ptab[(int64_t)view * p.nb + m];
        bin_a[m] = fabsf(a);
        bin_s2[m] = p.cyl == 1 ? p.s_axis[m] * p.s_axis[m] : 0.f;
    }
    if (tid < nrows) {
        const int t = t0 + tid;
        float cw = p.scale * (p.wview ? p.wview[view] : 1.f), tt2 = 0.f;
        if (p.cyl == 1) {                  // dso / sqrt(s^2 + t^2 + dsd^2)
            tt2 = p.t_axis[t] * p.t_axis[t] + p.cyl_dsd * p.cyl_dsd;
            cw *= p.cyl_dso;
        } else if (p.cyl == 2) {           // (dso/dsd) cos(s/dsd) / sqrt((t/dsd)^2 + 1)
            const float r = p.t_axis[t] / p.cyl_dsd;
            cw *= (p.cyl_dso / p.cyl_dsd) / sqrtf(r * r + 1.f);
        }
        row_c[tid] = fabsf(cw);
        row_t2[tid] = tt2;
    }
    __syncthreads();

    // flat index i = r * nb + m; r = floor((i + 0.5) / nb) is exact in FP32
    // for i < kBoundRows * kBoundMaxBins
    const float inv_nb = 1.f / (float)p.nb;
    const float* base = p.in + ((int64_t)view * p.nt + t0) * p.nb;
    const int n = nrows * p.nb;
    auto weight = [&](int r, int m) -> float {
        float w = row_c[r] * bin_a[m];
        if (p.cyl == 1) w *= rsqrtf(bin_s2[m] + row_t2[r]);
        return w;
    };
    float mx = 0.f;
    if (vec) {
        const float4* b4 = reinterpret_cast<const float4*>(base);
        const int n4 = n >> 2;
        for (int i0 = tid; i0 < n4; i0 += 256 * kBoundLoads) {
            float4 v[kBoundLoads];
#pragma unroll
            for (int k = 0; k < kBoundLoads; ++k) {
                const int i4 = i0 + 256 * k;
                v[k] = i4 < n4 ? b4[i4] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int k = 0; k < kBoundLoads; ++k) {
                const int i = 4 * (i0 + 256 * k);
                if (i >= n) break;
                const int r = (int)(((float)i + 0.5f) * inv_nb);
                const int m = i - r * p.nb;          // a multiple of 4 (nb % 4 == 0)
                // the 4 bins' factors as one 16-byte LDS read (conflict-free)
                const float4 a4 = *reinterpret_cast<const float4*>(bin_a + m);
                const float rc = row_c[r];
                float w0 = rc * a4.x, w1 = rc * a4.y, w2 = rc * a4.z, w3 = rc * a4.w;
                if (p.cyl == 1) {
                    const float4 s4 = *reinterpret_cast<const float4*>(bin_s2 + m);
                    const float t2 = row_t2[r];
                    w0 *= rsqrtf(s4.x + t2);
                    w1 *= rsqrtf(s4.y + t2);
                    w2 *= rsqrtf(s4.z + t2);
                    w3 *= rsqrtf(s4.w + t2);
                }
                mx = fmaxf(mx, fabsf(v[k].x) * w0);
                mx = fmaxf(mx, fabsf(v[k].y) * w1);
                mx = fmaxf(mx, fabsf(v[k].z) * w2);
                mx = fmaxf(mx, fabsf(v[k].w) * w3);
            }
        }
    } else {
        for (int i0 = tid; i0 < n; i0 += 256 * kBoundLoads) {
            float v[kBoundLoads];
#pragma unroll
            for (int k = 0; k < kBoundLoads; ++k) {
                const int i = i0 + 256 * k;
                v[k] = i < n ? base[i] : 0.f;
            }
#pragma unroll
            for (int k = 0; k < kBoundLoads; ++k) {
                const int i = i0 + 256 * k;
                if (i >= n) break;
                const int r = (int)(((float)i + 0.5f) * inv_nb);
                mx = fmaxf(mx, fabsf(v[k]) * weight(r, i - r * p.nb));
            }
        }
    }
    for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    __shared__ float red[4];
    if ((tid & 63) == 0) red[tid >> 6] = mx;
    __syncthreads();
    if (tid == 0) {
        mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        atomicMax(vbound + view, __float_as_uint(mx));     // mx >= 0: bit order = value order
    }
}

// Streaming range pass of the FFT form (scaled FP16): per view, max |input|
// times the view's weight bound |scale| x wmax x |wview[view]| (an upper
// bound of the weighted input, since every factor of the fused weight is
// bounded by its part and Parker weights by 1).  A looser bound than the
// per-sample weighted maximum only moves the power-of-two scale up, which
// FP16 stores exactly (no value falls out of FP16's normal range above
// 2^-24 of the view's largest).  One load and half a VALU max per sample:
// HBM-bound.
constexpr int kAbsRows = 64;     // rows per block of the streaming range pass
__global__ __launch_bounds__(256) void absmax_kernel(Params p, unsigned int* __restrict__ vbound,
                                                     int vec)
{
    const int tid = threadIdx.x;
    const int view = blockIdx.x;
    const int t0 = blockIdx.y * kAbsRows;
    const int nrows = min(kAbsRows, p.nt - t0);
    const float* base = p.in + ((int64_t)view * p.nt + t0) * p.nb;
    const int n = nrows * p.nb;
    float mx = 0.f;
    if (vec) {
        const float4* b4 = reinterpret_cast<const float4*>(base);
        const int n4 = n >> 2;
        for (int i0 = tid; i0 < n4; i0 += 256 * kBoundLoads) {
            float4 v[kBoundLoads];
#pragma unroll
            for (int k = 0; k < kBoundLoads; ++k) {
                const int i4 = i0 + 256 * k;
                v[k] = i4 < n4 ? b4[i4] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int k = 0; k < kBoundLoads; ++k)
                mx = fmaxf(fmaxf(mx, fmaxf(fabsf(v[k].x), fabsf(v[k].y))),
                           fmaxf(fabsf(v[k].z), fabsf(v[k].w)));
        }
    } else {
        for (int i0 = tid; i0 < n; i0 += 256 * kBoundLoads) {
            float v[kBoundLoads];
#pragma unroll
            for (int k = 0; k < kBoundLoads; ++k) {
                const int i = i0 + 256 * k;
                v[k] = i < n ? base[i] : 0.f;
            }
#pragma unroll
            for (int k = 0; k < kBoundLoads; ++k) mx = fmaxf(mx, fabsf(v[k]));
        }
    }
    for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    __shared__ float red[4];
    if ((tid & 63) == 0) red[tid >> 6] = mx;
    __syncthreads();
    if (tid == 0) {
        mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        const float w = fabsf(p.scale) * p.wmax * (p.wview ? fabsf(p.wview[view]) : 1.f);
        atomicMax(vbound + view, __float_as_uint(mx * w));     // >= 0: bit order = value order
    }
}

__global__ void zero_u32_kernel(unsigned int* p, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void parker_kernel(const double* __restrict__ gam, const double* __restrict__ bet,
                              int nb, int na, double gm, float* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < (int64_t)nb * na) out[i] = (float)parker_w(gam[i % nb], bet[i / nb], gm);
}

__global__ void weights_kernel(Params p, float* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < (int64_t)p.nb * p.ncols) out[i] = weight(p, (int)(i % p.nb), i / p.nb);
}

Params make_params(const FilterArgs& a)
{
    Params p{};   // every field not set below is zero / null
    p.in = a.in;
    p.nb = a.nb;
    p.ncols = a.ncols;
    p.nt = a.nt;
    p.g = a.g;
    p.npad = a.npad;
    p.nrows_out = a.nb + a.extra;
    p.out = a.out;
    p.scale = a.scale;
    p.wbin = a.wbin;
    p.wview = a.wview;
    p.cyl = a.cyl;
    p.cyl_dsd = a.cyl_dsd;
    p.cyl_dso = a.cyl_dso;
    p.s_axis = a.s_axis;
    p.t_axis = a.t_axis;
    p.ptab = a.ptab;
    p.vbound = a.vbound;
    p.g1 = a.g1;
    p.wmax = a.wmax;
    p.gscale = a.gscale;
    p.vscale = a.vscale;
    p.row_tiles_per_block = 0;
    p.fft = a.fft;
    p.fft_n = a.fft_n;
    p.wtab = a.wtab;
    return p;
}

}  // namespace

std::vector<float> fft_filter_tables(const float* Hk, int64_t npad, int64_t nb,
                                     const float* wbin, const double* s, int cyl, double dsd)
{
    // SGX_FILTER_TOEPLITZ=1 keeps every filter on the Toeplitz form (A/B runs)
    const char* env = std::getenv("SGX_FILTER_TOEPLITZ");
    if ((npad != 512 && npad != 1024 && npad != 2048) || nb > npad || (env && env[0] == '1'))
        return {};
    // npad = 512 runs as a 1024-point transform: circular convolution modulo
    // 512 equals convolution modulo 1024 with g extended periodically, whose
    // spectrum is g's 512-point spectrum on the even bins and 0 on the odd
    // ones (rows below 512 of the result are the reference's).
    const int N = npad < 1024 ? 1024 : (int)npad, M = N / 32, r = N / (int)npad;
    std::vector<float> t(5 * (size_t)N, 0.f);
    for (int64_t m = 0; m < nb; ++m) {
        float a = wbin ? wbin[m] : 1.f;
        if (cyl == 2) a *= (float)std::cos(s[m] / dsd);
        t[3 * N + m] = a;
        t[4 * N + m] = cyl == 1 ? (float)(s[m] * s[m]) : 0.f;
    }
    const double w = 2.0 * M_PI / (double)N;
    for (int k2 = 0; k2 < M; ++k2)
        for (int n1 = 0; n1 < 32; ++n1) {
            const double ph = w * (double)(n1 * k2);      // W^(n1 k2) = e^(-i ph)
            t[k2 * 32 + n1] = (float)std::cos(ph);
            t[N + k2 * 32 + n1] = (float)-std::sin(ph);
        }
    const int P = (int)npad;
    for (int k = 0; k < N; k += r) {
        const int q = k / r;
        t[2 * N + k] = (float)(0.5 * ((double)Hk[q] + (double)Hk[(P - q) % P]) / (double)P);
    }
    return t;
}

void launch_filter_weights(const FilterArgs& a, float* out, hipStream_t st)
{
    const Params p = make_params(a);
    const int64_t n = (int64_t)a.nb * a.ncols;
    if (n <= 0) return;
    hipLaunchKernelGGL(weights_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p,
                       out);
    SGX_HIP(hipGetLastError());
}

void launch_sino_filter(const FilterArgs& a, hipStream_t st)
{
    if (a.ncols <= 0) return;
    if (a.nb > kNbWide)
        throw Error(SGX_ERR_UNSUPPORTED, "sino_filter: nb > 8192 is not supported");
    const Params p = make_params(a);
    if (a.out_f16s && (a.out_bf16 || a.split_bf16 || !a.vbound || !a.vscale || a.nt < 1))
        throw Error(SGX_ERR_ARG, "sino_filter: scaled FP16 output takes FP32 or FP16 products");
    if (a.split_f16 && (a.split_bf16 || !a.vbound || !(a.gscale > 0.f)))
        throw Error(SGX_ERR_ARG, "sino_filter: split-FP16 products need the input bound and g scale");
    const int out = a.out_f16s ? OUT_F16S : a.out_bf16 ? OUT_BF16 : OUT_F32;
    if (a.fft) {
        // FP32 transforms; the split-product flags do not apply
        if ((a.fft_n != 1024 && a.fft_n != 2048) || p.nrows_out > a.npad || a.npad > a.fft_n)
            throw Error(SGX_ERR_ARG, "sino_filter: bad FFT-form table");
        switch (out) {
        case OUT_F32: launch_fft<OUT_F32>(p, st); break;
        case OUT_BF16: launch_fft<OUT_BF16>(p, st); break;
        default: launch_fft<OUT_F16S>(p, st); break;
        }
        return;
    }
    if (a.split_bf16) {
        switch (out) {
        case OUT_F32: launch_mode<M_BF16X3, OUT_F32>(p, st); break;
        default: launch_mode<M_BF16X3, OUT_BF16>(p, st); break;
        }
    } else if (a.split_f16) {
        switch (out) {
        case OUT_F32: launch_mode<M_F16X3, OUT_F32>(p, st); break;
        default: launch_mode<M_F16X3, OUT_F16S>(p, st); break;
        }
    } else {
        switch (out) {
        case OUT_F32: launch_mode<M_F32, OUT_F32>(p, st); break;
        case OUT_BF16: launch_mode<M_F32, OUT_BF16>(p, st); break;
        default: launch_mode<M_F32, OUT_F16S>(p, st); break;
        }
    }
}

void launch_filter_bound(const FilterArgs& a, unsigned int* vbound, hipStream_t st)
{
    if (a.ncols <= 0) return;
    if (a.nb > kNbWide)
        throw Error(SGX_ERR_UNSUPPORTED, "sino_filter: nb > 8192 is not supported");
    const int64_t na = a.ncols / a.nt;
    launch_zero_u32(vbound, na, st);
    const Params p = make_params(a);
    if ((a.nt + kBoundRows - 1) / kBoundRows > 65535)
        throw Error(SGX_ERR_UNSUPPORTED, "sino_filter: too many rows per view");
    const int vec = a.nb % 4 == 0 && (reinterpret_cast<uintptr_t>(p.in) & 15u) == 0;
    const dim3 grid((unsigned)na, (unsigned)((a.nt + kBoundRows - 1) / kBoundRows));
    if (a.fft && a.wmax > 0.f)
        hipLaunchKernelGGL(absmax_kernel,
                           dim3((unsigned)na, (unsigned)((a.nt + kAbsRows - 1) / kAbsRows)),
                           dim3(256), 0, st, p, vbound, vec);
    else if (a.nb > kNbSmall)
        hipLaunchKernelGGL(bound_kernel<kNbWide>, grid, dim3(256), 0, st, p, vbound, vec);
    else
        hipLaunchKernelGGL(bound_kernel<kNbSmall>, grid, dim3(256), 0, st, p, vbound, vec);
    SGX_HIP(hipGetLastError());
}

void launch_zero_u32(unsigned int* p, int64_t n, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p,
                       n);
    SGX_HIP(hipGetLastError());
}

void launch_parker(const double* gam, const double* bet, int nb, int na, double gamma_max,
                   float* out, hipStream_t st)
{
    const int64_t n = (int64_t)nb * na;
    if (n <= 0) return;
    hipLaunchKernelGGL(parker_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, gam,
                       bet, nb, na, gamma_max, out);
    SGX_HIP(hipGetLastError());
}

}  // namespace sgx
