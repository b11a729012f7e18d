// K8: the learned sinogram-domain correction -- a 1-D convolutional network
// applied along detector rows of the filtered sinogram -- with its training
// kernels, for gfx950.  No counterpart in the reference.
//
// Network (L layers, C channels, width K = 9, zero padding):
//   z_1 = conv(q) + b_1,  z_l = conv(gelu(z_{l-1})) + b_l  (l = 2..L),
//   q' = q + z_L           (one output channel: a residual correction)
// gelu(x) = x Phi(x) with the exact normal CDF Phi (erf).  Weights are kept
// in FP32 (the optimiser's master copy) and every kernel rounds them to BF16
// (round-to-nearest-even) as it stages them; all products accumulate in FP32.
//
// Layout: activations are (R, C, N) with N (detector bins) fastest; R rows
// are the views of all sinograms in a batch.
//
// conv_kernel       forward layer, or the data gradient of a layer (the same
//                   contraction with the channels swapped and the taps
//                   reversed), GELU applied to the input on load and GELU'
//                   applied to the output in the backward epilogue.
// conv_mfma_kernel  the same for 32 -> 32 channel layers on MFMA
//                   (v_mfma_f32_32x32x16_bf16): per 32 bins a 32 x 288 x 32
//                   product, K = (tap, input channel).  The weights are BF16
//                   (as everywhere); each FP32 activation is split exactly into
//                   three BF16 parts, so that the three products per K step
//                   are exact and only the FP32 accumulation rounds.
// wgrad_kernel      weight and bias gradients: per-block partial sums over a
//                   strided set of (row, bin tile) items, then a fixed-order
//                   reduction (deterministic).
// adam_kernel       Adam update of the FP32 master weights.
// mse_kernel        residual, loss partials and the loss gradient 2 (x - t) / n.

#include "sgx_internal.h"
#include "sgx.h"

namespace sgx {
namespace {

constexpr int kW = 9;                  // taps
constexpr int kHalf = kW / 2;
constexpr int kMaxC = 32;              // channels
constexpr int kTileN = 128;            // bins per block
constexpr int kThreads = 256;

__device__ __forceinline__ float bf16r(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

__device__ __forceinline__ float gelu(float x)
{
    return 0.5f * x * (1.f + erff(x * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad(float x)
{
    // Phi(x) + x phi(x)
    const float cdf = 0.5f * (1.f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * __expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// out[r, o, n] = bias[o] + sum_{i,k} Wt[o, i, k] * act(in[r, i, n + k - 4])
// forward  (TRANS = false): Wt[o, i, k] = W[o, i, k], act = gelu if GELU_IN;
// backward (TRANS = true):  the data gradient d/d in of a layer with weights
//                           W (Cout = ci, Cin = co): Wt[o, i, k] = W[i, o, 8 - k],
//                           no bias, and the result times gelu'(zprev[r, o, n])
//                           when zprev is given.
// resid (forward, optional): added to the output (one channel).
// Block: 128 bins of one row; thread = 2 bins x ceil(Cout / 4) channels.
template <bool TRANS, bool GELU_IN>
__global__ __launch_bounds__(kThreads) void conv_kernel(
    const float* __restrict__ in, int cin, int cout, int n, const float* __restrict__ w,
    const float* __restrict__ bias, const float* __restrict__ resid,
    const float* __restrict__ zprev, float* __restrict__ out)
{
    __shared__ float tile[kMaxC][kTileN + kW - 1];
    __shared__ float wt[kMaxC * kMaxC * kW];         // [o][i][k], BF16-rounded
    const int r = blockIdx.y;
    const int n0 = blockIdx.x * kTileN;
    const int tid = threadIdx.x;
    // weights
    for (int e = tid; e < cout * cin * kW; e += kThreads) {
        const int o = e / (cin * kW), rem = e % (cin * kW), i = rem / kW, k = rem % kW;
        const float v = TRANS ? w[(i * cout + o) * kW + (kW - 1 - k)] : w[e];
        wt[e] = bf16r(v);
    }
    // input tile with halo, zero outside [0, n); a thread's loads in flight
    // together
    const float* src = in + (int64_t)r * cin * n;
    constexpr int kSpan = kTileN + kW - 1;
    constexpr int kLoads = (kMaxC * kSpan + kThreads - 1) / kThreads;
    const int total = cin * kSpan;
    float xv[kLoads];
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
        const int e = tid + u * kThreads;
        const int i = e / kSpan, j = e % kSpan;
        const int m = n0 + j - kHalf;
        xv[u] = (e < total && m >= 0 && m < n) ? src[(int64_t)i * n + m] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
        const int e = tid + u * kThreads;
        if (e >= total) break;
        tile[e / kSpan][e % kSpan] = GELU_IN ? gelu(xv[u]) : xv[u];
    }
    __syncthreads();

    const int groups = (cout + 7) / 8;               // 8 output channels per thread
    const int pair = tid % (kTileN / 2), grp = tid / (kTileN / 2);
    if (grp >= groups) return;
    const int o0 = grp * 8, no = min(8, cout - o0);
    const int j0 = 2 * pair;                          // tile column of the first bin
    float acc0[8], acc1[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc0[c] = acc1[c] = 0.f;
    for (int i = 0; i < cin; ++i) {
        float x[kW + 1];
#pragma unroll
        for (int k = 0; k < kW + 1; ++k) x[k] = tile[i][j0 + k];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (c >= no) break;
            const float* wr = wt + ((o0 + c) * cin + i) * kW;
#pragma unroll
            for (int k = 0; k < kW; ++k) {
                acc0[c] = fmaf(wr[k], x[k], acc0[c]);
                acc1[c] = fmaf(wr[k], x[k + 1], acc1[c]);
            }
        }
    }
    float* dst = out + (int64_t)r * cout * n;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if (c >= no) break;
        const int o = o0 + c;
        const float b = (!TRANS && bias) ? bias[o] : 0.f;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int m = n0 + j0 + h;
            if (m >= n) continue;
            float v = (h ? acc1[c] : acc0[c]) + b;
            if (TRANS && zprev) v *= gelu_grad(zprev[((int64_t)r * cout + o) * n + m]);
            if (!TRANS && resid) v += resid[(int64_t)r * n + m];
            dst[(int64_t)o * n + m] = v;
        }
    }
}

// ---- MFMA form of the 32 -> 32 channel layers --------------------------------

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kMC = 32;                    // channels of the MFMA form
constexpr int kMK = kW * kMC;              // 288 = K of one output bin
constexpr int kMTile = 128;                // bins per block: 4 waves x 32
constexpr int kMPos = kMTile + kW - 1;     // input positions with the halo
constexpr int kMStride = 40;               // halves per position: 32 channels + 8 pad
                                           // (16-byte reads of 8 lanes hit distinct banks)

__device__ __forceinline__ uint16_t bf16_bits(float x)
{
    return (uint16_t)(__float_as_uint(bf16r(x)) >> 16);
}

// x = h + m + l exactly, each a BF16 (FP32 has 24 significant bits)
__device__ __forceinline__ void split3(float x, uint16_t& h, uint16_t& m, uint16_t& l)
{
    const float fh = bf16r(x);
    const float r1 = x - fh;
    const float fm = bf16r(r1);
    h = (uint16_t)(__float_as_uint(fh) >> 16);
    m = (uint16_t)(__float_as_uint(fm) >> 16);
    l = bf16_bits(r1 - fm);
}

// A operand of conv_mfma_kernel: wa[o][kk] = BF16 of Wt[o][i][k], kk = 32 k + i;
// forward Wt = W (cout x cin x 9), TRANS Wt[o][i][k] = W[i][o][8 - k].
template <bool TRANS>
__global__ void wprep_kernel(const float* __restrict__ w, uint16_t* __restrict__ wa)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kMC * kMK) return;
    const int o = e / kMK, kk = e % kMK, k = kk / kMC, i = kk % kMC;
    const float v = TRANS ? w[(i * kMC + o) * kW + (kW - 1 - k)] : w[(o * kMC + i) * kW + k];
    wa[e] = bf16_bits(v);
}

// out[r, o, n] as conv_kernel, cin = cout = 32.  Block: 128 bins of one row,
// wave w the 32 bins n0 + 32 w ..; lane (col, half) = (lane & 31, lane >> 5)
// holds A rows o = col, B columns n = col and K elements 8 half .. 8 half + 7
// of each 16-step; acc[v] is output channel 8 (v / 4) + 4 half + v % 4.
template <bool TRANS, bool GELU_IN>
__global__ __launch_bounds__(kThreads) void conv_mfma_kernel(
    const float* __restrict__ in, int n, const uint16_t* __restrict__ wa,
    const float* __restrict__ bias, const float* __restrict__ zprev, float* __restrict__ out)
{
    __shared__ __attribute__((aligned(16))) uint16_t xs[3][kMPos * kMStride];
    const int r = blockIdx.y;
    const int n0 = blockIdx.x * kMTile;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, half = lane >> 5;

    // input tile with halo (zero outside [0, n)), split, position-major; all
    // of a thread's loads in flight before the first is used
    const float* src = in + (int64_t)r * kMC * n;
    constexpr int kLoads = (kMC * kMPos + kThreads - 1) / kThreads;
    float xv[kLoads];
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
        const int e = tid + u * kThreads;
        const int i = e / kMPos, p = e % kMPos;
        const int m = n0 + p - kHalf;
        xv[u] = (e < kMC * kMPos && m >= 0 && m < n) ? src[(int64_t)i * n + m] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
        const int e = tid + u * kThreads;
        if (e >= kMC * kMPos) break;
        const int i = e / kMPos, p = e % kMPos;
        const float v = GELU_IN ? gelu(xv[u]) : xv[u];     // gelu(0) = 0 in the halo
        uint16_t h, md, l;
        split3(v, h, md, l);
        xs[0][p * kMStride + i] = h;
        xs[1][p * kMStride + i] = md;
        xs[2][p * kMStride + i] = l;
    }
    // this lane's A fragments (row o = col), all 18 K steps
    bf16x8 a[kMK / 16];
    const uint4* wrow = reinterpret_cast<const uint4*>(wa + col * kMK + 8 * half);
#pragma unroll
    for (int ks = 0; ks < kMK / 16; ++ks) a[ks] = __builtin_bit_cast(bf16x8, wrow[2 * ks]);
    __syncthreads();

    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
#pragma unroll
    for (int ks = 0; ks < kMK / 16; ++ks) {
        const int k = ks / 2, i0 = (ks & 1) * 16 + 8 * half;
        const int pos = 32 * wave + col + k;
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
            const bf16x8 b = __builtin_bit_cast(
                bf16x8, *reinterpret_cast<const uint4*>(&xs[pl][pos * kMStride + i0]));
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ks], b, acc, 0, 0, 0);
        }
    }

    const int m = n0 + 32 * wave + col;
    if (m >= n) return;
    float* dst = out + (int64_t)r * kMC * n + m;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int o = 8 * (v / 4) + 4 * half + (v % 4);
        float y = acc[v];
        if (!TRANS && bias) y += bias[o];
        if (TRANS && zprev) y *= gelu_grad(zprev[((int64_t)r * kMC + o) * n + m]);
        dst[(int64_t)o * n] = y;
    }
}

// ---- fused inference: all layers of a 32-channel network in one kernel ------
//
// Block = one row and T = 128 - 8 (L - 2) output bins.  Layer 1 (1 -> 32)
// runs on the VALU over 136 positions, the 32 -> 32 layers on MFMA over 128
// positions each (4 waves x 32), the last layer (32 -> 1) on the VALU;
// activations stay in LDS (exact three-way BF16 splits, FP32 before the last
// layer), so HBM sees only q in and q + f(q) out.  Each layer's output
// window starts 4 bins after its input's; positions a layer did not compute
// (past its 128) read as 0 and only reach outputs outside the final T.  As
// in the per-layer kernels, every activation outside [0, n) is 0 (zero
// padding), weights are BF16 and sums FP32.
constexpr int kNetMaxL = 16;
struct NetArgs {
    const float* q;
    float* out;
    int n;
    int layers;
    const float* w;                       // all parameters
    int woff[kNetMaxL], boff[kNetMaxL];
    uint16_t* wa;                         // A operands of layers 1 .. L-2
};

constexpr int kAStride = kMK + 8;         // halves per A row in LDS (conflict-free reads)

// the A operands (wprep_kernel<false> layout) of layers 1 .. L-2: block
// (x, l - 1) converts 256 weights of layer l
__global__ void wprep_net_kernel(NetArgs a)
{
    const int l = blockIdx.y + 1;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kMC * kMK) return;
    const int o = e / kMK, kk = e % kMK, k = kk / kMC, i = kk % kMC;
    a.wa[(int64_t)(l - 1) * kMC * kMK + e] = bf16_bits(a.w[a.woff[l] + (o * kMC + i) * kW + k]);
}

__global__ __launch_bounds__(kThreads) void net_infer_kernel(NetArgs a)
{
    __shared__ __attribute__((aligned(16))) uint16_t xs[3][kMPos * kMStride];
    // A operand of the current MFMA layer; after the last one, the FP32
    // gelu'd activations g[32][kMPos] of the last layer's input
    __shared__ __attribute__((aligned(16))) uint16_t abuf[kMC * kAStride];
    __shared__ float qt[kMPos + kW - 1];
    __shared__ float w1[kMC * kW], b1[kMC], wl[kMC * kW];
    __shared__ float red[kMTile];
    static_assert(sizeof(float) * kMC * kMPos <= sizeof(uint16_t) * kMC * kAStride, "g");
    float* g = reinterpret_cast<float*>(abuf);

    const int L = a.layers, n = a.n;
    const int T = kMTile - 8 * (L - 2);
    const int r = blockIdx.y, n0 = blockIdx.x * T;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, half = lane >> 5;
    int s = -kHalf * (L - 1);                 // layer 1's first output position - n0
    const float* qr = a.q + (int64_t)r * n;

    for (int j = tid; j < kMPos + kW - 1; j += kThreads) {
        const int m = n0 + s - kHalf + j;
        qt[j] = (m >= 0 && m < n) ? qr[m] : 0.f;
    }
    for (int e = tid; e < kMC * kW; e += kThreads) {
        w1[e] = bf16r(a.w[a.woff[0] + e]);            // W1[o][0][k]
        wl[e] = bf16r(a.w[a.woff[L - 1] + e]);        // WL[0][i][k]
    }
    if (tid < kMC) b1[tid] = a.w[a.boff[0] + tid];
    __syncthreads();

    // layer 1 (1 -> 32): z[o][p] at n0 + s + p, p < kMPos
    for (int e = tid; e < kMC * kMPos; e += kThreads) {
        const int p = e / kMC, o = e % kMC;
        const int m = n0 + s + p;
        float z = 0.f;
        if (m >= 0 && m < n) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < kW; ++k) acc = fmaf(w1[o * kW + k], qt[p + k], acc);
            z = acc + b1[o];
        }
        const float v = gelu(z);
        if (L == 2) {
            g[o * kMPos + p] = v;
        } else {
            uint16_t h, md, lo;
            split3(v, h, md, lo);
            xs[0][p * kMStride + o] = h;
            xs[1][p * kMStride + o] = md;
            xs[2][p * kMStride + o] = lo;
        }
    }

    // layers 1 .. L-2 (32 -> 32) on MFMA
    for (int l = 1; l <= L - 2; ++l) {
        const uint16_t* wsrc = a.wa + (int64_t)(l - 1) * kMC * kMK;
        for (int e = 8 * tid; e < kMC * kMK; e += 8 * kThreads) {
            const int o = e / kMK, kk = e % kMK;
            *reinterpret_cast<uint4*>(&abuf[o * kAStride + kk]) =
                *reinterpret_cast<const uint4*>(wsrc + e);
        }
        __syncthreads();                      // xs (previous layer) and A staged
        bf16x8 af[kMK / 16];
#pragma unroll
        for (int ks = 0; ks < kMK / 16; ++ks)
            af[ks] = __builtin_bit_cast(
                bf16x8, *reinterpret_cast<const uint4*>(&abuf[col * kAStride + 16 * ks + 8 * half]));
        floatx16 acc;
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[v] = 0.f;
#pragma unroll
        for (int ks = 0; ks < kMK / 16; ++ks) {
            const int k = ks / 2, i0 = (ks & 1) * 16 + 8 * half;
            const int pos = 32 * wave + col + k;
#pragma unroll
            for (int pl = 0; pl < 3; ++pl) {
                const bf16x8 b = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4*>(&xs[pl][pos * kMStride + i0]));
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[ks], b, acc, 0, 0, 0);
            }
        }
        __syncthreads();                      // every read of xs and abuf is done
        s += kHalf;                           // this layer's first output position - n0
        const int idx = 32 * wave + col;
        const int m = n0 + s + idx;
        const bool in = m >= 0 && m < n;
        const bool last_mfma = l == L - 2;
        const float* bias = a.w + a.boff[l];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int o = 8 * (v / 4) + 4 * half + (v % 4);
            const float y = gelu(in ? acc[v] + bias[o] : 0.f);
            if (last_mfma) {
                g[o * kMPos + idx] = y;
            } else {
                uint16_t h, md, lo;
                split3(y, h, md, lo);
                xs[0][idx * kMStride + o] = h;
                xs[1][idx * kMStride + o] = md;
                xs[2][idx * kMStride + o] = lo;
            }
        }
        // positions kMTile .. kMPos - 1 were not computed: 0
        for (int e = tid; e < kMC * (kMPos - kMTile); e += kThreads) {
            const int o = e / (kMPos - kMTile), p = kMTile + e % (kMPos - kMTile);
            if (last_mfma) {
                g[o * kMPos + p] = 0.f;
            } else {
                xs[0][p * kMStride + o] = 0;
                xs[1][p * kMStride + o] = 0;
                xs[2][p * kMStride + o] = 0;
            }
        }
    }
    __syncthreads();

    // last layer (32 -> 1) + bias + q: thread (bin m, channel half hh)
    const int m = tid % kMTile, hh = tid / kMTile;
    float acc = 0.f;
    if (m < T) {
        for (int i = 16 * hh; i < 16 * hh + 16; ++i)
#pragma unroll
            for (int k = 0; k < kW; ++k) acc = fmaf(wl[i * kW + k], g[i * kMPos + m + k], acc);
    }
    if (hh == 1) red[m] = acc;
    __syncthreads();
    if (hh == 0 && m < T && n0 + m < n) {
        float y = (acc + red[m]) + a.w[a.boff[L - 1]];
        y += qt[m + kHalf * L];
        a.out[(int64_t)r * n + n0 + m] = y;
    }
}

// Weight / bias gradient partials of one layer:
//   dW[o, i, k] = sum_{r, m} d[r, o, m] * act(x[r, i, m + k - 4]),  db[o] = sum d[r, o, m]
// Block b takes items (row, tile) b, b + gridDim.x, ...; thread t owns
// output channel o = t / 8 and input channels i = (t % 8) + 8 j (j < 4), all
// taps, and writes its partials to part[b][...] (cout*cin*9 + cout floats).
template <bool GELU_IN>
__global__ __launch_bounds__(kThreads) void wgrad_kernel(
    const float* __restrict__ d, const float* __restrict__ x, int rows, int cin, int cout,
    int n, float* __restrict__ part)
{
    __shared__ float dt[kMaxC][kTileN];
    __shared__ float xt[kMaxC][kTileN + kW - 1];
    const int tid = threadIdx.x;
    const int o = tid / 8, ig = tid % 8;
    const bool own = o < cout;
    float acc[4][kW], accb = 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < kW; ++k) acc[j][k] = 0.f;
    const int tiles = (n + kTileN - 1) / kTileN;
    const int64_t items = (int64_t)rows * tiles;
    for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const int r = (int)(it / tiles), n0 = (int)(it % tiles) * kTileN;
        __syncthreads();
        for (int e = tid; e < cout * kTileN; e += kThreads) {
            const int c = e / kTileN, j = e % kTileN, m = n0 + j;
            dt[c][j] = m < n ? d[((int64_t)r * cout + c) * n + m] : 0.f;
        }
        for (int e = tid; e < cin * (kTileN + kW - 1); e += kThreads) {
            const int i = e / (kTileN + kW - 1), j = e % (kTileN + kW - 1);
            const int m = n0 + j - kHalf;
            float v = 0.f;
            if (m >= 0 && m < n) {
                v = x[((int64_t)r * cin + i) * n + m];
                if (GELU_IN) v = gelu(v);
            }
            xt[i][j] = v;
        }
        __syncthreads();
        if (!own) continue;
        for (int j = 0; j < kTileN; ++j) {
            const float dv = dt[o][j];
            if (ig == 0) accb += dv;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int i = ig + 8 * jj;
                if (i >= cin) break;
#pragma unroll
                for (int k = 0; k < kW; ++k) acc[jj][k] = fmaf(dv, xt[i][j + k], acc[jj][k]);
            }
        }
    }
    if (!own) return;
    float* p = part + (int64_t)blockIdx.x * (cout * cin * kW + cout);
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        const int i = ig + 8 * jj;
        if (i >= cin) break;
#pragma unroll
        for (int k = 0; k < kW; ++k) p[(o * cin + i) * kW + k] = acc[jj][k];
    }
    if (ig == 0) p[cout * cin * kW + o] = accb;
}

// Weight / bias gradient partials of a 32 -> 32 layer on MFMA:
//   dW[o][i][k] = sum_{r,m} d[r][o][m] gelu(x[r][i][m + k - 4]),  db[o] = sum d[r][o][m]
// as per-item 32 x 32 x 128 products, K = the item's 128 bins: A = d (o x m),
// one N tile per tap k (B[m][i] = gelu(x)[i][m + k - 4]) and one of ones (the
// bias).  Both operands are split exactly into three BF16 parts; the six
// products down to 2^-16 relative are summed, so the result matches FP32
// products to within FP32 accumulation.  The tap shift makes the B rows start
// at any bin: two copies of the activations, the second shifted by one bin,
// keep every 8-element read on a 4-byte boundary.  Taps go to waves k % 4
// (wave 1 also takes the bias tile).  Block b takes items b, b + gridDim.x,
// ... and writes its partial sums as wgrad_kernel does.
constexpr int kGS = 136;          // halves per LDS row of the d and x tiles (68 dwords:
                                  // 16-byte reads of 8 lanes hit distinct banks)

__global__ __launch_bounds__(kThreads) void wgrad_mfma_kernel(
    const float* __restrict__ d, const float* __restrict__ x, int rows, int n,
    float* __restrict__ part)
{
    __shared__ __attribute__((aligned(16))) uint16_t ds[3][kMC * kGS];     // [plane][o][m]
    __shared__ __attribute__((aligned(16))) uint16_t xs2[2][3][kMC * kGS]; // [copy][plane][i][p]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, half = lane >> 5;
    const int tiles = (n + kMTile - 1) / kMTile;
    const int64_t items = (int64_t)rows * tiles;
    const int ntap = wave == 0 ? 3 : 2;             // taps wave, wave + 4 (, 8)
    floatx16 acc[3], accb;
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[t][v] = 0.f;
#pragma unroll
    for (int v = 0; v < 16; ++v) accb[v] = 0.f;
    bf16x8 ones;
#pragma unroll
    for (int j = 0; j < 8; ++j) ones[j] = (__bf16)1.0f;

    constexpr int kLd = kMC * kMTile / kThreads;     // 16 d loads per thread
    constexpr int kLx = (kMC * kGS + kThreads - 1) / kThreads;   // 17 x loads
    for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const int r = (int)(it / tiles), n0 = (int)(it % tiles) * kMTile;
        float dv[kLd], xv[kLx];
#pragma unroll
        for (int u = 0; u < kLd; ++u) {
            const int e = tid + u * kThreads, o = e / kMTile, m = n0 + e % kMTile;
            dv[u] = m < n ? d[((int64_t)r * kMC + o) * n + m] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < kLx; ++u) {
            const int e = tid + u * kThreads, i = e / kGS, m = n0 - kHalf + e % kGS;
            xv[u] = (e < kMC * kGS && m >= 0 && m < n) ? x[((int64_t)r * kMC + i) * n + m] : 0.f;
        }
        __syncthreads();                             // the previous item's reads are done
#pragma unroll
        for (int u = 0; u < kLd; ++u) {
            const int e = tid + u * kThreads, o = e / kMTile, m = e % kMTile;
            uint16_t h, md, l;
            split3(dv[u], h, md, l);
            ds[0][o * kGS + m] = h;
            ds[1][o * kGS + m] = md;
            ds[2][o * kGS + m] = l;
        }
#pragma unroll
        for (int u = 0; u < kLx; ++u) {
            const int e = tid + u * kThreads;
            if (e >= kMC * kGS) break;
            const int i = e / kGS, p = e % kGS;
            uint16_t h, md, l;
            split3(gelu(xv[u]), h, md, l);           // gelu(0) = 0 outside [0, n)
            xs2[0][0][i * kGS + p] = h;
            xs2[0][1][i * kGS + p] = md;
            xs2[0][2][i * kGS + p] = l;
            if (p > 0) {                             // copy 1: position p - 1
                xs2[1][0][i * kGS + p - 1] = h;
                xs2[1][1][i * kGS + p - 1] = md;
                xs2[1][2][i * kGS + p - 1] = l;
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kMTile / 16; ++ks) {
            const int m0 = 16 * ks + 8 * half;
            bf16x8 a[3];
#pragma unroll
            for (int pl = 0; pl < 3; ++pl)
                a[pl] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4*>(&ds[pl][col * kGS + m0]));
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                if (t >= ntap) break;
                const int k = wave + 4 * t;
                const int s0 = m0 + k;               // first position of the 8
                const int cp = s0 & 1;
                const uint32_t* src[3];
                bf16x8 b[3];
#pragma unroll
                for (int pl = 0; pl < 3; ++pl) {
                    src[pl] = reinterpret_cast<const uint32_t*>(&xs2[cp][pl][col * kGS + s0 - cp]);
                    const uint4 w4 = make_uint4(src[pl][0], src[pl][1], src[pl][2], src[pl][3]);
                    b[pl] = __builtin_bit_cast(bf16x8, w4);
                }
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[0], acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[1], acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], b[0], acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[2], acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[2], b[0], acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], b[1], acc[t], 0, 0, 0);
            }
            if (wave == 1) {
#pragma unroll
                for (int pl = 0; pl < 3; ++pl)
                    accb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[pl], ones, accb, 0, 0, 0);
            }
        }
    }
    // partials: acc[t][v] = dW[o][i = col][k = wave + 4 t], o = 8 (v / 4) + 4 half + v % 4
    float* p = part + (int64_t)blockIdx.x * (kMC * kMC * kW + kMC);
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        if (t >= ntap) break;
        const int k = wave + 4 * t;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int o = 8 * (v / 4) + 4 * half + (v % 4);
            p[(o * kMC + col) * kW + k] = acc[t][v];
        }
    }
    if (wave == 1 && col == 0) {
#pragma unroll
        for (int v = 0; v < 16; ++v) p[kMC * kMC * kW + 8 * (v / 4) + 4 * half + (v % 4)] = accb[v];
    }
}

// Weight / bias gradient partials of a layer with few (o, i) pairs (the
// 1 -> C and C -> 1 layers, P = cout * cin <= 128): thread t owns pair t % P
// and every (256 / P)-th bin of each item (phase t / P), all taps; the
// phases are added in order at the end of the block.
constexpr int kSmallPairs = 128;

template <bool GELU_IN>
__global__ __launch_bounds__(kThreads) void wgrad_small_kernel(
    const float* __restrict__ d, const float* __restrict__ x, int rows, int cin, int cout,
    int n, float* __restrict__ part)
{
    __shared__ float dt[kMaxC][kTileN];
    __shared__ float xt[kMaxC][kTileN + kW - 1];
    __shared__ float red[kThreads][kW + 1];
    const int tid = threadIdx.x;
    const int P = cout * cin, phases = kThreads / P;
    const int pair = tid % P, phase = tid / P;
    const bool own = phase < phases;
    const int o = pair / cin, i = pair % cin;
    float acc[kW], accb = 0.f;
#pragma unroll
    for (int k = 0; k < kW; ++k) acc[k] = 0.f;
    const int tiles = (n + kTileN - 1) / kTileN;
    const int64_t items = (int64_t)rows * tiles;
    constexpr int kSpan = kTileN + kW - 1;
    constexpr int kLd = (kMaxC * kTileN + kThreads - 1) / kThreads;
    constexpr int kLx = (kMaxC * kSpan + kThreads - 1) / kThreads;
    for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const int r = (int)(it / tiles), n0 = (int)(it % tiles) * kTileN;
        float dv[kLd], xv[kLx];
#pragma unroll
        for (int u = 0; u < kLd; ++u) {
            const int e = tid + u * kThreads, c = e / kTileN, m = n0 + e % kTileN;
            dv[u] = (c < cout && m < n) ? d[((int64_t)r * cout + c) * n + m] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < kLx; ++u) {
            const int e = tid + u * kThreads, c = e / kSpan, m = n0 - kHalf + e % kSpan;
            xv[u] = (c < cin && m >= 0 && m < n) ? x[((int64_t)r * cin + c) * n + m] : 0.f;
        }
        __syncthreads();                          // the previous item is consumed
#pragma unroll
        for (int u = 0; u < kLd; ++u) {
            const int e = tid + u * kThreads, c = e / kTileN;
            if (c < cout) dt[c][e % kTileN] = dv[u];
        }
#pragma unroll
        for (int u = 0; u < kLx; ++u) {
            const int e = tid + u * kThreads, c = e / kSpan;
            if (c < cin) xt[c][e % kSpan] = GELU_IN ? gelu(xv[u]) : xv[u];
        }
        __syncthreads();
        if (own) {
            for (int m = phase; m < kTileN; m += phases) {
                const float dvm = dt[o][m];
                if (i == 0) accb += dvm;
#pragma unroll
                for (int k = 0; k < kW; ++k) acc[k] = fmaf(dvm, xt[i][m + k], acc[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kW; ++k) red[tid][k] = acc[k];
    red[tid][kW] = accb;
    __syncthreads();
    if (tid >= P) return;
    float sw[kW], sb = 0.f;
#pragma unroll
    for (int k = 0; k < kW; ++k) sw[k] = 0.f;
    for (int ph = 0; ph < phases; ++ph) {
#pragma unroll
        for (int k = 0; k < kW; ++k) sw[k] += red[ph * P + tid][k];
        sb += red[ph * P + tid][kW];
    }
    float* pp = part + (int64_t)blockIdx.x * (cout * cin * kW + cout);
#pragma unroll
    for (int k = 0; k < kW; ++k) pp[(o * cin + i) * kW + k] = sw[k];
    if (i == 0) pp[cout * cin * kW + o] = sb;
}

// Sum the partials of `blocks` blocks in block order, add into grad (per
// element: the weights then the biases of one layer).
__global__ void reduce_kernel(const float* __restrict__ part, int blocks, int len,
                              float* __restrict__ gw, float* __restrict__ gb, int nw)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len) return;
    float s = 0.f;
    for (int b0 = 0; b0 < blocks; b0 += 8) {
        float v[8];                               // loads in flight together
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = b0 + u < blocks ? part[(int64_t)(b0 + u) * len + e] : 0.f;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (b0 + u < blocks) s += v[u];           // block order
    }
    if (e < nw) gw[e] += s;
    else gb[e - nw] += s;
}

__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void fill_kernel(float* __restrict__ p, int64_t n, float v)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// grad = scale * 2 (x - t); per-block partial sums of (x - t)^2 (fixed order).
__global__ __launch_bounds__(256) void mse_kernel(const float* __restrict__ x,
                                                  const float* __restrict__ t, int64_t n,
                                                  float scale, float* __restrict__ grad,
                                                  float* __restrict__ part)
{
    float s = 0.f;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float d = x[i] - t[i];
        s += d * d;
        if (grad) grad[i] = scale * 2.f * d;
    }
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

}  // namespace

int64_t nn_wprep_halves() { return (int64_t)kMC * kMK; }

// the MFMA form takes 32 -> 32 channel layers
static bool mfma_layer(int cin, int cout, const float* resid)
{
    return cin == kMC && cout == kMC && resid == nullptr;
}

void launch_conv(const float* in, int rows, int cin, int cout, int n, const float* w,
                 const float* bias, const float* resid, bool gelu_in, float* out,
                 uint16_t* wscratch, hipStream_t st)
{
    if (cin < 1 || cin > kMaxC || cout < 1 || cout > kMaxC)
        throw Error(SGX_ERR_UNSUPPORTED, "net: channels must be 1..32");
    if (mfma_layer(cin, cout, resid)) {
        hipLaunchKernelGGL(wprep_kernel<false>, dim3((kMC * kMK + 255) / 256), dim3(256), 0, st,
                           w, wscratch);
        SGX_HIP(hipGetLastError());
        dim3 g((unsigned)((n + kMTile - 1) / kMTile), (unsigned)rows);
        if (gelu_in)
            hipLaunchKernelGGL((conv_mfma_kernel<false, true>), g, dim3(kThreads), 0, st, in, n,
                               wscratch, bias, nullptr, out);
        else
            hipLaunchKernelGGL((conv_mfma_kernel<false, false>), g, dim3(kThreads), 0, st, in,
                               n, wscratch, bias, nullptr, out);
        SGX_HIP(hipGetLastError());
        return;
    }
    dim3 grid((unsigned)((n + kTileN - 1) / kTileN), (unsigned)rows);
    if (gelu_in)
        hipLaunchKernelGGL((conv_kernel<false, true>), grid, dim3(kThreads), 0, st, in, cin,
                           cout, n, w, bias, resid, nullptr, out);
    else
        hipLaunchKernelGGL((conv_kernel<false, false>), grid, dim3(kThreads), 0, st, in, cin,
                           cout, n, w, bias, resid, nullptr, out);
    SGX_HIP(hipGetLastError());
}

bool net_fused_ok(int layers, int channels)
{
    return channels == kMC && layers >= 2 && layers <= kNetMaxL;
}

void launch_net_prep(const float* w, int layers, const int64_t* woff, uint16_t* wa,
                     hipStream_t st)
{
    if (layers <= 2) return;
    NetArgs a{};
    a.layers = layers;
    a.w = w;
    for (int l = 0; l < layers; ++l) a.woff[l] = (int)woff[l];
    a.wa = wa;
    hipLaunchKernelGGL(wprep_net_kernel, dim3((kMC * kMK + 255) / 256, layers - 2), dim3(256), 0,
                       st, a);
    SGX_HIP(hipGetLastError());
}

void launch_net_infer(const float* q, int rows, int n, const float* w, int layers,
                      const int64_t* woff, const int64_t* boff, const uint16_t* wa, float* out,
                      hipStream_t st)
{
    if (!net_fused_ok(layers, kMC))
        throw Error(SGX_ERR_UNSUPPORTED, "fused inference: 32 channels, 2..16 layers");
    NetArgs a{};
    a.q = q;
    a.out = out;
    a.n = n;
    a.layers = layers;
    a.w = w;
    for (int l = 0; l < layers; ++l) {
        a.woff[l] = (int)woff[l];
        a.boff[l] = (int)boff[l];
    }
    a.wa = const_cast<uint16_t*>(wa);
    const int T = kMTile - 8 * (layers - 2);
    dim3 grid((unsigned)((n + T - 1) / T), (unsigned)rows);
    hipLaunchKernelGGL(net_infer_kernel, grid, dim3(kThreads), 0, st, a);
    SGX_HIP(hipGetLastError());
}

void launch_conv_dgrad(const float* dout, int rows, int cin, int cout, int n, const float* w,
                       const float* zprev, float* din, uint16_t* wscratch, hipStream_t st)
{
    // layer (cin -> cout) with weights w[cout][cin][9]: din has cin channels
    if (cin < 1 || cin > kMaxC || cout < 1 || cout > kMaxC)
        throw Error(SGX_ERR_UNSUPPORTED, "net: channels must be 1..32");
    if (mfma_layer(cin, cout, nullptr)) {
        hipLaunchKernelGGL(wprep_kernel<true>, dim3((kMC * kMK + 255) / 256), dim3(256), 0, st, w,
                           wscratch);
        SGX_HIP(hipGetLastError());
        dim3 g((unsigned)((n + kMTile - 1) / kMTile), (unsigned)rows);
        hipLaunchKernelGGL((conv_mfma_kernel<true, false>), g, dim3(kThreads), 0, st, dout, n,
                           wscratch, nullptr, zprev, din);
        SGX_HIP(hipGetLastError());
        return;
    }
    dim3 grid((unsigned)((n + kTileN - 1) / kTileN), (unsigned)rows);
    hipLaunchKernelGGL((conv_kernel<true, false>), grid, dim3(kThreads), 0, st, dout, cout, cin,
                       n, w, nullptr, nullptr, zprev, din);
    SGX_HIP(hipGetLastError());
}

int nn_wgrad_blocks(int rows, int n)
{
    const int64_t items = (int64_t)rows * ((n + kTileN - 1) / kTileN);
    return (int)std::min<int64_t>(items, 512);
}

void launch_conv_wgrad(const float* dout, const float* x, bool gelu_in, int rows, int cin,
                       int cout, int n, float* part, float* gw, float* gb, hipStream_t st)
{
    const int blocks = nn_wgrad_blocks(rows, n);
    if (gelu_in && mfma_layer(cin, cout, nullptr))
        hipLaunchKernelGGL(wgrad_mfma_kernel, dim3(blocks), dim3(kThreads), 0, st, dout, x, rows,
                           n, part);
    else if (cin * cout <= kSmallPairs && gelu_in)
        hipLaunchKernelGGL(wgrad_small_kernel<true>, dim3(blocks), dim3(kThreads), 0, st, dout,
                           x, rows, cin, cout, n, part);
    else if (cin * cout <= kSmallPairs)
        hipLaunchKernelGGL(wgrad_small_kernel<false>, dim3(blocks), dim3(kThreads), 0, st, dout,
                           x, rows, cin, cout, n, part);
    else if (gelu_in)
        hipLaunchKernelGGL((wgrad_kernel<true>), dim3(blocks), dim3(kThreads), 0, st, dout, x,
                           rows, cin, cout, n, part);
    else
        hipLaunchKernelGGL((wgrad_kernel<false>), dim3(blocks), dim3(kThreads), 0, st, dout, x,
                           rows, cin, cout, n, part);
    SGX_HIP(hipGetLastError());
    const int len = cout * cin * kW + cout;
    hipLaunchKernelGGL(reduce_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st,
                       part, blocks, len, gw, gb, cout * cin * kW);
    SGX_HIP(hipGetLastError());
}

void launch_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                 float b2, float eps, int step, hipStream_t st)
{
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w, g, m,
                       v, n, lr, b1, b2, eps, c1, c2);
    SGX_HIP(hipGetLastError());
}

void launch_fill(float* p, int64_t n, float v, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    SGX_HIP(hipGetLastError());
}

int nn_mse_blocks(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 512); }

void launch_mse(const float* x, const float* t, int64_t n, float scale, float* grad, float* part,
                hipStream_t st)
{
    hipLaunchKernelGGL(mse_kernel, dim3((unsigned)nn_mse_blocks(n)), dim3(256), 0, st, x, t, n,
                       scale, grad, part);
    SGX_HIP(hipGetLastError());
}

}  // namespace sgx
