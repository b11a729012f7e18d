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
    // input tile with halo, zero outside [0, n)
    const float* src = in + (int64_t)r * cin * n;
    for (int e = tid; e < cin * (kTileN + kW - 1); e += kThreads) {
        const int i = e / (kTileN + kW - 1), j = e % (kTileN + kW - 1);
        const int m = n0 + j - kHalf;
        float v = 0.f;
        if (m >= 0 && m < n) {
            v = src[(int64_t)i * n + m];
            if (GELU_IN) v = gelu(v);
        }
        tile[i][j] = v;
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

// Sum the partials of `blocks` blocks in block order, add into grad (per
// element: the weights then the biases of one layer).
__global__ void reduce_kernel(const float* __restrict__ part, int blocks, int len,
                              float* __restrict__ gw, float* __restrict__ gb, int nw)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len) return;
    float s = 0.f;
    for (int b = 0; b < blocks; ++b) s += part[(int64_t)b * len + e];
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

void launch_conv(const float* in, int rows, int cin, int cout, int n, const float* w,
                 const float* bias, const float* resid, bool gelu_in, float* out,
                 hipStream_t st)
{
    if (cin < 1 || cin > kMaxC || cout < 1 || cout > kMaxC)
        throw Error(SGX_ERR_UNSUPPORTED, "net: channels must be 1..32");
    dim3 grid((unsigned)((n + kTileN - 1) / kTileN), (unsigned)rows);
    if (gelu_in)
        hipLaunchKernelGGL((conv_kernel<false, true>), grid, dim3(kThreads), 0, st, in, cin,
                           cout, n, w, bias, resid, nullptr, out);
    else
        hipLaunchKernelGGL((conv_kernel<false, false>), grid, dim3(kThreads), 0, st, in, cin,
                           cout, n, w, bias, resid, nullptr, out);
    SGX_HIP(hipGetLastError());
}

void launch_conv_dgrad(const float* dout, int rows, int cin, int cout, int n, const float* w,
                       const float* zprev, float* din, hipStream_t st)
{
    // layer (cin -> cout) with weights w[cout][cin][9]: din has cin channels
    if (cin < 1 || cin > kMaxC || cout < 1 || cout > kMaxC)
        throw Error(SGX_ERR_UNSUPPORTED, "net: channels must be 1..32");
    dim3 grid((unsigned)((n + kTileN - 1) / kTileN), (unsigned)rows);
    hipLaunchKernelGGL((conv_kernel<true, false>), grid, dim3(kThreads), 0, st, dout, cout, cin,
                       n, w, nullptr, nullptr, zprev, din);
    SGX_HIP(hipGetLastError());
}

int nn_wgrad_blocks(int rows, int n)
{
    const int64_t items = (int64_t)rows * ((n + kTileN - 1) / kTileN);
    return (int)std::min<int64_t>(items, 1024);
}

void launch_conv_wgrad(const float* dout, const float* x, bool gelu_in, int rows, int cin,
                       int cout, int n, float* part, float* gw, float* gb, hipStream_t st)
{
    const int blocks = nn_wgrad_blocks(rows, n);
    if (gelu_in)
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
