// K6: branchless distance-driven (DD) fan-beam projector and its
// backprojector, for gfx950.
//
// Replaces project_bdd / backproject_bdd (src/sys2/bdd_2d.jl:17-220).  Per
// view beta (all in FP64, as the reference computes):
//   * tube (-DS0 sin b, DS0 cos b); the view maps every boundary onto the x
//     axis (axisCase, bdd_2d.jl:53-57) or the y axis, through the tube:
//       map2x(x1,y1,x2,y2) = x1 - y1 (x1-x2)/(y1-y2)     (bdd_2d.jl:260)
//       map2y(x1,y1,x2,y2) = y1 - x1 (y1-y2)/(x1-x2)     (bdd_2d.jl:249)
//   * detector cell scale L_k = |pSize / (detSize_k cos or sin(theta_k))|;
//   * projection: for every image row, the prefix integral of the row over
//     its mapped pixel boundaries (integrate1D, :229-241) is interpolated
//     linearly, held flat outside (Interpolations' Flat()), at the mapped
//     detector boundaries, and |differences| are summed over rows, then
//     scaled by L (:88-104).
//   * backprojection: the prefix integral of sino .* L over the detector
//     boundaries is interpolated at every row's mapped pixel boundaries and
//     |differences| are added to the image (:189-213).
//
// Layout.  Projection: a block owns one view and a group of 32 image rows.
// Each row's samples are read in one pass, the mapped boundaries and the
// prefix integral are formed in LDS (a block-wide scan), and each thread
// interpolates a strided set of detector boundaries with a binary search in
// LDS.  The group's per-cell sums go to a partial buffer; a second kernel
// adds the row groups in a fixed order and applies L (deterministic).
// Backprojection: a block owns one view and a group of rows; the detector
// prefix integral is built once per block in LDS and each thread
// interpolates pixel boundaries; image sums over views use FP64 atomics.

#include "sgx_internal.h"
#include "sgx.h"

namespace sgx {
namespace {

constexpr int kThreads = 256;
constexpr int kRows = 32;          // image rows per block

struct ViewGeo {
    double tx, ty;
    bool axis_x;
};

__device__ __forceinline__ ViewGeo view_geo(double beta, double DS0)
{
    ViewGeo v;
    double s, c;
    sincos(beta, &s, &c);
    v.tx = -DS0 * s;
    v.ty = DS0 * c;
    const double pi = 3.14159265358979323846;
    v.axis_x = !(((beta >= pi / 4) && (beta <= 3 * pi / 4)) ||
                 ((beta >= 5 * pi / 4) && (beta <= 7 * pi / 4)));
    return v;
}

// mapped detector boundary k (0 <= k <= nDet), bdd_2d.jl:28-29, 49-50, 60-68
__device__ __forceinline__ double det_boundary(const ViewGeo& v, double beta, double DSD,
                                               double DS0, double dSize, int nDet, int k)
{
    double s, c;
    sincos(beta, &s, &c);
    const double detX = (-(double)nDet / 2 + (double)k) * dSize;
    const double detY = -(DSD - DS0) - dSize / 2;
    const double rx = detX * c - detY * s;
    const double ry = detX * s + detY * c;
    return v.axis_x ? v.tx - v.ty * (v.tx - rx) / (v.ty - ry)
                    : v.ty - v.tx * (v.ty - ry) / (v.tx - rx);
}

// mapped pixel boundary j (0 <= j <= nPix) of row r in the orientation of
// img[row, :] (bdd_2d.jl:32-35, 60-68)
__device__ __forceinline__ double pix_boundary(const ViewGeo& v, int nPix, double pSize, int r,
                                               int j)
{
    if (v.axis_x) {
        const double pX = (-(double)nPix / 2 + (double)j) * pSize;
        const double pY = ((double)nPix / 2 - (double)r - 0.5) * pSize;
        return v.tx - v.ty * (v.tx - pX) / (v.ty - pY);
    }
    const double pX = (-(double)nPix / 2 + (double)r) * pSize;
    const double pY = ((double)nPix / 2 - (double)(nPix - j) - 0.5) * pSize;
    return v.ty - v.tx * (v.ty - pY) / (v.tx - pX);
}

// Interpolations.linear_interpolation(xs, ys, Flat()) on ascending knots.
__device__ __forceinline__ double interp_flat(const double* xs, const double* ys, int n, double x)
{
    if (x <= xs[0]) return ys[0];
    if (x >= xs[n - 1]) return ys[n - 1];
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (xs[mid] <= x) lo = mid;
        else hi = mid;
    }
    const double t = (x - xs[lo]) / (xs[lo + 1] - xs[lo]);
    return ys[lo] + t * (ys[lo + 1] - ys[lo]);
}

// Block-wide inclusive scan of LDS values v[0..n) (n <= 4096), in place.
__device__ void block_scan(double* v, double* tmp, int n)
{
    // per-thread serial chunk, then a scan of the chunk totals
    const int t = threadIdx.x;
    const int per = (n + kThreads - 1) / kThreads;
    const int b = t * per, e = min(n, b + per);
    double s = 0;
    for (int i = b; i < e; ++i) s += v[i];
    tmp[t] = s;
    __syncthreads();
    if (t == 0) {
        double a = 0;
        for (int i = 0; i < kThreads; ++i) {
            const double x = tmp[i];
            tmp[i] = a;
            a += x;
        }
    }
    __syncthreads();
    double a = tmp[t];
    for (int i = b; i < e; ++i) {
        a += v[i];
        v[i] = a;
    }
    __syncthreads();
}

template <typename TI>
__global__ __launch_bounds__(kThreads) void bdd_project_kernel(
    const TI* __restrict__ img, int nPix, int nDet, const double* __restrict__ angle, int nangle,
    double DSD, double DS0, double pSize, double dSize, double* __restrict__ partial)
{
    extern __shared__ double sm[];
    double* detm = sm;                         // nDet + 1
    double* rowm = detm + (nDet + 1);          // nPix + 1
    double* ppj = rowm + (nPix + 1);           // nPix + 1 (prefix integral)
    double* pdk = ppj + (nPix + 1);            // nDet + 1
    double* tmp = pdk + (nDet + 1);            // kThreads

    const int view = blockIdx.x;
    const int r0 = blockIdx.y * kRows;
    const double beta = angle[view];
    const ViewGeo v = view_geo(beta, DS0);
    for (int k = threadIdx.x; k <= nDet; k += kThreads)
        detm[k] = det_boundary(v, beta, DSD, DS0, dSize, nDet, k);

    // per-thread cell accumulators for cells k = tid, tid + 256, ...
    constexpr int kMaxCells = 16;              // nDet <= 4096
    double acc[kMaxCells];
#pragma unroll
    for (int i = 0; i < kMaxCells; ++i) acc[i] = 0;

    for (int r = r0; r < min(nPix, r0 + kRows); ++r) {
        __syncthreads();
        for (int j = threadIdx.x; j <= nPix; j += kThreads) rowm[j] = pix_boundary(v, nPix, pSize, r, j);
        __syncthreads();
        // img[row, :]: axis_x -> phantom[r, j]; else reverse(phantom', dims=2)[r, j]
        if (threadIdx.x == 0) ppj[0] = 0;
        for (int j = threadIdx.x; j < nPix; j += kThreads) {
            const double val = v.axis_x ? (double)img[r + (int64_t)nPix * j]
                                        : (double)img[(nPix - 1 - j) + (int64_t)nPix * r];
            ppj[j + 1] = val * (rowm[j + 1] - rowm[j]);
        }
        __syncthreads();
        block_scan(ppj + 1, tmp, nPix);
        for (int k = threadIdx.x; k <= nDet; k += kThreads)
            pdk[k] = interp_flat(rowm, ppj, nPix + 1, detm[k]);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kMaxCells; ++i) {
            const int k = threadIdx.x + i * kThreads;
            if (k < nDet) acc[i] += fabs(pdk[k + 1] - pdk[k]);
        }
    }
    double* out = partial + ((int64_t)blockIdx.y * nangle + view) * nDet;
#pragma unroll
    for (int i = 0; i < kMaxCells; ++i) {
        const int k = threadIdx.x + i * kThreads;
        if (k < nDet) out[k] = acc[i];
    }
}

__global__ void bdd_finish_kernel(const double* __restrict__ partial, int groups, int nDet,
                                  const double* __restrict__ angle, int nangle, double DSD,
                                  double DS0, double pSize, double dSize,
                                  double* __restrict__ sino)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)nangle * nDet) return;
    const int view = (int)(i / nDet), k = (int)(i % nDet);
    const double beta = angle[view];
    const ViewGeo v = view_geo(beta, DS0);
    const double d0 = det_boundary(v, beta, DSD, DS0, dSize, nDet, k);
    const double d1 = det_boundary(v, beta, DSD, DS0, dSize, nDet, k + 1);
    const double mid = (d0 + d1) / 2, dsz = d1 - d0;
    double L;
    if (v.axis_x) {
        const double th = atan(fabs(v.tx - mid) / fabs(v.ty));
        L = fabs(pSize / (dsz * cos(th)));
    } else {
        const double th = atan(1.0 / (fabs(v.ty - mid) / fabs(v.tx)));   // acot
        L = fabs(pSize / (dsz * sin(th)));
    }
    double s = 0;
    for (int g = 0; g < groups; ++g) s += partial[((int64_t)g * nangle + view) * nDet + k];
    sino[view + (int64_t)nangle * k] = s * L;                           // (nangle, nDet)
}

// Backprojection: block = (view, group of kRows image rows).
__global__ __launch_bounds__(kThreads) void bdd_back_kernel(
    const double* __restrict__ sino, int nPix, int nDet, const double* __restrict__ angle,
    int nangle, double DSD, double DS0, double pSize, double dSize, int ref_alias,
    double* __restrict__ image)
{
    extern __shared__ double sm[];
    double* xs = sm;                           // nDet + 1 knots (ascending)
    double* ys = xs + (nDet + 1);              // nDet + 1 prefix integral
    double* pdk = ys + (nDet + 1);             // nPix + 1
    double* tmp = pdk + (nPix + 1);            // kThreads
    double* cell = tmp + kThreads;             // nDet: sino .* L .* detSize, then detm

    const int view = blockIdx.x;
    const int r0 = blockIdx.y * kRows;
    const double beta = angle[view];
    const ViewGeo v = view_geo(beta, DS0);
    // detector boundaries (into xs), then per-cell sino * L * detSize
    for (int k = threadIdx.x; k <= nDet; k += kThreads)
        xs[k] = det_boundary(v, beta, DSD, DS0, dSize, nDet, k);
    __syncthreads();
    for (int k = threadIdx.x; k < nDet; k += kThreads) {
        const double d0 = xs[k], d1 = xs[k + 1];
        const double mid = (d0 + d1) / 2, dsz = d1 - d0;
        double L;
        if (v.axis_x) {
            const double th = atan(fabs(v.tx - mid) / fabs(v.ty));
            L = fabs(pSize / (dsz * cos(th)));
        } else {
            const double th = atan(1.0 / (fabs(v.ty - mid) / fabs(v.tx)));
            L = fabs(pSize / (dsz * sin(th)));
        }
        cell[k] = sino[view + (int64_t)nangle * k] * L * dsz;        // :196
    }
    __syncthreads();
    if (threadIdx.x == 0) ys[0] = 0;
    for (int k = threadIdx.x; k < nDet; k += kThreads) ys[k + 1] = cell[k];
    __syncthreads();
    block_scan(ys + 1, tmp, nDet);
    // descending detector boundaries: reverse knots and values (:198-202)
    const bool desc = xs[0] > xs[1];
    __syncthreads();                           // every thread has read xs[0], xs[1]
    if (desc) {
        for (int k = threadIdx.x; k <= nDet / 2; k += kThreads) {
            const int k2 = nDet - k;
            if (k < k2) {
                double a = xs[k]; xs[k] = xs[k2]; xs[k2] = a;
                double b = ys[k]; ys[k] = ys[k2]; ys[k2] = b;
            }
        }
    }
    __syncthreads();

    for (int r = r0; r < min(nPix, r0 + kRows); ++r) {
        for (int j = threadIdx.x; j <= nPix; j += kThreads)
            pdk[j] = interp_flat(xs, ys, nDet + 1, pix_boundary(v, nPix, pSize, r, j));
        __syncthreads();
        for (int j = threadIdx.x; j < nPix; j += kThreads) {
            const double val = fabs(pdk[j + 1] - pdk[j]);                 // :205
            // axis_x: image[r, j]; else reverse(tmp, dims=2)' -> image[nPix-1-j, r]
            const int64_t o = v.axis_x ? r + (int64_t)nPix * j : (nPix - 1 - j) + (int64_t)nPix * r;
            atomicAdd(image + o, val);
            // bdd_2d.jl:139-140: the first view's raw buffer is added once more
            if (ref_alias && view == 0) atomicAdd(image + r + (int64_t)nPix * j, val);
        }
        __syncthreads();
    }
}

__global__ void zero_f64_kernel(double* p, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

}  // namespace

void launch_bdd_backproject(const double* sino, int nPix, int nDet, const double* angle,
                            int nangle, double DSD, double DS0, double pSize, double dSize,
                            bool ref_alias, double* image, hipStream_t st)
{
    const size_t lds = sizeof(double) * (2 * (nDet + 1) + (nPix + 1) + kThreads + nDet);
    if (lds > 64 * 1024)
        throw Error(SGX_ERR_UNSUPPORTED, "backproject_bdd: needs 3 nDet + nPix + 259 <= 8192");
    const int64_t n = (int64_t)nPix * nPix;
    hipLaunchKernelGGL(zero_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       image, n);
    SGX_HIP(hipGetLastError());
    dim3 grid(nangle, (nPix + kRows - 1) / kRows);
    hipLaunchKernelGGL(bdd_back_kernel, grid, dim3(kThreads), lds, st, sino, nPix, nDet, angle,
                       nangle, DSD, DS0, pSize, dSize, (int)ref_alias, image);
    SGX_HIP(hipGetLastError());
}

int64_t bdd_project_partial_doubles(int nPix, int nDet, int nangle)
{
    return (int64_t)((nPix + kRows - 1) / kRows) * nangle * nDet;
}

void launch_bdd_project(const void* img, bool img_f32, int nPix, int nDet,
                        const double* angle, int nangle, double DSD, double DS0, double pSize,
                        double dSize, double* partial, double* sino, hipStream_t st)
{
    const int groups = (nPix + kRows - 1) / kRows;
    const size_t lds = sizeof(double) * (2 * (nDet + 1) + 2 * (nPix + 1) + kThreads);
    if (nDet > 16 * kThreads || lds > 64 * 1024)
        throw Error(SGX_ERR_UNSUPPORTED, "project_bdd: needs nDet <= 4096 and "
                                         "2 (nDet + nPix) + 260 <= 8192");
    dim3 grid(nangle, groups);
    if (img_f32)
        hipLaunchKernelGGL(bdd_project_kernel<float>, grid, dim3(kThreads), lds, st,
                           static_cast<const float*>(img), nPix, nDet, angle, nangle, DSD, DS0,
                           pSize, dSize, partial);
    else
        hipLaunchKernelGGL(bdd_project_kernel<double>, grid, dim3(kThreads), lds, st,
                           static_cast<const double*>(img), nPix, nDet, angle, nangle, DSD,
                           DS0, pSize, dSize, partial);
    SGX_HIP(hipGetLastError());
    const int64_t n = (int64_t)nangle * nDet;
    hipLaunchKernelGGL(bdd_finish_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       partial, groups, nDet, angle, nangle, DSD, DS0, pSize, dSize, sino);
    SGX_HIP(hipGetLastError());
}

}  // namespace sgx
