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
// Backprojection, owner-computes (no atomics, the reference's summation
// order).  A first kernel (one block per view) forms the view's mapped
// detector boundaries and the prefix integral of sino .* L .* detSize at them
// (ascending knots, as the reference's reversal makes them) and, for every
// 16 x 16 image tile, the range of knots that the tile's mapped pixel
// boundaries fall in (the mapping is monotone along each axis, so the tile's
// corners bound it; one knot of margin).  The second kernel gives each block
// one tile and walks the views in order: the tile's knots are staged in LDS,
// the 16 x 17 mapped pixel boundaries are interpolated once each, and every
// thread adds its pixel's |difference| to a register.  A boundary outside the
// staged knots (or a range wider than the LDS slot) is interpolated from the
// view's full table in global memory, so the result never depends on the
// range.  The reference's first-view quirk (bdd_2d.jl:139-140: the first
// view's row buffer aliases the image, so it is added once more in the
// untransposed frame) is an extra pass before view 0 in the identity frame.
// Each pixel's sum runs over the passes in the reference's order.

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

// Block-wide inclusive scan of LDS values v[0..n) (n <= 4096), in place:
// per-thread serial chunks, their totals scanned across each wave with
// shuffles, then the wave totals (tmp holds kThreads / 64 of them).
__device__ void block_scan(double* v, double* tmp, int n)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (n + kThreads - 1) / kThreads;
    const int b = t * per, e = min(n, b + per);
    double s = 0;
    for (int i = b; i < e; ++i) s += v[i];
    double x = s;                              // inclusive scan over the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) tmp[w] = x;
    double ex = __shfl_up(x, 1, 64);           // exclusive within the wave
    if (lane == 0) ex = 0;
    __syncthreads();
    double a = 0;
    for (int k = 0; k < w; ++k) a += tmp[k];
    a += ex;
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

constexpr int kBT = 16;             // owner tile edge (pixels)
constexpr int kKnotMax = 256;       // staged knots per (tile, pass)

// The view frame's rows [R0, R0 + kBT) and pixel boundaries [J0, J0 + kBT]
// that image tile (ta, tb) covers.  Identity frame: image[r, j]; transposed
// (views mapped on the y axis): image[nPix - 1 - j, r].
__device__ __forceinline__ void tile_frame(bool ident, int ta, int tb, int nPix, int& R0, int& J0)
{
    if (ident) {
        R0 = kBT * ta;
        J0 = kBT * tb;
    } else {
        R0 = kBT * tb;
        J0 = nPix - kBT - kBT * ta;
    }
}

// interp_flat on a view's full table of {x, P} knots (ascending)
__device__ __forceinline__ double interp_knots(const double2* __restrict__ kn, int n, double x)
{
    if (x <= kn[0].x) return kn[0].y;
    if (x >= kn[n - 1].x) return kn[n - 1].y;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (kn[mid].x <= x) lo = mid;
        else hi = mid;
    }
    const double2 p = kn[lo], q = kn[lo + 1];
    const double t = (x - p.x) / (q.x - p.x);
    return p.y + t * (q.y - p.y);
}

// Per view: knots[view] = {mapped detector boundary, prefix integral}
// (nDet + 1, ascending), and per (pass, tile) the knot range of the tile.
// Passes: [view 0 in the identity frame when ref_alias], then views in order.
__global__ __launch_bounds__(kThreads) void bdd_back_table_kernel(
    const double* __restrict__ sino, int nPix, int nDet, const double* __restrict__ angle,
    int nangle, double DSD, double DS0, double pSize, double dSize, int ref_alias, int tiles_a,
    int ntiles, double2* __restrict__ knots, int2* __restrict__ ranges)
{
    extern __shared__ double sm[];
    double* xs = sm;                           // nDet + 1 knots (ascending)
    double* ys = xs + (nDet + 1);              // nDet + 1 prefix integral
    double* tmp = ys + (nDet + 1);             // kThreads
    double* cell = tmp + kThreads;             // nDet: sino .* L .* detSize

    const int view = blockIdx.x;
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
    double2* kv = knots + (int64_t)view * (nDet + 1);
    for (int k = threadIdx.x; k <= nDet; k += kThreads) kv[k] = make_double2(xs[k], ys[k]);

    const int npasses_here = (ref_alias && view == 0) ? 2 : 1;
    for (int q = 0; q < npasses_here; ++q) {
        const bool alias = q == 1;
        const bool ident = alias || v.axis_x;
        const int pass = alias ? 0 : view + (ref_alias ? 1 : 0);
        for (int t = threadIdx.x; t < ntiles; t += kThreads) {
            int R0, J0;
            tile_frame(ident, t % tiles_a, t / tiles_a, nPix, R0, J0);
            const int r0 = max(R0, 0), r1 = min(R0 + kBT - 1, nPix - 1);
            const int j0 = max(J0, 0), j1 = min(J0 + kBT, nPix);
            double lo = 1e308, hi = -1e308;
            for (int c = 0; c < 4; ++c) {
                const double x = pix_boundary(v, nPix, pSize, (c & 1) ? r1 : r0, (c & 2) ? j1 : j0);
                lo = fmin(lo, x);
                hi = fmax(hi, x);
            }
            // largest k with xs[k] <= lo, smallest k with xs[k] >= hi
            int a = 0, b = nDet;
            while (b - a > 1) {
                const int m = (a + b) >> 1;
                if (xs[m] <= lo) a = m;
                else b = m;
            }
            const int klo = max(a - 1, 0);
            a = 0;
            b = nDet;
            while (b - a > 1) {
                const int m = (a + b) >> 1;
                if (xs[m] >= hi) b = m;
                else a = m;
            }
            const int khi = min(b + 1, nDet);
            ranges[(int64_t)pass * ntiles + t] = make_int2(klo, khi);
        }
    }
}

// Block = one 16 x 16 image tile, all passes in order (see above); one
// thread per mapped pixel boundary of the tile (16 x 17), the first 256 of
// them also own a pixel.
constexpr int kOwnThreads = kBT * (kBT + 1);
__global__ __launch_bounds__(kOwnThreads) void bdd_back_kernel(
    const double2* __restrict__ knots, const int2* __restrict__ ranges, int npass,
    int ref_alias, int nPix, int nDet, const double* __restrict__ angle, double DS0,
    double pSize, int tiles_a, int ntiles, double* __restrict__ image)
{
    __shared__ double2 kn[2][kKnotMax];
    __shared__ double P[2][kBT][kBT + 1];      // [row][boundary] of the view frame
    const int t = blockIdx.x;
    const int ta = t % tiles_a, tb = t / tiles_a;
    const int tid = threadIdx.x;
    const int tx = tid % kBT, ty = tid / kBT;
    const int a = kBT * ta + tx, b = kBT * tb + ty;      // image[a, b]
    const int nk = nDet + 1;
    double acc = 0;

    auto view_of = [&](int p) { return ref_alias ? max(p - 1, 0) : p; };
    auto stage = [&](int p, int buf) {
        const int2 rg = ranges[(int64_t)p * ntiles + t];
        const int m = rg.y - rg.x + 1;
        if (m <= kKnotMax) {
            const double2* src = knots + (int64_t)view_of(p) * nk + rg.x;
            for (int i = tid; i < m; i += kOwnThreads) kn[buf][i] = src[i];
        }
    };
    stage(0, 0);
    __syncthreads();
    for (int p = 0; p < npass; ++p) {
        const int buf = p & 1;
        const int view = view_of(p);
        const double beta = angle[view];
        const ViewGeo v = view_geo(beta, DS0);
        const bool ident = (ref_alias && p == 0) || v.axis_x;
        int R0, J0;
        tile_frame(ident, ta, tb, nPix, R0, J0);
        const int2 rg = ranges[(int64_t)p * ntiles + t];
        const int m = rg.y - rg.x + 1;
        const double2* full = knots + (int64_t)view * nk;
        const double2* st = kn[buf];
        const int rho = tid / (kBT + 1), be = tid % (kBT + 1);
        const int r = R0 + rho, j = J0 + be;
        if (r >= 0 && r < nPix && j >= 0 && j <= nPix) {
            const double x = pix_boundary(v, nPix, pSize, r, j);
            const bool in_lo = x > st[0].x || rg.x == 0;
            const bool in_hi = x < st[m - 1].x || rg.y == nDet;
            double val;
            if (m > kKnotMax || !in_lo || !in_hi) {
                val = interp_knots(full, nk, x);
            } else if (x <= st[0].x) {
                val = st[0].y;                   // the table's first knot: flat
            } else if (x >= st[m - 1].x) {
                val = st[m - 1].y;               // the table's last knot: flat
            } else {
                int lo = 0, hi = m - 1;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (st[mid].x <= x) lo = mid;
                    else hi = mid;
                }
                const double2 u = st[lo], w = st[lo + 1];
                const double tt = (x - u.x) / (w.x - u.x);
                val = u.y + tt * (w.y - u.y);
            }
            P[buf][rho][be] = val;
        }
        if (p + 1 < npass) stage(p + 1, buf ^ 1);
        __syncthreads();
        if (tid < kBT * kBT && a < nPix && b < nPix) {
            acc += ident ? fabs(P[buf][tx][ty + 1] - P[buf][tx][ty])
                         : fabs(P[buf][ty][kBT - tx] - P[buf][ty][kBT - 1 - tx]);
        }
    }
    if (tid < kBT * kBT && a < nPix && b < nPix) image[a + (int64_t)nPix * b] = acc;
}

}  // namespace

int64_t bdd_back_work_bytes(int nPix, int nDet, int nangle, bool ref_alias)
{
    const int64_t tiles = (int64_t)((nPix + kBT - 1) / kBT) * ((nPix + kBT - 1) / kBT);
    const int64_t knots = (int64_t)nangle * (nDet + 1) * (int64_t)sizeof(double2);
    return knots + (nangle + (ref_alias ? 1 : 0)) * tiles * (int64_t)sizeof(int2);
}

void launch_bdd_backproject(const double* sino, int nPix, int nDet, const double* angle,
                            int nangle, double DSD, double DS0, double pSize, double dSize,
                            bool ref_alias, double* image, void* work, hipStream_t st)
{
    const size_t lds = sizeof(double) * (2 * (nDet + 1) + kThreads + nDet);
    if (lds > 64 * 1024)
        throw Error(SGX_ERR_UNSUPPORTED, "backproject_bdd: needs 3 nDet + 258 <= 8192");
    const int tiles_a = (nPix + kBT - 1) / kBT;
    const int ntiles = tiles_a * tiles_a;
    const int npass = nangle + (ref_alias ? 1 : 0);
    double2* knots = static_cast<double2*>(work);
    int2* ranges = reinterpret_cast<int2*>(knots + (int64_t)nangle * (nDet + 1));
    hipLaunchKernelGGL(bdd_back_table_kernel, dim3(nangle), dim3(kThreads), lds, st, sino, nPix,
                       nDet, angle, nangle, DSD, DS0, pSize, dSize, (int)ref_alias, tiles_a,
                       ntiles, knots, ranges);
    SGX_HIP(hipGetLastError());
    hipLaunchKernelGGL(bdd_back_kernel, dim3(ntiles), dim3(kOwnThreads), 0, st, knots, ranges,
                       npass, (int)ref_alias, nPix, nDet, angle, DS0, pSize, tiles_a, ntiles,
                       image);
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
