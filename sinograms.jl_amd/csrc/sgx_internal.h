// Internal declarations shared by the libsgx translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "sgx.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace sgx {

// Error carrying an sgx status code; converted to a return value at the C ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SGX_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw ::sgx::Error(SGX_ERR_DEVICE, std::string(#call) + ": " +             \
                                                   hipGetErrorString(e_));             \
    } while (0)

// atan2(y, x) in FP32 for the arc detectors' gamma: the usual octant reduction
// onto atan(u), u = min(|x|, |y|) / max(|x|, |y|) in [0, 1] (one reciprocal),
// then u * P(u^2) with P the degree-9 minimax (relative error) polynomial
// of csrc/gen_atan_coeffs.py: 1.3 ulp from float64 atan in FP32 Horner form,
// about 20 VALU instructions against 45 for the library's atan2f.
// (0, 0) gives NaN, which fails every bin test.
__device__ __forceinline__ float atan_unit(float u)
{
    // u * P(u^2) = atan(u) for |u| <= 1, odd in u
    const float u2 = u * u;
    float p = -1.753933378e-03f;
    p = fmaf(p, u2, 1.072695851e-02f);
    p = fmaf(p, u2, -3.080462478e-02f);
    p = fmaf(p, u2, 5.755212903e-02f);
    p = fmaf(p, u2, -8.377317339e-02f);
    p = fmaf(p, u2, 1.094203144e-01f);
    p = fmaf(p, u2, -1.426193267e-01f);
    p = fmaf(p, u2, 1.999826580e-01f);
    p = fmaf(p, u2, -3.333328366e-01f);
    p = fmaf(p, u2, 1.0f);
    return u * p;
}

__device__ __forceinline__ float atan2_fast(float y, float x)
{
    const float ay = fabsf(y), ax = fabsf(x);
    const float mx = fmaxf(ay, ax), mn = fminf(ay, ax);
    float a = atan_unit(mn * __builtin_amdgcn_rcpf(mx));
    if (ay > ax) a = 1.57079637f - a;            // pi/2 - atan(1/t)
    if (x < 0.f) a = 3.14159274f - a;            // pi - atan(t)
    return copysignf(a, y);
}

// atan2(y, x) for |y| < x, bit for bit what atan2_fast gives there (the
// same reciprocal and polynomial; the polynomial is odd); anywhere else it
// returns +-pi/2 with the sign of y.  Those points are at least 45 degrees
// off the x axis: for a detector that spans less than +-45 degrees both
// answers map off the detector, on the same side.
__device__ __forceinline__ float atan2_narrow(float y, float x)
{
    const float a = atan_unit(y * __builtin_amdgcn_rcpf(x));
    return fabsf(y) < x ? a : copysignf(1.57079637f, y);
}

// Per-view geometry of the backprojectors, formed on the device in FP64 and
// stored as FP32 for the inner loops.
struct ViewCS {
    float c, s;
};

// --- launch wrappers (defined in the .hip kernel files) -------------------

// Per-view cos/sin from angles (FP64 on the device); with `mojette` both are
// divided by max(|cos|, |sin|), the mojette view's radial spacing over d.
void launch_view_trig(const double* betas, int64_t n, int64_t stride, ViewCS* out,
                      hipStream_t st, bool mojette = false);

// K3: cone-beam voxel-driven backprojection.
enum { SGX_Q_FMT_F32 = 0, SGX_Q_FMT_BF16 = 1, SGX_Q_FMT_F16S = 2 };
struct CbctArgs {
    const void* q;          // (pitch, nt, na) float, bf16 or scaled fp16, column-major;
                            // 16-bit data must be followed by >= 4 bytes of readable padding
    int fmt;                // SGX_Q_FMT_*
    const float* vscale;    // per used view, for SGX_Q_FMT_F16S (else null)
    int ns, nt;
    int pitch;              // elements per detector row (>= ns, a multiple of 4)
    int nviews;             // views used (after ia_skip)
    int64_t view_stride;    // elements between used views
    const ViewCS* cs;       // nviews
    const float* xds;       // nx, x / ds
    const float* yds;       // ny
    const float* zds;       // nz
    int nx, ny, nz;
    const uint8_t* mask;    // nx*ny*nz or null
    float* image;
    float dsd_ds, dso_ds, ws, wt, inv_dt_ds;
    bool is_arc;
    // per-(block, view) detector windows of the LDS-staged kernel, from
    // launch_cbct_windows; null selects the gather kernel
    const int4* win;
    // options the reference leaves commented out (cbct-back.jl:36-59, 180,
    // 188): per used view the source z / ds (t' = mag (z - zs)), or null for
    // zs = 0; and for the arc detector the t magnification of the ray that
    // meets the detector cylinder, dsd / sqrt(x_beta^2 + y_beta^2)
    const float* zs;
    bool arc_t_exact;
    // the lateral source offset / ds (r_loop = x_beta - so_ds), 0 for none
    float so_ds = 0.f;
    // scaled-FP16 data with windows: true takes the packed-FP16 kernel in
    // place of the staged kernel's FP32 interpolation
    bool f16_packed = false;
    // staged kernel's block form: 0 = 16 z per lane, 1 = the mirror form
    // (pairs z and -z; see cbct_mirror_ok), 2 = 8 z per lane, 3 = 8 z per
    // lane with windows of up to 64 bins x 32 rows (coarse voxels)
    int zform = 0;
    // a staged launch over a chunk of a plan's views (q, cs, vscale, zs and
    // win already offset to its first view): the window table's views per
    // block (0: nviews), and whether the launch adds its sum to image
    // instead of storing it (the chunks after the first)
    int win_stride = 0;
    bool accumulate = false;
};
void launch_cbct_back(const CbctArgs& a, hipStream_t st);
// Whether the staged kernel may take its mirror form for `a`: the z axis
// is symmetric (z[nz-1-k] = -z[k], as the caller checks on its axis), no
// per-view source z, the t centre on the middle row, nz a multiple of 16.
bool cbct_mirror_ok(const CbctArgs& a, bool z_symmetric);

// Batch (task-table) form of the LDS-staged cone-beam backprojector: one
// item per reconstruction (device array), tasks (item, voxel block); block b
// does task b; the items of one table share a kernel variant.
struct CbctItem {
    const void* q;
    const ViewCS* cs;
    const float* vscale;
    const float* xds;
    const float* yds;
    const float* zds;
    const uint8_t* mask;
    float* image;
    const int4* win;
    const float* zs;
    int64_t view_stride;
    int ns, nt, pitch, nviews, nx, ny, nz, arc_t_exact, narrow;
    float dsd_ds, dso_ds, ws, wt, inv_dt_ds, so_ds;
};
// kernel variant of `a`: the staged kernel when a.win is set, else the
// gather kernel
int cbct_variant(const CbctArgs& a);
CbctItem cbct_item(const CbctArgs& a);
// blocks of the variant's kernel for `a`
int64_t cbct_item_blocks(const CbctArgs& a);
void launch_cbct_queue(int variant, const CbctItem* items, const int2* tasks, int ntasks,
                       hipStream_t st);
// Entries of the window table the staged kernel takes for these arguments.
int64_t cbct_window_entries(const CbctArgs& a);
// Fill the window table (needs cs, axes, geometry and sizes of `a`).
void launch_cbct_windows(const CbctArgs& a, int4* win, hipStream_t st);
// Bytes of readable padding the 16-bit forms need after the last view.
constexpr size_t kQ16Pad = 16;

// Analytic phantom projections in FP64 (k_phantom.hip): ellipse rows
// (cx, cy, a, b, theta, rho) along 2-D rays of kind SGX_SINO_*; axis-aligned
// ellipsoid rows (cx, cy, cz, a, b, c, rho) along cone-beam rays from the
// source to the detector point (lx[is], ly[is], t[it]) of the unrotated view,
// rotated by beta, with an optional per-view source z (zs, or null).
void launch_radon_ellipses(int kind, const double* s, int nb, const double* betas, int na,
                           double dsd, double dso, const double* comp, int nc, double* out,
                           hipStream_t st);
void launch_radon_ellipsoids(const double* lx, const double* ly, int ns, const double* t,
                             int nt, const double* betas, const double* zs, int na, double dso,
                             const double* comp, int nc, double* out, hipStream_t st);

// K1/K2: 2-D pixel-driven backprojection with view chunks in LDS.
struct Back2dArgs {
    const float* q;         // (nb, na) float
    int nb;
    int nviews;             // views used (after ia_skip)
    int64_t view_stride;    // elements between used views
    const ViewCS* cs;       // nviews
    const float* xds;       // nx
    const float* yds;       // ny
    int nx, ny;
    const uint8_t* mask;    // nx*ny or null
    float* image;
    int geom;               // 0 par, 1 fan arc, 2 fan flat
    float wb;               // (nb+1)/2 + offset
    float dsd_ds, dso_ds, so_ds;
    float* partial;         // workspace for view-group partial images (may be null)
    int64_t partial_cap;    // floats available in partial
    // shape: views are split into groups until pixel tiles x groups reaches
    // want_blocks; at most vc_max views per LDS chunk (and 48 KB)
    int want_blocks = 2048;
    int vc_max = 64;
    // arc fan: bounds for the tile-relative form (back2d_rel_bounds; negative:
    // not known, the per-pixel form runs), and pixel_form = true to keep the
    // per-pixel form (the exact transpose pair of sgx_plan_back / _back_t)
    float rel_rmax = -1.f, rel_h = -1.f;
    bool pixel_form = false;
};
void launch_back2d(const Back2dArgs& a, hipStream_t st);

// Task tables of the batch launches are interleaved over kQueueXcds XCDs
// (workgroups are dispatched to the XCDs round-robin: block b runs on XCD
// b % kQueueXcds).
constexpr int kQueueXcds = 8;

// Batch (task-table) form of the 2-D backprojector: one item per
// reconstruction (device array), tasks (item, pixel tile); block b does
// task b.
struct Back2dItem {
    const float* q;
    const ViewCS* cs;
    const float* xds;
    const float* yds;
    const uint8_t* mask;
    float* image;
    int64_t view_stride;
    int nb, nviews, nx, ny, vc, wshift;
    float wb, dsd_ds, dso_ds, so_ds;
};
// kernel variant of `a` (items of one queue share it); set a.vc_max = 64
// first, the queue's LDS chunk
int back2d_variant(const Back2dArgs& a);
Back2dItem back2d_item(const Back2dArgs& a);
int back2d_item_tiles(const Back2dItem& it);
int back2d_item_tiles_x(const Back2dItem& it);
size_t back2d_item_lds(const Back2dItem& it);
void launch_back2d_queue(int variant, const Back2dItem* items, const int2* tasks, int ntasks,
                         size_t lds, hipStream_t st);
// Series terms of the tile-relative arc form launch_back2d takes for `a`
// (0: the per-pixel form).
int back2d_rel_terms(const Back2dArgs& a);
// Host bounds of the tile-relative arc form from the image axes and mask:
// the largest |(x, y)| / ds of a backprojected pixel and the largest
// half-diagonal / ds of a 16 x 16 tile.
void back2d_rel_bounds(const double* xc, int64_t nx, const double* yc, int64_t ny,
                       const uint8_t* mask, double ds, float& rmax, float& h);
int64_t back2d_partial_floats(int nb, int nviews, int nx, int ny, int want_blocks, int vc_max);
// Exact transpose of launch_back2d (view_stride = nb): qt (nb, nviews) from an
// image; uses the geometry fields of `a` (q, image and partial are unused).
void launch_back2d_t(const Back2dArgs& a, const float* img, float* qt, hipStream_t st);

// K8: the learned correction's kernels (csrc/k_nn.hip).  Activations are
// (rows, channels, n); weights (cout, cin, 9) FP32, rounded to BF16 in use.
// wscratch: nn_wprep_halves() device halves for the MFMA form's BF16 weights
// (32 -> 32 channel layers); stream-ordered reuse across layers is safe
int64_t nn_wprep_halves();
void launch_conv(const float* in, int rows, int cin, int cout, int n, const float* w,
                 const float* bias, const float* resid, bool gelu_in, float* out,
                 uint16_t* wscratch, hipStream_t st);
void launch_conv_dgrad(const float* dout, int rows, int cin, int cout, int n, const float* w,
                       const float* zprev, float* din, uint16_t* wscratch, hipStream_t st);
// Fused inference of a whole network (32 channels, 2..16 layers): q (rows, n)
// -> out = q + f(q), one launch.  wa: the BF16 A operands of the 32 -> 32
// layers, (layers - 2) * nn_wprep_halves() halves, written by launch_net_prep.
bool net_fused_ok(int layers, int channels);
void launch_net_prep(const float* w, int layers, const int64_t* woff, uint16_t* wa,
                     hipStream_t st);
void launch_net_infer(const float* q, int rows, int n, const float* w, int layers,
                      const int64_t* woff, const int64_t* boff, const uint16_t* wa, float* out,
                      hipStream_t st);
int nn_wgrad_blocks(int rows, int n);
void launch_conv_wgrad(const float* dout, const float* x, bool gelu_in, int rows, int cin,
                       int cout, int n, float* part, float* gw, float* gb, hipStream_t st);
void launch_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                 float b2, float eps, int step, hipStream_t st);
void launch_fill(float* p, int64_t n, float v, hipStream_t st);
int nn_mse_blocks(int64_t n);
void launch_mse(const float* x, const float* t, int64_t n, float scale, float* grad, float* part,
                hipStream_t st);

// K4 (+K5): Toeplitz row filter with fused weights on MFMA.
struct FilterArgs {
    const float* in;        // (nb, ncols)
    int nb;
    int64_t ncols;
    int nt;                 // rows per view (ncols = nt * na); 1 for 2-D
    const float* g;         // npad, g = real(ifft(Hk))
    int npad;
    int extra;              // zero rows appended to each output column
    void* out;              // (nb + extra, ncols), float or bf16
    bool out_bf16;
    bool out_f16s;          // scaled FP16: q / vscale[view]
    bool split_bf16;        // split-BF16x3 products (else FP32 MFMA)
    bool split_f16;         // split-FP16x3 products (needs vbound and gscale)
    float gscale;           // split-FP16: the power of two applied to g (g * gscale <= 2^14)
    // fused weights (all optional): w = scale * wbin[m] * cyl(m, t) * ptab[m, view]
    float scale;
    const float* wbin;      // nb or null
    const float* wview;     // na or null (per-view weight, e.g. parallel-beam Parker)
    int cyl;                // 0 none, 1 flat, 2 arc
    float cyl_dsd, cyl_dso;
    const float* s_axis;    // nb (cyl)
    const float* t_axis;    // nt (cyl)
    const float* ptab;      // (nb, na) short-scan Parker weights, or null
    // scaled FP16: vbound[view] = max |weighted input| of the view (float
    // bits); the output scale is the power of two vscale[view] >= that bound
    // times g1 / 2^14, so |q| / vscale <= 2^14 and the differences fit FP16
    const unsigned int* vbound;
    float g1;               // sum |g| over the band the filter uses
    float* vscale;          // written per view
    // FFT form (npad = 512, 1024 or 2048): the table of fft_filter_tables and
    // its transform size (1024 or 2048); null selects the Toeplitz form
    const float* fft = nullptr;
    int fft_n = 0;
    // FFT form, cone beam: the per-(bin, row) part of the fused weight,
    // wbin[m] times the cylinder weight of (s[m], t[row]), (nb, nt), or null
    const float* wtab = nullptr;
    // scaled FP16 on the FFT form: a bound of |wbin[m] x cylinder weight|
    // over all (bin, row); > 0 selects the streaming range pass, whose view
    // bound is max |input| x |scale| x wmax x |wview[view]| (Parker weights
    // are <= 1), instead of the per-sample weighted maximum
    float wmax = 0.f;
};
// Host: the FFT form's table for the response Hk, 5 x N floats with the
// transform size N = max(npad, 1024): the twiddles W_N^(n1 k2) as
// re[k2 * 32 + n1] then im, the spectrum (Hk[k] + Hk[-k]) / (2 npad) (at
// npad = 512 on the even bins of N, 0 on the odd ones), and the per-bin
// factors of the fused weight
// (wbin[m], times cos(s[m] / dsd) for the arc cylinder weight; s[m]^2 for the
// flat one), 0 past nb.  wbin and s may be null (no such factor).  Empty when
// npad is not 512, 1024 or 2048, or when the environment sets SGX_FILTER_TOEPLITZ=1.
std::vector<float> fft_filter_tables(const float* Hk, int64_t npad, int64_t nb,
                                     const float* wbin, const double* s, int cyl, double dsd);
void launch_sino_filter(const FilterArgs& a, hipStream_t st);
// The fused pre-weights of a filter launch, (nb, ncols) FP32.
void launch_filter_weights(const FilterArgs& a, float* out, hipStream_t st);

// Per-view max |weighted input| of a filter launch (float bits) into
// vbound[na], zeroed first: the bound the scaled-FP16 output is sized by.
void launch_filter_bound(const FilterArgs& a, unsigned int* vbound, hipStream_t st);
// Fill with zeros by a kernel (usable under stream capture).
void launch_zero_u32(unsigned int* p, int64_t n, hipStream_t st);

// K6: distance-driven projector (FP64).
int64_t bdd_project_partial_doubles(int nPix, int nDet, int nangle);
void launch_bdd_project(const void* img, bool img_f32, int nPix, int nDet,
                        const double* angle, int nangle, double DSD, double DS0, double pSize,
                        double dSize, double* partial, double* sino, hipStream_t st);

// K6 adjoint: distance-driven backprojector (FP64, atomics over views).
void launch_bdd_backproject(const double* sino, int nPix, int nDet, const double* angle,
                            int nangle, double DSD, double DS0, double pSize, double dSize,
                            bool ref_alias, double* image, void* work, hipStream_t st);
// bytes of the workspace of launch_bdd_backproject (per-view knot tables and
// per-(view, tile) knot ranges)
int64_t bdd_back_work_bytes(int nPix, int nDet, int nangle, bool ref_alias);

// Parker weights (nb, na) in FP64 on the device, stored FP32.
void launch_parker(const double* gam, const double* bet, int nb, int na, double gamma_max,
                   float* out, hipStream_t st);

}  // namespace sgx
