// CPU restatement of the Sinograms.jl reconstruction hot loops.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP library is
// checked against; it is never linked into, loaded by, or called from the
// product path.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
// of bench.py may load it.
//
// Every routine is a template instantiated for double (the Float64 oracle) and
// float (the "reference precision" mode, which follows the reference's
// Float32 evaluation order: no FMA contraction, same operand order, same
// floor/bounds logic).  Compile with -ffp-contract=off so the float mode is
// not fused behind our back.
//
// Layouts are Julia column-major: sino[ib, ia] at ib + nb*ia,
// proj[is, it, ia] at is + ns*(it + nt*ia), image[ix, iy, iz] with x fastest.
//
// Parity pins (see tests/test_oracle.py): unit-impulse views, bin-centre
// hits, constant-input closed forms, out-of-range bins, integral
// conservation on analytic disks/ellipsoids with the reference's own
// tolerances, brute-force DD overlap and the DD adjoint identity.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Julia's floor(Int, x) for finite x.
template <class T>
inline int64_t jfloor(T x) { return (int64_t)std::floor(x); }

// ---------------------------------------------------------------------------
// fbp_back_par! / fbp_back_par_xy   (src/fbp2/back-par.jl:206-315)
// sinphi/cosphi have length na_sub = ceil(na / ia_skip); views used are
// ia = 0, ia_skip, 2 ia_skip, ... (back-par.jl:225-227).
// ---------------------------------------------------------------------------
template <class T>
void back_par(const T* sino, int64_t nb, int64_t na, int64_t ia_skip,
              const double* sinphi, const double* cosphi, double ds, double offset,
              const double* xc, int64_t nx, const double* yc, int64_t ny,
              const uint8_t* mask, T* image)
{
    const int64_t na_sub = (na + ia_skip - 1) / ia_skip;
    const T wb = (T)(float)((nb + 1) / 2.0 + (double)(float)offset); // Toffset(...)
    std::vector<T> s(na_sub), c(na_sub), xd(nx), yd(ny);
    for (int64_t i = 0; i < na_sub; ++i) { s[i] = (T)sinphi[i]; c[i] = (T)cosphi[i]; }
    for (int64_t i = 0; i < nx; ++i) xd[i] = (T)xc[i] / (T)ds;            // xc / ds
    for (int64_t i = 0; i < ny; ++i) yd[i] = (T)yc[i] / (T)ds;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < nx * ny; ++p) {
        if (!mask[p]) { image[p] = 0; continue; }
        const T x = xd[p % nx], y = yd[p / nx];
        T pixel = 0;
        for (int64_t ja = 0; ja < na_sub; ++ja) {
            const T* q = sino + nb * (ja * ia_skip);
            const T bb = x * c[ja] + y * s[ja] + wb;                      // :293
            const int64_t il = jfloor(bb);                                // :304
            const int64_t ir = 1 + il;
            if (1 <= il && ir <= nb) {                                    // :307
                const T wr = bb - (T)il;
                const T wl = 1 - wr;
                pixel += wl * q[il - 1] + wr * q[ir - 1];                 // :310
            }
        }
        image[p] = pixel;
    }
}

// ---------------------------------------------------------------------------
// fbp_back_fan! / fbp_back_fan_xy   (src/fbp2/back-fan.jl:215-336)
// ---------------------------------------------------------------------------
template <class T>
void back_fan(const T* sino, int64_t nb, int64_t na, int64_t ia_skip,
              const double* sinb, const double* cosb,
              double dsd, double dso, double source_offset, int is_arc,
              double ds, double offset,
              const double* xc, int64_t nx, const double* yc, int64_t ny,
              const uint8_t* mask, T* image)
{
    const int64_t na_sub = (na + ia_skip - 1) / ia_skip;
    const T wb = (T)(float)((nb + 1) / 2.0 + (double)(float)offset);
    const T dsd_ds = (T)dsd / (T)ds, dso_ds = (T)dso / (T)ds, so_ds = (T)source_offset / (T)ds;
    std::vector<T> s(na_sub), c(na_sub), xd(nx), yd(ny);
    for (int64_t i = 0; i < na_sub; ++i) { s[i] = (T)sinb[i]; c[i] = (T)cosb[i]; }
    for (int64_t i = 0; i < nx; ++i) xd[i] = (T)xc[i] / (T)ds;
    for (int64_t i = 0; i < ny; ++i) yd[i] = (T)yc[i] / (T)ds;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < nx * ny; ++p) {
        if (!mask[p]) { image[p] = 0; continue; }
        const T x = xd[p % nx], y = yd[p / nx];
        T pixel = 0;
        for (int64_t ja = 0; ja < na_sub; ++ja) {
            const T* q = sino + nb * (ja * ia_skip);
            const T sb = s[ja], cb = c[ja];
            const T d_loop = dso_ds + x * sb - y * cb;                    // :302
            const T r_loop = x * cb + y * sb - so_ds;                     // :303
            T sprime, w2;
            if (is_arc) {
                sprime = dsd_ds * std::atan2(r_loop, d_loop);             // :306
                w2 = (dsd_ds * dsd_ds) / (d_loop * d_loop + r_loop * r_loop);
            } else {
                const T mag = dsd_ds / d_loop;                            // :309
                sprime = mag * r_loop;
                w2 = mag * mag;
            }
            const T bb = sprime + wb;                                     // :314
            const int64_t il = jfloor(bb);
            const int64_t ir = 1 + il;
            if (1 <= il && ir <= nb) {                                    // :328
                const T wr = bb - (T)il;
                const T wl = 1 - wr;
                pixel += (wl * q[il - 1] + wr * q[ir - 1]) * w2;          // :331
            }
        }
        image[p] = pixel;
    }
}

// ---------------------------------------------------------------------------
// cbct_back_fan! / cbct_back_fan_voxel   (src/fbp3/cbct-back.jl:95-213)
// Three options that the reference leaves commented out, off by default:
//   source_offset: the lateral source offset of the commented arc branch,
//   r_loop = x_beta - offset_source / ds (:157, :178), in s' and w2; applied
//   to the flat detector as fbp_back_fan_xy does (src/fbp2/back-fan.jl:303,
//   309-311: s' = mag r_loop);
//   zs (na source z positions, or null): t' = mag (z - zs[ia]) (:188);
//   arc_t_exact: for the arc detector, the t magnification of the ray that
//   meets the detector cylinder, dsd / sqrt(x_beta^2 + y_beta^2), instead of
//   the flat form dsd / y_beta (:180).
// ---------------------------------------------------------------------------
template <class T>
void cbct_back(const T* proj, int64_t ns, int64_t nt, int64_t na, int64_t ia_skip,
               const double* betas, double dsd, double dso, double ds, double dt,
               double offset_s, double offset_t, int is_arc,
               const double* xc, int64_t nx, const double* yc, int64_t ny,
               const double* zc, int64_t nz, const uint8_t* mask, T* image,
               const T* dnext = nullptr, const double* zs = nullptr, int arc_t_exact = 0,
               double source_offset = 0.0)
{
    // dnext (optional, same shape as proj): the stored row-pair differences;
    // the taps of row it+1 are then proj[., it] + dnext[., it] (a checker for
    // data kept as (value, difference) pairs), else proj[., it+1].
    const int64_t na_sub = (na + ia_skip - 1) / ia_skip;
    const T ws = (T)(float)((ns + 1) / 2.0 + (double)(float)offset_s);  // :119
    const T wt = (T)(float)((nt + 1) / 2.0 + (double)(float)offset_t);  // :120
    const T dt_ds = (T)dt / (T)ds, dsd_ds = (T)dsd / (T)ds, dso_ds = (T)dso / (T)ds;
    const T so_ds = (T)source_offset / (T)ds;
    std::vector<T> s(na_sub), c(na_sub), xd(nx), yd(ny), zd(nz), zsd(na_sub, (T)0);
    for (int64_t i = 0; i < na_sub; ++i) {
        const T b = (T)betas[i * ia_skip];
        s[i] = std::sin(b); c[i] = std::cos(b);                           // :126-127
        if (zs) zsd[i] = (T)zs[i * ia_skip] / (T)ds;
    }
    for (int64_t i = 0; i < nx; ++i) xd[i] = (T)xc[i] / (T)ds;            // :129-131
    for (int64_t i = 0; i < ny; ++i) yd[i] = (T)yc[i] / (T)ds;
    for (int64_t i = 0; i < nz; ++i) zd[i] = (T)zc[i] / (T)ds;
    const int64_t nxy = nx * ny;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < nxy * nz; ++p) {
        if (!mask[p]) { image[p] = 0; continue; }
        const T x = xd[p % nx], y = yd[(p / nx) % ny], z = zd[p / nxy];
        T voxel = 0;
        for (int64_t ja = 0; ja < na_sub; ++ja) {
            const T* q = proj + ns * nt * (ja * ia_skip);
            const T sb = s[ja], cb = c[ja];
            const T x_beta = x * cb + y * sb;                             // :174
            const T y_beta = dso_ds - (-x * sb + y * cb);                 // :175
            const T r_loop = source_offset != 0.0 ? x_beta - so_ds : x_beta;   // :178
            T sprime, mag, w2;
            if (is_arc) {
                sprime = dsd_ds * std::atan2(r_loop, y_beta);             // :179
                mag = dsd_ds / y_beta;
                w2 = (dsd_ds * dsd_ds) / (r_loop * r_loop + y_beta * y_beta);
            } else {
                mag = dsd_ds / y_beta;                                    // :183
                sprime = mag * r_loop;
                w2 = mag * mag;
            }
            const T magt = (is_arc && arc_t_exact)
                               ? dsd_ds / std::sqrt(r_loop * r_loop + y_beta * y_beta)
                               : mag;
            const T tprime = magt * (z - zsd[ja]);                        // :188-189
            const T bs = sprime + ws;
            const T bt = tprime / dt_ds + wt;                             // :192
            const int64_t is = jfloor(bs), it = jfloor(bt);
            if (1 <= is && is < ns && 1 <= it && it < nt) {               // :197
                const T wr = bs - (T)is, wl = 1 - wr;
                const T wu = bt - (T)it, wd = 1 - wu;
                const int64_t o0 = (is - 1) + ns * (it - 1);
                const T* r0 = q + o0;
                const T* d0 = dnext ? dnext + ns * nt * (ja * ia_skip) + o0 : nullptr;
                const T u0 = d0 ? r0[0] + d0[0] : r0[ns];
                const T u1 = d0 ? r0[1] + d0[1] : r0[ns + 1];
                const T p1 = wl * r0[0] + wr * r0[1];                     // :205
                const T p2 = wl * u0 + wr * u1;                           // :206
                voxel += w2 * (wd * p1 + wu * p2);                        // :208
            }
        }
        image[p] = voxel;
    }
}

// ---------------------------------------------------------------------------
// fbp_sino_filter (src/fbp/filter.jl:122-147) as an explicit circular
// convolution: zero-pad each column to npad, circularly convolve with
// g = real(ifft(Hk)), keep rows 1..nb, zero rows nb+1..nb+extra.
// g is formed here in double by a direct inverse DFT of the real, even Hk
// (no FFT library).  The float mode rounds the double result once, i.e. it
// is the correctly rounded Float32 answer (FFTW's Float32 error, ~1e-7 of
// the row norm, is far below every tolerance used against it).
// ---------------------------------------------------------------------------
template <class H>
void make_g(const H* Hk, int64_t npad, std::vector<double>& g)
{
    g.assign(npad, 0.0);
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t k = 0; k < npad; ++k) {
        double acc = 0;
        for (int64_t j = 0; j < npad; ++j) {
            const int64_t jk = (j * k) % npad;   // exact phase reduction
            acc += (double)Hk[j] * std::cos(two_pi * (double)jk / (double)npad);
        }
        g[k] = acc / (double)npad;
    }
}

template <class T>
void sino_filter(const T* in, int64_t nb, int64_t ncols, const T* Hk, int64_t npad,
                 int64_t extra, T* out)
{
    std::vector<double> g;
    make_g(Hk, npad, g);
    // gr[j] = g[(j - (nb-1)) mod npad] so out[n] = sum_m gr[n - m + nb - 1] in[m]
    std::vector<double> gr(2 * nb);
    for (int64_t j = 0; j < 2 * nb - 1; ++j) {
        int64_t k = (j - (nb - 1)) % npad;
        if (k < 0) k += npad;
        gr[j] = g[k];
    }
    const int64_t nout = nb + extra;
#pragma omp parallel
    {
        std::vector<double> col(nb);
#pragma omp for schedule(static)
        for (int64_t c = 0; c < ncols; ++c) {
            const T* x = in + nb * c;
            for (int64_t m = 0; m < nb; ++m) col[m] = (double)x[m];
            T* y = out + nout * c;
            for (int64_t n = 0; n < nb; ++n) {
                const double* gp = gr.data() + n + nb - 1;   // gp[-m] = g[(n-m) mod npad]
                double acc = 0;
                for (int64_t m = 0; m < nb; ++m) acc += gp[-m] * col[m];
                y[n] = (T)acc;
            }
            for (int64_t n = nb; n < nout; ++n) y[n] = 0;              // filter.jl:143
        }
    }
}

// ---------------------------------------------------------------------------
// project_bdd / backproject_bdd  (src/sys2/bdd_2d.jl:17-220), double only.
// phantom and image are (nPix, nPix) column-major; sinogram is (nangle, nDet)
// column-major, i.e. sino[proj, det] at proj + nangle*det (bdd_2d.jl:37).
// Interpolations.linear_interpolation(xs, ys, extrapolation_bc=Flat()) is
// restated as piecewise-linear interpolation on ascending knots with the end
// values held outside the knot range.
// ---------------------------------------------------------------------------
double interp_flat(const double* xs, const double* ys, int64_t n, double x)
{
    if (x <= xs[0]) return ys[0];
    if (x >= xs[n - 1]) return ys[n - 1];
    // largest k with xs[k] <= x
    int64_t lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) / 2;
        if (xs[mid] <= x) lo = mid; else hi = mid;
    }
    const double t = (x - xs[lo]) / (xs[lo + 1] - xs[lo]);
    return ys[lo] + t * (ys[lo + 1] - ys[lo]);
}

struct BddView {
    bool axis_x;            // axisCase
    std::vector<double> detm, L;
    double tx, ty;
};

void bdd_view_setup(double DSD, double DS0, double pSize, double dSize, int64_t nDet,
                    double beta, BddView& v)
{
    const int64_t nd1 = nDet + 1;
    v.tx = -DS0 * std::sin(beta);                                     // :45
    v.ty = DS0 * std::cos(beta);
    const double pi = 3.14159265358979323846;
    v.axis_x = !(((beta >= pi / 4) && (beta <= 3 * pi / 4)) ||
                 ((beta >= 5 * pi / 4) && (beta <= 7 * pi / 4)));     // :53-57
    v.detm.resize(nd1);
    const double detY = (-(DSD - DS0) - (dSize / 2));                 // :29
    const double cb = std::cos(beta), sb = std::sin(beta);
    for (int64_t k = 0; k < nd1; ++k) {
        const double detX = (-(double)nDet / 2 + (double)k) * dSize;  // :28
        const double rx = detX * cb - detY * sb;                      // :49
        const double ry = detX * sb + detY * cb;                      // :50
        v.detm[k] = v.axis_x ? v.tx - v.ty * (v.tx - rx) / (v.ty - ry)    // map2x :260
                             : v.ty - v.tx * (v.ty - ry) / (v.tx - rx);   // map2y :249
    }
    v.L.resize(nDet);
    for (int64_t n = 0; n < nDet; ++n) {                               // :71-86
        const double det_mid = (v.detm[n] + v.detm[n + 1]) / 2;
        const double dsz = v.detm[n + 1] - v.detm[n];
        if (v.axis_x) {
            const double th = std::atan(std::fabs(v.tx - det_mid) / std::fabs(v.ty));
            v.L[n] = std::fabs(pSize / (dsz * std::cos(th)));
        } else {
            const double th = std::atan(1.0 / (std::fabs(v.ty - det_mid) / std::fabs(v.tx))); // acot
            v.L[n] = std::fabs(pSize / (dsz * std::sin(th)));
        }
    }
}

// Mapped pixel boundary for "row" r (0-based) and boundary j (0-based) in the
// orientation the reference uses for img[row, :] (bdd_2d.jl:32-35, 60-68).
inline double bdd_pix_boundary(const BddView& v, int64_t nPix, double pSize, int64_t r, int64_t j)
{
    // 0-based boundary grids (ndgrid, transpose, reverse and the -1/2 shift of
    // bdd_2d.jl:32-35):  pixelX[i,j] = (-nPix/2 + j) pSize,
    //                    pixelY[i,j] = (nPix/2 - i - 1/2) pSize.
    if (v.axis_x) {
        // pixm = map2x.(tube, pixelX, pixelY); row r is pixm[r, :]
        const double pX = (-(double)nPix / 2 + (double)j) * pSize;
        const double pY = ((double)nPix / 2 - (double)r - 0.5) * pSize;
        return v.tx - v.ty * (v.tx - pX) / (v.ty - pY);
    } else {
        // pixm = reverse(M', dims=2) with M = map2y.(tube, pixelX, pixelY):
        // pixm[r, j] = M[nPix - j, r]
        const int64_t mi = nPix - j, mj = r;
        const double pX = (-(double)nPix / 2 + (double)mj) * pSize;
        const double pY = ((double)nPix / 2 - (double)mi - 0.5) * pSize;
        return v.ty - v.tx * (v.ty - pY) / (v.tx - pX);
    }
}

} // namespace

extern "C" {

#define ORC_EXPORT __attribute__((visibility("default")))

ORC_EXPORT void orc_back_par_f64(const double* sino, int64_t nb, int64_t na, int64_t ia_skip,
    const double* sinphi, const double* cosphi, double ds, double offset,
    const double* xc, int64_t nx, const double* yc, int64_t ny, const uint8_t* mask, double* image)
{ back_par<double>(sino, nb, na, ia_skip, sinphi, cosphi, ds, offset, xc, nx, yc, ny, mask, image); }

ORC_EXPORT void orc_back_par_f32(const float* sino, int64_t nb, int64_t na, int64_t ia_skip,
    const double* sinphi, const double* cosphi, double ds, double offset,
    const double* xc, int64_t nx, const double* yc, int64_t ny, const uint8_t* mask, float* image)
{ back_par<float>(sino, nb, na, ia_skip, sinphi, cosphi, ds, offset, xc, nx, yc, ny, mask, image); }

ORC_EXPORT void orc_back_fan_f64(const double* sino, int64_t nb, int64_t na, int64_t ia_skip,
    const double* sinb, const double* cosb, double dsd, double dso, double source_offset,
    int is_arc, double ds, double offset,
    const double* xc, int64_t nx, const double* yc, int64_t ny, const uint8_t* mask, double* image)
{ back_fan<double>(sino, nb, na, ia_skip, sinb, cosb, dsd, dso, source_offset, is_arc, ds, offset,
                   xc, nx, yc, ny, mask, image); }

ORC_EXPORT void orc_back_fan_f32(const float* sino, int64_t nb, int64_t na, int64_t ia_skip,
    const double* sinb, const double* cosb, double dsd, double dso, double source_offset,
    int is_arc, double ds, double offset,
    const double* xc, int64_t nx, const double* yc, int64_t ny, const uint8_t* mask, float* image)
{ back_fan<float>(sino, nb, na, ia_skip, sinb, cosb, dsd, dso, source_offset, is_arc, ds, offset,
                  xc, nx, yc, ny, mask, image); }

ORC_EXPORT void orc_cbct_back_f64(const double* proj, int64_t ns, int64_t nt, int64_t na, int64_t ia_skip,
    const double* betas, double dsd, double dso, double ds, double dt,
    double offset_s, double offset_t, int is_arc,
    const double* xc, int64_t nx, const double* yc, int64_t ny, const double* zc, int64_t nz,
    const uint8_t* mask, double* image)
{ cbct_back<double>(proj, ns, nt, na, ia_skip, betas, dsd, dso, ds, dt, offset_s, offset_t, is_arc,
                    xc, nx, yc, ny, zc, nz, mask, image); }

ORC_EXPORT void orc_cbct_back_f32(const float* proj, int64_t ns, int64_t nt, int64_t na, int64_t ia_skip,
    const double* betas, double dsd, double dso, double ds, double dt,
    double offset_s, double offset_t, int is_arc,
    const double* xc, int64_t nx, const double* yc, int64_t ny, const double* zc, int64_t nz,
    const uint8_t* mask, float* image)
{ cbct_back<float>(proj, ns, nt, na, ia_skip, betas, dsd, dso, ds, dt, offset_s, offset_t, is_arc,
                   xc, nx, yc, ny, zc, nz, mask, image); }

ORC_EXPORT void orc_cbct_back_opts_f64(const double* proj, int64_t ns, int64_t nt, int64_t na,
    int64_t ia_skip, const double* betas, double dsd, double dso, double ds, double dt,
    double offset_s, double offset_t, int is_arc,
    const double* xc, int64_t nx, const double* yc, int64_t ny, const double* zc, int64_t nz,
    const uint8_t* mask, const double* zs, int arc_t_exact, double* image)
{ cbct_back<double>(proj, ns, nt, na, ia_skip, betas, dsd, dso, ds, dt, offset_s, offset_t, is_arc,
                    xc, nx, yc, ny, zc, nz, mask, image, nullptr, zs, arc_t_exact); }

ORC_EXPORT void orc_cbct_back_opts_f32(const float* proj, int64_t ns, int64_t nt, int64_t na,
    int64_t ia_skip, const double* betas, double dsd, double dso, double ds, double dt,
    double offset_s, double offset_t, int is_arc,
    const double* xc, int64_t nx, const double* yc, int64_t ny, const double* zc, int64_t nz,
    const uint8_t* mask, const double* zs, int arc_t_exact, float* image)
{ cbct_back<float>(proj, ns, nt, na, ia_skip, betas, dsd, dso, ds, dt, offset_s, offset_t, is_arc,
                   xc, nx, yc, ny, zc, nz, mask, image, nullptr, zs, arc_t_exact); }

ORC_EXPORT void orc_cbct_back_opts2_f64(const double* proj, int64_t ns, int64_t nt, int64_t na,
    int64_t ia_skip, const double* betas, double dsd, double dso, double ds, double dt,
    double offset_s, double offset_t, int is_arc,
    const double* xc, int64_t nx, const double* yc, int64_t ny, const double* zc, int64_t nz,
    const uint8_t* mask, const double* zs, int arc_t_exact, double source_offset, double* image)
{ cbct_back<double>(proj, ns, nt, na, ia_skip, betas, dsd, dso, ds, dt, offset_s, offset_t, is_arc,
                    xc, nx, yc, ny, zc, nz, mask, image, nullptr, zs, arc_t_exact, source_offset); }

ORC_EXPORT void orc_cbct_back_opts2_f32(const float* proj, int64_t ns, int64_t nt, int64_t na,
    int64_t ia_skip, const double* betas, double dsd, double dso, double ds, double dt,
    double offset_s, double offset_t, int is_arc,
    const double* xc, int64_t nx, const double* yc, int64_t ny, const double* zc, int64_t nz,
    const uint8_t* mask, const double* zs, int arc_t_exact, double source_offset, float* image)
{ cbct_back<float>(proj, ns, nt, na, ia_skip, betas, dsd, dso, ds, dt, offset_s, offset_t, is_arc,
                   xc, nx, yc, ny, zc, nz, mask, image, nullptr, zs, arc_t_exact, source_offset); }

ORC_EXPORT void orc_cbct_back_rp_f64(const double* proj, const double* dnext, int64_t ns,
    int64_t nt, int64_t na, const double* betas, double dsd, double dso, double ds, double dt,
    int is_arc, const double* xc, int64_t nx, const double* yc, int64_t ny, const double* zc,
    int64_t nz, const uint8_t* mask, double* image)
{ cbct_back<double>(proj, ns, nt, na, 1, betas, dsd, dso, ds, dt, 0.0, 0.0, is_arc,
                    xc, nx, yc, ny, zc, nz, mask, image, dnext); }

ORC_EXPORT void orc_cbct_back_rp_f32(const float* proj, const float* dnext, int64_t ns,
    int64_t nt, int64_t na, const double* betas, double dsd, double dso, double ds, double dt,
    int is_arc, const double* xc, int64_t nx, const double* yc, int64_t ny, const double* zc,
    int64_t nz, const uint8_t* mask, float* image)
{ cbct_back<float>(proj, ns, nt, na, 1, betas, dsd, dso, ds, dt, 0.0, 0.0, is_arc,
                   xc, nx, yc, ny, zc, nz, mask, image, dnext); }

ORC_EXPORT void orc_filter_kernel(const double* Hk, int64_t npad, double* g)
{
    std::vector<double> gv;
    make_g(Hk, npad, gv);
    std::memcpy(g, gv.data(), sizeof(double) * npad);
}

ORC_EXPORT void orc_sino_filter_f64(const double* in, int64_t nb, int64_t ncols, const double* Hk,
    int64_t npad, int64_t extra, double* out)
{ sino_filter<double>(in, nb, ncols, Hk, npad, extra, out); }

ORC_EXPORT void orc_sino_filter_f32(const float* in, int64_t nb, int64_t ncols, const float* Hk,
    int64_t npad, int64_t extra, float* out)
{ sino_filter<float>(in, nb, ncols, Hk, npad, extra, out); }

// project_bdd: phantom (nPix,nPix) col-major -> sino (nangle, nDet) col-major.
ORC_EXPORT void orc_project_bdd(const double* phantom, int64_t nPix, int64_t nDet,
    const double* angle, int64_t nangle, double DSD, double DS0, double pSize, double dSize,
    double* sino)
{
#pragma omp parallel
    {
        BddView v;
        std::vector<double> rowm(nPix + 1), Ppj(nPix + 1), acc(nDet);
#pragma omp for schedule(dynamic)
        for (int64_t ip = 0; ip < nangle; ++ip) {
            bdd_view_setup(DSD, DS0, pSize, dSize, nDet, angle[ip], v);
            std::fill(acc.begin(), acc.end(), 0.0);
            for (int64_t r = 0; r < nPix; ++r) {
                for (int64_t j = 0; j <= nPix; ++j) rowm[j] = bdd_pix_boundary(v, nPix, pSize, r, j);
                // img[row, :] : axis_x -> phantom[r, j] ; else reverse(phantom', dims=2)[r, j]
                //                                           = phantom[nPix-1-j, r]
                Ppj[0] = 0;
                double P = 0;
                for (int64_t j = 0; j < nPix; ++j) {                          // integrate1D :229
                    const double val = v.axis_x ? phantom[r + nPix * j]
                                                : phantom[(nPix - 1 - j) + nPix * r];
                    P += val * (rowm[j + 1] - rowm[j]);
                    Ppj[j + 1] = P;
                }
                double prev = interp_flat(rowm.data(), Ppj.data(), nPix + 1, v.detm[0]);
                for (int64_t k = 0; k < nDet; ++k) {
                    const double cur = interp_flat(rowm.data(), Ppj.data(), nPix + 1, v.detm[k + 1]);
                    acc[k] += std::fabs(cur - prev);                          // :101
                    prev = cur;
                }
            }
            for (int64_t k = 0; k < nDet; ++k) sino[ip + nangle * k] = acc[k] * v.L[k];   // :104
        }
    }
}

// backproject_bdd: sino (nangle, nDet) col-major -> image (nPix,nPix) col-major.
// ref_alias != 0 reproduces the reference exactly: bdd_2d.jl:139-140 binds
// reconImgTmp to the same array as the initial reconImg, so the first view's
// raw per-row buffer is added once more (in the map-to-x orientation) on top
// of the regular sum.  ref_alias == 0 gives the plain sum over views.
ORC_EXPORT void orc_backproject_bdd(const double* sino, int64_t nPix, int64_t nDet,
    const double* angle, int64_t nangle, double DSD, double DS0, double pSize, double dSize,
    int ref_alias, double* image)
{
    std::memset(image, 0, sizeof(double) * nPix * nPix);
    // views are accumulated serially in the reference; per-view images are
    // summed here in view order for a deterministic result.
    std::vector<double> tmp((size_t)nangle * nPix * nPix);
    std::vector<double> first_raw(ref_alias ? (size_t)nPix * nPix : 0);
#pragma omp parallel
    {
        BddView v;
        std::vector<double> rowm(nPix + 1), Ppj(nDet + 1), xs(nDet + 1), ys(nDet + 1);
#pragma omp for schedule(dynamic)
        for (int64_t ip = 0; ip < nangle; ++ip) {
            bdd_view_setup(DSD, DS0, pSize, dSize, nDet, angle[ip], v);
            double* im = tmp.data() + (size_t)ip * nPix * nPix;
            // Ppj = integrate1D(sino[proj,:] .* L, detSize)                  :196
            Ppj[0] = 0;
            double P = 0;
            for (int64_t k = 0; k < nDet; ++k) {
                P += sino[ip + nangle * k] * v.L[k] * (v.detm[k + 1] - v.detm[k]);
                Ppj[k + 1] = P;
            }
            const bool desc = v.detm[0] > v.detm[1];                           // :198
            for (int64_t k = 0; k <= nDet; ++k) {
                xs[k] = desc ? v.detm[nDet - k] : v.detm[k];
                ys[k] = desc ? Ppj[nDet - k] : Ppj[k];
            }
            for (int64_t r = 0; r < nPix; ++r) {
                for (int64_t j = 0; j <= nPix; ++j) rowm[j] = bdd_pix_boundary(v, nPix, pSize, r, j);
                double prev = interp_flat(xs.data(), ys.data(), nDet + 1, rowm[0]);
                for (int64_t j = 0; j < nPix; ++j) {
                    const double cur = interp_flat(xs.data(), ys.data(), nDet + 1, rowm[j + 1]);
                    const double val = std::fabs(cur - prev);                  // :205
                    prev = cur;
                    // reconImgTmp[r, j]; axis_x: added as is; else reverse(tmp, dims=2)'
                    //   -> image[a, b] += tmp[b, nPix-1-a]  i.e. tmp[r, j] -> image[nPix-1-j, r]
                    if (v.axis_x) im[r + nPix * j] = val;
                    else im[(nPix - 1 - j) + nPix * r] = val;
                    if (ref_alias && ip == 0) first_raw[r + nPix * j] = val;
                }
            }
        }
    }
    for (int64_t ip = 0; ip < nangle; ++ip) {
        const double* im = tmp.data() + (size_t)ip * nPix * nPix;
        for (int64_t p = 0; p < nPix * nPix; ++p) image[p] += im[p];
        if (ref_alias && ip == 0)
            for (int64_t p = 0; p < nPix * nPix; ++p) image[p] += first_raw[p];
    }
}

} // extern "C"
