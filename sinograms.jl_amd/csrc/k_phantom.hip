// Analytic line integrals of ellipse and ellipsoid phantoms along the rays
// of a geometry, in FP64 (synthetic data for tests, the bench and the
// learned correction's training set).  For gfx950.
//
// The reference makes such data with ImagePhantoms' radon(rays(rg), ob) on
// the CPU; the ray conventions are Sinograms.jl's:
//   parallel  (r, phi) = (s, beta)                          src/geom/sino-geom.jl:99-104
//   fan       (r, phi) = (dso sin g, g + beta), g = s/dsd (arc) or atan(s/dsd)
//                                                           src/geom/sino-geom.jl:118-132
//   cone      source (-dso sin b, dso cos b, zs) to the detector point
//             arc (dsd sin g, dso - dsd cos g, t), g = s / dsd, or flat
//             (s, -dod, t), rotated by b and raised by zs  src/fbp3/cb_arc_to_par.jl:23-43
// (the cone form is the true source-to-detector line; it equals rays() of
// the arc detector; see the ray-convention tests).
//
// One thread per ray, the phantom in LDS, the sum over its components in
// FP64 (78.6 TFLOP/s of FP64 on MI355X: a ray costs ~40 FP64 operations per
// component, so these kernels are compute-bound and far from the HBM roof).

#include "sgx_internal.h"
#include "sgx.h"

namespace sgx {
namespace {

constexpr int kMaxComp = 512;          // phantom components per launch (LDS)

// Ellipse rows (cx, cy, a, b, theta, rho) -> (cx, cy, a, b, cos theta,
// sin theta, 2 rho a b); ellipsoid rows (cx, cy, cz, a, b, c, rho) ->
// (cx, cy, cz, 1/a, 1/b, 1/c, rho).
__global__ void radon_ellipses_kernel(int kind, const double* __restrict__ s, int nb,
                                      const double* __restrict__ betas, int na, double dsd,
                                      double dso, const double* __restrict__ comp, int nc,
                                      double* __restrict__ out)
{
    __shared__ double e[kMaxComp * 7];
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        const double* r = comp + 6 * i;
        double* d = e + 7 * i;
        d[0] = r[0]; d[1] = r[1]; d[2] = r[2]; d[3] = r[3];
        d[4] = cos(r[4]); d[5] = sin(r[4]);
        d[6] = 2.0 * r[5] * r[2] * r[3];
    }
    __syncthreads();
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)nb * na) return;
    const int ib = (int)(i % nb), ia = (int)(i / nb);
    double r = s[ib], phi = betas[ia];
    if (kind != SGX_SINO_PAR) {
        const double g = kind == SGX_SINO_FAN_ARC ? s[ib] / dsd : atan(s[ib] / dsd);
        r = dso * sin(g);
        phi += g;
    }
    double sp, cp;
    sincos(phi, &sp, &cp);
    double acc = 0.0;
    for (int k = 0; k < nc; ++k) {
        const double* d = e + 7 * k;
        const double tau = r - (d[0] * cp + d[1] * sp);
        const double cd = cp * d[4] + sp * d[5];           // cos(phi - theta)
        const double sd = sp * d[4] - cp * d[5];           // sin(phi - theta)
        const double s2 = d[2] * d[2] * cd * cd + d[3] * d[3] * sd * sd;
        const double in = s2 - tau * tau;
        if (in > 0.0) acc += d[6] * sqrt(in) / s2;
    }
    out[i] = acc;
}

__global__ void radon_ellipsoids_kernel(const double* __restrict__ lx,
                                        const double* __restrict__ ly, int ns,
                                        const double* __restrict__ t, int nt,
                                        const double* __restrict__ betas,
                                        const double* __restrict__ zs, int na, double dso,
                                        const double* __restrict__ comp, int nc,
                                        double* __restrict__ out)
{
    __shared__ double e[kMaxComp * 7];
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        const double* r = comp + 7 * i;
        double* d = e + 7 * i;
        d[0] = r[0]; d[1] = r[1]; d[2] = r[2];
        d[3] = 1.0 / r[3]; d[4] = 1.0 / r[4]; d[5] = 1.0 / r[5];
        d[6] = r[6];
    }
    __syncthreads();
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nst = (int64_t)ns * nt;
    if (i >= nst * na) return;
    const int is = (int)(i % ns), it = (int)((i / ns) % nt), ia = (int)(i / nst);
    double sb, cb;
    sincos(betas[ia], &sb, &cb);
    const double pz = zs ? zs[ia] : 0.0;
    const double px = -dso * sb, py = dso * cb;            // source
    const double dx = (lx[is] * cb - ly[is] * sb) - px;     // to the detector point
    const double dy = (lx[is] * sb + ly[is] * cb) - py;
    const double dz = t[it];
    const double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    const double ux = dx * inv, uy = dy * inv, uz = dz * inv;
    double acc = 0.0;
    for (int k = 0; k < nc; ++k) {
        const double* d = e + 7 * k;
        const double qx = (px - d[0]) * d[3], qy = (py - d[1]) * d[4], qz = (pz - d[2]) * d[5];
        const double vx = ux * d[3], vy = uy * d[4], vz = uz * d[5];
        const double A = vx * vx + vy * vy + vz * vz;
        const double B = 2.0 * (qx * vx + qy * vy + qz * vz);
        const double C = qx * qx + qy * qy + qz * qz - 1.0;
        const double disc = B * B - 4.0 * A * C;
        if (disc > 0.0) acc += d[6] * sqrt(disc) / A;      // chord length times density
    }
    out[i] = acc;
}

}  // namespace

void launch_radon_ellipses(int kind, const double* s, int nb, const double* betas, int na,
                           double dsd, double dso, const double* comp, int nc, double* out,
                           hipStream_t st)
{
    const int64_t n = (int64_t)nb * na;
    if (n <= 0) return;
    if (nc > kMaxComp) throw Error(SGX_ERR_UNSUPPORTED, "radon: at most 512 ellipses");
    hipLaunchKernelGGL(radon_ellipses_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       st, kind, s, nb, betas, na, dsd, dso, comp, nc, out);
    SGX_HIP(hipGetLastError());
}

void launch_radon_ellipsoids(const double* lx, const double* ly, int ns, const double* t,
                             int nt, const double* betas, const double* zs, int na, double dso,
                             const double* comp, int nc, double* out, hipStream_t st)
{
    const int64_t n = (int64_t)ns * nt * na;
    if (n <= 0) return;
    if (nc > kMaxComp) throw Error(SGX_ERR_UNSUPPORTED, "radon: at most 512 ellipsoids");
    if ((n + 255) / 256 >= (int64_t(1) << 31))
        throw Error(SGX_ERR_UNSUPPORTED, "radon: too many rays for one launch");
    hipLaunchKernelGGL(radon_ellipsoids_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                       0, st, lx, ly, ns, t, nt, betas, zs, na, dso, comp, nc, out);
    SGX_HIP(hipGetLastError());
}

}  // namespace sgx
