/*
 * sgx.h - C ABI of libsgx, the MI355X (gfx950) HIP library behind the
 * Sinograms.jl reconstruction hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns
 * an int status (SGX_OK = 0, negative on error) and records a message that
 * sgx_last_error() returns.  Arrays follow Julia's column-major layout:
 *   sino[ib, ia]        at ib + nb*ia
 *   proj[is, it, ia]    at is + ns*(it + nt*ia)
 *   image[ix, iy(, iz)] with x fastest.
 *
 * Pointer kinds.  Small per-view / per-axis arrays (angles, sin/cos tables,
 * pixel-centre axes, filter response) are always host memory and are copied
 * by the call.  The large arrays (sinogram / projections, image, mask) are
 * device memory unless SGX_HOST is set in flags, in which case the library
 * stages them through device buffers and copies the result back.
 *
 * Synchronisation.  Calls run on the context's stream and return after it
 * has finished, matching the reference's synchronous semantics, unless
 * SGX_ASYNC is set (device pointers only).
 *
 * Threading.  Contexts are independent; one context must not be used from
 * two threads at once.
 */
#ifndef SGX_H
#define SGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sgx_ctx sgx_ctx;
typedef struct sgx_plan sgx_plan;

/* status codes */
enum {
    SGX_OK = 0,
    SGX_ERR_ARG = -1,         /* null pointer, negative size, bad enum           */
    SGX_ERR_SIZE = -2,        /* reference throw("size mismatch"/"sino size")    */
    SGX_ERR_VALUE = -3,       /* reference error(...), e.g. nb + extra > npad     */
    SGX_ERR_DEVICE = -4,      /* HIP runtime failure                              */
    SGX_ERR_UNSUPPORTED = -5  /* option not supported by this build               */
};

/* flags */
#define SGX_HOST        0x1u   /* large arrays are host memory                    */
#define SGX_ASYNC       0x2u   /* return without synchronising (device ptrs only) */
#define SGX_Q_BF16      0x10u  /* filtered data held in bfloat16 (device ptrs)    */

/* geometry kinds (the reference's SinoPar, SinoFanArc, SinoFanFlat,
 * CtFanArc, CtFanFlat, SinoMoj; src/geom/type2.jl:95-159, type3.jl:145-184).
 * SGX_SINO_MOJ is the mojette geometry: ds is the pixel size d, and view
 * beta samples the radial axis at d max(|cos beta|, |sin beta|) (_d_moj,
 * src/geom/common.jl:128); its plans are the :mojette FBP (plan_fbp's `how`,
 * src/fbp2/plan2.jl:95, 129-130). */
enum {
    SGX_SINO_PAR = 0,
    SGX_SINO_FAN_ARC = 1,
    SGX_SINO_FAN_FLAT = 2,
    SGX_CT_FAN_ARC = 3,
    SGX_CT_FAN_FLAT = 4,
    SGX_SINO_MOJ = 5
};

/* precision forms of a plan */
enum {
    SGX_PREC_FP32 = 0,        /* FP32 filter (FP32 MFMA), FP32 filtered storage  */
    SGX_PREC_BF16 = 1,        /* split-BF16x3 MFMA filter, BF16 filtered storage  */
    SGX_PREC_F16S = 2         /* split-FP16x3 MFMA filter (power-of-two scaled
                                 operands), FP16 filtered storage scaled per view
                                 by a power of two (cone-beam plans)             */
};

/* ---- context ------------------------------------------------------------ */

int sgx_version(void);
int sgx_ctx_create(int device, sgx_ctx** out);
void sgx_ctx_destroy(sgx_ctx* ctx);
const char* sgx_last_error(const sgx_ctx* ctx);
/* Run subsequent calls on an existing HIP stream (hipStream_t, may be NULL for
 * the default stream).  The context never destroys a stream it did not make. */
int sgx_ctx_set_stream(sgx_ctx* ctx, void* hip_stream);
int sgx_ctx_synchronize(sgx_ctx* ctx);

/* Kernel choices of a context (no counterpart in the reference; the results
 * agree to FP32 rounding whatever the choice).
 *   SGX_OPT_CBCT_KERNEL: 0 = LDS-staged cone-beam backprojector (default),
 *                        1 = gather backprojector,
 *                        2 = LDS-staged, with scaled-FP16 plans interpolating
 *                            in packed FP16 arithmetic (otherwise as 0;
 *                            about twice the FP16 form's rounding error),
 *                        3 = LDS-staged with blocks 8 voxels deep (twice the
 *                            blocks: for volumes too small to fill the GPU;
 *                            the mirror form, where it applies, takes
 *                            precedence),
 *                        4 = as 3, with detector windows of up to 64 bins x
 *                            32 rows (coarse voxels, whose blocks span more
 *                            of the detector than the others' windows hold).
 *   SGX_OPT_CBCT_MIRROR: the staged backprojector's mirror form, which pairs
 *                        each voxel z > 0 with its image at -z where the
 *                        geometry is symmetric (z axis symmetric about 0, no
 *                        per-view source z, offset_t = 0, nz a multiple of
 *                        16), one detector position serving both:
 *                        1 = for scaled-FP16 plans (default; the autotuner
 *                            times it for the other forms),
 *                        2 = wherever the geometry allows, 0 = never.
 *   SGX_OPT_VIEW_CHUNKS: 1 to 8 (default 1): a cone-beam plan execute with
 *                        the staged backprojector (FP32 interpolation)
 *                        filters and backprojects its views in this many
 *                        chunks of at least 32 views; the filters of later
 *                        chunks run on a stream of the plan beside the
 *                        backprojections of earlier ones (in chunk order on
 *                        the caller's stream under stream capture, with the
 *                        same result).  The chunks' sums are added in chunk
 *                        order; 1 = one filter launch, then one
 *                        backprojection over all views.
 * Plans take the context's choices when they are created (and an autotuned
 * kernel after sgx_plan_autotune). */
enum { SGX_OPT_CBCT_KERNEL = 1, SGX_OPT_CBCT_MIRROR = 2, SGX_OPT_VIEW_CHUNKS = 3 };
int sgx_ctx_set_option(sgx_ctx* ctx, int option, int64_t value);
int sgx_ctx_get_option(const sgx_ctx* ctx, int option, int64_t* value);

/* ---- backprojectors ------------------------------------------------------ */

/* fbp_back_par!(image, sino, sinphi, cosphi, ds, offset, xc, yc, mask; ia_skip)
 * Replaces src/fbp2/back-par.jl:206-253 (+ fbp_back_par_xy :276-315).
 * sinphi, cosphi: length ceil(na / ia_skip) (already subset, as in the
 * reference).  mask may be NULL (all pixels).  Pixels outside the mask are 0.
 * Returns SGX_ERR_SIZE for the reference's throw("size mismatch"). */
int sgx_back_par(sgx_ctx* ctx, const float* sino, int64_t nb, int64_t na, int64_t ia_skip,
                 const double* sinphi, const double* cosphi, double ds, double offset,
                 const double* xc, int64_t nx, const double* yc, int64_t ny,
                 const uint8_t* mask, float* image, uint32_t flags);

/* fbp_back_fan!(image, sino, sinb, cosb, dsd, dso, source_offset, is_arc,
 *               ds, offset, xc, yc, mask; ia_skip)
 * Replaces src/fbp2/back-fan.jl:215-254 (+ fbp_back_fan_xy :280-336). */
int sgx_back_fan(sgx_ctx* ctx, const float* sino, int64_t nb, int64_t na, int64_t ia_skip,
                 const double* sinb, const double* cosb, double dsd, double dso,
                 double source_offset, int is_arc, double ds, double offset,
                 const double* xc, int64_t nx, const double* yc, int64_t ny,
                 const uint8_t* mask, float* image, uint32_t flags);

/* cbct_back_fan!(image, proj, betas, dsd, dso, ds, dt, offset_s, offset_t,
 *                is_arc, xc, yc, zc, mask; ia_skip)
 * Replaces src/fbp3/cbct-back.jl:95-144 (+ cbct_back_fan_voxel :150-213).
 * betas has length na (the full view list; ia_skip selects from it, as in
 * the reference).  Per-view sin/cos are formed on the device in FP64.
 * With SGX_Q_BF16, proj points to bfloat16 data. */
int sgx_cbct_back(sgx_ctx* ctx, const void* proj, int64_t ns, int64_t nt, int64_t na,
                  int64_t ia_skip, const double* betas, double dsd, double dso,
                  double ds, double dt, double offset_s, double offset_t, int is_arc,
                  const double* xc, int64_t nx, const double* yc, int64_t ny,
                  const double* zc, int64_t nz, const uint8_t* mask, float* image,
                  uint32_t flags);

/* Cone-beam options that the reference leaves commented out in cbct_back
 * (src/fbp3/cbct-back.jl:36-59 and 188: source_zs; :180: the arc detector's
 * t magnification; :36, :53, :157, :178: the source offset).  All zero (or a
 * NULL pointer to this struct) is the reference's geometry.
 *   source_zs      na source z positions (host, units of xc / ds, e.g. the
 *                  helical _source_zs of src/geom/ct-geom.jl:134-140), or
 *                  NULL for z = 0: t' = mag (z - source_zs[ia]);
 *   arc_t_exact    1: for the arc detector the t magnification of the ray
 *                  that meets the detector cylinder,
 *                  dsd / sqrt(r_loop^2 + y_beta^2), instead of the flat
 *                  detector's dsd / y_beta;
 *   source_offset  the lateral source offset (units of ds), rg.source_offset:
 *                  r_loop = x_beta - source_offset / ds in place of x_beta in
 *                  s' and w2, for both detectors, as fbp_back_fan_xy applies
 *                  it (src/fbp2/back-fan.jl:303-311). */
typedef struct {
    const double* source_zs;
    int32_t arc_t_exact;
    double source_offset;
} sgx_cbct_opts;

/* sgx_cbct_back with the options above; no counterpart in the reference (it
 * extends cbct_back, src/fbp3/cbct-back.jl:28-45, with what that file keeps
 * commented out).  source_zs is indexed like betas (ia_skip applies). */
int sgx_cbct_back_opts(sgx_ctx* ctx, const void* proj, int64_t ns, int64_t nt, int64_t na,
                       int64_t ia_skip, const double* betas, double dsd, double dso,
                       double ds, double dt, double offset_s, double offset_t, int is_arc,
                       const double* xc, int64_t nx, const double* yc, int64_t ny,
                       const double* zc, int64_t nz, const uint8_t* mask, float* image,
                       uint32_t flags, const sgx_cbct_opts* opts);

/* sgx_cbct_back without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Same computation and kernels as
 * sgx_cbct_back, with the call's geometry tables and the staged
 * backprojector's window table (16 bytes per 16 x 16 x 16 voxel block and
 * view) in a fixed 16 MiB device scratch array that the library holds on
 * every GPU it runs on until the process ends.  Device FP32 data only
 * (flags: 0 or SGX_ASYNC), ns a multiple of 4 and proj 16-byte aligned;
 * a call that needs more than the scratch returns SGX_ERR_UNSUPPORTED before
 * any launch.  Every call on one GPU shares that GPU's scratch: calls that
 * overlap on one GPU (other streams, other threads or contexts) race. */
int sgx_cbct_back_noalloc(sgx_ctx* ctx, const float* proj, int64_t ns, int64_t nt, int64_t na,
                          int64_t ia_skip, const double* betas, double dsd, double dso,
                          double ds, double dt, double offset_s, double offset_t, int is_arc,
                          const double* xc, int64_t nx, const double* yc, int64_t ny,
                          const double* zc, int64_t nz, const uint8_t* mask, float* image,
                          uint32_t flags);

/* ---- row filter ------------------------------------------------------------ */

/* fbp_sino_filter(sino, filter; extra)
 * Replaces src/fbp/filter.jl:122-147.  in is (nb, ncols), ncols = product of
 * the trailing dimensions; out is (nb + extra, ncols) with the extra rows 0.
 * Computed as the Toeplitz contraction out[n,c] = sum_m g[(n-m) mod npad] in[m,c]
 * with g = real(ifft(Hk)), on MFMA, or for SGX_PREC_FP32 at npad 512, 1024
 * or 2048 as the zero-padded FFT product itself.  prec selects SGX_PREC_FP32
 * (FP32 output) or SGX_PREC_BF16 (split-BF16x3 products, FP32 accumulation,
 * FP32 output).  The context keeps the device form of the last 8 filters
 * (by the contents of Hk, nb and prec): a repeated call with the same Hk does
 * no host transform and no upload.
 * Returns SGX_ERR_VALUE for the reference's error("nb + extra > npad"), and
 * SGX_ERR_UNSUPPORTED for nb > 8192 (the reference has no such limit; the
 * band of g lives in one workgroup's LDS).  Plans share this limit. */
int sgx_sino_filter(sgx_ctx* ctx, const float* in, int64_t nb, int64_t ncols,
                    const float* Hk, int64_t npad, int64_t extra, int prec,
                    float* out, uint32_t flags);

/* parker_weight_fan_short(nb, na, orbit, orbit_short, ar, gam, gamma_max)
 * Replaces src/fbp/parker.jl:62-108.  Writes the (nb, na) Float32 weights,
 * evaluated in FP64 on the device. */
int sgx_parker_fan_short(sgx_ctx* ctx, int64_t nb, int64_t na, const double* ar,
                         const double* gam, double gamma_max, float* out, uint32_t flags);

/* ---- distance-driven projector --------------------------------------------- */

/* project_bdd(phantom, geo)
 * Replaces src/sys2/bdd_2d.jl:17-110 (+ integrate1D :229-241, map2x/map2y
 * :249-263).  phantom is (nPix, nPix) column-major, Float64, or Float32 with
 * SGX_IN_F32; sino is (nangle, nDet) column-major Float64 (the reference's
 * output layout and type).  angle: nangle radians (host).  Computed in FP64. */
#define SGX_IN_F32      0x20u  /* project_bdd: phantom is Float32             */
int sgx_project_bdd(sgx_ctx* ctx, const void* phantom, int64_t nPix, int64_t nDet,
                    const double* angle, int64_t nangle, double DSD, double DS0, double pSize,
                    double dSize, double* sino, uint32_t flags);

/* backproject_bdd(sinogram, geo)
 * Replaces src/sys2/bdd_2d.jl:119-220.  sino is (nangle, nDet) Float64
 * column-major, image (nPix, nPix) Float64.  ref_alias = 1 reproduces the
 * reference exactly: bdd_2d.jl:139-140 binds reconImgTmp to the initial
 * reconImg, so the first view's per-row buffer is added once more; 0 gives
 * the plain sum over views (the exact adjoint of sgx_project_bdd).  Views
 * are summed with FP64 atomics (order not fixed). */
int sgx_backproject_bdd(sgx_ctx* ctx, const double* sino, int64_t nPix, int64_t nDet,
                        const double* angle, int64_t nangle, double DSD, double DS0,
                        double pSize, double dSize, int ref_alias, double* image,
                        uint32_t flags);

/* ---- analytic phantom projections (test and training data) -------------- */

/* Line integrals in Float64 of ellipses, rows (cx, cy, a, b, theta, rho)
 * (nell <= 512), along the rays of a 2-D geometry: kind SGX_SINO_PAR,
 * (r, phi) = (s[ib], betas[ia]) (src/geom/sino-geom.jl:99-104), or
 * SGX_SINO_FAN_ARC / SGX_SINO_FAN_FLAT, (r, phi) = (dso sin g, g + beta) with
 * g = s / dsd or atan(s / dsd) (sino-geom.jl:118-132).  s (nb), betas (na),
 * ell: host arrays; out: (nb, na) column-major, host memory with SGX_HOST,
 * else device memory.  No counterpart in the reference, whose tests make
 * these with ImagePhantoms' radon(rays(rg), ob) on the CPU. */
int sgx_radon_ellipses(sgx_ctx* ctx, int kind, const double* s, int64_t nb,
                       const double* betas, int64_t na, double dsd, double dso,
                       const double* ell, int64_t nell, double* out, uint32_t flags);

/* Cone-beam line integrals in Float64 of axis-aligned ellipsoids, rows
 * (cx, cy, cz, a, b, c, rho) (nell <= 512), along the lines from the source
 * (-dso sin b, dso cos b, zs) to the detector point of (s, t) (kind
 * SGX_CT_FAN_ARC: (dsd sin g, dso - dsd cos g, t), g = s / dsd; SGX_CT_FAN_FLAT:
 * (s, dso - dsd, t)), rotated by b and raised by zs; the arc case is the
 * line of cb_arc_to_par (src/fbp3/cb_arc_to_par.jl:23-43).  s (ns), t (nt),
 * betas (na), source_zs (na, or NULL for 0), ell: host arrays; out: (ns, nt, na)
 * column-major, host or device memory as above.  No counterpart in the
 * reference (see sgx_radon_ellipses). */
int sgx_radon_ellipsoids(sgx_ctx* ctx, int kind, const double* s, int64_t ns,
                         const double* t, int64_t nt, const double* betas,
                         const double* source_zs, int64_t na, double dsd, double dso,
                         const double* ell, int64_t nell, double* out, uint32_t flags);

/* ---- plans: the fused whole path ----------------------------------------- */

/* Geometry and weighting of a plan (plan_fbp, src/fbp2/plan2.jl:106-138 and
 * src/fbp3/plan3.jl:60-80).  Angles are the caller's _ar(rg) in radians. */
typedef struct {
    int32_t kind;          /* SGX_SINO_* or SGX_CT_*                              */
    int32_t parker;        /* 1: short-scan Parker weights (parker.jl:62-108)     */
    int64_t ns;            /* detector bins: nb (2-D) or ns (3-D)                 */
    int64_t nt;            /* detector rows (1 for 2-D)                           */
    int64_t na;            /* views                                               */
    double ds;             /* bin spacing: rg.d (2-D) or rg.ds (3-D)              */
    double dt;             /* row spacing (3-D; ignored for 2-D)                  */
    double offset_s;       /* rg.offset or rg.offset_s                            */
    double offset_t;       /* rg.offset_t (3-D)                                   */
    double dsd, dso;       /* fan / cone distances (ignored for SGX_SINO_PAR)     */
    double source_offset;  /* 2-D fan only (cbct_back ignores it, as reference)   */
    double gamma_max;      /* _gamma_max(rg), for Parker                          */
    double angle_weight;   /* _angle_weights(_ar(rg)) (d-angle.jl:209-220)        */
} sgx_geom;

/* Create a plan.  betas: na angles (radians).  Image grid: axes xc, yc (, zc
 * for 3-D; pass nz = 1 and zc = NULL for 2-D) and an optional host mask.
 * A SGX_SINO_MOJ plan filters each view with Hk / max(|cos|, |sin|) (the
 * response for the view's radial spacing, with Hk that of spacing d) and
 * backprojects with the view's spacing; Hk is fbp_filter for spacing d. 
 * Hk: the npad filter response of fbp_filter (host, Float32).  prec:
 * SGX_PREC_FP32 or SGX_PREC_BF16.  The plan owns device copies of all of
 * these plus the filtered-data buffer. */
int sgx_plan_create(sgx_ctx* ctx, const sgx_geom* geom, const double* betas,
                    const double* xc, int64_t nx, const double* yc, int64_t ny,
                    const double* zc, int64_t nz, const uint8_t* mask,
                    const float* Hk, int64_t npad, int prec, sgx_plan** out);
void sgx_plan_destroy(sgx_plan* plan);

/* The cone-beam options above for a cone-beam plan (the plan's backprojection
 * of every later execute); NULL restores the reference's geometry.  No
 * counterpart in the reference (see sgx_cbct_back_opts). */
int sgx_plan_set_cbct_opts(sgx_plan* plan, const sgx_cbct_opts* opts);

/* Per-view weight (na Float32, host) multiplied into the filter's operand,
 * e.g. parker_weight_par for a parallel-beam orbit that is not a multiple of
 * 180 degrees (src/fbp/parker.jl:26-55).  NULL removes it. */
int sgx_plan_set_view_weight(sgx_plan* plan, const float* wview);

/* fbp(plan, sino) / fdk(plan, proj): weight -> filter -> backproject.
 * Replaces src/fbp2/fbp.jl:24-86 and src/fbp3/fdk.jl:65-81.  sino is
 * (ns, nt, na) Float32 (nt = 1 for 2-D); image is (nx, ny, nz). */
int sgx_plan_execute(sgx_plan* plan, const float* sino, float* image, uint32_t flags);

/* The filtered, weighted data of the last execute (device pointer):
 * column-major (pitch, nt, na) (nt = 1 for 2-D plans) of the plan's storage
 * type (Float32, bfloat16 or scaled FP16 per the precision), where pitch =
 * sgx_plan_pitch (ns for 2-D plans, ns rounded up to a multiple of 4 for
 * cone-beam plans; the pad bins are 0). */
const void* sgx_plan_filtered(const sgx_plan* plan);

/* Elements per detector row of the stored filtered data. */
int64_t sgx_plan_pitch(const sgx_plan* plan);

/* Copy the stored filtered data of the last execute to host memory: `host`
 * receives ns*nt*na elements of the plan's storage type (4 bytes for
 * SGX_PREC_FP32, 2 bytes otherwise; column-major, pad bins dropped) and
 * `view_scales` (na floats, may be NULL) the per-view factor
 * that multiplies the stored values (1 except for SGX_PREC_F16S). */
int sgx_plan_read_filtered(sgx_plan* plan, void* host, float* view_scales);

/* Size in bytes of the stored filtered data (the layout of sgx_plan_filtered,
 * pad bins included). */
int64_t sgx_plan_stored_bytes(const sgx_plan* plan);

/* Copy the stored filtered data as-is (`bytes` must equal
 * sgx_plan_stored_bytes) to host memory. */
int sgx_plan_read_stored(sgx_plan* plan, void* host, int64_t bytes);

/* The (ns, nt, na) view weights the plan applies, as the reference's
 * plan.view_weight broadcast to full size (host Float32 out). */
int sgx_plan_view_weight(sgx_plan* plan, float* out);

/* Device time (ms) of the backprojection and filter launches of the last
 * execute, measured with HIP events on the plan's stream. */
int sgx_plan_last_times(const sgx_plan* plan, float* filter_ms, float* back_ms);

/* ---- autotuning and batches (fbp(plan, aa) serial slice loop,
 *      src/fbp2/fbp.jl:46-68, as one graph) --------------------------------- */

/* Time the plan's candidate kernel shapes on (sino, image) (device pointers;
 * `reps` timed executes each) and keep the fastest in the plan.  The choice
 * is cached in the context per size bucket (kind, precision and the
 * power-of-two ceilings of nx, ny, nz, ns, nt, na); with `use_cache` a plan
 * whose bucket is cached takes it without timing.  `image` holds the chosen
 * shape's result on return.  Candidates: cone-beam, the staged
 * backprojector with blocks 16 or 8 voxels deep (the latter also with wide
 * windows) and (symmetric geometries) its mirror form, the gather backprojector, and for scaled-FP16 plans the
 * packed-FP16 interpolation; 2-D, view-grouping targets {1024, 2048, 4096}
 * blocks x LDS view chunks {16, 32, 64}. */
int sgx_plan_autotune(sgx_plan* plan, const float* sino, float* image, int reps,
                      int use_cache);
/* The plan's kernel shape: values[0] = cone-beam kernel (0 staged, 1
 * gather, 2 staged with packed-FP16 interpolation, 3 the staged kernel's
 * mirror form, 4 the staged kernel with blocks 8 voxels deep, 5 the same
 * with wide windows), values[1] =
 * 2-D view-grouping target, values[2] = 2-D view chunk, values[3] = 1 when
 * autotuned. */
int sgx_plan_tuning(const sgx_plan* plan, int64_t values[4]);

/* A batch: many reconstructions (plan, input, output; device pointers) run as
 * one HIP graph.  The first execute after items were added autotunes the
 * plans not yet tuned (through the context cache) and captures every item
 * into the graph; later executes replay the graph with the same pointers.
 * Backprojections run as queue launches: the filter phases of the queued
 * items run first on parallel branches, then one launch per backprojector
 * variant walks a table of (item, pixel tile or voxel block) tasks, items
 * ordered by views (largest tasks first), one workgroup per task, dispatched
 * in table order.  Items whose plan
 * occurs more than once in the batch (they share the plan's filtered
 * buffer) run as whole plan executes on the parallel branches instead
 * (independent plans in parallel, items of one plan in order).  No memset nodes are captured.  A plan change
 * that the graph would not see (sgx_plan_set_cbct_opts,
 * sgx_plan_set_view_weight, sgx_plan_autotune) makes the next execute
 * capture again.  The reference's counterpart is the serial slice loop of
 * fbp(plan, aa), src/fbp2/fbp.jl:46-68. */
typedef struct sgx_batch sgx_batch;
int sgx_batch_create(sgx_ctx* ctx, sgx_batch** out);
void sgx_batch_destroy(sgx_batch* batch);
int sgx_batch_add(sgx_batch* batch, sgx_plan* plan, const float* sino, float* image);
/* flags: SGX_ASYNC returns without waiting for the graph;
 * SGX_BATCH_NO_TUNE keeps the plans' current shapes.  By default a capture
 * times one replay of the graph with queue launches and one of the graph of
 * whole plan executes (no queues) and keeps the faster (sgx_batch_choice);
 * SGX_BATCH_QUEUE or SGX_BATCH_NO_QUEUE takes that form without timing.  A
 * change of these flags between executes captures again. */
#define SGX_BATCH_NO_TUNE 0x100u
#define SGX_BATCH_NO_QUEUE 0x200u
#define SGX_BATCH_QUEUE 0x400u
int sgx_batch_execute(sgx_batch* batch, uint32_t flags);
/* Milliseconds of the last execute's graph launch (HIP events). */
int sgx_batch_last_ms(sgx_batch* batch, float* ms);
/* Number of items and of parallel branches of the captured graph. */
int sgx_batch_info(const sgx_batch* batch, int64_t* items, int64_t* branches);
/* The captured graph's queues: their number, the items they hold and their
 * tasks (all 0 without queues). */
int sgx_batch_queues(const sgx_batch* batch, int64_t* queues, int64_t* queued_items,
                     int64_t* tasks);
/* The timed choice of the last capture: milliseconds of a replay with queue
 * launches and without (-1 when not timed). */
int sgx_batch_choice(const sgx_batch* batch, float* ms_queues, float* ms_launches);

/* ---- learned sparse-view correction (no counterpart in the reference) -----
 * A 1-D convolutional network applied along detector rows of the filtered
 * sinogram of a 2-D plan, before backprojection: q' = q + f(q), f = L layers
 * of width-9 convolutions (1 -> C -> ... -> C -> 1 channels) with GELU
 * between them; weights kept in FP32, used rounded to BF16, FP32
 * accumulation.  Callers: training loops and sparse-view reconstruction;
 * plain FBP (f = 0, the initial state) is the fallback.  Device pointers. */

/* Weights + ramp filter of a 2-D FP32 plan into q (ns, na) (no backprojection). */
int sgx_plan_filter(sgx_plan* plan, const float* sino, float* q, uint32_t flags);
/* Backprojection of a given filtered sinogram q (ns, na) with the plan's geometry. */
int sgx_plan_back(sgx_plan* plan, const float* q, float* image, uint32_t flags);
/* The exact transpose of sgx_plan_back: q (ns, na) from an image. */
int sgx_plan_back_t(sgx_plan* plan, const float* image, float* q, uint32_t flags);

typedef struct sgx_net sgx_net;
/* layers 2..16, channels 1..32; hidden weights uniform(+-sqrt(6/fan_in)) from
 * `seed`, last layer and biases 0 (the initial correction is zero). */
int sgx_net_create(sgx_ctx* ctx, int layers, int channels, uint64_t seed, sgx_net** out);
void sgx_net_destroy(sgx_net* net);
int64_t sgx_net_num_params(const sgx_net* net);
/* Copy the FP32 parameters to (set = 0) or from (set = 1) host memory; per
 * layer the weights (cout, cin, 9) row-major, then the biases (cout). */
int sgx_net_params(sgx_net* net, float* host, int set);
/* The accumulated parameter gradients (host), same layout. */
int sgx_net_grads(sgx_net* net, float* host);
int sgx_net_zero_grad(sgx_net* net);
/* out (rows, n) = q + f(q) for q (rows, n); keep = 1 keeps the activations
 * (and the q pointer) for sgx_net_backward. */
int sgx_net_forward(sgx_net* net, const float* q, int64_t rows, int64_t n, float* out, int keep);
/* Add dL/dparams to the gradients, given gout = dL/dout of the kept forward. */
int sgx_net_backward(sgx_net* net, const float* gout);
/* One Adam step on the FP32 parameters (bias-corrected, step counted inside). */
int sgx_net_adam(sgx_net* net, float lr, float beta1, float beta2, float eps);
/* grad = scale * 2 (x - t) (grad may be NULL); *sum_sq = sum (x - t)^2. */
int sgx_mse(sgx_ctx* ctx, const float* x, const float* t, int64_t n, float scale, float* grad,
            double* sum_sq);

#ifdef __cplusplus
}
#endif

#endif /* SGX_H */
