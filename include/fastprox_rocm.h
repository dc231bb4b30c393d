/*
 * fastprox_rocm.h -- fastprox's plain-C ABI for image total variation, on
 * AMD Instinct GPUs (gfx950) through HIP.
 *
 * This header declares the subset of fastprox's public C interface that the
 * image-TV path binds (the reference's include/fastprox.h, cited per entry
 * below), with the same names, types, status codes and semantics, so that a
 * caller of the reference's TV operators links against libfastprox_rocm.so
 * unchanged. Arrays may live on the host (FP_DEVICE_CPU: staged through
 * device memory) or on the GPU (FP_DEVICE_ROCM, DLPack's kDLROCM, the value
 * the reference reserves at include/fastprox.h:53). Every solve runs on the
 * HIP device current when the workspace was created, on the workspace's
 * stream (fp_workspace_set_stream; NULL: the null stream).
 *
 * Extensions with no counterpart in the reference are prefixed fpr_.
 *
 * SPDX-License-Identifier: MIT
 */
#ifndef FASTPROX_ROCM_H
#define FASTPROX_ROCM_H

#include <stddef.h>
#include <stdint.h>

#define FP_VERSION_MAJOR 1
#define FP_VERSION_MINOR 0
#define FP_VERSION_PATCH 6
#define FP_ABI_VERSION 1

#define FP_MAX_NDIM 32

#define FP_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* include/fastprox.h:52 */
typedef enum { FP_F32 = 1, FP_F64 = 2, FP_I8 = 3, FP_I32 = 4, FP_I64 = 5 } fp_dtype;
/* include/fastprox.h:54, with the reserved kDLROCM value enabled. */
typedef enum { FP_DEVICE_CPU = 0, FP_DEVICE_CUDA = 2, FP_DEVICE_ROCM = 10 } fp_device;

/* include/fastprox.h:61-68 */
typedef struct {
    void          *data;
    fp_dtype       dtype;
    int            ndim;
    const int64_t *shape;
    const int64_t *strides; /* in elements; NULL: C-contiguous */
    fp_device      device;
} fp_array;

/* include/fastprox.h:70-80 */
typedef enum {
    FP_OK = 0,
    FP_EINVAL,
    FP_EDTYPE,
    FP_ESHAPE,
    FP_ENOMEM,
    FP_ENOTCONV,
    FP_EPARAM,
    FP_EINTERNAL,
    FP_EDEVICE
} fp_status;

typedef struct fp_op fp_op;
typedef struct fp_workspace fp_workspace;

/* include/fastprox.h:93-103 (axis, tie: accepted and ignored by image operators;
 * threads: ignored, the GPU runs the batch) */
typedef struct {
    int    axis;
    int    threads;
    double tol;
    int    max_iter;
    int    tie;
    int    keep_structure;
} fp_options;

/* include/fastprox.h:105-109 */
typedef struct {
    double  gap;
    int     iters;
    int64_t ties;
} fp_info;

/* include/fastprox.h:113 -- defaults: axis -1, threads 0, tol 1e-6, max_iter 1000. */
FP_API void fp_options_init(fp_options *opt);

/* include/fastprox.h:322 -- anisotropic TV of the last ndim (2 or 3) axes. On
 * the GPU it is the device method of fastprox.torch.tv_aniso
 * (bindings/python/fastprox/_tv_torch.py:229-283): dual FISTA on the per-edge
 * duals with restarts and the gap check every 4 iterations. */
FP_API fp_status fp_op_tv_aniso(int ndim, const fp_array *axis_weights, fp_op **out);
/* include/fastprox.h:332 -- with edge weights of shape (ndim, image shape). */
FP_API fp_status fp_op_tv_aniso_weighted(int ndim, const fp_array *axis_weights, const fp_array *edge_weights,
                                         fp_op **out);
/* include/fastprox.h:343 -- isotropic TV, solved by the reference's dual FISTA
 * (src/ops/tv_iso.cpp:424-543) with its schedule. */
FP_API fp_status fp_op_tv_iso(int ndim, const fp_array *axis_weights, fp_op **out);
/* include/fastprox.h:352 -- with pixel weights of the image shape. */
FP_API fp_status fp_op_tv_iso_weighted(int ndim, const fp_array *axis_weights, const fp_array *pixel_weights,
                                       fp_op **out);

/* include/fastprox.h:392-394 */
FP_API void        fp_op_free(fp_op *op);
FP_API int64_t     fp_op_nparams(const fp_op *op);
FP_API const char *fp_op_name(const fp_op *op);

/* include/fastprox.h:411-413 -- all device memory of a solve is allocated
 * here: fp_prox and fp_prox_steps on a workspace allocate nothing. */
FP_API fp_status fp_workspace_create(const fp_op *op, const fp_array *like, const fp_options *opt,
                                     fp_workspace **out);
FP_API void      fp_workspace_free(fp_workspace *ws);

/* include/fastprox.h:426-428 -- the dual of the last call, shape (ndim, *x.shape). */
FP_API fp_status fp_workspace_get_dual(const fp_workspace *ws, fp_array *out);
FP_API fp_status fp_workspace_set_dual(fp_workspace *ws, const fp_array *dual);
FP_API fp_status fp_workspace_set_warm(fp_workspace *ws, int on);

/* include/fastprox.h:437-438 */
FP_API fp_status fp_workspace_set_deriv_tol(fp_workspace *ws, double tol, int max_iter);
FP_API fp_status fp_workspace_deriv_info(const fp_workspace *ws, double *residual, int *iters);

/* include/fastprox.h:599-601. On a workspace whose stream is being
 * captured into a hipGraph (fp_workspace_set_stream, hipStreamIsCapturing),
 * fp_prox enqueues the whole static schedule that max_iter allows, each
 * step and epilogue returning at once for the images already finished (the
 * same iterates as an eager call), and neither waits nor reads the host: x
 * and out must be contiguous device arrays, info NULL, keep_structure off,
 * and the method dual FISTA (FPR_METHOD_ADBPG reads the host between its
 * steps); fp_prox_steps is refused. The per-image results are read from the
 * graph's last run by the first fp_workspace_info or fp_workspace_image_info
 * (which waits for the stream). docs/spec.md:890 of the reference. */
FP_API fp_status fp_prox(const fp_op *op, double gamma, const fp_array *x, fp_array *out, const fp_options *opt,
                         fp_workspace *ws, fp_info *info);
/* include/fastprox.h:613-614 -- one step per image. */
FP_API fp_status fp_prox_steps(const fp_op *op, const fp_array *gamma, const fp_array *x, fp_array *out,
                               const fp_options *opt, fp_workspace *ws);

/* include/fastprox.h:646-654 */
FP_API fp_status fp_jvp(const fp_op *op, const fp_workspace *ws, const fp_array *dx, double dgamma,
                        const fp_array *dparams, fp_array *dout);
FP_API fp_status fp_vjp(const fp_op *op, const fp_workspace *ws, const fp_array *g, fp_array *gx, double *ggamma,
                        fp_array *gparams);

/* include/fastprox.h:712-725 */
FP_API int       fp_device_available(fp_device dev);
FP_API fp_status fp_op_device_check(const fp_op *op, fp_device dev, int *row_parallel_only);
FP_API fp_status fp_workspace_set_stream(fp_workspace *ws, void *stream);
FP_API fp_status fp_workspace_get_stream(const fp_workspace *ws, void **stream);

/* include/fastprox.h:732, 769, 773-774, 779-780 */
FP_API fp_status fp_workspace_info(const fp_workspace *ws, fp_info *info);
FP_API fp_status fp_workspace_deriv_exact(const fp_workspace *ws, int *exact, double *smoothing);
/* include/fastprox.h:759-771: the kept call solved again from its saved input
 * and its dual (a warm start) to tol within max_iter; the derivatives then
 * read the refined point and the results and dual report it; out (x's shape
 * and dtype) receives the refined z when not NULL. */
FP_API fp_status fp_workspace_refine(const fp_op *op, fp_workspace *ws, double tol, int max_iter,
                                     fp_array *out, fp_info *info);
FP_API fp_status fp_workspace_image_info(const fp_workspace *ws, double *gaps, int *iters, double *zerrs,
                                         int64_t count);
FP_API fp_status fp_vjp_device(const fp_op *op, const fp_workspace *ws, const fp_array *g, fp_array *gx,
                               fp_array *ggamma, fp_array *gparams);

/* Structure of the derivative (include/fastprox.h:514-579) and kinks
 * (:440-511) at the point kept by the last prox call on a keep_structure
 * workspace (FP_EINVAL without one), computed on the device. The units are
 * the images (count: the batch's images, size: their entries); every entry is
 * FP_ENTRY_FREE (sign +1) with the label of its region within its image, in
 * the order of the regions' first entries; an image whose input holds a NaN
 * has -1 regions, labels -1 and sign 0; no unit has a constraint. Output
 * arrays have x's shape (labels FP_I32 or FP_I64, state, sign and mask FP_I8,
 * coef x's dtype), host or device, any strides; each may be NULL. Kinks: iso,
 * 1 at a pixel with |(W D z)_p| <= tol and |u_p| >= gamma alpha_p - tol;
 * aniso, bit a at p where the edge (p, p + e_a) has |jump| <= tol and its 1D
 * dual within tol of gamma w_a beta (edges and pixels of zero bound are never
 * kinks). fp_workspace_set_branch is not provided. */
typedef enum { FP_ENTRY_FREE = 0, FP_ENTRY_ZERO = 1, FP_ENTRY_LOWER = 2, FP_ENTRY_UPPER = 3 } fp_entry_state;
typedef enum { FP_STRUCT_NONE = 0, FP_STRUCT_PROJECTOR = 1, FP_STRUCT_WEIGHTED = 2, FP_STRUCT_CURVED = 3 } fp_struct_kind;
/* include/fastprox.h:534-550: iso TV is FP_STRUCT_CURVED, aniso FP_STRUCT_PROJECTOR. */
FP_API fp_status fp_op_struct_kind(const fp_op *op, int *kind);
/* include/fastprox.h:552-554 */
FP_API fp_status fp_workspace_units(const fp_workspace *ws, int64_t *count, int64_t *size);
/* include/fastprox.h:558 */
FP_API fp_status fp_workspace_states(const fp_workspace *ws, fp_array *state, int64_t *nfixed);
/* include/fastprox.h:569-570 */
FP_API fp_status fp_workspace_regions(const fp_workspace *ws, fp_array *labels, fp_array *sign,
                                      int64_t *nregions, int64_t count);
/* include/fastprox.h:578-579 */
FP_API fp_status fp_workspace_constraint(const fp_workspace *ws, fp_array *coef, int8_t *active, int64_t count);
/* include/fastprox.h:510-511 */
FP_API fp_status fp_workspace_kinks(const fp_workspace *ws, double tol, fp_array *mask, int64_t *count);
/* include/fastprox.h:485-512. Anisotropic TV: for the edges at a kink within
 * tol (the kinks of fp_workspace_kinks), side[p] = -1 joins the two sides of
 * the edge (p, p + e_a) and +1 keeps them apart, for the fp_jvp and fp_vjp
 * calls until the next prox call or side = NULL; 0 keeps the default rule.
 * side: FP_I8, x's shape, any strides, host or device, values -1..1
 * (FP_EPARAM otherwise). FP_EINVAL for isotropic TV and without a kept point;
 * a failed call leaves an earlier override in place. */
FP_API fp_status fp_workspace_set_branch(fp_workspace *ws, const fp_array *side, double tol);

/* include/fastprox.h:782-789 */
FP_API const char *fp_strerror(fp_status s);
FP_API int         fp_abi_version(void);
FP_API const char *fp_version_string(void);
FP_API int         fp_default_threads(void);

/* ------------------------------------------------------------------------
 * The plugin table (src/device/plugin.hpp:16-23 in the reference): the one
 * symbol a loader resolves, with the ABI number, the build string it must
 * match, the device count and the operator check.
 * ---------------------------------------------------------------------- */
typedef struct {
    int abi;
    const char *build;
    int (*device_count)(void);
    fp_status (*op_check)(const fp_op *op, int *row_parallel_only);
} fpr_plugin_table;
#define FPR_PLUGIN_ABI 1
FP_API const fpr_plugin_table *fastprox_rocm_plugin(void);

/* ------------------------------------------------------------------------
 * Extensions (no counterpart in the reference).
 * ---------------------------------------------------------------------- */

/* A short description of why the last failing fpr_ call on this thread failed. */
FP_API const char *fpr_last_error(void);

/* fp_prox without a workspace for callers that cannot allocate device memory
 * (code under graph capture, serving loops, harnesses that forbid it): no
 * counterpart in the reference. It allocates nothing: its device scratch (the
 * dual iterates, descriptors, states, tile lists, partial sums, the dual
 * output, a copy of x when out aliases x, the weights) is a fixed 64 MiB array
 * that every GPU the library runs on holds until the process ends. x and out:
 * contiguous device arrays on the current device, at most 256 images whose
 * buffers fit the scratch (FP_ENOMEM otherwise, with fpr_last_error saying
 * which limit); stream: a hipStream_t (NULL: the null stream); the call
 * blocks until the solve is done. Results as fp_prox's (info: the gaps
 * summed and the largest iteration count; no dual or structure is kept). The
 * kernel is fixed: 2D, the streaming kernel with 6 iterations per launch and
 * 16-row segments; 3D, the streaming kernel with 16-slice slabs. Calls that
 * overlap on one GPU share its scratch and race: serialise them. */
FP_API fp_status fpr_prox_static(const fp_op *op, double gamma, const fp_array *x, fp_array *out,
                                 const fp_options *opt, void *stream, fp_info *info);

/* The number of device and pinned host allocations (hipMalloc,
 * hipHostMalloc) the library has made in this process. Every device buffer is
 * allocated by a create call (fp_workspace_create; with keep_structure also
 * the derivative structure, its scratch and the queries' buffers;
 * fpr_batch_create, fpr_server_create, fpr_deriv_create): fp_prox,
 * fp_prox_steps, fp_jvp, fp_vjp, fp_vjp_device, fp_workspace_refine, the
 * dual calls, the structure queries and fp_workspace_set_branch leave the
 * count unchanged (the reference's device rule, docs/spec.md:887, and its
 * allocation counter, tests/c/test_core.cpp:40-67). No counterpart in the
 * reference's C ABI. */
FP_API int64_t fpr_alloc_count(void);

/* The step kernel of a workspace: generic = 1 selects the one-iteration
 * kernel; otherwise the register-tile kernel with tt iterations per launch
 * and ry rows per tile (FP_EINVAL for a combination not built). The default
 * is the autotuned choice for the workspace's shape and dtype. */
FP_API fp_status fpr_workspace_set_kernel(fp_workspace *ws, int generic, int tt, int ry);
FP_API fp_status fpr_workspace_get_kernel(const fp_workspace *ws, int *generic, int *tt, int *ry);
/* Launch statistics of the last solve on ws: step launches, host syncs. */
FP_API fp_status fpr_workspace_stats(const fp_workspace *ws, int *launches, int *host_syncs);

/* The method of an anisotropic workspace's solves. FPR_METHOD_ADBPG (the
 * default of workspaces for operators without edge weights) is the
 * reference CPU library's (src/ops/tv_aniso.cpp:259-616, 886-953): exact 1D
 * TV steps on every fibre, the accelerated dual block proximal gradient
 * method for two active axes and the three-block fast dual proximal gradient
 * method for three, its gap schedule, restarts and polish, so that
 * iterates, iteration counts and results follow the reference's fp_prox.
 * FPR_METHOD_FISTA is dual FISTA on the per-edge dual, the reference's
 * device path (bindings/python/fastprox/_tv_torch.py:209-282): the default of
 * batches (fpr_batch_*, and so of fastprox_amd.torch), and the method of
 * edge-weighted operators. FASTPROX_ROCM_ANISO=adbpg or fista sets the
 * default of new workspaces and batches. Setting a method builds its
 * solver's buffers (it allocates); the dual carries over. FP_EINVAL for an
 * isotropic workspace, or ADBPG on edge weights. No counterpart in the
 * reference's C ABI. */
#define FPR_METHOD_ADBPG 0
#define FPR_METHOD_FISTA 1
FP_API fp_status fpr_workspace_set_method(fp_workspace *ws, int method);
FP_API fp_status fpr_workspace_get_method(const fp_workspace *ws, int *method);
/* ADBPG workspaces: per image of the last solve, whether its polish
 * certified it (the regions read off the dual, z their closed-form means),
 * and the number of fibres whose 1D step took the taut string after
 * Condat's rescan budget (tv1d.cpp:264). */
FP_API fp_status fpr_workspace_polished(const fp_workspace *ws, int *polished, int64_t count, int *fallbacks);

/* A batch of images of uneven sizes, each with its own step and, for a
 * weighted batch, its own spatial weights, solved in one persistent solve
 * (the same dual FISTA per image as fp_prox). Every pointer is a device
 * pointer of the batch's dtype to contiguous C-order data. */
typedef struct fpr_batch fpr_batch;
typedef struct {
    const void *x;        /* the image (shape from fpr_batch_create) */
    void *z;              /* its output */
    void *dual;           /* its dual: ndim planes of the image's size, in the
                             form of fp_workspace_get_dual; read first when
                             fpr_batch_solve is called with warm = 1 */
    const void *weights;  /* weighted batches: pixel weights (iso, the
                             image's size) or edge weights (aniso, ndim
                             planes); else ignored */
    double gamma;         /* its step */
} fpr_image;
/* shapes: count rows of op's ndim extents. The batch allocates all device
 * memory its solves need; fpr_batch_solve allocates nothing. */
FP_API fp_status fpr_batch_create(const fp_op *op, fp_dtype dtype, int64_t count, const int64_t *shapes,
                                  int weighted, fpr_batch **out);
FP_API void      fpr_batch_free(fpr_batch *b);
/* Solves every image on stream (NULL: the null stream) and waits for it.
 * FP_ENOTCONV when an image stopped at max_iter; FP_EINVAL for an infinite
 * entry; FP_EPARAM for a step that is negative, not finite or out of range.
 * On a stream being captured into a hipGraph it enqueues the static schedule
 * instead, as fp_prox does, and returns FP_OK at once: the one captured
 * graph then solves the whole batch; fpr_batch_image_info and
 * fpr_batch_scale_exps read the graph's last run (waiting for the stream). */
FP_API fp_status fpr_batch_solve(fpr_batch *b, const fpr_image *images, const fp_options *opt, int warm,
                                 void *stream);
/* Per image, the results of the last solve (any pointer may be NULL). */
FP_API fp_status fpr_batch_image_info(const fpr_batch *b, double *gaps, int *iters, double *zerrs,
                                      int *converged, int64_t count);
/* Selecting a step kernel on an anisotropic batch (or workspace) selects
 * FPR_METHOD_FISTA, whose kernels they are. */
FP_API fp_status fpr_batch_set_kernel(fpr_batch *b, int generic, int tt, int ry);
/* The method of an anisotropic batch (as fpr_workspace_set_method):
 * FPR_METHOD_FISTA by default; FPR_METHOD_ADBPG for batches of one shape
 * without edge weights. */
FP_API fp_status fpr_batch_set_method(fpr_batch *b, int method);
FP_API fp_status fpr_batch_get_method(const fpr_batch *b, int *method);
FP_API fp_status fpr_batch_get_kernel(const fpr_batch *b, int *generic, int *tt, int *ry);
FP_API fp_status fpr_batch_stats(const fpr_batch *b, int *launches, int *host_syncs);
/* The last solve's pixel passes: sum over the images of the pixels times the
 * step launches that iterated the image (each launch reads and writes an
 * image's state once, whatever its iterations): the solve's algorithmic HBM
 * traffic is this times (1 + 4k) words, plus one word per pixel with pixel
 * weights. No counterpart in the reference. */
FP_API fp_status fpr_batch_pixel_passes(const fpr_batch *b, double *pixel_passes);
/* Kernel timing of the solves on b (HIP events on the solve's stream around
 * every step and epilogue launch): on = 1 enables it for later solves; the
 * last solve's summed step-kernel and epilogue milliseconds and the number of
 * step launches timed. */
FP_API fp_status fpr_batch_set_timing(fpr_batch *b, int on);
FP_API fp_status fpr_batch_timing(const fpr_batch *b, double *step_ms, double *epilogue_ms, int *launches);
/* With timing on: the time, in ms, during which at least one step kernel of
 * the last solve ran. A batch of many tiles runs as two halves on two
 * streams whose step kernels overlap, so this is the device's step time,
 * below the summed step_ms of fpr_batch_timing; with one stream the two
 * agree. */
FP_API fp_status fpr_batch_busy(const fpr_batch *b, double *busy_ms);
/* Each chunk of a solve's step and epilogue launches is captured once as a
 * hipGraph and replayed (on by default; timing turns it off for the timed
 * solves); on = 0 launches every kernel on its own. Results are the same
 * either way. No counterpart in the reference. */
FP_API fp_status fpr_batch_set_graph(fpr_batch *b, int on);
/* The tail kernel of a batch or workspace (2D, streaming kernels): once the
 * live tiles of a solve fall below frac of all its tiles, the remaining
 * launches use the streaming kernel with tt iterations and segments of ry
 * rows on its own tiling (tt at least the main kernel's). Results do not
 * change (every kernel computes the same iterates). frac = 0 turns it off;
 * the default comes from
 * FASTPROX_ROCM_TAIL ("stream:TT,SEG,FRAC") or the tuning. */
FP_API fp_status fpr_batch_set_tail(fpr_batch *b, int tt, int ry, double frac);
FP_API fp_status fpr_workspace_set_tail(fp_workspace *ws, int tt, int ry, double frac);
/* The last solve's power-of-two scale exponent per image (isotropic solves
 * far from unit scale run on 2^-e x; src/ops/tv_iso.cpp:367-377). */
FP_API fp_status fpr_batch_scale_exps(const fpr_batch *b, int *exps, int64_t count);
/* Copies the last solve's dual iterate u of every image to dst (device, the
 * batch's dtype): image i's ndim planes at dst + i ndim N. Anisotropic
 * batches: the per-edge dual |u_e| <= gamma beta_e that the derivatives read
 * (the C form of fpr_image.dual sums it along each axis). Uniform batches only
 * (FP_ESHAPE otherwise). Enqueued on stream; does not wait. */
FP_API fp_status fpr_batch_copy_dual_iterate(const fpr_batch *b, void *dst, void *stream);

/* Derivatives of a solved uniform batch, the device counterpart of the
 * keep_structure workspace for callers that hold the solution themselves
 * (the PyTorch wrapper's autograd; _tv_torch.py:286-585 Structure):
 *   x, z: the batch's inputs and outputs, B contiguous images of img_shape;
 *   dual: (B, ndim, *img) contiguous: isotropic, fpr_image.dual of the solve;
 *         anisotropic, fpr_batch_copy_dual_iterate's per-edge dual;
 *   spatial: pixel weights (B, *img), or edge weights (B, ndim, *img), or one
 *         map shared by every image (spatial_per_image = 0), or NULL;
 *   gammas, gaps (host, B): the solve's steps and gaps; scale_exps (host, B,
 *         or NULL for 0): fpr_batch_scale_exps.
 * The device arrays must outlive the handle. Creating it builds the structure
 * (the flat pixels or edges and their regions) on stream and waits. */
typedef struct fpr_deriv fpr_deriv;
FP_API fp_status fpr_deriv_create(const fp_op *op, fp_dtype dtype, int64_t count, const int64_t *img_shape,
                                  const void *x, const void *z, const void *dual, const void *spatial,
                                  int spatial_per_image, const double *gammas, const double *gaps,
                                  const int *scale_exps, void *stream, fpr_deriv **out);
FP_API void      fpr_deriv_free(fpr_deriv *d);
/* VJP of the cotangent g (device, B images): gx (device, B images); ggamma
 * (device, B doubles: one per image) or NULL; gw (host, ndim doubles: the axis
 * weights' gradient summed over the images) or NULL; gspatial (device
 * doubles shaped like the spatial weights; summed over the images for a
 * shared map) or NULL. tol 0: the default tolerance (the relative error of z
 * that each image's gap certifies, at least 1e-10). Waits for the stream. */
FP_API fp_status fpr_deriv_vjp(fpr_deriv *d, const void *g, void *gx, double *ggamma, double *gw, double *gspatial,
                               double tol, int max_iter, void *stream);
/* JVP along dx (device, B images, or NULL), dgamma (device, B doubles, or
 * NULL), dw (host, ndim doubles, or NULL), dspatial (device doubles shaped
 * like the spatial weights, or NULL): dz (device, B images). Waits. */
FP_API fp_status fpr_deriv_jvp(fpr_deriv *d, const void *dx, const double *dgamma, const double *dw,
                               const double *dspatial, void *dz, double tol, int max_iter, void *stream);
/* The last derivative's largest CG residual and iterations over the images
 * (0 for anisotropic TV: the region means are exact), and the number of
 * regions of the structure. */
FP_API fp_status fpr_deriv_info(const fpr_deriv *d, double *residual, int *iters, int64_t *regions);
FP_API fp_status fpr_deriv_image_info(const fpr_deriv *d, double *residuals, int *iters, int64_t count);

/* A serving loop: requests of different sizes, steps and tolerances solved
 * by continuous batching on one GPU. Each class (2D, 3D) has a table of
 * `slots` slots, a device pool of pool_bytes for the requests' buffers and its
 * own stream; the step kernels run over every live request of a class at
 * once, and a request is admitted as soon as a slot, its memory and its tiles
 * free up (first come, first served; kept duals are evicted, least recently
 * used first, when memory is short). A request's result is the same as the
 * plain solve's (fp_prox on that image alone). No counterpart in the
 * reference, which has no server; the solve per request is tv_iso.cpp's. */
typedef struct fpr_server fpr_server;
typedef struct {
    int ndim;                 /* 2 or 3: the class */
    int64_t shape[3];         /* the first ndim entries */
    const void *x;            /* device, the server's dtype; read at admission */
    const int64_t *strides;   /* element strides of x, or NULL for C order */
    void *z;                  /* device output (contiguous), or NULL; the
                                 server writes it in place at any time between
                                 the submit and the poll that returns the
                                 result, so the caller leaves it alone then */
    double gamma;
    double tol;               /* stop at gap <= tol |x|^2 / 2 */
    int max_iter;
    int keep_dual;            /* keep the dual resident after completion */
    int64_t warm_from;        /* start from the kept dual of that request
                                 (same shape), or -1 */
    void *stream;             /* the caller's hipStream_t (NULL: the null
                                 stream): x is read after the work enqueued on
                                 it before the submit, and z is written before
                                 the work enqueued on it after the poll that
                                 returns the request's result */
} fpr_request;
typedef struct {
    int64_t id;
    int status;               /* FP_OK, or FP_EINVAL for an infinite entry */
    int converged;
    int iters;
    double gap;
    double zerr;              /* certified relative error of z */
    double t_submit, t_admit, t_done;  /* seconds on the server's clock */
} fpr_result;
/* op2d / op3d: the operators of the two classes (NULL: class not served;
 * spatial weights are not served). */
FP_API fp_status fpr_server_create(const fp_op *op2d, const fp_op *op3d, fp_dtype dtype, size_t pool_bytes,
                                   int slots, fpr_server **out);
FP_API void      fpr_server_free(fpr_server *s);
/* Queues r; t_submit: its arrival on the server clock (< 0: now). */
FP_API fp_status fpr_server_submit(fpr_server *s, const fpr_request *r, double t_submit, int64_t *id);
/* One round of the loop (admissions, launches, retirements); busy: live
 * requests plus chunks in flight, queued: requests waiting. */
FP_API fp_status fpr_server_step(fpr_server *s, int *busy, int *queued);
/* Moves up to max finished requests' results to out; returns the count. A
 * result's z is ordered on the request's stream: work enqueued there after
 * this call sees it (a host read needs that stream synchronised first).
 * Allocates nothing. */
FP_API int64_t   fpr_server_poll(fpr_server *s, fpr_result *out, int64_t max);
FP_API double    fpr_server_clock(const fpr_server *s);
FP_API fp_status fpr_server_release(fpr_server *s, int64_t id);
FP_API fp_status fpr_server_stats(const fpr_server *s, int64_t *launches);
/* The step kernel serving k-dimensional requests (kind as fpr_batch_get_kernel:
 * 0 streaming 2D, 1 generic, 2 block, 3 streaming 3D; tt iterations per launch,
 * ry rows per segment or slices per slab). No counterpart in the reference. */
FP_API fp_status fpr_server_get_kernel(const fpr_server *s, int k, int *kind, int *tt, int *ry);

#ifdef __cplusplus
}
#endif

#endif /* FASTPROX_ROCM_H */
