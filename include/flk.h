/*
 * flk.h — C ABI of the MI355X linking-number engine (libflk.so).
 *
 * Plain C types only: host pointers, sizes, status codes.  The engine owns all
 * device memory; callers never see HIP or torch types.  One context = one HIP
 * stream on one device; a context is not thread-safe, separate contexts are
 * independent.  Every call is synchronous with respect to the caller.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository antequ/fastlinkingnumbers).
 */
#ifndef FLK_H_
#define FLK_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: 0 = success, negative = error (message via flk_last_error). */
#define FLK_OK 0
#define FLK_ERR_ARG (-1)
#define FLK_ERR_HIP (-2)
#define FLK_ERR_NOMEM (-3)
#define FLK_ERR_STATE (-4)
#define FLK_ERR_DEVICE (-5)

/* Linking-number methods: the --method switch, src/main.cpp:34-42. */
#define FLK_DIRECTSUM 0
#define FLK_BARNESHUT 1

/* Barnes–Hut opening criteria. */
#define FLK_CRIT_BETA 0 /* |r| > beta (r1 + r2), src/compute_link.cpp:573-574 */
#define FLK_CRIT_MLP 1  /* learned far-field error predictor (no reference counterpart) */

/* Per-pair result flags (bit set in flags[k]). */
#define FLK_PAIR_FALLBACK 1u /* recomputed by the FP64 direct sum */
#define FLK_PAIR_NAN 2u      /* the final sum is NaN */
#define FLK_PAIR_RETRY 4u    /* Barnes–Hut re-run with a larger beta */
#define FLK_PAIR_DIRECT 8u   /* Barnes–Hut request served by direct summation */

typedef struct flk_ctx flk_ctx;

typedef struct flk_params {
  int32_t method;       /* FLK_DIRECTSUM or FLK_BARNESHUT */
  int32_t criterion;    /* FLK_CRIT_BETA or FLK_CRIT_MLP */
  double beta_init;     /* 2.0,  src/define_args.h:51 */
  double beta_limit;    /* 10.0, src/define_args.h:54 */
  double eps_sq;        /* 1e-24, src/linking_number_certificate.cpp:24 */
  double target_error;  /* 0.1,   src/linking_number_certificate.cpp:25 */
  double err_scale;     /* 0.05,  src/linking_number_certificate.cpp:26 */
  double fallback_tol;  /* a pair whose sum is farther than this from an integer
                           is recomputed by the FP64 direct sum (<= 0 disables) */
  int64_t bh_direct_threshold; /* BH pairs with n1*n2 below this use the direct sum */
  int32_t leaf_size;    /* segments per Barnes–Hut leaf bucket (1 = reference) */
  int32_t fp64;         /* 1: direct sum in FP64 instead of FP32 */
  double mlp_log_tol;   /* MLP criterion: accept far field when the predicted
                           log10 |error| is below this */
  int32_t bh_batch_ksegs; /* Barnes–Hut pairs are processed in batches whose
                             curves total at most this many thousand segments
                             (0: sized from free device memory) */
  int32_t timing;       /* 1: time every Barnes–Hut level launch with HIP events
                           (ms_bh_level, ms_mlp); the per-launch
                           events cost ~15% of a traversal, so 0 by default */
  int32_t graph;        /* 1 (default): a direct-sum batch, or a whole Barnes–Hut
                           certificate (its direct-sum part, tree builds and
                           traversal passes), is captured in a HIP graph and
                           replayed when the same pairs come again; a
                           certificate with both parts runs its direct sum as
                           a branch of its own, on the context's second stream
                           (default priority), beside the Barnes–Hut batches;
                           0: launched eagerly every time */
  int32_t reserved[5];
} flk_params;

typedef struct flk_stats {
  double ms_total;          /* wall time of the last flk_linking_numbers call */
  double ms_kernel;         /* device time of its main kernel(s), HIP events */
  double ms_ds_kernel;      /* device time of the direct-sum tile kernel */
  double ms_bh_traverse;    /* wall time of the Barnes–Hut traversal (host clock) */
  double ms_bh_build;       /* wall time of tree + moment construction (host clock) */
  double ms_fallback;       /* device time of FP64 fallback recomputation */
  int64_t seg_pairs;        /* sum of n1*n2 over the requested pairs */
  int64_t ds_seg_pairs;     /* segment pairs evaluated by direct summation */
  int64_t bh_visits;        /* Barnes–Hut node-pair visits */
  int64_t bh_far;           /* far-field evaluations */
  int64_t bh_leaf_pairs;    /* exact segment pairs evaluated in leaf buckets */
  int64_t fallbacks;        /* pairs recomputed by the FP64 direct sum */
  int64_t retries;          /* pairs re-run with a larger beta */
  int32_t ds_launches;      /* direct-sum tile kernel launches */
  int32_t ds_waves;         /* waves per direct-sum launch */
  double ms_bh_level;       /* device time of the traversal level kernels (params.timing) */
  double ms_mlp;            /* device time of the level kernels that evaluated the learned criterion,
                               which runs inside them (params.timing) */
  int64_t mlp_evals;        /* learned-criterion network evaluations */
  int64_t bh_host_passes;   /* Barnes–Hut passes that ran level by level from the host
                               (per-launch timing, or the device-scheduled pass's
                               buffers overflowed) instead of all levels enqueued at once */
  int64_t graph_launches;   /* Barnes–Hut certificates run as one captured HIP graph (params.graph:
                               the direct-sum part, every tree build and both traversal passes);
                               their device time is ms_kernel, and ms_bh_build/ms_bh_traverse
                               stay 0 */
  double ms_graph_launch;   /* host time of that hipGraphLaunch call */
  int64_t reserved[2];
} flk_stats;

/* Fills *p with the reference defaults (method Barnes–Hut, beta 2/10, ...). */
int32_t flk_default_params(flk_params *p);

/* Context lifecycle.  No reference counterpart (the reference is CPU-only);
 * replaces the implicit OpenMP runtime of src/linking_number_certificate.cpp. */
int32_t flk_create(int32_t device, flk_ctx **out);
void flk_destroy(flk_ctx *ctx);
const char *flk_last_error(const flk_ctx *ctx);
int32_t flk_device_name(const flk_ctx *ctx, char *buf, int32_t len);

/* Uploads closed polylines: xyz is AoS FP64 (3 * offsets[ncurves] doubles),
 * curve c owns vertices [offsets[c], offsets[c+1]).  Replaces the per-pair
 * Curve::GetStartPoints / GetLoopedStartPoints copies made by
 * src/compute_link.cpp:204-205 and :131-132 (the vertex list of a discretized
 * curve is its start-point list, include/curve.h:105-108). */
int32_t flk_set_curves(flk_ctx *ctx, const double *xyz, const int64_t *offsets,
                       int32_t ncurves);

/* flk_set_curves for FP32 input (12 bytes per vertex over PCIe): the layout of
 * a BCC file's control points, which the reference widens to FP64 Curve
 * objects on import (src/model.cpp:299-345) before every computation.  The
 * result is identical to flk_set_curves on the same values widened to FP64;
 * flk_set_curves itself sends FP64 input as FP32 whenever every value
 * converts exactly. */
int32_t flk_set_curves_f32(flk_ctx *ctx, const float *xyz, const int64_t *offsets,
                           int32_t ncurves);

/* Curve bounding boxes of the uploaded polylines, ncurves x (min xyz, max xyz)
 * in FP64.  Replaces PrecomputeCurveBoundingBoxes,
 * src/potential_link_search.cpp:68-73 (polyline case of curve.cpp:331-343). */
int32_t flk_curve_boxes(flk_ctx *ctx, double *boxes_out);

/* Potential-link search over curve boxes (ncurves x 6 FP64; NULL = use the
 * boxes of the uploaded curves).  Writes the sorted unique pair list (i < j,
 * closed-interval overlap) into a context-owned array valid until the next
 * call.  Replaces GetPotentialLinksUniqueList,
 * src/potential_link_search.cpp:88-127. */
int32_t flk_potential_links(flk_ctx *ctx, const double *boxes, int32_t ncurves,
                            const int32_t **pairs, int64_t *npairs);

/* Linking numbers of curve pairs (2 * npairs indices into the uploaded curves).
 * lk[k] receives the summed Gauss integral of pair k, err_est[k] (may be NULL)
 * the Barnes–Hut error estimate (0 for direct summation), flags[k] (may be
 * NULL) FLK_PAIR_* bits.  Replaces the per-pair loops of
 * ComputeLinkingNumbersDirectSum (src/linking_number_certificate.cpp:241-338,
 * calling src/compute_link.cpp:126-280) and ComputeLinkingNumbersBarnesHut
 * (src/linking_number_certificate.cpp:102-229, calling
 * src/compute_link.cpp:294-934). */
int32_t flk_linking_numbers(flk_ctx *ctx, const int32_t *pairs, int64_t npairs,
                            const flk_params *params, double *lk, double *err_est,
                            uint32_t *flags);

/* No counterpart in the reference: flk_linking_numbers for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Direct summation only (params->method
 * must be FLK_DIRECTSUM; FP32, or FP64 with params->fp64), with the same
 * kernels, results and FP64 judge as flk_linking_numbers launched eagerly
 * (params->graph is ignored).  The batch's plan, per-task partials and
 * per-pair sums live in one fixed device array per GPU of
 * FLK_NOALLOC_SCRATCH_MB MiB, held from the first flk_create on that GPU until
 * the process ends; a batch may hold at most FLK_NOALLOC_MAX_PAIRS pairs and
 * FLK_NOALLOC_MAX_TASKS tasks (the plan has one task per 64R rows and column
 * chunk; FLK_ERR_ARG otherwise, naming the count).  The curve set must have
 * been uploaded before (flk_set_curves*, which allocates).  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (two contexts
 * driven from two host threads) race and must be serialised by the caller. */
#define FLK_NOALLOC_MAX_PAIRS 65536
#define FLK_NOALLOC_MAX_TASKS 1048576
#define FLK_NOALLOC_SCRATCH_MB 40
int32_t flk_linking_numbers_noalloc(flk_ctx *ctx, const int32_t *pairs, int64_t npairs,
                                    const flk_params *params, double *lk, double *err_est,
                                    uint32_t *flags);

/* Statistics of the last flk_linking_numbers call. */
int32_t flk_get_stats(const flk_ctx *ctx, flk_stats *out);

/* One node of a curve's Barnes–Hut tree, in input coordinates.  Moments are
 * about the node's box centre, with the reference's index convention
 * (src/compute_link.h:95-114): D[3k+j] = int r'_k (r - center)_j and
 * Q[9k+3i+j] = int r'_k (r - center)_i (r - center)_j; the norms feed the
 * Eq. 14 error estimate (src/compute_link.cpp:383-388). */
typedef struct flk_tree_node {
  double center[3];   /* box centre (FP32 hi + lo, as the traversal uses it) */
  double radius;      /* half box diagonal */
  double rsq;         /* radius squared as the traversal compares it */
  double c[3];        /* monopole, int r' */
  double D[9];
  double Q[27];
  double cn, dn, qn;  /* |c|, |D|_F / sqrt 3, sqrt(sum_k |Q_k|_F^2) / 3 */
  int32_t child[2];   /* node indices; -1 for a leaf bucket */
  int32_t seg0, nseg; /* range of the tree's segment order */
} flk_tree_node;

/* Builds the Barnes–Hut tree of uploaded curve `curve` with leaf buckets of at
 * most leaf_size segments -- the same trees and FP32 records the Barnes–Hut
 * path of flk_linking_numbers builds -- and copies it out: *nnodes nodes,
 * root first (nodes may be NULL to query the count; otherwise cap must hold
 * them), and seg_order[s] (may be NULL; n entries for a curve of n segments)
 * = the curve segment at position s of the tree's order.  A curve of fewer
 * than 2 vertices has no tree (*nnodes = 0).  Replaces MakeSegmentTrees and
 * BuildTreeMoments (src/compute_link.h:128-151, src/compute_link.cpp:294-389)
 * for callers that inspect the trees. */
int32_t flk_tree_moments(flk_ctx *ctx, int32_t curve, int32_t leaf_size, flk_tree_node *nodes,
                         int32_t cap, int32_t *nnodes, int32_t *seg_order);

/* Segment types of flk_discretize, in the order of the reference's
 * SegmentType (include/curve.h). */
#define FLK_SEG_BSPLINE 0
#define FLK_SEG_CATMULLROM 1
#define FLK_SEG_BEZIER 2
#define FLK_SEG_POLYLINE 3

typedef struct flk_disc_info {
  int64_t segments;         /* input segments */
  int64_t vertices;         /* output vertices over all curves */
  int32_t rounds;           /* bisection rounds run */
  int32_t small_curve;      /* when the input guard fails: the smallest segment */
  int32_t small_segment;
  int32_t pad;
  double small_size;        /* and its box diagonal */
  int64_t overlaps[55];     /* overlapping piece pairs found in each round */
  int64_t pieces;           /* output polyline segments over all curves */
  int64_t reserved[3];
} flk_disc_info;

/* Replaces Discretize (src/discretize.cpp:102-281): bisects the pieces of
 * every segment whose exact box overlaps a piece of a potentially linked
 * curve (pairs: i < j, as from flk_potential_links on the curves' spline
 * boxes) until no overlap is left, then returns each curve as a polyline of
 * its pieces' start points (an open curve also gets its last end point).
 * ctrl_xyz: control points, AoS, curve c = rows offsets[c]..offsets[c+1]
 * (NULL: the context's current curve set, whose offsets must be given);
 * segment_types: FLK_SEG_*; looped: 0/1 per curve.  On success the polylines
 * become the context's curve set, as after flk_set_curves on them (they stay
 * on the device: no copy back and up again), and *out_xyz / *out_offsets
 * receive a context-owned host copy (valid until the next call; either may be
 * NULL, and out_xyz NULL skips the copy).  *info (may be NULL) gets the
 * counts and, on the reference's failures (smallest input segment below
 * eps * mean |q| / sqrt(3), a new piece below it, 55 rounds with overlaps
 * left), the error details (status FLK_ERR_ARG, message as the reference's). */
int32_t flk_discretize(flk_ctx *ctx, const double *ctrl_xyz, const int64_t *offsets,
                       const int32_t *segment_types, const uint8_t *looped, int32_t ncurves,
                       const int32_t *pairs, int64_t npairs, const double **out_xyz,
                       const int64_t **out_offsets, flk_disc_info *info);

/* Result of flk_load_bcc.  offsets and looped are context-owned, valid until
 * the next flk_load_bcc. */
typedef struct flk_bcc_info {
  int32_t ncurves;
  int32_t segment_type;     /* FLK_SEG_* of every curve of the file */
  int64_t points;           /* control points over all curves */
  const int64_t *offsets;   /* ncurves + 1: curve c = points offsets[c]..offsets[c+1] */
  const uint8_t *looped;    /* ncurves: 1 for a closed loop (negative count in the file) */
  int64_t reserved[4];
} flk_bcc_info;

/* Reads a BCC file straight into the context's curve set, as
 * flk_set_curves_f32 on its control points would: the reference's
 * Model::ImportFromFile for BCC (src/model.cpp:280-364, same header checks and
 * messages) without the FP64 Curve objects.  The file is read in chunks into
 * pinned buffers whose DMA overlaps the reading of the next chunk.  For a
 * polyline file the curves are ready for flk_potential_links and
 * flk_linking_numbers; for a spline file they are the control polygons, which
 * flk_discretize(ctx, NULL, info->offsets, ...) discretizes in place.  A
 * file that fails to read (status FLK_ERR_ARG and the reference's message:
 * header, counts, truncated coordinates) leaves the previous curve set in
 * place; a device failure after the file was read leaves an empty one. */
int32_t flk_load_bcc(flk_ctx *ctx, const char *path, flk_bcc_info *info);

/* Launch shapes of the kernels.  No reference counterpart (the reference is
 * CPU-only): flk_autotune picks them by timing candidates on the caller's own
 * pairs, flk_set_tune installs a saved choice.  Every shape gives the same
 * results up to floating-point summation order. */
typedef struct flk_tune {
  int32_t ds_rows_per_lane;    /* FP32 direct-sum tile: rows per lane, 2, 4 or 8
                                  (a wave owns 64 * R consecutive rows) */
  int32_t ds_blocks_per_cu;    /* persistent direct-sum grid: 256-thread blocks
                                  per CU (0: as many as are resident at once) */
  int32_t ds_tasks_per_wave;   /* column chunking: target tasks per wave */
  int32_t bh_items_per_thread; /* Barnes–Hut level and leaf kernels: 2, 4 or 8 (0: by
                                  criterion, 8 for beta and 4 for the learned one) */
  int32_t ingest_threads;      /* host threads filling the upload's pinned
                                  staging buffers (0: half the cores, at most 8) */
  int32_t bh_level_capacity;   /* Barnes–Hut items per level buffer when all levels
                                  of a pass are enqueued at once (0: sized from free
                                  memory, at most 2^26); a pass whose level or leaf
                                  list outgrows it re-runs level by level from the host */
  int32_t reserved[2];
} flk_tune;

int32_t flk_get_tune(const flk_ctx *ctx, flk_tune *out);
int32_t flk_set_tune(flk_ctx *ctx, const flk_tune *tune);

/* Times candidate shapes for these pairs and parameters -- the direct-sum
 * shapes on a deterministic sample of the pairs' direct-sum work (every k-th
 * 64R-row block, about a tenth of it), the Barnes–Hut shapes on the
 * Barnes–Hut pairs -- keeps the
 * fastest in the context and writes it to *out (NULL allowed).  The
 * linking numbers themselves are not returned. */
int32_t flk_autotune(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, const flk_params *params,
                     flk_tune *out);

/* Learned opening criterion (no reference counterpart; it replaces the beta
 * test of src/compute_link.cpp:573-574 and :608).  Weights are FP32 masters of
 * the 12-64-64-1 MLP with hard-swish activations, h(z) = z min(max(z/6 + 1/2,
 * 0), 1): W1[64x12] b1[64] W2[64x64] b2[64] W3[64] b3[1] (5,057 floats), plus
 * 24 floats of feature normalisation (mean[12], std[12]).
 */
#define FLK_MLP_NPARAM 5057
#define FLK_MLP_NFEAT 12
int32_t flk_mlp_set_weights(flk_ctx *ctx, const float *params, const float *norm24);
int32_t flk_mlp_get_weights(const flk_ctx *ctx, float *params, float *norm24);
/* Forward pass of the FP16-MFMA inference kernel: x is n x 12 FP32 features
 * (raw, normalised inside), y receives n predicted log10 errors. */
int32_t flk_mlp_forward(flk_ctx *ctx, const float *x, int64_t n, float *y);
/* One full-batch gradient of the mean squared error over (x, y) with the
 * bf16-MFMA forward/backward kernels; grad receives FLK_MLP_NPARAM floats,
 * loss the mean loss. */
int32_t flk_mlp_gradient(flk_ctx *ctx, const float *x, const float *y, int64_t n,
                         float *grad, float *loss);
/* Adam training on the device: epochs over (x, y) with minibatch size batch,
 * learning rate lr, shuffling seed.  Returns the final training loss. */
int32_t flk_mlp_train(flk_ctx *ctx, const float *x, const float *y, int64_t n,
                      int32_t epochs, int32_t batch, float lr, uint64_t seed,
                      float *final_loss);
/* Node-pair training samples from the Barnes–Hut trees of curve pair (a, b):
 * up to cap samples of 12 features and the label log10 |exact - expansion|,
 * where exact is the direct sum over the two nodes' segments and expansion
 * the second-order far field.  Returns the number written in *count. */
int32_t flk_mlp_samples(flk_ctx *ctx, int32_t a, int32_t b, int64_t cap,
                        uint64_t seed, float *x, float *y, int64_t *count);

#ifdef __cplusplus
}
#endif

#endif /* FLK_H_ */
