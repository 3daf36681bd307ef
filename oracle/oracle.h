/*
 * CPU ORACLE — test infrastructure only.
 *
 * An FP64 C++17 + OpenMP restatement of the reference's linking-number
 * algorithms (antequ/fastlinkingnumbers, "verifycurves").  It is the checker
 * for the HIP engine and the CPU baseline timed beside it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path (fastlinkingnumbers_amd/) never links, imports or calls it.
 *
 * Every function cites the reference file:line whose semantics it restates.
 * How each part is pinned (reference certificates, closed forms, brute force,
 * independent quadrature) is listed in tests/test_oracle_*.py.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Segment types, same order as SegmentType in include/curve.h:35. */
enum { ORC_BSPLINE = 0, ORC_CATMULLROM = 1, ORC_BEZIER = 2, ORC_POLYLINE = 3 };

typedef struct orc_model orc_model;

const char *orc_last_error(void);

/* ---- model (include/model.h, src/model.cpp) ---- */
orc_model *orc_model_new(void);
void orc_model_free(orc_model *m);
int orc_model_add_curve(orc_model *m, const double *xyz, int64_t n, int seg_type,
                        int looped);
int orc_model_load_bcc(orc_model *m, const char *path);      /* model.cpp:280-364 */
int orc_model_save_bcc(const orc_model *m, const char *path); /* model.cpp:367-452 */
int orc_model_load_obj(orc_model *m, const char *path, int seg_type); /* model.cpp:124-264 */
int32_t orc_model_ncurves(const orc_model *m);
int64_t orc_model_curve_size(const orc_model *m, int32_t c);
int orc_model_curve_info(const orc_model *m, int32_t c, int *seg_type, int *looped);
int orc_model_get_curve(const orc_model *m, int32_t c, double *xyz_out);

/* ---- curve geometry (src/curve.cpp) ---- */
int orc_point(const double *q, int64_t n, int seg_type, int looped, int64_t seg,
              double t, double *out3);                       /* curve.cpp:156-184 */
int orc_segment_box(const double *q, int64_t n, int seg_type, int looped,
                    int64_t seg, double tlo, double thi, double *out6); /* curve.cpp:293-324 */
int orc_curve_boxes(const orc_model *m, double *out6n);      /* curve.cpp:331-343 */

/* ---- potential-link search (src/potential_link_search.cpp:88-127) ---- */
/* Writes the sorted unique pair list (i<j) into pairs_out (2*cap ints); returns
 * the pair count, or -1 on error.  use_tree selects the kd-BVH self join the
 * reference uses for >=50 curves; both must give the same set. */
int64_t orc_pls(const orc_model *m, int use_tree, int32_t *pairs_out, int64_t cap);

/* ---- discretization (src/discretize.cpp:102-281) ---- */
int orc_discretize(const orc_model *in, orc_model *out);

/* ---- direct summation (src/compute_link.cpp:126-280) ---- */
/* p: n1 vertices of closed polyline 1 (xyz AoS), q: n2 vertices of polyline 2. */
double orc_ds_one_arctan(const double *p, int64_t n1, const double *q, int64_t n2,
                         int parallel);
double orc_ds_fast(const double *p, int64_t n1, const double *q, int64_t n2);
/* Same sum as orc_ds_one_arctan restricted to rows [row0,row1) of curve 1;
 * used for bounded CPU-baseline samples. */
double orc_ds_rows(const double *p, int64_t n1, const double *q, int64_t n2,
                   int64_t row0, int64_t row1, int parallel);
/* Per-segment-pair Gauss term (one-arctan form, compute_link.cpp:151-175) for
 * segment p0->p1 against q0->q1; returns the linking-number contribution. */
double orc_segment_pair_term(const double *p0, const double *p1, const double *q0,
                             const double *q1);
/* Barnes–Hut leaf term (compute_link.cpp:576-606) from midpoints and tangents. */
double orc_bh_leaf_term(const double *m1, const double *t1, const double *m2,
                        const double *t2);

/* The pair loop of ComputeLinkingNumbersDirectSum
 * (src/linking_number_certificate.cpp:263-293): with ncurves >= 6 OpenMP runs
 * over pairs and each pair uses the running-product sum (DirectFast), else each
 * pair uses the one-arctan sum parallel over rows.  Curves are xyz AoS with
 * offsets[ncurves+1]; out receives one sum per pair. */
int orc_ds_pairs(const double *xyz, const int64_t *offsets, int32_t ncurves,
                 const int32_t *pairs, int64_t npairs, double *out);

/* ---- Barnes–Hut (src/compute_link.cpp:294-934) ---- */
/* One closed polyline pair.  beta is the opening parameter (itolsq = beta^2),
 * parallel selects the BFS(<75k)+DFS path, serial the float-truncated DFS.
 * err receives the accumulated Eq. 14 estimate. visits (may be NULL) receives
 * the number of node pairs visited. */
double orc_bh_pair(const double *p, int64_t n1, const double *q, int64_t n2,
                   double beta, int parallel, double *err, int64_t *visits);
/* Moment record of 45 doubles: Q0(9) Q1(9) Q2(9) D(9) c(3) mid(3) norms(3),
 * 3x3 matrices row-major, matching compute_link.h:95-114. */
int orc_tree_moments(const double *p, int64_t n, double *moments_out,
                     double *boxes_out, int32_t *children_out, int32_t *leaf_seg_out);
double orc_far_field(const double *m1, const double *m2, double r1, double r2,
                     double epssq, double *err);

/* ---- certificate (src/linking_number_certificate.cpp:102-394) ---- */
/* Fills trip_out with (i, j, lk) triplets sorted by (i, j), i<j, and returns the
 * count (or -1).  mean_err / max_err receive the printed off-integer errors;
 * lk_out (may be NULL, cap entries) receives the raw per-candidate-pair sums in
 * candidate order with cand_out (2 ints each) holding the pair. */
int64_t orc_certificate(const orc_model *m, int force_direct_sum, int discretize,
                        double beta_init, double beta_limit, int32_t *trip_out,
                        int64_t cap, double *mean_err, double *max_err,
                        double *lk_out, int32_t *cand_out, int64_t cand_cap,
                        int64_t *ncand);

#ifdef __cplusplus
}
#endif
