// Internal declarations of the MI355X linking-number engine (libflk.so).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <memory>
#include <vector>

#include "flk.h"

#define FLK_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw flk::Error(FLK_ERR_HIP, std::string(#call) + ": " +                \
                                        hipGetErrorString(e_));                \
  } while (0)

namespace flk {

struct Error {
  int32_t code;
  std::string msg;
  Error(int32_t c, std::string m) : code(c), msg(std::move(m)) {}
};

// Kernel launch that optionally carries start/stop events on the dispatch
// itself (hipExtLaunchKernelGGL); without events it is a plain launch.
template <typename K, typename... Args>
inline void launch_timed(hipEvent_t start, hipEvent_t stop, K kernel, dim3 grid, dim3 block, hipStream_t s,
                         Args... args) {
  if (start && stop) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, start, stop, 0, args...);
  else hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

// Allocations and frees of a tracked DevBuf (the Barnes–Hut scratch and the
// learned criterion's weights, which a certificate graph holds pointers into)
// bump this counter: such a graph is replayed only while it is unchanged.
inline std::atomic<uint64_t> g_devbuf_epoch{0};
// Set while this thread captures a graph: a DevBuf that would allocate then
// throws instead (every buffer is reserved before the capture begins).
inline thread_local bool t_capturing = false;

// Device allocation that grows on demand and is freed with its owner.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  bool tracked = false;  // (re)allocation bumps g_devbuf_epoch
  DevBuf() = default;
  explicit DevBuf(bool track) : tracked(track) {}
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes), tracked(o.tracked) {
    o.ptr = nullptr;
    o.bytes = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      bytes = o.bytes;
      o.ptr = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void reserve(size_t n) {
    if (n <= bytes) return;
    if (t_capturing) throw Error(FLK_ERR_STATE, "device allocation during graph capture");
    release();
    if (n == 0) return;
    FLK_HIP(hipMalloc(&ptr, n));
    bytes = n;
    if (tracked) ++g_devbuf_epoch;
  }
  template <typename T> T *as() const { return static_cast<T *>(ptr); }
  void release() {
    if (ptr) {
      (void)hipFree(ptr);
      if (tracked) ++g_devbuf_epoch;
    }
    ptr = nullptr;
    bytes = 0;
  }
};

// Pinned host allocation (grow-only): the end of a graph's read-back copies,
// which then run as DMA without staging through pageable memory.
struct HostBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  HostBuf() = default;
  HostBuf(const HostBuf &) = delete;
  HostBuf &operator=(const HostBuf &) = delete;
  HostBuf(HostBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) {
    o.ptr = nullptr;
    o.bytes = 0;
  }
  HostBuf &operator=(HostBuf &&o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      bytes = o.bytes;
      o.ptr = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~HostBuf() { release(); }
  void reserve(size_t n) {
    if (n <= bytes) return;
    if (t_capturing) throw Error(FLK_ERR_STATE, "host allocation during graph capture");
    release();
    if (n == 0) return;
    FLK_HIP(hipHostMalloc(&ptr, n, 0));
    bytes = n;
  }
  template <typename T> T *as() const { return static_cast<T *>(ptr); }
  void release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
};

// One segment-pair work description for the direct-sum tile kernel: rows
// [row0, row0 + 64*R) of the row curve against columns [col0, col0 + ncols)
// of the column curve.
struct DsPair {
  int64_t off_r, off_c;  // first vertex of each curve in the vertex arrays
  int32_t n_r, n_c;      // vertex (= segment) counts
};
struct DsTask {
  int32_t pair;
  int32_t row0;
  int32_t col0;
  int32_t ncols;
};

// Barnes–Hut node record (FP32, 160 bytes): centre, radius, monopole c, dipole
// D (row-major D[k][j]), quadrupole Q[k] symmetric in (i,j) stored as
// xx xy xz yy yz zz per k, norms, children and segment range.
struct BhNode {
  float ctr[3];
  float radius;
  float c[3];
  float cn;
  float D[9];
  float dn;
  float Q[18];
  float qn;
  int32_t child[2];     // >= 0 internal node index; < 0: leaf bucket -(k+1)
  int32_t seg0, nseg;   // contiguous range in the tree's segment order
  float rsq;            // radius squared as the reference computes it
  float pad[2];
};

struct Curves {
  int32_t ncurves = 0;
  std::vector<int64_t> off;      // host offsets into the device vertex arrays (n+1 per curve)
  std::vector<int32_t> n;        // vertices per curve
  std::vector<double> box;       // ncurves x 6, FP64, in input coordinates
  double origin[3] = {0, 0, 0};  // input = origin + scale^-1 * stored
  double scale = 1.0;            // power of two
  std::vector<int64_t> in_off;   // input offsets
  DevBuf raw;                    // input FP64 AoS (no closing vertex)
  DevBuf dpts;                   // double4 per vertex, shifted+scaled, +1 closing vertex per curve
  DevBuf hi, lo, ed;             // float4 per vertex (+1 closing vertex per curve)
  DevBuf d_off;                  // int64 device offsets (ncurves + 1)
  int64_t nvert_dev = 0;
  bool inexact = false;          // some stored coordinate needs the FP32 lo part
  int64_t bh_budget = 0;         // Barnes–Hut batch size in segments, fixed per curve set
};

// Pinned host staging buffers of the coordinate upload, two per worker
// thread, each with an event marking the end of its last DMA.
struct HostStage {
  static constexpr size_t kChunkBytes = size_t(8) << 20;
  std::vector<void *> buf;
  std::vector<hipEvent_t> done;
  std::vector<uint8_t> pending;  // a DMA from buf[i] may still run
  void ensure(int workers);
  void release();
};

// Barnes–Hut device scratch, kept by the context between calls (grow-only)
// so repeated certificates do not pay hipMalloc/hipFree for the trees, their
// FP64 build records and the traversal queues.
struct BhScratch {
  // forest
  DevBuf geo, mom, order, m64, range, child, flag, flag2, scan, keys0, keys1, ids0, tmp, seg, vtx, box, sp, rtiles, rcurve;
  // traversal: root items of a pass, re-run list, per-pair sums and results
  DevBuf roots_a, roots_b, beta2, cnt, leaves, pred, sum, err, items0, redo, outv, oute, outf;
  std::vector<std::unique_ptr<DevBuf>> levels;
  unsigned long long *host_cnt = nullptr;  // pinned, 2 words per level of a traversal round trip
  int64_t dev_cap = 0;                     // items per level buffer of the device-scheduled pass
  std::vector<unsigned long long> dev_counts;  // its per-level counts, read back once per pass
  BhScratch() {
    each([](DevBuf &b) { b.tracked = true; });
  }
  DevBuf &add_level() {
    levels.push_back(std::make_unique<DevBuf>(true));
    return *levels.back();
  }
  template <typename Fn> void each(Fn fn) {
    DevBuf *all[] = {&geo, &mom, &order, &m64, &range, &child, &flag, &flag2, &scan, &keys0, &keys1, &ids0,
                     &tmp, &seg, &vtx, &box, &sp, &rtiles, &rcurve, &roots_a, &roots_b, &beta2, &cnt, &leaves, &pred,
                     &sum, &err, &items0, &redo, &outv, &oute, &outf};
    for (DevBuf *b : all) fn(*b);
    for (auto &b : levels) fn(*b);
  }
  size_t bytes() {
    size_t n = 0;
    each([&](DevBuf &b) { n += b.bytes; });
    return n;
  }
  void release() {
    each([](DevBuf &b) { b.release(); });
    levels.clear();
    if (host_cnt) (void)hipHostFree(host_cnt);
    host_cnt = nullptr;
  }
};

// A captured direct-sum batch (uploads, tile kernel, per-pair reduction,
// read-back) for replay when the same pairs come again with the same shape
// and curves.  It owns its device buffers and host staging, so reallocations
// elsewhere cannot leave dangling pointers in the graph.
struct DsGraph {
  uint64_t key = 0;
  hipGraphExec_t exec = nullptr;
  int64_t work = 0;
  int32_t waves = 0;
  int64_t npairs = 0;
  int R = 2, blocks = 0;          // the plan's launch shape and precision
  bool fp64 = false, precise = false;
  std::vector<DsPair> dp;
  std::vector<DsTask> tasks;
  std::vector<int64_t> prefix, task0;
  std::vector<double> out;
  DevBuf d_pairs, d_tasks, d_cost, d_theta, d_corr, d_task0, d_out;
  HostBuf h_out;  // pinned end of the read-back (copied into out)
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key = 0;
  }
  ~DsGraph() { reset(); }
};

// Result of a device discretization (context-owned).
// A BCC file read into the context's curve set: segment type, per-curve
// offsets and loop flags.
struct BccLoad {
  int32_t type = 0;
  int64_t points = 0;
  std::vector<int64_t> offsets;
  std::vector<uint8_t> looped;
};

struct DiscResult {
  int64_t segments = 0, retired = 0;
  int32_t rounds = 0;
  std::vector<int64_t> overlaps;     // overlapping piece pairs per round
  std::vector<int64_t> out_offsets;  // ncurves + 1, into xyz / 3
  std::vector<double> xyz;           // polyline vertices, AoS (when copied to the host)
  int32_t small_curve = -1, small_segment = -1;  // the input guard's smallest segment
  double small_size = 0.0;
};

struct MlpState {
  std::vector<float> params;  // FLK_MLP_NPARAM FP32 masters
  std::vector<float> norm;    // mean[12], std[12]
  bool loaded = false;
  DevBuf d_params{true}, d_norm{true}, d_bf16;
  DevBuf d_image{true};  // the level kernel's LDS image of the network (mlp_upload)
};

// A whole Barnes–Hut or hybrid certificate captured into one graph (bh.hip).
struct CertGraph;

}  // namespace flk

struct flk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  // a captured hybrid certificate's direct-sum branch (default priority)
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int num_cus = 0;
  std::string name;
  std::string err;
  flk::Curves curves;
  flk_stats stats{};
  std::vector<int32_t> pls_pairs;
  flk::MlpState mlp;
  flk::BhScratch bh;
  flk::HostStage stage;
  // default shapes: measured best on the 1,600-loop chain mail (R = 2 with
  // every block resident: 880 ms against 931 ms for R = 4, 1,068 ms for R = 8)
  flk_tune tune{2, 0, 8, 0, 0, 0, {0, 0}};
  flk::DsGraph ds_graph[2];      // FP32, FP64 batches
  std::shared_ptr<flk::CertGraph> cert_graph;  // last Barnes–Hut certificate graph
  flk::DiscResult disc;          // last flk_discretize result
  flk::BccLoad bcc;              // last flk_load_bcc result
  flk::DevBuf bcc_land;          // FP32 landing area of flk_load_bcc
  uint64_t curves_version = 0;   // bumped by every flk_set_curves
  void *noalloc_scratch = nullptr;  // this device's fixed direct-sum array (ds.hip)
  // scratch
  flk::DevBuf d_pairs, d_tasks, d_cost, d_task_theta, d_task_corr, d_pair_task0,
      d_pair_out, d_pair_corr, d_boxes, d_pls, d_count, d_misc;
};

namespace flk {

// Direct sum over a set of curve pairs.  theta_out[k] receives the sum of the
// per-segment-pair half solid angles (radians) and corr_out[k] the integer
// branch corrections; the linking number is theta/(2 pi) + corr.
void ds_run(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, bool fp64,
            std::vector<double> &lk_out, float *ms_kernel, bool use_graph = false);
// Direct-sum kernel time (ms) for every `stride`-th 64R-row block of the
// pairs' work with launch shape `t` (FP32); *work receives the segment pairs
// of the sample.  Results are discarded.
float ds_time_sample(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, const flk_tune &t, int64_t stride,
                     int64_t *work);
// Segment pairs of the pairs' direct-sum work.
int64_t ds_work(flk_ctx *ctx, const int32_t *pairs, int64_t npairs);
// A direct-sum batch recorded inside a larger graph: ds_capture_prepare plans
// the pairs into G's host arrays, reserves G's buffers and uploads the plan
// (call it before the capture begins); ds_capture_enqueue enqueues the tile
// kernel, the per-pair reduction and the read-back into G.h_out, which the
// caller copies into G.out after the launch.  share: the fraction of the
// tile kernel's resident blocks per CU its grid takes (below 1 when it runs
// beside other kernels; a tuned ds_blocks_per_cu takes precedence).
void ds_capture_prepare(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, bool fp64, DsGraph &G,
                        double share = 1.0);
void ds_capture_enqueue(flk_ctx *ctx, DsGraph &G, hipStream_t s);

// Device-wide exclusive prefix sum (sort.hip); in and out may alias; temp
// holds scan_temp_bytes(n).  Kernels only: capturable into a graph.
size_t scan_temp_bytes(int64_t n);
void exclusive_scan(const int32_t *in, int32_t *out, int64_t n, int32_t *temp, hipStream_t s);
// Stable LSD radix sort of (key, value) pairs on bits [0, end_bit): keys and
// vals are the input (overwritten), keys_alt and vals_alt scratch of n
// entries, vals_out (distinct from vals and vals_alt) receives the values in
// key order; the sorted keys are not kept.  temp holds
// radix_temp_bytes(n, table_tiles).  With a device tile table (from
// segmented_tiles) the sort is segmented: every key stays within its segment
// of the input, sorted there.  Kernels only: capturable into a graph.
struct RadixTile {
  int32_t start, count;  // keys [start, start + count), count <= 4096
  int32_t cbase, stride; // digit d's count at cbase + d * stride
};
size_t radix_temp_bytes(int64_t n, int64_t table_tiles = 0);
// Tiles of the segments [seg_off[c], seg_off[c+1]) (host table); returns
// their number.
int64_t segmented_tiles(const int64_t *seg_off, int64_t nseg, std::vector<RadixTile> &table);
template <typename K>
void radix_sort_pairs(K *keys, K *keys_alt, uint32_t *vals, uint32_t *vals_alt, uint32_t *vals_out, int64_t n,
                      int end_bit, void *temp, hipStream_t s, const RadixTile *table = nullptr,
                      int64_t table_tiles = 0);
// Device address of this GPU's fixed direct-sum array (looked up once per
// device; call with the device current).
void *noalloc_scratch_init(int device);
// ds_run launched eagerly with every device array carved from the fixed
// array: no device allocation (limits FLK_NOALLOC_MAX_PAIRS / _TASKS).
void ds_run_noalloc(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, bool fp64, std::vector<double> &lk_out,
                    float *ms_kernel);

// Spline-to-polyline discretization of the given curves against the given
// potentially linked pairs (see disc.hip); the polylines become the context's
// curve set, and res.xyz a host copy of them when host_copy is set.
void discretize_run(flk_ctx *ctx, const double *ctrl, const int64_t *offsets, const int32_t *types,
                    const uint8_t *looped, int32_t nc, const int32_t *pairs, int64_t npairs, DiscResult &res,
                    bool host_copy);

void bh_run(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, const flk_params &prm,
            std::vector<double> &lk, std::vector<double> &err, std::vector<uint32_t> &flags);

// The Barnes–Hut tree of one curve (as bh_run builds it), copied to the host
// in input coordinates, and the curve segment at each position of its order.
void tree_export(flk_ctx *ctx, int32_t curve, int32_t leaf, std::vector<flk_tree_node> &nodes,
                 std::vector<int32_t> &seg_order);

void pls_run(flk_ctx *ctx, const double *boxes, int32_t ncurves, std::vector<int32_t> &out);

void curve_boxes_run(flk_ctx *ctx, std::vector<double> &boxes);

// A new curve set: validates the offsets and sizes the device arrays; the
// caller then fills curves.raw (input FP64 AoS) and calls curves_finish, which
// computes the boxes, origin and scale and the staged vertex arrays.  Returns
// the vertex count.
int64_t curves_layout(flk_ctx *ctx, const int64_t *offsets, int32_t ncurves);
void curves_finish(flk_ctx *ctx);
// Resets to the empty curve set (after a failure part-way through a new one).
void curves_clear(flk_ctx *ctx);

// Copies n host xyz triples (FP32, or FP64 sent as FP32 when every value fits
// exactly) to dst (device FP64 AoS) through the pinned staging buffers; land
// is device scratch for the FP32 copy (>= 12 n bytes).  Enqueued on the
// context's stream; the host array is no longer read on return.  Returns the
// bytes sent.
int64_t ingest_points(flk_ctx *ctx, const void *src, bool src_f32, int64_t n, double *dst, float *land,
                      size_t land_bytes);
// dst[i] = (double)land[i], i < nval, on the context's stream.
void widen_launch(flk_ctx *ctx, const float *land, int64_t nval, double *dst);

// Reads a BCC file into the context's curve set (see ingest.hip).
void load_bcc(flk_ctx *ctx, const char *path, BccLoad &out);
void convert_vertices_launch(hipStream_t s, const double4 *d, int64_t n, float4 *hi, float4 *lo,
                             float4 *ed, int32_t *inexact);

int64_t mlp_samples(flk_ctx *ctx, int32_t a, int32_t b, int64_t cap, uint64_t seed, float *x,
                    float *y);
void mlp_forward(flk_ctx *ctx, const float *x, int64_t n, float *y);
void mlp_upload(flk_ctx *ctx);
// Forward pass on device buffers (features n x 12, predictions n).
void mlp_forward_device(flk_ctx *ctx, const float *dx, int64_t n, float *dy);
float mlp_gradient(flk_ctx *ctx, const float *x, const float *y, int64_t n, float *grad);
float mlp_train(flk_ctx *ctx, const float *x, const float *y, int64_t n, int epochs, int batch,
                float lr, uint64_t seed);

}  // namespace flk
