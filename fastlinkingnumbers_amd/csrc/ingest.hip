// Host-to-device ingest of curve coordinates.
//
// Coordinates arrive as FP64 (the reference keeps every curve as
// Eigen::Vector3d control points, include/curve.h:184) or as FP32 (the BCC
// payload, src/model.cpp:299-304, and the FP32 synthetic workloads).  They
// cross PCIe through pinned staging buffers: worker threads fill the buffers
// while the DMA of the buffers filled before runs, and FP64 input is narrowed
// to FP32 on the way whenever every value converts exactly -- the case for any
// model whose coordinates came from FP32 data, for which it halves the bytes
// sent.  One value that does not convert exactly sends the whole array as
// FP64 instead.  On the device a widening kernel writes the FP64 array the
// rest of the engine reads, so both routes give identical bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>

#include "flk_internal.h"

namespace flk {
namespace {

__global__ __launch_bounds__(256) void widen_f32(const float *__restrict__ in, int64_t n,
                                                 double *__restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (double)in[i];
}

// Narrows src[0..n) into dst; returns false if some value changes.  NaN never
// compares equal, so a NaN keeps the FP64 route (which carries it unchanged).
bool narrow(const double *src, int64_t n, float *dst) {
  bool ok = true;
  for (int64_t i = 0; i < n; ++i) {
    const float f = (float)src[i];
    ok &= (double)f == src[i];
    dst[i] = f;
  }
  return ok;
}

}  // namespace

void HostStage::ensure(int workers) {
  const size_t want = 2 * (size_t)workers;
  while (buf.size() < want) {
    void *p = nullptr;
    FLK_HIP(hipHostMalloc(&p, kChunkBytes, hipHostMallocDefault));
    hipEvent_t e = nullptr;
    const hipError_t rc = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (rc != hipSuccess) {
      (void)hipHostFree(p);
      FLK_HIP(rc);
    }
    buf.push_back(p);
    done.push_back(e);
    pending.push_back(0);
  }
}

void HostStage::release() {
  for (size_t i = 0; i < buf.size(); ++i) {
    if (pending[i]) (void)hipEventSynchronize(done[i]);
    (void)hipEventDestroy(done[i]);
    (void)hipHostFree(buf[i]);
  }
  buf.clear();
  done.clear();
  pending.clear();
}

void widen_launch(flk_ctx *ctx, const float *land, int64_t nval, double *dst) {
  if (nval <= 0) return;
  const int64_t blocks = std::min<int64_t>((nval + 255) / 256, 64 * (int64_t)std::max(ctx->num_cus, 1));
  hipLaunchKernelGGL(widen_f32, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, land, nval, dst);
  FLK_HIP(hipGetLastError());
}

int ingest_workers(const flk_ctx *ctx) {
  if (ctx->tune.ingest_threads > 0) return ctx->tune.ingest_threads;
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::min(8u, std::max(1u, hw / 2));
}

int64_t ingest_points(flk_ctx *ctx, const void *src, bool src_f32, int64_t n, double *dst, float *land,
                      size_t land_bytes) {
  if (n <= 0) return 0;
  hipStream_t s = ctx->stream;
  HostStage &st = ctx->stage;
  const int64_t nval = 3 * n;
  if (land_bytes < sizeof(float) * (size_t)nval) throw Error(FLK_ERR_STATE, "ingest landing buffer too small");
  // one pass with FP32 on the wire (FP32 input, or FP64 input narrowed), and
  // an FP64 pass if the narrowing met a value that does not fit
  for (int pass = src_f32 ? 1 : 0; pass < 2; ++pass) {
    const bool wire_f32 = src_f32 || pass == 0;
    const size_t esz = wire_f32 ? sizeof(float) : sizeof(double);
    const int64_t per = (int64_t)(HostStage::kChunkBytes / esz);
    const int64_t nchunks = (nval + per - 1) / per;
    const int workers = (int)std::min<int64_t>(ingest_workers(ctx), nchunks);
    st.ensure(workers);
    std::atomic<bool> inexact{false};
    std::atomic<int> hip_err{(int)hipSuccess};
    auto work = [&](int t) {
      int slot = 0;
      for (int64_t k = t; k < nchunks; k += workers) {
        if (inexact.load(std::memory_order_relaxed) || hip_err.load(std::memory_order_relaxed) != hipSuccess)
          return;
        const size_t b = 2 * (size_t)t + slot;
        slot ^= 1;
        if (st.pending[b]) {
          const hipError_t rc = hipEventSynchronize(st.done[b]);
          st.pending[b] = 0;
          if (rc != hipSuccess) {
            hip_err = (int)rc;
            return;
          }
        }
        const int64_t v0 = k * per, v1 = std::min(nval, v0 + per);
        if (src_f32) {
          std::memcpy(st.buf[b], static_cast<const float *>(src) + v0, sizeof(float) * (v1 - v0));
        } else if (wire_f32) {
          if (!narrow(static_cast<const double *>(src) + v0, v1 - v0, static_cast<float *>(st.buf[b]))) {
            inexact = true;
            return;
          }
        } else {
          std::memcpy(st.buf[b], static_cast<const double *>(src) + v0, sizeof(double) * (v1 - v0));
        }
        void *to = wire_f32 ? (void *)(land + v0) : (void *)(dst + v0);
        hipError_t rc = hipMemcpyAsync(to, st.buf[b], esz * (v1 - v0), hipMemcpyHostToDevice, s);
        if (rc == hipSuccess) rc = hipEventRecord(st.done[b], s);
        if (rc != hipSuccess) {
          hip_err = (int)rc;
          return;
        }
        st.pending[b] = 1;
      }
    };
    if (workers == 1) {
      work(0);
    } else {
      std::vector<std::thread> pool;
      pool.reserve(workers);
      for (int t = 0; t < workers; ++t) pool.emplace_back(work, t);
      for (auto &th : pool) th.join();
    }
    if (hip_err != (int)hipSuccess) FLK_HIP((hipError_t)hip_err.load());
    if (inexact) {
      // the staged FP32 copies still in flight only read pinned buffers and
      // write the landing area; wait for them before reusing either
      FLK_HIP(hipStreamSynchronize(s));
      for (auto &p : st.pending) p = 0;
      continue;
    }
    if (wire_f32) widen_launch(ctx, land, nval, dst);
    return (int64_t)esz * nval;
  }
  return 0;  // not reached
}

// BCC file (reference src/model.cpp:266-364): a 64-byte header, then per
// curve an int32 control-point count (negative: closed loop) and that many
// FP32 xyz triples.  The coordinates are read straight into the pinned
// staging buffers, the counts stripped on the way, and each full buffer is
// sent while the next one is read; once the counts are known the curve set is
// laid out and the FP32 array widened on the device.
void load_bcc(flk_ctx *ctx, const char *path, BccLoad &out) {
  out = BccLoad{};
  std::FILE *f = std::fopen(path, "rb");
  if (!f) throw Error(FLK_ERR_ARG, std::string("Error opening input file: ") + path);
  struct Closer {
    std::FILE *f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  auto fail = [&](const char *why) { throw Error(FLK_ERR_ARG, std::string(why) + ": " + path); };
#pragma pack(push, 1)
  struct Header {
    char sign[3];
    unsigned char bytes;
    char type[2];
    char dims;
    char up;
    uint64_t ncurves;
    uint64_t npoints;
    char info[40];
  } h;
#pragma pack(pop)
  static_assert(sizeof(Header) == 64, "BCC header is 64 bytes");
  if (std::fread(&h, sizeof h, 1, f) != 1) fail("Incomplete header");
  if (h.sign[0] != 'B' || h.sign[1] != 'C' || h.sign[2] != 'C') fail("Invalid header");
  if (h.bytes != 0x44) fail("Invalid bytecount. We only support 32-bit integers and floats.");
  if (h.dims != 3) fail("Invalid dimensions. We only support data in 3D Cartesian space.");
  if (h.type[0] == 'C' && h.type[1] == '0') out.type = FLK_SEG_CATMULLROM;
  else if (h.type[0] == 'B' && h.type[1] == 'S') out.type = FLK_SEG_BSPLINE;
  else if (h.type[0] == 'B' && h.type[1] == 'Z') out.type = FLK_SEG_BEZIER;
  else if (h.type[0] == 'P' && h.type[1] == 'L') out.type = FLK_SEG_POLYLINE;
  else fail("Invalid valid curve segment type. We only support C0, BS, BZ, and PL.");
  if (h.ncurves > (uint64_t)INT32_MAX - 1) fail("Too many curves");
  const int32_t nc = (int32_t)h.ncurves;
  // the coordinates cannot exceed what the file holds after the counts
  if (std::fseek(f, 0, SEEK_END) != 0) fail("Cannot read file size");
  const long long fsize = std::ftell(f);
  if (std::fseek(f, (long)sizeof h, SEEK_SET) != 0 || fsize < 0) fail("Cannot read file size");
  const long long coord_bytes = fsize - (long long)sizeof h - 4LL * nc;
  if (coord_bytes < 0) fail("Cannot read curve control point count");
  hipStream_t s = ctx->stream;
  ctx->bcc_land.reserve(std::max<size_t>((size_t)coord_bytes, 16));
  char *land = static_cast<char *>(ctx->bcc_land.ptr);
  HostStage &st = ctx->stage;
  st.ensure(1);
  int b = 0;
  size_t fill = 0;
  int64_t sent = 0;  // bytes handed to DMA so far
  auto flush = [&]() {
    if (fill == 0) return;
    FLK_HIP(hipMemcpyAsync(land + sent, st.buf[b], fill, hipMemcpyHostToDevice, s));
    FLK_HIP(hipEventRecord(st.done[b], s));
    st.pending[b] = 1;
    sent += (int64_t)fill;
    fill = 0;
    b ^= 1;
    if (st.pending[b]) {  // the other buffer's DMA must end before it is refilled
      FLK_HIP(hipEventSynchronize(st.done[b]));
      st.pending[b] = 0;
    }
  };
  if (st.pending[b]) {
    FLK_HIP(hipEventSynchronize(st.done[b]));
    st.pending[b] = 0;
  }
  out.offsets.assign(1, 0);
  out.looped.assign(nc, 0);
  for (int32_t c = 0; c < nc; ++c) {
    int32_t cnt;
    if (std::fread(&cnt, sizeof cnt, 1, f) != 1) fail("Cannot read curve control point count");
    out.looped[c] = cnt < 0 ? 1 : 0;
    const int64_t n = cnt < 0 ? -(int64_t)cnt : (int64_t)cnt;
    if (12 * n > coord_bytes - (sent + (int64_t)fill)) fail("Invalid number of control points read");
    size_t left = (size_t)(12 * n);
    while (left > 0) {
      const size_t take = std::min(left, HostStage::kChunkBytes - fill);
      if (std::fread(static_cast<char *>(st.buf[b]) + fill, 1, take, f) != take)
        fail("Invalid number of control points read");
      fill += take;
      left -= take;
      if (fill == HostStage::kChunkBytes) flush();
    }
    out.offsets.push_back(out.offsets.back() + n);
  }
  flush();
  out.points = out.offsets.back();
  // the curve set: layout from the counts, then the FP32 array widened
  try {
    curves_layout(ctx, out.offsets.data(), nc);
    widen_launch(ctx, reinterpret_cast<const float *>(land), 3 * out.points, ctx->curves.raw.as<double>());
    curves_finish(ctx);
  } catch (...) {
    curves_clear(ctx);
    throw;
  }
}

}  // namespace flk
