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
#include <cstring>
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
    if (wire_f32) {
      const int64_t blocks = std::min<int64_t>((nval + 255) / 256, 64 * (int64_t)std::max(ctx->num_cus, 1));
      hipLaunchKernelGGL(widen_f32, dim3((unsigned)blocks), dim3(256), 0, s, land, nval, dst);
      FLK_HIP(hipGetLastError());
    }
    return (int64_t)esz * nval;
  }
  return 0;  // not reached
}

}  // namespace flk
