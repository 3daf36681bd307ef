// Device-wide prefix sum and stable LSD radix sort for gfx950, written so a
// caller can capture them into a HIP graph: no memsets, no host reads, no
// allocation and no waits between blocks (every block finishes on its own).
//
// Prefix sum: three launches over tiles of 4,096 values (per-tile totals, one
// block scanning the totals, per-tile scans with their offsets), or one block
// for short inputs.
//
// Radix sort: 8-bit digits, least significant first, tiles of 4,096 keys
// (or, segmented, tiles that never straddle a segment boundary).
// Per digit pass: a count kernel writes every tile's digit histogram into a
// digit-major array (digit d of tile t at d * tiles + t), the prefix sum turns
// it into each (digit, tile) run's first output position, and a scatter
// kernel ranks its tile's keys -- each wave walks 64 consecutive keys per step
// and ranks equal digits with eight ballots, in key order, so the sort is
// stable -- reorders the tile by digit in LDS and writes each digit's run
// contiguously.  The last pass writes only the values (the caller wants the
// permutation, not the sorted keys).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "flk_internal.h"

namespace flk {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;      // values per block tile
constexpr int64_t kSingleBlockMax = 4 * kTile;  // shorter inputs: one block

__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

__device__ __forceinline__ int wave_inclusive(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

// Exclusive scan of one value per thread over the 256-thread block; *total
// receives the block's sum.  sw: 4 ints of LDS.  Contains block barriers.
__device__ __forceinline__ int block_exclusive(int v, int *total, int *sw) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int inc = wave_inclusive(v);
  if (lane == 63) sw[w] = inc;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kBlock / 64; ++k) {
    const int t = sw[k];
    off += k < w ? t : 0;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// Sum of each 4,096-value tile.
__global__ __launch_bounds__(kBlock) void scan_tile_sums(const int32_t *__restrict__ in, int64_t n,
                                                         int32_t *__restrict__ part) {
  __shared__ int sw[4];
  const int64_t base = (int64_t)blockIdx.x * kTile;
  int v = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int64_t g = base + i * kBlock + threadIdx.x;
    if (g < n) v += in[g];
  }
  int tot = 0;
  (void)block_exclusive(v, &tot, sw);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// A tile of m <= kTile values into LDS (zeros past m), every load of a
// thread issued before the first is waited for.
__device__ __forceinline__ void tile_to_lds(const int32_t *in, int m, int32_t *buf) {
  int v[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int k = i * kBlock + threadIdx.x;
    v[i] = k < m ? in[k] : 0;
  }
#pragma unroll
  for (int i = 0; i < kItems; ++i) buf[i * kBlock + threadIdx.x] = v[i];
}

// Exclusive scan of the LDS tile in place, plus carry; *tot receives the
// tile's sum.  Each thread owns 16 consecutive values (four 16-byte LDS
// accesses).  Contains block barriers.
__device__ __forceinline__ void scan_lds_tile(int32_t *buf, int carry, int *tot, int *sw) {
  int4 *b4 = reinterpret_cast<int4 *>(buf) + threadIdx.x * (kItems / 4);
  int v[kItems], s = 0;
#pragma unroll
  for (int q = 0; q < kItems / 4; ++q) {
    const int4 x = b4[q];
    v[4 * q] = x.x;
    v[4 * q + 1] = x.y;
    v[4 * q + 2] = x.z;
    v[4 * q + 3] = x.w;
  }
#pragma unroll
  for (int i = 0; i < kItems; ++i) s += v[i];
  int run = carry + block_exclusive(s, tot, sw);
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int x = v[i];
    v[i] = run;
    run += x;
  }
#pragma unroll
  for (int q = 0; q < kItems / 4; ++q) b4[q] = make_int4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
}

__device__ __forceinline__ void lds_to_tile(const int32_t *buf, int m, int32_t *out) {
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int k = i * kBlock + threadIdx.x;
    if (k < m) out[k] = buf[k];
  }
}

// One block, exclusive scan of in[0, n) into out (may alias), tile by tile
// with a running carry.
__global__ __launch_bounds__(kBlock) void scan_one_block(const int32_t *in, int32_t *out, int64_t n) {
  __shared__ int sw[4];
  __shared__ int4 buf4[kTile / 4];
  int32_t *buf = reinterpret_cast<int32_t *>(buf4);
  int carry = 0;
  for (int64_t c0 = 0; c0 < n; c0 += kTile) {
    const int m = (int)min<int64_t>(kTile, n - c0);
    tile_to_lds(in + c0, m, buf);
    __syncthreads();
    int tot = 0;
    scan_lds_tile(buf, carry, &tot, sw);
    __syncthreads();
    lds_to_tile(buf, m, out + c0);
    __syncthreads();  // buf is refilled by the next tile
    carry += tot;
  }
}

// Exclusive scan of each tile (in and out may alias), offset by its tile's
// scanned total.
__global__ __launch_bounds__(kBlock) void scan_tiles(const int32_t *in, int32_t *out, int64_t n,
                                                     const int32_t *__restrict__ part) {
  __shared__ int sw[4];
  __shared__ int4 buf4[kTile / 4];
  int32_t *buf = reinterpret_cast<int32_t *>(buf4);
  const int64_t base = (int64_t)blockIdx.x * kTile;
  const int m = (int)min<int64_t>(kTile, n - base);
  tile_to_lds(in + base, m, buf);
  __syncthreads();
  int tot = 0;
  scan_lds_tile(buf, part[blockIdx.x], &tot, sw);
  __syncthreads();
  lds_to_tile(buf, m, out + base);
}

// A block's tile: its first key, key count, and where its digit counts go
// (digit d at cbase + d * stride).  Without a tile table the tiles are the
// consecutive 4,096-key pieces of the input and the counts digit-major over
// all tiles; a segmented sort passes a table whose tiles never straddle a
// segment, with each segment's counts in a block of their own (digit-major
// within it), so the prefix sum keeps every key inside its segment.
__device__ __forceinline__ RadixTile tile_of(const RadixTile *__restrict__ table, int64_t n, int64_t tiles) {
  if (table) return table[blockIdx.x];
  const int64_t start = (int64_t)blockIdx.x * kTile;
  return RadixTile{(int32_t)start, (int32_t)min<int64_t>(kTile, n - start), (int32_t)blockIdx.x, (int32_t)tiles};
}

// Digit histogram of one tile: per-wave LDS counters, summed into
// counts[cbase + d * stride].  A wave reads 64 consecutive keys per step and
// counts runs of equal digits in consecutive lanes with one LDS add per run:
// keys that arrive nearly sorted (the high digits of Morton codes along a
// curve) would otherwise serialise on one counter.
template <typename K>
__global__ __launch_bounds__(kBlock) void radix_count(const K *__restrict__ keys, int64_t n, int shift,
                                                      const RadixTile *__restrict__ table, int64_t tiles,
                                                      int32_t *__restrict__ counts) {
  __shared__ int hist[kBlock / 64][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 256 * (kBlock / 64); i += kBlock) (&hist[0][0])[i] = 0;
  __syncthreads();
  const RadixTile tl = tile_of(table, n, tiles);
  const uint64_t above = lane == 63 ? 0ull : (~0ull << (lane + 1));
#pragma unroll 4
  for (int i = 0; i < kItems; ++i) {
    const int o = i * kBlock + (int)threadIdx.x;
    const bool ok = o < tl.count;
    const int d = ok ? (int)((keys[(int64_t)tl.start + o] >> shift) & 255) : -1;
    const int dp = __shfl_up(d, 1, 64);
    // a run starts at lane 0 and wherever the digit changes; invalid lanes
    // (the tile's tail, always the last lanes) form a run that is not added
    const uint64_t starts = __ballot(lane == 0 || d != dp);
    const uint64_t later = starts & above;
    const int next = later ? __builtin_ctzll(later) : 64;
    if (ok && (lane == 0 || d != dp)) atomicAdd(&hist[w][d], next - lane);
  }
  __syncthreads();
  const int d = threadIdx.x;
  counts[(int64_t)tl.cbase + (int64_t)d * tl.stride] = hist[0][d] + hist[1][d] + hist[2][d] + hist[3][d];
}

// Stable scatter of one tile by the digit at `shift`.  offsets[cbase + d *
// stride] is the first output position of the tile's keys with digit d.  Wave w ranks
// keys [w * 1024, (w + 1) * 1024) of the tile, 64 per step in key order: the
// lanes holding the same digit find each other with eight ballots, take the
// wave's running count of that digit, and the lowest of them advances it.
template <typename K, bool kLast>
__global__ __launch_bounds__(kBlock) void radix_scatter(const K *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                        int64_t n, int shift, const RadixTile *__restrict__ table,
                                                        int64_t tiles, const int32_t *__restrict__ offsets,
                                                        K *__restrict__ kout, uint32_t *__restrict__ vout) {
  constexpr int kWaves = kBlock / 64;
  __shared__ int wcnt[kWaves][256];
  __shared__ int dstart[256];
  __shared__ int gbase[256];
  __shared__ int sw[4];
  __shared__ K skey[kTile];
  __shared__ uint32_t sval[kTile];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 256 * kWaves; i += kBlock) (&wcnt[0][0])[i] = 0;
  __syncthreads();
  const RadixTile tl = tile_of(table, n, tiles);
  const int wbase = w * 64 * kItems;  // the wave's first key within the tile
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  K key[kItems];
  uint32_t val[kItems];
  int rank[kItems];
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const int o = wbase + r * 64 + lane;
    const bool ok = o < tl.count;
    key[r] = ok ? kin[(int64_t)tl.start + o] : K(0);
    val[r] = ok ? vin[(int64_t)tl.start + o] : 0u;
  }
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const bool ok = wbase + r * 64 + lane < tl.count;
    const int d = (int)((key[r] >> shift) & 255);
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t m = __ballot((d >> b) & 1);
      peers &= ((d >> b) & 1) ? m : ~m;
    }
    const int before = __popcll(peers & below);
    const int seen = ok ? wcnt[w][d] : 0;
    wave_lds_order();  // every lane has read its digit's count before one of them advances it
    if (ok && before == 0) wcnt[w][d] = seen + __popcll(peers);
    wave_lds_order();
    rank[r] = ok ? seen + before : -1;
  }
  __syncthreads();
  {
    const int d = threadIdx.x;  // one digit per thread
    int c[kWaves], t = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
      c[k] = wcnt[k][d];
      t += c[k];
    }
    int all = 0;
    const int start = block_exclusive(t, &all, sw);
    int run = start;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
      wcnt[k][d] = run;
      run += c[k];
    }
    dstart[d] = start;
    gbase[d] = offsets[(int64_t)tl.cbase + (int64_t)d * tl.stride];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kItems; ++r)
    if (rank[r] >= 0) {
      const int pos = wcnt[w][(int)((key[r] >> shift) & 255)] + rank[r];
      skey[pos] = key[r];
      sval[pos] = val[r];
    }
  __syncthreads();
  const int m = tl.count;
  for (int i = threadIdx.x; i < m; i += kBlock) {
    const K k = skey[i];
    const int d = (int)((k >> shift) & 255);
    const int64_t dst = (int64_t)gbase[d] + (i - dstart[d]);
    if (!kLast) kout[dst] = k;
    vout[dst] = sval[i];
  }
}

int64_t tiles_of(int64_t n) { return (n + kTile - 1) / kTile; }

}  // namespace

size_t scan_temp_bytes(int64_t n) { return sizeof(int32_t) * (size_t)(tiles_of(std::max<int64_t>(n, 1)) + 16); }

void exclusive_scan(const int32_t *in, int32_t *out, int64_t n, int32_t *temp, hipStream_t s) {
  if (n <= 0) return;
  if (n <= kSingleBlockMax) {
    hipLaunchKernelGGL(scan_one_block, dim3(1), dim3(kBlock), 0, s, in, out, n);
    FLK_HIP(hipGetLastError());
    return;
  }
  const int64_t t = tiles_of(n);
  hipLaunchKernelGGL(scan_tile_sums, dim3((unsigned)t), dim3(kBlock), 0, s, in, n, temp);
  hipLaunchKernelGGL(scan_one_block, dim3(1), dim3(kBlock), 0, s, (const int32_t *)temp, temp, t);
  hipLaunchKernelGGL(scan_tiles, dim3((unsigned)t), dim3(kBlock), 0, s, in, out, n, (const int32_t *)temp);
  FLK_HIP(hipGetLastError());
}

size_t radix_temp_bytes(int64_t n, int64_t table_tiles) {
  const int64_t cells = 256 * std::max<int64_t>(tiles_of(std::max<int64_t>(n, 1)), table_tiles);
  return sizeof(int32_t) * (size_t)cells + scan_temp_bytes(cells);
}

int64_t segmented_tiles(const int64_t *seg_off, int64_t nseg, std::vector<RadixTile> &table) {
  table.clear();
  for (int64_t c = 0; c < nseg; ++c) {
    const int64_t len = seg_off[c + 1] - seg_off[c];
    const int64_t nt = (len + kTile - 1) / kTile;
    const int64_t first = (int64_t)table.size();
    for (int64_t j = 0; j < nt; ++j)
      table.push_back(RadixTile{(int32_t)(seg_off[c] + j * kTile), (int32_t)std::min<int64_t>(kTile, len - j * kTile),
                                (int32_t)(256 * first + j), (int32_t)nt});
  }
  return (int64_t)table.size();
}

template <typename K>
void radix_sort_pairs(K *keys, K *keys_alt, uint32_t *vals, uint32_t *vals_alt, uint32_t *vals_out, int64_t n,
                      int end_bit, void *temp, hipStream_t s, const RadixTile *table, int64_t table_tiles) {
  if (n <= 0) return;
  if (n >= (int64_t)INT32_MAX) throw Error(FLK_ERR_ARG, "radix sort: too many keys");
  const int64_t tiles = table ? table_tiles : tiles_of(n);
  if (tiles <= 0) return;
  const int64_t cells = 256 * tiles;
  int32_t *counts = static_cast<int32_t *>(temp);
  int32_t *part = counts + cells;
  const int passes = std::max(1, (end_bit + 7) / 8);
  K *kbuf[2] = {keys, keys_alt};
  uint32_t *vbuf[2] = {vals, vals_alt};
  for (int p = 0; p < passes; ++p) {
    const K *ki = kbuf[p & 1];
    const uint32_t *vi = vbuf[p & 1];
    hipLaunchKernelGGL(radix_count<K>, dim3((unsigned)tiles), dim3(kBlock), 0, s, ki, n, 8 * p, table, tiles,
                       counts);
    FLK_HIP(hipGetLastError());
    exclusive_scan(counts, counts, cells, part, s);
    if (p + 1 == passes)
      hipLaunchKernelGGL((radix_scatter<K, true>), dim3((unsigned)tiles), dim3(kBlock), 0, s, ki, vi, n, 8 * p,
                         table, tiles, (const int32_t *)counts, (K *)nullptr, vals_out);
    else
      hipLaunchKernelGGL((radix_scatter<K, false>), dim3((unsigned)tiles), dim3(kBlock), 0, s, ki, vi, n, 8 * p,
                         table, tiles, (const int32_t *)counts, kbuf[(p + 1) & 1], vbuf[(p + 1) & 1]);
    FLK_HIP(hipGetLastError());
  }
}

template void radix_sort_pairs<uint32_t>(uint32_t *, uint32_t *, uint32_t *, uint32_t *, uint32_t *, int64_t, int,
                                         void *, hipStream_t, const RadixTile *, int64_t);
template void radix_sort_pairs<uint64_t>(uint64_t *, uint64_t *, uint32_t *, uint32_t *, uint32_t *, int64_t, int,
                                         void *, hipStream_t, const RadixTile *, int64_t);

}  // namespace flk
