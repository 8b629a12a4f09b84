// sfb_sort.hpp -- stable radius order of a batch plan's points, made of plain
// kernels so that it can be captured in a hipGraph (no library sort, no
// memset).
//
// Each catalogue is sorted within its own range of the point table by a
// 16-bit radius bucket, key = floor(r 2^16 / R) (r <= R after the point
// preparation), with two least-significant-digit passes of 8 bits.  A pass is
// a per-block digit histogram, a per-catalogue exclusive scan over (digit,
// block) and a stable scatter: a block walks its chunk in order, 256 points at
// a time, and ranks equal digits by wave ballots, so that equal keys keep their
// input order and the result does not depend on scheduling.  Finally the
// point table is gathered in the new order.
//
// Consecutive points then have radii within one bucket of sorted order, so
// the Bessel values of a 32-point matrix-core block have similar magnitudes:
// the matrix cores' FP32 sum of a block loses precision relative to its
// largest product, and without the order a single small-radius point of large
// j_l(k r) in a block of small values sets that scale.
#pragma once

#include "sfb_common.hpp"

namespace sfb {

constexpr int SORT_IPT = 16;              // points per thread of a block's chunk
constexpr int SORT_CHUNK = 256 * SORT_IPT;  // points per block

// One catalogue: its range [base, base + n) of the point table, its blocks
// (nb = ceil(n / SORT_CHUNK)) and where its 256 x nb histogram starts.
struct SortSeg {
  int64_t base, n, hbase;
  int32_t nb, pad;
};

// 256-bin histogram of digit (key >> shift) & 255 of block blockIdx.x of
// catalogue blockIdx.y, stored digit-major: hist[hbase + d nb + b].
__global__ __launch_bounds__(256) void sort_hist_kernel(const SortSeg *__restrict__ segs,
                                                        const uint16_t *__restrict__ keys, int shift,
                                                        uint32_t *__restrict__ hist) {
  const SortSeg sg = segs[blockIdx.y];
  const int b = blockIdx.x;
  if (b >= sg.nb) return;  // block-uniform
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t p0 = (int64_t)b * SORT_CHUNK;
  for (int k = 0; k < SORT_IPT; ++k) {
    const int64_t i = p0 + k * 256 + threadIdx.x;
    if (i < sg.n) atomicAdd(&h[(keys[sg.base + i] >> shift) & 255], 1u);
  }
  __syncthreads();
  hist[sg.hbase + (int64_t)threadIdx.x * sg.nb + b] = h[threadIdx.x];
}

// In-place exclusive scan of one catalogue's histogram (blockIdx.x = catalogue):
// each thread sums a contiguous run, the 1024 run sums are scanned in LDS.
__global__ __launch_bounds__(1024) void sort_scan_kernel(const SortSeg *__restrict__ segs, uint32_t *__restrict__ hist) {
  const SortSeg sg = segs[blockIdx.x];
  const int tid = threadIdx.x;
  const int64_t total = 256LL * sg.nb;
  uint32_t *hh = hist + sg.hbase;
  const int64_t per = (total + 1023) / 1024;
  const int64_t lo = min<int64_t>(total, tid * per), hi = min<int64_t>(total, lo + per);
  uint32_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += hh[i];
  __shared__ uint32_t sums[1024];
  sums[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = tid >= off ? sums[tid - off] : 0u;
    __syncthreads();
    sums[tid] += v;
    __syncthreads();
  }
  uint32_t run = sums[tid] - s;
  for (int64_t i = lo; i < hi; ++i) {
    const uint32_t v = hh[i];
    hh[i] = run;
    run += v;
  }
}

// Stable scatter of one pass.  kin/iin: keys and point indices (relative to
// the catalogue; iin NULL = identity) in the current order; kout/iout: the
// order by digit (key >> shift) & 255, equal digits in current order.
__global__ __launch_bounds__(256) void sort_scatter_kernel(const SortSeg *__restrict__ segs,
                                                           const uint16_t *__restrict__ kin,
                                                           const int32_t *__restrict__ iin, int shift,
                                                           const uint32_t *__restrict__ hist,
                                                           uint16_t *__restrict__ kout, int32_t *__restrict__ iout) {
  const SortSeg sg = segs[blockIdx.y];
  const int b = blockIdx.x;
  if (b >= sg.nb) return;  // block-uniform
  __shared__ uint32_t run[256];    // next slot per digit (relative to the catalogue)
  __shared__ uint32_t wc[4][256];  // the current 256 points' count per (wave, digit)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  run[tid] = hist[sg.hbase + (int64_t)tid * sg.nb + b];
  const int64_t p0 = (int64_t)b * SORT_CHUNK;
  for (int k = 0; k < SORT_IPT; ++k) {
    // each thread clears and (below) reads only its own column tid of wc
    for (int w = 0; w < 4; ++w) wc[w][tid] = 0u;
    __syncthreads();  // run and the cleared counts visible to all
    const int64_t i = p0 + k * 256 + tid;
    const bool ok = i < sg.n;
    uint16_t key = 0;
    int32_t idx = 0;
    int d = 0;
    if (ok) {
      key = kin[sg.base + i];
      idx = iin ? iin[sg.base + i] : (int32_t)i;
      d = (key >> shift) & 255;
    }
    // the lanes of this wave holding the same digit (lanes past the end excluded)
    unsigned long long peers = __ballot(ok);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool set = (d >> bit) & 1;
      const unsigned long long bm = __ballot(ok && set);
      peers &= set ? bm : ~bm;
    }
    const int rank = __popcll(peers & ((1ull << lane) - 1ull));
    if (ok && rank == 0) wc[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();  // the waves' counts written
    if (ok) {
      uint32_t pos = run[d] + (uint32_t)rank;
      for (int w = 0; w < wave; ++w) pos += wc[w][d];
      kout[sg.base + pos] = key;
      iout[sg.base + pos] = idx;
    }
    __syncthreads();  // every read of run and wc done
    run[tid] += wc[0][tid] + wc[1][tid] + wc[2][tid] + wc[3][tid];
  }
}

// out[a][base + i] = in[a][base + idx[base + i]] for the six arrays of the
// point table (blockIdx.y = catalogue).
__global__ void sort_gather_kernel(const SortSeg *__restrict__ segs, const int32_t *__restrict__ idx,
                                   const double *__restrict__ in, double *__restrict__ out, int64_t ntot) {
  const SortSeg sg = segs[blockIdx.y];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < sg.n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = sg.base + i, j = sg.base + idx[o];
#pragma unroll
    for (int a = 0; a < 6; ++a) out[a * ntot + o] = in[a * ntot + j];
  }
}

}  // namespace sfb
