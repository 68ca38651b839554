// scx_prims.hip -- device-wide primitives for gfx950 (wave64): exclusive scan of
// u32 arrays and the stable multi-bucket compaction used for plane-major contact
// ordering and colour-major contact ordering.
#include "scx_internal.h"
#include "scx_device.h"

namespace scx {

// ---------------------------------------------------------------------------
// Exclusive scan: reduce tiles -> scan tile sums (one workgroup) -> rescan tiles.
// out has n + 1 entries; out[n] = total. out may alias in.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                        uint32_t* __restrict__ tile_sums) {
  const size_t base = (size_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t i = base + (size_t)k * kBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  s = block_reduce_sum_u32(s);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = s;
}

// Single workgroup: in-place exclusive scan of nb tile sums, total -> tile_sums[nb].
__global__ void __launch_bounds__(kBlock) k_scan_top(uint32_t* __restrict__ tile_sums, uint32_t nb) {
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nb; base += kScanTile) {
    uint32_t v[kScanItems];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      const uint32_t i = base + threadIdx.x * kScanItems + k;
      v[k] = (i < nb) ? tile_sums[i] : 0u;
      local += v[k];
    }
    uint32_t total;
    uint32_t excl = block_exclusive_scan_u32(local, &total);
    uint32_t run = carry + excl;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      const uint32_t i = base + threadIdx.x * kScanItems + k;
      if (i < nb) tile_sums[i] = run;
      run += v[k];
    }
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) tile_sums[nb] = carry;
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(const uint32_t* in, uint32_t* out, uint32_t n,
                                                       const uint32_t* __restrict__ tile_offsets) {
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t local = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t i = base + k;
    v[k] = (i < n) ? in[i] : 0u;
    local += v[k];
  }
  uint32_t total;
  uint32_t run = tile_offsets[blockIdx.x] + block_exclusive_scan_u32(local, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = tile_offsets[gridDim.x];
}

size_t scan_tmp_words(uint32_t n) { return (size_t)(n + kScanTile - 1) / kScanTile + 1; }

hipError_t launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, hipStream_t s) {
  const uint32_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 0) {
    // empty input: out[0] = 0
    return hipMemsetAsync(out, 0, sizeof(uint32_t), s);
  }
  k_scan_reduce<<<nb, kBlock, 0, s>>>(in, n, tmp);
  k_scan_top<<<1, kBlock, 0, s>>>(tmp, nb);
  k_scan_apply<<<nb, kBlock, 0, s>>>(in, out, n, tmp);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stable multi-bucket compaction. Item i belongs to every bucket b whose bit is
// set in masks[i] (b < nbuckets <= 64). Output: for each bucket in order, the
// indices of its items in increasing order (so plane-major / colour-major lists
// keep the reference's body / contact order). counts: nbuckets * nblk words, then
// scanned in place; bucket b starts at counts[b * nblk].
// ---------------------------------------------------------------------------
// One ballot per (bucket, wave) is kept in LDS: the count pass sums popcounts,
// the scatter pass ranks each member with the lanes below it.
__global__ void __launch_bounds__(kBlock) k_mb_count(const unsigned long long* __restrict__ masks, uint32_t n,
                                                     uint32_t nbuckets, uint32_t* __restrict__ counts) {
  __shared__ uint32_t wc[kMaxColours][kBlock / 64];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const unsigned long long m = (i < n) ? masks[i] : 0ull;
  const int wave = threadIdx.x / 64;
  for (uint32_t b = 0; b < nbuckets; ++b) {
    const unsigned long long bal = __ballot((m >> b) & 1ull);
    if ((threadIdx.x & 63) == 0) wc[b][wave] = (uint32_t)__popcll(bal);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nbuckets; b += kBlock) {
    uint32_t s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += wc[b][w];
    counts[(size_t)b * gridDim.x + blockIdx.x] = s;
  }
}

__global__ void __launch_bounds__(kBlock) k_mb_scatter(const unsigned long long* __restrict__ masks, uint32_t n,
                                                       uint32_t nbuckets, const uint32_t* __restrict__ offsets,
                                                       uint32_t* __restrict__ out_idx, uint32_t cap,
                                                       uint32_t* __restrict__ overflow) {
  __shared__ unsigned long long wb[kMaxColours][kBlock / 64];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const unsigned long long m = (i < n) ? masks[i] : 0ull;
  const int wave = threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t b = 0; b < nbuckets; ++b) {
    const unsigned long long bal = __ballot((m >> b) & 1ull);
    if (lane == 0) wb[b][wave] = bal;
  }
  __syncthreads();
  if (m == 0ull) return;
  for (uint32_t b = 0; b < nbuckets; ++b) {
    if (!((m >> b) & 1ull)) continue;
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += (uint32_t)__popcll(wb[b][w]);
    const uint32_t prefix = (uint32_t)__popcll(wb[b][wave] & lt);
    const uint32_t pos = offsets[(size_t)b * gridDim.x + blockIdx.x] + before + prefix;
    if (pos < cap)
      out_idx[pos] = i;
    else
      atomicOr(overflow, 4u);
  }
}

// bucket_offsets[b] = start of bucket b, bucket_offsets[nbuckets] = total.
__global__ void k_gather_bucket_offsets(const uint32_t* __restrict__ counts, uint32_t nblk, uint32_t nbuckets,
                                        uint32_t* __restrict__ bucket_offsets) {
  for (uint32_t b = threadIdx.x; b <= nbuckets; b += blockDim.x)
    bucket_offsets[b] = counts[(size_t)b * nblk];
}

hipError_t launch_multibucket(scx_ctx* c, const unsigned long long* masks, uint32_t n, uint32_t nbuckets,
                              uint32_t* out_idx, uint32_t out_cap, uint32_t* bucket_offsets) {
  hipStream_t s = c->stream;
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  if (nblk == 0 || nbuckets == 0) {
    if (bucket_offsets) return hipMemsetAsync(bucket_offsets, 0, sizeof(uint32_t) * (nbuckets + 1), s);
    return hipSuccess;
  }
  uint32_t* counts = c->d.mbcount;
  const uint32_t total = nbuckets * nblk;
  k_mb_count<<<nblk, kBlock, 0, s>>>(masks, n, nbuckets, counts);
  hipError_t e = launch_exclusive_scan(counts, counts, total, c->d.scan_tmp, s);
  if (e != hipSuccess) return e;
  k_mb_scatter<<<nblk, kBlock, 0, s>>>(masks, n, nbuckets, counts, out_idx, out_cap, &c->d.ctr->overflow);
  if (bucket_offsets) k_gather_bucket_offsets<<<1, 64, 0, s>>>(counts, nblk, nbuckets, bucket_offsets);
  return hipGetLastError();
}

}  // namespace scx
