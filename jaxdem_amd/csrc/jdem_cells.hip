// jdem_cells.hip -- (a) the cell list on gfx950: cell hash, an LDS-staged LSD
// radix sort of (hash, particle id), cell start/end tables, the neighbour
// list of DynamicCellList.create_neighbor_list / NaiveSimulator
// .create_neighbor_list, the history remap of NeighborList, and float64 /
// min-max reductions.
//
// Reference: colliders/cell_list.py:37-96, 500-595; colliders/_partition.py:
// 55-150; colliders/naive.py:122-185; colliders/neighbor_list.py:29-54;
// domains/free.py:42-58.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "jdem_internal.h"

namespace jdem {

// ---------------------------------------------------------------------------
// cell coordinates (cell_list.py:55-60), clamped into [0, dims-1]
// ---------------------------------------------------------------------------
template <bool PER, class Real>
JD int cell_coord(Real x, int a, const DomPT<Real>& dom, const GridPT<Real>& g) {
  Real f;
  if (PER) {
    Real t = (x - dom.anchor[a]) / dom.box[a];
    t = t - floor_(t);  // jnp.remainder(t, 1.0)
    f = floor_(t * (Real)g.dims[a]);
  } else {
    f = floor_((x - dom.anchor[a]) / g.cell[a]);
  }
  f = fmin_(fmax_(f, Real(0)), (Real)(g.dims[a] - 1));
  return (int)f;
}

template <int D, bool PER, class Real>
__global__ void __launch_bounds__(kBlock) k_hash(const Real* __restrict__ pos, int64_t n, DomPT<Real> dom,
                                                  GridPT<Real> g, uint32_t* __restrict__ keys) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int h = 0;
#pragma unroll
  for (int a = 0; a < D; ++a) h += cell_coord<PER>(pos[D * i + a], a, dom, g) * g.strides[a];
  keys[i] = (uint32_t)h;
}

template <class Real>
hipError_t hash_run(const Real* pos, int64_t n, int dim, const DomPT<Real>& dom, const GridPT<Real>& g,
                    uint32_t* keys, hipStream_t s) {
  if (n == 0) return hipSuccess;
  dim3 grd(nblocks(n)), blk(kBlock);
  bool per = g.periodic != 0;
  if (dim == 3) {
    if (per)
      hipLaunchKernelGGL((k_hash<3, true, Real>), grd, blk, 0, s, pos, n, dom, g, keys);
    else
      hipLaunchKernelGGL((k_hash<3, false, Real>), grd, blk, 0, s, pos, n, dom, g, keys);
  } else {
    if (per)
      hipLaunchKernelGGL((k_hash<2, true, Real>), grd, blk, 0, s, pos, n, dom, g, keys);
    else
      hipLaunchKernelGGL((k_hash<2, false, Real>), grd, blk, 0, s, pos, n, dom, g, keys);
  }
  return hipGetLastError();
}
hipError_t launch_hash(const float* pos, int64_t n, int dim, const DomP& dom, const GridP& g, uint32_t* keys,
                       hipStream_t s) {
  return hash_run(pos, n, dim, dom, g, keys, s);
}
hipError_t launch_hash(const double* pos, int64_t n, int dim, const DomP64& dom, const GridP64& g, uint32_t* keys,
                       hipStream_t s) {
  return hash_run(pos, n, dim, dom, g, keys, s);
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits.  A tile of RS_TILE keys per block:
//   hist:    per-tile digit counts (LDS atomics) -> hist[digit][tile]
//   scan:    exclusive scan of hist in digit-major order (one block)
//   scatter: per 256-key row of the tile, each key's stable rank among equal
//            digits (wavefront ballot match + per-wave counts in LDS), then
//            a scatter to offset[digit][tile] + running count.
// ---------------------------------------------------------------------------
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = kBlock * RS_ITEMS;
constexpr int RS_WAVES = kBlock / 64;

// n_dev (when set) holds the key count on the device, n is then an upper
// bound that sized the launch; gate (when set) skips the sort while 0.
__global__ void __launch_bounds__(kBlock) k_radix_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                        uint32_t* __restrict__ hist, int ntiles,
                                                        const int64_t* __restrict__ n_dev,
                                                        const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  if (n_dev) n = *n_dev;
  __shared__ uint32_t bins[256];
  const int tid = threadIdx.x;
  // one tile per workgroup, or (a grid smaller than the tiles) several in turn
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t base = (int64_t)tile * RS_TILE;
    if (base >= n) {  // a tile past the keys counts nothing
      hist[(int64_t)tid * ntiles + tile] = 0u;
      continue;
    }
    bins[tid] = 0;
    __syncthreads();
#pragma unroll 4
    for (int it = 0; it < RS_ITEMS; ++it) {
      int64_t idx = base + (int64_t)it * kBlock + tid;
      if (idx < n) atomicAdd(&bins[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)tid * ntiles + tile] = bins[tid];
    __syncthreads();  // the next tile clears bins
  }
}

// Parallel exclusive scan of `total` uint32 values in three launches:
// per-chunk totals (coalesced), a one-block scan of the totals, then a
// per-chunk scan in LDS plus the chunk offset.
constexpr int SC_ITEMS = 16;
constexpr int SC_CHUNK = kBlock * SC_ITEMS;

__global__ void __launch_bounds__(kBlock) k_scan_totals(const uint32_t* __restrict__ data, int64_t total,
                                                         uint32_t* __restrict__ sums,
                                                         const int32_t* __restrict__ gate = nullptr) {
  if (gate && *gate == 0) return;
  __shared__ uint32_t sh[kBlock];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * SC_CHUNK;
  uint32_t s = 0;
#pragma unroll
  for (int it = 0; it < SC_ITEMS; ++it) {
    const int64_t k = base + (int64_t)it * kBlock + tid;
    if (k < total) s += data[k];
  }
  sh[tid] = s;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (tid < off) sh[tid] += sh[tid + off];
    __syncthreads();
  }
  if (tid == 0) sums[blockIdx.x] = sh[0];
}

// exclusive scan of nsums (<= 1024 * 1024) values in one block
__global__ void __launch_bounds__(1024) k_scan_sums(uint32_t* __restrict__ sums, int nsums,
                                                     const int32_t* __restrict__ gate = nullptr) {
  if (gate && *gate == 0) return;
  __shared__ uint32_t part[1024];
  const int tid = threadIdx.x;
  const int per = (nsums + 1023) / 1024;
  const int b = tid * per;
  const int e = min(b + per, nsums);
  uint32_t s = 0;
  for (int k = b; k < e; ++k) s += sums[k];
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = (tid >= off) ? part[tid - off] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = (tid == 0) ? 0u : part[tid - 1];
  for (int k = b; k < e; ++k) {
    const uint32_t v = sums[k];
    sums[k] = run;
    run += v;
  }
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(uint32_t* __restrict__ data, int64_t total,
                                                        const uint32_t* __restrict__ sums,
                                                        const int32_t* __restrict__ gate = nullptr) {
  if (gate && *gate == 0) return;
  __shared__ uint32_t tile[SC_CHUNK];
  __shared__ uint32_t tsum[kBlock];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * SC_CHUNK;
#pragma unroll
  for (int it = 0; it < SC_ITEMS; ++it) {
    const int64_t k = base + (int64_t)it * kBlock + tid;
    tile[it * kBlock + tid] = (k < total) ? data[k] : 0u;
  }
  __syncthreads();
  // thread t scans its contiguous run tile[t*SC_ITEMS .. +SC_ITEMS)
  uint32_t s = 0;
#pragma unroll
  for (int it = 0; it < SC_ITEMS; ++it) s += tile[tid * SC_ITEMS + it];
  tsum[tid] = s;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    const uint32_t v = (tid >= off) ? tsum[tid - off] : 0u;
    __syncthreads();
    tsum[tid] += v;
    __syncthreads();
  }
  uint32_t run = sums[blockIdx.x] + ((tid == 0) ? 0u : tsum[tid - 1]);
#pragma unroll
  for (int it = 0; it < SC_ITEMS; ++it) {
    const uint32_t v = tile[tid * SC_ITEMS + it];
    tile[tid * SC_ITEMS + it] = run;
    run += v;
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < SC_ITEMS; ++it) {
    const int64_t k = base + (int64_t)it * kBlock + tid;
    if (k < total) data[k] = tile[it * kBlock + tid];
  }
}

// Exclusive scan of `total` uint32 values in one block of 1024 threads
// (kept for totals small enough that one launch is cheaper).
__global__ void __launch_bounds__(1024) k_scan_exclusive(uint32_t* __restrict__ data, int64_t total,
                                                          const int32_t* __restrict__ gate = nullptr) {
  if (gate && *gate == 0) return;
  __shared__ uint32_t part[1024];
  const int tid = threadIdx.x;
  const int64_t chunk = (total + 1023) / 1024;
  const int64_t b = tid * chunk;
  const int64_t e = (b + chunk < total) ? b + chunk : total;
  uint32_t s = 0;
  for (int64_t k = b; k < e; ++k) s += data[k];
  part[tid] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan of the 1024 partial sums
  for (int off = 1; off < 1024; off <<= 1) {
    uint32_t v = (tid >= off) ? part[tid - off] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = (tid == 0) ? 0u : part[tid - 1];
  for (int64_t k = b; k < e; ++k) {
    uint32_t v = data[k];
    data[k] = run;
    run += v;
  }
}

__global__ void __launch_bounds__(kBlock) k_radix_scatter(const uint32_t* __restrict__ kin,
                                                           const int32_t* __restrict__ vin,
                                                           uint32_t* __restrict__ kout, int32_t* __restrict__ vout,
                                                           int64_t n, int shift, const uint32_t* __restrict__ offs,
                                                           int ntiles, const int64_t* __restrict__ n_dev,
                                                           const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  if (n_dev) n = *n_dev;
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[RS_WAVES][256];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint64_t lt = (1ull << lane) - 1ull;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
  const int64_t base = (int64_t)tile * RS_TILE;
  if (base >= n) break;  // tiles past the keys move nothing (and later tiles lie further past)
  run[tid] = offs[(int64_t)tid * ntiles + tile];
  for (int it = 0; it < RS_ITEMS; ++it) {
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    const int64_t idx = base + (int64_t)it * kBlock + tid;
    const bool valid = idx < n;
    const uint32_t key = valid ? kin[idx] : 0u;
    const int32_t val = valid ? (vin ? vin[idx] : (int32_t)idx) : 0;
    const uint32_t d = (key >> shift) & 255u;
    const uint64_t vmask = __ballot(valid);
    uint64_t peers = vmask;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(valid && bit);
      peers &= bit ? bb : (vmask & ~bb);
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pre = 0;
      for (int w = 0; w < wave; ++w) pre += wcnt[w][d];
      const uint32_t dst = run[d] + pre + rank;
      kout[dst] = key;
      vout[dst] = val;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) add += wcnt[w][tid];
    run[tid] += add;
    __syncthreads();
  }
  }
}

size_t radix_hist_words(int64_t n) {
  const int ntiles = (int)((n + RS_TILE - 1) / RS_TILE);
  const int64_t total = (int64_t)ntiles * 256;
  const int nchunks = (int)((total + SC_CHUNK - 1) / SC_CHUNK);
  return (size_t)(total > 0 ? total : 1) + (size_t)(nchunks + 1);
}

hipError_t radix_sort_bufs(const SortBufs& sb, const uint32_t* keys_in, int64_t n, int nbits, uint32_t** keys_out,
                           int32_t** vals_out, hipStream_t s, const int64_t* n_dev, const int32_t* gate) {
  const int ntiles = (int)((n + RS_TILE - 1) / RS_TILE);
  const int64_t total = (int64_t)ntiles * 256;
  const int nchunks = (int)((total + SC_CHUNK - 1) / SC_CHUNK);
  uint32_t* kb[2] = {sb.k0, sb.k1};
  int32_t* vb[2] = {sb.v0, sb.v1};
  uint32_t* hist = sb.hist;
  int passes = (nbits + 7) / 8;
  if (passes < 1) passes = 1;
  const uint32_t* kcur = keys_in;
  const int32_t* vcur = nullptr;  // first pass: value = index
  int out = 0;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    if (n > 0) {
      // with the count on the device the launch is capped and the tiles are
      // walked in turn: an idle (gated or empty) sort dispatches few workgroups
      const int grid = n_dev ? std::min(ntiles, kCappedGrid) : ntiles;
      hipLaunchKernelGGL(k_radix_hist, dim3(grid), dim3(kBlock), 0, s, kcur, n, shift, hist, ntiles, n_dev, gate);
      if (nchunks <= 1) {
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, hist, total, gate);
      } else {
        uint32_t* sums = hist + total;
        hipLaunchKernelGGL(k_scan_totals, dim3(nchunks), dim3(kBlock), 0, s, hist, total, sums, gate);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, nchunks, gate);
        hipLaunchKernelGGL(k_scan_apply, dim3(nchunks), dim3(kBlock), 0, s, hist, total, sums, gate);
      }
      hipLaunchKernelGGL(k_radix_scatter, dim3(grid), dim3(kBlock), 0, s, kcur, vcur, kb[out], vb[out], n, shift,
                         hist, ntiles, n_dev, gate);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    kcur = kb[out];
    vcur = vb[out];
    out ^= 1;
  }
  *keys_out = (uint32_t*)kcur;
  *vals_out = (int32_t*)vcur;
  return hipSuccess;
}

hipError_t radix_sort_pairs(jdem_ctx* ctx, const uint32_t* keys_in, int64_t n, int nbits, uint32_t** keys_out,
                            int32_t** vals_out, hipStream_t s, const int64_t* n_dev, const int32_t* gate) {
  const size_t nn = (size_t)(n > 0 ? n : 1);
  if (!ensure(ctx, ctx->keys[0], sizeof(uint32_t) * nn) || !ensure(ctx, ctx->keys[1], sizeof(uint32_t) * nn) ||
      !ensure(ctx, ctx->vals[0], sizeof(int32_t) * nn) || !ensure(ctx, ctx->vals[1], sizeof(int32_t) * nn) ||
      !ensure(ctx, ctx->hist, sizeof(uint32_t) * radix_hist_words(n)))
    return hipErrorOutOfMemory;
  SortBufs sb;
  sb.k0 = (uint32_t*)ctx->keys[0].p;
  sb.k1 = (uint32_t*)ctx->keys[1].p;
  sb.v0 = (int32_t*)ctx->vals[0].p;
  sb.v1 = (int32_t*)ctx->vals[1].p;
  sb.hist = (uint32_t*)ctx->hist.p;
  sb.cur = nullptr;
  return radix_sort_bufs(sb, keys_in, n, nbits, keys_out, vals_out, s, n_dev, gate);
}

// ---------------------------------------------------------------------------
// counting sort by cell.  Placement through atomic cursors puts a cell's ids
// in arbitrary order inside the cell's range; the rank pass then writes each
// id at its cell start + the number of smaller ids in the cell, which is the
// order of a stable sort of (key, id).  Cells hold a few particles, so the
// rank pass reads a handful of ids per element.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_cs_zero(int32_t* __restrict__ cnt, int32_t* __restrict__ cur,
                                                     int64_t n_cells, const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > n_cells) return;
  cnt[c] = 0;
  if (c < n_cells) cur[c] = 0;
}

__global__ void __launch_bounds__(kBlock) k_cs_count(const uint32_t* __restrict__ keys, int64_t n,
                                                      int32_t* __restrict__ cnt, const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicAdd(&cnt[keys[i]], 1);
}

__global__ void __launch_bounds__(kBlock) k_cs_place(const uint32_t* __restrict__ keys, int64_t n,
                                                      const int32_t* __restrict__ starts, int32_t* __restrict__ cur,
                                                      uint32_t* __restrict__ tkey, int32_t* __restrict__ tid,
                                                      const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  const int32_t slot = starts[k] + atomicAdd(&cur[k], 1);
  tkey[slot] = k;
  tid[slot] = (int32_t)i;
}

__global__ void __launch_bounds__(kBlock) k_cs_rank(const uint32_t* __restrict__ tkey, const int32_t* __restrict__ tid,
                                                     int64_t n, const int32_t* __restrict__ starts,
                                                     uint32_t* __restrict__ sk, int32_t* __restrict__ perm,
                                                     const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t k = tkey[j];
  const int32_t id = tid[j];
  const int32_t s = starts[k], e = starts[k + 1];
  int32_t r = 0;
  for (int32_t m = s; m < e; ++m) r += tid[m] < id ? 1 : 0;
  sk[s + r] = k;
  perm[s + r] = id;
}

__global__ void __launch_bounds__(kBlock) k_cs_ends(const int32_t* __restrict__ starts, int64_t n_cells,
                                                     int32_t* __restrict__ cend) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n_cells) cend[c] = starts[c + 1];
}
hipError_t launch_starts_to_ends(const int32_t* starts, int64_t n_cells, int32_t* cend, hipStream_t s) {
  if (n_cells <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cs_ends, dim3(nblocks(n_cells)), dim3(kBlock), 0, s, starts, n_cells, cend);
  return hipGetLastError();
}

size_t cell_sort_sum_words(int64_t n_cells) {
  const int64_t total = n_cells + 1;
  return (size_t)((total + SC_CHUNK - 1) / SC_CHUNK) + 1;
}

// keys_in may be sb.k1: it is read only before the rank pass writes the
// sorted keys there.
hipError_t cell_sort_bufs(const SortBufs& sb, const uint32_t* keys_in, int64_t n, int64_t n_cells, uint32_t** keys_out,
                          int32_t** vals_out, int32_t* starts, hipStream_t s, const int32_t* gate) {
  const int64_t total = n_cells + 1;
  const int nchunks = (int)((total + SC_CHUNK - 1) / SC_CHUNK);
  uint32_t* tkey = sb.k0;
  int32_t* tid = sb.v0;
  uint32_t* sk = sb.k1;
  int32_t* perm = sb.v1;
  int32_t* cur = sb.cur;
  hipLaunchKernelGGL(k_cs_zero, dim3(nblocks(total)), dim3(kBlock), 0, s, starts, cur, n_cells, gate);
  if (n > 0) hipLaunchKernelGGL(k_cs_count, dim3(nblocks(n)), dim3(kBlock), 0, s, keys_in, n, starts, gate);
  uint32_t* cnt = (uint32_t*)starts;  // counts are non-negative: scan them as uint32
  if (nchunks <= 1) {
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, cnt, total, gate);
  } else {
    uint32_t* sums = sb.hist;
    hipLaunchKernelGGL(k_scan_totals, dim3(nchunks), dim3(kBlock), 0, s, cnt, total, sums, gate);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, nchunks, gate);
    hipLaunchKernelGGL(k_scan_apply, dim3(nchunks), dim3(kBlock), 0, s, cnt, total, sums, gate);
  }
  if (n > 0) {
    hipLaunchKernelGGL(k_cs_place, dim3(nblocks(n)), dim3(kBlock), 0, s, keys_in, n, starts, cur, tkey, tid, gate);
    hipLaunchKernelGGL(k_cs_rank, dim3(nblocks(n)), dim3(kBlock), 0, s, tkey, tid, n, starts, sk, perm, gate);
  }
  *keys_out = sk;
  *vals_out = perm;
  return hipGetLastError();
}

// ---- counting sort with device-side counts (batch rebuilds) --------------
// Each kernel returns at once while *gate == 0 and walks its range in turn
// over a grid capped at kCappedGrid workgroups.
__global__ void __launch_bounds__(kBlock) k_csd_zero(int32_t* __restrict__ cnt, int32_t* __restrict__ cur,
                                                      const int64_t* __restrict__ c_dev,
                                                      const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int64_t nc = *c_dev;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c <= nc; c += (int64_t)gridDim.x * blockDim.x) {
    cnt[c] = 0;
    if (c < nc) cur[c] = 0;
  }
}

__global__ void __launch_bounds__(kBlock) k_csd_count(const uint32_t* __restrict__ keys,
                                                       const int64_t* __restrict__ n_dev, int32_t* __restrict__ cnt,
                                                       const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int64_t n = *n_dev;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[keys[i]], 1);
}

// exclusive scan of *c_dev + 1 values: chunk totals, one block over the
// totals, chunk scans plus offsets (k_scan_totals / _sums / _apply with the
// length read on the device)
__global__ void __launch_bounds__(kBlock) k_scan_totals_dev(const uint32_t* __restrict__ data,
                                                             const int64_t* __restrict__ c_dev,
                                                             uint32_t* __restrict__ sums,
                                                             const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ uint32_t sh[kBlock];
  const int tid = threadIdx.x;
  const int64_t total = *c_dev + 1;
  const int64_t nchunks = (total + SC_CHUNK - 1) / SC_CHUNK;
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int64_t base = ch * SC_CHUNK;
    uint32_t s = 0;
#pragma unroll
    for (int it = 0; it < SC_ITEMS; ++it) {
      const int64_t k = base + (int64_t)it * kBlock + tid;
      if (k < total) s += data[k];
    }
    sh[tid] = s;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
      if (tid < off) sh[tid] += sh[tid + off];
      __syncthreads();
    }
    if (tid == 0) sums[ch] = sh[0];
    __syncthreads();  // the next chunk overwrites sh
  }
}

__global__ void __launch_bounds__(1024) k_scan_sums_dev(uint32_t* __restrict__ sums, const int64_t* __restrict__ c_dev,
                                                         const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ uint32_t part[1024];
  const int tid = threadIdx.x;
  const int nsums = (int)((*c_dev + 1 + SC_CHUNK - 1) / SC_CHUNK);
  const int per = (nsums + 1023) / 1024;
  const int b = tid * per;
  const int e = min(b + per, nsums);
  uint32_t s = 0;
  for (int k = b; k < e; ++k) s += sums[k];
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = (tid >= off) ? part[tid - off] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = (tid == 0) ? 0u : part[tid - 1];
  for (int k = b; k < e; ++k) {
    const uint32_t v = sums[k];
    sums[k] = run;
    run += v;
  }
}

__global__ void __launch_bounds__(kBlock) k_scan_apply_dev(uint32_t* __restrict__ data,
                                                            const int64_t* __restrict__ c_dev,
                                                            const uint32_t* __restrict__ sums,
                                                            const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ uint32_t tile[SC_CHUNK];
  __shared__ uint32_t tsum[kBlock];
  const int tid = threadIdx.x;
  const int64_t total = *c_dev + 1;
  const int64_t nchunks = (total + SC_CHUNK - 1) / SC_CHUNK;
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int64_t base = ch * SC_CHUNK;
#pragma unroll
    for (int it = 0; it < SC_ITEMS; ++it) {
      const int64_t k = base + (int64_t)it * kBlock + tid;
      tile[it * kBlock + tid] = (k < total) ? data[k] : 0u;
    }
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (int it = 0; it < SC_ITEMS; ++it) s += tile[tid * SC_ITEMS + it];
    tsum[tid] = s;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
      const uint32_t v = (tid >= off) ? tsum[tid - off] : 0u;
      __syncthreads();
      tsum[tid] += v;
      __syncthreads();
    }
    uint32_t run = sums[ch] + ((tid == 0) ? 0u : tsum[tid - 1]);
#pragma unroll
    for (int it = 0; it < SC_ITEMS; ++it) {
      const uint32_t v = tile[tid * SC_ITEMS + it];
      tile[tid * SC_ITEMS + it] = run;
      run += v;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SC_ITEMS; ++it) {
      const int64_t k = base + (int64_t)it * kBlock + tid;
      if (k < total) data[k] = tile[it * kBlock + tid];
    }
    __syncthreads();  // the next chunk overwrites tile and tsum
  }
}

__global__ void __launch_bounds__(kBlock) k_csd_place(const uint32_t* __restrict__ keys,
                                                       const int64_t* __restrict__ n_dev,
                                                       const int32_t* __restrict__ starts, int32_t* __restrict__ cur,
                                                       uint32_t* __restrict__ tkey, int32_t* __restrict__ tid,
                                                       const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int64_t n = *n_dev;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t k = keys[i];
    const int32_t slot = starts[k] + atomicAdd(&cur[k], 1);
    tkey[slot] = k;
    tid[slot] = (int32_t)i;
  }
}

__global__ void __launch_bounds__(kBlock) k_csd_rank(const uint32_t* __restrict__ tkey,
                                                      const int32_t* __restrict__ tid,
                                                      const int64_t* __restrict__ n_dev,
                                                      const int32_t* __restrict__ starts, uint32_t* __restrict__ sk,
                                                      int32_t* __restrict__ perm, const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int64_t n = *n_dev;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t k = tkey[j];
    const int32_t id = tid[j];
    const int32_t s = starts[k], e = starts[k + 1];
    int32_t r = 0;
    for (int32_t m = s; m < e; ++m) r += tid[m] < id ? 1 : 0;
    sk[s + r] = k;
    perm[s + r] = id;
  }
}

hipError_t cell_sort_dev(jdem_ctx* ctx, const uint32_t* keys_in, int64_t n, int64_t n_cells, const int64_t* n_dev,
                         const int64_t* c_dev, uint32_t** keys_out, int32_t** vals_out, hipStream_t s,
                         const int32_t* gate) {
  const size_t nn = (size_t)(n > 0 ? n : 1);
  const int64_t nc = n_cells > 0 ? n_cells : 1;
  const int64_t max_chunks = (nc + 1 + SC_CHUNK - 1) / SC_CHUNK;
  if (max_chunks > 1024 * 1024) return hipErrorInvalidValue;
  if (!ensure(ctx, ctx->keys[0], sizeof(uint32_t) * nn) || !ensure(ctx, ctx->keys[1], sizeof(uint32_t) * nn) ||
      !ensure(ctx, ctx->vals[0], sizeof(int32_t) * nn) || !ensure(ctx, ctx->vals[1], sizeof(int32_t) * nn) ||
      !ensure(ctx, ctx->ccur, sizeof(int32_t) * (size_t)nc) ||
      !ensure(ctx, ctx->ccnt, sizeof(int32_t) * (size_t)(nc + 1)) ||
      !ensure(ctx, ctx->hist, sizeof(uint32_t) * (size_t)(max_chunks + 1)))
    return hipErrorOutOfMemory;
  uint32_t* tkey = (uint32_t*)ctx->keys[0].p;
  int32_t* tid = (int32_t*)ctx->vals[0].p;
  uint32_t* sk = (uint32_t*)ctx->keys[1].p;
  int32_t* perm = (int32_t*)ctx->vals[1].p;
  int32_t* cnt = (int32_t*)ctx->ccnt.p;
  int32_t* cur = (int32_t*)ctx->ccur.p;
  uint32_t* sums = (uint32_t*)ctx->hist.p;
  const unsigned gc = std::min<unsigned>(nblocks(nc + 1), (unsigned)kCappedGrid);
  const unsigned gn = std::min<unsigned>(nblocks((int64_t)nn), (unsigned)kCappedGrid);
  const unsigned gk = (unsigned)std::min<int64_t>(max_chunks, (int64_t)kCappedGrid);
  hipLaunchKernelGGL(k_csd_zero, dim3(gc), dim3(kBlock), 0, s, cnt, cur, c_dev, gate);
  hipLaunchKernelGGL(k_csd_count, dim3(gn), dim3(kBlock), 0, s, keys_in, n_dev, cnt, gate);
  hipLaunchKernelGGL(k_scan_totals_dev, dim3(gk), dim3(kBlock), 0, s, (const uint32_t*)cnt, c_dev, sums, gate);
  hipLaunchKernelGGL(k_scan_sums_dev, dim3(1), dim3(1024), 0, s, sums, c_dev, gate);
  hipLaunchKernelGGL(k_scan_apply_dev, dim3(gk), dim3(kBlock), 0, s, (uint32_t*)cnt, c_dev, sums, gate);
  hipLaunchKernelGGL(k_csd_place, dim3(gn), dim3(kBlock), 0, s, keys_in, n_dev, cnt, cur, tkey, tid, gate);
  hipLaunchKernelGGL(k_csd_rank, dim3(gn), dim3(kBlock), 0, s, tkey, tid, n_dev, cnt, sk, perm, gate);
  *keys_out = sk;
  *vals_out = perm;
  return hipGetLastError();
}

hipError_t cell_sort_pairs(jdem_ctx* ctx, const uint32_t* keys_in, int64_t n, int64_t n_cells, uint32_t** keys_out,
                           int32_t** vals_out, int32_t* starts, hipStream_t s, const int32_t* gate) {
  const size_t nn = (size_t)(n > 0 ? n : 1);
  if (!ensure(ctx, ctx->keys[0], sizeof(uint32_t) * nn) || !ensure(ctx, ctx->keys[1], sizeof(uint32_t) * nn) ||
      !ensure(ctx, ctx->vals[0], sizeof(int32_t) * nn) || !ensure(ctx, ctx->vals[1], sizeof(int32_t) * nn) ||
      !ensure(ctx, ctx->ccur, sizeof(int32_t) * (size_t)(n_cells > 0 ? n_cells : 1)) ||
      !ensure(ctx, ctx->hist, sizeof(uint32_t) * cell_sort_sum_words(n_cells)))
    return hipErrorOutOfMemory;
  SortBufs sb;
  sb.k0 = (uint32_t*)ctx->keys[0].p;
  sb.k1 = (uint32_t*)ctx->keys[1].p;
  sb.v0 = (int32_t*)ctx->vals[0].p;
  sb.v1 = (int32_t*)ctx->vals[1].p;
  sb.hist = (uint32_t*)ctx->hist.p;
  sb.cur = (int32_t*)ctx->ccur.p;
  return cell_sort_bufs(sb, keys_in, n, n_cells, keys_out, vals_out, starts, s, gate);
}

// ---------------------------------------------------------------------------
// cell tables
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3(nblocks(n)), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) k_copy_f32(float* __restrict__ d, const float* __restrict__ s, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}
hipError_t launch_copy_f32(float* dst, const float* src, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_f32, dim3(nblocks(n)), dim3(kBlock), 0, s, dst, src, n);
  return hipGetLastError();
}

// start[h] = first sorted index of cell h, end[h] = one past its last.
__global__ void __launch_bounds__(kBlock) k_cell_ranges(const uint32_t* __restrict__ sk, int64_t n,
                                                         int32_t* __restrict__ cstart, int32_t* __restrict__ cend) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t k = sk[s];
  if (s == 0 || sk[s - 1] != k) cstart[k] = (int32_t)s;
  if (s == n - 1 || sk[s + 1] != k) cend[k] = (int32_t)(s + 1);
}
hipError_t launch_cell_ranges(const uint32_t* skeys, int64_t n, int32_t* cstart, int32_t* cend, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cell_ranges, dim3(nblocks(n)), dim3(kBlock), 0, s, skeys, n, cstart, cend);
  return hipGetLastError();
}

// Sorted copy (x, y, z, r) of the particles: contiguous candidate reads.
__global__ void __launch_bounds__(kBlock) k_gather_spos(const int32_t* __restrict__ perm,
                                                         const float* __restrict__ pos, const float* __restrict__ rad,
                                                         int64_t n, int dim, float4* __restrict__ spos) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int64_t i = perm[s];
  float4 v;
  v.x = pos[dim * i];
  v.y = pos[dim * i + 1];
  v.z = (dim == 3) ? pos[dim * i + 2] : 0.0f;
  v.w = rad[i];
  spos[s] = v;
}
hipError_t launch_gather_spos(const int32_t* perm, const float* pos, const float* rad, int64_t n, int dim,
                              float4* spos, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_spos, dim3(nblocks(n)), dim3(kBlock), 0, s, perm, pos, rad, n, dim, spos);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// neighbour lists
// ---------------------------------------------------------------------------
JD bool bonded(const int32_t* bond_id, int bond_w, int64_t i, int32_t j) {
  for (int w = 0; w < bond_w; ++w)
    if (bond_id[i * bond_w + w] == j) return true;
  return false;
}

// Highest stencil offset kept along an axis by _dedup_stencil_hashes
// (cell_list.py:92-96): on a periodic axis with d cells the offset o in
// [-sr, sr] is a first occurrence iff o + sr < d.
JD int stencil_hi(int sr, int d, bool per) { return per ? min(sr, d - sr - 1) : sr; }

template <int D>
JD Vec<D> v4(float4 a);
template <>
JD Vec<3> v4<3>(float4 a) { return {a.x, a.y, a.z}; }
template <>
JD Vec<2> v4<2>(float4 a) { return {a.x, a.y}; }

// DynamicCellList.create_neighbor_list (cell_list.py:500-595): thread per
// sorted index; rows in stencil order, ascending id within a cell.
template <int D, bool PER, class Real>
__global__ void __launch_bounds__(kBlock) k_nlist_cells(NLBuildT<Real> b) {
  constexpr bool kF32 = std::is_same<Real, float>::value;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.n) return;
  const int32_t i = b.perm[s];
  Vec<D, Real> xi;
  if constexpr (kF32)
    xi = v4<D>(b.spos[s]);
  else  // float64: the particles' own positions
    xi = ld<D>(b.pos, i);
  const uint32_t key = (uint32_t)b.skeys[s];
  int c[3];
  c[0] = (int)(key % (uint32_t)b.grid.dims[0]);
  c[1] = (int)((key / (uint32_t)b.grid.dims[0]) % (uint32_t)b.grid.dims[1]);
  c[2] = (D == 3) ? (int)(key / ((uint32_t)b.grid.dims[0] * (uint32_t)b.grid.dims[1])) : 0;
  const int sr = b.grid.sr;
  int hi[3];
  for (int a = 0; a < D; ++a) hi[a] = stencil_hi(sr, b.grid.dims[a], PER);
  int32_t* row = b.nl + (int64_t)i * b.M;
  int count = 0;
  const bool chk_bond = (!b.isb) && b.bond_w > 0;
  for (int ox = -sr; ox <= hi[0]; ++ox) {
    int cx = c[0] + ox;
    if (PER) {
      cx = cx % b.grid.dims[0];
      if (cx < 0) cx += b.grid.dims[0];
    } else if (cx < 0 || cx >= b.grid.dims[0])
      continue;
    for (int oy = -sr; oy <= hi[1]; ++oy) {
      int cy = c[1] + oy;
      if (PER) {
        cy = cy % b.grid.dims[1];
        if (cy < 0) cy += b.grid.dims[1];
      } else if (cy < 0 || cy >= b.grid.dims[1])
        continue;
      const int zlo = (D == 3) ? -sr : 0, zhi = (D == 3) ? hi[2] : 0;
      for (int oz = zlo; oz <= zhi; ++oz) {
        int cz = 0;
        if (D == 3) {
          cz = c[2] + oz;
          if (PER) {
            cz = cz % b.grid.dims[2];
            if (cz < 0) cz += b.grid.dims[2];
          } else if (cz < 0 || cz >= b.grid.dims[2])
            continue;
        }
        const int h = cx * b.grid.strides[0] + cy * b.grid.strides[1] + (D == 3 ? cz * b.grid.strides[2] : 0);
        const int k0 = b.cstart[h], k1 = b.cend[h];
        for (int k = k0; k < k1; ++k) {
          const int32_t j = b.perm[k];
          if (b.clump ? b.clump[j] == b.clump[i] : j == i) continue;
          Vec<D, Real> xj;
          if constexpr (kF32)
            xj = v4<D>(b.spos[k]);
          else
            xj = ld<D>(b.pos, j);
          if (!(dist_sq<PER>(xi, xj, b.dom) <= b.cutoff_sq)) continue;
          if (chk_bond && bonded(b.bond_id, b.bond_w, i, j)) continue;
          if (count < b.M) row[count] = j;
          ++count;
        }
      }
    }
  }
  for (int t = count; t < b.M; ++t) row[t] = -1;
  if (count > b.M) atomicOr(b.overflow, 1);
}

// NaiveSimulator.create_neighbor_list (naive.py:122-185): descending ids,
// _displacement (multiply by the inverse box).
template <int D, bool PER, class Real>
__global__ void __launch_bounds__(kBlock) k_nlist_naive(NLBuildT<Real> b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  const Vec<D, Real> xi = ld<D>(b.pos, i);
  int32_t* row = b.nl + i * b.M;
  int count = 0;
  const bool chk_bond = (!b.isb) && b.bond_w > 0;
  for (int64_t j = b.n - 1; j >= 0; --j) {
    if (b.clump ? b.clump[j] == b.clump[i] : j == i) continue;
    const Vec<D, Real> d = disp_<PER>(xi, ld<D>(b.pos, j), b.dom);
    if (!(dot(d, d) <= b.cutoff_sq)) continue;
    if (chk_bond && bonded(b.bond_id, b.bond_w, i, (int32_t)j)) continue;
    if (count < b.M) row[count] = (int32_t)j;
    ++count;
  }
  for (int t = count; t < b.M; ++t) row[t] = -1;
  if (count > b.M) atomicOr(b.overflow, 1);
}

template <class Real>
hipError_t nlist_cells_run(const NLBuildT<Real>& b, hipStream_t s) {
  if (b.n == 0) return hipSuccess;
  dim3 grd(nblocks(b.n)), blk(kBlock);
  bool per = b.grid.periodic != 0;
  if (b.dim == 3) {
    if (per)
      hipLaunchKernelGGL((k_nlist_cells<3, true, Real>), grd, blk, 0, s, b);
    else
      hipLaunchKernelGGL((k_nlist_cells<3, false, Real>), grd, blk, 0, s, b);
  } else {
    if (per)
      hipLaunchKernelGGL((k_nlist_cells<2, true, Real>), grd, blk, 0, s, b);
    else
      hipLaunchKernelGGL((k_nlist_cells<2, false, Real>), grd, blk, 0, s, b);
  }
  return hipGetLastError();
}

template <class Real>
hipError_t nlist_naive_run(const NLBuildT<Real>& b, hipStream_t s) {
  if (b.n == 0) return hipSuccess;
  dim3 grd(nblocks(b.n)), blk(kBlock);
  bool per = b.dom.kind == JDEM_DOMAIN_PERIODIC;
  if (b.dim == 3) {
    if (per)
      hipLaunchKernelGGL((k_nlist_naive<3, true, Real>), grd, blk, 0, s, b);
    else
      hipLaunchKernelGGL((k_nlist_naive<3, false, Real>), grd, blk, 0, s, b);
  } else {
    if (per)
      hipLaunchKernelGGL((k_nlist_naive<2, true, Real>), grd, blk, 0, s, b);
    else
      hipLaunchKernelGGL((k_nlist_naive<2, false, Real>), grd, blk, 0, s, b);
  }
  return hipGetLastError();
}

hipError_t launch_nlist_cells(const NLBuild& b, hipStream_t s) { return nlist_cells_run(b, s); }
hipError_t launch_nlist_naive(const NLBuild& b, hipStream_t s) { return nlist_naive_run(b, s); }
hipError_t launch_nlist_cells(const NLBuildT<double>& b, hipStream_t s) { return nlist_cells_run(b, s); }
hipError_t launch_nlist_naive(const NLBuildT<double>& b, hipStream_t s) { return nlist_naive_run(b, s); }

// _remap_history_array (neighbor_list.py:29-54): a new slot inherits the
// history of the first old slot with the same partner id.
__global__ void __launch_bounds__(kBlock) k_remap_hist(int64_t n, int M, int W32, int dim,
                                                        const int32_t* __restrict__ old_nl,
                                                        const uint32_t* __restrict__ old_live,
                                                        const float* __restrict__ old_xi,
                                                        const int32_t* __restrict__ new_nl,
                                                        uint32_t* __restrict__ new_live, float* __restrict__ new_xi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t* orow = old_nl + i * M;
  const int32_t* nrow = new_nl + i * M;
  for (int w = 0; w < W32; ++w) {
    uint32_t bits = 0;
    for (int sl = w * 32; sl < M && sl < (w + 1) * 32; ++sl) {
      const int32_t j = nrow[sl];
      if (j < 0) continue;
      int t = -1;
      for (int u = 0; u < M; ++u)
        if (orow[u] == j) {
          t = u;
          break;
        }
      if (t < 0) continue;
      if (!((old_live[i * W32 + (t >> 5)] >> (t & 31)) & 1u)) continue;
      for (int c = 0; c < dim; ++c) new_xi[(i * M + sl) * dim + c] = old_xi[(i * M + t) * dim + c];
      bits |= 1u << (sl & 31);
    }
    new_live[i * W32 + w] = bits;
  }
}

hipError_t launch_remap_hist(int64_t n, int M, int W32, int dim, const int32_t* old_nl, const uint32_t* old_live,
                             const float* old_xi, const int32_t* new_nl, uint32_t* new_live, float* new_xi,
                             hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_remap_hist, dim3(nblocks(n)), dim3(kBlock), 0, s, n, M, W32, dim, old_nl, old_live, old_xi,
                     new_nl, new_live, new_xi);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// FreeDomain.apply (free.py:42-58): per-block min(x - r) / max(x + r).
template <class Real>
__global__ void __launch_bounds__(kBlock) k_box_minmax(const Real* __restrict__ pos, const Real* __restrict__ rad,
                                                        int64_t n, int dim, Real* __restrict__ part,
                                                        const Real* __restrict__ pos_p) {
  __shared__ Real smin[3][kBlock], smax[3][kBlock];
  const int tid = threadIdx.x;
  Real lo[3] = {Real(INFINITY), Real(INFINITY), Real(INFINITY)};
  Real hi[3] = {-Real(INFINITY), -Real(INFINITY), -Real(INFINITY)};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + tid; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    Real r = rad[i];
    if (pos_p) r = (dim == 3 ? norm_(ld<3>(pos_p, i)) : norm_(ld<2>(pos_p, i))) + r;  // clumps: |pos_p| + rad
    for (int a = 0; a < dim; ++a) {
      const Real x = pos[dim * i + a];
      lo[a] = fmin_(lo[a], x - r);
      hi[a] = fmax_(hi[a], x + r);
    }
  }
  for (int a = 0; a < 3; ++a) {
    smin[a][tid] = lo[a];
    smax[a][tid] = hi[a];
  }
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (tid < off)
      for (int a = 0; a < 3; ++a) {
        smin[a][tid] = fmin_(smin[a][tid], smin[a][tid + off]);
        smax[a][tid] = fmax_(smax[a][tid], smax[a][tid + off]);
      }
    __syncthreads();
  }
  if (tid == 0)
    for (int a = 0; a < 3; ++a) {
      part[6 * blockIdx.x + a] = smin[a][0];
      part[6 * blockIdx.x + 3 + a] = smax[a][0];
    }
}

hipError_t launch_box_minmax(const float* pos, const float* rad, int64_t n, int dim, float* part, int nparts,
                             hipStream_t s, const float* pos_p) {
  hipLaunchKernelGGL((k_box_minmax<float>), dim3(nparts), dim3(kBlock), 0, s, pos, rad, n, dim, part, pos_p);
  return hipGetLastError();
}
hipError_t launch_box_minmax(const double* pos, const double* rad, int64_t n, int dim, double* part, int nparts,
                             hipStream_t s, const double* pos_p) {
  hipLaunchKernelGGL((k_box_minmax<double>), dim3(nparts), dim3(kBlock), 0, s, pos, rad, n, dim, part, pos_p);
  return hipGetLastError();
}

// Deterministic float64 sum of `ncomp` interleaved components (stride).
__global__ void __launch_bounds__(kBlock) k_sum_f64_part(const double* __restrict__ in, int64_t n, int stride,
                                                          int ncomp, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int tid = threadIdx.x;
  for (int c = 0; c < ncomp; ++c) {
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + tid; i < n; i += (int64_t)gridDim.x * blockDim.x)
      acc += in[i * stride + c];
    sh[tid] = acc;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
      if (tid < off) sh[tid] += sh[tid + off];
      __syncthreads();
    }
    if (tid == 0) part[(int64_t)blockIdx.x * ncomp + c] = sh[0];
    __syncthreads();
  }
}
__global__ void __launch_bounds__(kBlock) k_sum_f64_final(const double* __restrict__ part, int nparts, int ncomp,
                                                           double* __restrict__ out) {
  __shared__ double sh[kBlock];
  const int tid = threadIdx.x;
  for (int c = 0; c < ncomp; ++c) {
    double acc = 0.0;
    for (int b = tid; b < nparts; b += kBlock) acc += part[(int64_t)b * ncomp + c];
    sh[tid] = acc;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
      if (tid < off) sh[tid] += sh[tid + off];
      __syncthreads();
    }
    if (tid == 0) out[c] = sh[0];
    __syncthreads();
  }
}
hipError_t launch_sum_f64(const double* in, int64_t n, int stride, int ncomp, double* partial, double* out,
                          int nparts, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_f64_part, dim3(nparts), dim3(kBlock), 0, s, in, n, stride, ncomp, partial);
  hipLaunchKernelGGL(k_sum_f64_final, dim3(1), dim3(kBlock), 0, s, partial, nparts, ncomp, out);
  return hipGetLastError();
}

}  // namespace jdem
