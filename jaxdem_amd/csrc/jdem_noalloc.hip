// jdem_noalloc.hip -- jdem_cell_build_noalloc: the spatial partition of
// jdem_cell_build (hash + stable sort by cell + cell start/end) for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  Every scratch array of
// the sort is carved out of one fixed device array, `scratch`, of
// kNoallocScratchBytes; each GPU holds its own copy from the time the
// library's code object is loaded until the process ends.
//
// All calls on one GPU share that GPU's array and it keeps the last call's
// contents: calls that overlap on one GPU (two streams, two threads) race.
// Every word a call reads is written earlier in the same call (the counting
// sort's counts and cursors are cleared by a kernel, not hipMemset), so a
// call never depends on what an earlier call left behind.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "jdem_internal.h"

using namespace jdem;

namespace {
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

// Device address of `scratch` on each device (the host-side name of the
// array is not a device address and must never reach a kernel).
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

int fail(jdem_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

#define NACHK(ctx, expr)                                                                      \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail((ctx), JDEM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

__global__ void __launch_bounds__(kBlock) k_copy_i32(int32_t* __restrict__ d, const int32_t* __restrict__ s,
                                                      int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

hipError_t copy_i32(int32_t* d, const int32_t* s, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_i32, dim3(nblocks(n)), dim3(kBlock), 0, st, d, s, n);
  return hipGetLastError();
}

// Byte offsets of the sort's arrays inside `scratch`.
struct Layout {
  bool counted;  // counting sort (else LSD radix sort)
  size_t hk, k0, k1, v0, v1, hist, cur, starts, total;
};

Layout layout_for(int64_t n, int64_t n_cells) {
  Layout L = {};
  const size_t nn = 4 * (size_t)(n > 0 ? n : 1);
  L.counted = use_cell_sort(n, n_cells);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += up256(bytes);
    return at;
  };
  L.hk = take(nn);
  L.k0 = take(nn);
  L.v0 = take(nn);
  if (L.counted) {
    // the sorted keys and ids go straight to the caller's outputs
    L.cur = take(4 * (size_t)(n_cells > 0 ? n_cells : 1));
    L.hist = take(4 * cell_sort_sum_words(n_cells));
    L.starts = take(4 * ((size_t)n_cells + 1));
  } else {
    L.k1 = take(nn);
    L.v1 = take(nn);
    L.hist = take(4 * radix_hist_words(n));
  }
  L.total = off;
  return L;
}

}  // namespace

namespace jdem {

// Look up the device address of `scratch` once per device, with that device
// current (called by jdem_ctx_create).
int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return JDEM_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return JDEM_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return JDEM_EHIP;
  g_scratch_ptr[device] = p;
  return JDEM_OK;
}

}  // namespace jdem

extern "C" {

size_t jdem_noalloc_scratch_bytes(void) { return kNoallocScratchBytes; }

int jdem_cell_build_noalloc(jdem_ctx* ctx, const jdem_state* st, const jdem_domain* dom, const jdem_grid* grid,
                            int32_t* perm, int32_t* sorted_hash, int32_t* cell_start, int32_t* cell_end) {
  if (!ctx) return JDEM_EINVAL;
  if (!st || !dom || !grid || !perm || !sorted_hash || !cell_start || !cell_end)
    return fail(ctx, JDEM_EINVAL, "null argument");
  if (st->n < 0 || (st->n > 0 && !st->pos) || (st->dim != 2 && st->dim != 3))
    return fail(ctx, JDEM_EINVAL, "bad state");
  if (grid->n_cells <= 0 || grid->n_cells > (int64_t)INT32_MAX)
    return fail(ctx, JDEM_EUNSUPPORTED, "cell grid size out of the int32 hash range");
  const int64_t n = st->n;
  const int64_t n_cells = grid->n_cells;
  const Layout L = layout_for(n, n_cells);
  if (L.total > kNoallocScratchBytes)
    return fail(ctx, JDEM_EINVAL,
                "jdem_cell_build_noalloc: " + std::to_string(n) + " particles on " + std::to_string(n_cells) +
                    " cells need " + std::to_string(L.total) + " bytes of scratch, more than the fixed " +
                    std::to_string(kNoallocScratchBytes) + " bytes; use jdem_cell_build");
  if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch_ptr[ctx->device])
    return fail(ctx, JDEM_EHIP, "no-alloc scratch not set up on this device");
  NACHK(ctx, hipSetDevice(ctx->device));
  unsigned char* base = (unsigned char*)g_scratch_ptr[ctx->device];
  hipStream_t s = ctx->stream;
  const DomP dp = dom_of(*dom);
  const GridP gp = grid_of(*grid);
  uint32_t* hk = (uint32_t*)(base + L.hk);
  NACHK(ctx, launch_hash(st->pos, n, st->dim, dp, gp, hk, s));
  SortBufs sb;
  sb.k0 = (uint32_t*)(base + L.k0);
  sb.v0 = (int32_t*)(base + L.v0);
  sb.hist = (uint32_t*)(base + L.hist);
  uint32_t* sk = nullptr;
  int32_t* pm = nullptr;
  if (L.counted) {
    sb.k1 = (uint32_t*)sorted_hash;
    sb.v1 = perm;
    sb.cur = (int32_t*)(base + L.cur);
    int32_t* starts = (int32_t*)(base + L.starts);
    NACHK(ctx, cell_sort_bufs(sb, hk, n, n_cells, &sk, &pm, starts, s));
    // starts of every cell (an empty cell starts where the next one does)
    NACHK(ctx, copy_i32(cell_start, starts, n_cells, s));
    NACHK(ctx, launch_starts_to_ends(starts, n_cells, cell_end, s));
  } else {
    sb.k1 = (uint32_t*)(base + L.k1);
    sb.v1 = (int32_t*)(base + L.v1);
    sb.cur = nullptr;
    NACHK(ctx, radix_sort_bufs(sb, hk, n, nbits_for_cells(n_cells), &sk, &pm, s));
    if (n > 0) {
      NACHK(ctx, copy_i32(sorted_hash, (const int32_t*)sk, n, s));
      NACHK(ctx, copy_i32(perm, pm, n, s));
    }
    NACHK(ctx, launch_fill_i32(cell_start, n_cells, 0, s));
    NACHK(ctx, launch_fill_i32(cell_end, n_cells, 0, s));
    NACHK(ctx, launch_cell_ranges(sk, n, cell_start, cell_end, s));
  }
  return JDEM_OK;
}

}  // extern "C"
