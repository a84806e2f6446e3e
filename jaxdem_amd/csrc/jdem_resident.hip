// jdem_resident.hip -- the NeighborList time step on a resident, cell-sorted
// working set (the hot path of System.step with the NeighborList collider,
// system.py:60-82 + neighbor_list.py:58-131, 594-669).
//
// Layout in HBM for the duration of one jdem_system_step call:
//   P4[b] = (x, y, z, r)   V4[b] = (v, m)   W4[b] = (omega)   Q4[b] = q
//   I4[b] = (inertia, bits: species << 16 | mat_id << 1 | fixed)
// 2D records use their free lanes so that the step kernel reads neither I4
// nor old4: P4 = (x, y, x_build, r), V4 = (vx, vy, y_build, m),
// W4 = (omega, inertia, bits, spring offset).  I4 and old4 are kept as in 3D
// for the other kernels.
// in "work" order b: particles sorted by (cell, original id) at every list
// rebuild, so a particle's candidates are its memory neighbours.  orig[b] is
// the original index of work slot b.  The neighbour list is slot-major
// nl[s*N + b] (work indices) and ends at its first -1; one 64-bit live mask
// per particle and 64 slots marks the slots that carry a tangential spring.
// The springs themselves are stored by rank: xi[k*N + b] is the spring of
// the k-th live slot of particle b, so the rows a wavefront touches are the
// first few, densely used ones.  xi is double-buffered (read one, write the
// other) because a step's new ranks can overtake its old ones.  Sorting by (cell, original id) makes each row the
// reference's row (stencil order, ascending original id within a cell), so
// sums and history match the CPU oracle bit for bit.
//
// The caller's arrays (original order, JaxDEM layouts) are read once on entry
// and written once on exit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "jdem_internal.h"

namespace jdem {

// ---------------------------------------------------------------------------
// float4 helpers
// ---------------------------------------------------------------------------
template <int D>
JD Vec<D> xyz(float4 a);
template <>
JD Vec<3> xyz<3>(float4 a) { return {a.x, a.y, a.z}; }
template <>
JD Vec<2> xyz<2>(float4 a) { return {a.x, a.y}; }
JD float4 f4(Vec<3> v, float w) { return make_float4(v.x, v.y, v.z, w); }
JD float4 f4(Vec<2> v, float w) { return make_float4(v.x, v.y, 0.0f, w); }
template <int D>
JD Ang<D> ang4(float4 a);
template <>
JD Vec<3> ang4<3>(float4 a) { return {a.x, a.y, a.z}; }
template <>
JD Ang1 ang4<2>(float4 a) { return {a.x}; }
JD float4 f4a(Vec<3> v) { return make_float4(v.x, v.y, v.z, 0.0f); }
JD float4 f4a(Ang1 v) { return make_float4(v.z, 0.0f, 0.0f, 0.0f); }
JD Quat q4(float4 a) { return {a.x, a.y, a.z, a.w}; }
JD float4 f4q(Quat q) { return make_float4(q.w, q.x, q.y, q.z); }

// ---------------------------------------------------------------------------
// entry / exit conversions
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_in_state(InOut a) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.n) return;
  const int64_t i = a.orig[b];
  const int D = a.dim;
  const int A = D == 3 ? 3 : 1;
  float p[3] = {0, 0, 0}, v[3] = {0, 0, 0}, w[3] = {0, 0, 0}, f[3] = {0, 0, 0}, t[3] = {0, 0, 0}, I[3] = {0, 0, 0},
        o[3] = {0, 0, 0};
  for (int c = 0; c < D; ++c) {
    p[c] = a.pos[D * i + c];
    v[c] = a.vel[D * i + c];
    f[c] = a.force[D * i + c];
    o[c] = a.old_user ? a.old_user[D * i + c] : 0.0f;
  }
  for (int c = 0; c < A; ++c) {
    w[c] = a.ang[A * i + c];
    t[c] = a.torque[A * i + c];
    I[c] = a.inertia[A * i + c];
  }
  const int bits = ((a.species ? a.species[i] : 0) << 16) | ((a.mat_id ? a.mat_id[i] : 0) << 1) | (a.fixed[i] ? 1 : 0);
  if (a.iso_flag && D == 3) {
    const float r = a.rad[i], Is = (0.4f * a.mass[i]) * (r * r);
    if (a.fixed[i] || I[0] != Is || I[1] != Is || I[2] != Is) *a.iso_flag = 0;
  }
  if (D == 2) {
    float ox = o[0], oy = o[1];
    if (a.old_keep) {
      const float4 ok = a.old4[b];
      ox = ok.x;
      oy = ok.y;
    }
    a.P4[b] = make_float4(p[0], p[1], ox, a.rad[i]);
    a.V4[b] = make_float4(v[0], v[1], oy, a.mass[i]);
    a.W4[b] = make_float4(w[0], I[0], __int_as_float(bits), 0.0f);
  } else {
    a.P4[b] = make_float4(p[0], p[1], p[2], a.rad[i]);
    a.V4[b] = make_float4(v[0], v[1], v[2], a.mass[i]);
    a.W4[b] = make_float4(w[0], w[1], w[2], 0.0f);
  }
  a.Q4[b] = reinterpret_cast<const float4*>(a.q)[i];
  a.I4[b] = make_float4(I[0], I[1], I[2], __int_as_float(bits));
  a.F4[b] = make_float4(f[0], f[1], f[2], 0.0f);
  a.T4[b] = make_float4(t[0], t[1], t[2], 0.0f);
  if (!a.old_keep) a.old4[b] = make_float4(o[0], o[1], o[2], 0.0f);
}

// user (N, M) list of original ids -> slot-major work list; history likewise
__global__ void __launch_bounds__(kBlock) k_in_nlist(InOut a) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.n) return;
  const int64_t i = a.orig[b];
  const int W32 = (a.M + 31) / 32;
  for (int w = 0; w < a.W64; ++w) a.live[(int64_t)w * a.n + b] = 0ull;
  int rank = 0;
  for (int s = 0; s < a.M; ++s) {
    int32_t j = a.nl_user[i * a.M + s];
    if (j >= a.n) j = -1;  // not a particle of this state: treated as padding
    a.nl[(int64_t)s * a.n + b] = j >= 0 ? a.wpos[j] : -1;
    if (a.xi_user && j >= 0) {
      const bool lv = (a.live_user[i * W32 + (s >> 5)] >> (s & 31)) & 1u;
      if (lv) {
        float x[3] = {0, 0, 0};
        for (int c = 0; c < a.dim; ++c) x[c] = a.xi_user[(i * a.M + s) * a.dim + c];
        a.xi[(int64_t)rank * a.n + b] = make_float4(x[0], x[1], x[2], 0.0f);
        ++rank;
        a.live[(int64_t)(s >> 6) * a.n + b] |= 1ull << (s & 63);
      }
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_out_state(InOut a) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.n) return;
  const int64_t i = a.orig[b];
  const int D = a.dim;
  const int A = D == 3 ? 3 : 1;
  const float4 p = a.P4[b], v = a.V4[b], w = a.W4[b], f = a.F4[b], t = a.T4[b], o = a.old4[b];
  const float pp[3] = {p.x, p.y, p.z}, vv[3] = {v.x, v.y, v.z}, ww[3] = {w.x, w.y, w.z}, ff[3] = {f.x, f.y, f.z},
              tt[3] = {t.x, t.y, t.z}, oo[3] = {o.x, o.y, o.z};
  for (int c = 0; c < D; ++c) {
    a.pos[D * i + c] = pp[c];
    a.vel[D * i + c] = vv[c];
    a.force[D * i + c] = ff[c];
    if (a.old_user) a.old_user[D * i + c] = oo[c];
  }
  for (int c = 0; c < A; ++c) {
    a.ang[A * i + c] = ww[c];
    a.torque[A * i + c] = tt[c];
  }
  reinterpret_cast<float4*>(a.q)[i] = a.Q4[b];
}

__global__ void __launch_bounds__(kBlock) k_out_nlist(InOut a) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.n) return;
  const int64_t i = a.orig[b];
  const int W32 = (a.M + 31) / 32;
  uint32_t words[8];
  for (int w = 0; w < W32 && w < 8; ++w) words[w] = 0u;
  bool ended = false;  // work rows end at their first -1
  int rank = 0;
  for (int s = 0; s < a.M; ++s) {
    const int32_t j = ended ? -1 : a.nl[(int64_t)s * a.n + b];
    ended = j < 0;
    a.nl_user[i * a.M + s] = j >= 0 ? a.orig[j] : -1;
    if (a.xi_user) {
      const bool lv = j >= 0 && ((a.live[(int64_t)(s >> 6) * a.n + b] >> (s & 63)) & 1ull);
      if (lv) {
        const float4 x = a.xi[(int64_t)rank * a.n + b];
        ++rank;
        const float xx[3] = {x.x, x.y, x.z};
        for (int c = 0; c < a.dim; ++c) a.xi_user[(i * a.M + s) * a.dim + c] = xx[c];
        if ((s >> 5) < 8) words[s >> 5] |= 1u << (s & 31);
      }
    }
  }
  if (a.xi_user)
    for (int w = 0; w < W32 && w < 8; ++w) a.live_user[i * W32 + w] = words[w];
}

// positions at the last build, work order -> the caller's (N, D)
__global__ void __launch_bounds__(kBlock) k_out_old(InOut a) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.n) return;
  const int64_t i = a.orig[b];
  const float4 o = a.old4[b];
  const float oo[3] = {o.x, o.y, o.z};
  for (int c = 0; c < a.dim; ++c) a.old_user[a.dim * i + c] = oo[c];
}

// Slot-major rows up to and including their first -1 (every reader of a work
// row stops there), so a working set is copied in proportion to its entries.
__global__ void __launch_bounds__(kBlock) k_copy_rows(const int32_t* __restrict__ src, int32_t* __restrict__ dst,
                                                       int64_t n, int M) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  for (int s = 0; s < M; ++s) {
    const int32_t v = src[(int64_t)s * n + b];
    dst[(int64_t)s * n + b] = v;
    if (v < 0) break;
  }
}

// ---------------------------------------------------------------------------
// rebuild: hash in original order, stable sort, permutation of the work set
// ---------------------------------------------------------------------------
template <bool PER>
JD int ccoord(float x, int a, const DomP& dom, const GridP& g) {
  float f;
  if (PER) {
    float t = (x - dom.anchor[a]) / dom.box[a];
    t = t - floorf(t);
    f = floorf(t * (float)g.dims[a]);
  } else {
    f = floorf((x - dom.anchor[a]) / g.cell[a]);
  }
  f = fminf(fmaxf(f, 0.0f), (float)(g.dims[a] - 1));
  return (int)f;
}

// Work-space cell key: the LAST axis varies fastest (z in 3D, y in 2D), so
// the stencil's fastest axis ('ij' meshgrid order, cell_list.py:425-427)
// walks consecutive keys.  Any linearisation gives the same rows (stencil
// order, ascending id within a cell); this one makes a stencil column one
// contiguous range of the sorted work set.
template <int D>
JD uint32_t wkey(const int* c, const GridP& g) {
  if (D == 3) return (uint32_t)((c[0] * g.dims[1] + c[1]) * g.dims[2] + c[2]);
  return (uint32_t)(c[0] * g.dims[1] + c[1]);
}

// keys[orig[b]] = work cell key of P4[b]; wpos[orig[b]] = b
template <int D, bool PER>
__global__ void __launch_bounds__(kBlock) k_hash_work(const float4* __restrict__ P4, const int32_t* __restrict__ orig,
                                                       int64_t n, DomP dom, GridP g, uint32_t* __restrict__ keys,
                                                       int32_t* __restrict__ wpos) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const float4 p = P4[b];
  const float x[3] = {p.x, p.y, p.z};
  int c[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; ++a) c[a] = ccoord<PER>(x[a], a, dom, g);
  const int32_t i = orig[b];
  keys[i] = wkey<D>(c, g);
  wpos[i] = (int32_t)b;
}

// seg[b] = the system whose particle range [offsets[s], offsets[s + 1])
// holds b
__global__ void __launch_bounds__(kBlock) k_seg_ids(const int64_t* __restrict__ offsets, int n_seg, int64_t n,
                                                     int32_t* __restrict__ seg) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  int lo = 0, hi = n_seg;  // largest s with offsets[s] <= b
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offsets[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  seg[b] = lo;
}

// Dense cell starts: starts[h] = first sorted slot with key >= h, for h in
// [0, n_cells]; an empty cell gets the start of the next occupied one, so
// consecutive cells h1..h2 hold the slots [starts[h1], starts[h2 + 1]).
// Thread s (0 <= s <= n) writes the cells in (sk[s-1], sk[s]], the gap
// before slot s, so every cell is written once.
__global__ void __launch_bounds__(kBlock) k_cell_starts(const uint32_t* __restrict__ sk, int64_t n, int64_t n_cells,
                                                         int32_t* __restrict__ starts) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s > n) return;
  const int64_t lo = (s == 0) ? 0 : (int64_t)sk[s - 1] + 1;
  const int64_t hi = (s == n) ? n_cells : (int64_t)sk[s];
  for (int64_t h = lo; h <= hi; ++h) starts[h] = (int32_t)s;
}

// new work slot b' takes old slot src = wpos[new_orig[b']]; inv[src] = b'
__global__ void __launch_bounds__(kBlock) k_permute(int64_t n, const int32_t* __restrict__ new_orig,
                                                     const int32_t* __restrict__ wpos, const float4* __restrict__ P,
                                                     const float4* __restrict__ V, const float4* __restrict__ W,
                                                     const float4* __restrict__ Q, const float4* __restrict__ I,
                                                     float4* __restrict__ P2, float4* __restrict__ V2,
                                                     float4* __restrict__ W2, float4* __restrict__ Q2,
                                                     float4* __restrict__ I2, int32_t* __restrict__ orig2,
                                                     int32_t* __restrict__ src_out, int32_t* __restrict__ inv_out,
                                                     float4* __restrict__ old2, int dim) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const int32_t i = new_orig[b];
  const int32_t s = wpos[i];
  float4 p = P[s];
  float4 v = V[s];
  if (old2) {
    old2[b] = p;  // the positions of this build
    if (dim == 2) {  // and in the 2D records' free lanes
      p.z = p.x;
      v.z = p.y;
    }
  }
  P2[b] = p;
  V2[b] = v;
  W2[b] = W[s];
  Q2[b] = Q[s];
  I2[b] = I[s];
  orig2[b] = i;
  src_out[b] = s;
  inv_out[s] = (int32_t)b;
}

// Batch rebuilds.  mode of system s: KEEP when it did not ask for a
// rebuild, FRESH for its first build, REMAP otherwise; its build count grows
// and its flag is cleared.
__global__ void k_seg_modes(int n_seg, int32_t* __restrict__ seg_flag, int32_t* __restrict__ n_build,
                            int32_t* __restrict__ seg_mode, const int32_t* __restrict__ gate,
                            int32_t* __restrict__ events) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0 && events && gate && *gate) *events += 1;
  if (s >= n_seg) return;
  const int f = seg_flag[s] != 0;
  seg_mode[s] = f ? (n_build[s] > 0 ? SEG_REMAP : SEG_FRESH) : SEG_KEEP;
  n_build[s] += f;
  seg_flag[s] = 0;
}
// Call start, before the first step kernel: a system without a list asks
// for a build (and raises *gate when set), every other flag starts clear.
__global__ void k_seg_init_flags(int n_seg, const int32_t* __restrict__ n_build, int32_t* __restrict__ seg_flag,
                                 int32_t* __restrict__ gate) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_seg) return;
  const bool unbuilt = n_build[s] == 0;
  seg_flag[s] = unbuilt ? 1 : 0;
  if (unbuilt && gate) *gate = 1;
}

// ---- batch rebuilds in place -----------------------------------------------
// Only the systems that asked for a rebuild are re-sorted and rebuilt, each
// inside its own range of work slots; the others are not touched.  Flagged
// particles are numbered compactly: system s's particles take the compact
// range [cpref[s], cpref[s + 1]) in original order, its cells the compact
// key range starting at ccref[s].

// Exclusive prefix sums over the systems of the flagged particle and cell
// counts (one block).
// Wavefronts (64 work slots) that hold particles of system s.
JD int64_t seg_waves(const int64_t* __restrict__ offsets, int s) {
  const int64_t a = offsets[s], e = offsets[s + 1];
  return e > a ? ((e - 1) >> 6) - (a >> 6) + 1 : 0;
}

// Exclusive prefix sums over the systems of the flagged particle, cell and
// wavefront counts (one block).
__global__ void __launch_bounds__(1024) k_seg_prefix(int n_seg, const int32_t* __restrict__ seg_mode,
                                                      const int64_t* __restrict__ offsets,
                                                      const SegP* __restrict__ segp, int64_t* __restrict__ cpref,
                                                      int64_t* __restrict__ ccref, int64_t* __restrict__ wpref) {
  __shared__ int64_t sp[1024], sc[1024], sw[1024];
  const int t = threadIdx.x;
  const int per = (n_seg + 1023) / 1024;
  const int s0 = t * per, s1 = min(n_seg, s0 + per);
  int64_t np = 0, nc = 0, nw = 0;
  for (int s = s0; s < s1; ++s)
    if (seg_mode[s] != SEG_KEEP) {
      np += offsets[s + 1] - offsets[s];
      nc += segp[s].grid.n_cells;
      nw += seg_waves(offsets, s);
    }
  sp[t] = np;
  sc[t] = nc;
  sw[t] = nw;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan
    const int64_t ap = t >= off ? sp[t - off] : 0, ac = t >= off ? sc[t - off] : 0, aw = t >= off ? sw[t - off] : 0;
    __syncthreads();
    sp[t] += ap;
    sc[t] += ac;
    sw[t] += aw;
    __syncthreads();
  }
  int64_t p = sp[t] - np, c = sc[t] - nc, w = sw[t] - nw;
  for (int s = s0; s < s1; ++s) {
    cpref[s] = p;
    ccref[s] = c;
    wpref[s] = w;
    if (seg_mode[s] != SEG_KEEP) {
      p += offsets[s + 1] - offsets[s];
      c += segp[s].grid.n_cells;
      w += seg_waves(offsets, s);
    }
  }
  if (t == 1023) {
    cpref[n_seg] = sp[t];
    ccref[n_seg] = sc[t];
    wpref[n_seg] = sw[t];
  }
}

// The flagged system that holds compact index k of a prefix over the
// systems (cpref, wpref): the last s with pref[s] <= k (systems that are
// not flagged add nothing to the prefix, so the next flagged one shares
// their value and comes after them).
JD int seg_of_compact(const int64_t* __restrict__ pref, int n_seg, int64_t k) {
  int lo = 0, hi = n_seg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pref[mid] <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// compact key of every particle of a flagged system, at its compact index
template <int D>
__global__ void __launch_bounds__(kBlock) k_hash_flagged(RB r) {
  if (r.gate && *r.gate == 0) return;
  const int64_t fn = r.cpref[r.n_seg];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < fn; k += (int64_t)gridDim.x * blockDim.x) {
  const int s = seg_of_compact(r.cpref, r.n_seg, k);
  const int64_t b = r.offsets[s] + (k - r.cpref[s]);
  const SegP& sp = r.segp[s];
  const float4 p = r.P[b];
  const float x[3] = {p.x, p.y, p.z};
  int c[3] = {0, 0, 0};
  const bool per = sp.dom.kind == JDEM_DOMAIN_PERIODIC;
#pragma unroll
  for (int a = 0; a < D; ++a) c[a] = per ? ccoord<true>(x[a], a, sp.dom, sp.grid) : ccoord<false>(x[a], a, sp.dom, sp.grid);
  const int64_t i = r.cpref[s] + (r.orig[b] - r.offsets[s]);
  r.keys_c[i] = (uint32_t)(r.ccref[s] + wkey<D>(c, sp.grid));
  r.slot_c[i] = (int32_t)b;
  }
}

// sorted position k -> staging (k) of the particle's fields and old list
// Flagged particles of the rebuild: r.fn, or (r.fn < 0) the device count.
JD int64_t rb_fn(const RB& r) { return r.fn >= 0 ? r.fn : r.cpref[r.n_seg]; }

// The staging arrays hold the compact index k at stride n (rows, live words
// and springs of k at c * n + k).
__global__ void __launch_bounds__(kBlock) k_place_gather(RB r) {
  if (r.gate && *r.gate == 0) return;
  const int64_t fn = rb_fn(r);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < fn; k += (int64_t)gridDim.x * blockDim.x) {
  const int32_t ob = r.slot_c[r.perm[k]];
  const int s = r.seg[ob];
  const int64_t nb = r.offsets[s] + (k - r.cpref[s]);
  r.tP[k] = r.P[ob];
  r.tV[k] = r.V[ob];
  r.tW[k] = r.W[ob];
  r.tQ[k] = r.Q[ob];
  r.tI[k] = r.I[ob];
  r.torig[k] = r.orig[ob];
  r.cs[k] = s;
  r.inv[ob] = (int32_t)nb;
  if (r.M > 0) {
    int c = 0;
    for (; c < r.M; ++c) {
      const int32_t j = r.nl[(int64_t)c * r.n + ob];
      r.tnl[(int64_t)c * r.n + k] = j;
      if (j < 0) break;
    }
    if (r.live) {
      int nlive = 0;
      for (int w = 0; w < r.W64; ++w) {
        const uint64_t bits = r.live[(int64_t)w * r.n + ob];
        r.tlive[(int64_t)w * r.n + k] = bits;
        nlive += __popcll(bits);
      }
      for (int q = 0; q < nlive; ++q) r.txi[(int64_t)q * r.n + k] = r.xi[(int64_t)q * r.n + ob];
    }
  }
  }
}

// staging (k) -> the particle's new work slot in its system's range
__global__ void __launch_bounds__(kBlock) k_place_scatter(RB r) {
  if (r.gate && *r.gate == 0) return;
  const int64_t fn = rb_fn(r);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < fn; k += (int64_t)gridDim.x * blockDim.x) {
  const int s = r.cs[k];
  const int64_t nb = r.offsets[s] + (k - r.cpref[s]);
  float4 p = r.tP[k];
  float4 v = r.tV[k];
  r.old4[nb] = p;  // build position
  if (r.dim == 2) {  // and in the 2D records' free lanes
    p.z = p.x;
    v.z = p.y;
  }
  r.P[nb] = p;
  r.V[nb] = v;
  r.W[nb] = r.tW[k];
  r.Q[nb] = r.tQ[k];
  r.I[nb] = r.tI[k];
  r.orig[nb] = r.torig[k];
  r.cidx[nb] = (int32_t)k;
  }
}

// cell starts of the flagged systems: thread k fills the cells of its system
// from the one after its predecessor's key to its own key (the last particle
// of a system also fills to the end of the system's grid)
__global__ void __launch_bounds__(kBlock) k_cell_starts_flagged(RB r, const uint32_t* __restrict__ sk,
                                                                 int32_t* __restrict__ starts) {
  if (r.gate && *r.gate == 0) return;
  const int64_t fn = rb_fn(r);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < fn; k += (int64_t)gridDim.x * blockDim.x) {
  const int s = r.cs[k];
  const SegP& sp = r.segp[s];
  const int64_t first = r.cpref[s], last = r.cpref[s + 1] - 1;
  const int64_t key = (int64_t)sk[k] - r.ccref[s];
  const int64_t lo = (k == first) ? 0 : (int64_t)sk[k - 1] - r.ccref[s] + 1;
  const int32_t slot = (int32_t)(r.offsets[s] + (k - first));
  for (int64_t h = lo; h <= key; ++h) starts[sp.cell_base + h] = slot;
  if (k == last)
    for (int64_t h = key + 1; h <= sp.grid.n_cells; ++h) starts[sp.cell_base + h] = (int32_t)r.offsets[s + 1];
  }
}

// neighbour list in work space (cell_list.py:500-595).  The stencil is
// walked in the reference's 'ij' order (cell_list.py:425-427), last axis
// fastest; a run of stencil cells along that axis is one contiguous range of
// work slots, [starts[h_first], starts[h_last + 1]), split in two where a
// periodic axis wraps.  Candidates within a cell come in ascending original
// id, the order of the reference's stable hash sort.

JD bool bonded_o(const int32_t* bond_id, int bond_w, int64_t i, int32_t j) {
  for (int w = 0; w < bond_w; ++w)
    if (bond_id[i * bond_w + w] == j) return true;
  return false;
}

// Cells of the stencil along one axis, in stencil order: the offsets
// -sr..hi (hi = min(sr, dims - sr - 1) on a periodic axis, which drops the
// offsets that would revisit a cell) give at most two runs of consecutive
// cells, [r0a, r0b] then [r1a, r1b] (empty when a > b).
JD void stencil_runs(bool per, int c, int sr, int dims, int& r0a, int& r0b, int& r1a, int& r1b) {
  if (per) {
    const int cnt = sr + min(sr, dims - sr - 1) + 1;  // in [1, dims]
    int start = (c - sr) % dims;
    if (start < 0) start += dims;
    const int last = start + cnt - 1;
    r0a = start;
    r0b = min(last, dims - 1);
    r1a = 0;
    r1b = last - dims;
  } else {
    r0a = max(c - sr, 0);
    r0b = min(c + sr, dims - 1);
    r1a = 0;
    r1b = -1;
  }
}

// Offsets along a slow axis; returns the cell or -1 when it is off the grid.
JD int stencil_cell(bool per, int c, int o, int dims) {
  int x = c + o;
  if (per) {
    x %= dims;
    if (x < 0) x += dims;
    return x;
  }
  return (x < 0 || x >= dims) ? -1 : x;
}

// Number of set bits of the live words (w0..w3) below slot s.
JD int rank_below(uint64_t w0, uint64_t w1, uint64_t w2, uint64_t w3, int s) {
  const int ws = s >> 6;
  const uint64_t below = (1ull << (s & 63)) - 1ull;
  int r = 0;
  r += ws > 0 ? __popcll(w0) : __popcll(w0 & below);
  if (ws >= 1) r += ws > 1 ? __popcll(w1) : __popcll(w1 & below);
  if (ws >= 2) r += ws > 2 ? __popcll(w2) : __popcll(w2 & below);
  if (ws >= 3) r += __popcll(w3 & below);
  return r;
}

// History of a freshly built row (neighbor_list.py:29-54): the history of
// new pair (b, j) is the old history of the same two particles, zero when
// they were not listed together.  Only live old slots carry history, so each
// one is looked up in the new row (row(s), s < nrow): old slot t of src[b]
// lists old work slot o, now at inv[o].  A pair occurs at most once per row,
// so the lookups are one-to-one.  The first sweep finds the new live slots,
// the second copies each spring from its old rank to its new one.  Rows hold
// at most 256 slots (four live words).
template <class Row>
JD int find_in_row(int32_t tgt, int nrow, Row row) {
  for (int s = 0; s < nrow; ++s)
    if (row(s) == tgt) return s;
  return -1;
}

// exclusive prefix sum over the 64 lanes of a wave; every lane must call it
JD int wave_exclusive_scan(int v, int lane) {
  int incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int t = __shfl_up(incl, d);
    if (lane >= d) incl += t;
  }
  return incl - v;
}

// With new_W set, every lane of the wave must call it (inactive lanes with
// active = false): the new springs are placed per wavefront after a
// wave-wide scan of the new counts.
template <class Row>
JD void remap_row(const NLW& a, int64_t b, bool active, int nrow, Row row, bool remap) {
  if (!a.live) return;
  const int64_t n = a.n;
  uint64_t out0 = 0ull, out1 = 0ull, out2 = 0ull, out3 = 0ull;
  // the old row of the particle now at b: at slot src[b] of the previous
  // buffers, or (batches) at compact index cidx[b] of the staged rows
  const bool go = active && remap;
  const int64_t ob = go ? (a.cidx ? (int64_t)a.cidx[b] : (int64_t)a.src[b]) : 0;
  const int64_t os = a.cidx ? a.ostride : n;
  // start of the old springs, read before any lane writes new offsets (old_W
  // and new_W are one buffer when the rebuild does not re-sort)
  const int64_t obase = (go && a.old_W) ? (ob & ~(int64_t)63) * a.M + __float_as_int(a.old_W[ob].w) : 0;
  // visit(t, o_slot_new): each live old slot t whose partner is listed
  // again, with the rank of t among the live old slots
  auto sweep = [&](auto&& f) {
    int rank_old = 0;
    for (int w = 0; w < a.W64; ++w) {
      uint64_t bits = a.old_live[(int64_t)w * os + ob];
      while (bits) {
        const int t = w * 64 + (__ffsll((unsigned long long)bits) - 1);
        bits &= bits - 1ull;
        const int32_t o = a.old_nl[(int64_t)t * os + ob];
        const int s = o < 0 ? -1 : find_in_row(a.inv[o], nrow, row);
        if (s >= 0) f(s, rank_old);
        ++rank_old;
      }
    }
  };
  if (go)
    sweep([&](int s, int) {
      const uint64_t bit = 1ull << (s & 63);
      const int ws = s >> 6;
      out0 |= ws == 0 ? bit : 0ull;
      out1 |= ws == 1 ? bit : 0ull;
      out2 |= ws == 2 ? bit : 0ull;
      out3 |= ws == 3 ? bit : 0ull;
    });
  int64_t wbase = 0;
  if (a.new_W) {
    const int cnt = __popcll(out0) + __popcll(out1) + __popcll(out2) + __popcll(out3);
    const int off = wave_exclusive_scan(cnt, (int)(threadIdx.x & 63));
    if (active) a.new_W[b].w = __int_as_float(off);
    wbase = (b & ~(int64_t)63) * a.M + off;
  }
  if (go)
    sweep([&](int s, int rank_old) {
      const int rank_new = rank_below(out0, out1, out2, out3, s);
      const int64_t oi = a.old_W ? obase + rank_old : (int64_t)rank_old * os + ob;
      a.xi[a.new_W ? wbase + rank_new : (int64_t)rank_new * n + b] = a.old_xi[oi];
    });
  if (!active) return;
  a.live[b] = out0;
  if (a.W64 > 1) a.live[n + b] = out1;
  if (a.W64 > 2) a.live[2 * n + b] = out2;
  if (a.W64 > 3) a.live[3 * n + b] = out3;
}

// A row ends at its first -1 (or at M slots); the slots after the
// terminator are not written.
JD void end_row(const NLW& a, int64_t b, bool active, int count) {
  if (active && count < a.M) a.nl[(int64_t)count * a.n + b] = -1;
  if (active && count > a.M) atomicOr(a.overflow, 1);
  remap_row(a, b, active, min(count, a.M), [&](int s) { return a.nl[(int64_t)s * a.n + b]; }, a.remap != 0);
}



// The first kRowLds slots of each row are collected in LDS, slot-major, and
// written out one slot at a time for the whole workgroup, so the stores of a
// slot are contiguous; later slots go straight to memory.
constexpr int kRowLds = 16;

// PM: 0 = one non-periodic system, 1 = one periodic system, 2 = a segmented
// batch, where each work slot takes the grid, domain and cutoff of its system.
template <int D, int PM>
__global__ void __launch_bounds__(kBlock) k_nlist_work(NLW a) {
  if (a.gate && *a.gate == 0) return;  // uniform over the grid
  __shared__ int32_t rows[kRowLds][kBlock];
  __shared__ int row_max;
  const int tid = threadIdx.x;
  // one tile of kBlock slots per workgroup, or (a grid smaller than the
  // tiles, as batches launch it) several in turn.  A batch with wpref walks
  // the wavefronts of the flagged systems only: a tile is four of them.
  const bool by_wave = PM == 2 && a.wpref;
  const int64_t nwaves = by_wave ? a.wpref[a.n_seg] : 0;
  const int64_t ntiles = by_wave ? (nwaves + kBlock / 64 - 1) / (kBlock / 64) : (a.n + kBlock - 1) / kBlock;
  for (int64_t tile = xcd_tile(blockIdx.x, gridDim.x); tile < ntiles; tile += gridDim.x) {
  int64_t b = tile * blockDim.x + tid;
  if (by_wave) {
    const int64_t cw = tile * (kBlock / 64) + (tid >> 6);
    if (cw < nwaves) {
      const int s = seg_of_compact(a.wpref, a.n_seg, cw);
      b = ((a.offsets[s] >> 6) + (cw - a.wpref[s])) * 64 + (tid & 63);
    } else {
      b = a.n;
    }
  }
  // a batch workgroup whose systems all keep their rows has nothing to do
  if (PM == 2 && !__syncthreads_or(b < a.n && a.seg_mode[a.seg[b]] != SEG_KEEP)) continue;
  if (tid == 0) row_max = 0;
  __syncthreads();
  bool remap = a.remap != 0;
  if (PM == 2 && b < a.n) {
    const int mode = a.seg_mode[a.seg[b]];
    if (mode == SEG_KEEP) b = a.n;  // the system keeps its rows (the barrier below is still reached)
    remap = mode == SEG_REMAP;
  }
  const bool active = b < a.n;
  int count = 0;
  if (active) {
    GridP grid;
    DomP dom;
    float cutoff_sq;
    uint32_t cbase = 0, key;
    if (PM == 2) {
      const int sg = a.seg[b];
      const SegP& sp = a.segp[sg];
      grid = sp.grid;
      dom = sp.dom;
      cutoff_sq = sp.cutoff_sq;
      cbase = sp.cell_base;
      key = a.skeys[a.cidx[b]] - (uint32_t)a.ccref[sg];  // batches sort the flagged systems compactly
    } else {
      grid = a.grid;
      dom = a.dom;
      cutoff_sq = a.cutoff_sq;
      key = a.skeys[b];
    }
    const bool PER = PM == 2 ? dom.kind == JDEM_DOMAIN_PERIODIC : PM == 1;
    const Vec<D> xi = xyz<D>(a.P4[b]);
    const int d0 = grid.dims[0], d1 = grid.dims[1], d2 = (D == 3) ? grid.dims[2] : 1;
    int c0, c1, cf;
    if (D == 3) {
      cf = (int)(key % (uint32_t)d2);
      const uint32_t r = key / (uint32_t)d2;
      c1 = (int)(r % (uint32_t)d1);
      c0 = (int)(r / (uint32_t)d1);
    } else {
      cf = (int)(key % (uint32_t)d1);
      c0 = (int)(key / (uint32_t)d1);
      c1 = 0;
    }
    const int sr = grid.sr;
    const int dF = (D == 3) ? d2 : d1;
    int r0a, r0b, r1a, r1b;
    stencil_runs(PER, cf, sr, dF, r0a, r0b, r1a, r1b);
    const int hi0 = PER ? min(sr, d0 - sr - 1) : sr;
    const int lo1 = (D == 3) ? -sr : 0;
    const int hi1 = (D == 3) ? (PER ? min(sr, d1 - sr - 1) : sr) : 0;
    const bool chk_bond = (!a.isb) && a.bond_w > 0;
    const int64_t ib = chk_bond ? (int64_t)a.orig[b] : 0;
    const int32_t* __restrict__ starts = a.starts;
    for (int ox = -sr; ox <= hi0; ++ox) {
      const int cx = stencil_cell(PER, c0, ox, d0);
      if (cx < 0) continue;
      for (int oy = lo1; oy <= hi1; ++oy) {
        int64_t base = cbase;
        if (D == 3) {
          const int cy = stencil_cell(PER, c1, oy, d1);
          if (cy < 0) continue;
          base += (int64_t)(cx * d1 + cy) * d2;
        } else {
          base += (int64_t)cx * d1;
        }
        // both runs' bounds are loaded before either is walked
        const int k0 = starts[base + r0a], k0e = starts[base + r0b + 1];
        const int k1 = (r1b >= r1a) ? starts[base + r1a] : 0;
        const int k1e = (r1b >= r1a) ? starts[base + r1b + 1] : 0;
#pragma unroll
        for (int run = 0; run < 2; ++run) {
          const int kb = run ? k1 : k0, ke = run ? k1e : k0e;
          // candidates in groups of four, loaded back to back
          for (int kg = kb; kg < ke; kg += 4) {
            float4 pk[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) pk[u] = (kg + u < ke) ? a.P4[kg + u] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int k = kg + u;
              if (k >= ke) break;
              if (k == b) continue;
              const float dsq =
                  PER ? dist_sq<true>(xi, xyz<D>(pk[u]), dom) : dist_sq<false>(xi, xyz<D>(pk[u]), dom);
              if (!(dsq <= cutoff_sq)) continue;
              if (chk_bond && bonded_o(a.bond_id, a.bond_w, ib, a.orig[k])) continue;
              if (count < kRowLds)
                rows[count][tid] = k;
              else if (count < a.M)
                a.nl[(int64_t)count * a.n + b] = k;
              ++count;
            }
          }
        }
      }
    }
    atomicMax(&row_max, min(count, kRowLds));
  }
  __syncthreads();
  const int c = min(count, a.M);
  if (active) {
    const int staged = min(row_max, a.M);
    for (int s = 0; s < staged; ++s)
      if (s < c) a.nl[(int64_t)s * a.n + b] = rows[s][tid];
    if (count < a.M) a.nl[(int64_t)count * a.n + b] = -1;
    if (count > a.M) atomicOr(a.overflow, 1);
  }
  remap_row(a, b, active, c, [&](int s) { return s < kRowLds ? rows[s][tid] : a.nl[(int64_t)s * a.n + b]; },
            remap);
  __syncthreads();  // the next tile reuses rows and row_max
  }
}



__global__ void __launch_bounds__(kBlock) k_copy_f4(float4* __restrict__ d, const float4* __restrict__ s, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

__global__ void __launch_bounds__(kBlock) k_iota(int32_t* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (int32_t)i;
}


// ---------------------------------------------------------------------------
// spring layouts: by rank <-> per wavefront
// ---------------------------------------------------------------------------

JD bool wave_touched(int64_t b, int64_t n, const int32_t* seg, const int32_t* seg_mode) {
  if (!seg_mode) return true;
  const bool mine = b < n && seg_mode[seg[b]] != SEG_KEEP;
  return __ballot(mine) != 0ull;
}

__global__ void __launch_bounds__(kBlock) k_pack_springs(int64_t n, int M, int W64, const float4* __restrict__ rank,
                                                          const uint64_t* __restrict__ live, float4* __restrict__ wave,
                                                          float4* __restrict__ W, const int32_t* __restrict__ seg,
                                                          const int32_t* __restrict__ seg_mode,
                                                          const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int lane = threadIdx.x & 63;
  const int64_t npad = (n + 63) & ~(int64_t)63;  // whole waves: every lane of a wave loops alike
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < npad; b += stride) {
    if (!wave_touched(b, n, seg, seg_mode)) continue;  // wave-uniform
    int cnt = 0;
    if (b < n)
      for (int w = 0; w < W64; ++w) cnt += __popcll(live[(int64_t)w * n + b]);
    const int off = wave_exclusive_scan(cnt, lane);
    if (b >= n) continue;
    float4* dst = wave + (b & ~(int64_t)63) * M + off;
    for (int k = 0; k < cnt; ++k) dst[k] = rank[(int64_t)k * n + b];
    W[b].w = __int_as_float(off);
  }
}

// The same conversions in a batch rebuild, over the wavefronts that hold a
// flagged system's particles only (a wavefront shared by two flagged
// systems is converted twice, to the same values).
__global__ void __launch_bounds__(kBlock) k_pack_springs_seg(int64_t n, int M, int W64, const float4* __restrict__ rank,
                                                              const uint64_t* __restrict__ live,
                                                              float4* __restrict__ wave, float4* __restrict__ W,
                                                              const int64_t* __restrict__ wpref,
                                                              const int64_t* __restrict__ offsets, int n_seg,
                                                              const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int lane = threadIdx.x & 63;
  const int64_t nt = wpref[n_seg] * 64;  // whole waves: every lane of a wave loops alike
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += stride) {
    const int64_t cw = t >> 6;
    const int s = seg_of_compact(wpref, n_seg, cw);
    const int64_t b = ((offsets[s] >> 6) + (cw - wpref[s])) * 64 + lane;
    int cnt = 0;
    if (b < n)
      for (int w = 0; w < W64; ++w) cnt += __popcll(live[(int64_t)w * n + b]);
    const int off = wave_exclusive_scan(cnt, lane);
    if (b >= n) continue;
    float4* dst = wave + (b & ~(int64_t)63) * M + off;
    for (int k = 0; k < cnt; ++k) dst[k] = rank[(int64_t)k * n + b];
    W[b].w = __int_as_float(off);
  }
}

__global__ void __launch_bounds__(kBlock) k_unpack_springs_seg(int64_t n, int M, int W64,
                                                                const float4* __restrict__ wave,
                                                                const uint64_t* __restrict__ live,
                                                                const float4* __restrict__ W, float4* __restrict__ rank,
                                                                const int64_t* __restrict__ wpref,
                                                                const int64_t* __restrict__ offsets, int n_seg,
                                                                const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int lane = threadIdx.x & 63;
  const int64_t nt = wpref[n_seg] * 64;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += stride) {
    const int64_t cw = t >> 6;
    const int s = seg_of_compact(wpref, n_seg, cw);
    const int64_t b = ((offsets[s] >> 6) + (cw - wpref[s])) * 64 + lane;
    if (b >= n) continue;
    int cnt = 0;
    for (int w = 0; w < W64; ++w) cnt += __popcll(live[(int64_t)w * n + b]);
    const float4* src = wave + (b & ~(int64_t)63) * M + __float_as_int(W[b].w);
    for (int k = 0; k < cnt; ++k) rank[(int64_t)k * n + b] = src[k];
  }
}

__global__ void __launch_bounds__(kBlock) k_unpack_springs(int64_t n, int M, int W64, const float4* __restrict__ wave,
                                                            const uint64_t* __restrict__ live,
                                                            const float4* __restrict__ W, float4* __restrict__ rank,
                                                            const int32_t* __restrict__ seg,
                                                            const int32_t* __restrict__ seg_mode,
                                                            const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int64_t npad = (n + 63) & ~(int64_t)63;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < npad; b += stride) {
    if (!wave_touched(b, n, seg, seg_mode) || b >= n) continue;
    int cnt = 0;
    for (int w = 0; w < W64; ++w) cnt += __popcll(live[(int64_t)w * n + b]);
    const float4* src = wave + (b & ~(int64_t)63) * M + __float_as_int(W[b].w);
    for (int k = 0; k < cnt; ++k) rank[(int64_t)k * n + b] = src[k];
  }
}

// ---------------------------------------------------------------------------
// the fused step on the work set
// ---------------------------------------------------------------------------

// Minimum waves per SIMD asked of the compiler: six (80 VGPRs) holds the
// spring-friction law and the integrators; the Cundall-Strack and
// Hertz-friction laws in 3D need five, and one fewer in a segmented batch,
// whose per-system domain takes registers too, or with rows of more than 64
// slots, whose slot masks take four words.
template <int D, int LAW, int PM, int MW>
constexpr int ws_waves() {
  const bool heavy = D == 3 && (LAW == JDEM_LAW_CUNDALL_STRACK || LAW == JDEM_LAW_HERTZ_FRICTION ||
                                LawTraits<LAW>::routed);
  const bool seg3 = D == 3 && PM == 2 && LAW != JDEM_LAW_SPRING && LAW != JDEM_LAW_HERTZ;
  const bool wide = MW > 1 && LawTraits<LAW>::hist;
  const int w = heavy ? (PM == 2 ? 4 : 5) : (seg3 ? 5 : 6);
  return wide ? w - 1 : w;
}

// Domain._displacement for mode PM: 0 / 1 = one non-periodic / periodic
// system, 2 = the domain of the slot's system in a segmented batch.
template <int PM, int D>
JD Vec<D> disp_pm(Vec<D> a, Vec<D> b, const DomP& d) {
  if (PM == 2) return d.kind == JDEM_DOMAIN_PERIODIC ? disp_<true>(a, b, d) : disp_<false>(a, b, d);
  return disp_<PM == 1>(a, b, d);
}

// LDS of the lane-balanced force pass: per wave, kPairCap pair descriptors
// (a wave with more marked slots takes them in several chunks) and one round
// of 64 pair results (force + contact flag, torque, spring).
constexpr int kPairCap = 128;
// pass 1 keeps the partner index of a lane's first kStash marked slots in LDS,
// so the pair evaluation need not reload it from the list
constexpr int kStash = 4;

// orders the LDS traffic of one wave around the exchange of pair results
JD void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// MW = 64-slot words of a row the kernel handles (1 when M <= 64, else 4):
// the per-lane slot masks then take 2 * MW registers
// BS = workgroup size (the tile width: 256, or 128 for batches when the
// autotuner prefers it)
template <int D, int LAW, int PM, int MW, int BS = kBlock>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(ws_waves<D, LAW, PM, MW>(), 8))) k_ws_step(WS a) {
  __shared__ int pdesc[BS / 64][kPairCap];
  __shared__ int pjx[BS / 64][kPairCap];
  __shared__ int pstash[BS / 64][kStash][64];
  __shared__ float4 pres[BS / 64][3][64];
  if (a.flag_rd && *a.flag_rd) {
    if (a.flag_wr) *a.flag_wr = *a.flag_rd;
    return;
  }
  // every lane of a wave stays to the end of the force pass, which works
  // wave-wide; a lane past the end reads the last particle and writes nothing
  const int64_t b0 = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const bool active = b0 < a.n;
  const int64_t b = active ? b0 : a.n - 1;
  DomP dom;
  float dt, trigger;
  if (PM == 2) {
    const SegP& sp = a.segp[a.seg[b]];
    dom = sp.dom;
    dt = sp.dt;
    trigger = sp.trigger;
  } else {
    dom = a.dom;
    dt = a.dt;
    trigger = a.trigger;
  }
  // the force pass needs only this particle's position and radius; the
  // integration reloads them with the velocities and inertia after it,
  // which keeps them out of the registers the pair evaluation needs
  const float4 p4 = a.P[b];
  Side<D> me;
  me.x = xyz<D>(p4);
  me.r = p4.w;
  Vec<D> F;
  Ang<D> T;
  int spring_off = -1;  // start of this slot's new springs in its wave's region
  if (a.do_force) {
    F = zero_vec<D>();
    T = zero_ang<D>();
    // Pass 1 marks the slots whose pair is within reach of a contact,
    // d^2 <= R^2 * 1.0001 (a margin over the law's own overlap test); pass 2
    // evaluates the law on the marked slots only, in slot order, so a wave
    // runs the law as often as its busiest lane has contacts.  Pass 1 loads
    // PF list entries, then their positions, back to back before using any.
    constexpr int PF = 4;
    uint64_t cm[MW];
#pragma unroll
    for (int q = 0; q < MW; ++q) cm[q] = 0ull;
    int n_marked = 0;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    bool done = false;
    for (int s0 = 0; s0 < a.M && !done; s0 += PF) {
      int32_t js[PF];
      float4 pjs[PF];
#pragma unroll
      for (int k = 0; k < PF; ++k) js[k] = (s0 + k < a.M) ? a.nl[(int64_t)(s0 + k) * a.n + b] : -1;
      // the row ends at its first -1: what follows is stale
#pragma unroll
      for (int k = 1; k < PF; ++k) js[k] = js[k - 1] < 0 ? -1 : js[k];
#pragma unroll
      for (int k = 0; k < PF; ++k) pjs[k] = (js[k] >= 0) ? a.P[js[k]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int k = 0; k < PF; ++k) {
        if (js[k] < 0) {
          done = true;
          break;
        }
        const float4 pj = pjs[k];
        const Vec<D> rij = disp_pm<PM>(me.x, xyz<D>(pj), dom);
        const float d_sq = dot(rij, rij);
        const float R = me.r + pj.w;
        if (d_sq > (R * R) * 1.0001f) continue;
        if (n_marked < kStash) pstash[wv][n_marked][lane] = js[k];
        ++n_marked;
        const int s = s0 + k;
        const uint64_t bit = 1ull << (s & 63);
        const int ws = s >> 6;
#pragma unroll
        for (int q = 0; q < MW; ++q) cm[q] |= ws == q ? bit : 0ull;
      }
    }
    const int n_words = LawTraits<LAW>::hist ? a.W64 : (a.M + 63) >> 6;
    int nc = 0;
#pragma unroll
    for (int q = 0; q < MW; ++q) {
      if (!active) cm[q] = 0ull;
      nc += __popcll(cm[q]);
    }
    // Pass 2, lane-balanced: the wave's marked slots are numbered lane by
    // lane (slot order within a lane) and dealt out 64 at a time, one pair
    // per lane, so the law runs ceil(total / 64) times per wave instead of
    // as often as the busiest lane has contacts.  Each owner then adds its
    // pairs' results in slot order: the sums, springs and live bits are
    // those of one lane walking its own row.
    int incl = nc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(incl, d);
      if (lane >= d) incl += t;
    }
    const int total = __shfl(incl, 63);
    const int off = incl - nc;
    spring_off = off;
    uint64_t ln[MW];
#pragma unroll
    for (int q = 0; q < MW; ++q) ln[q] = 0ull;
    int rank_new = 0;  // springs are stored by rank
    for (int cb = 0; cb < total; cb += kPairCap) {  // wave-uniform
      // descriptors of pairs cb .. cb + kPairCap - 1: owner lane | slot << 6
      // | (old spring's place in the wave's region + 1) << 14, and the
      // partner index
      if (off < cb + kPairCap && off + nc > cb) {
        // old springs: this slot's start in the wave's spring region
        int idx = off, rank_base = LawTraits<LAW>::hist ? __float_as_int(a.W[b].w) : 0, c = 0;
#pragma unroll
        for (int w = 0; w < MW; ++w) {
          uint64_t cmw = w < n_words ? cm[w] : 0ull;
          uint64_t live_old = 0ull;
          if constexpr (LawTraits<LAW>::hist)
            if (w < n_words) live_old = a.live[(int64_t)w * a.n + b];
#pragma unroll 1
          while (cmw) {
            const int sl = __ffsll((unsigned long long)cmw) - 1;
            cmw &= cmw - 1ull;
            int k = -1;
            if constexpr (LawTraits<LAW>::hist)
              if ((live_old >> sl) & 1ull) k = rank_base + __popcll(live_old & ((1ull << sl) - 1ull));
            if (idx >= cb && idx < cb + kPairCap) {
              pdesc[wv][idx - cb] = lane | ((w * 64 + sl) << 6) | ((k + 1) << 14);
              // the LDS read is unconditional, so the two sources stay two
              // loads (not one flat load through a selected pointer)
              int jj = pstash[wv][min(c, kStash - 1)][lane];
              if (c >= kStash) jj = a.nl[(int64_t)(w * 64 + sl) * a.n + b];
              pjx[wv][idx - cb] = jj;
            }
            ++idx;
            ++c;
          }
          if constexpr (LawTraits<LAW>::hist) rank_base += __popcll(live_old);
        }
      }
      const int tc = min(total - cb, kPairCap);
      for (int r0 = 0; r0 < tc; r0 += 64) {
        wave_sync();
        if (r0 + lane < tc) {
          const int dsc = pdesc[wv][r0 + lane];
          const int32_t j = pjx[wv][r0 + lane];
          const int64_t bo = b0 - lane + (dsc & 63);
          const int k = (dsc >> 14) - 1;
          // every load of the pair is issued before any is used
          const float4 po = a.P[bo], pj = a.P[j];
          float4 vo, vj, wo, wj, xo;
          if constexpr (LawTraits<LAW>::kin) {
            vo = a.V[bo];
            vj = a.V[j];
            wo = a.W[bo];
            wj = a.W[j];
          }
          if constexpr (LawTraits<LAW>::hist) xo = a.xi_rd[(b0 - lane) * a.M + (k >= 0 ? k : 0)];
          DomP dom_o;
          float dt_o;
          if (PM == 2) {
            const SegP& sp = a.segp[a.seg[bo]];
            dom_o = sp.dom;
            dt_o = sp.dt;
          } else {
            dom_o = dom;
            dt_o = dt;
          }
          Side<D> mo, o;
          mo.x = xyz<D>(po);
          mo.r = po.w;
          o.x = xyz<D>(pj);
          o.r = pj.w;
          mo.mat = 0;
          o.mat = 0;
          mo.sp = 0;
          o.sp = 0;
          if (a.law.n_mat > 1 || (LawTraits<LAW>::routed && a.law.n_species > 1)) {
            const int bo_bits = D == 2 ? __float_as_int(a.W[bo].z) : __float_as_int(a.I[bo].w);
            const int j_bits = D == 2 ? __float_as_int(a.W[j].z) : __float_as_int(a.I[j].w);
            mo.mat = (bo_bits >> 1) & kMatMask;
            o.mat = (j_bits >> 1) & kMatMask;
            mo.sp = (int)((unsigned)bo_bits >> 16);
            o.sp = (int)((unsigned)j_bits >> 16);
          }
          if constexpr (LawTraits<LAW>::kin) {
            mo.v = xyz<D>(vo);
            mo.m = vo.w;
            mo.w = ang4<D>(wo);
            o.v = xyz<D>(vj);
            o.m = vj.w;
            o.w = ang4<D>(wj);
          }
          Vec<D> xold = zero_vec<D>();
          if constexpr (LawTraits<LAW>::hist)
            if (k >= 0) xold = xyz<D>(xo);
          const Vec<D> rij = disp_pm<PM>(mo.x, o.x, dom_o);
          const float d_sq = dot(rij, rij);
          const PairOut<D> r = law_eval<D, LAW>(a.law, mo, o, rij, d_sq, xold, dt_o);
          pres[wv][0][lane] = f4(r.F, r.contact ? 1.0f : 0.0f);
          if constexpr (LawTraits<LAW>::torque) pres[wv][1][lane] = f4a(r.T);
          if constexpr (LawTraits<LAW>::hist) pres[wv][2][lane] = f4(r.xi, 0.0f);
        }
        wave_sync();
        const int q0 = cb + r0, q1 = min(off + nc, q0 + 64);
#pragma unroll 1
        for (int q = max(off, q0); q < q1; ++q) {
          const float4 fr = pres[wv][0][q - q0];
          F = F + xyz<D>(fr);
          if constexpr (LawTraits<LAW>::torque) T = T + ang4<D>(pres[wv][1][q - q0]);
          if constexpr (LawTraits<LAW>::hist) {
            if (fr.w != 0.0f) {
              const int sl = (pdesc[wv][q - cb] >> 6) & 255;
              // new springs from the slot's marked-pair prefix: the wave's
              // springs fill one contiguous run of its region
              a.xi_wr[(b0 - lane) * a.M + off + rank_new] = pres[wv][2][q - q0];
              ++rank_new;
              const uint64_t bit = 1ull << (sl & 63);
              const int ws = sl >> 6;
#pragma unroll
              for (int qq = 0; qq < MW; ++qq) ln[qq] |= ws == qq ? bit : 0ull;
            }
          }
        }
      }
      wave_sync();  // the next chunk's descriptors overwrite these
    }
    if constexpr (LawTraits<LAW>::hist)
      if (active)
#pragma unroll
        for (int w = 0; w < MW; ++w)
          if (w < n_words) {  // most words are unchanged from the last step: skip their stores
            const int64_t li = (int64_t)w * a.n + b;
            if (a.live[li] != ln[w]) a.live[li] = ln[w];
          }
  } else {
    F = xyz<D>(a.F[b]);
    T = ang4<D>(a.T[b]);
  }
  if (!active) return;
  const float4 p4b = a.P[b], v4 = a.V[b], w4 = a.W[b];
  float4 i4;
  if constexpr (D == 2) {
    i4 = make_float4(w4.y, 0.0f, 0.0f, w4.z);
  } else if (a.iso && *a.iso) {
    const float Is = (0.4f * v4.w) * (p4b.w * p4b.w);
    i4 = make_float4(Is, Is, Is, __int_as_float(0));
  } else {
    i4 = a.I[b];
  }
  const int bits = __float_as_int(i4.w);
  Vec<D> x = xyz<D>(p4b), v = xyz<D>(v4);
  Ang<D> w = ang4<D>(w4);
  Quat q = q4(a.Q[b]);
  const Ang<D> I = ang4<D>(i4);
  const float m = v4.w;
  const bool fixed = (bits & 1) != 0;
  const float freef = fixed ? 0.0f : 1.0f;
  if (a.do_after) {
    if (a.has_fm) {
      const int64_t i = a.orig[b];
      Vec<D> fp = a.ext_f ? ld<D>(a.ext_f, i) : zero_vec<D>();
      Vec<D> fc = a.ext_fc ? ld<D>(a.ext_fc, i) : zero_vec<D>();
      Vec<D> g;
      if constexpr (D == 3)
        g = Vec<3>{a.g[0], a.g[1], a.g[2]};
      else
        g = Vec<2>{a.g[0], a.g[1]};
      fc = fc + g * (m / 1.0f);
      F = (F + fp) + fc;
      if (a.ext_t) T = T + lda<D>(a.ext_t, i);
    }
    if (a.lin == JDEM_LIN_VERLET) {
      const float s = (dt * 0.5f) / m;
      v = v + (F * s) * freef;
    } else if (a.lin == JDEM_LIN_EULER) {
      const float s = dt / m;
      v = v + (F * s) * freef;
      x = x + v * dt;
    }
    if (a.rot == JDEM_ROT_VERLET_SPIRAL)
      vvs_after<D>(q, w, T, I, dt, freef);
    else if (a.rot == JDEM_ROT_SPIRAL)
      spiral_after<D>(q, w, T, I, dt, freef);
  }
  if (a.do_before) {
    if (dom.kind == JDEM_DOMAIN_REFLECT) reflect_apply<D>(dom, x, v, q, F, m, p4b.w, fixed, dt);
    if (a.lin == JDEM_LIN_VERLET) {
      const float s = (dt * 0.5f) / m;
      v = v + (F * s) * freef;
      x = x + v * dt;
    }
    if (a.rot == JDEM_ROT_VERLET_SPIRAL) vvs_before<D>(q, w, T, I, dt, freef);
  }
  float4 p_out = f4(x, p4b.w), v_out = f4(v, m);
  float4 w_out = f4a(w);
  if constexpr (D == 2) {  // build position, inertia and bits ride along
    p_out.z = p4b.z;
    v_out.z = v4.z;
    w_out.y = w4.y;
    w_out.z = w4.z;
  }
  a.P2[b] = p_out;
  a.V2[b] = v_out;
  // the springs' place travels with the angular velocity (its free lane)
  w_out.w = __int_as_float(spring_off >= 0 ? spring_off : __float_as_int(w4.w));
  a.W2[b] = w_out;
  a.Q[b] = f4q(q);
  if (a.store_ft) {
    a.F[b] = f4(F, 0.0f);
    a.T[b] = f4a(T);
  }
  if (a.do_before && a.check_disp) {
    Vec<D> xb;
    if constexpr (D == 2)
      xb = Vec<2>{p4b.z, v4.z};
    else
      xb = xyz<D>(a.old4[b]);
    const Vec<D> d = x - xb;
    if (dot(d, d) > trigger) {
      *a.flag_wr = a.step_idx + 1;
      if (PM == 2) a.seg_flag[a.seg[b]] = 1;
    }
  }
}

template <int D, int LAW, int MW>
void launch_ws_mw(const WS& a, bool per, hipStream_t s) {
  dim3 g(nblocks(a.n)), b(kBlock);
  if (a.segp && a.block == 128)
    hipLaunchKernelGGL((k_ws_step<D, LAW, 2, MW, 128>), dim3(nblocks(a.n, 128)), dim3(128), 0, s, a);
  else if (a.segp)
    hipLaunchKernelGGL((k_ws_step<D, LAW, 2, MW>), g, b, 0, s, a);
  else if (per)
    hipLaunchKernelGGL((k_ws_step<D, LAW, 1, MW>), g, b, 0, s, a);
  else
    hipLaunchKernelGGL((k_ws_step<D, LAW, 0, MW>), g, b, 0, s, a);
}
template <int D, int LAW>
hipError_t launch_ws_law(const WS& a, bool per, hipStream_t s) {
  const int words = max(a.W64, (a.M + 63) >> 6);
  if (words <= 1)
    launch_ws_mw<D, LAW, 1>(a, per, s);
  else
    launch_ws_mw<D, LAW, 4>(a, per, s);
  return hipGetLastError();
}

template <int D>
hipError_t launch_ws_dim(const WS& a, int law, bool per, hipStream_t s) {
  switch (law) {
    case JDEM_LAW_SPRING:
      return launch_ws_law<D, JDEM_LAW_SPRING>(a, per, s);
    case JDEM_LAW_HERTZ:
      return launch_ws_law<D, JDEM_LAW_HERTZ>(a, per, s);
    case JDEM_LAW_CUNDALL_STRACK:
      return launch_ws_law<D, JDEM_LAW_CUNDALL_STRACK>(a, per, s);
    case JDEM_LAW_SPRING_FRICTION:
      return launch_ws_law<D, JDEM_LAW_SPRING_FRICTION>(a, per, s);
    case JDEM_LAW_HERTZ_FRICTION:
      return launch_ws_law<D, JDEM_LAW_HERTZ_FRICTION>(a, per, s);
    case kLawRouted:
      return launch_ws_law<D, kLawRouted>(a, per, s);
    case kLawRoutedHist:
      return launch_ws_law<D, kLawRoutedHist>(a, per, s);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t ws_launch_pack_springs(int64_t n, int M, int W64, const float4* rank, const uint64_t* live, float4* wave,
                                  float4* W, const int32_t* seg, const int32_t* seg_mode, hipStream_t s,
                                  const int32_t* gate, const int64_t* wpref, const int64_t* offsets, int n_seg) {
  if (n == 0 || M == 0) return hipSuccess;
  if (wpref)
    hipLaunchKernelGGL(k_pack_springs_seg, dim3(std::min<unsigned>(nblocks(n + 64 * (int64_t)n_seg), (unsigned)kCappedGrid)),
                       dim3(kBlock), 0, s, n, M, W64, rank, live, wave, W, wpref, offsets, n_seg, gate);
  else
    hipLaunchKernelGGL(k_pack_springs, dim3(gate ? std::min<unsigned>(nblocks(n), (unsigned)kCappedGrid) : nblocks(n)),
                       dim3(kBlock), 0, s, n, M, W64, rank, live, wave, W, seg, seg_mode, gate);
  return hipGetLastError();
}
hipError_t ws_launch_unpack_springs(int64_t n, int M, int W64, const float4* wave, const uint64_t* live,
                                    const float4* W, float4* rank, const int32_t* seg, const int32_t* seg_mode,
                                    hipStream_t s, const int32_t* gate, const int64_t* wpref,
                                    const int64_t* offsets, int n_seg) {
  if (n == 0 || M == 0) return hipSuccess;
  if (wpref)
    hipLaunchKernelGGL(k_unpack_springs_seg,
                       dim3(std::min<unsigned>(nblocks(n + 64 * (int64_t)n_seg), (unsigned)kCappedGrid)), dim3(kBlock),
                       0, s, n, M, W64, wave, live, W, rank, wpref, offsets, n_seg, gate);
  else
    hipLaunchKernelGGL(k_unpack_springs,
                       dim3(gate ? std::min<unsigned>(nblocks(n), (unsigned)kCappedGrid) : nblocks(n)), dim3(kBlock),
                       0, s, n, M, W64, wave, live, W, rank, seg, seg_mode, gate);
  return hipGetLastError();
}

hipError_t ws_launch_step(const WS& a, int dim, int law, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  const bool per = a.dom.kind == JDEM_DOMAIN_PERIODIC;
  return dim == 3 ? launch_ws_dim<3>(a, law, per, s) : launch_ws_dim<2>(a, law, per, s);
}

}  // namespace jdem

// ---------------------------------------------------------------------------
// launch wrappers of the conversion and rebuild kernels
// ---------------------------------------------------------------------------
namespace jdem {

__global__ void __launch_bounds__(kBlock) k_inverse(const int32_t* __restrict__ orig, int64_t n,
                                                     int32_t* __restrict__ wpos) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < n) wpos[orig[b]] = (int32_t)b;
}

// NaiveSimulator.create_neighbor_list in work space (descending original id;
// the work order is the original order when the secondary collider is naive)
template <int D, bool PER>
__global__ void __launch_bounds__(kBlock) k_nlist_work_naive(NLW a) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = b < a.n;  // every lane reaches end_row (wave-wide scan)
  int count = 0;
  if (active) {
    const Vec<D> xi = xyz<D>(a.P4[b]);
    const bool chk_bond = (!a.isb) && a.bond_w > 0;
    for (int64_t k = a.n - 1; k >= 0; --k) {
      if (k == b) continue;
      const Vec<D> d = disp_<PER>(xi, xyz<D>(a.P4[k]), a.dom);
      if (!(dot(d, d) <= a.cutoff_sq)) continue;
      if (chk_bond && bonded_o(a.bond_id, a.bond_w, a.orig[b], a.orig[k])) continue;
      if (count < a.M) a.nl[(int64_t)count * a.n + b] = (int32_t)k;
      ++count;
    }
  }
  end_row(a, b, active, count);
}

hipError_t ws_launch_in(const InOut& io, bool with_nl, hipStream_t s) {
  dim3 g(nblocks(io.n)), b(kBlock);
  hipLaunchKernelGGL(k_in_state, g, b, 0, s, io);
  if (with_nl) hipLaunchKernelGGL(k_in_nlist, g, b, 0, s, io);
  return hipGetLastError();
}
hipError_t ws_launch_out(const InOut& io, bool with_nl, hipStream_t s) {
  dim3 g(nblocks(io.n)), b(kBlock);
  hipLaunchKernelGGL(k_out_state, g, b, 0, s, io);
  if (with_nl) hipLaunchKernelGGL(k_out_nlist, g, b, 0, s, io);
  return hipGetLastError();
}
__global__ void k_gate_end(int32_t* __restrict__ gate) { *gate = 0; }
__global__ void k_gated_zero(const int32_t* __restrict__ gate, int32_t* __restrict__ word) {
  if (*gate) *word = 0;
}
hipError_t ws_launch_gate_end(int32_t* gate, hipStream_t s) {
  hipLaunchKernelGGL(k_gate_end, dim3(1), dim3(1), 0, s, gate);
  return hipGetLastError();
}
hipError_t ws_launch_gated_zero(const int32_t* gate, int32_t* word, hipStream_t s) {
  hipLaunchKernelGGL(k_gated_zero, dim3(1), dim3(1), 0, s, gate, word);
  return hipGetLastError();
}
hipError_t ws_launch_export(const InOut& io, hipStream_t s) {
  if (io.n == 0) return hipSuccess;
  dim3 g(nblocks(io.n)), b(kBlock);
  if (io.old_user) hipLaunchKernelGGL(k_out_old, g, b, 0, s, io);
  if (io.nl_user && io.M > 0) hipLaunchKernelGGL(k_out_nlist, g, b, 0, s, io);
  return hipGetLastError();
}
hipError_t ws_launch_copy_rows(const int32_t* src, int32_t* dst, int64_t n, int M, hipStream_t s) {
  if (n == 0 || M == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_rows, dim3(nblocks(n)), dim3(kBlock), 0, s, src, dst, n, M);
  return hipGetLastError();
}
hipError_t ws_launch_inverse(const int32_t* orig, int64_t n, int32_t* wpos, hipStream_t s) {
  hipLaunchKernelGGL(k_inverse, dim3(nblocks(n)), dim3(kBlock), 0, s, orig, n, wpos);
  return hipGetLastError();
}
hipError_t ws_launch_iota(int32_t* p, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_iota, dim3(nblocks(n)), dim3(kBlock), 0, s, p, n);
  return hipGetLastError();
}
hipError_t ws_launch_copy4(float4* d, const float4* src, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_f4, dim3(nblocks(n)), dim3(kBlock), 0, s, d, src, n);
  return hipGetLastError();
}
hipError_t ws_launch_hash(const float4* P4, const int32_t* orig, int64_t n, int dim, const DomP& dom, const GridP& g,
                          uint32_t* keys, int32_t* wpos, hipStream_t s) {
  dim3 gr(nblocks(n)), b(kBlock);
  const bool per = g.periodic != 0;
  if (dim == 3) {
    if (per)
      hipLaunchKernelGGL((k_hash_work<3, true>), gr, b, 0, s, P4, orig, n, dom, g, keys, wpos);
    else
      hipLaunchKernelGGL((k_hash_work<3, false>), gr, b, 0, s, P4, orig, n, dom, g, keys, wpos);
  } else {
    if (per)
      hipLaunchKernelGGL((k_hash_work<2, true>), gr, b, 0, s, P4, orig, n, dom, g, keys, wpos);
    else
      hipLaunchKernelGGL((k_hash_work<2, false>), gr, b, 0, s, P4, orig, n, dom, g, keys, wpos);
  }
  return hipGetLastError();
}
hipError_t ws_launch_seg_ids(const int64_t* offsets, int n_seg, int64_t n, int32_t* seg, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_seg_ids, dim3(nblocks(n)), dim3(kBlock), 0, s, offsets, n_seg, n, seg);
  return hipGetLastError();
}
hipError_t ws_launch_permute(int64_t n, const int32_t* new_orig, const int32_t* wpos, const float4* P,
                             const float4* V, const float4* W, const float4* Q, const float4* I, float4* P2,
                             float4* V2, float4* W2, float4* Q2, float4* I2, int32_t* orig2, int32_t* src,
                             int32_t* inv, hipStream_t s, float4* old2, int dim) {
  hipLaunchKernelGGL(k_permute, dim3(nblocks(n)), dim3(kBlock), 0, s, n, new_orig, wpos, P, V, W, Q, I, P2, V2, W2,
                     Q2, I2, orig2, src, inv, old2, dim);
  return hipGetLastError();
}
hipError_t ws_launch_seg_prefix(int n_seg, const int32_t* seg_mode, const int64_t* offsets, const SegP* segp,
                               int64_t* cpref, int64_t* ccref, int64_t* wpref, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_prefix, dim3(1), dim3(1024), 0, s, n_seg, seg_mode, offsets, segp, cpref, ccref, wpref);
  return hipGetLastError();
}
hipError_t ws_launch_hash_flagged(const RB& r, int dim, hipStream_t s) {
  if (r.n == 0) return hipSuccess;
  const dim3 g(r.gate ? std::min<unsigned>(nblocks(r.n), (unsigned)kCappedGrid) : nblocks(r.n));
  if (dim == 3)
    hipLaunchKernelGGL(k_hash_flagged<3>, g, dim3(kBlock), 0, s, r);
  else
    hipLaunchKernelGGL(k_hash_flagged<2>, g, dim3(kBlock), 0, s, r);
  return hipGetLastError();
}
hipError_t ws_launch_place(const RB& r, const uint32_t* sk, int32_t* starts, hipStream_t s) {
  if (r.fn == 0) return hipSuccess;
  // a device count: a capped grid that walks the flagged particles in turn
  const dim3 g(r.fn >= 0 ? nblocks(r.fn) : std::min<unsigned>(nblocks(r.n), (unsigned)kCappedGrid)), b(kBlock);
  hipLaunchKernelGGL(k_place_gather, g, b, 0, s, r);
  hipLaunchKernelGGL(k_place_scatter, g, b, 0, s, r);
  hipLaunchKernelGGL(k_cell_starts_flagged, g, b, 0, s, r, sk, starts);
  return hipGetLastError();
}
hipError_t ws_launch_seg_modes(int n_seg, int32_t* seg_flag, int32_t* n_build, int32_t* seg_mode, hipStream_t s,
                               const int32_t* gate, int32_t* events) {
  hipLaunchKernelGGL(k_seg_modes, dim3(nblocks(n_seg)), dim3(kBlock), 0, s, n_seg, seg_flag, n_build, seg_mode, gate,
                     events);
  return hipGetLastError();
}
hipError_t ws_launch_seg_init_flags(int n_seg, const int32_t* n_build, int32_t* seg_flag, hipStream_t s,
                                   int32_t* gate) {
  hipLaunchKernelGGL(k_seg_init_flags, dim3(nblocks(n_seg)), dim3(kBlock), 0, s, n_seg, n_build, seg_flag, gate);
  return hipGetLastError();
}
hipError_t ws_launch_cell_starts(const uint32_t* skeys, int64_t n, int64_t n_cells, int32_t* starts, hipStream_t s) {
  hipLaunchKernelGGL(k_cell_starts, dim3(nblocks(n + 1)), dim3(kBlock), 0, s, skeys, n, n_cells, starts);
  return hipGetLastError();
}
hipError_t ws_launch_nlist(const NLW& a, int dim, bool naive, hipStream_t s) {
  dim3 g(nblocks(a.n)), b(kBlock);
  const bool per = naive ? a.dom.kind == JDEM_DOMAIN_PERIODIC : a.grid.periodic != 0;
  if (naive) {
    if (dim == 3) {
      if (per)
        hipLaunchKernelGGL((k_nlist_work_naive<3, true>), g, b, 0, s, a);
      else
        hipLaunchKernelGGL((k_nlist_work_naive<3, false>), g, b, 0, s, a);
    } else {
      if (per)
        hipLaunchKernelGGL((k_nlist_work_naive<2, true>), g, b, 0, s, a);
      else
        hipLaunchKernelGGL((k_nlist_work_naive<2, false>), g, b, 0, s, a);
    }
  } else if (a.segp) {
    const dim3 gs(a.gate ? std::min<unsigned>(g.x, (unsigned)kCappedGrid) : g.x);
    if (dim == 3)
      hipLaunchKernelGGL((k_nlist_work<3, 2>), gs, b, 0, s, a);
    else
      hipLaunchKernelGGL((k_nlist_work<2, 2>), gs, b, 0, s, a);
  } else {
    if (dim == 3) {
      if (per)
        hipLaunchKernelGGL((k_nlist_work<3, 1>), g, b, 0, s, a);
      else
        hipLaunchKernelGGL((k_nlist_work<3, 0>), g, b, 0, s, a);
    } else {
      if (per)
        hipLaunchKernelGGL((k_nlist_work<2, 1>), g, b, 0, s, a);
      else
        hipLaunchKernelGGL((k_nlist_work<2, 0>), g, b, 0, s, a);
    }
  }
  return hipGetLastError();
}
// FreeDomain box of float4 positions (free.py:42-58)
__global__ void __launch_bounds__(kBlock) k_box_minmax4(const float4* __restrict__ P, int64_t n,
                                                         float* __restrict__ part) {
  __shared__ float smin[3][kBlock], smax[3][kBlock];
  const int tid = threadIdx.x;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + tid; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float4 p = P[i];
    const float x[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], x[a] - p.w);
      hi[a] = fmaxf(hi[a], x[a] + p.w);
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
        smin[a][tid] = fminf(smin[a][tid], smin[a][tid + off]);
        smax[a][tid] = fmaxf(smax[a][tid], smax[a][tid + off]);
      }
    __syncthreads();
  }
  if (tid == 0)
    for (int a = 0; a < 3; ++a) {
      part[6 * blockIdx.x + a] = smin[a][0];
      part[6 * blockIdx.x + 3 + a] = smax[a][0];
    }
}
hipError_t ws_launch_box(const float4* P, int64_t n, float* part, int nparts, hipStream_t s) {
  hipLaunchKernelGGL(k_box_minmax4, dim3(nparts), dim3(kBlock), 0, s, P, n, part);
  return hipGetLastError();
}

}  // namespace jdem
