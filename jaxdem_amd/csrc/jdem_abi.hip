// jdem_abi.hip -- the C ABI of libjdem_hip.so (include/jdem.h): contexts,
// scratch, and the host-side drivers that sequence the kernels of
// jdem_cells.hip and jdem_step.hip on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "jdem_internal.h"

using namespace jdem;

#ifndef JDEM_TARGET
#define JDEM_TARGET "gfx950"
#endif

std::vector<Buf*> jdem_ctx::bufs() {
  std::vector<Buf*> all = {&pos2,    &vel2,   &ang2,     &keys[0], &keys[1], &vals[0], &vals[1], &hist,
                           &cstart,  &cend,   &spos,     &nl_new,  &xi_new,  &live_new, &flags,  &red,
                           &nl_copy, &xi_copy, &live_copy, &wF,     &wT,      &wOld,     &wWpos, &wSrc,
                           &wInv,    &wKeys,  &fireV,    &fireDev, &segp,  &segid, &segflag, &wXiRank,
                           &route,   &ccnt,   &cl_start, &cl_members, &plab, &ppr, &ftot, &ttot, &crefl, &pvl, &pvaux};
  for (int k = 0; k < 2; ++k)
    for (Buf* b : {&wP[k], &wV[k], &wW[k], &wQ[k], &wI[k], &wOrig[k], &wNl[k], &wXi[k], &wLive[k]}) all.push_back(b);
  return all;
}

size_t jdem_ctx::scratch_bytes() {
  size_t s = 0;
  for (const Buf* b : bufs()) s += b->bytes;
  return s;
}

namespace jdem {
namespace {
constexpr size_t kPoolMaxBuffers = 24;
}

void pool_give(jdem_ctx* ctx, Buf& b) {
  if (!b.p) return;
  std::lock_guard<std::mutex> lock(ctx->pool_mu);
  ctx->pool.push_back(b);
  b = Buf{};
  if (ctx->pool.size() > kPoolMaxBuffers) {
    // drop the smallest pooled buffer; the stream may still use it
    size_t k = 0;
    for (size_t i = 1; i < ctx->pool.size(); ++i)
      if (ctx->pool[i].bytes < ctx->pool[k].bytes) k = i;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->pool[k].p);
    ctx->pool.erase(ctx->pool.begin() + (long)k);
  }
}

void pool_trim(jdem_ctx* ctx) {
  std::lock_guard<std::mutex> lock(ctx->pool_mu);
  if (ctx->pool.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (Buf& b : ctx->pool) (void)hipFree(b.p);
  ctx->pool.clear();
}

// The smallest pooled buffer of at least `bytes` and at most about four
// times that (a small request does not take a large buffer).
static bool pool_take(jdem_ctx* ctx, Buf& b, size_t bytes) {
  std::lock_guard<std::mutex> lock(ctx->pool_mu);
  long best = -1;
  for (size_t i = 0; i < ctx->pool.size(); ++i) {
    const size_t have = ctx->pool[i].bytes;
    if (have < bytes || have > 4 * bytes + (size_t(1) << 20)) continue;
    if (best < 0 || have < ctx->pool[(size_t)best].bytes) best = (long)i;
  }
  if (best < 0) return false;
  b = ctx->pool[(size_t)best];
  ctx->pool.erase(ctx->pool.begin() + best);
  return true;
}

bool ensure(jdem_ctx* ctx, Buf& b, size_t bytes) {
  if (bytes == 0) bytes = 4;
  if (b.bytes >= bytes) return true;
  if (!b.p && pool_take(ctx, b, bytes)) return true;
  if (b.p) {
    // the stream may still use the old buffer
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return false;
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes + bytes / 8;  // headroom for slowly growing callers
  if (hipMalloc(&b.p, want) != hipSuccess) {
    b.p = nullptr;
    return false;
  }
  b.bytes = want;
  return true;
}
DomP dom_of(const jdem_domain& d) {
  DomP p;
  p.kind = d.kind;
  for (int a = 0; a < 3; ++a) {
    p.box[a] = d.box_size[a];
    p.inv_box[a] = d.inv_box_size[a];
    p.anchor[a] = d.anchor[a];
  }
  p.restitution = d.restitution;
  return p;
}

GridP grid_of(const jdem_grid& g) {
  GridP p;
  for (int a = 0; a < 3; ++a) {
    p.dims[a] = g.dims[a] > 0 ? g.dims[a] : 1;
    p.strides[a] = g.strides[a];
    p.cell[a] = g.cell_size[a];
  }
  p.sr = g.search_range;
  p.periodic = g.periodic;
  p.n_cells = g.n_cells;
  return p;
}

DomP64 dom_of(const jdem_domain_f64& d) {
  DomP64 p;
  p.kind = d.kind;
  for (int a = 0; a < 3; ++a) {
    p.box[a] = d.box_size[a];
    p.inv_box[a] = d.inv_box_size[a];
    p.anchor[a] = d.anchor[a];
  }
  p.restitution = d.restitution;
  return p;
}

GridP64 grid_of(const jdem_grid_f64& g) {
  GridP64 p;
  for (int a = 0; a < 3; ++a) {
    p.dims[a] = g.dims[a] > 0 ? g.dims[a] : 1;
    p.strides[a] = g.strides[a];
    p.cell[a] = g.cell_size[a];
  }
  p.sr = g.search_range;
  p.periodic = g.periodic;
  p.n_cells = g.n_cells;
  return p;
}

}  // namespace jdem

namespace {

int fail(jdem_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail((ctx), e_ == hipErrorOutOfMemory ? JDEM_ENOMEM : JDEM_EHIP,                 \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                         \
  } while (0)

#define ENSURE(ctx, buf, bytes)                                                   \
  do {                                                                            \
    if (!ensure((ctx), (buf), (bytes)))                                           \
      return fail((ctx), JDEM_ENOMEM, "scratch allocation failed: " #buf);        \
  } while (0)

// Scalar type of an ABI law struct: jdem_law is float32, jdem_law_f64 float64.
template <class L>
struct LawScalar;
template <>
struct LawScalar<jdem_law> {
  using T = float;
};
template <>
struct LawScalar<jdem_law_f64> {
  using T = double;
};
template <class L>
using LawOf = LawPT<typename LawScalar<L>::T>;

template <class L>
LawOf<L> law_of_elem(const L& l) {
  LawOf<L> p;
  std::memset(&p, 0, sizeof(p));
  p.kind = l.kind;
  p.n_mat = l.n_mat;
  p.t0 = l.t0;
  p.t1 = l.t1;
  p.t2 = l.t2;
  p.m0 = l.m0;
  p.m1 = l.m1;
  p.m2 = l.m2;
  p.m3 = l.m3;
  p.p0 = l.p0;
  p.p1 = l.p1;
  return p;
}

bool is_history_kind(int k) { return k == JDEM_LAW_SPRING_FRICTION || k == JDEM_LAW_HERTZ_FRICTION; }

// Does the law carry a tangential spring per pair?
template <class L>
bool law_has_history(const L& l) {
  if (l.kind != JDEM_LAW_ROUTED) return is_history_kind(l.kind);
  for (int k = 0; k < l.n_terms; ++k)
    if (is_history_kind(l.route_terms[k].kind)) return true;
  return false;
}

// Entries of a routed law's species table: n_species^2, or 1 without species.
template <class L>
int route_entries(const L& l) {
  return l.n_species > 1 ? l.n_species * l.n_species : 1;
}

// Checks a routed law and copies its terms and species offsets to the
// context's device table when they differ from the last copy.
template <class L>
int prepare_route(jdem_ctx* ctx, const L& l, int64_t n) {
  if (l.n_species < 0 || l.n_species > kMatMask) return fail(ctx, JDEM_EINVAL, "routed law: n_species out of range");
  if (l.n_terms < 1 || !l.route_off || !l.route_terms) return fail(ctx, JDEM_EINVAL, "routed law without terms");
  if (l.n_species > 1 && n > 0 && !l.species) return fail(ctx, JDEM_EINVAL, "routed law: species array missing");
  const int E = route_entries(l);
  if (l.route_off[0] != 0 || l.route_off[E] != l.n_terms) return fail(ctx, JDEM_EINVAL, "routed law: bad route_off");
  for (int e = 0; e < E; ++e) {
    if (l.route_off[e + 1] <= l.route_off[e]) return fail(ctx, JDEM_EINVAL, "routed law: an entry has no term");
    int n_hist = 0;
    for (int k = l.route_off[e]; k < l.route_off[e + 1]; ++k) {
      const L& t = l.route_terms[k];
      if (t.kind < JDEM_LAW_NONE || t.kind > JDEM_LAW_HERTZ_FRICTION)
        return fail(ctx, JDEM_EINVAL, "routed law: a term is not an elementary law");
      if (t.kind != JDEM_LAW_NONE) {
        if (t.n_mat != l.n_mat) return fail(ctx, JDEM_EINVAL, "routed law: a term's n_mat differs from the law's");
        if (!t.t0 || (t.kind == JDEM_LAW_CUNDALL_STRACK && (!t.m0 || !t.m1 || !t.m2 || !t.m3)) ||
            (is_history_kind(t.kind) && (!t.t1 || !t.t2)))
          return fail(ctx, JDEM_EINVAL, "routed law: a term's tables are missing");
      }
      n_hist += is_history_kind(t.kind);
    }
    if (n_hist > 1) return fail(ctx, JDEM_EUNSUPPORTED, "routed law: an entry with more than one history law");
  }
  // device table: the terms (LawPT of the law's scalar type), then the
  // offsets (int32)
  const size_t tb = sizeof(LawOf<L>) * (size_t)l.n_terms;
  std::vector<char> host(tb + sizeof(int32_t) * (size_t)(E + 1));
  LawOf<L>* terms = reinterpret_cast<LawOf<L>*>(host.data());
  for (int k = 0; k < l.n_terms; ++k) terms[k] = law_of_elem(l.route_terms[k]);
  std::memcpy(host.data() + tb, l.route_off, sizeof(int32_t) * (size_t)(E + 1));
  if (host != ctx->route_host || !ctx->route.p) {
    ENSURE(ctx, ctx->route, host.size());
    // the stream's earlier kernels may still read the previous table
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpy(ctx->route.p, host.data(), host.size(), hipMemcpyHostToDevice));
    ctx->route_host.swap(host);
  }
  return JDEM_OK;
}

// Kernel-side law of a call: a routed law points at the context's device
// table (prepare_route ran in check_system) and takes the kernel kind with
// or without the history term.
template <class L>
LawOf<L> law_of(jdem_ctx* ctx, const L& l) {
  LawOf<L> p = law_of_elem(l);
  if (l.kind == JDEM_LAW_ROUTED) {
    p.kind = law_has_history(l) ? kLawRoutedHist : kLawRouted;
    p.n_species = l.n_species;
    p.route_terms = reinterpret_cast<const LawOf<L>*>(ctx->route.p);
    p.route_off = reinterpret_cast<const int32_t*>(static_cast<const char*>(ctx->route.p) +
                                                   sizeof(LawOf<L>) * (size_t)l.n_terms);
    p.species = l.n_species > 1 ? l.species : nullptr;
  }
  return p;
}

int check_system(jdem_ctx* ctx, const jdem_system* sys, bool batch = false, bool allow_resident = false) {
  if (!ctx) return JDEM_EINVAL;
  if (!sys) return fail(ctx, JDEM_EINVAL, "null system");
  const jdem_state& s = sys->state;
  if (s.dim != 2 && s.dim != 3) return fail(ctx, JDEM_EINVAL, "dim must be 2 or 3");
  if (s.n < 0 || s.n > (int64_t)INT32_MAX) return fail(ctx, JDEM_EINVAL, "particle count out of range");
  if (s.n > 0 && (!s.pos || !s.vel || !s.ang_vel || !s.q || !s.force || !s.torque || !s.rad || !s.mass ||
                  !s.inertia || !s.fixed))
    return fail(ctx, JDEM_EINVAL, "null state array");
  if (s.bond_width > 0 && !s.bond_id) return fail(ctx, JDEM_EINVAL, "bond_width > 0 but bond_id is null");
  if (sys->law.n_mat > 1 && !s.mat_id) return fail(ctx, JDEM_EINVAL, "n_mat > 1 but mat_id is null");
  if (s.clump_id || s.pos_p) {
    if (s.n > 0 && (!s.clump_id || !s.pos_p)) return fail(ctx, JDEM_EINVAL, "clumps need both pos_p and clump_id");
    if (batch) return fail(ctx, JDEM_EUNSUPPORTED, "batches with clumps");
    if (sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST && (sys->nlist.resident_in || sys->nlist.resident_out))
      return fail(ctx, JDEM_EUNSUPPORTED, "clumps with a resident NeighborList working set");
  }
  const int lk = sys->law.kind;
  if (sys->collider_kind != JDEM_COLLIDER_NONE &&
      !(lk >= JDEM_LAW_SPRING && lk <= JDEM_LAW_ROUTED))
    return fail(ctx, JDEM_EINVAL, "unknown law kind");
  if (sys->law.n_mat > kMatMask) return fail(ctx, JDEM_EINVAL, "more than 32767 materials");
  if (lk == JDEM_LAW_ROUTED) {
    if (sys->law.n_mat < 1) return fail(ctx, JDEM_EINVAL, "routed law: n_mat < 1");
    const int rc = prepare_route(ctx, sys->law, s.n);
    if (rc) return rc;
  } else if (sys->law.n_mat < 1 || !sys->law.t0) {
    if (sys->collider_kind != JDEM_COLLIDER_NONE && lk != JDEM_LAW_CUNDALL_STRACK)
      return fail(ctx, JDEM_EINVAL, "law table t0 missing");
  }
  if (sys->domain.kind < JDEM_DOMAIN_FREE || sys->domain.kind > JDEM_DOMAIN_REFLECT)
    return fail(ctx, JDEM_EINVAL, "unknown domain kind");
  const bool hist_law = law_has_history(sys->law);
  if (hist_law && sys->collider_kind != JDEM_COLLIDER_NONE) {
    if (sys->collider_kind != JDEM_COLLIDER_NEIGHBOR_LIST)
      return fail(ctx, JDEM_EINVAL, "history laws require the NeighborList collider");
    const jdem_resident* R = sys->nlist.resident_in;
    if (R ? !R->hist : (!sys->nlist.hist_xi || !sys->nlist.hist_live))
      return fail(ctx, JDEM_EINVAL, "history law without history buffers");
  }
  if (sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST) {
    const jdem_nlist& L = sys->nlist;
    if (L.max_neighbors < 0) return fail(ctx, JDEM_EINVAL, "max_neighbors < 0");
    if ((L.resident_in || L.resident_out) && !allow_resident)
      return fail(ctx, JDEM_EUNSUPPORTED, "resident working sets are used by jdem_system_step and jdem_rollout only");
    if (L.resident_in) {
      const jdem_resident* R = L.resident_in;
      if (R->ctx != ctx || R->n != s.n || R->M != L.max_neighbors || R->dim != s.dim)
        return fail(ctx, JDEM_EINVAL, "resident working set does not match the system (context, N, M or dim)");
    } else if (s.n > 0 && L.max_neighbors > 0 && (!L.nl || !L.old_pos)) {
      return fail(ctx, JDEM_EINVAL, "neighbour-list buffers missing");
    }
    if (L.secondary != JDEM_COLLIDER_CELL_LIST && L.secondary != JDEM_COLLIDER_NAIVE)
      return fail(ctx, JDEM_EINVAL, "secondary collider must be CellList or naive");
  }
  if (!batch && (sys->collider_kind == JDEM_COLLIDER_CELL_LIST ||
                 (sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST && sys->nlist.secondary == JDEM_COLLIDER_CELL_LIST))) {
    const jdem_grid& g = sys->collider_kind == JDEM_COLLIDER_CELL_LIST ? sys->grid : sys->nlist.grid;
    if (g.n_cells <= 0 || g.n_cells > (int64_t)INT32_MAX)
      return fail(ctx, JDEM_EUNSUPPORTED, "cell grid size out of the int32 hash range");
    if (g.search_range < 1) return fail(ctx, JDEM_EINVAL, "search_range must be >= 1");
  }
  return JDEM_OK;
}

StepArgs base_args(jdem_ctx* ctx, const jdem_system* sys) {
  StepArgs a;
  std::memset(&a, 0, sizeof(a));
  const jdem_state& s = sys->state;
  a.n = s.n;
  a.bond_w = s.bond_width;
  a.isb = sys->interact_same_bond_id;
  a.pos_in = s.pos;
  a.vel_in = s.vel;
  a.ang_in = s.ang_vel;
  a.pos_out = s.pos;
  a.vel_out = s.vel;
  a.ang_out = s.ang_vel;
  a.q = s.q;
  a.force = s.force;
  a.torque = s.torque;
  a.rad = s.rad;
  a.mass = s.mass;
  a.inertia = s.inertia;
  a.mat_id = s.mat_id;
  a.bond_id = s.bond_id;
  a.fixed = s.fixed;
  a.ext_f = sys->ext_force;
  a.ext_fc = sys->ext_force_com;
  a.ext_t = sys->ext_torque;
  for (int c = 0; c < 3; ++c) a.g[c] = sys->gravity[c];
  a.has_fm = (sys->gravity[0] != 0.0f || sys->gravity[1] != 0.0f || sys->gravity[2] != 0.0f || a.ext_f ||
              a.ext_fc || a.ext_t)
                 ? 1
                 : 0;
  a.law = law_of(ctx, sys->law);
  a.dom = dom_of(sys->domain);
  a.dt = sys->dt;
  a.lin = sys->lin_kind;
  a.rot = sys->rot_kind;
  a.M = sys->nlist.max_neighbors;
  a.W32 = (a.M + 31) / 32;
  a.nl = sys->nlist.nl;
  a.hist_xi = sys->nlist.hist_xi;
  a.hist_live = sys->nlist.hist_live;
  return a;
}

struct Cells {
  int32_t* perm;
  int32_t* skeys;
  int32_t* cstart;
  int32_t* cend;
  float4* spos;
};

int nbits_for(int64_t n_cells) { return nbits_for_cells(n_cells); }

// Hash + sort + cell ranges + sorted (x, r) of `pos` (cell_list.py:37-87);
// in float64 the kernels read the partners' own arrays and no (x, r) copy
// is made.
template <class Real>
int build_cells(jdem_ctx* ctx, const Real* pos, const Real* rad, int64_t n, int dim, const DomPT<Real>& dom,
                const GridPT<Real>& g, Cells* out) {
  const size_t nn = (size_t)(n > 0 ? n : 1);
  ENSURE(ctx, ctx->keys[0], 4 * nn);
  ENSURE(ctx, ctx->keys[1], 4 * nn);
  ENSURE(ctx, ctx->vals[0], 4 * nn);
  ENSURE(ctx, ctx->vals[1], 4 * nn);
  ENSURE(ctx, ctx->cstart, 4 * ((size_t)g.n_cells + 1));
  ENSURE(ctx, ctx->cend, 4 * (size_t)g.n_cells);
  ENSURE(ctx, ctx->spos, 16 * nn);
  ENSURE(ctx, ctx->hist, 4 * 256 * (size_t)((n + 4095) / 4096 + 1));
  uint32_t* hk = (uint32_t*)ctx->keys[1].p;
  HIPCHK(ctx, launch_hash(pos, n, dim, dom, g, hk, ctx->stream));
  uint32_t* sk = nullptr;
  int32_t* perm = nullptr;
  int32_t* cs = (int32_t*)ctx->cstart.p;
  int32_t* ce = (int32_t*)ctx->cend.p;
  if (use_cell_sort(n, g.n_cells)) {
    // starts of every cell (an empty cell starts where the next one does)
    HIPCHK(ctx, cell_sort_pairs(ctx, hk, n, g.n_cells, &sk, &perm, cs, ctx->stream));
    HIPCHK(ctx, launch_starts_to_ends(cs, g.n_cells, ce, ctx->stream));
  } else {
    HIPCHK(ctx, radix_sort_pairs(ctx, hk, n, nbits_for(g.n_cells), &sk, &perm, ctx->stream));
    HIPCHK(ctx, launch_fill_i32(cs, g.n_cells, 0, ctx->stream));
    HIPCHK(ctx, launch_fill_i32(ce, g.n_cells, 0, ctx->stream));
    HIPCHK(ctx, launch_cell_ranges(sk, n, cs, ce, ctx->stream));
  }
  if constexpr (std::is_same<Real, float>::value)
    HIPCHK(ctx, launch_gather_spos(perm, pos, rad, n, dim, (float4*)ctx->spos.p, ctx->stream));
  out->perm = perm;
  out->skeys = (int32_t*)sk;
  out->cstart = cs;
  out->cend = ce;
  out->spos = (float4*)ctx->spos.p;
  return JDEM_OK;
}

// Device words in ctx->flags: [0], [1] rebuild-pending ping-pong, [2] list
// overflow, [3] displacement flag of the standalone check, [4] rebuild gate
// and [5] rebuild events of a batch, [6] default-inertia flag of a call.
int ensure_flags(jdem_ctx* ctx) {
  ENSURE(ctx, ctx->flags, 64);
  if (!ctx->host_flags) {
    if (hipHostMalloc((void**)&ctx->host_flags, 64, hipHostMallocDefault) != hipSuccess) {
      ctx->host_flags = nullptr;
      return fail(ctx, JDEM_ENOMEM, "pinned flag allocation failed");
    }
  }
  return JDEM_OK;
}

int read_flags(jdem_ctx* ctx, int count) {
  HIPCHK(ctx, hipMemcpyAsync(ctx->host_flags, ctx->flags.p, sizeof(int32_t) * count, hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return JDEM_OK;
}

// Build a neighbour list of `pos` into nl_out (N, M) with the cell list or
// the naive search; overflow accumulates into flags[2].
int build_nlist(jdem_ctx* ctx, const jdem_system* sys, int kind, const jdem_grid& grid, const float* pos,
                float cutoff, int M, int32_t* nl_out) {
  const jdem_state& s = sys->state;
  NLBuild b;
  std::memset(&b, 0, sizeof(b));
  b.n = s.n;
  b.dim = s.dim;
  b.M = M;
  b.cutoff_sq = cutoff * cutoff;
  b.isb = sys->interact_same_bond_id;
  b.bond_w = s.bond_width;
  b.bond_id = s.bond_id;
  b.pos = pos;
  b.dom = dom_of(sys->domain);
  b.nl = nl_out;
  b.overflow = (int32_t*)ctx->flags.p + 2;
  b.clump = s.clump_id;
  if (kind == JDEM_COLLIDER_CELL_LIST) {
    Cells c;
    b.grid = grid_of(grid);
    int rc = build_cells(ctx, pos, s.rad, s.n, s.dim, b.dom, b.grid, &c);
    if (rc) return rc;
    b.perm = c.perm;
    b.skeys = c.skeys;
    b.cstart = c.cstart;
    b.cend = c.cend;
    b.spos = c.spos;
    HIPCHK(ctx, launch_nlist_cells(b, ctx->stream));
  } else {
    HIPCHK(ctx, launch_nlist_naive(b, ctx->stream));
  }
  return JDEM_OK;
}

// NeighborList rebuild at positions `pos` (neighbor_list.py:58-131, 545-589):
// new list from the secondary collider, history remapped by partner id,
// old_pos <- pos, n_build_times += 1.  flags[2] receives the overflow.
int nl_rebuild(jdem_ctx* ctx, jdem_system* sys, const float* pos) {
  jdem_nlist& L = sys->nlist;
  const jdem_state& s = sys->state;
  const int64_t n = s.n;
  const int M = L.max_neighbors;
  const int W32 = (M + 31) / 32;
  const float lc = L.cutoff + L.skin;
  HIPCHK(ctx, launch_fill_i32((int32_t*)ctx->flags.p + 2, 1, 0, ctx->stream));
  if (M > 0 && n > 0) {
    ENSURE(ctx, ctx->nl_new, 4 * (size_t)n * M);
    int32_t* nl_new = (int32_t*)ctx->nl_new.p;
    int rc = build_nlist(ctx, sys, L.secondary, L.grid, pos, lc, M, nl_new);
    if (rc) return rc;
    if (L.hist_xi && L.hist_live) {
      if (L.n_build_times == 0) {
        HIPCHK(ctx, launch_fill_i32((int32_t*)L.hist_live, n * W32, 0, ctx->stream));
      } else {
        ENSURE(ctx, ctx->xi_new, 4 * (size_t)n * M * s.dim);
        ENSURE(ctx, ctx->live_new, 4 * (size_t)n * W32);
        HIPCHK(ctx, launch_remap_hist(n, M, W32, s.dim, L.nl, L.hist_live, L.hist_xi, nl_new,
                                      (uint32_t*)ctx->live_new.p, (float*)ctx->xi_new.p, ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(L.hist_xi, ctx->xi_new.p, 4 * (size_t)n * M * s.dim, hipMemcpyDeviceToDevice,
                                   ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(L.hist_live, ctx->live_new.p, 4 * (size_t)n * W32, hipMemcpyDeviceToDevice,
                                   ctx->stream));
      }
    }
    HIPCHK(ctx, hipMemcpyAsync(L.nl, nl_new, 4 * (size_t)n * M, hipMemcpyDeviceToDevice, ctx->stream));
  }
  if (n > 0) HIPCHK(ctx, launch_copy_f32(L.old_pos, pos, n * s.dim, ctx->stream));
  L.n_build_times += 1;
  return JDEM_OK;
}

template <class Real>
__global__ void k_disp_check(const Real* __restrict__ pos, const Real* __restrict__ old, int64_t n, int dim,
                             Real trigger, int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Real d2;
  const Real dx = pos[dim * i] - old[dim * i];
  const Real dy = pos[dim * i + 1] - old[dim * i + 1];
  if (dim == 3) {
    const Real dz = pos[dim * i + 2] - old[dim * i + 2];
    d2 = (dx * dx + dy * dy) + dz * dz;
  } else {
    d2 = dx * dx + dy * dy;
  }
  if (d2 > trigger) *flag = 1;
}

float nl_trigger(const jdem_nlist& L) {
  float sk = L.skin * L.skin;
  return sk / 4.0f;
}

// _check_and_rebuild on its own (neighbor_list.py:58-131).
int nl_check_and_rebuild(jdem_ctx* ctx, jdem_system* sys, bool* rebuilt) {
  jdem_nlist& L = sys->nlist;
  const jdem_state& s = sys->state;
  *rebuilt = false;
  int rc = ensure_flags(ctx);
  if (rc) return rc;
  bool need = L.n_build_times == 0;
  if (!need && s.n > 0) {
    HIPCHK(ctx, launch_fill_i32((int32_t*)ctx->flags.p + 3, 1, 0, ctx->stream));
    hipLaunchKernelGGL(k_disp_check<float>, dim3(nblocks(s.n)), dim3(kBlock), 0, ctx->stream, s.pos, L.old_pos, s.n, s.dim,
                       nl_trigger(L), (int32_t*)ctx->flags.p + 3);
    HIPCHK(ctx, hipGetLastError());
    rc = read_flags(ctx, 4);
    if (rc) return rc;
    need = ctx->host_flags[3] != 0;
  }
  if (need) {
    rc = nl_rebuild(ctx, sys, s.pos);
    if (rc) return rc;
    rc = read_flags(ctx, 3);
    if (rc) return rc;
    L.overflow = ctx->host_flags[2] != 0;
    *rebuilt = true;
  }
  return JDEM_OK;
}

// ---- HIP-event timing ----------------------------------------------------
hipEvent_t take_event(jdem_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  // Timing-only events: without the system-scope fence, recording one does
  // not write back and invalidate the caches, which would slow the kernel
  // that follows and distort what is measured.
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
  return e;
}

// Returns the span index or -1 when timing is off.
int span_begin(jdem_ctx* ctx, int kind, int64_t step) {
  if (!ctx->timing) return -1;
  jdem_ctx::TimedSpan s;
  s.start = take_event(ctx);
  s.stop = take_event(ctx);
  s.kind = kind;
  s.step = step;
  if (!s.start || !s.stop) return -1;
  (void)hipEventRecord(s.start, ctx->stream);
  ctx->spans.push_back(s);
  return (int)ctx->spans.size() - 1;
}

void span_end(jdem_ctx* ctx, int idx) {
  if (idx < 0) return;
  (void)hipEventRecord(ctx->spans[idx].stop, ctx->stream);
}

// ---- resident NeighborList engine driver ---------------------------------
// System.step with the NeighborList collider on the cell-sorted work set of
// jdem_resident.hip; see that file for the layout.
// A segmented batch of independent systems (jdem_batch_step): the system of
// every particle and the systems' parameters, on the device.
struct SegRun {
  int n_seg;
  const int32_t* seg;
  const SegP* segp;
  const int64_t* offsets;  // particle ranges (device, n_seg + 1)
  int64_t* cpref;          // flagged-particle prefix (device, n_seg + 1)
  int64_t* ccref;          // flagged-cell prefix (device, n_seg + 1)
  int64_t* wpref;          // flagged-wavefront prefix (device, n_seg + 1)
  int64_t total_cells;
  int32_t* n_build;   // builds of each system (device)
  int32_t* seg_flag;  // system asks for a rebuild (device)
  int32_t* seg_mode;  // what the current rebuild does to it (device)
  bool any_unbuilt;   // some system has no list yet
  int launch_mode;    // jdem_batch.launch_mode
  int tile;           // jdem_batch.tile_width
  int spec_min_gap;   // jdem_batch.spec_min_gap (0 -> 2)
};

// n steps of a NeighborList system on the resident working set.  With
// frames, the steps are split into fr->n_frames segments of fr->steps[f]
// steps, each ended as System.step ends, and the state after segment f is
// written to frame f (caller order) before the next segment starts.
int resident_nl_step(jdem_ctx* ctx, jdem_system* sys, int64_t n, jdem_step_stats* S, const SegRun* sg = nullptr,
                     const jdem_frames* fr = nullptr) {
  jdem_state& s = sys->state;
  jdem_nlist& L = sys->nlist;
  const int64_t N = s.n;
  const int D = s.dim;
  const int M = L.max_neighbors;
  const int W64 = std::max(1, (M + 63) / 64);
  const jdem_resident* R = L.resident_in;  // working set of an earlier call
  const bool hist = R ? R->hist : (L.hist_xi && L.hist_live);
  const bool naive2 = L.secondary == JDEM_COLLIDER_NAIVE;
  if (M > 256) return fail(ctx, JDEM_EUNSUPPORTED, "max_neighbors > 256 on the resident path");
  if (!naive2 && !sg && sys->domain.kind == JDEM_DOMAIN_FREE)
    return fail(ctx, JDEM_EUNSUPPORTED, "NeighborList over a CellList with a FreeDomain (per-step grid)");
  const int64_t n_cells = sg ? sg->total_cells : L.grid.n_cells;
  const size_t n4 = 16 * (size_t)N, ni = 4 * (size_t)N;
  const size_t Npad = ((size_t)N + 63) & ~(size_t)63;  // whole wavefronts of spring regions
  for (int k = 0; k < 2; ++k) {
    ENSURE(ctx, ctx->wP[k], n4);
    ENSURE(ctx, ctx->wV[k], n4);
    ENSURE(ctx, ctx->wW[k], n4);
    ENSURE(ctx, ctx->wQ[k], n4);
    ENSURE(ctx, ctx->wI[k], n4);
    ENSURE(ctx, ctx->wOrig[k], ni);
    ENSURE(ctx, ctx->wNl[k], 4 * (size_t)N * (size_t)std::max(M, 1));
    ENSURE(ctx, ctx->wXi[k], hist ? 16 * Npad * (size_t)std::max(M, 1) : 16);
    ENSURE(ctx, ctx->wLive[k], 8 * (size_t)N * W64);
  }
  ENSURE(ctx, ctx->wXiRank, hist ? 16 * (size_t)N * (size_t)std::max(M, 1) : 16);
  ENSURE(ctx, ctx->wF, n4);
  ENSURE(ctx, ctx->wT, n4);
  ENSURE(ctx, ctx->wOld, n4);
  ENSURE(ctx, ctx->wWpos, ni);
  ENSURE(ctx, ctx->wSrc, ni);
  ENSURE(ctx, ctx->wInv, ni);
  ENSURE(ctx, ctx->wKeys, ni);
  ENSURE(ctx, ctx->keys[0], ni);
  ENSURE(ctx, ctx->keys[1], ni);
  ENSURE(ctx, ctx->vals[0], ni);
  ENSURE(ctx, ctx->vals[1], ni);
  ENSURE(ctx, ctx->hist, 4 * 256 * (size_t)((N + 4095) / 4096 + 1));
  if (!naive2) ENSURE(ctx, ctx->cstart, 4 * ((size_t)n_cells + 1));
  auto F4 = [](jdem::Buf& b) { return (float4*)b.p; };
  auto I32 = [](jdem::Buf& b) { return (int32_t*)b.p; };
  auto U64 = [](jdem::Buf& b) { return (uint64_t*)b.p; };
  hipStream_t st = ctx->stream;
  const DomP dom = dom_of(sys->domain);
  const GridP g = grid_of(L.grid);
  int cur = 0, set = 0, nlc = 0;
  int xc = 0;  // springs: read wXi[xc], write wXi[1 - xc]

  // --- entry: order the work set by (cell, original id), convert inputs ---
  // (from a resident working set: its work order, list and springs as they are)
  if (R) {
    HIPCHK(ctx, hipMemcpyAsync(ctx->wOrig[0].p, R->orig.p, ni, hipMemcpyDeviceToDevice, st));
  } else if (!naive2 && !sg) {
    HIPCHK(ctx, launch_hash(s.pos, N, D, dom, g, (uint32_t*)ctx->wKeys.p, st));
    uint32_t* sk = nullptr;
    int32_t* perm = nullptr;
    if (use_cell_sort(N, g.n_cells))
      HIPCHK(ctx, cell_sort_pairs(ctx, (uint32_t*)ctx->wKeys.p, N, g.n_cells, &sk, &perm, I32(ctx->cstart), st));
    else
      HIPCHK(ctx, radix_sort_pairs(ctx, (uint32_t*)ctx->wKeys.p, N, nbits_for(g.n_cells), &sk, &perm, st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->wOrig[0].p, perm, ni, hipMemcpyDeviceToDevice, st));
  } else {
    HIPCHK(ctx, ws_launch_iota(I32(ctx->wOrig[0]), N, st));
  }
  HIPCHK(ctx, ws_launch_inverse(I32(ctx->wOrig[0]), N, I32(ctx->wWpos), st));
  InOut io;
  std::memset(&io, 0, sizeof(io));
  io.n = N;
  io.dim = D;
  io.M = M;
  io.W64 = W64;
  io.pos = s.pos;
  io.vel = s.vel;
  io.ang = s.ang_vel;
  io.q = s.q;
  io.force = s.force;
  io.torque = s.torque;
  io.rad = s.rad;
  io.mass = s.mass;
  io.inertia = s.inertia;
  io.mat_id = sys->law.n_mat > 1 ? s.mat_id : nullptr;
  io.species = sys->law.kind == JDEM_LAW_ROUTED && sys->law.n_species > 1 ? sys->law.species : nullptr;
  io.fixed = s.fixed;
  io.nl_user = L.nl;
  io.old_user = L.old_pos;
  io.xi_user = hist ? L.hist_xi : nullptr;
  io.live_user = hist ? L.hist_live : nullptr;
  io.orig = I32(ctx->wOrig[0]);
  io.wpos = I32(ctx->wWpos);
  io.P4 = F4(ctx->wP[0]);
  io.V4 = F4(ctx->wV[0]);
  io.W4 = F4(ctx->wW[0]);
  io.Q4 = F4(ctx->wQ[0]);
  io.I4 = F4(ctx->wI[0]);
  io.F4 = F4(ctx->wF);
  io.T4 = F4(ctx->wT);
  io.old4 = F4(ctx->wOld);
  io.nl = I32(ctx->wNl[0]);
  io.xi = F4(ctx->wXiRank);
  io.live = U64(ctx->wLive[0]);
  // 3D spheres with one material, none fixed and the default inertia skip
  // the inertia record in the step kernel (flags[6], decided at entry)
  int32_t* iso = nullptr;
  if (D == 3 && sys->law.n_mat == 1 && !io.species) {
    iso = (int32_t*)ctx->flags.p + 6;
    HIPCHK(ctx, launch_fill_i32(iso, 1, 1, st));
  }
  io.iso_flag = iso;
  if (R) {
    io.old_user = nullptr;
    io.old_keep = 1;
    io.nl_user = nullptr;
    io.xi_user = nullptr;
    io.live_user = nullptr;
    HIPCHK(ctx, hipMemcpyAsync(ctx->wOld.p, R->old4.p, n4, hipMemcpyDeviceToDevice, st));
    HIPCHK(ctx, ws_launch_in(io, false, st));
    if (M > 0) {
      HIPCHK(ctx, ws_launch_copy_rows((const int32_t*)R->nl.p, I32(ctx->wNl[0]), N, M, st));
      HIPCHK(ctx, hipMemcpyAsync(ctx->wLive[0].p, R->live.p, 8 * (size_t)N * W64, hipMemcpyDeviceToDevice, st));
    }
  } else {
    HIPCHK(ctx, ws_launch_in(io, M > 0, st));
  }
  // the step kernel keeps springs per wavefront
  if (hist && M > 0)
    HIPCHK(ctx, ws_launch_pack_springs(N, M, W64, R ? (const float4*)R->xi.p : F4(ctx->wXiRank),
                                       U64(ctx->wLive[0]), F4(ctx->wXi[0]), F4(ctx->wW[0]), nullptr, nullptr, st));

  // --- rebuild at the current positions (neighbor_list.py:545-589) ---
  // Batches: only the systems that asked for a rebuild are re-sorted and
  // rebuilt, each within its own range of work slots; the work order, rows,
  // springs and build positions of the other systems stay where they are.
  // Batches decide their rebuilds on the device.  After every step kernel a
  // chain of kernels runs, each returning at once unless the gate word is
  // set (by the step kernel when a particle of some system moved more than
  // half its skin, or at the call start by a system without a list); the
  // chain rebuilds exactly the flagged systems, each inside its own range
  // of work slots, sized from device counts (flagged particles, compact
  // cells), and clears the gate.  The work order, rows, springs and build
  // positions of the other systems stay where they are.  No host round trip.
  int32_t* gate = (int32_t*)ctx->flags.p + 4;
  int32_t* events = (int32_t*)ctx->flags.p + 5;
  auto rebuild_batch = [&]() -> int {
    const int sp = span_begin(ctx, 2, 0);
    HIPCHK(ctx, ws_launch_gated_zero(gate, (int32_t*)ctx->flags.p + 2, st));  // overflow of this build
    HIPCHK(ctx, ws_launch_seg_modes(sg->n_seg, sg->seg_flag, sg->n_build, sg->seg_mode, st, gate, events));
    HIPCHK(ctx, ws_launch_seg_prefix(sg->n_seg, sg->seg_mode, sg->offsets, sg->segp, sg->cpref, sg->ccref, sg->wpref,
                                     st));
    // springs of the wavefronts that hold a rebuilt system: per wave -> by rank
    if (hist && M > 0)
      HIPCHK(ctx, ws_launch_unpack_springs(N, M, W64, F4(ctx->wXi[xc]), U64(ctx->wLive[nlc]), F4(ctx->wW[cur]),
                                           F4(ctx->wXiRank), sg->seg, sg->seg_mode, st, gate, sg->wpref, sg->offsets,
                                           sg->n_seg));
    RB r;
    std::memset(&r, 0, sizeof(r));
    r.n = N;
    r.fn = -1;  // the flagged-particle count is cpref[n_seg] on the device
    r.n_seg = sg->n_seg;
    r.dim = D;
    r.gate = gate;
    r.M = M;
    r.W64 = W64;
    r.seg = sg->seg;
    r.segp = sg->segp;
    r.seg_mode = sg->seg_mode;
    r.offsets = sg->offsets;
    r.cpref = sg->cpref;
    r.ccref = sg->ccref;
    r.P = F4(ctx->wP[cur]);
    r.V = F4(ctx->wV[cur]);
    r.W = F4(ctx->wW[cur]);
    r.Q = F4(ctx->wQ[set]);
    r.I = F4(ctx->wI[set]);
    r.old4 = F4(ctx->wOld);
    r.orig = I32(ctx->wOrig[set]);
    r.nl = I32(ctx->wNl[nlc]);
    r.live = hist ? U64(ctx->wLive[nlc]) : nullptr;
    r.xi = F4(ctx->wXiRank);
    r.keys_c = (uint32_t*)ctx->wKeys.p;
    r.slot_c = I32(ctx->cend);
    // staging: buffers that hold nothing live during a batch rebuild
    r.tP = F4(ctx->wP[1 - cur]);
    r.tV = F4(ctx->wV[1 - cur]);
    r.tW = F4(ctx->wW[1 - cur]);
    r.tQ = F4(ctx->wQ[1 - set]);
    r.tI = F4(ctx->wI[1 - set]);
    r.torig = I32(ctx->wOrig[1 - set]);
    r.cs = I32(ctx->wWpos);
    r.tnl = I32(ctx->wNl[1 - nlc]);
    r.tlive = U64(ctx->wLive[1 - nlc]);
    r.txi = F4(ctx->wXi[1 - xc]);
    r.inv = I32(ctx->wInv);
    r.cidx = I32(ctx->wSrc);
    HIPCHK(ctx, ws_launch_hash_flagged(r, D, st));
    uint32_t* sk = nullptr;
    int32_t* perm = nullptr;
    // counting sort by compact cell, sized for every particle and every cell
    // of the batch; the kernels read the flagged particle and cell counts on
    // the device
    HIPCHK(ctx, cell_sort_dev(ctx, r.keys_c, N, std::max<int64_t>(sg->total_cells, 1), sg->cpref + sg->n_seg,
                              sg->ccref + sg->n_seg, &sk, &perm, st, gate));
    r.perm = perm;
    HIPCHK(ctx, ws_launch_place(r, sk, I32(ctx->cstart), st));
    if (M > 0) {
      NLW b;
      std::memset(&b, 0, sizeof(b));
      b.n = N;
      b.M = M;
      b.isb = sys->interact_same_bond_id;
      b.bond_w = s.bond_width;
      b.bond_id = s.bond_id;
      b.orig = I32(ctx->wOrig[set]);
      b.P4 = F4(ctx->wP[cur]);
      b.skeys = sk;  // compact sorted keys: slot b's key is sk[cidx[b]]
      b.ccref = sg->ccref;
      b.cidx = I32(ctx->wSrc);
      b.starts = I32(ctx->cstart);
      b.grid = g;
      b.dom = dom;
      b.seg = sg->seg;
      b.segp = sg->segp;
      b.seg_mode = sg->seg_mode;
      b.nl = I32(ctx->wNl[nlc]);
      b.overflow = (int32_t*)ctx->flags.p + 2;
      b.gate = gate;
      b.wpref = sg->wpref;
      b.offsets = sg->offsets;
      b.n_seg = sg->n_seg;
      if (hist) {
        b.W64 = W64;
        b.inv = I32(ctx->wInv);
        b.ostride = N;
        b.old_nl = r.tnl;
        b.old_xi = r.txi;
        b.old_live = r.tlive;
        b.xi = F4(ctx->wXiRank);
        b.live = U64(ctx->wLive[nlc]);
      }
      HIPCHK(ctx, ws_launch_nlist(b, D, false, st));
    }
    if (hist && M > 0)
      HIPCHK(ctx, ws_launch_pack_springs(N, M, W64, F4(ctx->wXiRank), U64(ctx->wLive[nlc]), F4(ctx->wXi[xc]),
                                         F4(ctx->wW[cur]), sg->seg, sg->seg_mode, st, gate, sg->wpref, sg->offsets,
                                         sg->n_seg));
    HIPCHK(ctx, ws_launch_gate_end(gate, st));
    span_end(ctx, sp);
    return JDEM_OK;
  };

  auto rebuild = [&]() -> int {
    if (sg) return rebuild_batch();
    const int sp = span_begin(ctx, 2, 0);
    HIPCHK(ctx, launch_fill_i32((int32_t*)ctx->flags.p + 2, 1, 0, st));
    const int cur_before = cur;  // the W buffer that holds the springs' offsets
    int32_t* src = I32(ctx->wSrc);
    int32_t* inv = I32(ctx->wInv);
    uint32_t* sk = nullptr;
    if (!naive2) {
      HIPCHK(ctx, ws_launch_hash(F4(ctx->wP[cur]), I32(ctx->wOrig[set]), N, D, dom, g, (uint32_t*)ctx->wKeys.p,
                                 I32(ctx->wWpos), st));
      int32_t* new_orig = nullptr;
      // the counting sort also writes the cell starts
      const bool counted = use_cell_sort(N, n_cells);
      if (counted)
        HIPCHK(ctx, cell_sort_pairs(ctx, (uint32_t*)ctx->wKeys.p, N, n_cells, &sk, &new_orig, I32(ctx->cstart), st));
      else
        HIPCHK(ctx, radix_sort_pairs(ctx, (uint32_t*)ctx->wKeys.p, N, nbits_for(n_cells), &sk, &new_orig, st));
      HIPCHK(ctx, ws_launch_permute(N, new_orig, I32(ctx->wWpos), F4(ctx->wP[cur]), F4(ctx->wV[cur]),
                                    F4(ctx->wW[cur]), F4(ctx->wQ[set]), F4(ctx->wI[set]), F4(ctx->wP[1 - cur]),
                                    F4(ctx->wV[1 - cur]), F4(ctx->wW[1 - cur]), F4(ctx->wQ[1 - set]),
                                    F4(ctx->wI[1 - set]), I32(ctx->wOrig[1 - set]), src, inv, st, F4(ctx->wOld), D));
      cur = 1 - cur;
      set = 1 - set;
      if (!counted) HIPCHK(ctx, ws_launch_cell_starts(sk, N, n_cells, I32(ctx->cstart), st));
    } else {
      HIPCHK(ctx, ws_launch_iota(src, N, st));
      HIPCHK(ctx, ws_launch_iota(inv, N, st));
    }
    if (M > 0) {
      NLW b;
      std::memset(&b, 0, sizeof(b));
      b.n = N;
      b.M = M;
      const float lc = L.cutoff + L.skin;
      b.cutoff_sq = lc * lc;
      b.isb = sys->interact_same_bond_id;
      b.bond_w = s.bond_width;
      b.bond_id = s.bond_id;
      b.orig = I32(ctx->wOrig[set]);
      b.P4 = F4(ctx->wP[cur]);
      b.skeys = sk;
      b.starts = I32(ctx->cstart);
      b.grid = g;
      b.dom = dom;
      b.nl = I32(ctx->wNl[1 - nlc]);
      b.overflow = (int32_t*)ctx->flags.p + 2;
      if (hist) {
        b.W64 = W64;
        b.remap = L.n_build_times > 0;
        b.src = src;
        b.inv = inv;
        b.old_nl = I32(ctx->wNl[nlc]);
        b.old_xi = F4(ctx->wXi[xc]);  // per wave, read through the old offsets
        b.old_W = F4(ctx->wW[cur_before]);
        b.old_live = U64(ctx->wLive[nlc]);
        b.xi = F4(ctx->wXi[1 - xc]);  // per wave, offsets into the new W
        b.new_W = F4(ctx->wW[cur]);
        b.live = U64(ctx->wLive[1 - nlc]);
      }
      HIPCHK(ctx, ws_launch_nlist(b, D, naive2, st));
      nlc = 1 - nlc;
      if (hist) xc = 1 - xc;
    }
    if (naive2) HIPCHK(ctx, ws_launch_copy4(F4(ctx->wOld), F4(ctx->wP[cur]), N, st));  // else the permute wrote them
    L.n_build_times += 1;
    span_end(ctx, sp);
    return JDEM_OK;
  };

  int32_t* flags = (int32_t*)ctx->flags.p;
  float fbox[3] = {0, 0, 0}, fanc[3] = {0, 0, 0};
  bool free_done = false;
  std::vector<int> cur_at, xc_at;
  int64_t nseg = n;  // steps of the current segment

  // speculative step kernels (a batch in launch mode 2 while its rebuilds are rare)
  bool spec = sg && sg->launch_mode == 2;
  auto launch_t = [&](int64_t t) -> int {
    if (!sg && sys->domain.kind == JDEM_DOMAIN_FREE && t == nseg - 1 && !free_done) {
      const int nparts = 64;
      ENSURE(ctx, ctx->red, 4 * 6 * (size_t)nparts);
      HIPCHK(ctx, ws_launch_box(F4(ctx->wP[cur]), N, (float*)ctx->red.p, nparts, st));
      float hp[6 * 64];
      HIPCHK(ctx, hipMemcpyAsync(hp, ctx->red.p, sizeof(hp), hipMemcpyDeviceToHost, st));
      HIPCHK(ctx, hipStreamSynchronize(st));
      for (int a = 0; a < D; ++a) {
        float lo = INFINITY, hi = -INFINITY;
        for (int b = 0; b < nparts; ++b) {
          lo = std::min(lo, hp[6 * b + a]);
          hi = std::max(hi, hp[6 * b + 3 + a]);
        }
        fbox[a] = hi - lo;
        fanc[a] = lo;
      }
      free_done = true;
      S->host_syncs++;
    }
    WS a;
    std::memset(&a, 0, sizeof(a));
    a.n = N;
    a.M = M;
    a.W64 = W64;
    a.iso = iso;
    a.law = law_of(ctx, sys->law);
    a.dom = dom;
    a.dt = sys->dt;
    a.lin = sys->lin_kind;
    a.rot = sys->rot_kind;
    for (int c = 0; c < 3; ++c) a.g[c] = sys->gravity[c];
    a.ext_f = sys->ext_force;
    a.ext_fc = sys->ext_force_com;
    a.ext_t = sys->ext_torque;
    a.has_fm = (a.g[0] != 0.0f || a.g[1] != 0.0f || a.g[2] != 0.0f || a.ext_f || a.ext_fc || a.ext_t) ? 1 : 0;
    a.orig = I32(ctx->wOrig[set]);
    cur_at[t] = cur;
    a.P = F4(ctx->wP[cur]);
    a.V = F4(ctx->wV[cur]);
    a.W = F4(ctx->wW[cur]);
    a.P2 = F4(ctx->wP[1 - cur]);
    a.V2 = F4(ctx->wV[1 - cur]);
    a.W2 = F4(ctx->wW[1 - cur]);
    a.Q = F4(ctx->wQ[set]);
    a.I = F4(ctx->wI[set]);
    a.F = F4(ctx->wF);
    a.T = F4(ctx->wT);
    a.old4 = F4(ctx->wOld);
    a.nl = I32(ctx->wNl[nlc]);
    xc_at[t] = xc;
    a.xi_rd = F4(ctx->wXi[xc]);
    a.xi_wr = F4(ctx->wXi[1 - xc]);
    a.live = U64(ctx->wLive[nlc]);
    a.do_force = t >= 1;
    a.do_after = t >= 1;
    a.do_before = t < nseg;
    a.store_ft = t == nseg;
    a.trigger = nl_trigger(L);
    a.check_disp = a.do_before;
    a.flag_rd = t >= 1 ? flags + (t & 1) : nullptr;
    a.flag_wr = flags + ((t + 1) & 1);
    a.step_idx = (int32_t)t;
    if (sg && !spec) {
      // no speculation: the rebuild chain after this kernel acts on the gate
      a.flag_rd = nullptr;
      a.flag_wr = gate;
    }
    if (sg) {
      a.block = sg->tile;
      a.seg = sg->seg;
      a.segp = sg->segp;
      a.seg_flag = sg->seg_flag;
    }
    const int sp = span_begin(ctx, 0, t);
    hipError_t e = ws_launch_step(a, D, a.law.kind, st);
    span_end(ctx, sp);
    if (e != hipSuccess) return fail(ctx, JDEM_EHIP, std::string("resident step kernel: ") + hipGetErrorString(e));
    cur = 1 - cur;
    if (a.do_force) xc = 1 - xc;  // only a force pass rewrites the springs
    return JDEM_OK;
  };

  // Batch, gated chain: step kernel, rebuild chain, step kernel, ... from
  // step t (the gate and the systems' flags clear), the chain after every
  // step but the last.
  auto gated_steps = [&](int64_t t, int64_t ns) -> int {
    auto step_and_chain = [&](int64_t tt) -> int {
      int rc3 = launch_t(tt);
      if (rc3) return rc3;
      return tt < ns ? rebuild() : JDEM_OK;
    };
    // The middle steps (force, integration and the next step's start, then
    // the gated rebuild chain) repeat with the same arguments every two
    // steps: a chunk of kGraphSteps of them is captured once in a hipGraph
    // on a private stream and replayed, which removes the host's kernel
    // launches (~27 per step) from the critical path.  Short calls and
    // timed calls (HIP events per kernel) launch one by one.
    constexpr int64_t kGraphSteps = 50;  // even: the ping-pong buffers return to their start
    if (sg->launch_mode == 0 && !ctx->timing && ns - t >= 2 * kGraphSteps) {
      if (!ctx->gstream) {
        HIPCHK(ctx, hipStreamCreateWithFlags(&ctx->gstream, hipStreamNonBlocking));
        // ordinary (fenced) events: they order the two streams' memory traffic
        HIPCHK(ctx, hipEventCreateWithFlags(&ctx->gevent[0], hipEventDisableTiming));
        HIPCHK(ctx, hipEventCreateWithFlags(&ctx->gevent[1], hipEventDisableTiming));
      }
      hipEvent_t ev_in = ctx->gevent[0], ev_out = ctx->gevent[1];
      HIPCHK(ctx, hipEventRecord(ev_in, st));
      HIPCHK(ctx, hipStreamWaitEvent(ctx->gstream, ev_in, 0));
      const hipStream_t caller = st;
      st = ctx->gstream;
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      HIPCHK(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc3 = JDEM_OK;
      for (int64_t k = 0; k < kGraphSteps && rc3 == JDEM_OK; ++k) rc3 = step_and_chain(t + k);
      const hipError_t ec = hipStreamEndCapture(st, &graph);
      if (rc3) {
        st = caller;
        if (graph) (void)hipGraphDestroy(graph);
        return rc3;
      }
      if (ec != hipSuccess) {
        st = caller;
        return fail(ctx, JDEM_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
      }
      const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        st = caller;
        return fail(ctx, JDEM_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      }
      hipError_t el = hipSuccess;
      for (; t + kGraphSteps <= ns - 1 && el == hipSuccess; t += kGraphSteps) el = hipGraphLaunch(exec, st);
      (void)hipEventRecord(ev_out, st);
      st = caller;
      (void)hipStreamWaitEvent(st, ev_out, 0);
      (void)hipGraphExecDestroy(exec);
      if (el != hipSuccess) return fail(ctx, JDEM_EHIP, std::string("hipGraphLaunch: ") + hipGetErrorString(el));
    }
    for (; t <= ns; ++t) {
      const int rc3 = step_and_chain(t);
      if (rc3) return rc3;
    }
    return JDEM_OK;
  };

  bool rebuilt_here = false;
  auto run_segment = [&](int64_t ns) -> int {
  nseg = ns;
  cur_at.assign((size_t)ns + 2, 0);
  xc_at.assign((size_t)ns + 2, 0);
  if (sg) {
    // flags[5] counts the chain runs that rebuilt something
    HIPCHK(ctx, launch_fill_i32(flags, 6, 0, st));
    HIPCHK(ctx, ws_launch_seg_init_flags(sg->n_seg, sg->n_build, sg->seg_flag, st, gate));
    int64_t next = 0;
    spec = sg->launch_mode == 2;
    if (spec) {
      // Speculative chunks, as for one system: each step kernel checks the
      // word its predecessor wrote and does nothing once a system asked for
      // a rebuild; the host reads the words once per chunk, runs the rebuild
      // chain (its gate raised from the host) and resumes at the flagged
      // step.  Once rebuilds come close together (two intervals in a row
      // of spec_min_gap steps or fewer) the rest of the call runs the gated
      // chain instead, which needs no host read.  A rebuild found by a host
      // read costs about two idle chains (the read and the chain), so
      // speculation pays while rebuilds are more than ~2 steps apart.
      const int64_t kSpecMin = sg->spec_min_gap;
      int rc2 = launch_t(0);
      if (rc2) return rc2;
      if (sg->any_unbuilt) {  // systems without a list build before the first force
        rc2 = rebuild();
        if (rc2) return rc2;
        HIPCHK(ctx, launch_fill_i32(flags, 2, 0, st));  // that chain also served any system kernel 0 flagged
      }
      next = 1;
      int64_t chunk_len = 32, last_rebuild_at = 0;
      int close = 0;
      while (next <= ns && spec) {
        const int64_t cnt = std::min<int64_t>(chunk_len, ns - next + 1);
        for (int64_t tt = next; tt < next + cnt; ++tt) {
          rc2 = launch_t(tt);
          if (rc2) return rc2;
        }
        rc2 = read_flags(ctx, 2);
        if (rc2) return rc2;
        S->host_syncs++;
        const int32_t v = std::max(ctx->host_flags[0], ctx->host_flags[1]);
        if (v == 0) {
          next += cnt;
          chunk_len = std::min<int64_t>(chunk_len * 2, 1024);
          continue;
        }
        // kernels v.. of this chunk did no work; the data is where kernel v read it
        if (v < next + cnt) {
          cur = cur_at[v];
          xc = xc_at[v];
        }
        HIPCHK(ctx, launch_fill_i32(gate, 1, 1, st));
        rc2 = rebuild();
        if (rc2) return rc2;
        HIPCHK(ctx, launch_fill_i32(flags, 2, 0, st));
        if (last_rebuild_at > 0) {
          close = v - last_rebuild_at <= kSpecMin ? close + 1 : 0;
          chunk_len = std::max<int64_t>(2, v - last_rebuild_at + 1);
        }
        last_rebuild_at = v;
        next = v;
        if (close >= 2) spec = false;  // the chain ran: gate and system flags are clear
      }
    }
    if (next <= ns) {
      if (next == 0) {
        // step 0 (integration start, displacement check) then the chain
        // builds the systems without a list and any that step 0 flagged
        int rc2 = launch_t(0);
        if (rc2) return rc2;
        if (ns > 0) {
          rc2 = rebuild();
          if (rc2) return rc2;
        }
        next = 1;
      }
      if (next <= ns) {
        const int rc2 = gated_steps(next, ns);
        if (rc2) return rc2;
      }
    }
    const int rc2 = read_flags(ctx, 6);
    if (rc2) return rc2;
    S->host_syncs++;
    const int32_t ev = ctx->host_flags[5];
    S->rebuilds += ev;
    L.n_build_times += ev;
    if (ev > 0) rebuilt_here = true;
    return JDEM_OK;
  }
  HIPCHK(ctx, launch_fill_i32(flags, 4, 0, st));
  int rc = launch_t(0);
  if (rc) return rc;
  bool pending = L.n_build_times == 0;
  int64_t next = 1;
  // speculative chunk: a chunk that runs into a rebuild wastes its remaining
  // kernels (each returns at once), one that does not saves a host read
  int64_t chunk_len = 32;
  int64_t last_rebuild_at = 0;
  while (next <= ns) {
    if (pending) {
      rc = rebuild();
      if (rc) return rc;
      HIPCHK(ctx, launch_fill_i32(flags, 2, 0, st));
      S->rebuilds++;
      rebuilt_here = true;
      if (last_rebuild_at > 0) chunk_len = std::max<int64_t>(2, next - last_rebuild_at + 1);
      last_rebuild_at = next;
      pending = false;
    }
    const int64_t cnt = std::min<int64_t>(chunk_len, ns - next + 1);
    const size_t chunk_spans = ctx->spans.size();
    for (int64_t t = next; t < next + cnt; ++t) {
      rc = launch_t(t);
      if (rc) return rc;
    }
    rc = read_flags(ctx, 3);
    if (rc) return rc;
    S->host_syncs++;
    const int32_t v = std::max(ctx->host_flags[0], ctx->host_flags[1]);
    if (v == 0) {
      next += cnt;
      chunk_len = std::min<int64_t>(chunk_len * 2, 1024);
    } else {
      for (size_t k = chunk_spans; k < ctx->spans.size(); ++k) {
        auto& sp = ctx->spans[k];
        if (sp.kind == 0 && sp.step >= v) sp.kind = 1;
      }
      // kernels v.. of this chunk did no work: the data is where kernel v read
      // it.  When the chunk's last kernel raised the flag (v == next + cnt),
      // kernel v was never launched and `cur` already points at the data.
      if (v < next + cnt) {
        cur = cur_at[v];
        xc = xc_at[v];
      }
      next = v;
      pending = true;
    }
  }
  return JDEM_OK;
  };

  // the state after each segment, written to its frame in caller order
  auto export_frame = [&](int64_t f) -> int {
    const int A = D == 3 ? 3 : 1;
    InOut fo = io;
    fo.pos = fr->pos + (size_t)f * N * D;
    fo.vel = fr->vel + (size_t)f * N * D;
    fo.ang = fr->ang_vel + (size_t)f * N * A;
    fo.q = fr->q + (size_t)f * N * 4;
    fo.force = fr->force + (size_t)f * N * D;
    fo.torque = fr->torque + (size_t)f * N * A;
    fo.old_user = nullptr;
    fo.orig = I32(ctx->wOrig[set]);
    fo.P4 = F4(ctx->wP[cur]);
    fo.V4 = F4(ctx->wV[cur]);
    fo.W4 = F4(ctx->wW[cur]);
    fo.Q4 = F4(ctx->wQ[set]);
    fo.I4 = F4(ctx->wI[set]);
    fo.old4 = F4(ctx->wOld);
    HIPCHK(ctx, ws_launch_out(fo, false, st));
    return JDEM_OK;
  };
  int rc = JDEM_OK;
  if (!fr) {
    rc = run_segment(n);
    if (rc) return rc;
  } else {
    for (int64_t f = 0; f < fr->n_frames; ++f) {
      if (fr->steps[f] > 0) {
        rc = run_segment(fr->steps[f]);
        if (rc) return rc;
      }
      rc = export_frame(f);
      if (rc) return rc;
    }
  }
  if (rebuilt_here) L.overflow = ctx->host_flags[2] != 0;
  S->overflow = L.overflow;
  // --- exit: back to the caller's order and layouts ---
  if (L.resident_out) {
    // the state in the caller's order; the list, build positions and springs
    // stay in a new working set (jdem_resident_export converts on demand)
    io.orig = I32(ctx->wOrig[set]);
    io.old4 = F4(ctx->wOld);
    io.P4 = F4(ctx->wP[cur]);
    io.V4 = F4(ctx->wV[cur]);
    io.W4 = F4(ctx->wW[cur]);
    io.Q4 = F4(ctx->wQ[set]);
    io.I4 = F4(ctx->wI[set]);
    io.old_user = nullptr;
    HIPCHK(ctx, ws_launch_out(io, false, st));
    if (hist && M > 0)
      HIPCHK(ctx, ws_launch_unpack_springs(N, M, W64, F4(ctx->wXi[xc]), U64(ctx->wLive[nlc]), F4(ctx->wW[cur]),
                                           F4(ctx->wXiRank), nullptr, nullptr, st));
    jdem_resident* res = new (std::nothrow) jdem_resident();
    if (!res) return fail(ctx, JDEM_ENOMEM, "resident working set");
    res->ctx = ctx;
    res->n = N;
    res->M = M;
    res->W64 = W64;
    res->dim = D;
    res->hist = hist;
    // hand the buffers over; ensure() refills the emptied slots from the pool
    std::swap(res->orig, ctx->wOrig[set]);
    std::swap(res->old4, ctx->wOld);
    std::swap(res->nl, ctx->wNl[nlc]);
    std::swap(res->live, ctx->wLive[nlc]);
    if (hist && M > 0) std::swap(res->xi, ctx->wXiRank);
    *L.resident_out = res;
    if (!R) {
      // The first working set of a system reserves its successor's buffers
      // now: the next call starts from this one while it is still alive, and
      // its own working set is handed back to the pool only after that call.
      ENSURE(ctx, ctx->wOrig[set], ni);
      ENSURE(ctx, ctx->wOld, n4);
      ENSURE(ctx, ctx->wNl[nlc], 4 * (size_t)N * (size_t)std::max(M, 1));
      ENSURE(ctx, ctx->wLive[nlc], 8 * (size_t)N * W64);
      if (hist && M > 0) ENSURE(ctx, ctx->wXiRank, 16 * (size_t)N * (size_t)std::max(M, 1));
    }
    if (free_done) {
      for (int a = 0; a < D; ++a) {
        sys->domain.box_size[a] = fbox[a];
        sys->domain.anchor[a] = fanc[a];
        sys->domain.inv_box_size[a] = 1.0f / fbox[a];
        S->free_box[a] = fbox[a];
        S->free_anchor[a] = fanc[a];
      }
    }
    S->steps = n;
    return JDEM_OK;
  }
  io.orig = I32(ctx->wOrig[set]);
  io.old4 = F4(ctx->wOld);
  io.P4 = F4(ctx->wP[cur]);
  io.V4 = F4(ctx->wV[cur]);
  io.W4 = F4(ctx->wW[cur]);
  io.Q4 = F4(ctx->wQ[set]);
  io.I4 = F4(ctx->wI[set]);
  io.nl = I32(ctx->wNl[nlc]);
  io.live = U64(ctx->wLive[nlc]);
  io.xi = F4(ctx->wXiRank);
  if (hist && M > 0)
    HIPCHK(ctx, ws_launch_unpack_springs(N, M, W64, F4(ctx->wXi[xc]), U64(ctx->wLive[nlc]), F4(ctx->wW[cur]),
                                         F4(ctx->wXiRank), nullptr, nullptr, st));
  HIPCHK(ctx, ws_launch_out(io, M > 0, st));
  if (free_done) {
    for (int a = 0; a < D; ++a) {
      sys->domain.box_size[a] = fbox[a];
      sys->domain.anchor[a] = fanc[a];
      sys->domain.inv_box_size[a] = 1.0f / fbox[a];
      S->free_box[a] = fbox[a];
      S->free_anchor[a] = fanc[a];
    }
  }
  S->steps = n;
  return JDEM_OK;
}

int src_of(int collider_kind) {
  switch (collider_kind) {
    case JDEM_COLLIDER_NAIVE:
      return SRC_NAIVE;
    case JDEM_COLLIDER_CELL_LIST:
      return SRC_CELLS;
    case JDEM_COLLIDER_NEIGHBOR_LIST:
      return SRC_NLIST;
    default:
      return 0;
  }
}

// Clumps (defined with the float64 helpers at the end of this file).
int clump_step(jdem_ctx* ctx, jdem_system* sys, int64_t n, jdem_step_stats* S);
int clump_force(jdem_ctx* ctx, jdem_system* sys);
int clump_reduce(jdem_ctx* ctx, jdem_system* sys, int ncomp, double* host_out, bool kinetic);
int clump_step(jdem_ctx* ctx, jdem_system_f64* sys, int64_t n, jdem_step_stats* S);
int clump_force(jdem_ctx* ctx, jdem_system_f64* sys);
int clump_reduce(jdem_ctx* ctx, jdem_system_f64* sys, int ncomp, double* host_out, bool kinetic);

// The calls that do not take clumps refuse them.
int reject_clumps(jdem_ctx* ctx, const jdem_state& s, const char* what) {
  if (s.clump_id || s.pos_p) return fail(ctx, JDEM_EUNSUPPORTED, std::string(what) + " with clumps");
  return JDEM_OK;
}

// Fill the source fields of `a` for the collider (building cells if needed).
int prepare_source(jdem_ctx* ctx, jdem_system* sys, StepArgs& a, const float* pos) {
  if (sys->collider_kind == JDEM_COLLIDER_CELL_LIST) {
    Cells c;
    a.grid = grid_of(sys->grid);
    int rc = build_cells(ctx, pos, sys->state.rad, sys->state.n, sys->state.dim, a.dom, a.grid, &c);
    if (rc) return rc;
    a.perm = c.perm;
    a.skeys = c.skeys;
    a.cstart = c.cstart;
    a.cend = c.cend;
    a.spos = c.spos;
  }
  return JDEM_OK;
}

}  // namespace

extern "C" {

int jdem_abi_version(void) { return JDEM_ABI_VERSION; }

const char* jdem_build_target(void) { return JDEM_TARGET; }

int jdem_ctx_create(int device, void* stream, jdem_ctx** out) {
  if (!out) return JDEM_EINVAL;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return JDEM_EHIP;
  if (device < 0 || device >= count) return JDEM_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return JDEM_EHIP;
  // the fixed scratch of jdem_cell_build_noalloc, once per device
  if (noalloc_init(device) != JDEM_OK) return JDEM_EHIP;
  jdem_ctx* c = new (std::nothrow) jdem_ctx();
  if (!c) return JDEM_ENOMEM;
  c->device = device;
  c->stream = (hipStream_t)stream;
  *out = c;
  return JDEM_OK;
}

int jdem_ctx_destroy(jdem_ctx* ctx) {
  if (!ctx) return JDEM_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (Buf* b : ctx->bufs())
    if (b->p) (void)hipFree(b->p);
  pool_trim(ctx);
  if (ctx->gstream) {
    (void)hipStreamDestroy(ctx->gstream);
    (void)hipEventDestroy(ctx->gevent[0]);
    (void)hipEventDestroy(ctx->gevent[1]);
  }
  if (ctx->host_flags) (void)hipHostFree(ctx->host_flags);
  for (auto& s : ctx->spans) {
    (void)hipEventDestroy(s.start);
    (void)hipEventDestroy(s.stop);
  }
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  delete ctx;
  return JDEM_OK;
}

int jdem_ctx_set_stream(jdem_ctx* ctx, void* stream) {
  if (!ctx) return JDEM_EINVAL;
  ctx->stream = (hipStream_t)stream;
  return JDEM_OK;
}

const char* jdem_last_error(const jdem_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int jdem_ctx_scratch_bytes(const jdem_ctx* ctx, size_t* bytes) {
  if (!ctx || !bytes) return JDEM_EINVAL;
  *bytes = const_cast<jdem_ctx*>(ctx)->scratch_bytes();
  return JDEM_OK;
}

int jdem_ctx_trim(jdem_ctx* ctx) {
  if (!ctx) return JDEM_EINVAL;
  (void)hipSetDevice(ctx->device);
  pool_trim(ctx);
  return JDEM_OK;
}

int jdem_resident_release(jdem_ctx* ctx, jdem_resident* res) {
  if (!ctx) return JDEM_EINVAL;
  if (!res) return JDEM_OK;
  if (res->ctx != ctx) return fail(ctx, JDEM_EINVAL, "working set of another context");
  for (Buf* b : {&res->orig, &res->old4, &res->nl, &res->live, &res->xi}) pool_give(ctx, *b);
  delete res;
  return JDEM_OK;
}

int jdem_resident_export(jdem_ctx* ctx, const jdem_resident* res, int32_t* nl, float* old_pos, float* hist_xi,
                         uint32_t* hist_live) {
  if (!ctx) return JDEM_EINVAL;
  if (!res) return fail(ctx, JDEM_EINVAL, "null working set");
  if (res->ctx != ctx) return fail(ctx, JDEM_EINVAL, "working set of another context");
  if (res->n == 0) return JDEM_OK;
  (void)hipSetDevice(ctx->device);
  InOut io;
  std::memset(&io, 0, sizeof(io));
  io.n = res->n;
  io.dim = res->dim;
  io.M = res->M;
  io.W64 = res->W64;
  io.orig = (const int32_t*)res->orig.p;
  io.old4 = (float4*)res->old4.p;
  io.nl = (int32_t*)res->nl.p;
  io.live = (uint64_t*)res->live.p;
  io.xi = (float4*)res->xi.p;
  io.old_user = old_pos;
  io.nl_user = nl;
  // the springs come with the list they belong to
  io.xi_user = (res->hist && nl) ? hist_xi : nullptr;
  io.live_user = (res->hist && nl) ? hist_live : nullptr;
  if (io.xi_user && !io.live_user) return fail(ctx, JDEM_EINVAL, "hist_xi without hist_live");
  HIPCHK(ctx, ws_launch_export(io, ctx->stream));
  return JDEM_OK;
}

int jdem_ctx_set_timing(jdem_ctx* ctx, int enable) {
  if (!ctx) return JDEM_EINVAL;
  ctx->timing = enable != 0;
  return JDEM_OK;
}

int jdem_ctx_get_timing(jdem_ctx* ctx, jdem_timing* out) {
  if (!ctx || !out) return JDEM_EINVAL;
  std::memset(out, 0, sizeof(*out));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& s : ctx->spans) {
    float ms = 0.0f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, s.start, s.stop));
    if (s.kind == 0) {
      out->step_kernel_ms += ms;
      out->step_kernels++;
    } else if (s.kind == 1) {
      out->idle_kernel_ms += ms;
      out->idle_kernels++;
    } else {
      out->rebuild_ms += ms;
      out->rebuilds++;
    }
    ctx->event_pool.push_back(s.start);
    ctx->event_pool.push_back(s.stop);
  }
  ctx->spans.clear();
  return JDEM_OK;
}

int jdem_cell_build(jdem_ctx* ctx, const jdem_state* st, const jdem_domain* dom, const jdem_grid* grid,
                    int32_t* perm, int32_t* sorted_hash, int32_t* cell_start, int32_t* cell_end) {
  if (!ctx) return JDEM_EINVAL;
  if (!st || !dom || !grid || !perm || !sorted_hash || !cell_start || !cell_end)
    return fail(ctx, JDEM_EINVAL, "null argument");
  if (grid->n_cells <= 0 || grid->n_cells > (int64_t)INT32_MAX)
    return fail(ctx, JDEM_EUNSUPPORTED, "cell grid size out of the int32 hash range");
  (void)hipSetDevice(ctx->device);
  Cells c;
  int rc = build_cells(ctx, st->pos, st->rad, st->n, st->dim, dom_of(*dom), grid_of(*grid), &c);
  if (rc) return rc;
  const size_t n4 = 4 * (size_t)st->n;
  if (st->n > 0) {
    HIPCHK(ctx, hipMemcpyAsync(perm, c.perm, n4, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(sorted_hash, c.skeys, n4, hipMemcpyDeviceToDevice, ctx->stream));
  }
  HIPCHK(ctx, hipMemcpyAsync(cell_start, c.cstart, 4 * (size_t)grid->n_cells, hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(cell_end, c.cend, 4 * (size_t)grid->n_cells, hipMemcpyDeviceToDevice, ctx->stream));
  return JDEM_OK;
}

int jdem_neighbor_list(jdem_ctx* ctx, const jdem_system* sys, int32_t collider_kind, const jdem_grid* grid,
                       float cutoff, int32_t max_neighbors, int32_t* nl, int32_t* overflow) {
  if (!ctx) return JDEM_EINVAL;
  jdem_system tmp = *sys;
  tmp.collider_kind = JDEM_COLLIDER_NONE;
  int rc = check_system(ctx, &tmp);
  if (rc) return rc;
  rc = reject_clumps(ctx, sys->state, "jdem_neighbor_list");
  if (rc) return rc;
  if (max_neighbors < 0 || (max_neighbors > 0 && sys->state.n > 0 && !nl))
    return fail(ctx, JDEM_EINVAL, "bad neighbour-list output");
  if (collider_kind != JDEM_COLLIDER_CELL_LIST && collider_kind != JDEM_COLLIDER_NAIVE)
    return fail(ctx, JDEM_EINVAL, "collider must be CellList or naive");
  if (collider_kind == JDEM_COLLIDER_CELL_LIST && (!grid || grid->n_cells <= 0 || grid->n_cells > (int64_t)INT32_MAX))
    return fail(ctx, JDEM_EUNSUPPORTED, "cell grid size out of the int32 hash range");
  (void)hipSetDevice(ctx->device);
  rc = ensure_flags(ctx);
  if (rc) return rc;
  HIPCHK(ctx, launch_fill_i32((int32_t*)ctx->flags.p + 2, 1, 0, ctx->stream));
  if (max_neighbors > 0 && sys->state.n > 0) {
    jdem_grid g0;
    std::memset(&g0, 0, sizeof(g0));
    rc = build_nlist(ctx, sys, collider_kind, grid ? *grid : g0, sys->state.pos, cutoff, max_neighbors, nl);
    if (rc) return rc;
  }
  rc = read_flags(ctx, 3);
  if (rc) return rc;
  if (overflow) *overflow = ctx->host_flags[2] != 0;
  return JDEM_OK;
}

int jdem_compute_force(jdem_ctx* ctx, jdem_system* sys) {
  int rc = check_system(ctx, sys);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  if (sys->collider_kind == JDEM_COLLIDER_NONE) {
    const jdem_state& s = sys->state;
    HIPCHK(ctx, hipMemsetAsync(s.force, 0, 4 * (size_t)s.n * s.dim, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(s.torque, 0, 4 * (size_t)s.n * (s.dim == 3 ? 3 : 1), ctx->stream));
    return JDEM_OK;
  }
  if (sys->state.clump_id) return clump_force(ctx, sys);
  if (sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST) {
    bool rebuilt = false;
    rc = nl_check_and_rebuild(ctx, sys, &rebuilt);
    if (rc) return rc;
  }
  StepArgs a = base_args(ctx, sys);
  rc = prepare_source(ctx, sys, a, sys->state.pos);
  if (rc) return rc;
  a.do_force = 1;
  a.write_kin = 0;
  HIPCHK(ctx, launch_step(a, sys->state.dim, a.law.kind, src_of(sys->collider_kind), ctx->stream));
  return JDEM_OK;
}

static int reduce_f64(jdem_ctx* ctx, const jdem_system* sys, int ncomp, double* host_out,
                      hipError_t (*launch)(const StepArgs&, int, int, int, hipStream_t), bool by_collider) {
  const jdem_state& s = sys->state;
  const int64_t n = s.n;
  const int nparts = 256;
  ENSURE(ctx, ctx->red, 8 * ((size_t)(n > 0 ? n : 1) * ncomp + (size_t)nparts * ncomp + ncomp));
  double* per = (double*)ctx->red.p;
  double* part = per + (size_t)(n > 0 ? n : 1) * ncomp;
  double* out = part + (size_t)nparts * ncomp;
  StepArgs a = base_args(ctx, sys);
  jdem_system* ms = const_cast<jdem_system*>(sys);
  if (by_collider) {
    int rc = prepare_source(ctx, ms, a, s.pos);
    if (rc) return rc;
  }
  a.e_out = per;
  if (n > 0) {
    if (by_collider) {
      HIPCHK(ctx, launch(a, s.dim, a.law.kind, src_of(sys->collider_kind), ctx->stream));
    } else {
      HIPCHK(ctx, launch_kinetic(a, s.dim, ctx->stream));
    }
    HIPCHK(ctx, launch_sum_f64(per, n, ncomp, ncomp, part, out, nparts, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(host_out, out, 8 * (size_t)ncomp, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  } else {
    for (int c = 0; c < ncomp; ++c) host_out[c] = 0.0;
  }
  return JDEM_OK;
}

int jdem_potential_energy(jdem_ctx* ctx, jdem_system* sys, double* energy) {
  int rc = check_system(ctx, sys);
  if (rc) return rc;
  if (!energy) return fail(ctx, JDEM_EINVAL, "null output");
  (void)hipSetDevice(ctx->device);
  if (sys->collider_kind == JDEM_COLLIDER_NONE) {
    *energy = 0.0;
    return JDEM_OK;
  }
  if (sys->state.clump_id) return clump_reduce(ctx, sys, 1, energy, false);
  if (sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST) {
    bool rebuilt = false;
    rc = nl_check_and_rebuild(ctx, sys, &rebuilt);
    if (rc) return rc;
  }
  return reduce_f64(ctx, sys, 1, energy, launch_energy, true);
}

int jdem_contact_virial(jdem_ctx* ctx, jdem_system* sys, double* w) {
  int rc = check_system(ctx, sys);
  if (rc) return rc;
  rc = reject_clumps(ctx, sys->state, "the contact virial");
  if (rc) return rc;
  if (!w) return fail(ctx, JDEM_EINVAL, "null output");
  (void)hipSetDevice(ctx->device);
  const int D = sys->state.dim;
  if (sys->collider_kind == JDEM_COLLIDER_NONE) {
    for (int c = 0; c < D * D; ++c) w[c] = 0.0;
    return JDEM_OK;
  }
  // a NeighborList that is unbuilt or stale is refreshed first, as
  // get_pair_forces_and_ids does through create_neighbor_list (contacts.py:58-62)
  if (sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST) {
    bool rebuilt = false;
    rc = nl_check_and_rebuild(ctx, sys, &rebuilt);
    if (rc) return rc;
  }
  return reduce_f64(ctx, sys, D * D, w, launch_virial, true);
}

int jdem_kinetic_energy_parts(jdem_ctx* ctx, const jdem_system* sys, double* translational, double* rotational) {
  int rc = check_system(ctx, sys);
  if (rc) return rc;
  if (!translational || !rotational) return fail(ctx, JDEM_EINVAL, "null output");
  (void)hipSetDevice(ctx->device);
  double parts[2] = {0.0, 0.0};
  rc = sys->state.clump_id ? clump_reduce(ctx, const_cast<jdem_system*>(sys), 2, parts, true)
                           : reduce_f64(ctx, sys, 2, parts, nullptr, false);
  if (rc) return rc;
  *translational = parts[0];
  *rotational = parts[1];
  return JDEM_OK;
}

int jdem_kinetic_energy(jdem_ctx* ctx, const jdem_system* sys, double* energy) {
  if (!energy) return ctx ? fail(ctx, JDEM_EINVAL, "null output") : JDEM_EINVAL;
  double t = 0.0, r = 0.0;
  const int rc = jdem_kinetic_energy_parts(ctx, sys, &t, &r);
  if (rc) return rc;
  *energy = t + r;
  return JDEM_OK;
}

static int free_box(jdem_ctx* ctx, const float* pos, const jdem_system* sys, float box[3], float anchor[3]) {
  const jdem_state& s = sys->state;
  const int nparts = 64;
  ENSURE(ctx, ctx->red, 4 * 6 * (size_t)nparts);
  HIPCHK(ctx, launch_box_minmax(pos, s.rad, s.n, s.dim, (float*)ctx->red.p, nparts, ctx->stream, s.pos_p));
  float hp[6 * 64];
  HIPCHK(ctx, hipMemcpyAsync(hp, ctx->red.p, sizeof(hp), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int a = 0; a < s.dim; ++a) {
    float lo = INFINITY, hi = -INFINITY;
    for (int b = 0; b < nparts; ++b) {
      lo = std::min(lo, hp[6 * b + a]);
      hi = std::max(hi, hp[6 * b + 3 + a]);
    }
    box[a] = hi - lo;
    anchor[a] = lo;
  }
  return JDEM_OK;
}

int jdem_step_stage(jdem_ctx* ctx, jdem_system* sys, int32_t stage) {
  int rc = check_system(ctx, sys);
  if (rc) return rc;
  rc = reject_clumps(ctx, sys->state, "jdem_step_stage");
  if (rc) return rc;
  if (stage < JDEM_STAGE_DOMAIN_APPLY || stage > JDEM_STAGE_ROT_AFTER)
    return fail(ctx, JDEM_EINVAL, "unknown stage");
  (void)hipSetDevice(ctx->device);
  if (stage == JDEM_STAGE_DOMAIN_APPLY && sys->domain.kind == JDEM_DOMAIN_FREE) {
    if (sys->state.n == 0) return JDEM_OK;
    float box[3], anc[3];
    rc = free_box(ctx, sys->state.pos, sys, box, anc);
    if (rc) return rc;
    for (int a = 0; a < sys->state.dim; ++a) {
      sys->domain.box_size[a] = box[a];
      sys->domain.anchor[a] = anc[a];
      sys->domain.inv_box_size[a] = 1.0f / box[a];
    }
    return JDEM_OK;
  }
  StepArgs a = base_args(ctx, sys);
  HIPCHK(ctx, launch_stage(a, sys->state.dim, stage, ctx->stream));
  return JDEM_OK;
}

int jdem_batch_step(jdem_ctx* ctx, jdem_system* sys, const jdem_batch* batch, int64_t n, jdem_step_stats* stats) {
  int rc = check_system(ctx, sys, true, true);
  if (rc) return rc;
  if (!batch || batch->n_systems < 1 || !batch->offsets || !batch->domains || !batch->grids || !batch->dt ||
      !batch->cutoff || !batch->skin || !batch->n_build_times)
    return fail(ctx, JDEM_EINVAL, "incomplete batch description");
  if (n < 0) return fail(ctx, JDEM_EINVAL, "negative step count");
  if (batch->launch_mode < 0 || batch->launch_mode > 2) return fail(ctx, JDEM_EINVAL, "launch_mode must be 0, 1 or 2");
  if (batch->tile_width != 0 && batch->tile_width != 128 && batch->tile_width != 256)
    return fail(ctx, JDEM_EINVAL, "tile_width must be 128 or 256");
  if (batch->spec_min_gap < 0) return fail(ctx, JDEM_EINVAL, "spec_min_gap must be >= 0");
  if (sys->collider_kind != JDEM_COLLIDER_NEIGHBOR_LIST || sys->nlist.secondary != JDEM_COLLIDER_CELL_LIST)
    return fail(ctx, JDEM_EUNSUPPORTED, "batches run the NeighborList collider over a CellList");
  const int B = batch->n_systems;
  const int64_t N = sys->state.n;
  if (batch->offsets[0] != 0 || batch->offsets[B] != N) return fail(ctx, JDEM_EINVAL, "offsets do not cover the state");
  std::vector<SegP> host((size_t)B);
  int64_t base = 0;
  for (int k = 0; k < B; ++k) {
    if (batch->offsets[k + 1] < batch->offsets[k]) return fail(ctx, JDEM_EINVAL, "offsets must not decrease");
    const jdem_domain& d = batch->domains[k];
    if (d.kind != JDEM_DOMAIN_PERIODIC && d.kind != JDEM_DOMAIN_REFLECT)
      return fail(ctx, JDEM_EUNSUPPORTED, "batched systems need a periodic or reflect domain");
    const jdem_grid& g = batch->grids[k];
    if (g.n_cells <= 0 || g.search_range < 1) return fail(ctx, JDEM_EINVAL, "bad cell grid of a batched system");
    SegP& p = host[k];
    std::memset(&p, 0, sizeof(p));
    p.dom = dom_of(d);
    p.grid = grid_of(g);
    p.cell_base = (uint32_t)base;
    p.dt = batch->dt[k];
    const float lc = batch->cutoff[k] + batch->skin[k];
    p.cutoff_sq = lc * lc;
    const float sk = batch->skin[k] * batch->skin[k];
    p.trigger = sk / 4.0f;
    base += g.n_cells;
    if (base > (int64_t)UINT32_MAX) return fail(ctx, JDEM_EUNSUPPORTED, "batch has more than 2^32 cells");
  }
  (void)hipSetDevice(ctx->device);
  rc = ensure_flags(ctx);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  ENSURE(ctx, ctx->segp, sizeof(SegP) * (size_t)B + 4 * 8 * ((size_t)B + 1));
  ENSURE(ctx, ctx->cend, 4 * (size_t)std::max<int64_t>(N, 1));
  ENSURE(ctx, ctx->segid, 4 * (size_t)std::max<int64_t>(N, 1));
  ENSURE(ctx, ctx->segflag, 3 * 4 * (size_t)B);
  SegP* dsegp = (SegP*)ctx->segp.p;
  int64_t* doff = (int64_t*)((char*)ctx->segp.p + sizeof(SegP) * (size_t)B);
  int32_t* dbuild = (int32_t*)ctx->segflag.p;
  bool any_unbuilt = false;
  for (int k = 0; k < B; ++k) any_unbuilt |= batch->n_build_times[k] == 0;
  HIPCHK(ctx, hipMemcpyAsync(dsegp, host.data(), sizeof(SegP) * (size_t)B, hipMemcpyHostToDevice, st));
  HIPCHK(ctx, hipMemcpyAsync(doff, batch->offsets, 8 * ((size_t)B + 1), hipMemcpyHostToDevice, st));
  HIPCHK(ctx, hipMemcpyAsync(dbuild, batch->n_build_times, 4 * (size_t)B, hipMemcpyHostToDevice, st));
  HIPCHK(ctx, ws_launch_seg_ids(doff, B, N, (int32_t*)ctx->segid.p, st));
  // the host arrays above are read by the copies before this call returns
  HIPCHK(ctx, hipStreamSynchronize(st));
  SegRun sg;
  sg.n_seg = B;
  sg.seg = (const int32_t*)ctx->segid.p;
  sg.segp = dsegp;
  sg.total_cells = base;
  sg.offsets = doff;
  sg.cpref = doff + (B + 1);
  sg.ccref = doff + 2 * (B + 1);
  sg.wpref = doff + 3 * (B + 1);
  sg.n_build = dbuild;
  sg.seg_flag = dbuild + B;
  sg.seg_mode = dbuild + 2 * B;
  sg.any_unbuilt = any_unbuilt;
  sg.launch_mode = batch->launch_mode;
  sg.tile = batch->tile_width == 128 ? 128 : 256;
  sg.spec_min_gap = batch->spec_min_gap > 0 ? batch->spec_min_gap : 2;
  jdem_step_stats local;
  jdem_step_stats* S = stats ? stats : &local;
  std::memset(S, 0, sizeof(*S));
  if (N == 0 || n == 0) {
    S->steps = n;
    return JDEM_OK;
  }
  rc = resident_nl_step(ctx, sys, n, S, &sg);
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpyAsync(batch->n_build_times, dbuild, 4 * (size_t)B, hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipStreamSynchronize(st));
  return JDEM_OK;
}

int jdem_minimize(jdem_ctx* ctx, jdem_system* sys, const jdem_fire* fire, int64_t max_steps, double pe_tol,
                  double pe_diff_tol, double force_tol, jdem_minimize_stats* stats) {
  int rc = check_system(ctx, sys);
  if (rc) return rc;
  if (!fire) return fail(ctx, JDEM_EINVAL, "null FIRE arguments");
  if (max_steps < 0) return fail(ctx, JDEM_EINVAL, "max_steps < 0");
  rc = reject_clumps(ctx, sys->state, "minimize");
  if (rc) return rc;
  if (law_has_history(sys->law)) return fail(ctx, JDEM_EUNSUPPORTED, "minimize with a history law");
  if (sys->ext_force || sys->ext_force_com || sys->ext_torque)
    return fail(ctx, JDEM_EUNSUPPORTED, "minimize with external force arrays");
  (void)hipSetDevice(ctx->device);
  jdem_state& s = sys->state;
  const int64_t N = s.n;
  const int D = s.dim, A = D == 3 ? 3 : 1;
  hipStream_t st = ctx->stream;
  const int nparts = std::max(1u, nblocks(N));
  ENSURE(ctx, ctx->fireV, 4 * (size_t)std::max<int64_t>(N, 1) * (D + A));
  ENSURE(ctx, ctx->fireDev, sizeof(FireDev) + 8 * (size_t)nparts);
  ENSURE(ctx, ctx->red, 8 * ((size_t)std::max<int64_t>(N, 1) + 256 + 1));
  FireArgs f;
  std::memset(&f, 0, sizeof(f));
  f.n = N;
  f.dim = D;
  f.p.dt0 = (float)fire->dt;
  f.p.alpha_init = (float)fire->alpha_init;
  f.p.f_inc = (float)fire->f_inc;
  f.p.f_dec = (float)fire->f_dec;
  f.p.f_alpha = (float)fire->f_alpha;
  // dt * dt_{max,min}_scale are Python floats in the reference, rounded once
  f.p.dt_max = (float)(fire->dt * fire->dt_max_scale);
  f.p.dt_min = (float)(fire->dt * fire->dt_min_scale);
  f.p.n_min = fire->n_min;
  f.p.n_bad_max = fire->n_bad_max;
  f.st = (FireDev*)ctx->fireDev.p;
  f.part = (double*)((char*)ctx->fireDev.p + sizeof(FireDev));
  f.nparts = nparts;
  f.force = s.force;
  f.torque = s.torque;
  f.fixed = s.fixed;
  f.vel_pos = (float*)ctx->fireV.p;
  f.vel_rot = f.vel_pos + (size_t)std::max<int64_t>(N, 1) * D;
  f.pos = s.pos;
  f.q = s.q;
  for (int c = 0; c < 3; ++c) f.g[c] = sys->gravity[c];
  f.mass = s.mass;
  double* per = (double*)ctx->red.p;
  double* part = per + std::max<int64_t>(N, 1);
  double* out = part + 256;
  f.e_out = per;
  const bool grav = sys->gravity[0] != 0.0f || sys->gravity[1] != 0.0f || sys->gravity[2] != 0.0f;
  // optimiser state: zero velocities, dt = dt0, alpha = alpha_init (optimizers.py:231-239)
  HIPCHK(ctx, hipMemsetAsync(ctx->fireV.p, 0, 4 * (size_t)std::max<int64_t>(N, 1) * (D + A), st));
  FireDev h0;
  std::memset(&h0, 0, sizeof(h0));
  h0.dt = (float)fire->dt;
  h0.alpha = (float)fire->alpha_init;
  HIPCHK(ctx, hipMemcpyAsync(f.st, &h0, sizeof(h0), hipMemcpyHostToDevice, st));

  // eval_step (routines.py:228-245): force, ForceManager, energy, max |grad|
  double pe = 0.0, gmax = 0.0;
  auto eval = [&]() -> int {
    if (sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST) {
      bool rebuilt = false;
      int r = nl_check_and_rebuild(ctx, sys, &rebuilt);
      if (r) return r;
    }
    StepArgs a = base_args(ctx, sys);
    if (sys->collider_kind == JDEM_COLLIDER_NONE) {
      HIPCHK(ctx, hipMemsetAsync(s.force, 0, 4 * (size_t)N * D, st));
      HIPCHK(ctx, hipMemsetAsync(s.torque, 0, 4 * (size_t)N * A, st));
    } else if (N > 0) {
      int r = prepare_source(ctx, sys, a, s.pos);
      if (r) return r;
      a.do_force = 1;
      HIPCHK(ctx, launch_step(a, D, a.law.kind, src_of(sys->collider_kind), st));
    }
    if (N > 0) HIPCHK(ctx, launch_stage(a, D, JDEM_STAGE_FORCE_MANAGER, st));
    HIPCHK(ctx, hipMemsetAsync(&f.st->max_grad_bits, 0, sizeof(unsigned), st));
    HIPCHK(ctx, launch_fire_grad_max(f, st));
    double h[2] = {0.0, 0.0};
    if (N > 0) {
      if (sys->collider_kind != JDEM_COLLIDER_NONE) {
        a.e_out = per;
        HIPCHK(ctx, launch_energy(a, D, a.law.kind, src_of(sys->collider_kind), st));
      } else {
        HIPCHK(ctx, hipMemsetAsync(per, 0, 8 * (size_t)N, st));
      }
      if (grav) HIPCHK(ctx, launch_fire_gravity_energy(f, st));
      HIPCHK(ctx, launch_sum_f64(per, N, 1, 1, part, out, 256, st));
      HIPCHK(ctx, hipMemcpyAsync(&h[0], out, 8, hipMemcpyDeviceToHost, st));
    }
    unsigned bits = 0;
    HIPCHK(ctx, hipMemcpyAsync(&bits, &f.st->max_grad_bits, sizeof(bits), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    float g;
    std::memcpy(&g, &bits, sizeof(g));
    pe = h[0];
    gmax = (double)g;
    return JDEM_OK;
  };

  rc = eval();
  if (rc) return rc;
  double prev = INFINITY;
  int64_t steps = 0;
  const double tiny = 2.2250738585072014e-308;  // smallest normal float64
  const double dN = (double)std::max<int64_t>(N, 1);
  while (steps < max_steps) {
    const bool conv_pe = std::fabs(pe / dN) <= pe_tol;
    const double denom = std::max(std::max(std::fabs(pe), std::fabs(prev)), tiny);
    const bool conv_rel = std::fabs(pe - prev) / denom < pe_diff_tol;
    const bool conv_f = gmax <= force_tol;
    if (conv_pe || conv_rel || conv_f) break;
    HIPCHK(ctx, launch_fire_power(f, st));
    HIPCHK(ctx, launch_fire_decide(f, st));
    HIPCHK(ctx, launch_fire_move(f, st));
    HIPCHK(ctx, launch_fire_commit(f, st));  // dt, alpha <- the values just decided
    prev = pe;
    rc = eval();
    if (rc) return rc;
    ++steps;
  }
  if (stats) {
    FireDev h;
    HIPCHK(ctx, hipMemcpyAsync(&h, f.st, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    stats->steps = steps;
    stats->pe = pe / dN;
    stats->max_grad = gmax;
    stats->power = h.power;
    stats->overflow = sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST ? sys->nlist.overflow : 0;
  }
  return JDEM_OK;
}

int jdem_precise_minimize(jdem_ctx* ctx, jdem_precise* p, const jdem_fire* fire, int64_t max_steps, double pe_tol,
                          double pe_diff_tol, double force_tol, jdem_minimize_stats* stats, double* virial) {
  if (!ctx) return JDEM_EINVAL;
  if (!p || !fire) return fail(ctx, JDEM_EINVAL, "null argument");
  const int D = p->dim;
  const int64_t N = p->n;
  if ((D != 2 && D != 3) || N < 0 || N > (int64_t)INT32_MAX) return fail(ctx, JDEM_EINVAL, "bad dim or n");
  if (N > 0 && (!p->cidx || !p->off || !p->rad || !p->force)) return fail(ctx, JDEM_EINVAL, "null array");
  int64_t n_cells = 1;
  for (int a = 0; a < D; ++a) {
    if (p->dims[a] < 3 || !(p->cell[a] > 0.0f)) return fail(ctx, JDEM_EINVAL, "need dims >= 3 and cell > 0");
    n_cells *= p->dims[a];
  }
  if (n_cells > (int64_t)INT32_MAX) return fail(ctx, JDEM_EUNSUPPORTED, "too many cells");
  if (max_steps < 0) return fail(ctx, JDEM_EINVAL, "max_steps < 0");
  (void)hipSetDevice(ctx->device);
  hipStream_t st = ctx->stream;
  const int A = D == 3 ? 3 : 1;
  const size_t Nn = (size_t)std::max<int64_t>(N, 1);
  const int nparts = std::max(1u, nblocks(N));
  // scratch: FIRE velocities and zero torque / fixed / rotation velocities,
  // per-sphere energies and virials, cell keys and starts
  ENSURE(ctx, ctx->fireV, 4 * Nn * (D + 2 * A) + Nn);
  ENSURE(ctx, ctx->fireDev, sizeof(FireDev) + 8 * (size_t)nparts + 64 + sizeof(PrecStop));
  ENSURE(ctx, ctx->red, 8 * (Nn * (1 + D * D) + 256 * D * D + D * D + 8));
  ENSURE(ctx, ctx->wKeys, 4 * Nn);
  ENSURE(ctx, ctx->cstart, 4 * ((size_t)n_cells + 1));
  ENSURE(ctx, ctx->spos, 16 * Nn);
  float* vpos = (float*)ctx->fireV.p;
  float* vrot = vpos + Nn * D;
  float* zt = vrot + Nn * A;
  uint8_t* zfix = (uint8_t*)(zt + Nn * A);
  HIPCHK(ctx, hipMemsetAsync(ctx->fireV.p, 0, 4 * Nn * (D + 2 * A) + Nn, st));
  FireArgs f;
  std::memset(&f, 0, sizeof(f));
  f.n = N;
  f.dim = D;
  f.p.dt0 = (float)fire->dt;
  f.p.alpha_init = (float)fire->alpha_init;
  f.p.f_inc = (float)fire->f_inc;
  f.p.f_dec = (float)fire->f_dec;
  f.p.f_alpha = (float)fire->f_alpha;
  f.p.dt_max = (float)(fire->dt * fire->dt_max_scale);
  f.p.dt_min = (float)(fire->dt * fire->dt_min_scale);
  f.p.n_min = fire->n_min;
  f.p.n_bad_max = fire->n_bad_max;
  f.st = (FireDev*)ctx->fireDev.p;
  f.part = (double*)((char*)ctx->fireDev.p + sizeof(FireDev));
  f.nparts = nparts;
  f.force = p->force;
  f.torque = zt;
  f.fixed = zfix;
  f.vel_pos = vpos;
  f.vel_rot = vrot;
  int32_t* moved = (int32_t*)((char*)ctx->fireDev.p + sizeof(FireDev) + 8 * (size_t)nparts);
  FireDev h0;
  std::memset(&h0, 0, sizeof(h0));
  h0.dt = (float)fire->dt;
  h0.alpha = (float)fire->alpha_init;
  HIPCHK(ctx, hipMemcpyAsync(f.st, &h0, sizeof(h0), hipMemcpyHostToDevice, st));
  double* per = (double*)ctx->red.p;
  double* wper = per + Nn;
  double* part = wper + Nn * D * D;
  double* out = part + 256 * D * D;
  PrecArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = N;
  a.dim = D;
  for (int q = 0; q < 3; ++q) {
    a.dims[q] = q < D ? p->dims[q] : 1;
    a.h[q] = q < D ? p->cell[q] : 1.0f;
    a.hlo[q] = q < D ? p->cell_lo[q] : 0.0f;
  }
  a.cidx = p->cidx;
  a.off = p->off;
  a.rad = p->rad;
  a.k = p->k;
  a.keys = (uint32_t*)ctx->wKeys.p;
  a.starts = (int32_t*)ctx->cstart.p;
  a.sp = (float4*)ctx->spos.p;
  a.force = p->force;
  a.e_out = per;
  a.w_out = nullptr;
  a.moved = moved;
  a.gmax_bits = &f.st->max_grad_bits;
  // With the counting sort the re-sort is decided on the device: its kernels
  // run only while a sphere has changed cell, so a step needs no host read
  // of the flag.  The radix fallback keeps the host-side decision.
  const bool gated = use_cell_sort(N, n_cells);
  a.gated = gated ? 1 : 0;
  // The Verlet list (with the counting sort only): pairs within r_i + r_j +
  // skin; a partner is at most 1 + ceil(skin / h) cells away on an axis
  // when h >= 2 r_max, so skin <= h keeps the build stencil at half-width
  // 2 (cell differences in one byte of the entry), which must not wrap onto
  // itself; slots take the entry's low 24 bits.
  int32_t* vovf = nullptr;
  if (gated && N > 0 && N < (int64_t)1 << 24 && p->skin > 0.0f && p->max_neighbors > 0) {
    float hmin = p->cell[0];
    for (int q = 1; q < D; ++q) hmin = std::min(hmin, p->cell[q]);
    const int sr = 1 + (int)std::ceil(p->skin / hmin);
    bool fits = sr <= 2;
    for (int q = 0; q < D; ++q) fits = fits && p->dims[q] >= 2 * sr + 1;
    if (fits) {
      ENSURE(ctx, ctx->pvl, 4 * Nn * (size_t)p->max_neighbors);
      ENSURE(ctx, ctx->pvaux, 4 * Nn * (size_t)D + 64);
      a.vl = (int32_t*)ctx->pvl.p;
      a.vm = p->max_neighbors;
      a.vsr = sr;
      a.vcut = p->skin;
      a.vtrig = 0.25f * p->skin * p->skin;
      a.boff = (float*)ctx->pvaux.p;
      vovf = (int32_t*)(a.boff + D * Nn);
      a.vovf = vovf;
      HIPCHK(ctx, launch_fill_i32(vovf, 1, 0, st));
    }
  }
  auto list_overflow = [&]() -> int {
    int32_t ov = 0;
    if (!vovf) return JDEM_OK;
    HIPCHK(ctx, hipMemcpyAsync(&ov, vovf, sizeof(ov), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    return ov ? fail(ctx, JDEM_EINVAL, "a neighbour row exceeds max_neighbors") : JDEM_OK;
  };
  bool sorted = false;
  double pe = 0.0, gmax = 0.0;
  auto eval = [&](bool with_virial) -> int {
    if ((gated || !sorted) && N > 0) {  // the cell list, rebuilt when a sphere changed cell
      HIPCHK(ctx, launch_prec_keys(a, st));
      uint32_t* sk = nullptr;
      int32_t* perm = nullptr;
      if (gated) {
        HIPCHK(ctx, cell_sort_pairs(ctx, a.keys, N, n_cells, &sk, &perm, (int32_t*)ctx->cstart.p, st, moved));
      } else {
        HIPCHK(ctx, radix_sort_pairs(ctx, a.keys, N, nbits_for(n_cells), &sk, &perm, st));
        HIPCHK(ctx, ws_launch_cell_starts(sk, N, n_cells, (int32_t*)ctx->cstart.p, st));
      }
      a.perm = perm;
      sorted = true;
    }
    a.w_out = with_virial ? wper : nullptr;
    HIPCHK(ctx, launch_prec_gather(a, st));
    if (a.vl) HIPCHK(ctx, launch_prec_vbuild(a, st));
    const int sp = span_begin(ctx, 0, 0);  // the force kernel, for its roofline
    HIPCHK(ctx, a.vl ? launch_prec_vforce(a, st) : launch_prec_force(a, st));
    span_end(ctx, sp);
    if (N == 0) HIPCHK(ctx, hipMemsetAsync(&f.st->max_grad_bits, 0, sizeof(unsigned), st));  // else the gather clears it
    HIPCHK(ctx, launch_fire_grad_max(f, st));
    double h = 0.0;
    if (N > 0) {
      HIPCHK(ctx, launch_sum_f64(per, N, 1, 1, part, out, 256, st));
      HIPCHK(ctx, hipMemcpyAsync(&h, out, 8, hipMemcpyDeviceToHost, st));
    }
    unsigned bits = 0;
    HIPCHK(ctx, hipMemcpyAsync(&bits, &f.st->max_grad_bits, sizeof(bits), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    float g;
    std::memcpy(&g, &bits, sizeof(g));
    pe = h;
    gmax = (double)g;
    return list_overflow();
  };
  // the first evaluation sorts: the flag starts set
  HIPCHK(ctx, launch_fill_i32(moved, 1, 1, st));
  int rc = eval(false);
  if (rc) return rc;
  double prev = INFINITY;
  int64_t steps = 0;
  const double tiny = 2.2250738585072014e-308;
  const double dN = (double)std::max<int64_t>(N, 1);
  // progress: a new low of max|F|, or of the energy by more than 1e-13 of it
  double best_g = gmax, best_e = pe;
  int64_t best_at = 0;
  if (gated && N > 0) {
    // The stopping tests run on the device (k_prec_check before every
    // update; every kernel after a stop returns at once), so the host reads
    // the loop's state once per chunk of iterations, not once per iteration.
    PrecStop* ps = (PrecStop*)((char*)ctx->fireDev.p + sizeof(FireDev) + 8 * (size_t)nparts + 64);
    PrecStop h_ps;
    std::memset(&h_ps, 0, sizeof(h_ps));
    h_ps.prev = INFINITY;
    h_ps.first = 1;
    HIPCHK(ctx, hipMemcpyAsync(ps, &h_ps, sizeof(h_ps), hipMemcpyHostToDevice, st));
    PrecStopP pp;
    pp.pe_tol = pe_tol;
    pp.pe_diff_tol = pe_diff_tol;
    pp.force_tol = force_tol;
    pp.dN = dN;
    pp.max_steps = max_steps;
    pp.stall_steps = p->stall_steps;
    a.stop = &ps->stop;
    f.stop = &ps->stop;
    constexpr int kChunk = 64;
    while (true) {
      for (int k = 0; k < kChunk; ++k) {
        HIPCHK(ctx, launch_prec_check(ps, out, &f.st->max_grad_bits, pp, st));
        HIPCHK(ctx, launch_fire_power(f, st));
        HIPCHK(ctx, launch_fire_decide(f, st));
        HIPCHK(ctx, launch_prec_move(f, a, st));
        HIPCHK(ctx, launch_fire_commit(f, st));
        // the evaluation (cell list re-sorted while a sphere changed cell)
        HIPCHK(ctx, launch_prec_keys(a, st));
        uint32_t* sk = nullptr;
        int32_t* perm = nullptr;
        HIPCHK(ctx, cell_sort_pairs(ctx, a.keys, N, n_cells, &sk, &perm, (int32_t*)ctx->cstart.p, st, moved));
        a.perm = perm;
        HIPCHK(ctx, launch_prec_gather(a, st));
        if (a.vl) HIPCHK(ctx, launch_prec_vbuild(a, st));
        const int sp = span_begin(ctx, 0, k);
        HIPCHK(ctx, a.vl ? launch_prec_vforce(a, st) : launch_prec_force(a, st));
        span_end(ctx, sp);
        HIPCHK(ctx, launch_fire_grad_max(f, st));
        HIPCHK(ctx, launch_sum_f64(per, N, 1, 1, part, out, 256, st));
      }
      HIPCHK(ctx, hipMemcpyAsync(&h_ps, ps, sizeof(h_ps), hipMemcpyDeviceToHost, st));
      HIPCHK(ctx, hipStreamSynchronize(st));
      rc = list_overflow();
      if (rc) return rc;
      if (h_ps.stop) break;
    }
    steps = h_ps.steps;
    pe = h_ps.pe;
    gmax = h_ps.gmax;
    a.stop = nullptr;
    f.stop = nullptr;
  }
  while (!(gated && N > 0) && steps < max_steps) {
    const bool conv_pe = std::fabs(pe / dN) <= pe_tol;
    const double denom = std::max(std::max(std::fabs(pe), std::fabs(prev)), tiny);
    const bool conv_rel = std::fabs(pe - prev) / denom < pe_diff_tol;
    if (gmax < best_g) best_g = gmax, best_at = steps;
    if (pe < best_e - 1e-13 * std::fabs(best_e)) best_e = pe, best_at = steps;
    const bool stalled = p->stall_steps > 0 && steps - best_at >= p->stall_steps;
    if (conv_pe || conv_rel || gmax <= force_tol || stalled) break;
    HIPCHK(ctx, launch_fire_power(f, st));
    HIPCHK(ctx, launch_fire_decide(f, st));
    if (!gated) HIPCHK(ctx, launch_fill_i32(moved, 1, 0, st));
    HIPCHK(ctx, launch_prec_move(f, a, st));
    HIPCHK(ctx, launch_fire_commit(f, st));
    prev = pe;
    if (!gated) {
      // a sphere that left its cell needs a new cell list before the evaluation
      int32_t mv = 0;
      HIPCHK(ctx, hipMemcpyAsync(&mv, moved, sizeof(mv), hipMemcpyDeviceToHost, st));
      HIPCHK(ctx, hipStreamSynchronize(st));
      if (mv) sorted = false;
    }
    rc = eval(false);
    if (rc) return rc;
    ++steps;
  }
  if (a.vl) {
    // list mode: fold the offsets back into [0, h) (a rebuild) for the caller
    HIPCHK(ctx, launch_fill_i32(moved, 1, 1, st));
    rc = eval(virial != nullptr);
    if (rc) return rc;
  }
  if (virial) {
    if (!a.vl) rc = eval(true);  // same configuration: the virial of the final state
    if (rc) return rc;
    double w9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (N > 0) {
      HIPCHK(ctx, launch_sum_f64(wper, N, D * D, D * D, part, out, 256, st));
      HIPCHK(ctx, hipMemcpyAsync(w9, out, 8 * (size_t)(D * D), hipMemcpyDeviceToHost, st));
      HIPCHK(ctx, hipStreamSynchronize(st));
    }
    for (int q = 0; q < D * D; ++q) virial[q] = w9[q];
  }
  if (stats) {
    FireDev h;
    HIPCHK(ctx, hipMemcpyAsync(&h, f.st, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    stats->steps = steps;
    stats->pe = pe / dN;
    stats->max_grad = gmax;
    stats->power = h.power;
    stats->overflow = 0;
  }
  return JDEM_OK;
}

int jdem_system_step(jdem_ctx* ctx, jdem_system* sys, int64_t n, jdem_step_stats* stats) {
  int rc = check_system(ctx, sys, false, true);
  if (rc) return rc;
  if (n < 0) return fail(ctx, JDEM_EINVAL, "n < 0");
  jdem_step_stats local;
  std::memset(&local, 0, sizeof(local));
  jdem_step_stats* S = stats ? stats : &local;
  std::memset(S, 0, sizeof(*S));
  if (n == 0) return JDEM_OK;
  // A CellList in a FreeDomain hashes on the box FreeDomain.apply fits to the
  // spheres every step (free.py:42-58), so its grid is rebuilt per step.
  const bool free_cells = sys->collider_kind == JDEM_COLLIDER_CELL_LIST && sys->domain.kind == JDEM_DOMAIN_FREE;
  (void)hipSetDevice(ctx->device);
  rc = ensure_flags(ctx);
  if (rc) return rc;
  jdem_state& s = sys->state;
  const int64_t N = s.n;
  const int D = s.dim;
  const int A = D == 3 ? 3 : 1;
  if (N == 0) {
    S->steps = n;
    return JDEM_OK;
  }
  if (s.clump_id) return clump_step(ctx, sys, n, S);
  if (sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST) return resident_nl_step(ctx, sys, n, S);
  ENSURE(ctx, ctx->pos2, 4 * (size_t)N * D);
  ENSURE(ctx, ctx->vel2, 4 * (size_t)N * D);
  ENSURE(ctx, ctx->ang2, 4 * (size_t)N * A);
  float* P[2] = {s.pos, (float*)ctx->pos2.p};
  float* V[2] = {s.vel, (float*)ctx->vel2.p};
  float* W[2] = {s.ang_vel, (float*)ctx->ang2.p};
  // kernel t (0..n) writes buffer wr(t) so that the last one writes the
  // caller's arrays; it reads the buffer kernel t-1 wrote.
  auto wr = [&](int64_t t) { return (int)((n - t) & 1); };
  auto rd = [&](int64_t t) { return t == 0 ? 0 : wr(t - 1); };
  const bool is_nl = sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST;
  const int src = src_of(sys->collider_kind);
  int32_t* flags = (int32_t*)ctx->flags.p;
  HIPCHK(ctx, launch_fill_i32(flags, 4, 0, ctx->stream));
  bool free_done = false;
  float fbox[3] = {0, 0, 0}, fanc[3] = {0, 0, 0};

  jdem_system fsys;  // free_cells: the system with this step's fitted box and grid
  float bx[3] = {0, 0, 0}, an[3] = {0, 0, 0};
  auto launch_t = [&](int64_t t) -> int {
    const int r = rd(t), w = wr(t);
    jdem_system* ss = sys;
    if (free_cells && N > 0 && t <= 1) {
      // step 1's box: the input positions, before kernel 0 (which may drift
      // them in place) runs
      if (t == 0) {
        int rc2 = free_box(ctx, P[r], sys, bx, an);
        if (rc2) return rc2;
        S->host_syncs++;
      }
    } else if (free_cells && N > 0) {
      // step t applies the domain before its drift (system.py:230-260): the
      // box of the positions kernel t-1 read, still held in this kernel's
      // output buffer for t >= 2
      int rc2 = free_box(ctx, P[w], sys, bx, an);
      if (rc2) return rc2;
      S->host_syncs++;
    }
    if (free_cells && N > 0 && t >= 1) {
      // the grid on the fitted box (_partition.py:55-99, non-periodic:
      // ceil(box / cell) cells)
      fsys = *sys;
      jdem_grid& g = fsys.grid;
      double cells = 1.0;
      for (int q = 0; q < D; ++q) {
        fsys.domain.box_size[q] = bx[q];
        fsys.domain.inv_box_size[q] = 1.0f / bx[q];
        fsys.domain.anchor[q] = an[q];
        const float c = std::ceil(bx[q] / g.cell_size[q]);
        g.dims[q] = c >= 1.0f ? (int32_t)std::min(c, 2.0e9f) : 1;
        cells *= (double)g.dims[q];
      }
      if (cells > (double)INT32_MAX) return fail(ctx, JDEM_EUNSUPPORTED, "cell grid size out of the int32 hash range");
      g.strides[0] = 1;
      for (int q = 1; q < D; ++q) g.strides[q] = g.strides[q - 1] * g.dims[q - 1];
      g.n_cells = (int64_t)cells;
      ss = &fsys;
    }
    StepArgs a = base_args(ctx, ss);
    a.pos_in = P[r];
    a.vel_in = V[r];
    a.ang_in = W[r];
    a.pos_out = P[w];
    a.vel_out = V[w];
    a.ang_out = W[w];
    a.do_force = t >= 1;
    a.do_after = t >= 1;
    a.do_before = t < n;
    a.write_kin = 1;
    a.step_idx = (int32_t)t;
    if (sys->domain.kind == JDEM_DOMAIN_FREE && t == n - 1 && !free_done) {
      // the last FreeDomain.apply (step n) sees the positions kernel n-1 reads
      int rc2 = free_box(ctx, P[r], sys, fbox, fanc);
      if (rc2) return rc2;
      free_done = true;
      S->host_syncs++;
    }
    if (a.do_force && sys->collider_kind != JDEM_COLLIDER_NONE) {
      const int sp = sys->collider_kind == JDEM_COLLIDER_CELL_LIST ? span_begin(ctx, 2, t) : -1;
      int rc2 = prepare_source(ctx, ss, a, P[r]);
      if (rc2) return rc2;
      span_end(ctx, sp);
    }
    int src_eff = src;
    int law_eff = a.law.kind;
    if (sys->collider_kind == JDEM_COLLIDER_NONE) {
      a.force_zero = 1;  // the "" collider (colliders/__init__.py:218-222)
      src_eff = SRC_NAIVE;
      law_eff = JDEM_LAW_SPRING;
    }
    if (is_nl) {
      a.flag_rd = t >= 1 ? flags + (t & 1) : nullptr;
      a.flag_wr = flags + ((t + 1) & 1);
      a.check_disp = a.do_before ? 1 : 0;
      a.old_pos = sys->nlist.old_pos;
      a.trigger = nl_trigger(sys->nlist);
    }
    const int sp = span_begin(ctx, 0, t);
    hipError_t e = launch_step(a, D, law_eff, src_eff, ctx->stream);
    span_end(ctx, sp);
    if (e != hipSuccess) return fail(ctx, JDEM_EHIP, std::string("step kernel: ") + hipGetErrorString(e));
    return JDEM_OK;
  };

  if (!is_nl) {
    for (int64_t t = 0; t <= n; ++t) {
      rc = launch_t(t);
      if (rc) return rc;
    }
  } else {
    // Speculative chunks: every step kernel checks the rebuild word set by its
    // predecessor and turns into a no-op when a rebuild is pending; the host
    // reads the words once per chunk, rebuilds, and resumes at that step.
    jdem_nlist& L = sys->nlist;
    rc = launch_t(0);
    if (rc) return rc;
    bool pending = L.n_build_times == 0;
    int64_t next = 1;
    int64_t chunk_len = 8;
    int64_t last_rebuild_at = 0;
    int64_t interval = 0;
    bool rebuilt_here = false;
    while (next <= n) {
      if (pending) {
        const int sp = span_begin(ctx, 2, next);
        rc = nl_rebuild(ctx, sys, P[rd(next)]);
        if (rc) return rc;
        span_end(ctx, sp);
        HIPCHK(ctx, launch_fill_i32(flags, 2, 0, ctx->stream));
        S->rebuilds++;
        rebuilt_here = true;
        if (last_rebuild_at > 0) interval = next - last_rebuild_at;
        last_rebuild_at = next;
        pending = false;
        if (interval > 0) chunk_len = std::max<int64_t>(2, interval - 1);
      }
      const int64_t cnt = std::min<int64_t>(chunk_len, n - next + 1);
      const size_t chunk_spans = ctx->spans.size();
      for (int64_t t = next; t < next + cnt; ++t) {
        rc = launch_t(t);
        if (rc) return rc;
      }
      rc = read_flags(ctx, 3);
      if (rc) return rc;
      S->host_syncs++;
      const int32_t v = std::max(ctx->host_flags[0], ctx->host_flags[1]);
      if (v == 0) {
        next += cnt;
        chunk_len = std::min<int64_t>(chunk_len * 2, 1024);
      } else {
        // kernels v .. next+cnt-1 found the rebuild pending and did no work
        for (size_t k = chunk_spans; k < ctx->spans.size(); ++k) {
          auto& s = ctx->spans[k];
          if (s.kind == 0 && s.step >= v) s.kind = 1;
        }
        next = v;
        pending = true;
      }
    }
    // overflow of the most recent build (neighbor_list.py:582-589)
    if (rebuilt_here) L.overflow = ctx->host_flags[2] != 0;
    S->overflow = L.overflow;
  }
  if (free_done) {
    for (int a = 0; a < D; ++a) {
      sys->domain.box_size[a] = fbox[a];
      sys->domain.anchor[a] = fanc[a];
      sys->domain.inv_box_size[a] = 1.0f / fbox[a];
      S->free_box[a] = fbox[a];
      S->free_anchor[a] = fanc[a];
    }
  }
  S->steps = n;
  return JDEM_OK;
}

int jdem_rollout(jdem_ctx* ctx, jdem_system* sys, const jdem_frames* fr, jdem_step_stats* stats) {
  int rc = check_system(ctx, sys, false, true);
  if (rc) return rc;
  if (!fr || fr->n_frames < 0 || (fr->n_frames > 0 && !fr->steps)) return fail(ctx, JDEM_EINVAL, "bad frames");
  if (sys->collider_kind != JDEM_COLLIDER_NEIGHBOR_LIST)
    return fail(ctx, JDEM_EUNSUPPORTED, "rollout frames need the NeighborList collider");
  if (sys->domain.kind == JDEM_DOMAIN_FREE) return fail(ctx, JDEM_EUNSUPPORTED, "rollout frames with a FreeDomain");
  int64_t n = 0;
  for (int64_t f = 0; f < fr->n_frames; ++f) {
    if (fr->steps[f] < 0) return fail(ctx, JDEM_EINVAL, "negative frame stride");
    n += fr->steps[f];
  }
  jdem_state& s = sys->state;
  if (fr->n_frames > 0 && s.n > 0 &&
      (!fr->pos || !fr->vel || !fr->ang_vel || !fr->q || !fr->force || !fr->torque))
    return fail(ctx, JDEM_EINVAL, "null frame array");
  jdem_step_stats local;
  std::memset(&local, 0, sizeof(local));
  jdem_step_stats* S = stats ? stats : &local;
  std::memset(S, 0, sizeof(*S));
  if (s.n == 0 || fr->n_frames == 0) {
    S->steps = n;
    return JDEM_OK;
  }
  (void)hipSetDevice(ctx->device);
  rc = ensure_flags(ctx);
  if (rc) return rc;
  return resident_nl_step(ctx, sys, n, S, nullptr, fr);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// (f) the plain step in float64: the kernels of jdem_system_step,
// jdem_compute_force and the energies instantiated in double.
// ---------------------------------------------------------------------------
namespace {

int check_system64(jdem_ctx* ctx, const jdem_system_f64* sys) {
  if (!ctx) return JDEM_EINVAL;
  if (!sys) return fail(ctx, JDEM_EINVAL, "null system");
  const jdem_state_f64& s = sys->state;
  if (s.dim != 2 && s.dim != 3) return fail(ctx, JDEM_EINVAL, "dim must be 2 or 3");
  if (s.n < 0 || s.n > (int64_t)INT32_MAX) return fail(ctx, JDEM_EINVAL, "particle count out of range");
  if (s.n > 0 && (!s.pos || !s.vel || !s.ang_vel || !s.q || !s.force || !s.torque || !s.rad || !s.mass ||
                  !s.inertia || !s.fixed))
    return fail(ctx, JDEM_EINVAL, "null state array");
  if (s.bond_width > 0 && !s.bond_id) return fail(ctx, JDEM_EINVAL, "bond_width > 0 but bond_id is null");
  if (sys->law.n_mat > 1 && !s.mat_id) return fail(ctx, JDEM_EINVAL, "n_mat > 1 but mat_id is null");
  if (s.clump_id || s.pos_p) {
    if (s.n > 0 && (!s.clump_id || !s.pos_p)) return fail(ctx, JDEM_EINVAL, "clumps need both pos_p and clump_id");
  }
  const int ck = sys->collider_kind;
  if (ck != JDEM_COLLIDER_NONE && ck != JDEM_COLLIDER_NAIVE && ck != JDEM_COLLIDER_CELL_LIST &&
      ck != JDEM_COLLIDER_NEIGHBOR_LIST)
    return fail(ctx, JDEM_EINVAL, "unknown collider kind");
  if (ck == JDEM_COLLIDER_NEIGHBOR_LIST) {
    const jdem_nlist_f64& L = sys->nlist;
    if (L.max_neighbors < 0) return fail(ctx, JDEM_EINVAL, "max_neighbors < 0");
    if (s.n > 0 && L.max_neighbors > 0 && (!L.nl || !L.old_pos))
      return fail(ctx, JDEM_EINVAL, "neighbour-list buffers missing");
    if (L.secondary != JDEM_COLLIDER_CELL_LIST && L.secondary != JDEM_COLLIDER_NAIVE)
      return fail(ctx, JDEM_EINVAL, "secondary collider must be CellList or naive");
  }
  const int lk = sys->law.kind;
  if (ck != JDEM_COLLIDER_NONE && !(lk >= JDEM_LAW_SPRING && lk <= JDEM_LAW_ROUTED))
    return fail(ctx, JDEM_EINVAL, "unknown law kind");
  if (sys->law.n_mat > kMatMask) return fail(ctx, JDEM_EINVAL, "more than 32767 materials");
  if (lk == JDEM_LAW_ROUTED) {
    if (sys->law.n_mat < 1) return fail(ctx, JDEM_EINVAL, "routed law: n_mat < 1");
    const int rc = prepare_route(ctx, sys->law, s.n);
    if (rc) return rc;
  } else if (sys->law.n_mat < 1 || !sys->law.t0) {
    if (ck != JDEM_COLLIDER_NONE && lk != JDEM_LAW_CUNDALL_STRACK)
      return fail(ctx, JDEM_EINVAL, "law table t0 missing");
  }
  if (ck != JDEM_COLLIDER_NONE && law_has_history(sys->law))
    return fail(ctx, JDEM_EUNSUPPORTED, "float64: the history laws run in float32 only");
  if (sys->domain.kind < JDEM_DOMAIN_FREE || sys->domain.kind > JDEM_DOMAIN_REFLECT)
    return fail(ctx, JDEM_EINVAL, "unknown domain kind");
  if (ck == JDEM_COLLIDER_CELL_LIST ||
      (ck == JDEM_COLLIDER_NEIGHBOR_LIST && sys->nlist.secondary == JDEM_COLLIDER_CELL_LIST)) {
    const jdem_grid_f64& g = ck == JDEM_COLLIDER_CELL_LIST ? sys->grid : sys->nlist.grid;
    if (g.n_cells <= 0 || g.n_cells > (int64_t)INT32_MAX)
      return fail(ctx, JDEM_EUNSUPPORTED, "cell grid size out of the int32 hash range");
    if (g.search_range < 1) return fail(ctx, JDEM_EINVAL, "search_range must be >= 1");
  }
  return JDEM_OK;
}

StepArgs64 base_args64(jdem_ctx* ctx, const jdem_system_f64* sys) {
  StepArgs64 a;
  std::memset(&a, 0, sizeof(a));
  const jdem_state_f64& s = sys->state;
  a.n = s.n;
  a.bond_w = s.bond_width;
  a.isb = sys->interact_same_bond_id;
  a.pos_in = s.pos;
  a.vel_in = s.vel;
  a.ang_in = s.ang_vel;
  a.pos_out = s.pos;
  a.vel_out = s.vel;
  a.ang_out = s.ang_vel;
  a.q = s.q;
  a.force = s.force;
  a.torque = s.torque;
  a.rad = s.rad;
  a.mass = s.mass;
  a.inertia = s.inertia;
  a.mat_id = s.mat_id;
  a.bond_id = s.bond_id;
  a.fixed = s.fixed;
  a.ext_f = sys->ext_force;
  a.ext_fc = sys->ext_force_com;
  a.ext_t = sys->ext_torque;
  for (int c = 0; c < 3; ++c) a.g[c] = sys->gravity[c];
  a.has_fm = (sys->gravity[0] != 0.0 || sys->gravity[1] != 0.0 || sys->gravity[2] != 0.0 || a.ext_f || a.ext_fc ||
              a.ext_t)
                 ? 1
                 : 0;
  a.law = law_of(ctx, sys->law);
  a.dom = dom_of(sys->domain);
  a.dt = sys->dt;
  a.lin = sys->lin_kind;
  a.rot = sys->rot_kind;
  a.M = sys->nlist.max_neighbors;
  a.W32 = (a.M + 31) / 32;
  a.nl = sys->nlist.nl;
  return a;
}

// Candidate pairs within `cutoff` into nl_out (N, M) through the secondary
// collider `kind` (cell_list.py:500-595, naive.py:122-185), float64.
int build_nlist64(jdem_ctx* ctx, const jdem_system_f64* sys, int kind, const jdem_grid_f64& grid, const double* pos,
                  double cutoff, int M, int32_t* nl_out) {
  const jdem_state_f64& s = sys->state;
  NLBuildT<double> b;
  std::memset(&b, 0, sizeof(b));
  b.n = s.n;
  b.dim = s.dim;
  b.M = M;
  b.cutoff_sq = cutoff * cutoff;
  b.isb = sys->interact_same_bond_id;
  b.bond_w = s.bond_width;
  b.bond_id = s.bond_id;
  b.pos = pos;
  b.dom = dom_of(sys->domain);
  b.nl = nl_out;
  b.overflow = (int32_t*)ctx->flags.p + 2;
  b.clump = s.clump_id;
  if (kind == JDEM_COLLIDER_CELL_LIST) {
    Cells c;
    b.grid = grid_of(grid);
    int rc = build_cells(ctx, pos, s.rad, s.n, s.dim, b.dom, b.grid, &c);
    if (rc) return rc;
    b.perm = c.perm;
    b.skeys = c.skeys;
    b.cstart = c.cstart;
    b.cend = c.cend;
    HIPCHK(ctx, launch_nlist_cells(b, ctx->stream));
  } else {
    HIPCHK(ctx, launch_nlist_naive(b, ctx->stream));
  }
  return JDEM_OK;
}

// NeighborList rebuild at positions `pos` in float64 (neighbor_list.py:
// 58-131, 545-589): new list, old_pos <- pos, n_build_times += 1; flags[2]
// receives the overflow.
int nl_rebuild64(jdem_ctx* ctx, jdem_system_f64* sys, const double* pos) {
  jdem_nlist_f64& L = sys->nlist;
  const jdem_state_f64& s = sys->state;
  const int64_t n = s.n;
  const int M = L.max_neighbors;
  const double lc = L.cutoff + L.skin;
  HIPCHK(ctx, launch_fill_i32((int32_t*)ctx->flags.p + 2, 1, 0, ctx->stream));
  if (M > 0 && n > 0) {
    ENSURE(ctx, ctx->nl_new, 4 * (size_t)n * M);
    int32_t* nl_new = (int32_t*)ctx->nl_new.p;
    int rc = build_nlist64(ctx, sys, L.secondary, L.grid, pos, lc, M, nl_new);
    if (rc) return rc;
    HIPCHK(ctx, hipMemcpyAsync(L.nl, nl_new, 4 * (size_t)n * M, hipMemcpyDeviceToDevice, ctx->stream));
  }
  if (n > 0 && pos != L.old_pos)
    HIPCHK(ctx, hipMemcpyAsync(L.old_pos, pos, 8 * (size_t)n * s.dim, hipMemcpyDeviceToDevice, ctx->stream));
  L.n_build_times += 1;
  return JDEM_OK;
}

// (skin / 2)^2 of the displacement check (neighbor_list.py:78-85).
double nl_trigger64(const jdem_nlist_f64& L) { return (L.skin * L.skin) / 4.0; }

// _check_and_rebuild on its own (neighbor_list.py:58-131), float64.
int nl_check_and_rebuild64(jdem_ctx* ctx, jdem_system_f64* sys) {
  jdem_nlist_f64& L = sys->nlist;
  const jdem_state_f64& s = sys->state;
  int rc = ensure_flags(ctx);
  if (rc) return rc;
  bool need = L.n_build_times == 0;
  if (!need && s.n > 0) {
    HIPCHK(ctx, launch_fill_i32((int32_t*)ctx->flags.p + 3, 1, 0, ctx->stream));
    hipLaunchKernelGGL(k_disp_check<double>, dim3(nblocks(s.n)), dim3(kBlock), 0, ctx->stream, s.pos, L.old_pos, s.n,
                       s.dim, nl_trigger64(L), (int32_t*)ctx->flags.p + 3);
    HIPCHK(ctx, hipGetLastError());
    rc = read_flags(ctx, 4);
    if (rc) return rc;
    need = ctx->host_flags[3] != 0;
  }
  if (need) {
    rc = nl_rebuild64(ctx, sys, s.pos);
    if (rc) return rc;
    rc = read_flags(ctx, 3);
    if (rc) return rc;
    L.overflow = ctx->host_flags[2] != 0;
  }
  return JDEM_OK;
}

int prepare_source64(jdem_ctx* ctx, const jdem_system_f64* sys, StepArgs64& a, const double* pos) {
  if (sys->collider_kind == JDEM_COLLIDER_CELL_LIST) {
    Cells c;
    a.grid = grid_of(sys->grid);
    int rc = build_cells(ctx, pos, sys->state.rad, sys->state.n, sys->state.dim, a.dom, a.grid, &c);
    if (rc) return rc;
    a.perm = c.perm;
    a.skeys = c.skeys;
    a.cstart = c.cstart;
    a.cend = c.cend;
    a.spos = nullptr;
  }
  return JDEM_OK;
}

// FreeDomain.apply (free.py:42-58) in float64.
int free_box64(jdem_ctx* ctx, const double* pos, const jdem_system_f64* sys, double box[3], double anchor[3]) {
  const jdem_state_f64& s = sys->state;
  const int nparts = 64;
  ENSURE(ctx, ctx->red, 8 * 6 * (size_t)nparts);
  HIPCHK(ctx, launch_box_minmax(pos, s.rad, s.n, s.dim, (double*)ctx->red.p, nparts, ctx->stream, s.pos_p));
  std::vector<double> hp(6 * nparts);
  HIPCHK(ctx, hipMemcpyAsync(hp.data(), ctx->red.p, 8 * hp.size(), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int a = 0; a < s.dim; ++a) {
    double lo = INFINITY, hi = -INFINITY;
    for (int b = 0; b < nparts; ++b) {
      lo = std::min(lo, hp[6 * b + a]);
      hi = std::max(hi, hp[6 * b + 3 + a]);
    }
    box[a] = hi - lo;
    anchor[a] = lo;
  }
  return JDEM_OK;
}

// Per-particle float64 terms summed on the device: the potential energy
// (collider pairs) or the two kinetic parts.
int reduce64(jdem_ctx* ctx, const jdem_system_f64* sys, int ncomp, double* host_out, bool kinetic) {
  const jdem_state_f64& s = sys->state;
  const int64_t n = s.n;
  const int nparts = 256;
  ENSURE(ctx, ctx->red, 8 * ((size_t)(n > 0 ? n : 1) * ncomp + (size_t)nparts * ncomp + ncomp));
  double* per = (double*)ctx->red.p;
  double* part = per + (size_t)(n > 0 ? n : 1) * ncomp;
  double* out = part + (size_t)nparts * ncomp;
  StepArgs64 a = base_args64(ctx, sys);
  if (!kinetic) {
    int rc = prepare_source64(ctx, sys, a, s.pos);
    if (rc) return rc;
  }
  a.e_out = per;
  if (n > 0) {
    if (kinetic) {
      HIPCHK(ctx, launch_kinetic(a, s.dim, ctx->stream));
    } else {
      HIPCHK(ctx, launch_energy(a, s.dim, a.law.kind, src_of(sys->collider_kind), ctx->stream));
    }
    HIPCHK(ctx, launch_sum_f64(per, n, ncomp, ncomp, part, out, nparts, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(host_out, out, 8 * (size_t)ncomp, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  } else {
    for (int c = 0; c < ncomp; ++c) host_out[c] = 0.0;
  }
  return JDEM_OK;
}

}  // namespace

extern "C" {

int jdem_system_step_f64(jdem_ctx* ctx, jdem_system_f64* sys, int64_t n, jdem_step_stats* stats) {
  int rc = check_system64(ctx, sys);
  if (rc) return rc;
  if (n < 0) return fail(ctx, JDEM_EINVAL, "n < 0");
  jdem_step_stats local;
  std::memset(&local, 0, sizeof(local));
  jdem_step_stats* S = stats ? stats : &local;
  std::memset(S, 0, sizeof(*S));
  if (n == 0) return JDEM_OK;
  if (sys->collider_kind == JDEM_COLLIDER_CELL_LIST && sys->domain.kind == JDEM_DOMAIN_FREE)
    return fail(ctx, JDEM_EUNSUPPORTED, "CellList with a FreeDomain (per-step grid) in float64 is not implemented");
  (void)hipSetDevice(ctx->device);
  jdem_state_f64& s = sys->state;
  const int64_t N = s.n;
  const int D = s.dim;
  const int A = D == 3 ? 3 : 1;
  if (N == 0) {
    S->steps = n;
    return JDEM_OK;
  }
  if (s.clump_id) return clump_step(ctx, sys, n, S);
  ENSURE(ctx, ctx->pos2, 8 * (size_t)N * D);
  ENSURE(ctx, ctx->vel2, 8 * (size_t)N * D);
  ENSURE(ctx, ctx->ang2, 8 * (size_t)N * A);
  double* P[2] = {s.pos, (double*)ctx->pos2.p};
  double* V[2] = {s.vel, (double*)ctx->vel2.p};
  double* W[2] = {s.ang_vel, (double*)ctx->ang2.p};
  // kernel t (0..n) writes buffer wr(t), so that the last one writes the
  // caller's arrays; it reads the buffer kernel t-1 wrote (as
  // jdem_system_step).
  auto wr = [&](int64_t t) { return (int)((n - t) & 1); };
  auto rd = [&](int64_t t) { return t == 0 ? 0 : wr(t - 1); };
  bool free_done = false;
  double fbox[3] = {0, 0, 0}, fanc[3] = {0, 0, 0};
  // NeighborList: kernel t-1 checks the positions it writes against the
  // build positions (flags[3]); the host reads that word before kernel t and
  // rebuilds the list at those positions when it is set (neighbor_list.py:
  // 58-131 runs the same check inside compute_force).
  const bool is_nl = sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST;
  bool rebuilt_here = false;
  if (is_nl) {
    rc = ensure_flags(ctx);
    if (rc) return rc;
    HIPCHK(ctx, launch_fill_i32((int32_t*)ctx->flags.p, 4, 0, ctx->stream));
  }
  for (int64_t t = 0; t <= n; ++t) {
    if (is_nl && t >= 1) {
      bool need = sys->nlist.n_build_times == 0;
      if (!need) {
        rc = read_flags(ctx, 4);
        if (rc) return rc;
        S->host_syncs++;
        need = ctx->host_flags[3] != 0;
      }
      if (need) {
        rc = nl_rebuild64(ctx, sys, P[rd(t)]);
        if (rc) return rc;
        HIPCHK(ctx, launch_fill_i32((int32_t*)ctx->flags.p + 3, 1, 0, ctx->stream));
        S->rebuilds++;
        rebuilt_here = true;
      }
    }
    StepArgs64 a = base_args64(ctx, sys);
    const int r = rd(t), w = wr(t);
    a.pos_in = P[r];
    a.vel_in = V[r];
    a.ang_in = W[r];
    a.pos_out = P[w];
    a.vel_out = V[w];
    a.ang_out = W[w];
    a.do_force = t >= 1;
    a.do_after = t >= 1;
    a.do_before = t < n;
    a.write_kin = 1;
    a.step_idx = (int32_t)t;
    if (sys->domain.kind == JDEM_DOMAIN_FREE && t == n - 1 && !free_done) {
      // the last FreeDomain.apply (step n) sees the positions kernel n-1 reads
      rc = free_box64(ctx, P[r], sys, fbox, fanc);
      if (rc) return rc;
      free_done = true;
      S->host_syncs++;
    }
    if (a.do_force && sys->collider_kind != JDEM_COLLIDER_NONE) {
      rc = prepare_source64(ctx, sys, a, P[r]);
      if (rc) return rc;
    }
    int src_eff = src_of(sys->collider_kind);
    int law_eff = a.law.kind;
    if (sys->collider_kind == JDEM_COLLIDER_NONE) {
      a.force_zero = 1;  // the "" collider (colliders/__init__.py:218-222)
      src_eff = SRC_NAIVE;
      law_eff = JDEM_LAW_SPRING;
    }
    if (is_nl && a.do_before) {
      a.check_disp = 1;
      a.old_pos = sys->nlist.old_pos;
      a.trigger = nl_trigger64(sys->nlist);
      a.flag_wr = (int32_t*)ctx->flags.p + 3;
    }
    const hipError_t e = launch_step(a, D, law_eff, src_eff, ctx->stream);
    if (e != hipSuccess) return fail(ctx, JDEM_EHIP, std::string("step kernel: ") + hipGetErrorString(e));
  }
  if (is_nl) {
    // overflow of the most recent build (neighbor_list.py:582-589)
    if (rebuilt_here) {
      rc = read_flags(ctx, 3);
      if (rc) return rc;
      sys->nlist.overflow = ctx->host_flags[2] != 0;
    }
    S->overflow = sys->nlist.overflow;
  }
  if (free_done) {
    for (int a = 0; a < D; ++a) {
      sys->domain.box_size[a] = fbox[a];
      sys->domain.anchor[a] = fanc[a];
      sys->domain.inv_box_size[a] = 1.0 / fbox[a];
      S->free_box[a] = fbox[a];
      S->free_anchor[a] = fanc[a];
    }
  }
  S->steps = n;
  return JDEM_OK;
}

int jdem_compute_force_f64(jdem_ctx* ctx, jdem_system_f64* sys) {
  int rc = check_system64(ctx, sys);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  const jdem_state_f64& s = sys->state;
  if (sys->collider_kind == JDEM_COLLIDER_NONE) {
    HIPCHK(ctx, hipMemsetAsync(s.force, 0, 8 * (size_t)s.n * s.dim, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(s.torque, 0, 8 * (size_t)s.n * (s.dim == 3 ? 3 : 1), ctx->stream));
    return JDEM_OK;
  }
  if (s.clump_id) return clump_force(ctx, sys);
  if (sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST) {
    rc = nl_check_and_rebuild64(ctx, sys);
    if (rc) return rc;
  }
  StepArgs64 a = base_args64(ctx, sys);
  rc = prepare_source64(ctx, sys, a, s.pos);
  if (rc) return rc;
  a.do_force = 1;
  a.write_kin = 0;
  HIPCHK(ctx, launch_step(a, s.dim, a.law.kind, src_of(sys->collider_kind), ctx->stream));
  return JDEM_OK;
}

int jdem_potential_energy_f64(jdem_ctx* ctx, jdem_system_f64* sys, double* energy) {
  int rc = check_system64(ctx, sys);
  if (rc) return rc;
  if (!energy) return fail(ctx, JDEM_EINVAL, "null output");
  (void)hipSetDevice(ctx->device);
  if (sys->collider_kind == JDEM_COLLIDER_NONE) {
    *energy = 0.0;
    return JDEM_OK;
  }
  if (sys->state.clump_id) return clump_reduce(ctx, sys, 1, energy, false);
  if (sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST) {
    rc = nl_check_and_rebuild64(ctx, sys);
    if (rc) return rc;
  }
  return reduce64(ctx, sys, 1, energy, false);
}

int jdem_kinetic_energy_parts_f64(jdem_ctx* ctx, const jdem_system_f64* sys, double* translational,
                                  double* rotational) {
  int rc = check_system64(ctx, sys);
  if (rc) return rc;
  if (!translational || !rotational) return fail(ctx, JDEM_EINVAL, "null output");
  (void)hipSetDevice(ctx->device);
  double parts[2] = {0.0, 0.0};
  rc = sys->state.clump_id ? clump_reduce(ctx, const_cast<jdem_system_f64*>(sys), 2, parts, true)
                           : reduce64(ctx, sys, 2, parts, true);
  if (rc) return rc;
  *translational = parts[0];
  *rotational = parts[1];
  return JDEM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Clumps: rigid bodies of spheres on the plain step (state.py:230-304,
// force_manager.py:341-425, naive.py:210-233, cell_list.py:437-464), for
// both scalar types.  A step is five launches: members' lab positions and
// levers, the pair forces at the lab positions (same-clump pairs masked,
// lever torque added), ForceManager's member totals, their sum over each
// clump, and the integrators (after-force of this step, domain.apply and
// before-force of the next) on the clump state every member holds.
// ---------------------------------------------------------------------------
namespace {

StepArgs args_of(jdem_ctx* ctx, const jdem_system* sys) { return base_args(ctx, sys); }
StepArgs64 args_of(jdem_ctx* ctx, const jdem_system_f64* sys) { return base_args64(ctx, sys); }
int source_of(jdem_ctx* ctx, jdem_system* sys, StepArgs& a, const float* pos) {
  return prepare_source(ctx, sys, a, pos);
}
int source_of(jdem_ctx* ctx, jdem_system_f64* sys, StepArgs64& a, const double* pos) {
  return prepare_source64(ctx, sys, a, pos);
}
int box_of(jdem_ctx* ctx, const float* pos, const jdem_system* sys, float box[3], float anchor[3]) {
  return free_box(ctx, pos, sys, box, anchor);
}
int box_of(jdem_ctx* ctx, const double* pos, const jdem_system_f64* sys, double box[3], double anchor[3]) {
  return free_box64(ctx, pos, sys, box, anchor);
}

// R(q) pos_p and pos_c + R(q) pos_p of every member into the context's
// buffers (a.pos_in holds pos_c).
template <class Real, class Sys>
int clump_lab(jdem_ctx* ctx, const Sys* sys, StepArgsT<Real>& a) {
  const auto& s = sys->state;
  const size_t nd = sizeof(Real) * (size_t)s.n * s.dim;
  ENSURE(ctx, ctx->plab, nd);
  ENSURE(ctx, ctx->ppr, nd);
  a.plab = (Real*)ctx->plab.p;
  a.ppr = (Real*)ctx->ppr.p;
  a.clump = s.clump_id;
  a.pos_p = s.pos_p;
  HIPCHK(ctx, launch_clump_lab(a, s.dim, ctx->stream));
  return JDEM_OK;
}

// The members of each clump in ascending index (a stable counting sort of
// the ids, which lie in [0, N)) and the buffers of the member totals.
template <class Real, class Sys>
int clump_members(jdem_ctx* ctx, const Sys* sys, StepArgsT<Real>& a) {
  const auto& s = sys->state;
  const int64_t N = s.n;
  const int A = s.dim == 3 ? 3 : 1;
  ENSURE(ctx, ctx->cl_start, 4 * ((size_t)N + 1));
  ENSURE(ctx, ctx->cl_members, 4 * (size_t)N);
  ENSURE(ctx, ctx->ftot, sizeof(Real) * (size_t)N * s.dim);
  ENSURE(ctx, ctx->ttot, sizeof(Real) * (size_t)N * A);
  uint32_t* sk = nullptr;
  int32_t* perm = nullptr;
  HIPCHK(ctx, cell_sort_pairs(ctx, reinterpret_cast<const uint32_t*>(s.clump_id), N, N, &sk, &perm,
                              (int32_t*)ctx->cl_start.p, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->cl_members.p, perm, 4 * (size_t)N, hipMemcpyDeviceToDevice, ctx->stream));
  a.clump = s.clump_id;
  a.cl_start = (const int32_t*)ctx->cl_start.p;
  a.cl_members = (const int32_t*)ctx->cl_members.p;
  a.ftot = (Real*)ctx->ftot.p;
  a.ttot = (Real*)ctx->ttot.p;
  return JDEM_OK;
}

int rebuild_of(jdem_ctx* ctx, jdem_system* sys, const float* pos) { return nl_rebuild(ctx, sys, pos); }
int rebuild_of(jdem_ctx* ctx, jdem_system_f64* sys, const double* pos) { return nl_rebuild64(ctx, sys, pos); }
float trigger_of(const jdem_nlist& L) { return nl_trigger(L); }
double trigger_of(const jdem_nlist_f64& L) { return nl_trigger64(L); }

// _check_and_rebuild on the members' lab positions (neighbor_list.py:58-131
// compares state.pos, the lab positions, with the build positions).
template <class Real, class Sys>
int clump_nl_check(jdem_ctx* ctx, Sys* sys, const Real* plab, bool* rebuilt) {
  auto& L = sys->nlist;
  const auto& s = sys->state;
  int rc = ensure_flags(ctx);
  if (rc) return rc;
  bool need = L.n_build_times == 0;
  if (!need) {
    HIPCHK(ctx, launch_fill_i32((int32_t*)ctx->flags.p + 3, 1, 0, ctx->stream));
    hipLaunchKernelGGL(k_disp_check<Real>, dim3(nblocks(s.n)), dim3(kBlock), 0, ctx->stream, plab, L.old_pos, s.n,
                       s.dim, trigger_of(L), (int32_t*)ctx->flags.p + 3);
    HIPCHK(ctx, hipGetLastError());
    rc = read_flags(ctx, 4);
    if (rc) return rc;
    need = ctx->host_flags[3] != 0;
  }
  if (need) {
    rc = rebuild_of(ctx, sys, plab);
    if (rc) return rc;
    *rebuilt = true;
  }
  return JDEM_OK;
}

template <class Sys>
int clump_nl_overflow(jdem_ctx* ctx, Sys* sys) {
  int rc = read_flags(ctx, 3);
  if (rc) return rc;
  sys->nlist.overflow = ctx->host_flags[2] != 0;
  return JDEM_OK;
}

// Pair forces of the members at their lab positions (a from clump_lab).
template <class Real, class Sys>
int clump_pair_forces(jdem_ctx* ctx, Sys* sys, const StepArgsT<Real>& a0) {
  StepArgsT<Real> f = a0;
  f.pos_in = a0.plab;
  f.do_force = 1;
  f.do_after = 0;
  f.do_before = 0;
  f.write_kin = 0;
  int law = f.law.kind, src = src_of(sys->collider_kind);
  if (sys->collider_kind == JDEM_COLLIDER_NONE) {
    f.force_zero = 1;  // the "" collider (colliders/__init__.py:218-222)
    law = JDEM_LAW_SPRING;
    src = SRC_NAIVE;
  } else if (sys->collider_kind == JDEM_COLLIDER_CELL_LIST) {
    int rc = source_of(ctx, sys, f, a0.plab);
    if (rc) return rc;
  }
  HIPCHK(ctx, launch_step(f, sys->state.dim, law, src, ctx->stream));
  return JDEM_OK;
}

template <class Real, class Sys>
int clump_step_t(jdem_ctx* ctx, Sys* sys, int64_t n, jdem_step_stats* S) {
  auto& s = sys->state;
  const int D = s.dim;
  StepArgsT<Real> a0 = args_of(ctx, sys);
  int rc = clump_members(ctx, sys, a0);
  if (rc) return rc;
  const size_t nd = sizeof(Real) * (size_t)s.n * D;
  ENSURE(ctx, ctx->plab, nd);
  ENSURE(ctx, ctx->ppr, nd);
  a0.plab = (Real*)ctx->plab.p;
  a0.ppr = (Real*)ctx->ppr.p;
  a0.pos_p = s.pos_p;
  const bool none = sys->collider_kind == JDEM_COLLIDER_NONE;
  const bool free_dom = sys->domain.kind == JDEM_DOMAIN_FREE;
  const bool reflect = sys->domain.kind == JDEM_DOMAIN_REFLECT;
  if (reflect) {
    ENSURE(ctx, ctx->crefl, sizeof(Real) * (size_t)s.n * refl_scratch_words(D));
    a0.rscr = (Real*)ctx->crefl.p;
  }
  bool free_done = false;
  Real fbox[3] = {0, 0, 0}, fanc[3] = {0, 0, 0};
  // the integrators on the clump state, in place (no pair reads); with a
  // reflective domain, domain.apply runs as the clump kernels before the
  // before-force half (the sphere form in the step kernel is skipped)
  auto integrate = [&](bool after, bool before) -> int {
    if (reflect && before) {
      if (after) {
        int r2 = [&]() -> int {
          StepArgsT<Real> a = a0;
          a.do_force = 0;
          a.do_after = 1;
          a.do_before = 0;
          a.write_kin = 1;
          a.has_fm = 0;
          a.ppr = nullptr;
          HIPCHK(ctx, launch_step(a, D, none ? JDEM_LAW_SPRING : a.law.kind,
                                  none ? SRC_NAIVE : src_of(sys->collider_kind), ctx->stream));
          return JDEM_OK;
        }();
        if (r2) return r2;
        after = false;
      }
      HIPCHK(ctx, launch_clump_reflect(a0, D, ctx->stream));
    }
    StepArgsT<Real> a = a0;
    if (reflect) a.dom.kind = JDEM_DOMAIN_FREE;
    a.do_force = 0;
    a.do_after = after;
    a.do_before = before;
    a.write_kin = 1;
    a.has_fm = 0;  // ForceManager ran in the clump kernels
    a.ppr = nullptr;
    const hipError_t e =
        launch_step(a, D, none ? JDEM_LAW_SPRING : a.law.kind, none ? SRC_NAIVE : src_of(sys->collider_kind),
                    ctx->stream);
    if (e != hipSuccess) return fail(ctx, JDEM_EHIP, std::string("step kernel: ") + hipGetErrorString(e));
    return JDEM_OK;
  };
  // FreeDomain.apply of step n sees the positions after step n - 1
  auto free_at = [&]() -> int {
    int r2 = box_of(ctx, s.pos, sys, fbox, fanc);
    if (r2) return r2;
    free_done = true;
    S->host_syncs++;
    return JDEM_OK;
  };
  if (free_dom && n == 1) {
    rc = free_at();
    if (rc) return rc;
  }
  rc = integrate(false, true);  // domain.apply + before-force of step 1
  if (rc) return rc;
  const bool is_nl = sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST;
  bool rebuilt = false;
  for (int64_t t = 1; t <= n; ++t) {
    HIPCHK(ctx, launch_clump_lab(a0, D, ctx->stream));
    if (is_nl) {
      bool r2 = false;
      rc = clump_nl_check(ctx, sys, a0.plab, &r2);
      if (rc) return rc;
      S->host_syncs++;
      if (r2) {
        S->rebuilds++;
        rebuilt = true;
      }
    }
    rc = clump_pair_forces(ctx, sys, a0);
    if (rc) return rc;
    HIPCHK(ctx, launch_clump_forces(a0, D, ctx->stream));
    if (free_dom && t == n - 1) {
      rc = free_at();
      if (rc) return rc;
    }
    rc = integrate(true, t < n);
    if (rc) return rc;
  }
  if (is_nl) {
    if (rebuilt) {
      rc = clump_nl_overflow(ctx, sys);  // of the most recent build
      if (rc) return rc;
    }
    S->overflow = sys->nlist.overflow;
  }
  if (free_done) {
    for (int a = 0; a < D; ++a) {
      sys->domain.box_size[a] = fbox[a];
      sys->domain.anchor[a] = fanc[a];
      sys->domain.inv_box_size[a] = Real(1) / fbox[a];
      S->free_box[a] = fbox[a];
      S->free_anchor[a] = fanc[a];
    }
  }
  S->steps = n;
  return JDEM_OK;
}

template <class Real, class Sys>
int clump_force_t(jdem_ctx* ctx, Sys* sys) {
  StepArgsT<Real> a = args_of(ctx, sys);
  int rc = clump_lab(ctx, sys, a);
  if (rc) return rc;
  if (sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST) {
    bool rebuilt = false;
    rc = clump_nl_check(ctx, sys, a.plab, &rebuilt);
    if (rc) return rc;
    if (rebuilt) {
      rc = clump_nl_overflow(ctx, sys);
      if (rc) return rc;
    }
  }
  return clump_pair_forces(ctx, sys, a);
}

template <class Real, class Sys>
int clump_reduce_t(jdem_ctx* ctx, Sys* sys, int ncomp, double* host_out, bool kinetic) {
  const auto& s = sys->state;
  const int64_t n = s.n;
  const int nparts = 256;
  if (n == 0) {
    for (int c = 0; c < ncomp; ++c) host_out[c] = 0.0;
    return JDEM_OK;
  }
  StepArgsT<Real> a = args_of(ctx, sys);
  int rc = kinetic ? clump_members(ctx, sys, a) : clump_lab(ctx, sys, a);
  if (rc) return rc;
  ENSURE(ctx, ctx->red, 8 * ((size_t)n * ncomp + (size_t)nparts * ncomp + ncomp));
  double* per = (double*)ctx->red.p;
  double* part = per + (size_t)n * ncomp;
  double* out = part + (size_t)nparts * ncomp;
  a.e_out = per;
  if (kinetic) {
    HIPCHK(ctx, launch_kinetic(a, s.dim, ctx->stream));
  } else {
    a.pos_in = a.plab;
    if (sys->collider_kind == JDEM_COLLIDER_CELL_LIST) {
      rc = source_of(ctx, sys, a, a.plab);
      if (rc) return rc;
    } else if (sys->collider_kind == JDEM_COLLIDER_NEIGHBOR_LIST) {
      bool rebuilt = false;
      rc = clump_nl_check(ctx, sys, a.plab, &rebuilt);
      if (rc) return rc;
      if (rebuilt) {
        rc = clump_nl_overflow(ctx, sys);
        if (rc) return rc;
      }
    }
    HIPCHK(ctx, launch_energy(a, s.dim, a.law.kind, src_of(sys->collider_kind), ctx->stream));
  }
  HIPCHK(ctx, launch_sum_f64(per, n, ncomp, ncomp, part, out, nparts, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(host_out, out, 8 * (size_t)ncomp, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return JDEM_OK;
}

int clump_step(jdem_ctx* ctx, jdem_system* sys, int64_t n, jdem_step_stats* S) {
  return clump_step_t<float>(ctx, sys, n, S);
}
int clump_force(jdem_ctx* ctx, jdem_system* sys) { return clump_force_t<float>(ctx, sys); }
int clump_reduce(jdem_ctx* ctx, jdem_system* sys, int ncomp, double* host_out, bool kinetic) {
  return clump_reduce_t<float>(ctx, sys, ncomp, host_out, kinetic);
}
int clump_step(jdem_ctx* ctx, jdem_system_f64* sys, int64_t n, jdem_step_stats* S) {
  return clump_step_t<double>(ctx, sys, n, S);
}
int clump_force(jdem_ctx* ctx, jdem_system_f64* sys) { return clump_force_t<double>(ctx, sys); }
int clump_reduce(jdem_ctx* ctx, jdem_system_f64* sys, int ncomp, double* host_out, bool kinetic) {
  return clump_reduce_t<double>(ctx, sys, ncomp, host_out, kinetic);
}

}  // namespace
